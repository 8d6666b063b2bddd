// cubo_shift.hip — C ABI of the shifted mask stack and its adjoint (include/cubo_hip.h, "shifted stack").
//
// The host side of the bilinear shift (Sources/CuboVirtual/ShiftedImaging.swift:43-73, 109-122): the origin
// map's mean (NumPy's pairwise order), each position's integer shift wrapped into the detector and its
// fractions, the positions sorted by integer shift into tiles, and the rolled weight tables the tiles need.
// The device does the rest: the matrix-core pass for Y or for the origin adjoint, and the pixel-major pass
// for the mask adjoint, both re-reading the resident cube.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <thread>
#include <vector>

#include "../../include/cubo_hip.h"
#include "cube.hpp"
#include "errors.hpp"
#include "host_math.hpp"
#include "shift_kernels.hpp"
#include "stack_prep.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::hip_fail;
using cubo::pairwise_sum;
using cubo::wrap_shift;

struct cubo_hip_shift_plan {
    cubo_hip_cube *cube;
    int k, kp;
    // A cube the matrix-core path does not take (Float32, or a pattern length not a multiple of 8) runs the forward
    // pass on the vector ALUs in Double (shift_valu); its adjoint is not offered.
    bool valu = false;
    float *vmasks = nullptr;         // [k][ndet]
    int32_t *vshift = nullptr;       // [npos][2]
    double *vwr = nullptr, *vwc = nullptr;
    int nchunks;
    int any_lo;
    void *base = nullptr;            // split weights, one 32-mask tile [chunk][kstep][hi, lo][lane][8]
    double *inv = nullptr;           // [k]
    int64_t slot_stride;             // 16-byte entries per rolled table
    // shift-class preparation of the last origin map
    std::vector<double> o_row, o_col;
    bool prepared = false;
    bool prepared_rounded = false;   // the classes are of MaskShift.rounded (one roll, no blend)
    int ntiles = 0, nslots = 0;
    std::vector<int32_t> h_shift;    // [npos][2]
    std::vector<float> h_wts;        // [npos][4]
    std::vector<double> h_wr, h_wc;
    void *wroll = nullptr; size_t wroll_slots = 0;
    int32_t *perm = nullptr, *shift = nullptr, *slot_shifts = nullptr;
    cubo::ShiftTile *tiles = nullptr; size_t tiles_cap = 0;
    // split tiles (the classes' partial tiles, each over nsplit chunk ranges) after the whole ones
    int first_split = 0, split_groups = 0, nsplit = 1;
    double *part = nullptr; size_t part_cap = 0;
    // dY/dwr and dY/dwc of the last bilinear forward ([2][npos][k]), valid while the origin map (prepared) and the
    // cube (its generation) are those of that forward: the origin adjoint then needs no pass over the cube
    double *jac = nullptr; size_t jac_cap = 0;
    bool keep_jac = false;           // cubo_hip_shift_keep_jacobian
    bool jac_valid = false;
    uint64_t jac_generation = 0;
    int64_t jac_uses = 0;            // origin adjoints taken from kept rows since the last kernel-time query
    double *wr = nullptr, *wc = nullptr;
    float *wts = nullptr;
    // per-call buffers
    double *yhost = nullptr, *y = nullptr;   // Y [npos][k]: pinned host memory the kernels write, its device address
    double *dy = nullptr, *g = nullptr;      // g: [2][npos]
    double *dyhost = nullptr;                // pinned staging of dL/dY
    float *dy32 = nullptr;
    void *dyw = nullptr;             // dL/dY split into FP16 hi/lo for the matrix-core mask adjoint
    size_t dyw_bytes = 0;
    double *dy_inv = nullptr;
    double *rpartial = nullptr, *rdmask = nullptr;   // the matrix-core adjoint's [groups][32][ndet] and [32][ndet]
    int rgroups = 0;
    // batches of 16 positions of one column shift for shift_grad_t (built from the last origin map)
    bool batches_ready = false;
    int64_t nbatch = 0;
    int32_t *bslots = nullptr, *bic = nullptr, *border = nullptr;   // [nbatch][16][2], [nbatch], [nbatch][16]
    size_t bcap = 0;
    double *partial = nullptr, *dmask = nullptr;
    int ngroups = 0;
    hipEvent_t ev[6] = {};
    double ms[3] = {0, 0, 0};        // forward stack, origin adjoint, mask adjoint
    int64_t calls = 0;
};

namespace {

void free_plan(cubo_hip_shift_plan *p) {
    if (!p) return;
    DeviceGuard g(p->cube->device);
    (void)hipStreamSynchronize(p->cube->stream);
    void *bufs[] = {p->base, p->inv, p->wroll, p->perm, p->shift, p->slot_shifts, p->tiles, p->part, p->jac, p->wr, p->wc, p->wts,
                    p->dy, p->g, p->dy32, p->partial, p->dmask, p->vmasks, p->vshift, p->vwr, p->vwc,
                    p->dyw, p->dy_inv, p->rpartial, p->rdmask, p->bslots, p->bic, p->border};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (p->yhost) (void)hipHostFree(p->yhost);
    if (p->dyhost) (void)hipHostFree(p->dyhost);
    for (hipEvent_t e : p->ev)
        if (e) (void)hipEventDestroy(e);
    delete p;
}

// Y in mapped pinned host memory: the forward kernels write it over PCIe while they stream the cube.
hipError_t ensure_y(cubo_hip_shift_plan *p, size_t ny) {
    if (p->yhost) return hipSuccess;
    hipError_t e = hipHostMalloc((void **)&p->yhost, ny * 8, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&p->y, p->yhost, 0);
    return e;
}

// One pass over dL/dY [npos][k] on host threads: copied into the pinned staging buffer, checked finite, and each
// column's largest magnitude. False if a value is not finite.
bool stage_dy(cubo_hip_shift_plan *p, const double *dy, int64_t npos, std::vector<double> &colmax) {
    const int k = p->k, T = (int)std::min<int64_t>(cubo::reader_threads(), std::max<int64_t>(1, npos / 4096));
    std::vector<std::vector<double>> mx(T, std::vector<double>(k, 0.0));
    std::vector<char> ok(T, 1);
    auto run = [&](int t) {
        const int64_t i0 = npos * t / T, i1 = npos * (t + 1) / T;
        double *m = mx[t].data();
        bool fin = true;
        for (int64_t i = i0; i < i1; ++i)
            for (int kk = 0; kk < k; ++kk) {
                const double v = dy[(size_t)i * k + kk];
                p->dyhost[(size_t)i * k + kk] = v;
                fin = fin && std::isfinite(v);
                m[kk] = std::max(m[kk], std::fabs(v));
            }
        ok[t] = fin;
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(run, t);
    run(0);
    for (auto &th : pool) th.join();
    colmax.assign(k, 0.0);
    for (int t = 0; t < T; ++t)
        for (int kk = 0; kk < k; ++kk) colmax[kk] = std::max(colmax[kk], mx[t][kk]);
    return std::all_of(ok.begin(), ok.end(), [](char v) { return v != 0; });
}

// The caller's copy of Y, on several host threads.
void copy_y(const cubo_hip_shift_plan *p, double *y, size_t ny) {
    cubo::parallel_pieces(ny * 8, cubo::reader_threads(), [&](size_t off, size_t len) {
        memcpy((char *)y + off, (const char *)p->yhost + off, len);
        return true;
    });
}

template <typename T>
hipError_t grow(T **p, size_t *cap, size_t n) {
    if (*p && *cap >= n) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) *cap = n;
    return e;
}

// Shift classes, tiles and rolled weight tables for an origin map (skipped when it is the previous one).
// rounded: MaskShift.rounded (ShiftedImaging.swift:66): the roll by wrap(round_half_even(d)), no blend; else
// MaskShift.bilinear (:58-64): floor and fractions.
int prepare(cubo_hip_shift_plan *p, const double *orow, const double *ocol, bool rounded = false) {
    cubo_hip_cube *c = p->cube;
    const int64_t npos = c->npos;
    const int qy = (int)c->shape[2], qx = (int)c->shape[3];
    if (p->prepared && p->prepared_rounded == rounded && std::equal(orow, orow + npos, p->o_row.begin()) &&
        std::equal(ocol, ocol + npos, p->o_col.begin()))
        return CUBO_HIP_OK;
    for (int64_t i = 0; i < npos; ++i)
        if (!std::isfinite(orow[i]) || !std::isfinite(ocol[i]))
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the origin map must be finite (position %lld)", (long long)i);
    p->prepared = false;
    p->jac_valid = false;
    const double mr = pairwise_sum(orow, npos) / (double)npos, mc = pairwise_sum(ocol, npos) / (double)npos;
    p->h_shift.resize(2 * npos);
    p->h_wts.resize(4 * npos);
    p->h_wr.resize(npos);
    p->h_wc.resize(npos);
    std::vector<int32_t> key(npos);
    for (int64_t i = 0; i < npos; ++i) {
        const double dr = orow[i] - mr, dc = ocol[i] - mc;
        const double fr = rounded ? std::nearbyint(dr) : std::floor(dr), fc = rounded ? std::nearbyint(dc) : std::floor(dc);
        const double wr = rounded ? 0.0 : dr - fr, wc = rounded ? 0.0 : dc - fc;
        const int ir = wrap_shift(fr, qy), ic = wrap_shift(fc, qx);
        p->h_shift[2 * i] = ir;
        p->h_shift[2 * i + 1] = ic;
        p->h_wr[i] = wr;
        p->h_wc[i] = wc;
        p->h_wts[4 * i] = (float)((1 - wr) * (1 - wc));
        p->h_wts[4 * i + 1] = (float)(wr * (1 - wc));
        p->h_wts[4 * i + 2] = (float)((1 - wr) * wc);
        p->h_wts[4 * i + 3] = (float)(wr * wc);
        key[i] = ir * qx + ic;
    }
    // counting sort by class
    std::vector<int64_t> count((size_t)qy * qx + 1, 0);
    for (int64_t i = 0; i < npos; ++i) ++count[key[i] + 1];
    for (size_t i = 1; i < count.size(); ++i) count[i] += count[i - 1];
    std::vector<int32_t> perm(npos);
    {
        std::vector<int64_t> at(count.begin(), count.end() - 1);
        for (int64_t i = 0; i < npos; ++i) perm[at[key[i]]++] = (int32_t)i;
    }
    constexpr int TP = 32 * cubo::kShiftWaves;
    std::vector<cubo::ShiftTile> tiles;
    std::map<int, int> slot_of;
    std::vector<int32_t> slot_shifts;
    auto slot = [&](int tr, int tc) {
        const int kk = tr * qx + tc;
        auto it = slot_of.find(kk);
        if (it != slot_of.end()) return it->second;
        const int s = (int)slot_shifts.size() / 2;
        slot_of[kk] = s;
        slot_shifts.push_back(tr);
        slot_shifts.push_back(tc);
        return s;
    };
    for (int kk = 0; kk < qy * qx; ++kk) {
        const int64_t b = count[kk], e = count[kk + 1];
        if (b == e) continue;
        const int ir = kk / qx, ic = kk % qx;
        const int s0 = slot(ir, ic);
        const int s1 = rounded ? s0 : slot((ir + 1) % qy, ic), s2 = rounded ? s0 : slot(ir, (ic + 1) % qx),
                  s3 = rounded ? s0 : slot((ir + 1) % qy, (ic + 1) % qx);
        for (int64_t t = b; t < e; t += TP) {
            cubo::ShiftTile tl{};
            tl.start = (int32_t)t;
            tl.count = (int32_t)std::min<int64_t>(TP, e - t);
            tl.slot[0] = s0; tl.slot[1] = s1; tl.slot[2] = s2; tl.slot[3] = s3;
            tl.c0 = 0;
            tl.c1 = p->nchunks;
            tl.part = -1;
            tiles.push_back(tl);
        }
    }
    DeviceGuard g(c->device);
    // One block per compute unit at a time: when the whole tiles fill at least one round, the classes' partial
    // tiles would add a round of their own, so each is split into nsplit chunk ranges (their partial outputs added
    // afterwards in split order) and they run after the whole tiles.
    const int cus = cubo::compute_units(c->device);
    {
        std::vector<cubo::ShiftTile> whole, partial;
        for (const auto &t : tiles) (t.count == TP ? whole : partial).push_back(t);
        p->first_split = p->split_groups = 0;
        p->nsplit = 1;
        if ((int64_t)whole.size() >= cus && !partial.empty() && p->nchunks >= 2) {
            const int ns = std::min(16, p->nchunks);
            tiles = whole;
            p->first_split = (int)whole.size();
            p->split_groups = (int)partial.size();
            p->nsplit = ns;
            for (size_t gi = 0; gi < partial.size(); ++gi)
                for (int sp = 0; sp < ns; ++sp) {
                    cubo::ShiftTile tl = partial[gi];
                    tl.c0 = (int32_t)((int64_t)p->nchunks * sp / ns);
                    tl.c1 = (int32_t)((int64_t)p->nchunks * (sp + 1) / ns);
                    tl.part = (int32_t)(gi * ns + sp);
                    tiles.push_back(tl);
                }
        }
    }
    const int nslots = (int)slot_shifts.size() / 2;
    hipError_t e = hipSuccess;
    size_t cap_tiles = p->tiles_cap;
    if (!p->perm) {
        e = hipMalloc((void **)&p->perm, npos * 4);
        if (e == hipSuccess) e = hipMalloc((void **)&p->shift, npos * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->wr, npos * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->wc, npos * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->wts, npos * 16);
        if (e == hipSuccess) e = hipMalloc((void **)&p->slot_shifts, (size_t)qy * qx * 8);
        if (e != hipSuccess) return hip_fail(e, "allocating the shift classes");
    }
    e = grow(&p->tiles, &cap_tiles, tiles.size());
    p->tiles_cap = cap_tiles;
    if (e == hipSuccess && p->split_groups > 0)
        e = grow(&p->part, &p->part_cap, (size_t)p->split_groups * p->nsplit * TP * 3 * std::max(p->k, 1));
    if (e == hipSuccess) e = grow((char **)&p->wroll, &p->wroll_slots, (size_t)nslots * p->slot_stride * 16);
    if (e != hipSuccess) return hip_fail(e, "allocating the rolled weights");
    hipStream_t s = c->stream;
    e = hipMemcpyAsync(p->perm, perm.data(), npos * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->shift, p->h_shift.data(), npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->wr, p->h_wr.data(), npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->wc, p->h_wc.data(), npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->wts, p->h_wts.data(), npos * 16, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->tiles, tiles.data(), tiles.size() * sizeof(cubo::ShiftTile), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->slot_shifts, slot_shifts.data(), slot_shifts.size() * 4,
                                            hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = cubo::launch_roll_weights(p->base, p->wroll, p->slot_shifts, nslots, p->nchunks, c->ndet, qy,
                                                       qx, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);      // the host vectors above may change on the next call
    if (e != hipSuccess) return hip_fail(e, "preparing the shift classes");
    p->ntiles = (int)tiles.size();
    p->nslots = nslots;
    p->batches_ready = false;
    p->o_row.assign(orow, orow + npos);
    p->o_col.assign(ocol, ocol + npos);
    p->prepared = true;
    p->prepared_rounded = rounded;
    return CUBO_HIP_OK;
}

// The positions in batches of 16 of one column shift for shift_grad_t: a counting sort by ic (stable in position
// order), each shift's last batch padded with slots that repeat its first position and carry no dL/dY (order -1).
int prepare_batches(cubo_hip_shift_plan *p) {
    if (p->batches_ready) return CUBO_HIP_OK;
    cubo_hip_cube *c = p->cube;
    const int64_t npos = c->npos;
    const int qy = (int)c->shape[2], qx = (int)c->shape[3];
    std::vector<int64_t> count((size_t)qx + 1, 0);
    for (int64_t i = 0; i < npos; ++i) ++count[p->h_shift[2 * i + 1] + 1];
    int64_t nbatch = 0;
    for (int ic = 0; ic < qx; ++ic) nbatch += (count[ic + 1] + 15) / 16;
    for (size_t i = 1; i < count.size(); ++i) count[i] += count[i - 1];
    std::vector<int32_t> sorted(npos);
    {
        std::vector<int64_t> at(count.begin(), count.end() - 1);
        for (int64_t i = 0; i < npos; ++i) sorted[at[p->h_shift[2 * i + 1]]++] = (int32_t)i;
    }
    std::vector<int32_t> slots((size_t)nbatch * 32), bic((size_t)nbatch), order((size_t)nbatch * 16);
    int64_t b = 0;
    for (int ic = 0; ic < qx; ++ic) {
        for (int64_t t = count[ic]; t < count[ic + 1]; t += 16, ++b) {
            bic[b] = ic;
            for (int sl = 0; sl < 16; ++sl) {
                const bool live = t + sl < count[ic + 1];
                const int32_t pos = sorted[live ? t + sl : t];
                slots[(size_t)b * 32 + 2 * sl] = (int32_t)((uint32_t)pos * (uint32_t)qy);   // the first row
                slots[(size_t)b * 32 + 2 * sl + 1] = p->h_shift[2 * (size_t)pos];
                order[(size_t)b * 16 + sl] = live ? pos : -1;
            }
        }
    }
    DeviceGuard g(c->device);
    hipError_t e = hipSuccess;
    if (p->bcap < (size_t)nbatch) {
        for (int32_t **q : {&p->bslots, &p->bic, &p->border}) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
        p->bcap = 0;
        e = hipMalloc((void **)&p->bslots, (size_t)nbatch * 32 * 4);
        if (e == hipSuccess) e = hipMalloc((void **)&p->bic, (size_t)nbatch * 4);
        if (e == hipSuccess) e = hipMalloc((void **)&p->border, (size_t)nbatch * 16 * 4);
        if (e != hipSuccess) return hip_fail(e, "allocating the adjoint's batches");
        p->bcap = (size_t)nbatch;
    }
    hipStream_t s = c->stream;
    e = hipMemcpyAsync(p->bslots, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->bic, bic.data(), bic.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->border, order.data(), order.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);     // the host vectors go out of scope
    if (e != hipSuccess) return hip_fail(e, "uploading the adjoint's batches");
    p->nbatch = nbatch;
    p->batches_ready = true;
    return CUBO_HIP_OK;
}

// Of the group counts that give at least one block per block slot (per_cu of them on each compute unit), the one whose
// blocks fill their last round best.
int group_count(int64_t bands, int64_t units, int cus, int per_cu) {
    cus *= per_cu;
    int ngroups = 1;
    double best = -1;
    for (int64_t ng = std::max<int64_t>(1, (cus + bands - 1) / bands); ng <= std::min<int64_t>(64, units); ++ng) {
        const int64_t blocks = bands * ng, rounds = (blocks + cus - 1) / cus;
        const double fill = (double)blocks / (double)(rounds * cus);
        if (fill > best + 0.02) { best = fill; ngroups = (int)ng; }
    }
    return ngroups;
}

cubo::ShiftLaunch launch_args(const cubo_hip_shift_plan *p) {
    const cubo_hip_cube *c = p->cube;
    cubo::ShiftLaunch a{};
    a.cube = (const uint16_t *)c->data;
    a.npos = c->npos;
    a.ndet = c->ndet;
    a.nchunks = p->nchunks;
    a.k = p->k;
    a.any_lo = p->any_lo;
    a.row_chunks = c->shape[3] % 64 == 0 ? (int32_t)(c->shape[3] / 64) : 0;
    a.wroll = p->wroll;
    a.slot_stride = p->slot_stride;
    a.inv_scale = p->inv;
    a.tiles = p->tiles;
    a.perm = p->perm;
    a.wr = p->wr;
    a.wc = p->wc;
    a.part = p->part;
    return a;
}

// The stack kernel over every tile, then the split tiles' partial outputs added.
hipError_t run_stack(const cubo_hip_shift_plan *p, const cubo::ShiftLaunch &a, int pass, bool rounded, hipStream_t s) {
    hipError_t e = cubo::launch_shift_stack(a, p->ntiles, pass, rounded, s);
    if (e == hipSuccess) e = cubo::launch_shift_split_sum(a, p->first_split, p->split_groups, p->nsplit, pass, s);
    return e;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms : 0.f;
}

}  // namespace

extern "C" {

int cubo_hip_shift_plan_create(cubo_hip_cube *c, const float *masks, int32_t k, cubo_hip_shift_plan **out) {
    if (!c || !masks || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (k < 1 || k > 32) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "%d masks (1..32)", k);
    if (int st = cubo::require_resident(c, "the shifted stack")) return st;
    if (c->shape[2] > 65535 || c->shape[3] > 65535)
        return fail(CUBO_HIP_ERR_UNSUPPORTED, "the shifted stack takes detectors up to 65535 pixels a side");
    for (int64_t i = 0; i < (int64_t)k * c->ndet; ++i)
        if (!std::isfinite(masks[i])) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "a mask needs finite weights");
    auto *p = new (std::nothrow) cubo_hip_shift_plan();
    if (!p) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    p->cube = c;
    p->k = k;
    p->kp = k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : k <= 24 ? 24 : 32;
    DeviceGuard g(c->device);
    if (c->dtype != CUBO_HIP_U16 || c->ndet % 8 != 0) {
        p->valu = true;
        hipError_t e = hipMalloc((void **)&p->vmasks, (size_t)k * c->ndet * 4);
        if (e == hipSuccess) e = hipMemcpy(p->vmasks, masks, (size_t)k * c->ndet * 4, hipMemcpyHostToDevice);
        for (hipEvent_t &ev : p->ev)
            if (e == hipSuccess) e = hipEventCreate(&ev);
        if (e != hipSuccess) { free_plan(p); return hip_fail(e, "preparing the shifted stack"); }
        *out = p;
        return CUBO_HIP_OK;
    }
    cubo::SplitWeights sw = cubo::split_weights(masks, k, c->ndet);
    p->nchunks = sw.nchunks;
    p->any_lo = sw.tile_lo[0];
    p->slot_stride = (int64_t)sw.nchunks * 4 * 2 * 64;
    hipError_t e = hipMalloc(&p->base, sw.wb.size() * sizeof(_Float16));
    if (e == hipSuccess) e = hipMemcpy(p->base, sw.wb.data(), sw.wb.size() * sizeof(_Float16), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&p->inv, k * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(p->inv, sw.inv.data(), k * sizeof(double), hipMemcpyHostToDevice);
    for (hipEvent_t &ev : p->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e != hipSuccess) { free_plan(p); return hip_fail(e, "preparing the shifted stack"); }
    *out = p;
    return CUBO_HIP_OK;
}

namespace {

// The vector-ALU forward pass (any cube): the shifts of MaskShift.rounded / .bilinear per position, then shift_valu.
int forward_valu(cubo_hip_shift_plan *p, const double *orow, const double *ocol, double *y, bool rounded) {
    cubo_hip_cube *c = p->cube;
    const int64_t npos = c->npos;
    const int qy = (int)c->shape[2], qx = (int)c->shape[3];
    for (int64_t i = 0; i < npos; ++i)
        if (!std::isfinite(orow[i]) || !std::isfinite(ocol[i]))
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the origin map must be finite (position %lld)", (long long)i);
    const double mr = pairwise_sum(orow, npos) / (double)npos, mc = pairwise_sum(ocol, npos) / (double)npos;
    std::vector<int32_t> sh(2 * npos);
    std::vector<double> wr(npos), wc(npos);
    for (int64_t i = 0; i < npos; ++i) {
        const double dr = orow[i] - mr, dc = ocol[i] - mc;
        const double fr = rounded ? std::nearbyint(dr) : std::floor(dr), fc = rounded ? std::nearbyint(dc) : std::floor(dc);
        wr[i] = rounded ? 0.0 : dr - fr;
        wc[i] = rounded ? 0.0 : dc - fc;
        sh[2 * i] = wrap_shift(fr, qy);
        sh[2 * i + 1] = wrap_shift(fc, qx);
    }
    DeviceGuard g(c->device);
    hipStream_t s = c->stream;
    const size_t ny = (size_t)npos * p->k;
    hipError_t e = hipSuccess;
    if (!p->vshift) e = hipMalloc((void **)&p->vshift, (size_t)npos * 8);
    if (e == hipSuccess && !p->vwr) e = hipMalloc((void **)&p->vwr, (size_t)npos * 8);
    if (e == hipSuccess && !p->vwc) e = hipMalloc((void **)&p->vwc, (size_t)npos * 8);
    if (e == hipSuccess) e = ensure_y(p, ny);
    if (e == hipSuccess) e = hipMemcpyAsync(p->vshift, sh.data(), (size_t)npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->vwr, wr.data(), (size_t)npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->vwc, wc.data(), (size_t)npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(p->ev[0], s);
    if (e == hipSuccess) {
        cubo::ShiftValu a{};
        a.cube = c->data;
        a.u16 = c->dtype == CUBO_HIP_U16;
        a.npos = npos;
        a.qy = qy;
        a.qx = qx;
        a.k = p->k;
        a.masks = p->vmasks;
        a.shift = p->vshift;
        a.wr = p->vwr;
        a.wc = p->vwc;
        a.y = p->y;
        e = cubo::launch_shift_valu(a, rounded, s);
    }
    if (e == hipSuccess) e = hipEventRecord(p->ev[1], s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);    // the host vectors go out of scope
    else (void)hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "the shifted stack");
    copy_y(p, y, ny);
    p->ms[0] += elapsed(p->ev[0], p->ev[1]);
    return CUBO_HIP_OK;
}

int forward(cubo_hip_shift_plan *p, const double *origin_row, const double *origin_col, double *y, bool rounded) {
    if (!p || !origin_row || !origin_col || !y) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (p->valu) return forward_valu(p, origin_row, origin_col, y, rounded);
    if (int st = prepare(p, origin_row, origin_col, rounded)) return st;
    cubo_hip_cube *c = p->cube;
    DeviceGuard g(c->device);
    const size_t ny = (size_t)c->npos * p->k;
    CUBO_HIP_TRY(ensure_y(p, ny), "allocating the stack output");
    // When asked to, the bilinear forward also keeps the Jacobian rows for the origin adjoint of the same map.
    const int pass = p->keep_jac && !rounded ? cubo::kShiftYJac : cubo::kShiftY;
    p->jac_valid = false;
    if (pass == cubo::kShiftYJac) CUBO_HIP_TRY(grow(&p->jac, &p->jac_cap, 2 * ny), "allocating the Jacobian rows");
    cubo::ShiftLaunch a = launch_args(p);
    a.y = p->y;
    a.jac = p->jac;
    CUBO_HIP_TRY(hipEventRecord(p->ev[0], c->stream), "hipEventRecord");
    CUBO_HIP_TRY(run_stack(p, a, pass, rounded, c->stream), "launching the shifted stack");
    CUBO_HIP_TRY(hipEventRecord(p->ev[1], c->stream), "hipEventRecord");
    CUBO_HIP_TRY(hipStreamSynchronize(c->stream), "waiting for the shifted stack");
    copy_y(p, y, ny);
    p->ms[0] += elapsed(p->ev[0], p->ev[1]);
    if (pass == cubo::kShiftYJac) {
        p->jac_valid = true;
        p->jac_generation = c->generation;
    }
    return CUBO_HIP_OK;
}

}  // namespace

int cubo_hip_shift_forward(cubo_hip_shift_plan *p, const double *origin_row, const double *origin_col, double *y) {
    return forward(p, origin_row, origin_col, y, false);
}

int cubo_hip_shift_forward_rounded(cubo_hip_shift_plan *p, const double *origin_row, const double *origin_col,
                                   double *y) {
    return forward(p, origin_row, origin_col, y, true);
}

int cubo_hip_shift_backward(cubo_hip_shift_plan *p, const double *origin_row, const double *origin_col, const double *dy,
                            double *dmasks, double *d_origin_row, double *d_origin_col) {
    if (!p || !origin_row || !origin_col || !dy) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if ((d_origin_row == nullptr) != (d_origin_col == nullptr))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "ask for both origin gradients or neither");
    if (p->valu)
        return fail(CUBO_HIP_ERR_UNSUPPORTED, "the adjoint of the shifted stack needs a uint16 cube whose pattern "
                    "length is a multiple of 8");
    cubo_hip_cube *c = p->cube;
    const int64_t npos = c->npos, ndet = c->ndet;
    const size_t ny = (size_t)npos * p->k;
    DeviceGuard g(c->device);
    if (!p->dyhost) CUBO_HIP_TRY(hipHostMalloc((void **)&p->dyhost, ny * 8, hipHostMallocDefault), "allocating dL/dY");
    std::vector<double> colmax;
    if (!stage_dy(p, dy, npos, colmax)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "dL/dY must be finite");
    if (int st = prepare(p, origin_row, origin_col)) return st;
    hipStream_t s = c->stream;
    if (!p->dy) CUBO_HIP_TRY(hipMalloc((void **)&p->dy, ny * 8), "allocating dL/dY");
    CUBO_HIP_TRY(hipMemcpyAsync(p->dy, p->dyhost, ny * 8, hipMemcpyHostToDevice, s), "uploading dL/dY");
    if (d_origin_row) {
        if (!p->g) CUBO_HIP_TRY(hipMalloc((void **)&p->g, (size_t)npos * 16), "allocating the origin gradient");
        CUBO_HIP_TRY(hipEventRecord(p->ev[2], s), "hipEventRecord");
        if (p->jac_valid && p->jac_generation == c->generation) {
            // the forward on this origin map kept dY/dwr and dY/dwc: one dot product per position
            CUBO_HIP_TRY(cubo::launch_shift_origin_from_jac(p->jac, p->dy, npos, p->k, p->g, p->g + npos, s),
                         "launching the origin adjoint");
            ++p->jac_uses;
        } else {
            cubo::ShiftLaunch a = launch_args(p);
            a.dy = p->dy;
            a.gr = p->g;
            a.gc = p->g + npos;
            CUBO_HIP_TRY(run_stack(p, a, cubo::kShiftOrigin, false, s), "launching the origin adjoint");
        }
        CUBO_HIP_TRY(hipEventRecord(p->ev[3], s), "hipEventRecord");
    }
    if (dmasks && cubo::shift_grad_rows_supported((int)c->shape[3])) {
        // dL/dY times each corner weight, split for the matrix cores on the device (split_dy): each column scaled
        // by a power of two so its largest value lies in [1024, 2048), then FP16 hi + lo in the B-operand order
        const int64_t nks = (npos + 15) / 16;
        std::vector<double> inv(32, 0.0), scale(32, 0.0);
        for (int kk = 0; kk < p->k; ++kk)
            if (colmax[kk] > 0) {
                const int ex = (int)std::floor(std::log2(colmax[kk]));
                scale[kk] = std::ldexp(1.0, 10 - ex);
                inv[kk] = std::ldexp(1.0, ex - 10);
            }
        const int qy = (int)c->shape[2], qx = (int)c->shape[3];
        const int cus = cubo::compute_units(c->device);
        // batches of one column shift, staged transposed; the kernel indexes the cube's rows in 32 bits
        const bool tposed = cubo::shift_grad_t_supported(qy, qx) && (uint64_t)npos * (uint64_t)qy <= 0xffffffffull;
        if (tposed)
            if (int st = prepare_batches(p)) return st;
        const int64_t units = tposed ? p->nbatch : nks;
        const int band = tposed ? cubo::shift_grad_t_band(qx) : cubo::shift_grad_rows_band(qx);
        const int ngroups = group_count((qy + band - 1) / band, units, cus, tposed ? cubo::shift_grad_t_blocks_per_cu(qx) : 1);
        const size_t dyw_bytes = (size_t)units * 4 * 2 * 64 * 8 * sizeof(_Float16);
        if (p->dyw_bytes < dyw_bytes) {
            if (p->dyw) (void)hipFree(p->dyw);
            p->dyw = nullptr;
            p->dyw_bytes = 0;
            CUBO_HIP_TRY(hipMalloc(&p->dyw, dyw_bytes), "allocating dL/dY");
            p->dyw_bytes = dyw_bytes;
        }
        if (!p->dy_inv) CUBO_HIP_TRY(hipMalloc((void **)&p->dy_inv, 64 * 8), "allocating dL/dY scales");
        if (ngroups > p->rgroups) {
            if (p->rpartial) (void)hipFree(p->rpartial);
            p->rpartial = nullptr;
            p->rgroups = 0;
            CUBO_HIP_TRY(hipMalloc((void **)&p->rpartial, (size_t)ngroups * 32 * ndet * 8), "allocating partials");
            p->rgroups = ngroups;
        }
        if (!p->rdmask) CUBO_HIP_TRY(hipMalloc((void **)&p->rdmask, (size_t)32 * ndet * 8), "allocating dL/dmasks");
        inv.insert(inv.end(), scale.begin(), scale.end());          // dy_inv: [32] 1 / scale, then [32] scale
        CUBO_HIP_TRY(hipMemcpyAsync(p->dy_inv, inv.data(), 64 * 8, hipMemcpyHostToDevice, s), "uploading dL/dY scales");
        CUBO_HIP_TRY(cubo::launch_split_dy(p->dy, p->wr, p->wc, npos, p->k, p->dy_inv + 32, tposed ? p->border : nullptr,
                                           units * 16, p->dyw, s),
                     "splitting dL/dY");
        CUBO_HIP_TRY(hipEventRecord(p->ev[4], s), "hipEventRecord");
        if (tposed) {
            cubo::ShiftGradT gt{};
            gt.cube = (const uint16_t *)c->data;
            gt.qy = qy;
            gt.qx = qx;
            gt.nbatch = p->nbatch;
            gt.slots = p->bslots;
            gt.bic = p->bic;
            gt.dyw = p->dyw;
            gt.dy_inv = p->dy_inv;
            gt.partial = p->rpartial;
            gt.ngroups = ngroups;
            gt.bpg = (p->nbatch + ngroups - 1) / ngroups;
            CUBO_HIP_TRY(cubo::launch_shift_grad_t(gt, s), "launching the mask adjoint");
        } else {
            cubo::ShiftGradRows gw{};
            gw.cube = (const uint16_t *)c->data;
            gw.npos = npos;
            gw.qy = qy;
            gw.qx = qx;
            gw.shift = p->shift;
            gw.dyw = p->dyw;
            gw.dy_inv = p->dy_inv;
            gw.partial = p->rpartial;
            gw.ngroups = ngroups;
            gw.ppg = ((nks + ngroups - 1) / ngroups) * 16;
            CUBO_HIP_TRY(cubo::launch_shift_grad_rows(gw, s), "launching the mask adjoint");
        }
        CUBO_HIP_TRY(cubo::launch_sum_groups(p->rpartial, ngroups, (int64_t)32 * ndet, p->rdmask, s),
                     "launching the mask adjoint sum");
        CUBO_HIP_TRY(hipEventRecord(p->ev[5], s), "hipEventRecord");
        CUBO_HIP_TRY(hipMemcpyAsync(dmasks, p->rdmask, (size_t)p->k * ndet * 8, hipMemcpyDeviceToHost, s),
                     "reading dL/dmasks");
        CUBO_HIP_TRY(hipStreamSynchronize(s), "waiting for the mask adjoint");   // the host vectors go out of scope
    } else if (dmasks) {
        // dL/dY in FP32, padded to the kernel's mask count
        std::vector<float> d32((size_t)npos * p->kp, 0.f);
        for (int64_t i = 0; i < npos; ++i)
            for (int kk = 0; kk < p->k; ++kk) d32[(size_t)i * p->kp + kk] = (float)dy[(size_t)i * p->k + kk];
        const int cus = cubo::compute_units(c->device);
        const int64_t pixel_blocks = (ndet + 255) / 256;
        const int ngroups = (int)std::max<int64_t>(1, std::min<int64_t>({(4 * cus + pixel_blocks - 1) / pixel_blocks,
                                                                         256, npos}));
        if (!p->dy32) CUBO_HIP_TRY(hipMalloc((void **)&p->dy32, d32.size() * 4), "allocating dL/dY");
        if (ngroups > p->ngroups) {
            if (p->partial) (void)hipFree(p->partial);
            p->partial = nullptr;
            CUBO_HIP_TRY(hipMalloc((void **)&p->partial, (size_t)ngroups * p->kp * ndet * 8), "allocating partials");
            p->ngroups = ngroups;
        }
        if (!p->dmask) CUBO_HIP_TRY(hipMalloc((void **)&p->dmask, (size_t)p->kp * ndet * 8), "allocating dL/dmasks");
        CUBO_HIP_TRY(hipMemcpyAsync(p->dy32, d32.data(), d32.size() * 4, hipMemcpyHostToDevice, s), "uploading dL/dY");
        cubo::ShiftGradW gw{};
        gw.cube = (const uint16_t *)c->data;
        gw.npos = npos;
        gw.qy = (int32_t)c->shape[2];
        gw.qx = (int32_t)c->shape[3];
        gw.kp = p->kp;
        gw.shift = p->shift;
        gw.wts = p->wts;
        gw.dy = p->dy32;
        gw.partial = p->partial;
        gw.ngroups = ngroups;
        gw.ppg = (npos + ngroups - 1) / ngroups;
        CUBO_HIP_TRY(hipEventRecord(p->ev[4], s), "hipEventRecord");
        CUBO_HIP_TRY(cubo::launch_shift_grad_w(gw, s), "launching the mask adjoint");
        CUBO_HIP_TRY(cubo::launch_sum_groups(p->partial, ngroups, (int64_t)p->kp * ndet, p->dmask, s),
                     "launching the mask adjoint sum");
        CUBO_HIP_TRY(hipEventRecord(p->ev[5], s), "hipEventRecord");
        CUBO_HIP_TRY(hipMemcpyAsync(dmasks, p->dmask, (size_t)p->k * ndet * 8, hipMemcpyDeviceToHost, s),
                     "reading dL/dmasks");
    }
    std::vector<double> hg;
    if (d_origin_row) {
        hg.resize(2 * npos);
        CUBO_HIP_TRY(hipMemcpyAsync(hg.data(), p->g, (size_t)npos * 16, hipMemcpyDeviceToHost, s),
                     "reading the origin gradient");
    }
    CUBO_HIP_TRY(hipStreamSynchronize(s), "waiting for the adjoint");
    if (d_origin_row) {
        // d = origin - mean(origin): dL/d origin_q = g_q - mean(g)
        double mr = 0, mc = 0;
        for (int64_t i = 0; i < npos; ++i) { mr += hg[i]; mc += hg[npos + i]; }
        mr /= (double)npos;
        mc /= (double)npos;
        for (int64_t i = 0; i < npos; ++i) {
            d_origin_row[i] = hg[i] - mr;
            d_origin_col[i] = hg[npos + i] - mc;
        }
        p->ms[1] += elapsed(p->ev[2], p->ev[3]);
    }
    if (dmasks) p->ms[2] += elapsed(p->ev[4], p->ev[5]);
    ++p->calls;
    return CUBO_HIP_OK;
}

int cubo_hip_shift_kernel_time(cubo_hip_shift_plan *p, double *ms3, int64_t *calls) {
    if (!p || !ms3 || !calls) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int i = 0; i < 3; ++i) { ms3[i] = p->ms[i]; p->ms[i] = 0; }
    *calls = p->calls;
    p->calls = 0;
    return CUBO_HIP_OK;
}

int cubo_hip_shift_keep_jacobian(cubo_hip_shift_plan *p, int32_t keep) {
    if (!p) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL plan");
    if (keep != 0 && keep != 1) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "keep: 0 or 1, not %d", keep);
    p->keep_jac = keep == 1;
    if (!p->keep_jac) {
        p->jac_valid = false;
        if (p->jac) {
            DeviceGuard g(p->cube->device);
            (void)hipStreamSynchronize(p->cube->stream);
            (void)hipFree(p->jac);
            p->jac = nullptr;
            p->jac_cap = 0;
        }
    }
    return CUBO_HIP_OK;
}

int cubo_hip_shift_jacobian_uses(cubo_hip_shift_plan *p, int64_t *uses) {
    if (!p || !uses) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *uses = p->jac_uses;
    p->jac_uses = 0;
    return CUBO_HIP_OK;
}

int cubo_hip_shift_plan_destroy(cubo_hip_shift_plan *p) {
    free_plan(p);
    return CUBO_HIP_OK;
}

}  // extern "C"
