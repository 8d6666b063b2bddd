// cubo_pattern.hip — C ABI of the selected-area and centre-shifted patterns (include/cubo_hip.h, "patterns").
//
// VirtualDiffraction.pattern(of:statistic:scan:shift:) for the mean and the max over a mask or weights selection
// (Sources/CuboVirtual/SelectedDiffraction.swift:57-179, 225-271) and meanAndMax(of:shift:) with a rounded
// shift (ShiftedImaging.swift:125-167), on the resident cube. The session checks the selection as the reference
// does, uploads the selected positions (or the weights) and the per-position rolls once, and then takes scan-row
// tiles in order like the fused reduction: one pattern launch per tile over the tile's selected patterns, and a
// small kernel that folds the tile's group partials, in order, into running sums and maxima. The last tile writes
// the result into mapped pinned host memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/cubo_hip.h"
#include "cube.hpp"
#include "errors.hpp"
#include "host_math.hpp"
#include "dpc_kernels.hpp"
#include "pattern_kernels.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::hip_fail;

struct cubo_hip_pattern_session {
    cubo_hip_cube *cube;
    cubo::PatternKind kind;
    bool masked;                     // a list of selected positions
    std::vector<int64_t> prefix;     // masked: selected positions before position p, [npos + 1]
    double count;                    // the mean's divisor
    int cus;
    int64_t max_groups;
    int64_t next_row;
    bool started;
    // device
    int32_t *list = nullptr;
    double *w = nullptr;
    uint32_t *shift = nullptr;
    void *part_sum = nullptr, *part_max = nullptr;
    void *run_sum = nullptr, *run_max = nullptr;
    double *out_host = nullptr, *out_dev = nullptr;   // mapped pinned: the pattern
    // PatternShift.fourier: per-position shifts, fftfreq tables, twiddles and the batch's complex patterns
    bool fourier = false;
    bool fused = false;              // the one-kernel path (power-of-two detectors); else batches of FFTs
    int batch_max = 0;
    double *sr = nullptr, *sc = nullptr, *fr = nullptr, *fc = nullptr;
    double2 *tw_r = nullptr, *tw_c = nullptr, *fdata = nullptr, *ftmp = nullptr;
};

namespace {

void free_pattern(cubo_hip_pattern_session *p) {
    if (!p) return;
    DeviceGuard g(p->cube->device);
    (void)hipStreamSynchronize(p->cube->stream);
    void *bufs[] = {p->list, p->w, p->shift, p->part_sum, p->part_max, p->run_sum, p->run_max, p->sr, p->sc, p->fr,
                    p->fc, p->tw_r, p->tw_c, p->fdata, p->ftmp};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (p->out_host) (void)hipHostFree(p->out_host);
    delete p;
}

// Neumaier's sum in order (CenterOfMass.swift:83-92): the correction skipped where the sum is not finite.
double neumaier(const double *x, int64_t n) {
    double sum = 0.0, corr = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double t = sum + x[i];
        if (std::isfinite(t)) corr += std::fabs(sum) >= std::fabs(x[i]) ? (sum - t) + x[i] : (x[i] - t) + sum;
        sum = t;
    }
    return sum + corr;
}

// Groups of one launch: about 8 blocks of 4 waves per compute unit over all pixel blocks, at least 64 entries per
// group (16 per wave); the integer sums stay below 2^32 with at most 65536 entries per group.
void plan(const cubo_hip_pattern_session *p, int64_t count, int *ngroups, int64_t *ppg) {
    const int64_t bp = cubo::pattern_block_pixels(p->kind.vec);
    const int64_t blocks = (p->cube->ndet + bp - 1) / bp;
    int64_t g = std::max<int64_t>(1, ((int64_t)8 * p->cus + blocks - 1) / blocks);
    g = std::min<int64_t>(g, std::max<int64_t>(1, count / 64));
    g = std::min<int64_t>(g, p->max_groups);
    int64_t per = (count + g - 1) / g;
    if (p->kind.integer) per = std::min<int64_t>(per, 65536);
    *ppg = per;
    *ngroups = (int)((count + per - 1) / per);
}

int check_request(const cubo_hip_cube *c, const cubo_hip_pattern_request *q) {
    if (q->statistic != CUBO_HIP_PATTERN_MEAN && q->statistic != CUBO_HIP_PATTERN_MAX)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "statistic %d is neither the mean nor the max", q->statistic);
    if (q->take && q->weights)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "a selection is a mask or weights, not both");
    if ((q->origin_row == nullptr) != (q->origin_col == nullptr))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "a shift needs both origin maps");
    if (q->max_ties != CUBO_HIP_TIES_FIRST && q->max_ties != CUBO_HIP_TIES_LATER)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "max_ties %d is neither first nor later", q->max_ties);
    if (q->shift_kind != CUBO_HIP_SHIFT_ROUNDED && q->shift_kind != CUBO_HIP_SHIFT_FOURIER)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "shift_kind %d is neither rounded nor Fourier", q->shift_kind);
    if (q->origin_row && q->shift_kind == CUBO_HIP_SHIFT_FOURIER &&
        (c->shape[2] > cubo::kFftMax || c->shape[3] > cubo::kFftMax))
        return fail(CUBO_HIP_ERR_UNSUPPORTED, "the Fourier shift takes detectors up to %d pixels a side", cubo::kFftMax);
    if (c->dtype != CUBO_HIP_U16 && c->dtype != CUBO_HIP_F32)
        return fail(CUBO_HIP_ERR_UNSUPPORTED, "the cube's element type is neither uint16 nor float");
    if (c->npos > INT32_MAX) return fail(CUBO_HIP_ERR_UNSUPPORTED, "more than 2^31 - 1 scan positions");
    if (q->origin_row && (c->shape[2] > 65535 || c->shape[3] > 65535))
        return fail(CUBO_HIP_ERR_UNSUPPORTED, "a shifted pattern takes detectors up to 65535 pixels a side");
    const int64_t npos = c->npos;
    if (q->weights)   // SelectedDiffraction.swift:76
        for (int64_t i = 0; i < npos; ++i)
            if (!std::isfinite(q->weights[i])) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the weights must be finite");
    if (q->origin_row)   // VirtualImaging.require (ShiftedImaging.swift:99-106)
        for (int64_t i = 0; i < npos; ++i)
            if (!std::isfinite(q->origin_row[i]) || !std::isfinite(q->origin_col[i]))
                return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the origin map has non-finite values");
    return CUBO_HIP_OK;
}

}  // namespace

extern "C" {

int cubo_hip_pattern_path(cubo_hip_cube *c, const cubo_hip_pattern_request *q, int *path) {
    if (!c || !q || !path) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = check_request(c, q)) return st;
    *path = (c->dtype == CUBO_HIP_U16 && !q->weights) ? 1 : 2;
    return CUBO_HIP_OK;
}

int cubo_hip_pattern_begin(cubo_hip_cube *c, const cubo_hip_pattern_request *q, cubo_hip_pattern_session **out) {
    if (!c || !q || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (int st = check_request(c, q)) return st;
    const int64_t npos = c->npos, ndet = c->ndet;
    const int qy = (int)c->shape[2], qx = (int)c->shape[3];
    cubo_hip_pattern_session *p = new (std::nothrow) cubo_hip_pattern_session();
    if (!p) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    p->cube = c;
    const bool u16 = c->dtype == CUBO_HIP_U16, shifted = q->origin_row != nullptr;
    p->fourier = shifted && q->shift_kind == CUBO_HIP_SHIFT_FOURIER;
    p->kind.u16 = u16;
    p->kind.integer = u16 && !q->weights && !p->fourier;
    p->kind.stat = q->statistic == CUBO_HIP_PATTERN_MEAN ? cubo::kPatternMean : cubo::kPatternMax;
    p->kind.vec = cubo::pattern_vec(u16, ndet, qx, shifted);
    p->kind.shift = shifted;
    p->kind.first_wins = q->max_ties == CUBO_HIP_TIES_FIRST;
    p->masked = q->take != nullptr;
    p->cus = cubo::compute_units(c->device);

    // the selection, its size and the mean's divisor (SelectedDiffraction.swift:64-102, 268)
    std::vector<int32_t> list;
    int64_t entries = npos;
    p->count = (double)npos;
    if (p->masked) {
        p->prefix.assign(npos + 1, 0);
        for (int64_t i = 0; i < npos; ++i) {
            p->prefix[i + 1] = p->prefix[i] + (q->take[i] ? 1 : 0);
            if (q->take[i]) list.push_back((int32_t)i);
        }
        entries = (int64_t)list.size();
        p->count = (double)entries;
    } else if (q->weights) {
        p->count = neumaier(q->weights, npos);
    }
    if (entries == 0) { free_pattern(p); return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the selection holds no scan position"); }
    if (q->weights && q->statistic == CUBO_HIP_PATTERN_MEAN && p->count == 0) {
        free_pattern(p);
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the weights sum to zero");
    }
    // the rolls: the map's mean minus the origin, rounded half to even and wrapped (ShiftedImaging.swift:143-156);
    // the Fourier shift keeps the unrounded shift and NumPy's fftfreq (ShiftedImaging.swift:169-172)
    std::vector<uint32_t> rolls;
    std::vector<double> vsr, vsc, vfr, vfc;
    if (p->fourier) {
        const double mr = cubo::pairwise_sum(q->origin_row, npos) / (double)npos;
        const double mc = cubo::pairwise_sum(q->origin_col, npos) / (double)npos;
        vsr.resize(npos);
        vsc.resize(npos);
        for (int64_t i = 0; i < npos; ++i) { vsr[i] = mr - q->origin_row[i]; vsc[i] = mc - q->origin_col[i]; }
        auto fftfreq = [](int n) {
            std::vector<double> f(n);
            const double scale = 1.0 / (double)n;
            const int positive = (n - 1) / 2 + 1;
            for (int k = 0; k < n; ++k) f[k] = (double)(k < positive ? k : k - n) * scale;
            return f;
        };
        vfr = fftfreq(qy);
        vfc = fftfreq(qx);
    } else if (shifted) {
        const double mr = cubo::pairwise_sum(q->origin_row, npos) / (double)npos;
        const double mc = cubo::pairwise_sum(q->origin_col, npos) / (double)npos;
        rolls.resize(npos);
        for (int64_t i = 0; i < npos; ++i) {
            const int sr = cubo::wrap_shift(std::nearbyint(mr - q->origin_row[i]), qy);
            const int sc = cubo::wrap_shift(std::nearbyint(mc - q->origin_col[i]), qx);
            rolls[i] = (uint32_t)sr << 16 | (uint32_t)sc;
        }
    }

    DeviceGuard g(c->device);
    hipStream_t s = c->stream;
    {
        const int64_t bp = cubo::pattern_block_pixels(p->kind.vec);
        const int64_t blocks = (ndet + bp - 1) / bp;
        int64_t gmax = std::max<int64_t>(1, ((int64_t)8 * p->cus + blocks - 1) / blocks);
        if (p->kind.integer) gmax = std::max<int64_t>(gmax, (entries + 65535) / 65536);
        if (p->fourier && cubo::fourier_fused_supported(qy, qx)) {
            p->fused = true;
            gmax = std::max<int64_t>(gmax, p->cus);
        }
        p->max_groups = gmax;
    }
    const bool mean = p->kind.stat == cubo::kPatternMean;
    const size_t sum_b = p->kind.integer ? 4 : 16, max_b = p->kind.integer ? 4 : 8;
    const size_t run_sb = p->kind.integer ? 8 : 16, run_mb = p->kind.integer ? 4 : 8;
    hipError_t e = hipSuccess;
    if (p->masked) e = hipMalloc((void **)&p->list, list.size() * 4);
    if (e == hipSuccess && q->weights) e = hipMalloc((void **)&p->w, npos * 8);
    if (e == hipSuccess && shifted && !p->fourier) e = hipMalloc((void **)&p->shift, npos * 4);
    std::vector<double2> twr, twc;
    if (p->fused) {
        // one twiddle table for both axes: exp(-2 pi i m / max(qy, qx)), m < max
        twr = cubo::fft_twiddles(std::max(qy, qx));
        if (e == hipSuccess) e = hipMalloc((void **)&p->sr, npos * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->sc, npos * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->tw_r, twr.size() * 16);
    } else if (p->fourier) {
        // batches of at most 256 MiB of complex patterns (twice: the FFT's transpose buffer)
        p->batch_max = (int)std::max<int64_t>(1, std::min<int64_t>(4096, ((int64_t)256 << 20) / (ndet * 16)));
        twr = cubo::fft_twiddles(qy);
        twc = cubo::fft_twiddles(qx);
        if (e == hipSuccess) e = hipMalloc((void **)&p->sr, npos * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->sc, npos * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->fr, qy * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->fc, qx * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&p->tw_r, qy * 16);
        if (e == hipSuccess) e = hipMalloc((void **)&p->tw_c, qx * 16);
        if (e == hipSuccess) e = hipMalloc((void **)&p->fdata, (size_t)p->batch_max * ndet * 16);
        if (e == hipSuccess) e = hipMalloc((void **)&p->ftmp, (size_t)p->batch_max * ndet * 16);
    }
    if (e == hipSuccess && mean) e = hipMalloc(&p->part_sum, (size_t)p->max_groups * ndet * sum_b);
    if (e == hipSuccess && !mean) e = hipMalloc(&p->part_max, (size_t)p->max_groups * ndet * max_b);
    if (e == hipSuccess && mean) e = hipMalloc(&p->run_sum, ndet * run_sb);
    if (e == hipSuccess && !mean) e = hipMalloc(&p->run_max, ndet * run_mb);
    if (e == hipSuccess) e = hipHostMalloc((void **)&p->out_host, ndet * 8, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&p->out_dev, p->out_host, 0);
    if (e != hipSuccess) { free_pattern(p); return hip_fail(e, "allocating the pattern buffers"); }
    if (p->masked) e = hipMemcpyAsync(p->list, list.data(), list.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && q->weights) e = hipMemcpyAsync(p->w, q->weights, npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && shifted && !p->fourier)
        e = hipMemcpyAsync(p->shift, rolls.data(), npos * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && p->fourier) e = hipMemcpyAsync(p->sr, vsr.data(), npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && p->fourier) e = hipMemcpyAsync(p->sc, vsc.data(), npos * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && p->fused) e = hipMemcpyAsync(p->tw_r, twr.data(), twr.size() * 16, hipMemcpyHostToDevice, s);
    const bool batched = p->fourier && !p->fused;
    if (e == hipSuccess && batched) e = hipMemcpyAsync(p->fr, vfr.data(), qy * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && batched) e = hipMemcpyAsync(p->fc, vfc.data(), qx * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && batched) e = hipMemcpyAsync(p->tw_r, twr.data(), qy * 16, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && batched) e = hipMemcpyAsync(p->tw_c, twc.data(), qx * 16, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);   // the host vectors go out of scope
    if (e != hipSuccess) { free_pattern(p); return hip_fail(e, "uploading the selection"); }
    *out = p;
    return CUBO_HIP_OK;
}

int cubo_hip_pattern_rows(cubo_hip_pattern_session *p, int64_t row0, int64_t nrows) {
    if (!p) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL pattern");
    cubo_hip_cube *c = p->cube;
    if (int st = cubo::rows_range(c, row0, nrows)) return st;
    if (row0 != p->next_row)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "tiles must come in scan order: expected row %lld, got %lld",
                    (long long)p->next_row, (long long)row0);
    if (nrows == 0) return CUBO_HIP_OK;
    DeviceGuard g(c->device);
    const int64_t pos0 = row0 * c->shape[1], pos1 = (row0 + nrows) * c->shape[1];
    const int64_t first = p->masked ? p->prefix[pos0] : pos0;
    const int64_t count = (p->masked ? p->prefix[pos1] : pos1) - first;
    const bool finish = row0 + nrows == c->shape[0];
    hipStream_t s = c->stream;
    if (p->fourier) {
        const void *tile = nullptr;
        if (int st = cubo::rows_data(c, row0, nrows, &tile)) return st;
        const void *base = (const void *)((uintptr_t)tile - (uintptr_t)pos0 * (uintptr_t)(c->ndet * cubo::element_size(c->dtype)));
        CUBO_HIP_TRY(c->timer.start(s), "hipEventRecord");
        if (p->fused && (count > 0 || finish)) {
            int64_t ngroups = 0;
            if (count > 0) {
                ngroups = std::min<int64_t>(p->max_groups, count);
                const int64_t ppg = (count + ngroups - 1) / ngroups;
                ngroups = (count + ppg - 1) / ppg;
                cubo::FourierFused ff{};
                ff.cube = base;
                ff.u16 = c->dtype == CUBO_HIP_U16;
                ff.qy = (int)c->shape[2];
                ff.qx = (int)c->shape[3];
                ff.ndet = c->ndet;
                ff.first = first;
                ff.count = count;
                ff.list = p->list;
                ff.w = p->w;
                ff.sr = p->sr;
                ff.sc = p->sc;
                ff.tw = p->tw_r;
                ff.ngroups = (int)ngroups;
                ff.ppg = ppg;
                ff.part_sum = (double2 *)p->part_sum;
                ff.part_max = (double *)p->part_max;
                ff.stat = p->kind.stat;
                ff.first_wins = p->kind.first_wins;
                CUBO_HIP_TRY(cubo::launch_fourier_fused(ff, s), "launching the Fourier-shifted patterns");
            }
            CUBO_HIP_TRY(cubo::launch_pattern_accumulate(p->kind, p->part_sum, p->part_max, (int)ngroups, c->ndet,
                                                         p->run_sum, p->run_max, !p->started, finish, p->count,
                                                         p->out_dev, s),
                         "launching the pattern totals");
            p->started = true;
        }
        for (int64_t b0 = 0; !p->fused && (b0 < count || (count == 0 && b0 == 0 && finish)); b0 += p->batch_max) {
            const int nb = (int)std::min<int64_t>(p->batch_max, count - b0);
            if (nb > 0) {
                cubo::FourierBatch fb{};
                fb.cube = base;
                fb.u16 = c->dtype == CUBO_HIP_U16;
                fb.ndet = c->ndet;
                fb.qy = (int)c->shape[2];
                fb.qx = (int)c->shape[3];
                fb.first = first + b0;
                fb.batch = nb;
                fb.list = p->list;
                fb.w = p->w;
                fb.sr = p->sr;
                fb.sc = p->sc;
                fb.fr = p->fr;
                fb.fc = p->fc;
                fb.data = p->fdata;
                fb.part_sum = (double2 *)p->part_sum;
                fb.part_max = (double *)p->part_max;
                fb.stat = p->kind.stat;
                fb.first_wins = p->kind.first_wins;
                CUBO_HIP_TRY(cubo::launch_fourier_expand(fb, s), "expanding the patterns");
                CUBO_HIP_TRY(cubo::launch_fft2_batched(p->fdata, p->ftmp, nb, fb.qy, fb.qx, p->tw_r, p->tw_c, false, s),
                             "the forward FFT");
                CUBO_HIP_TRY(cubo::launch_fourier_phase(fb, s), "the phase ramp");
                CUBO_HIP_TRY(cubo::launch_fft2_batched(p->fdata, p->ftmp, nb, fb.qy, fb.qx, p->tw_r, p->tw_c, true, s),
                             "the inverse FFT");
                CUBO_HIP_TRY(cubo::launch_fourier_reduce(fb, s), "reducing the batch");
            }
            const bool last = b0 + p->batch_max >= count;
            CUBO_HIP_TRY(cubo::launch_pattern_accumulate(p->kind, p->part_sum, p->part_max, nb > 0 ? 1 : 0, c->ndet,
                                                         p->run_sum, p->run_max, !p->started, finish && last, p->count,
                                                         p->out_dev, s),
                         "launching the pattern totals");
            p->started = true;
            if (count == 0) break;
        }
        CUBO_HIP_TRY(c->timer.stop(s), "hipEventRecord");
        p->next_row = row0 + nrows;
        return CUBO_HIP_OK;
    }
    int ngroups = 0;
    // The kernel addresses position p at base + p * pattern bytes; for a streamed cube the base is the ring slot
    // moved back by the tile's first position (only positions of this tile are read).
    const void *tile = nullptr;
    if (int st = cubo::rows_data(c, row0, nrows, &tile)) return st;
    const void *cube_base = (const void *)((uintptr_t)tile - (uintptr_t)pos0 * (uintptr_t)(c->ndet * cubo::element_size(c->dtype)));
    if (count > 0) {
        int64_t ppg = 0;
        plan(p, count, &ngroups, &ppg);
        cubo::PatternLaunch a{};
        a.cube = cube_base;
        a.ndet = c->ndet;
        a.qy = (int)c->shape[2];
        a.qx = (int)c->shape[3];
        a.first = first;
        a.count = count;
        a.list = p->list;
        a.w = p->w;
        a.shift = p->shift;
        a.ngroups = ngroups;
        a.ppg = ppg;
        a.sum = p->part_sum;
        a.max = p->part_max;
        CUBO_HIP_TRY(c->timer.start(s), "hipEventRecord");
        CUBO_HIP_TRY(cubo::launch_pattern(a, p->kind, s), "launching the pattern kernel");
        CUBO_HIP_TRY(c->timer.stop(s), "hipEventRecord");
    }
    if (count > 0 || finish) {
        CUBO_HIP_TRY(cubo::launch_pattern_accumulate(p->kind, p->part_sum, p->part_max, ngroups, c->ndet, p->run_sum,
                                                     p->run_max, !p->started, finish, p->count, p->out_dev, s),
                     "launching the pattern totals");
        p->started = true;
    }
    p->next_row = row0 + nrows;
    return CUBO_HIP_OK;
}

int cubo_hip_pattern_results(cubo_hip_pattern_session *p, double *out) {
    if (!p || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    cubo_hip_cube *c = p->cube;
    if (p->next_row != c->shape[0])
        return fail(CUBO_HIP_ERR_INCOMPLETE, "the pattern covered %lld of %lld scan rows", (long long)p->next_row,
                    (long long)c->shape[0]);
    DeviceGuard g(c->device);
    CUBO_HIP_TRY(hipStreamSynchronize(c->stream), "waiting for the pattern");
    memcpy(out, p->out_host, c->ndet * sizeof(double));
    return CUBO_HIP_OK;
}

int cubo_hip_pattern_reset(cubo_hip_pattern_session *p) {
    if (!p) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL pattern");
    DeviceGuard g(p->cube->device);
    CUBO_HIP_TRY(hipStreamSynchronize(p->cube->stream), "waiting for the pattern");
    p->next_row = 0;
    p->started = false;
    return CUBO_HIP_OK;
}

int cubo_hip_pattern_cancel(cubo_hip_pattern_session *p) {
    free_pattern(p);
    return CUBO_HIP_OK;
}

int cubo_hip_pattern_release(cubo_hip_pattern_session *p) { return cubo_hip_pattern_cancel(p); }

int cubo_hip_pattern_end(cubo_hip_pattern_session *p, double *out) {
    if (!p) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL pattern");
    const int st = cubo_hip_pattern_results(p, out);
    free_pattern(p);
    return st;
}

int cubo_hip_pattern(cubo_hip_cube *c, const cubo_hip_pattern_request *q, double *out) {
    cubo_hip_pattern_session *p = nullptr;
    if (int st = cubo_hip_pattern_begin(c, q, &p)) return st;
    if (int st = cubo_hip_pattern_rows(p, 0, c->shape[0])) { free_pattern(p); return st; }
    return cubo_hip_pattern_end(p, out);
}

}  // extern "C"
