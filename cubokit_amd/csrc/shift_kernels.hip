// shift_kernels.hip — the shifted mask stack and its adjoint (the bilinear shift of
// Sources/CuboVirtual/ShiftedImaging.swift:43-73, 109-122, differentiated by hand).
//
// Forward and origin adjoint: one matrix-core pass over the cube. A tile holds positions of one integer
// shift class, so the four corner contractions C_c = X * roll(W, s_c) share their weights across the tile:
// the data are the exact uint16 split (x & 0xF800) + (x & 0x07FF), the rolled weights the split-FP16 hi/lo
// of the power-of-two scaled masks, the products exact in FP32, the corner sums flushed to FP64 every 256
// pixels and combined there:
//   forward   Y   = sum_c w_c C_c
//   backward  dY/dwr = (1 - wc)(C_1 - C_0) + wc (C_3 - C_2),  dY/dwc = (1 - wr)(C_2 - C_0) + wr (C_3 - C_1),
//             g_r(p) = sum_k dL/dY[p, k] dY/dwr[p, k]   (d wr / d origin_r = 1 inside a bilinear cell).
// The forward can also keep the Jacobian rows dY/dwr, dY/dwc (2 x positions x masks Doubles, not cube-sized):
// the origin adjoint of the same origin map is then a dot product per position, with no pass over the cube.
// Mask adjoint: a second pass with the positions as the reduction dimension: shift_grad_t (rows of 64-256 pixels,
// batches of one column shift staged transposed in LDS), shift_grad_rows (other multiples of 32) or shift_grad_w
// (any width, vector ALUs).
#include "device_sync.hpp"
#include "shift_kernels.hpp"

namespace cubo {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short us2 __attribute__((ext_vector_type(2)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
#define CUBO_GLOBAL __attribute__((address_space(1)))

constexpr int kBK = 64;            // pixels per chunk
constexpr int kFlush = 4;          // chunks per FP64 flush (256 pixels)
constexpr int kTileEntries = 4 * 2 * 64;   // 16-byte weight entries of one chunk of one 32-mask tile

__device__ __forceinline__ uint32_t cvt_pair(uint32_t w) {
    const us2 u = __builtin_bit_cast(us2, w);
    const h2 h = __builtin_convertvector(u, h2);
    return __builtin_bit_cast(uint32_t, h);
}

// NC corners: 4 for the bilinear shift, 1 for the rounded one (MaskShift.rounded: one roll, weight 1).
// D > 0 (the detector's rows are D whole chunks): the corners one row down, (ir + 1, ic) and (ir + 1, ic + 1),
// hold the weights of the corners above them D chunks earlier (a roll by one row moves every pixel qx = 64 D
// places), so only corners 0 and 2 are loaded and LDS keeps a ring of D + 2 chunks of them; the lower corners
// read the ring D chunks back. The ring's slots of chunks -D .. -1 hold the last D chunks (the roll wraps).
// PASS: kShiftY (Y), kShiftOrigin (gr, gc from dL/dY) or kShiftYJac (Y and the Jacobian rows: Y in FP64 as the
// forward's, the two rows across flushes in FP32, whose bar is the corner sums' scale, not the rows' own size).
template <int WAVES, int PASS, int NC, int D>
__global__ __launch_bounds__(64 * WAVES) void shift_stack_mfma(ShiftLaunch a) {
    constexpr bool BWD = PASS == kShiftOrigin, JAC = PASS == kShiftYJac;
    static_assert(NC == 4 || (NC == 1 && PASS == kShiftY && D == 0), "the adjoint needs the four corners");
    constexpr int kCorners = NC;
    constexpr bool RING = D > 0;
    constexpr int RS = RING ? D + 2 : 2;                     // chunk buffers in LDS
    constexpr int NCL = RING ? 2 : kCorners;                 // corners loaded per chunk
    constexpr int WCH = NCL * kTileEntries;                  // entries of one chunk buffer
    constexpr int TP = 32 * WAVES;
    constexpr int WPT = WCH / (64 * WAVES);
    static_assert(WCH % (64 * WAVES) == 0, "every thread carries the same number of weight entries");
    __shared__ half8 wlds[RS][WCH];
    // the buffer of chunk c (c >= -D)
    auto buf_of = [](int c) { return RING ? (c + RS) % RS : c & 1; };
    __shared__ double wts[2][TP];                              // the positions' bilinear fractions wr, wc

    const ShiftTile t = a.tiles[blockIdx.x];
    const int c0 = t.c0, c1 = t.c1;                            // the tile's chunk range
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = wave * 32 + (lane & 31);                     // this lane's position slot (A-operand row)
    const bool live = q < t.count;
    const int64_t p = a.perm[t.start + (live ? q : 0)];
    const int nchunks = a.nchunks;
    const CUBO_GLOBAL uint16_t *xrow = (const CUBO_GLOBAL uint16_t *)a.cube + p * a.ndet;
    const CUBO_GLOBAL half8 *wsrc = (const CUBO_GLOBAL half8 *)a.wroll;

    for (int i = tid; i < TP; i += 64 * WAVES) {
        const int64_t pp = a.perm[t.start + min(i, t.count - 1)];
        wts[0][i] = a.wr[pp];
        wts[1][i] = a.wc[pp];
    }

    // Unconditional loads (a pixel past the pattern reads the row start; its weights are zero), as in the
    // mask-stack kernel: a load under a lane mask would make every wait drain the prefetch.
    auto load_x = [&](u32x4 (&x)[4], int chunk) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int64_t px = (int64_t)chunk * kBK + 32 * (lane >> 5) + 8 * s;
            x[s] = *(const CUBO_GLOBAL u32x4 *)(xrow + (px < a.ndet ? px : 0));
        }
    };
    // Entry e of a chunk's block copy belongs to corner e / kTileEntries (a compile-time constant per j).
    // The forward keeps two chunks of weights in registers (fetched two chunks ahead); the adjoint, short of
    // registers, one.
    constexpr bool DEEP = PASS == kShiftY || (JAC && D > 0);   // registers allow it
    half8 wreg[DEEP ? 2 : 1][WPT];
    auto fetch_w = [&](int set, int chunk) {
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            const int e = tid + j * 64 * WAVES;
            const int ci = (j * 64 * WAVES) / kTileEntries + (64 * WAVES >= kTileEntries ? 0 : tid / kTileEntries);
            const int corner = RING ? 2 * ci : ci;
            wreg[set][j] = wsrc[(size_t)t.slot[corner] * a.slot_stride + (size_t)chunk * kTileEntries + (e % kTileEntries)];
        }
    };
    auto put_w = [&](int set, int buf) {
#pragma unroll
        for (int j = 0; j < WPT; ++j) wlds[buf][tid + j * 64 * WAVES] = wreg[set][j];
    };

    constexpr int NACC = BWD ? 2 : 1;
    float16v acc[kCorners];
    double accd[NACC][16];
    float jacc[JAC ? 2 : 1][16];                               // JAC: dY/dwr, dY/dwc
#pragma unroll
    for (int c = 0; c < kCorners; ++c)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
#pragma unroll
    for (int m = 0; m < NACC; ++m)
#pragma unroll
        for (int r = 0; r < 16; ++r) accd[m][r] = 0.0;
    // Register r of lane l holds row 8 (r / 4) + 4 (l / 32) + r % 4 of the wave's 32 positions, mask column l % 32.
    auto slot_of = [&](int r) { return wave * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); };
#pragma unroll
    for (int m = 0; m < (JAC ? 2 : 1); ++m)
#pragma unroll
        for (int r = 0; r < 16; ++r) jacc[m][r] = 0.f;

    const bool any_lo = a.any_lo != 0;
    // The corner sums of a flush combined in FP64 through d1 = c1 - c0, d2 = c2 - c0 and the second difference
    // e = c0 - c1 - c2 + c3:  dY/dwr = d1 + wc e,  dY/dwc = d2 + wr e,  Y = c0 + wr dY/dwr + wc d2
    // (the blend (1-wr)(1-wc) c0 + wr(1-wc) c1 + (1-wr)wc c2 + wr wc c3 in four multiply-adds).
    auto flush = [&]() {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = slot_of(r);
            if constexpr (NC == 1) {
                accd[0][r] += (double)acc[0][r];
            } else {
                const double wr = wts[0][row], wc = wts[1][row];
                const double c0 = acc[0][r], c1 = acc[1][r], c2 = acc[2][r], c3 = acc[3][r];
                const double d1 = c1 - c0, d2 = c2 - c0, e = (c3 - c2) - d1;
                const double jr = __builtin_fma(wc, e, d1);
                if constexpr (BWD) {
                    accd[0][r] += jr;
                    accd[1][r] += __builtin_fma(wr, e, d2);
                } else {
                    accd[0][r] += __builtin_fma(wc, d2, __builtin_fma(wr, jr, c0));
                    if constexpr (JAC) {
                        jacc[0][r] += (float)jr;
                        jacc[1][r] += (float)__builtin_fma(wr, e, d2);
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < kCorners; ++c)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
    };
    auto compute = [&](const u32x4 (&x)[4], int chunk) {
        uint32_t any = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s) any |= (x[s].x | x[s].y | x[s].z | x[s].w) & 0xf800f800u;
        const bool use_hi = __builtin_amdgcn_ballot_w64(any != 0) != 0;
        const half8 *wl = wlds[buf_of(chunk)];
        const half8 *wd = RING ? wlds[buf_of(chunk - D)] : wl;    // the lower corners' weights
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            half8 bh[kCorners], bl[kCorners];
#pragma unroll
            for (int c = 0; c < kCorners; ++c) {
                const half8 *wc = RING ? ((c & 1) ? wd : wl) + (c >> 1) * kTileEntries : wl + c * kTileEntries;
                bh[c] = wc[(s * 2 + 0) * 64 + lane];
                bl[c] = wc[(s * 2 + 1) * 64 + lane];
            }
            const uint32_t w[4] = {x[s].x, x[s].y, x[s].z, x[s].w};
            u32x4 lo, hi;
            lo.x = cvt_pair(w[0] & 0x07ff07ffu); lo.y = cvt_pair(w[1] & 0x07ff07ffu);
            lo.z = cvt_pair(w[2] & 0x07ff07ffu); lo.w = cvt_pair(w[3] & 0x07ff07ffu);
            const half8 alo = __builtin_bit_cast(half8, lo);
#pragma unroll
            for (int c = 0; c < kCorners; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bh[c], acc[c], 0, 0, 0);
            if (any_lo) {
#pragma unroll
                for (int c = 0; c < kCorners; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bl[c], acc[c], 0, 0, 0);
            }
            if (use_hi) {
                hi.x = cvt_pair(w[0] & 0xf800f800u); hi.y = cvt_pair(w[1] & 0xf800f800u);
                hi.z = cvt_pair(w[2] & 0xf800f800u); hi.w = cvt_pair(w[3] & 0xf800f800u);
                const half8 ahi = __builtin_bit_cast(half8, hi);
#pragma unroll
                for (int c = 0; c < kCorners; ++c)
                    acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bh[c], acc[c], 0, 0, 0);
                if (any_lo) {
#pragma unroll
                    for (int c = 0; c < kCorners; ++c)
                        acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bl[c], acc[c], 0, 0, 0);
                }
            }
        }
        if ((chunk + 1) % kFlush == 0 || chunk + 1 == c1) flush();
    };

    // As the mask stack: weights through LDS (double-buffered, one barrier per chunk), data two chunks ahead in
    // registers, weight loads issued before data loads. The weights of chunk c + 1 go to LDS at the end of step
    // c; in the forward they were fetched during step c - 1 (a whole step of latency), in the adjoint during
    // step c.
    u32x4 X0[4], X1[4], X2[4];
    if constexpr (RING) {
        // the D chunks before c0 (the last D chunks of the pattern for c0 = 0: the roll wraps) into the slots of
        // chunks c0 - D .. c0 - 1, before the first data loads
#pragma unroll
        for (int v = -D; v < 0; ++v) {
            fetch_w(0, (((c0 + v) % nchunks) + nchunks) % nchunks);
            put_w(0, buf_of(c0 + v));
        }
    }
    fetch_w(0, c0);
    __builtin_amdgcn_sched_barrier(0);
    load_x(X0, c0);
    load_x(X1, min(c0 + 1, c1 - 1));
    put_w(0, buf_of(c0));
    if constexpr (DEEP) fetch_w(1, min(c0 + 1, c1 - 1));
    __syncthreads();                                   // also publishes wts
    // cur: the register set holding chunk c + 1 (the forward alternates: set (c + 1 - c0) & 1)
    auto step = [&](const u32x4 (&xc)[4], u32x4 (&xn)[4], int c, int cur) {
        if constexpr (DEEP) fetch_w(cur ^ 1, min(c + 2, c1 - 1));
        else fetch_w(0, min(c + 1, c1 - 1));
        __builtin_amdgcn_sched_barrier(0);
        load_x(xn, min(c + 2, c1 - 1));
        compute(xc, c);
        put_w(DEEP ? cur : 0, buf_of(c + 1));
        lds_barrier();
    };
    int c = c0;
    for (; c + 6 <= c1; c += 6) {
        step(X0, X2, c, 1);
        step(X1, X0, c + 1, 0);
        step(X2, X1, c + 2, 1);
        step(X0, X2, c + 3, 0);
        step(X1, X0, c + 4, 1);
        step(X2, X1, c + 5, 0);
    }
    if (c < c1) step(X0, X2, c, 1);
    if (c + 1 < c1) step(X1, X0, c + 1, 0);
    if (c + 2 < c1) step(X2, X1, c + 2, 1);
    if (c + 3 < c1) step(X0, X2, c + 3, 0);
    if (c + 4 < c1) step(X1, X0, c + 4, 1);

    const int col = lane & 31;
    const double inv = col < a.k ? ((const CUBO_GLOBAL double *)a.inv_scale)[col] : 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = slot_of(r);
        const bool rl = row < t.count;
        const int64_t pr = a.perm[t.start + (rl ? row : 0)];
        if constexpr (JAC) {
            if (rl && col < a.k) {
                const double y = accd[0][r] * inv, jr = (double)jacc[0][r] * inv, jc = (double)jacc[1][r] * inv;
                if (t.part < 0) {
                    ((CUBO_GLOBAL double *)a.y)[pr * a.k + col] = y;
                    ((CUBO_GLOBAL double *)a.jac)[pr * a.k + col] = jr;
                    ((CUBO_GLOBAL double *)a.jac)[(a.npos + pr) * a.k + col] = jc;
                } else {                               // one chunk range of a split tile: [part][256][3][k]
                    CUBO_GLOBAL double *q = (CUBO_GLOBAL double *)a.part + ((int64_t)t.part * 256 + row) * 3 * a.k + col;
                    q[0] = y;
                    q[a.k] = jr;
                    q[2 * a.k] = jc;
                }
            }
        } else if constexpr (BWD) {
            double g0 = 0.0, g1 = 0.0;
            if (rl && col < a.k) {
                const double d = ((const CUBO_GLOBAL double *)a.dy)[pr * a.k + col] * inv;
                g0 = d * accd[0][r];
                g1 = d * accd[1][r];
            }
            // sum over the 32 mask columns (the lanes of this half-wave), in a fixed tree order
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) {
                g0 += __shfl_xor(g0, off, 64);
                g1 += __shfl_xor(g1, off, 64);
            }
            if (rl && col == 0) {
                if (t.part < 0) {
                    ((CUBO_GLOBAL double *)a.gr)[pr] = g0;
                    ((CUBO_GLOBAL double *)a.gc)[pr] = g1;
                } else {                               // one chunk range of a split tile
                    ((CUBO_GLOBAL double *)a.part)[((int64_t)t.part * 256 + row) * 2] = g0;
                    ((CUBO_GLOBAL double *)a.part)[((int64_t)t.part * 256 + row) * 2 + 1] = g1;
                }
            }
        } else {
            if (rl && col < a.k) {
                if (t.part < 0) ((CUBO_GLOBAL double *)a.y)[pr * a.k + col] = accd[0][r] * inv;
                else ((CUBO_GLOBAL double *)a.part)[((int64_t)t.part * 256 + row) * a.k + col] = accd[0][r] * inv;
            }
        }
    }
}

// The split tiles' partial outputs added in split order: for split tile group g (tiles first_split + g nsplit ..,
// one position range), Y[p][k] (kShiftY), gr, gc[p] (kShiftOrigin) or Y and the Jacobian rows (kShiftYJac) of each
// of its positions.
__global__ void shift_split_sum(ShiftLaunch a, int first_split, int nsplit, int pass) {
    const ShiftTile t = a.tiles[first_split + blockIdx.x * nsplit];
    const int q = threadIdx.x;
    if (q >= t.count) return;
    const int64_t pr = a.perm[t.start + q];
    const double *part = a.part;
    const int64_t base = (int64_t)blockIdx.x * nsplit * 256;
    if (pass == kShiftOrigin) {
        double g0 = 0.0, g1 = 0.0;
        for (int s = 0; s < nsplit; ++s) {
            g0 += part[(base + (int64_t)s * 256 + q) * 2];
            g1 += part[(base + (int64_t)s * 256 + q) * 2 + 1];
        }
        a.gr[pr] = g0;
        a.gc[pr] = g1;
    } else if (pass == kShiftYJac) {
        for (int m = 0; m < 3; ++m)
            for (int kk = 0; kk < a.k; ++kk) {
                double v = 0.0;
                for (int s = 0; s < nsplit; ++s) v += part[((base + (int64_t)s * 256 + q) * 3 + m) * a.k + kk];
                if (m == 0) a.y[pr * a.k + kk] = v;
                else a.jac[((m - 1) * a.npos + pr) * a.k + kk] = v;
            }
    } else {
        for (int kk = 0; kk < a.k; ++kk) {
            double y = 0.0;
            for (int s = 0; s < nsplit; ++s) y += part[(base + (int64_t)s * 256 + q) * a.k + kk];
            a.y[pr * a.k + kk] = y;
        }
    }
}

// The origin adjoint from kept Jacobian rows: gr[p] = sum_k dy[p][k] jac[0][p][k], gc[p] likewise with jac[1].
// A block takes 64 positions: the products are formed with coalesced reads of the block's 64 k consecutive
// elements of each array, then each position's k products are added in mask order.
constexpr int kJacPositions = 64;
__global__ __launch_bounds__(256) void shift_origin_from_jac(const double *jac, const double *dy, int64_t npos, int k,
                                                             double *gr, double *gc) {
    __shared__ double pr[kJacPositions * 32], pc[kJacPositions * 32];
    const int64_t p0 = (int64_t)blockIdx.x * kJacPositions;
    const int np = (int)min<int64_t>(kJacPositions, npos - p0);
    const int n = np * k;
    const double *d = dy + p0 * k, *jr = jac + p0 * k, *jc = jac + (npos + p0) * k;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const double v = d[e];
        pr[e] = v * jr[e];
        pc[e] = v * jc[e];
    }
    __syncthreads();
    const int q = threadIdx.x;
    if (q < np) {
        double g0 = 0.0, g1 = 0.0;
        for (int kk = 0; kk < k; ++kk) {
            g0 += pr[q * k + kk];
            g1 += pc[q * k + kk];
        }
        gr[p0 + q] = g0;
        gc[p0 + q] = g1;
    }
}

__global__ void roll_weights(const half8 *base, half8 *out, const int32_t *shifts, int nchunks, int64_t ndet, int qy,
                             int qx) {
    // one thread per (slot, 16-byte entry); the entry's 8 halves are gathered one by one
    const int64_t per_slot = (int64_t)nchunks * kTileEntries;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    if (e >= per_slot) return;
    const int tr = shifts[2 * slot], tc = shifts[2 * slot + 1];
    const int chunk = (int)(e / kTileEntries), rem = (int)(e % kTileEntries);
    const int st = rem / 128, hl = (rem / 64) & 1, l = rem & 63;
    const _Float16 *b16 = reinterpret_cast<const _Float16 *>(base);
    half8 v;
#pragma unroll
    for (int el = 0; el < 8; ++el) {
        const int64_t u = (int64_t)chunk * kBK + 32 * (l >> 5) + 8 * st + el;
        _Float16 h = (_Float16)0.f;
        if (u < ndet) {
            const int ui = (int)(u / qx), uj = (int)(u % qx);
            int si = ui - tr, sj = uj - tc;
            si += si < 0 ? qy : 0;
            sj += sj < 0 ? qx : 0;
            const int64_t su = (int64_t)si * qx + sj;
            const int sch = (int)(su / kBK), sh = (int)((su % kBK) / 32), sst = (int)((su % 32) / 8), sel = (int)(su % 8);
            const int sl = sh * 32 + (l & 31);
            h = b16[((((int64_t)sch * 4 + sst) * 2 + hl) * 64 + sl) * 8 + sel];
        }
        v[el] = h;
    }
    out[(int64_t)slot * per_slot + e] = v;
}

// One thread per pixel v = (a, b); the block's positions are uniform, so their shift, corner weights and
// dL/dY come in scalar registers.
template <int KP>
__global__ __launch_bounds__(256) void shift_grad_w(ShiftGradW g) {
    const int qy = g.qy, qx = g.qx;
    const int64_t ndet = (int64_t)qy * qx;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int group = blockIdx.y;
    const bool live = v < ndet;
    const int va = live ? (int)(v / qx) : 0, vb = live ? (int)(v % qx) : 0;
    const int64_t p0 = (int64_t)group * g.ppg;
    const int64_t p1 = min(g.npos, p0 + g.ppg);
    const int32_t *__restrict__ shift = g.shift;
    const float *__restrict__ wts = g.wts;
    const float *__restrict__ dy = g.dy;

    float acc[KP];
    double accd[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) { acc[k] = 0.f; accd[k] = 0.0; }
    // U positions per round: their 4 U corner loads are issued before any is used, so a wave waits for memory
    // once per U positions; FP32 sums go to FP64 every 16 positions.
    constexpr int U = 8;
    auto corners = [&](int64_t p, float (&x)[4]) {
        const int ir = shift[2 * p], ic = shift[2 * p + 1];
        int r0 = va + ir; r0 -= r0 >= qy ? qy : 0;
        int r1 = r0 + 1;  r1 -= r1 >= qy ? qy : 0;
        int c0 = vb + ic; c0 -= c0 >= qx ? qx : 0;
        int c1 = c0 + 1;  c1 -= c1 >= qx ? qx : 0;
        const uint16_t *xp = g.cube + p * ndet;
        x[0] = xp[r0 * qx + c0];
        x[1] = xp[r1 * qx + c0];
        x[2] = xp[r0 * qx + c1];
        x[3] = xp[r1 * qx + c1];
    };
    auto accumulate = [&](int64_t p, const float (&x)[4]) {
        const float xs = wts[4 * p] * x[0] + wts[4 * p + 1] * x[1] + wts[4 * p + 2] * x[2] + wts[4 * p + 3] * x[3];
#pragma unroll
        for (int k = 0; k < KP; ++k) acc[k] = __builtin_fmaf(xs, dy[p * KP + k], acc[k]);
    };
    auto flush = [&]() {
#pragma unroll
        for (int k = 0; k < KP; ++k) { accd[k] += (double)acc[k]; acc[k] = 0.f; }
    };
    int64_t p = p0;
    int n = 0;
    for (; p + U <= p1; p += U) {
        float x[U][4];
#pragma unroll
        for (int q = 0; q < U; ++q) corners(p + q, x[q]);
#pragma unroll
        for (int q = 0; q < U; ++q) accumulate(p + q, x[q]);
        n += U;
        if (n >= 16) { flush(); n = 0; }
    }
    for (; p < p1; ++p) {
        float x[4];
        corners(p, x);
        accumulate(p, x);
    }
    if (!live) return;
#pragma unroll
    for (int k = 0; k < KP; ++k) g.partial[((int64_t)group * KP + k) * ndet + v] = accd[k] + (double)acc[k];
}

// RB output rows per block (RB x W waves): the RB + 1 input rows each position's band reads are staged once.
template <int W, int RB>
__global__ __launch_bounds__(64 * W * RB) void shift_grad_rows(ShiftGradRows g) {
    constexpr int QX = 32 * W;                      // detector columns
    constexpr int QXP = QX + 8;                     // a staged row: the row, then its first 8 pixels again
    constexpr int VPR = QX / 8;                     // 16-byte vectors per row
    constexpr int SR = RB + 1;                      // staged rows per position
    constexpr int VPB = 16 * SR * VPR;              // row vectors per batch of 16 positions
    constexpr int T = 64 * W * RB;
    constexpr int BVB = 4 * 2 * 64;                 // 16-byte B-operand entries per batch (corner, hi/lo, lane)
    constexpr int PER = (VPB + T - 1) / T, BPER = (BVB + T - 1) / T;
    __shared__ __align__(16) uint16_t st[2][16][SR][QXP];
    __shared__ half8 sb[2][BVB];
    __shared__ int sic[2][16];
    const int tid = threadIdx.x, lane = tid & 63, hf = lane >> 5;
    const int wv = tid >> 6, rw = wv / W, wave = wv % W;   // output row a0 + rw, columns 32 wave ..
    const int qy = g.qy;
    const int64_t ndet = (int64_t)qy * QX;
    const int a0 = blockIdx.x * RB, arow = a0 + rw, group = blockIdx.y;
    const int64_t p_begin = (int64_t)group * g.ppg, p_end = min(g.npos, p_begin + g.ppg);
    const CUBO_GLOBAL uint16_t *cube = (const CUBO_GLOBAL uint16_t *)g.cube;
    const CUBO_GLOBAL half8 *dyw = (const CUBO_GLOBAL half8 *)g.dyw;

    // two batches in registers (set 0, 1), so a batch's loads have a whole step to arrive
    u32x4 reg[2][PER];
    half8 breg[2][BPER];
    int reg_ic[2] = {0, 0};
    // the batch's rows and B operands into register set `set`; positions past the cube repeat the last one (their
    // B operands are zero)
    auto load = [&](int set, int64_t p0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int v = tid + k * T;
            if (VPB % T == 0 || v < VPB) {
                const int j = v / (SR * VPR), rem = v % (SR * VPR), row = rem / VPR, col8 = rem % VPR;
                const int64_t p = min(p0 + j, g.npos - 1);
                const int ir = g.shift[2 * p];
                int r = a0 + ir + row;                  // < 3 qy: a0 < qy, ir < qy, row <= RB <= qy
                r -= r >= qy ? qy : 0;
                r -= r >= qy ? qy : 0;
                reg[set][k] = *(const CUBO_GLOBAL u32x4 *)(cube + p * ndet + (int64_t)r * QX + col8 * 8);
            }
        }
        const int64_t ks = p0 >> 4;
#pragma unroll
        for (int k = 0; k < BPER; ++k)
            if (BVB % T == 0 || tid + k * T < BVB) breg[set][k] = dyw[ks * BVB + tid + k * T];
        if (tid < 16) reg_ic[set] = g.shift[2 * min(p0 + tid, g.npos - 1) + 1];
    };
    auto store = [&](int set, int b) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int v = tid + k * T;
            if (VPB % T == 0 || v < VPB) {
                const int j = v / (SR * VPR), rem = v % (SR * VPR), row = rem / VPR, col8 = rem % VPR;
                *(u32x4 *)&st[b][j][row][col8 * 8] = reg[set][k];
                if (col8 == 0) *(u32x4 *)&st[b][j][row][QX] = reg[set][k];   // the wrap: pixel QX reads pixel 0
            }
        }
#pragma unroll
        for (int k = 0; k < BPER; ++k)
            if (BVB % T == 0 || tid + k * T < BVB) sb[b][tid + k * T] = breg[set][k];
        if (tid < 16) sic[b][tid] = reg_ic[set];
    };

    float16v acc;
    double accd[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) { acc[r] = 0.f; accd[r] = 0.0; }
    const int col = 32 * wave + (lane & 31);
    int nk = 0;
    auto compute = [&](int b) {
        // x[c][e / 2]: corner c's values for the lane's pixel and positions 8 hf + e, e + 1 (one per half); corners
        // in the order (0,0), (1,0), (0,1), (1,1) of the B operands. Each position's pixel c0 = col + ic (mod qx)
        // and its right neighbour come from one aligned dword pair of the staged row (padded for the wrap).
        uint32_t x[4][4];
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            uint32_t pr[2][2];                      // [position e, e + 1][row 0, 1]: (x[c0], x[c0 + 1])
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int j = 8 * hf + e + h;
                int c0 = col + sic[b][j];
                c0 -= c0 >= QX ? QX : 0;
                const uint32_t sh = (uint32_t)(c0 & 1) * 16u;
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const uint32_t *w = (const uint32_t *)&st[b][j][rw + r][0] + (c0 >> 1);
                    pr[h][r] = __builtin_amdgcn_alignbit(w[1], w[0], sh);
                }
            }
            // corner (r, 0): the low halves of positions e, e + 1; corner (r, 1): the high halves
            x[0][e >> 1] = __builtin_amdgcn_perm(pr[1][0], pr[0][0], 0x05040100u);
            x[1][e >> 1] = __builtin_amdgcn_perm(pr[1][1], pr[0][1], 0x05040100u);
            x[2][e >> 1] = __builtin_amdgcn_perm(pr[1][0], pr[0][0], 0x07060302u);
            x[3][e >> 1] = __builtin_amdgcn_perm(pr[1][1], pr[0][1], 0x07060302u);
        }
        uint32_t any = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) any |= x[c][q] & 0xf800f800u;
        const bool use_hi = __builtin_amdgcn_ballot_w64(any != 0) != 0;   // wave-uniform
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const half8 bh = sb[b][(c * 2 + 0) * 64 + lane], bl = sb[b][(c * 2 + 1) * 64 + lane];
            u32x4 lo;
            lo.x = cvt_pair(x[c][0] & 0x07ff07ffu); lo.y = cvt_pair(x[c][1] & 0x07ff07ffu);
            lo.z = cvt_pair(x[c][2] & 0x07ff07ffu); lo.w = cvt_pair(x[c][3] & 0x07ff07ffu);
            const half8 alo = __builtin_bit_cast(half8, lo);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bl, acc, 0, 0, 0);
            if (use_hi) {
                u32x4 hi;
                hi.x = cvt_pair(x[c][0] & 0xf800f800u); hi.y = cvt_pair(x[c][1] & 0xf800f800u);
                hi.z = cvt_pair(x[c][2] & 0xf800f800u); hi.w = cvt_pair(x[c][3] & 0xf800f800u);
                const half8 ahi = __builtin_bit_cast(half8, hi);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bl, acc, 0, 0, 0);
            }
        }
        if (++nk == 2) {
#pragma unroll
            for (int r = 0; r < 16; ++r) { accd[r] += (double)acc[r]; acc[r] = 0.f; }
            nk = 0;
        }
    };

    if (p_begin < p_end) {
        load(0, p_begin);
        store(0, 0);
        if (p_begin + 16 < p_end) load(1, p_begin + 16);
        __syncthreads();
        // step(p0, b, cur): batch p0 is in LDS buffer b, batch p0 + 16 in register set cur; the loads of batch
        // p0 + 32 go to the other set while batch p0 is multiplied
        auto step = [&](int64_t p0, int b, int cur) {
            if (p0 + 32 < p_end) load(cur ^ 1, p0 + 32);   // uniform
            compute(b);
            if (p0 + 16 < p_end) store(cur, b ^ 1);
            lds_barrier();
        };
        int64_t p0 = p_begin;
        for (; p0 + 16 < p_end; p0 += 32) {
            step(p0, 0, 1);
            step(p0 + 16, 1, 0);
        }
        if (p0 < p_end) step(p0, 0, 1);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) accd[r] += (double)acc[r];
    // register r of lane l: output pixel (arow, 32 wave + 8 (r / 4) + 4 (l / 32) + r % 4), mask l % 32
    if (arow >= qy) return;                          // the band's last row past the detector (odd qy)
    const int mask = lane & 31;
    const double scale = g.dy_inv[mask];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t pix = (int64_t)arow * QX + 32 * wave + 8 * (r >> 2) + 4 * hf + (r & 3);
        g.partial[((int64_t)group * 32 + mask) * ndet + pix] = accd[r] * scale;
    }
}

// The mask adjoint with the patterns staged transposed. A batch is 16 positions of one column shift ic (padding
// slots repeat a position of the batch and have zero dL/dY). Block (band, g) owns output rows a0 .. a0 + RB - 1
// and the batches of group g; wave w owns two 32-pixel tiles of output row a0 + w / WPR. Per batch the SR = RB + 1
// pattern rows the band reads, (a0 + ir_s + rho) mod qy for slot s, are staged in LDS as FP16, transposed to
// [rho][column][slot], in the exact split x = (x & 0xFC00) + (x & 0x03FF). Every A operand (8 positions of one
// pixel) is then one aligned 16-byte read; the column shift is applied by the reader:
//   corner (r, cc) of output pixel (a0 + rw, b) reads [rw + r][(b + cc + ic) mod qx][8 hf .. 8 hf + 7].
// Entries are 32 bytes with a 32-byte gap after every 8, and the two 16-byte halves of an entry swap places in
// every other group of 8: the staging writes (8 slot pairs x 8 consecutive 16-byte pieces per wave, at offsets fixed
// at compile time) and the operand reads then fall on distinct banks. The high parts are staged, in a second phase
// behind a second barrier, only for batches where some count is 1024 or more.
template <int WPR, int RB>
__global__ __launch_bounds__(64 * WPR * RB) void shift_grad_t(ShiftGradT g) {
    constexpr int QX = 64 * WPR;
    constexpr int SR = RB + 1;
    constexpr int T = 64 * WPR * RB;
    constexpr int RS = 8 * (QX + QX / 8);            // dwords per staged row (entries 0 .. QX - 1 and the gaps)
    constexpr int PLANE = SR * RS;
    constexpr int C8 = QX / 8;                       // 16-byte pieces per pattern row
    constexpr int TASKS = 8 * SR * C8;               // (slot pair, staged row, piece)
    constexpr int PER = (TASKS + T - 1) / T;
    constexpr int BVB = 4 * 2 * 64;                  // B-operand entries per batch (corner, hi/lo, lane)
    constexpr int BPER = (BVB + T - 1) / T;
    static_assert(T % 64 == 0, "whole waves");
    __shared__ __align__(16) uint32_t slo[2][PLANE];
    __shared__ __align__(16) uint32_t shi[PLANE];
    __shared__ half8 sb[2][BVB];
    constexpr int WAVES = T / 64;
    static_assert(WAVES <= 8, "the flags are read as two 16-byte words");
    __shared__ __align__(16) int32_t hiflag[2][8];   // per buffer and wave: whether the wave saw high parts
    // byte offset in a plane of half h of column v mod QX of staged row rho, v < 2 QX: a lane's A-operand addresses
    // are then one table read at its column plus the batch's column shift
    __shared__ int32_t atab[RB][2][2 * QX];
    const int tid = threadIdx.x, lane = tid & 63, hf = lane >> 5, n = lane & 31;
    const int wave = tid >> 6, rw = wave / WPR, wc = wave % WPR;
    const int qy = g.qy;
    const int64_t ndet = (int64_t)qy * QX;
    const int a0 = blockIdx.x * RB, arow = a0 + rw, group = blockIdx.y;
    const int64_t b_begin = (int64_t)group * g.bpg, b_end = min(g.nbatch, b_begin + g.bpg);
    const CUBO_GLOBAL uint16_t *cube = (const CUBO_GLOBAL uint16_t *)g.cube;
    const CUBO_GLOBAL half8 *dyw = (const CUBO_GLOBAL half8 *)g.dyw;
    // per slot pair: the first row of each position's pattern (position x qy) and its row shift
    const CUBO_GLOBAL u32x4 *slots = (const CUBO_GLOBAL u32x4 *)g.slots;

    // dword address of half h of column j of staged row rho
    auto ent = [](int rho, int j, int h) { return rho * RS + 8 * (j + (j >> 3)) + 4 * (h ^ ((j >> 3) & 1)); };
    for (int i = tid; i < RB * 4 * QX; i += T) {
        const int v = i % (2 * QX), h = (i / (2 * QX)) & 1, rho = i / (4 * QX);
        (&atab[0][0][0])[i] = 4 * ent(rho, v >= QX ? v - QX : v, h);
    }

    // The staging tasks of this thread: v = tid + k T; its slot pair q = v & 7 is the same for every k.
    const int q = tid & 7;
    u32x4 reg[2][PER][2];
    half8 breg[2][BPER];
    int reg_ic[2] = {0, 0};
    u32x4 rec = {0u, 0u, 0u, 0u};                    // slot records of the next batch to load
    // Loads are unconditional (a batch past the group re-reads the group's last one), so every wait counts the same
    // loads on every path and never waits for the prefetch issued after the loads it needs.
    auto fetch_rec = [&](int64_t bi) { rec = slots[min(bi, b_end - 1) * 8 + q]; };
    // (a0 + rho) mod qy of each task (rho <= RB <= qy); a row a0 + rho + ir is then one subtraction away
    uint32_t arho[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int rho = ((tid + k * T) >> 3) / C8;
        arho[k] = (uint32_t)(a0 + rho >= qy ? a0 + rho - qy : a0 + rho);
    }
    auto load = [&](int set, int64_t bi) {
        // the slot record and everything issued before it (the previous batch's loads) have arrived: one wait before
        // the branches of the partial task, so no path waits for the loads issued below
        __builtin_amdgcn_s_waitcnt(0x0f70);          // vmcnt(0)
        const uint32_t f0 = rec.x, ir0 = rec.y, f1 = rec.z, ir1 = rec.w;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int v = tid + k * T;
            if (TASKS % T == 0 || v < TASKS) {
                const int c = (v >> 3) % C8;
                // r < 2 qy; min(r, r - qy) as unsigned is r mod qy; the cube's row index fits 32 bits
                uint32_t r0 = arho[k] + ir0, r1 = arho[k] + ir1;
                r0 = min(r0, r0 - (uint32_t)qy) + f0;
                r1 = min(r1, r1 - (uint32_t)qy) + f1;
                reg[set][k][0] = *(const CUBO_GLOBAL u32x4 *)(cube + 8 * c + (uint64_t)r0 * QX);
                reg[set][k][1] = *(const CUBO_GLOBAL u32x4 *)(cube + 8 * c + (uint64_t)r1 * QX);
            }
        }
#pragma unroll
        for (int k = 0; k < BPER; ++k)
            if (BVB % T == 0 || tid + k * T < BVB) breg[set][k] = dyw[bi * BVB + tid + k * T];
        reg_ic[set] = g.bic[bi];
        fetch_rec(bi + 1);
    };
    // The FP16 parts of a pair of counts: low 10 bits exactly through the magic number 1024 (0x6400 | v is
    // 1024 + v), the high 6 bits (multiples of 1024 up to 64512) by conversion.
    uint32_t magic = 0x64006400u;
    asm volatile("" : "+v"(magic));                  // a register operand: the mask and the OR in one instruction
    auto part_lo = [&](uint32_t w) {
        const h2 m = __builtin_bit_cast(h2, (w & 0x03ff03ffu) | magic);
        const h2 r = m - (h2){(_Float16)1024.f, (_Float16)1024.f};
        return __builtin_bit_cast(uint32_t, r);
    };
    auto part_hi = [](uint32_t w) { return cvt_pair(w & 0xfc00fc00u); };
    // Task k of this thread into one plane of the staged rows from register set `set`; returns the bits of counts
    // of 1024 or more.
    auto stage_task = [&](int set, int k, uint32_t *dst, bool hi) {
        const int v = tid + k * T;
        const int w = v >> 3, c = w % C8, rho = w / C8;
        const u32x4 x0 = reg[set][k][0], x1 = reg[set][k][1];
        const uint32_t r0[4] = {x0.x, x0.y, x0.z, x0.w}, r1[4] = {x1.x, x1.y, x1.z, x1.w};
        uint32_t h0[4], h1[4], any = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            any |= (r0[i] | r1[i]) & 0xfc00fc00u;
            h0[i] = hi ? part_hi(r0[i]) : part_lo(r0[i]);
            h1[i] = hi ? part_hi(r1[i]) : part_lo(r1[i]);
        }
        // column 8 c + e, slots 2q (low half) and 2q + 1 (high half): offsets 8 e from the piece's first entry
        uint32_t *d = dst + ent(rho, 8 * c, q >> 2) + (q & 3);
#pragma unroll
        for (int e = 0; e < 8; ++e)
            d[8 * e] = __builtin_amdgcn_perm(h1[e >> 1], h0[e >> 1], (e & 1) ? 0x07060302u : 0x05040100u);
        return any;
    };
    constexpr int KFULL = TASKS / T;                 // tasks every thread has; the last one only the first threads
    // The whole plane: the full tasks, then the partial one (a branch uniform over waves, as T is whole waves and
    // the partial task count TASKS % T is too).
    static_assert((TASKS % T) % 64 == 0, "the partial task covers whole waves");
    auto stage_rest = [&](int set, uint32_t *dst, bool hi) {
        uint32_t any = 0;
        if (TASKS % T != 0 && tid + KFULL * T < TASKS) any |= stage_task(set, KFULL, dst, hi);
        return any;
    };
    // The batch's B operands and the wave's high-part flag (every lane writes the same value).
    auto store_misc = [&](int set, int b, uint32_t any) {
#pragma unroll
        for (int k = 0; k < BPER; ++k)
            if (BVB % T == 0 || tid + k * T < BVB) sb[b][tid + k * T] = breg[set][k];
        hiflag[b][wave] = __builtin_amdgcn_ballot_w64(any != 0) != 0;
    };
    // Whether the batch in buffer b has high parts (uniform; every staging of a buffer rewrites its flags).
    auto has_hi = [&](int b) {
        const u32x4 f0 = *(const u32x4 *)&hiflag[b][0], f1 = *(const u32x4 *)&hiflag[b][4];
        const uint32_t f[8] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w};
        uint32_t h = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) h |= f[w];
        return __builtin_amdgcn_readfirstlane(h) != 0;
    };

    float16v acc[2];
    double accd[2][16];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) { acc[t][r] = 0.f; accd[t][r] = 0.0; }
    int nk = 0;
    half8 bh[4], bl[4];
    int aaddr[2][2], anext[2][2];
    const int32_t *tab = &atab[rw][hf][64 * wc + n];
    // A-operand byte offsets of tile t and column corner cc in staged row rw (row rw + 1 is RS dwords further) for
    // column shift ic
    auto addresses = [&](int ic, int (&dst)[2][2]) {
        const int32_t *tb = tab + ic;
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) dst[t][cc] = tb[32 * t + cc];
    };
    // The B operands of the batch in buffer b.
    auto operands = [&](int b) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            bh[c] = sb[b][(c * 2 + 0) * 64 + lane];
            bl[c] = sb[b][(c * 2 + 1) * 64 + lane];
        }
    };
    auto mfma_tile = [&](const uint32_t *plane, int t) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const int c = r + 2 * cc;
                const half8 a = *(const half8 *)((const char *)plane + aaddr[t][cc] + 4 * r * RS);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, bh[c], acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, bl[c], acc[t], 0, 0, 0);
            }
    };
    auto flush = [&]() {
        if (++nk == 2) {
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int r = 0; r < 16; ++r) { accd[t][r] += (double)acc[t][r]; acc[t][r] = 0.f; }
            nk = 0;
        }
    };

    if (b_begin < b_end) {
        fetch_rec(b_begin);
        load(0, b_begin);
        load(1, min(b_begin + 1, b_end - 1));
        {
            uint32_t any = 0;
#pragma unroll
            for (int k = 0; k < KFULL; ++k) any |= stage_task(0, k, slo[0], false);
            any |= stage_rest(0, slo[0], false);
            store_misc(0, 0, any);
        }
        __syncthreads();
        bool hi_cur = has_hi(0);                     // the high-part flag of the batch being multiplied
        if (hi_cur) {
#pragma unroll
            for (int k = 0; k < KFULL; ++k) stage_task(0, k, shi, true);
            stage_rest(0, shi, true);
            __syncthreads();
        }
        addresses(reg_ic[0], aaddr);
        // step(bi, b, cur): batch bi is in LDS buffer b, batch bi + 1 (if STORE) in register set cur; the loads of
        // batch bi + 2 go to the other set while batch bi is multiplied. The staging of batch bi + 1 sits between the
        // matrix-core instructions of batch bi (one basic block), so the two overlap, and so does the table read of
        // batch bi + 1's A addresses.
        auto step = [&](int64_t bi, int b, int cur, bool store) {
            load(cur ^ 1, min(bi + 2, b_end - 1));
            operands(b);
            uint32_t any = 0;
            mfma_tile(slo[b], 0);
            if (store) {
#pragma unroll
                for (int k = 0; k < KFULL; ++k) any |= stage_task(cur, k, slo[b ^ 1], false);
            }
            mfma_tile(slo[b], 1);
            if (store) {
                any |= stage_rest(cur, slo[b ^ 1], false);
                store_misc(cur, b ^ 1, any);
                addresses(reg_ic[cur], anext);
            }
            if (hi_cur) {
                mfma_tile(shi, 0);
                mfma_tile(shi, 1);
            }
            flush();
            lds_barrier();
            if (store) {
                hi_cur = has_hi(b ^ 1);
                if (hi_cur) {                        // the high parts of batch bi + 1
#pragma unroll
                    for (int k = 0; k < KFULL; ++k) stage_task(cur, k, shi, true);
                    stage_rest(cur, shi, true);
                    lds_barrier();
                }
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int cc = 0; cc < 2; ++cc) aaddr[t][cc] = anext[t][cc];
            }
        };
        int64_t bi = b_begin;
        for (; bi + 2 < b_end; bi += 2) {
            step(bi, 0, 1, true);
            step(bi + 1, 1, 0, true);
        }
        if (bi + 1 < b_end) {
            step(bi, 0, 1, true);
            step(bi + 1, 1, 0, false);
        } else if (bi < b_end) {
            step(bi, 0, 1, false);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) accd[t][r] += (double)acc[t][r];
    // register r of lane l: output pixel (arow, 64 wc + 32 t + 8 (r / 4) + 4 (l / 32) + r % 4), mask l % 32
    if (arow >= qy) return;                          // the band's last rows past the detector
    const double scale = g.dy_inv[n];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t pix = (int64_t)arow * QX + 64 * wc + 32 * t + 8 * (r >> 2) + 4 * hf + (r & 3);
            g.partial[((int64_t)group * 32 + n) * ndet + pix] = accd[t][r] * scale;
        }
}

template <bool U16, bool ROUNDED>
__global__ __launch_bounds__(256) void shift_valu(ShiftValu a) {
#pragma clang fp contract(off)   // the blend, the products and TwoSum round as written
    const int lane = threadIdx.x & 63;
    const int64_t pos = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pos >= a.npos) return;                      // whole waves; no block barrier
    const int qy = a.qy, qx = a.qx;
    const int64_t ndet = (int64_t)qy * qx;
    const int ir = a.shift[2 * pos], ic = a.shift[2 * pos + 1];
    const double wr = a.wr[pos], wc = a.wc[pos];
    const double w00 = (1 - wr) * (1 - wc), w10 = wr * (1 - wc), w01 = (1 - wr) * wc, w11 = wr * wc;
    for (int k = 0; k < a.k; ++k) {
        const float *m = a.masks + (int64_t)k * ndet;
        double s = 0.0, c = 0.0;
        for (int64_t u = lane; u < ndet; u += 64) {
            const int i = (int)(u / qx), j = (int)(u % qx);
            // roll(mask, dr, dc)[i, j] = mask[(i - dr) mod qy, (j - dc) mod qx]
            int i0 = i - ir; i0 += i0 < 0 ? qy : 0;
            int j0 = j - ic; j0 += j0 < 0 ? qx : 0;
            double mv;
            if constexpr (ROUNDED) {
                mv = (double)m[(int64_t)i0 * qx + j0];
            } else {
                const int i1 = i0 == 0 ? qy - 1 : i0 - 1, j1 = j0 == 0 ? qx - 1 : j0 - 1;
                mv = (double)m[(int64_t)i0 * qx + j0] * w00 + (double)m[(int64_t)i1 * qx + j0] * w10 +
                     (double)m[(int64_t)i0 * qx + j1] * w01 + (double)m[(int64_t)i1 * qx + j1] * w11;
            }
            double x;
            if constexpr (U16) x = (double)((const uint16_t *)a.cube)[pos * ndet + u];
            else x = (double)((const float *)a.cube)[pos * ndet + u];
            const double v = x * mv;
            const double t = s + v, bb = t - s;
            c += (s - (t - bb)) + (v - bb);
            s = t;
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const double os = __shfl_xor(s, off, 64), oc = __shfl_xor(c, off, 64);
            const double t = s + os, bb = t - s;
            c += ((s - (t - bb)) + (os - bb)) + oc;
            s = t;
        }
        if (lane == 0) a.y[pos * a.k + k] = __builtin_isfinite(s) ? s + c : s;
    }
}

__global__ void sum_groups(const double *partial, int ngroups, int64_t n, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int gi = 0; gi < ngroups; ++gi) s += partial[(int64_t)gi * n + i];
    out[i] = s;
}

}  // namespace

hipError_t launch_shift_split_sum(const ShiftLaunch &a, int first_split, int ngroups, int nsplit, int pass,
                                  hipStream_t s) {
    if (ngroups <= 0) return hipSuccess;
    hipLaunchKernelGGL(shift_split_sum, dim3((unsigned)ngroups), dim3(256), 0, s, a, first_split, nsplit, pass);
    return hipGetLastError();
}

hipError_t launch_shift_stack(const ShiftLaunch &a, int ntiles, int pass, bool rounded, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    if (rounded && pass != kShiftY) return hipErrorInvalidValue;
    const dim3 grid(ntiles), block(64 * kShiftWaves);
    // the weight ring for detectors whose rows are 1, 2 or 4 whole chunks (64, 128 or 256 columns)
#define CUBO_SHIFT_FOUR(PASS)                                                                            \
    switch (a.row_chunks) {                                                                              \
        case 1: hipLaunchKernelGGL((shift_stack_mfma<kShiftWaves, PASS, 4, 1>), grid, block, 0, s, a); break; \
        case 2: hipLaunchKernelGGL((shift_stack_mfma<kShiftWaves, PASS, 4, 2>), grid, block, 0, s, a); break; \
        case 4: hipLaunchKernelGGL((shift_stack_mfma<kShiftWaves, PASS, 4, 4>), grid, block, 0, s, a); break; \
        default: hipLaunchKernelGGL((shift_stack_mfma<kShiftWaves, PASS, 4, 0>), grid, block, 0, s, a); break; \
    }
    if (rounded) {
        hipLaunchKernelGGL((shift_stack_mfma<kShiftWaves, kShiftY, 1, 0>), grid, block, 0, s, a);
    } else if (pass == kShiftOrigin) {
        CUBO_SHIFT_FOUR(kShiftOrigin)
    } else if (pass == kShiftYJac) {
        CUBO_SHIFT_FOUR(kShiftYJac)
    } else {
        CUBO_SHIFT_FOUR(kShiftY)
    }
#undef CUBO_SHIFT_FOUR
    return hipGetLastError();
}

hipError_t launch_shift_origin_from_jac(const double *jac, const double *dy, int64_t npos, int k, double *gr, double *gc,
                                        hipStream_t s) {
    if (npos <= 0) return hipSuccess;
    if (k < 1 || k > 32) return hipErrorInvalidValue;
    hipLaunchKernelGGL(shift_origin_from_jac, dim3((unsigned)((npos + kJacPositions - 1) / kJacPositions)), dim3(256), 0,
                       s, jac, dy, npos, k, gr, gc);
    return hipGetLastError();
}

bool shift_grad_rows_supported(int qx) { return qx % 32 == 0 && qx >= 32 && qx <= 256; }

int shift_grad_rows_band(int qx) { return qx <= 128 ? 3 : 1; }

namespace {
// One block per batch of 16 slots: their positions, fractions and dL/dY rows staged in LDS, then thread (corner c,
// lane l) forms the 8 entries e of its half8 (slot 8 (l / 32) + e, mask l % 32) and writes the hi and lo parts as one
// 16-byte store each.
constexpr int kSplitMaxK = 32;
__global__ __launch_bounds__(256) void split_dy(const double *dy, const double *wr, const double *wc, int64_t npos,
                                                int k, const double *scale, const int32_t *order, half8 *dyw) {
    __shared__ double sdy[16][kSplitMaxK];
    __shared__ double sr[16], sq[16];
    __shared__ int64_t sp[16];
    const int64_t ks = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < 16) {
        const int64_t slot = ks * 16 + tid;
        const int64_t pos = order ? (int64_t)order[slot] : slot;       // order: -1 for a padding slot
        const bool live = pos >= 0 && pos < npos;
        sp[tid] = live ? pos : -1;
        sr[tid] = live ? wr[pos] : 0.0;
        sq[tid] = live ? wc[pos] : 0.0;
    }
    __syncthreads();
    for (int v = tid; v < 16 * kSplitMaxK; v += 256) {
        const int e = v / kSplitMaxK, kk = v % kSplitMaxK;
        sdy[e][kk] = (sp[e] >= 0 && kk < k) ? dy[sp[e] * k + kk] : 0.0;
    }
    __syncthreads();
    const int c = tid >> 6, l = tid & 63, kk = l & 31;
    half8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int sl = 8 * (l >> 5) + e;
        double v = 0.0;
        if (kk < k && sp[sl] >= 0) {
            const double r = sr[sl], q = sq[sl];
            const double w = c == 0 ? (1.0 - r) * (1.0 - q) : c == 1 ? r * (1.0 - q) : c == 2 ? (1.0 - r) * q : r * q;
            v = sdy[sl][kk] * w * scale[kk];
        }
        hi[e] = (_Float16)(float)v;
        lo[e] = (_Float16)(float)(v - (double)(float)hi[e]);
    }
    dyw[((ks * 4 + c) * 2 + 0) * 64 + l] = hi;
    dyw[((ks * 4 + c) * 2 + 1) * 64 + l] = lo;
}
}  // namespace

hipError_t launch_split_dy(const double *dy, const double *wr, const double *wc, int64_t npos, int k,
                           const double *scale, const int32_t *order, int64_t nslots, void *dyw, hipStream_t s) {
    if (k > kSplitMaxK) return hipErrorInvalidValue;
    const int64_t nbatch = (nslots + 15) / 16;
    hipLaunchKernelGGL(split_dy, dim3((unsigned)nbatch), dim3(256), 0, s, dy, wr, wc, npos, k, scale, order,
                       (half8 *)dyw);
    return hipGetLastError();
}

// Output rows per block: blocks of 8 waves (6 for rows of 192 pixels), one per compute unit. (Blocks of 4 waves,
// two per compute unit, bands of half the rows: 3.85 instead of 3.55 ms on 128-pixel rows, the pattern rows re-read
// 1.5 instead of 1.25 times.)
int shift_grad_t_band(int qx) {
    switch (qx) {
        case 64: return 8;
        case 128: return 4;
        case 192: return 2;
        case 256: return 2;
        default: return 0;
    }
}

int shift_grad_t_blocks_per_cu(int) { return 1; }

bool shift_grad_t_supported(int qy, int qx) { return shift_grad_t_band(qx) > 0 && qy >= shift_grad_t_band(qx); }

hipError_t launch_shift_grad_t(const ShiftGradT &a, hipStream_t s) {
    if (!shift_grad_t_supported(a.qy, a.qx) || a.ngroups <= 0) return hipErrorInvalidValue;
    const int rb = shift_grad_t_band(a.qx);
    const dim3 grid((unsigned)((a.qy + rb - 1) / rb), (unsigned)a.ngroups);
    switch (a.qx) {
        case 64: hipLaunchKernelGGL((shift_grad_t<1, 8>), grid, dim3(512), 0, s, a); break;
        case 128: hipLaunchKernelGGL((shift_grad_t<2, 4>), grid, dim3(512), 0, s, a); break;
        case 192: hipLaunchKernelGGL((shift_grad_t<3, 2>), grid, dim3(384), 0, s, a); break;
        default: hipLaunchKernelGGL((shift_grad_t<4, 2>), grid, dim3(512), 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_shift_grad_rows(const ShiftGradRows &a, hipStream_t s) {
    if (!shift_grad_rows_supported(a.qx) || a.ngroups <= 0) return hipErrorInvalidValue;
    // three output rows per block up to 128-pixel rows (at most 12 waves, one block of them per compute unit),
    // one above
    const int rb = shift_grad_rows_band(a.qx);
    const dim3 grid((unsigned)((a.qy + rb - 1) / rb), (unsigned)a.ngroups);
    switch (a.qx / 32) {
    case 1: hipLaunchKernelGGL((shift_grad_rows<1, 3>), grid, dim3(192), 0, s, a); break;
    case 2: hipLaunchKernelGGL((shift_grad_rows<2, 3>), grid, dim3(384), 0, s, a); break;
    case 3: hipLaunchKernelGGL((shift_grad_rows<3, 3>), grid, dim3(576), 0, s, a); break;
    case 4: hipLaunchKernelGGL((shift_grad_rows<4, 3>), grid, dim3(768), 0, s, a); break;
    case 5: hipLaunchKernelGGL((shift_grad_rows<5, 1>), grid, dim3(320), 0, s, a); break;
    case 6: hipLaunchKernelGGL((shift_grad_rows<6, 1>), grid, dim3(384), 0, s, a); break;
    case 7: hipLaunchKernelGGL((shift_grad_rows<7, 1>), grid, dim3(448), 0, s, a); break;
    default: hipLaunchKernelGGL((shift_grad_rows<8, 1>), grid, dim3(512), 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_shift_valu(const ShiftValu &a, bool rounded, hipStream_t s) {
    if (a.npos <= 0) return hipSuccess;
    const dim3 grid((unsigned)((a.npos + 3) / 4));
    if (a.u16) {
        if (rounded) hipLaunchKernelGGL((shift_valu<true, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((shift_valu<true, false>), grid, dim3(256), 0, s, a);
    } else {
        if (rounded) hipLaunchKernelGGL((shift_valu<false, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((shift_valu<false, false>), grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_roll_weights(const void *base, void *out, const int32_t *shifts, int nslots, int nchunks, int64_t ndet,
                               int qy, int qx, hipStream_t s) {
    if (nslots <= 0) return hipSuccess;
    const int64_t per_slot = (int64_t)nchunks * kTileEntries;
    const dim3 grid((unsigned)((per_slot + 255) / 256), (unsigned)nslots);
    hipLaunchKernelGGL(roll_weights, grid, dim3(256), 0, s, (const half8 *)base, (half8 *)out, shifts, nchunks, ndet, qy,
                       qx);
    return hipGetLastError();
}

hipError_t launch_shift_grad_w(const ShiftGradW &a, hipStream_t s) {
    const int64_t ndet = (int64_t)a.qy * a.qx;
    const dim3 grid((unsigned)((ndet + 255) / 256), (unsigned)a.ngroups);
    switch (a.kp) {
        case 4: hipLaunchKernelGGL(shift_grad_w<4>, grid, dim3(256), 0, s, a); break;
        case 8: hipLaunchKernelGGL(shift_grad_w<8>, grid, dim3(256), 0, s, a); break;
        case 16: hipLaunchKernelGGL(shift_grad_w<16>, grid, dim3(256), 0, s, a); break;
        case 24: hipLaunchKernelGGL(shift_grad_w<24>, grid, dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL(shift_grad_w<32>, grid, dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sum_groups(const double *partial, int ngroups, int64_t n, double *out, hipStream_t s) {
    hipLaunchKernelGGL(sum_groups, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partial, ngroups, n, out);
    return hipGetLastError();
}

}  // namespace cubo
