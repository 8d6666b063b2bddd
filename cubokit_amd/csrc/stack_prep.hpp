// stack_prep.hpp — host preparation of mask-stack weights for the split-FP16 matrix-core kernel.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace cubo {

// Weights of one 64-mask pass in MFMA B-operand order [chunk][tile][kstep][hi, lo][lane][8 halves]: lane l
// holds mask 32*tile + l%32 and pixels 64*chunk + 32*(l/32) + 8*kstep + e, the pixels the A operand's lane
// holds for its row. Each column is scaled by 2^(10 - floor(log2 max|w|)) so its largest weight lies in
// [1024, 2048), then split into FP16 hi = fp16(v) and lo = fp16(v - hi).
struct SplitWeights {
    std::vector<_Float16> wb;      // passes * nchunks * ntiles * 4 * 2 * 64 * 8
    std::vector<double> inv;       // k: 1 / scale
    std::vector<int> tile_lo;      // passes * 2: the tile has a non-zero lo part
    int ntiles = 0, passes = 0, nchunks = 0;
};

// One 32-mask tile for at most 32 masks, else two per pass of 64.
inline SplitWeights split_weights(const float *masks, int k, int64_t ndet) {
    SplitWeights s;
    s.ntiles = k <= 32 ? 1 : 2;
    s.passes = (k + 63) / 64;
    s.nchunks = (int)((ndet + 63) / 64);
    std::vector<double> scale(k, 1.0);
    s.inv.assign(k, 1.0);
    for (int m = 0; m < k; ++m) {
        double mx = 0;
        for (int64_t i = 0; i < ndet; ++i) mx = std::max(mx, (double)std::fabs(masks[(size_t)m * ndet + i]));
        if (mx > 0) {
            const int ex = (int)std::floor(std::log2(mx));
            scale[m] = std::ldexp(1.0, 10 - ex);
            s.inv[m] = std::ldexp(1.0, ex - 10);
        }
    }
    const size_t per_chunk = (size_t)s.ntiles * 4 * 2 * 64 * 8;
    const size_t per_pass = per_chunk * s.nchunks;
    s.wb.assign(per_pass * s.passes, (_Float16)0.f);
    s.tile_lo.assign((size_t)s.passes * 2, 0);
    for (int pass = 0; pass < s.passes; ++pass)
        for (int ch = 0; ch < s.nchunks; ++ch)
            for (int t = 0; t < s.ntiles; ++t)
                for (int st = 0; st < 4; ++st)
                    for (int l = 0; l < 64; ++l)
                        for (int el = 0; el < 8; ++el) {
                            const int m = pass * 64 + t * 32 + (l & 31);
                            const int64_t px = (int64_t)ch * 64 + 32 * (l >> 5) + 8 * st + el;
                            double v = 0;
                            if (m < k && px < ndet) v = (double)masks[(size_t)m * ndet + px] * scale[m];
                            const _Float16 hi = (_Float16)(float)v;
                            const _Float16 lo = (_Float16)(float)(v - (double)(float)hi);
                            const size_t base =
                                pass * per_pass + (((((size_t)ch * s.ntiles + t) * 4 + st) * 2) * 64 + l) * 8 + el;
                            s.wb[base] = hi;
                            s.wb[base + 64 * 8] = lo;
                            if ((float)lo != 0.f) s.tile_lo[pass * 2 + t] = 1;
                        }
    return s;
}

}  // namespace cubo
