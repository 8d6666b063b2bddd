// host_math.hpp — host-side helpers shared by the C ABI translation units: NumPy's pairwise sum (the mean of an
// origin map), the wrap of an integral shift into a detector axis, and a pool for large host copies.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace cubo {

// NumPy's pairwise float64 sum (numpy/core/src/umath/loops_utils.h.src pairwise_sum): fewer than 8 values
// in order; up to 128 in eight interleaved partials; above, split at a multiple of 8 and recurse.
inline double pairwise_sum(const double *a, int64_t n) {
    if (n < 8) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += a[i];
        return s;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        int64_t i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) s += a[i];
        return s;
    }
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
}

// An integral shift wrapped into [0, n), reduced modulo n in Double first so that any finite value is safe
// (ShiftedImaging.swift:108-112).
inline int wrap_shift(double v, int n) {
    double m = std::fmod(v, (double)n);
    if (m < 0) m += (double)n;
    return ((int)m) % n;
}

// Threads for large host copies and file reads: the hardware threads, at most 16.
inline int reader_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, h ? h : 1u));
}

// Runs fn(offset, bytes) over [0, n) in `threads` contiguous pieces (4 KiB aligned), in parallel.
template <typename Fn>
bool parallel_pieces(size_t n, int threads, Fn fn) {
    const size_t piece = std::max<size_t>(((n / threads) + 4095) & ~(size_t)4095, 4096);
    std::vector<std::thread> pool;
    std::vector<char> ok((n + piece - 1) / piece, 1);
    size_t k = 0;
    for (size_t off = 0; off < n; off += piece, ++k) {
        const size_t len = std::min(piece, n - off);
        if (off + len >= n && pool.empty()) { ok[k] = fn(off, len); break; }   // one piece: no thread
        pool.emplace_back([&, off, len, k] { ok[k] = fn(off, len); });
    }
    for (auto &t : pool) t.join();
    return std::all_of(ok.begin(), ok.end(), [](char v) { return v != 0; });
}

}  // namespace cubo
