// timing.hpp — device time of selected launches on a stream, from HIP events.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

namespace cubo {

// Events around launches whose device time a caller may ask for later (cubo_hip_cube_kernel_time,
// cubo_hip_batch_time). Events are recycled rather than created per launch, and a caller that never asks holds
// at most kMaxPending pairs: past that, the oldest half is folded into running totals (those launches are long
// complete). A start without a stop (a launch that failed in between) is recycled by the next start.
class LaunchTimer {
public:
    static constexpr size_t kMaxPending = 256;

    LaunchTimer() = default;
    LaunchTimer(const LaunchTimer &) = delete;
    LaunchTimer &operator=(const LaunchTimer &) = delete;
    ~LaunchTimer() { release(); }

    // Records the start of one launch on s.
    hipError_t start(hipStream_t s) {
        if (!open_) {
            if (hipError_t e = take(&open_)) return e;
        }
        return hipEventRecord(open_, s);
    }

    // Records its end on s; the launch then counts in total().
    hipError_t stop(hipStream_t s) {
        if (!open_) return hipErrorInvalidHandle;
        hipEvent_t end = nullptr;
        if (hipError_t e = take(&end)) return e;
        if (hipError_t e = hipEventRecord(end, s)) { spare_.push_back(end); return e; }
        pending_.emplace_back(open_, end);
        open_ = nullptr;
        if (pending_.size() > kMaxPending) return fold(kMaxPending / 2);
        return hipSuccess;
    }

    // The summed device time (ms) and count of the launches since the last call; waits for them.
    hipError_t total(double *ms, int64_t *launches) {
        const hipError_t e = fold(pending_.size());
        *ms = folded_ms_;
        *launches = folded_;
        folded_ms_ = 0;
        folded_ = 0;
        return e;
    }

    // Destroys every event (the stream must be idle).
    void release() {
        for (auto &p : pending_) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
        for (hipEvent_t e : spare_) (void)hipEventDestroy(e);
        if (open_) (void)hipEventDestroy(open_);
        pending_.clear();
        spare_.clear();
        open_ = nullptr;
    }

private:
    hipError_t take(hipEvent_t *e) {
        if (!spare_.empty()) {
            *e = spare_.back();
            spare_.pop_back();
            return hipSuccess;
        }
        return hipEventCreate(e);
    }

    // The n oldest pending launches into the totals; their events go back to the spares.
    hipError_t fold(size_t n) {
        hipError_t err = hipSuccess;
        size_t i = 0;
        for (; i < n && i < pending_.size(); ++i) {
            float t = 0;
            hipError_t e = hipEventSynchronize(pending_[i].second);
            if (e == hipSuccess) e = hipEventElapsedTime(&t, pending_[i].first, pending_[i].second);
            if (e == hipSuccess) {
                folded_ms_ += t;
                ++folded_;
            } else if (err == hipSuccess) {
                err = e;
            }
            spare_.push_back(pending_[i].first);
            spare_.push_back(pending_[i].second);
        }
        pending_.erase(pending_.begin(), pending_.begin() + (ptrdiff_t)i);
        return err;
    }

    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_;
    std::vector<hipEvent_t> spare_;
    hipEvent_t open_ = nullptr;
    double folded_ms_ = 0;
    int64_t folded_ = 0;
};

}  // namespace cubo
