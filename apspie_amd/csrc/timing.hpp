// hipEvent timing of the dominant kernel of an algorithm call (enabled by the
// context option "profiling"); the sum and launch count land in the stats.
#pragma once

#include <vector>

#include "common.hpp"

namespace apspie {

// The event pairs come from a pool kept by the context (created on first
// use, destroyed with it): a call records into them without creating any.
class KernelTimer {
  public:
    explicit KernelTimer(Ctx *c, bool allowed = true) : c_(c), on_(allowed && c->profiling) {}
    // an interval around `launches` back-to-back launches of the kernel (a
    // span over several launches keeps the events from separating them)
    void begin() {
        if (!on_) return;
        auto &pool = c_->timer_events;
        if (used_ == pool.size()) {
            std::pair<hipEvent_t, hipEvent_t> p{nullptr, nullptr};
            APSPIE_HIP(hipEventCreate(&p.first));
            const hipError_t e = hipEventCreate(&p.second);
            if (e != hipSuccess) (void)hipEventDestroy(p.first);
            APSPIE_HIP(e);
            pool.push_back(p);
        }
        APSPIE_HIP(hipEventRecord(pool[used_].first, c_->stream));
        ++used_;
    }
    void end(int64_t launches = 1) {
        if (!on_) return;
        APSPIE_HIP(hipEventRecord(c_->timer_events[used_ - 1].second, c_->stream));
        n_.push_back(launches);
    }
    // sums the recorded intervals and their launch counts into the stats
    void collect() {
        if (!on_) return;
        double total = 0;
        int64_t n = 0;
        for (size_t i = 0; i < used_; ++i) {
            const auto &p = c_->timer_events[i];
            APSPIE_HIP(hipEventSynchronize(p.second));
            float ms = 0;
            APSPIE_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            total += ms;
            n += n_[i];
        }
        c_->stats.kernel_ms += total;
        c_->stats.kernel_launches += n;
    }

  private:
    Ctx *c_;
    bool on_;
    size_t used_ = 0;
    std::vector<int64_t> n_;
};

}  // namespace apspie
