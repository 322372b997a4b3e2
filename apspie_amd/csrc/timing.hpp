// hipEvent timing of the dominant kernel of an algorithm call (enabled by the
// context option "profiling"); the sum and launch count land in the stats.
#pragma once

#include <vector>

#include "common.hpp"

namespace apspie {

class KernelTimer {
  public:
    explicit KernelTimer(Ctx *c) : c_(c), on_(c->profiling) {}
    ~KernelTimer() {
        for (auto &p : ev_) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }
    void begin() {
        if (!on_) return;
        std::pair<hipEvent_t, hipEvent_t> p;
        APSPIE_HIP(hipEventCreate(&p.first));
        APSPIE_HIP(hipEventCreate(&p.second));
        APSPIE_HIP(hipEventRecord(p.first, c_->stream));
        ev_.push_back(p);
    }
    void end() {
        if (!on_) return;
        APSPIE_HIP(hipEventRecord(ev_.back().second, c_->stream));
    }
    // sums the first `limit` recorded intervals (all if negative) into the
    // context stats; launches past the limit are the early-exit ones issued
    // after the work ran out
    void collect(int64_t limit = -1) {
        if (!on_) return;
        double total = 0;
        int64_t n = 0;
        for (auto &p : ev_) {
            if (limit >= 0 && n >= limit) break;
            APSPIE_HIP(hipEventSynchronize(p.second));
            float ms = 0;
            APSPIE_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            total += ms;
            ++n;
        }
        c_->stats.kernel_ms += total;
        c_->stats.kernel_launches += n;
    }

  private:
    Ctx *c_;
    bool on_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_;
};

}  // namespace apspie
