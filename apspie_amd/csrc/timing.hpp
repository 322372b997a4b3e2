// hipEvent timing of the dominant kernel of an algorithm call (enabled by the
// context option "profiling"); the sum and launch count land in the stats.
#pragma once

#include <vector>

#include "common.hpp"

namespace apspie {

class KernelTimer {
  public:
    explicit KernelTimer(Ctx *c, bool allowed = true) : c_(c), on_(allowed && c->profiling) {}
    ~KernelTimer() {
        for (auto &p : ev_) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }
    // an interval around `launches` back-to-back launches of the kernel (a
    // span over several launches keeps the events from separating them)
    void begin() {
        if (!on_) return;
        std::pair<hipEvent_t, hipEvent_t> p;
        APSPIE_HIP(hipEventCreate(&p.first));
        APSPIE_HIP(hipEventCreate(&p.second));
        APSPIE_HIP(hipEventRecord(p.first, c_->stream));
        ev_.push_back(p);
    }
    void end(int64_t launches = 1) {
        if (!on_) return;
        APSPIE_HIP(hipEventRecord(ev_.back().second, c_->stream));
        n_.push_back(launches);
    }
    // sums the recorded intervals and their launch counts into the stats
    void collect() {
        if (!on_) return;
        double total = 0;
        int64_t n = 0;
        for (size_t i = 0; i < ev_.size(); ++i) {
            APSPIE_HIP(hipEventSynchronize(ev_[i].second));
            float ms = 0;
            APSPIE_HIP(hipEventElapsedTime(&ms, ev_[i].first, ev_[i].second));
            total += ms;
            n += n_[i];
        }
        c_->stats.kernel_ms += total;
        c_->stats.kernel_launches += n;
    }

  private:
    Ctx *c_;
    bool on_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_;
    std::vector<int64_t> n_;
};

}  // namespace apspie
