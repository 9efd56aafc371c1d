// ktimer.hpp -- optional per-kernel timing with HIP events recorded on the
// stream the kernels run on.  Disabled by default (zero cost: one branch).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"

namespace lsdtt_gpu {

class KernelTimer {
public:
    ~KernelTimer()
    {
        for (auto& e : pool_) (void)hipEventDestroy(e);
    }
    bool enabled = false;

    // bracket one launch of kernel `name` on stream s
    void begin(const char* name, hipStream_t s)
    {
        if (!enabled) return;
        pending_ = name;
        hipEvent_t a = take();
        LSDTT_HIP(hipEventRecord(a, s));
        cur_ = a;
    }
    void end(hipStream_t s, double bytes)
    {
        if (!enabled) return;
        hipEvent_t b = take();
        LSDTT_HIP(hipEventRecord(b, s));
        rec_.push_back({pending_, cur_, b, bytes});
    }
    // resolve recorded pairs (after the stream has drained)
    void collect()
    {
        for (auto& r : rec_) {
            float ms = 0.0f;
            LSDTT_HIP(hipEventSynchronize(r.b));
            LSDTT_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            auto& t = totals_[r.name];
            t.ms += ms;
            t.launches += 1;
            t.bytes += r.bytes;
            free_.push_back(r.a);
            free_.push_back(r.b);
        }
        rec_.clear();
    }
    void reset()
    {
        collect();
        totals_.clear();
    }
    struct Total {
        double ms = 0.0, bytes = 0.0;
        long long launches = 0;
    };
    const std::map<std::string, Total>& totals() const { return totals_; }

private:
    struct Rec {
        std::string name;
        hipEvent_t a, b;
        double bytes;
    };
    hipEvent_t take()
    {
        if (free_.empty()) {
            hipEvent_t e;
            LSDTT_HIP(hipEventCreate(&e));
            pool_.push_back(e);
            return e;
        }
        hipEvent_t e = free_.back();
        free_.pop_back();
        return e;
    }
    std::vector<hipEvent_t> pool_, free_;
    std::vector<Rec> rec_;
    std::map<std::string, Total> totals_;
    std::string pending_;
    hipEvent_t cur_{};
};

}  // namespace lsdtt_gpu
