// Launch-level instrumentation: when enabled, instrumented launches record a pair of HIP
// events on their own stream together with the launch's algorithmic work (flops and bytes,
// counted on the host from the instances the launch processes). Reports sum event times.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <string>
#include <vector>

namespace msx {

bool profiling();

// Host copy of a per-instance active mask. A solver that checks convergence one iteration
// late issues launches before it knows which instances they will process: their work refers
// to a snapshot that is filled (and marked ready) when that mask arrives.
struct ActiveSnap {
    std::vector<int> on;
    std::atomic<bool> ready{true};
};

// Algorithmic work of one launch: c plus per[b] for every instance b active in snap (every
// instance when snap is null).
struct Work {
    double c = 0;
    std::vector<double> per;
    std::shared_ptr<const ActiveSnap> snap;
    Work() = default;
    Work(double v) : c(v) {}
    bool ready() const { return !snap || snap->ready.load(std::memory_order_acquire); }
    double value() const {
        double v = c;
        for (size_t b = 0; b < per.size(); ++b)
            if (!snap || (b < snap->on.size() && snap->on[b])) v += per[b];
        return v;
    }
};
inline Work operator+(Work a, double b) {
    a.c += b;
    return a;
}

// RAII scope around one kernel launch.
class ProfScope {
  public:
    ProfScope(const char *name, hipStream_t s, Work flops, Work bytes);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

  private:
    long long id_ = -1;
    hipStream_t s_ = 0;
};

}  // namespace msx
