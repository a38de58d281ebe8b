// Launch-level instrumentation: when enabled, instrumented launches record a pair of HIP
// events on their own stream together with the launch's algorithmic work (flops and bytes,
// counted on the host from the instances the launch processes). Reports sum event times.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace msx {

bool profiling();

// RAII scope around one kernel launch.
class ProfScope {
  public:
    ProfScope(const char *name, hipStream_t s, double flops, double bytes);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

  private:
    int slot_ = -1;
    hipStream_t s_ = 0;
};

}  // namespace msx
