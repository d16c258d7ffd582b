// RAII-style event bracket for one kernel launch (see timing.hip).
#pragma once

#include <string>

#include "common.hpp"

namespace coh {

class KernelTimer {
 public:
  KernelTimer(hipStream_t st, const char* name, double bytes, double flops, double ops);
  ~KernelTimer();  // drops a bracket that was never stopped (a launch that failed)
  KernelTimer(const KernelTimer&) = delete;
  KernelTimer& operator=(const KernelTimer&) = delete;
  void stop();

 private:
  hipStream_t st_;
  bool active_;
  std::string name_;
  double bytes_ = 0, flops_ = 0, ops_ = 0;
  hipEvent_t start_{}, stop_{};
};

}  // namespace coh
