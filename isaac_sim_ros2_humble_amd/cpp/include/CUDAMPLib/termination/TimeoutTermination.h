// CUDAMPLib timeout termination (cudampl/CUDAMPLib/include/termination/TimeoutTermination.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <chrono>
#include <cstdio>
#include <memory>

#include "../base/Termination.h"

namespace CUDAMPLib {

// TimeoutTermination.cpp:1-15: elapsed ms >= timeout * 1000
class TimeoutTermination : public BaseTermination {
  public:
    TimeoutTermination(float timeout_seconds) : timeout_ms_((long long)(timeout_seconds * 1000.0f)) {}
    void reset() override { t0_ = std::chrono::high_resolution_clock::now(); }
    bool checkTerminationCondition() override {
        auto dt = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::high_resolution_clock::now() - t0_);
        return dt.count() >= timeout_ms_;
    }
    void printTerminationReason() override { std::printf("Timeout reached\n"); }

  private:
    long long timeout_ms_;
    std::chrono::high_resolution_clock::time_point t0_ = std::chrono::high_resolution_clock::now();
};
typedef std::shared_ptr<TimeoutTermination> TimeoutTerminationPtr;

}  // namespace CUDAMPLib
