// CUDAMPLib step termination (cudampl/CUDAMPLib/include/termination/StepTermination.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <cstdio>
#include <memory>

#include "../base/Termination.h"

namespace CUDAMPLib {

// StepTermination.cpp:1-13: terminate when the step count exceeds the limit
class StepTermination : public BaseTermination {
  public:
    StepTermination(int step_limit) : step_limit_(step_limit) {}
    void reset() override { step_count_ = 0; }
    bool checkTerminationCondition() override { return ++step_count_ > step_limit_; }
    void printTerminationReason() override { std::printf("Step limit reached\n"); }

  private:
    int step_limit_, step_count_ = 0;
};
typedef std::shared_ptr<StepTermination> StepTerminationPtr;

}  // namespace CUDAMPLib
