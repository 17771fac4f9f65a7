// CUDAMPLib single-kernel RRT-Connect planner (cudampl/CUDAMPLib/include/planners/cRRTC.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>

#include "../base/Planner.h"

namespace CUDAMPLib {

// Single-kernel parallel RRT-Connect (cRRTC.cu / CPRRTC semantics).  Each termination step runs
// one GPU solve of up to max_iterations samples; solve() stops at the first success or when the
// termination condition fires (failure reason "MeetTerminationCondition").
class cRRTC : public BasePlanner {
  public:
    cRRTC(BaseSpacePtr space) : BasePlanner(space) {}
    void setMotionTask(BaseTaskPtr task, bool get_full_path = true) override {
        task_ = task;
        get_full_path_ = get_full_path;
    }
    void solve(BaseTerminationPtr termination_condition) override;
    void setMaxIterations(int n) { max_iterations_ = n; }

  protected:
    BaseTaskPtr task_;
    int max_iterations_ = 10000;
};
typedef std::shared_ptr<cRRTC> cRRTCPtr;

}  // namespace CUDAMPLib
