// CUDAMPLib planner base class (cudampl/CUDAMPLib/include/base/Planner.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>

#include "Space.h"
#include "Task.h"
#include "Termination.h"

namespace CUDAMPLib {

class BasePlanner {
  public:
    BasePlanner(BaseSpacePtr space) { space_ = space; }
    virtual ~BasePlanner() {}
    virtual void setMotionTask(BaseTaskPtr task, bool get_full_path) = 0;
    virtual void solve(BaseTerminationPtr termination_condition) = 0;
    BaseSpacePtr getSpace() { return space_; }

  protected:
    BaseSpacePtr space_;
    bool get_full_path_ = true;
};
typedef std::shared_ptr<BasePlanner> BasePlannerPtr;

}  // namespace CUDAMPLib
