// CUDAMPLib space base class (cudampl/CUDAMPLib/include/base/Space.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>
#include <vector>

#include "Constraint.h"
#include "States.h"

namespace CUDAMPLib {

class BaseSpace {
  public:
    virtual ~BaseSpace() {}
    virtual BaseStatesPtr sample(int num_of_config) = 0;
    virtual void checkStates(const BaseStatesPtr& states, std::vector<bool>& state_feasibility) = 0;
    virtual void checkStates(const BaseStatesPtr& states) = 0;
    virtual bool checkMotions(const BaseStatesPtr& states1, const BaseStatesPtr& states2,
                              std::vector<bool>& motion_feasibility, std::vector<float>& motion_costs) = 0;
    virtual void interpolate(const BaseStatesPtr& from_states, const BaseStatesPtr& to_states, float max_distance) = 0;
    virtual BaseStatesPtr getPathFromWaypoints(const BaseStatesPtr& waypoints) = 0;
    virtual BaseStateManagerPtr createStateManager() = 0;
    size_t getDim() const { return dim_; }

  protected:
    size_t dim_ = 0;
};
typedef std::shared_ptr<BaseSpace> BaseSpacePtr;

}  // namespace CUDAMPLib
