// CUDAMPLib self collision (cudampl/CUDAMPLib/include/constraints/SelfCollisionConstraint.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <vector>

#include "../base/Constraint.h"

namespace CUDAMPLib {

// SelfCollisionConstraint(name, sphere_map, radius, enables[L][L])
class SelfCollisionConstraint : public BaseConstraint {
  public:
    SelfCollisionConstraint(const std::string& name, const std::vector<int>& sphere_map,
                            const std::vector<float>& radius, const std::vector<std::vector<bool>>& enables)
        : BaseConstraint(name, false), enables_(enables) {
        (void)sphere_map;
        (void)radius;
    }
    std::vector<std::vector<bool>> enables_;
};
typedef std::shared_ptr<SelfCollisionConstraint> SelfCollisionConstraintPtr;

}  // namespace CUDAMPLib
