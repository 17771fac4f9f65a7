// CUDAMPLib joint bounds (cudampl/CUDAMPLib/include/constraints/BoundaryConstraint.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <vector>

#include "../base/Constraint.h"

namespace CUDAMPLib {

// BoundaryConstraint(name, lower, upper, active)
class BoundaryConstraint : public BaseConstraint {
  public:
    BoundaryConstraint(const std::string& name, const std::vector<float>& lower, const std::vector<float>& upper,
                       const std::vector<bool>& active)
        : BaseConstraint(name, false), lower_(lower), upper_(upper), active_(active) {}
    std::vector<float> lower_, upper_;
    std::vector<bool> active_;
};
typedef std::shared_ptr<BoundaryConstraint> BoundaryConstraintPtr;

}  // namespace CUDAMPLib
