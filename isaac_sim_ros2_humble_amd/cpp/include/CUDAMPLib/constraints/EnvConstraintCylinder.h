// CUDAMPLib finite cylinder obstacles (cudampl/CUDAMPLib/include/constraints/EnvConstraintCylinder.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <vector>

#include "../base/Constraint.h"

namespace CUDAMPLib {

// EnvConstraintCylinder(name, pos[N][3], orientation R[N][9], radius[N], height[N])
class EnvConstraintCylinder : public BaseConstraint {
  public:
    EnvConstraintCylinder(const std::string& name, const std::vector<std::vector<float>>& pos,
                          const std::vector<std::vector<float>>& orientation, const std::vector<float>& radius,
                          const std::vector<float>& height)
        : BaseConstraint(name, false), pos_(pos), rot_(orientation), radius_(radius), height_(height) {}
    std::vector<std::vector<float>> pos_, rot_;
    std::vector<float> radius_, height_;
};
typedef std::shared_ptr<EnvConstraintCylinder> EnvConstraintCylinderPtr;

}  // namespace CUDAMPLib
