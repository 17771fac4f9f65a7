// CUDAMPLib sphere obstacles (cudampl/CUDAMPLib/include/constraints/EnvConstraintSphere.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <vector>

#include "../base/Constraint.h"

namespace CUDAMPLib {

// EnvConstraintSphere(name, pos[N][3], radius[N])  (EnvConstraintSphere.cu:10-33)
class EnvConstraintSphere : public BaseConstraint {
  public:
    EnvConstraintSphere(const std::string& name, const std::vector<std::vector<float>>& pos,
                        const std::vector<float>& radius)
        : BaseConstraint(name, false), pos_(pos), radius_(radius) {}
    std::vector<std::vector<float>> pos_;
    std::vector<float> radius_;
};
typedef std::shared_ptr<EnvConstraintSphere> EnvConstraintSpherePtr;

}  // namespace CUDAMPLib
