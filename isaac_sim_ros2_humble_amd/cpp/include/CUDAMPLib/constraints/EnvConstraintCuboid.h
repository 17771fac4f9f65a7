// CUDAMPLib oriented box obstacles (cudampl/CUDAMPLib/include/constraints/EnvConstraintCuboid.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <vector>

#include "../base/Constraint.h"

namespace CUDAMPLib {

// EnvConstraintCuboid(name, pos[N][3], orientation R[N][9] row major, max[N][3], min[N][3])
class EnvConstraintCuboid : public BaseConstraint {
  public:
    EnvConstraintCuboid(const std::string& name, const std::vector<std::vector<float>>& pos,
                        const std::vector<std::vector<float>>& orientation, const std::vector<std::vector<float>>& max,
                        const std::vector<std::vector<float>>& min)
        : BaseConstraint(name, false), pos_(pos), rot_(orientation), max_(max), min_(min) {}
    std::vector<std::vector<float>> pos_, rot_, max_, min_;
};
typedef std::shared_ptr<EnvConstraintCuboid> EnvConstraintCuboidPtr;

}  // namespace CUDAMPLib
