// CUDAMPLib constraint base class (cudampl/CUDAMPLib/include/base/Constraint.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>
#include <string>

namespace CUDAMPLib {

class BaseConstraint {
  public:
    BaseConstraint(const std::string& name, bool projectable) : constraint_name(name), is_projectable(projectable) {}
    virtual ~BaseConstraint() {}
    std::string getName() const { return constraint_name; }
    bool isProjectable() const { return is_projectable; }

  protected:
    std::string constraint_name;
    bool is_projectable;
};
typedef std::shared_ptr<BaseConstraint> BaseConstraintPtr;

}  // namespace CUDAMPLib
