// CUDAMPLib termination base class (cudampl/CUDAMPLib/include/base/Termination.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>

namespace CUDAMPLib {

class BaseTermination {
  public:
    virtual ~BaseTermination() {}
    virtual void reset() = 0;
    virtual bool checkTerminationCondition() = 0;
    virtual void printTerminationReason() = 0;
};
typedef std::shared_ptr<BaseTermination> BaseTerminationPtr;

}  // namespace CUDAMPLib
