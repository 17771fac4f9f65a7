// CUDAMPLib task base class (cudampl/CUDAMPLib/include/base/Task.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>
#include <string>
#include <vector>

namespace CUDAMPLib {

class BaseTask {
  public:
    BaseTask() : has_solution(false) {}
    virtual ~BaseTask() {}
    virtual std::vector<std::vector<float>> getStartStatesVector() = 0;
    virtual std::vector<std::vector<float>> getGoalStatesVector() = 0;
    virtual void setSolution(const std::vector<std::vector<float>>& solution) = 0;
    bool hasSolution() const { return has_solution; }
    void setFailureReason(const std::string& reason) { failure_reason = reason; }
    std::string getFailureReason() const { return failure_reason; }

  protected:
    bool has_solution;
    std::string failure_reason;
};
typedef std::shared_ptr<BaseTask> BaseTaskPtr;

}  // namespace CUDAMPLib
