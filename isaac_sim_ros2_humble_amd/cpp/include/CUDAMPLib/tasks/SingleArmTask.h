// CUDAMPLib single-arm task (cudampl/CUDAMPLib/include/tasks/SingleArmTask.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>
#include <vector>

#include "../base/Space.h"
#include "../base/Task.h"

namespace CUDAMPLib {

class SingleArmTask : public BaseTask {
  public:
    SingleArmTask(const std::vector<std::vector<float>>& start_joint_values,
                  const std::vector<std::vector<float>>& goal_joint_values)
        : start_(start_joint_values), goal_(goal_joint_values) {}
    // Goal region (SingleArmTask.cpp:12-37): the free states among goal_config_sampling_attempts
    // samples of the goal-region space, without repeats (componentwise within 1e-5).  The
    // region's constraints are those a SingleArmSpace checks; there is no projection step.
    SingleArmTask(const std::vector<std::vector<float>>& start_joint_values, const BaseSpacePtr& goal_region,
                  int goal_config_sampling_attempts = 50);
    std::vector<std::vector<float>> getStartStatesVector() override { return start_; }
    std::vector<std::vector<float>> getGoalStatesVector() override { return goal_; }
    void setSolution(const std::vector<std::vector<float>>& solution) override {
        has_solution = true;
        solution_ = solution;
    }
    std::vector<std::vector<float>> getSolution() { return solution_; }

  private:
    std::vector<std::vector<float>> start_, goal_, solution_;
};
typedef std::shared_ptr<SingleArmTask> SingleArmTaskPtr;

}  // namespace CUDAMPLib
