// CUDAMPLib single-arm space (cudampl/CUDAMPLib/include/spaces/SingleArmSpace.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../mp_eigen.h"
#include "../base/Space.h"
#include "../states/SingleArmStates.h"

struct mp_robot;
struct mp_scene;

namespace CUDAMPLib {

class SingleArmSpace : public BaseSpace {
  public:
    SingleArmSpace(size_t dim, const std::vector<BaseConstraintPtr>& constraints, const std::vector<int>& joint_types,
                   const std::vector<mpx::Isometry3d>& joint_poses, const std::vector<mpx::Vector3d>& joint_axes,
                   const std::vector<int>& link_parent_link_maps,
                   const std::vector<int>& self_collision_spheres_to_link_map,
                   const std::vector<std::vector<float>>& self_collision_spheres_pos_in_link,
                   const std::vector<float>& self_collision_spheres_radius, const std::vector<bool>& active_joint_map,
                   const std::vector<float>& lower, const std::vector<float>& upper,
                   const std::vector<float>& default_joint_values, const std::vector<std::string>& link_names,
                   float resolution = 0.02f);
    ~SingleArmSpace() override;
    const std::vector<float>& getLower() const { return lower_bound; }
    const std::vector<float>& getUpper() const { return upper_bound; }
    BaseStatesPtr sample(int num_of_config) override;
    void checkStates(const BaseStatesPtr& states, std::vector<bool>& state_feasibility) override;
    void checkStates(const BaseStatesPtr& states) override;
    bool checkMotions(const BaseStatesPtr& states1, const BaseStatesPtr& states2, std::vector<bool>& motion_feasibility,
                      std::vector<float>& motion_costs) override;
    void interpolate(const BaseStatesPtr& from_states, const BaseStatesPtr& to_states, float max_distance) override;
    BaseStatesPtr getPathFromWaypoints(const BaseStatesPtr& waypoints) override;
    BaseStateManagerPtr createStateManager() override;
    BaseStatesPtr createStatesFromVector(const std::vector<std::vector<float>>& joint_values);
    BaseStatesPtr createStatesFromVectorFull(const std::vector<std::vector<float>>& joint_values);
    std::vector<std::vector<float>> getJointVectorInActiveJointsFromStates(const BaseStatesPtr& states);
    float getResolution() const { return resolution_; }
    // access for the planners of this library
    mp_robot* robot() const { return robot_; }
    mp_scene* scene() const { return scene_; }
    // n samples of the space's stream written to device memory d_q[n][dim] (as sample(n))
    void sampleInto(float* d_q, int n);
    // Space.h sampleConfigurations: the same into a caller's device buffer
    void sampleConfigurations(float* d_configurations, int num_of_config) { sampleInto(d_configurations, num_of_config); }
    // the space's sample stream (mp_sample's seed and the index of its next sample), for planners
    // that draw from it on the device; advanceSampleStream skips the n samples they drew
    unsigned long long sampleStreamSeed() const { return seed_; }
    unsigned long long sampleStreamOffset() const { return sample_offset_; }
    void advanceSampleStream(unsigned long long n) { sample_offset_ += n; }
    // The generated HIP module of this robot (FK, sphere placement, self / environment checks
    // and the planning kernels), compiled with hipRTC for gfx950.  Both reference hooks
    // (generateFKKernelSourceCode, generateRobotCollisionModelSourceCode) return it.
    std::string generateFKKernelSourceCode() const;
    std::string generateRobotCollisionModelSourceCode() const { return generateFKKernelSourceCode(); }

  private:
    std::vector<float> lower_bound, upper_bound, defaults_;
    std::vector<bool> active_;
    std::vector<std::string> link_names_;
    int num_spheres_ = 0;
    float resolution_;
    unsigned long long sample_offset_ = 0;
    unsigned long long seed_ = 0x5eed;
    mp_robot* robot_ = nullptr;
    mp_scene* scene_ = nullptr;
    // per constraint (in the order given): a scene holding only that constraint (its obstacles,
    // or its joint bounds) and the cost row of mp_constraint_costs that is its cost
    std::vector<mp_scene*> con_scenes_;
    std::vector<int> con_rows_;
};
typedef std::shared_ptr<SingleArmSpace> SingleArmSpacePtr;

}  // namespace CUDAMPLib
