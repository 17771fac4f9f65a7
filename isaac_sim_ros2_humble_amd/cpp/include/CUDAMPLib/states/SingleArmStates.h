// CUDAMPLib single-arm states and their manager (cudampl/CUDAMPLib/include/states/SingleArmStates.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../mp_eigen.h"
#include "../base/States.h"

struct mp_robot;

namespace CUDAMPLib {

// Joint states of one arm: active-joint values in a device buffer [n][dof]; full-joint views are
// produced on the host (inactive joints at their default values).
class SingleArmStates : public BaseStates {
  public:
    SingleArmStates(const std::vector<std::vector<float>>& active_values, const std::vector<bool>& active_map,
                    const std::vector<float>& defaults);
    ~SingleArmStates() override;
    // Forward kinematics of every state: link poses and sphere centres into device buffers
    // (SingleArmStates::update, SingleArmStates.cu:696-723).  The collision checks fuse FK and
    // need no update; the pose and centre getters run it on first use after a change.
    void update();
    float* getJointStatesCuda() { return d_q_; }  // [n][dof] active joints
    std::vector<std::vector<float>> getJointStatesHost() const { return q_; }
    std::vector<std::vector<float>> getJointStatesFullHost() const;
    void filterStates(const std::vector<bool>& keep);
    float* getLinkPosesInBaseLinkCuda();                // [n][L][16] row-major 4x4 per link
    float* getSelfCollisionSpheresPosInBaseLinkCuda();  // [n][S][3]
    std::vector<std::vector<std::vector<float>>> getSelfCollisionSpheresPosInBaseLinkHost();
    std::vector<std::vector<mpx::Isometry3d>> getLinkPosesInBaseLinkHost();
    std::vector<mpx::Isometry3d> getLinkPoseInBaseLinkHost(const std::string& link_name);

  private:
    std::vector<std::vector<float>> q_;
    std::vector<bool> active_;
    std::vector<float> defaults_;
    float* d_q_ = nullptr;
    // the robot of the space that created these states (the space outlives its states)
    const mp_robot* robot_ = nullptr;
    std::vector<std::string> link_names_;
    int num_spheres_ = 0;
    float* d_poses_ = nullptr;
    float* d_centres_ = nullptr;
    bool fk_current_ = false;
    void upload();
    void free_fk();
    friend class SingleArmSpace;
};
typedef std::shared_ptr<SingleArmStates> SingleArmStatesPtr;

class SingleArmSpace;

class SingleArmStateManager : public BaseStateManager {
  public:
    explicit SingleArmStateManager(SingleArmSpace* space) : space_(space) {}
    void clear() override;
    std::vector<int> add_states(const BaseStatesPtr& states) override;
    int find_k_nearest_neighbors(int k, const BaseStatesPtr& query_states,
                                 const std::vector<std::vector<int>>& group_indexs,
                                 std::vector<std::vector<int>>& neighbors_index) override;
    void find_the_nearest_neighbors(const BaseStatesPtr& query_states, const std::vector<std::vector<int>>& group_indexs,
                                    std::vector<std::vector<int>>& neighbors_index) override;
    BaseStatesPtr get_states(const std::vector<int>& states_index) override;
    BaseStatesPtr concatinate_states(const std::vector<BaseStatesPtr>& states) override;
    void interpolateToStates(const BaseStatesPtr& states, std::vector<int>& nearest_neighbors_indexes,
                             float max_distance) override;

  private:
    SingleArmSpace* space_;  // the space that created the manager; it must outlive the manager
    std::vector<std::vector<float>> q_;
};
typedef std::shared_ptr<SingleArmStateManager> SingleArmStateManagerPtr;

}  // namespace CUDAMPLib
