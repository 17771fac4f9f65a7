// CUDAMPLib planner API (cudampl/CUDAMPLib/include/{base,spaces,states,constraints,tasks,
// termination,planners}) re-implemented over the MI355X C ABI.  Class names, constructor
// arguments and the methods the reference's callers use keep their reference signatures; the
// per-constraint CUDA kernels, NVRTC and the host-side graph code are replaced by the fused HIP
// kernels behind include/mp_abi.h.  The reference include paths (<spaces/SingleArmSpace.h>, ...)
// forward to this header.
#pragma once
#include <chrono>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../mp_eigen.h"

struct mp_robot;
struct mp_scene;

namespace CUDAMPLib {

// ------------------------------------------------------------------------------- constraints
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

// EnvConstraintSphere(name, pos[N][3], radius[N])  (EnvConstraintSphere.cu:10-33)
class EnvConstraintSphere : public BaseConstraint {
  public:
    EnvConstraintSphere(const std::string& name, const std::vector<std::vector<float>>& pos,
                        const std::vector<float>& radius)
        : BaseConstraint(name, false), pos_(pos), radius_(radius) {}
    std::vector<std::vector<float>> pos_;
    std::vector<float> radius_;
};
// EnvConstraintCuboid(name, pos[N][3], orientation R[N][9] row major, max[N][3], min[N][3])
class EnvConstraintCuboid : public BaseConstraint {
  public:
    EnvConstraintCuboid(const std::string& name, const std::vector<std::vector<float>>& pos,
                        const std::vector<std::vector<float>>& orientation, const std::vector<std::vector<float>>& max,
                        const std::vector<std::vector<float>>& min)
        : BaseConstraint(name, false), pos_(pos), rot_(orientation), max_(max), min_(min) {}
    std::vector<std::vector<float>> pos_, rot_, max_, min_;
};
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
// SelfCollisionConstraint(name, sphere_map, radius, enables[L][L])
class SelfCollisionConstraint : public BaseConstraint {
  public:
    SelfCollisionConstraint(const std::string& name, const std::vector<int>& sphere_map,
                            const std::vector<float>& radius, const std::vector<std::vector<bool>>& enables)
        : BaseConstraint(name, false), enables_(enables) {
        (void)sphere_map;
        (void)radius;
    }
    std::vector<std::vector<bool>> enables_;
};
// BoundaryConstraint(name, lower, upper, active)
class BoundaryConstraint : public BaseConstraint {
  public:
    BoundaryConstraint(const std::string& name, const std::vector<float>& lower, const std::vector<float>& upper,
                       const std::vector<bool>& active)
        : BaseConstraint(name, false), lower_(lower), upper_(upper), active_(active) {}
    std::vector<float> lower_, upper_;
    std::vector<bool> active_;
};
typedef std::shared_ptr<EnvConstraintSphere> EnvConstraintSpherePtr;
typedef std::shared_ptr<EnvConstraintCuboid> EnvConstraintCuboidPtr;
typedef std::shared_ptr<EnvConstraintCylinder> EnvConstraintCylinderPtr;
typedef std::shared_ptr<SelfCollisionConstraint> SelfCollisionConstraintPtr;
typedef std::shared_ptr<BoundaryConstraint> BoundaryConstraintPtr;

// ------------------------------------------------------------------------------------ states
class BaseStates {
  public:
    virtual ~BaseStates() {}
    int getNumOfStates() const { return num_of_states_; }
    bool isValid() const { return valid_; }
    // total cost per state after checkStates(states) (BaseStates::getTotalCostsHost)
    std::vector<float> getTotalCostsHost() const { return total_costs_; }

  protected:
    int num_of_states_ = 0;
    bool valid_ = true;
    std::vector<float> total_costs_;
    friend class SingleArmSpace;
};
typedef std::shared_ptr<BaseStates> BaseStatesPtr;

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

// Node store of the planners (States.h:218-283, SingleArmStates.h:185-221): joint values of
// every state added, addressed by the index add_states returned.  Nearest neighbours run on the
// device (mp_nn_argmin over the group's states, ties to the earlier member of the group).
class BaseStateManager {
  public:
    virtual ~BaseStateManager() {}
    virtual void clear() = 0;
    virtual std::vector<int> add_states(const BaseStatesPtr& states) = 0;
    int get_num_of_states() const { return num_of_states_; }
    // the k nearest members of each group for each query (ascending distance; fewer when a
    // group is smaller), returns k
    virtual int find_k_nearest_neighbors(int k, const BaseStatesPtr& query_states,
                                         const std::vector<std::vector<int>>& group_indexs,
                                         std::vector<std::vector<int>>& neighbors_index) = 0;
    // neighbors_index[query][group] = the nearest member of group_indexs[group]
    virtual void find_the_nearest_neighbors(const BaseStatesPtr& query_states,
                                            const std::vector<std::vector<int>>& group_indexs,
                                            std::vector<std::vector<int>>& neighbors_index) = 0;
    virtual BaseStatesPtr get_states(const std::vector<int>& states_index) = 0;
    virtual BaseStatesPtr concatinate_states(const std::vector<BaseStatesPtr>& states) = 0;
    // states[i] <- within max_distance of stored state nearest_neighbors_indexes[i]
    virtual void interpolateToStates(const BaseStatesPtr& states, std::vector<int>& nearest_neighbors_indexes,
                                     float max_distance) = 0;

  protected:
    int num_of_states_ = 0;
};
typedef std::shared_ptr<BaseStateManager> BaseStateManagerPtr;

// ------------------------------------------------------------------------------------ spaces
class BaseSpace {
  public:
    virtual ~BaseSpace() {}
    virtual BaseStatesPtr sample(int num_of_config) = 0;
    virtual void checkStates(const BaseStatesPtr& states, std::vector<bool>& state_feasibility) = 0;
    virtual void checkStates(const BaseStatesPtr& states) = 0;
    virtual bool checkMotions(const BaseStatesPtr& states1, const BaseStatesPtr& states2,
                              std::vector<bool>& motion_feasibility, std::vector<float>& motion_costs) = 0;
    virtual void interpolate(const BaseStatesPtr& from_states, const BaseStatesPtr& to_states, float max_distance) = 0;
    virtual BaseStatesPtr getPathFromWaypoints(const BaseStatesPtr& waypoints) = 0;
    virtual BaseStateManagerPtr createStateManager() = 0;
    size_t getDim() const { return dim_; }

  protected:
    size_t dim_ = 0;
};
typedef std::shared_ptr<BaseSpace> BaseSpacePtr;

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
};
typedef std::shared_ptr<SingleArmSpace> SingleArmSpacePtr;

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

// --------------------------------------------------------------------------------- util
std::vector<std::vector<float>> interpolateVectors(const std::vector<float>& v1, const std::vector<float>& v2,
                                                   float resolution);
// vectors of vec without those equal (componentwise within tol) to an earlier kept one
// (util.cpp:250-267)
std::vector<std::vector<float>> removeRedundantVectors(const std::vector<std::vector<float>>& vec, float tol);

// ------------------------------------------------------------------------------------- tasks
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

// ------------------------------------------------------------------------------- termination
class BaseTermination {
  public:
    virtual ~BaseTermination() {}
    virtual void reset() = 0;
    virtual bool checkTerminationCondition() = 0;
    virtual void printTerminationReason() = 0;
};
typedef std::shared_ptr<BaseTermination> BaseTerminationPtr;

// StepTermination.cpp:1-13: terminate when the step count exceeds the limit
class StepTermination : public BaseTermination {
  public:
    StepTermination(int step_limit) : step_limit_(step_limit) {}
    void reset() override { step_count_ = 0; }
    bool checkTerminationCondition() override { return ++step_count_ > step_limit_; }
    void printTerminationReason() override { std::printf("Step limit reached\n"); }

  private:
    int step_limit_, step_count_ = 0;
};
// TimeoutTermination.cpp:1-15: elapsed ms >= timeout * 1000
class TimeoutTermination : public BaseTermination {
  public:
    TimeoutTermination(float timeout_seconds) : timeout_ms_((long long)(timeout_seconds * 1000.0f)) {}
    void reset() override { t0_ = std::chrono::high_resolution_clock::now(); }
    bool checkTerminationCondition() override {
        auto dt = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::high_resolution_clock::now() - t0_);
        return dt.count() >= timeout_ms_;
    }
    void printTerminationReason() override { std::printf("Timeout reached\n"); }

  private:
    long long timeout_ms_;
    std::chrono::high_resolution_clock::time_point t0_ = std::chrono::high_resolution_clock::now();
};
typedef std::shared_ptr<StepTermination> StepTerminationPtr;
typedef std::shared_ptr<TimeoutTermination> TimeoutTerminationPtr;

// ---------------------------------------------------------------------------------- planners
class BasePlanner {
  public:
    BasePlanner(BaseSpacePtr space) { space_ = space; }
    virtual ~BasePlanner() {}
    virtual void setMotionTask(BaseTaskPtr task, bool get_full_path) = 0;
    virtual void solve(BaseTerminationPtr termination_condition) = 0;
    BaseSpacePtr getSpace() { return space_; }

  protected:
    BaseSpacePtr space_;
    bool get_full_path_ = true;
};
typedef std::shared_ptr<BasePlanner> BasePlannerPtr;

// Single-kernel parallel RRT-Connect (cRRTC.cu / CPRRTC semantics).  Each termination step runs
// one GPU solve of up to max_iterations samples; solve() stops at the first success or when the
// termination condition fires (failure reason "MeetTerminationCondition").
class cRRTC : public BasePlanner {
  public:
    cRRTC(BaseSpacePtr space) : BasePlanner(space) {}
    void setMotionTask(BaseTaskPtr task, bool get_full_path = true) override {
        task_ = task;
        get_full_path_ = get_full_path;
    }
    void solve(BaseTerminationPtr termination_condition) override;
    void setMaxIterations(int n) { max_iterations_ = n; }

  protected:
    BaseTaskPtr task_;
    int max_iterations_ = 10000;
};
// RRG (RRG.h:12-130, RRG.cpp:9-544): batched extend-connect roadmap between a start group and a
// goal group.  Each iteration samples k states, steers them toward their nearest node of one
// group (the groups alternate), keeps the free ones that reach the nearest node of either group
// by a free motion, and adds them with those edges; a state that reaches both groups closes a
// path, which Dijkstra recovers.  A side whose samples reach none of its nodes halves its travel
// distance (floor 0.1) and goes again.  Sampling, nearest neighbours, steering and the state and
// motion checks run on the device in batches; the graph (a few thousand vertices) is host-side.
class RRG : public BasePlanner {
  public:
    RRG(BaseSpacePtr space) : BasePlanner(space) {}
    void setMotionTask(BaseTaskPtr task, bool get_full_path = true) override;
    void solve(BaseTerminationPtr termination_condition) override;
    void setMaxTravelDistance(float d) { max_travel_distance_ = d; }
    void setSampleAttemptsInEachIteration(int n) { sample_attempts_ = n; }
    // states of the start group and of the goal group (RRG.cpp:531-544)
    void getStartAndGoalGroupStates(BaseStatesPtr& start_group_states, BaseStatesPtr& goal_group_states);
    // the roadmap's states (start states, goal states, then added states in order of addition)
    BaseStateManagerPtr getStateManager();

  private:
    struct Edge {
        int to;
        float weight;
    };
    BaseTaskPtr task_;
    float max_travel_distance_ = 0.5f;
    int sample_attempts_ = 30;
    std::vector<std::vector<float>> node_;  // joint values per vertex (empty for the two roots)
    std::vector<std::vector<Edge>> adj_;
    std::vector<int> members_[3];           // vertices of group 1 (start) and 2 (goal), in order
    int start_root_ = -1, goal_root_ = -1;
    int add_vertex(const std::vector<float>& q);
    void add_edge(int a, int b, float weight);
    std::vector<int> shortest_path() const;
};
typedef std::shared_ptr<cRRTC> cRRTCPtr;
typedef std::shared_ptr<RRG> RRGPtr;

}  // namespace CUDAMPLib
