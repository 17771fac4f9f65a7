// CUDAMPLib RRG planner (cudampl/CUDAMPLib/include/planners/RRG.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>
#include <vector>

#include "../base/Planner.h"
#include "../base/States.h"

namespace CUDAMPLib {

// RRG (RRG.h:12-130, RRG.cpp:9-544): batched extend-connect roadmap between a start group and a
// goal group.  Each iteration samples k states, steers them toward their nearest node of one
// group (the groups alternate), keeps the free ones that reach the nearest node of either group
// by a free motion, and adds them with those edges; a state that reaches both groups closes a
// path, which Dijkstra recovers.  A side whose samples reach none of its nodes halves its travel
// distance (floor 0.1) and goes again.  The loop runs on the device (mp_roadmap_*): the roadmap
// stays in device memory and each pass is three launches with no host round trip; the graph is
// read back once at the end, where Dijkstra recovers the path.
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
typedef std::shared_ptr<RRG> RRGPtr;

}  // namespace CUDAMPLib
