// CUDAMPLib states and state manager base classes (cudampl/CUDAMPLib/include/base/States.h), over the MI355X C ABI (include/mp_abi.h).
#pragma once
#include <memory>
#include <vector>

namespace CUDAMPLib {

class BaseStates {
  public:
    virtual ~BaseStates() {}
    int getNumOfStates() const { return num_of_states_; }
    bool isValid() const { return valid_; }
    // total cost per state after checkStates(states) (BaseStates::getTotalCostsHost)
    std::vector<float> getTotalCostsHost() const { return total_costs_; }
    // the cost of each constraint of the space, in the space's constraint order, per state:
    // [state][constraint] (BaseStates::getCostsHost, States.cu:65-78), after checkStates(states)
    std::vector<std::vector<float>> getCostsHost() const { return costs_; }

  protected:
    int num_of_states_ = 0;
    bool valid_ = true;
    std::vector<float> total_costs_;
    std::vector<std::vector<float>> costs_;
    friend class SingleArmSpace;
};
typedef std::shared_ptr<BaseStates> BaseStatesPtr;

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

}  // namespace CUDAMPLib
