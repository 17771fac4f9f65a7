// CUDAMPLib API over the C ABI (reference: cudampl/CUDAMPLib/src).
#include <cmath>
#include <cstring>
#include <iostream>
#include <stdexcept>

#include "CUDAMPLib/mp_cudampl.h"
#include "mp_abi.h"
#include "mp_cpp_util.h"

namespace CUDAMPLib {

namespace {
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { mpcpp::check(mp_device_alloc(&p, bytes ? bytes : 1), "mp_device_alloc"); }
    ~DevBuf() {
        if (p) mp_device_free(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

std::vector<mpx::Vector3d> to_vec3(const std::vector<std::vector<float>>& v) {
    std::vector<mpx::Vector3d> out;
    for (const auto& p : v) {
        if (p.size() != 3) throw std::runtime_error("sphere positions must have 3 components");
        mpx::Vector3d x;
        x(0) = p[0];
        x(1) = p[1];
        x(2) = p[2];
        out.push_back(x);
    }
    return out;
}

SingleArmStatesPtr as_states(const BaseStatesPtr& s) {
    auto p = std::dynamic_pointer_cast<SingleArmStates>(s);
    if (!p) throw std::runtime_error("states are not SingleArmStates");
    return p;
}
}  // namespace

// ------------------------------------------------------------------------------------ states
SingleArmStates::SingleArmStates(const std::vector<std::vector<float>>& active_values, const std::vector<bool>& active_map,
                                 const std::vector<float>& defaults)
    : q_(active_values), active_(active_map), defaults_(defaults) {
    num_of_states_ = (int)q_.size();
    upload();
}

SingleArmStates::~SingleArmStates() {
    if (d_q_) mp_device_free(d_q_);
}

void SingleArmStates::upload() {
    if (d_q_) mp_device_free(d_q_);
    d_q_ = nullptr;
    size_t dof = q_.empty() ? 0 : q_[0].size();
    std::vector<float> flat;
    flat.reserve(q_.size() * dof);
    for (const auto& r : q_) {
        if (r.size() != dof) throw std::runtime_error("states of different dimensions");
        flat.insert(flat.end(), r.begin(), r.end());
    }
    void* p = nullptr;
    if (mp_device_alloc(&p, std::max<size_t>(flat.size() * sizeof(float), 4)) != MP_OK) {
        valid_ = false;  // reference: isValid() == false on allocation failure
        return;
    }
    d_q_ = (float*)p;
    if (!flat.empty()) mpcpp::check(mp_memcpy_h2d(d_q_, flat.data(), flat.size() * sizeof(float), nullptr), "h2d");
}

std::vector<std::vector<float>> SingleArmStates::getJointStatesFullHost() const {
    std::vector<std::vector<float>> out;
    for (const auto& r : q_) {
        std::vector<float> f(active_.size());
        size_t k = 0;
        for (size_t j = 0; j < active_.size(); j++) f[j] = active_[j] ? r[k++] : defaults_[j];
        out.push_back(f);
    }
    return out;
}

void SingleArmStates::filterStates(const std::vector<bool>& keep) {
    if (keep.size() != q_.size()) throw std::runtime_error("filterStates: size mismatch");
    std::vector<std::vector<float>> q;
    for (size_t i = 0; i < q_.size(); i++)
        if (keep[i]) q.push_back(q_[i]);
    q_.swap(q);
    num_of_states_ = (int)q_.size();
    upload();
}

// ------------------------------------------------------------------------------------ space
SingleArmSpace::SingleArmSpace(size_t dim, const std::vector<BaseConstraintPtr>& constraints,
                               const std::vector<int>& joint_types, const std::vector<mpx::Isometry3d>& joint_poses,
                               const std::vector<mpx::Vector3d>& joint_axes, const std::vector<int>& link_parent_link_maps,
                               const std::vector<int>& sphere_map, const std::vector<std::vector<float>>& sphere_pos,
                               const std::vector<float>& sphere_radius, const std::vector<bool>& active_joint_map,
                               const std::vector<float>& lower, const std::vector<float>& upper,
                               const std::vector<float>& default_joint_values,
                               const std::vector<std::string>& link_names, float resolution)
    : lower_bound(lower), upper_bound(upper), defaults_(default_joint_values), active_(active_joint_map),
      resolution_(resolution) {
    dim_ = dim;
    mpcpp::ensure_init();
    // self-collision pairs come from a SelfCollisionConstraint; without one no pair is checked
    std::vector<std::vector<bool>> enables(joint_types.size(), std::vector<bool>(joint_types.size(), false));
    std::vector<float> s, b, c;
    for (const auto& con : constraints) {
        if (auto sc = std::dynamic_pointer_cast<SelfCollisionConstraint>(con)) enables = sc->enables_;
        if (auto e = std::dynamic_pointer_cast<EnvConstraintSphere>(con)) {
            for (size_t i = 0; i < e->pos_.size(); i++) s.insert(s.end(), {e->pos_[i][0], e->pos_[i][1], e->pos_[i][2], e->radius_[i]});
        }
        if (auto e = std::dynamic_pointer_cast<EnvConstraintCuboid>(con)) {
            for (size_t i = 0; i < e->pos_.size(); i++) {
                b.insert(b.end(), e->pos_[i].begin(), e->pos_[i].end());
                b.insert(b.end(), e->rot_[i].begin(), e->rot_[i].end());
                b.insert(b.end(), e->min_[i].begin(), e->min_[i].end());
                b.insert(b.end(), e->max_[i].begin(), e->max_[i].end());
            }
        }
        if (auto e = std::dynamic_pointer_cast<EnvConstraintCylinder>(con)) {
            for (size_t i = 0; i < e->pos_.size(); i++) {
                c.insert(c.end(), e->pos_[i].begin(), e->pos_[i].end());
                c.insert(c.end(), e->rot_[i].begin(), e->rot_[i].end());
                c.insert(c.end(), {e->radius_[i], e->height_[i]});
            }
        }
    }
    if (b.size() % 18 || c.size() % 14) throw std::runtime_error("malformed obstacle arrays");
    mpcpp::RobotArrays a("space", joint_types, joint_poses, joint_axes, link_parent_link_maps, sphere_map,
                         to_vec3(sphere_pos), sphere_radius, active_joint_map, lower, upper, default_joint_values,
                         link_names.size(), enables);
    if (a.dof != (int)dim) throw std::runtime_error("SingleArmSpace: dim does not match the active joints");
    mpcpp::check(mp_robot_create(&a.desc, &robot_), "mp_robot_create");
    mpcpp::check(mp_scene_create(s.data(), (int)(s.size() / 4), b.data(), (int)(b.size() / 18), c.data(),
                                 (int)(c.size() / 14), nullptr, 0, &scene_),
                 "mp_scene_create");
}

SingleArmSpace::~SingleArmSpace() {
    if (scene_) mp_scene_destroy(scene_);
    if (robot_) mp_robot_destroy(robot_);
}

BaseStatesPtr SingleArmSpace::createStatesFromVector(const std::vector<std::vector<float>>& joint_values) {
    for (const auto& r : joint_values)
        if (r.size() != dim_) throw std::runtime_error("createStatesFromVector: expected active joint values");
    return std::make_shared<SingleArmStates>(joint_values, active_, defaults_);
}

BaseStatesPtr SingleArmSpace::createStatesFromVectorFull(const std::vector<std::vector<float>>& joint_values) {
    std::vector<std::vector<float>> act;
    for (const auto& r : joint_values) {
        if (r.size() != active_.size()) throw std::runtime_error("createStatesFromVectorFull: expected full joint values");
        std::vector<float> a;
        for (size_t j = 0; j < r.size(); j++)
            if (active_[j]) a.push_back(r[j]);
        act.push_back(a);
    }
    return createStatesFromVector(act);
}

std::vector<std::vector<float>> SingleArmSpace::getJointVectorInActiveJointsFromStates(const BaseStatesPtr& states) {
    return as_states(states)->getJointStatesHost();
}

BaseStatesPtr SingleArmSpace::sample(int n) {
    if (n < 0) throw std::runtime_error("sample: negative count");
    DevBuf d((size_t)n * dim_ * sizeof(float));
    mpcpp::check(mp_sample(robot_, seed_, sample_offset_, n, (float*)d.p, nullptr), "mp_sample");
    sample_offset_ += (unsigned long long)n;
    std::vector<float> h((size_t)n * dim_);
    if (n) mpcpp::check(mp_memcpy_d2h(h.data(), d.p, h.size() * sizeof(float), nullptr), "d2h");
    std::vector<std::vector<float>> q(n);
    for (int i = 0; i < n; i++) q[i].assign(h.begin() + i * dim_, h.begin() + (i + 1) * dim_);
    return createStatesFromVector(q);
}

void SingleArmSpace::checkStates(const BaseStatesPtr& states, std::vector<bool>& feas) {
    auto st = as_states(states);
    int n = st->getNumOfStates();
    feas.assign(n, false);
    if (n == 0) return;
    DevBuf v(n);
    mpcpp::check(mp_check_states(robot_, scene_, st->getJointStatesCuda(), n, (uint8_t*)v.p, nullptr, nullptr),
                 "mp_check_states");
    std::vector<uint8_t> h(n);
    mpcpp::check(mp_memcpy_d2h(h.data(), v.p, n, nullptr), "d2h");
    for (int i = 0; i < n; i++) feas[i] = h[i] != 0;
}

void SingleArmSpace::checkStates(const BaseStatesPtr& states) {
    auto st = as_states(states);
    int n = st->getNumOfStates();
    st->total_costs_.assign(n, 0.0f);
    if (n == 0) return;
    DevBuf c((size_t)n * sizeof(float));
    mpcpp::check(mp_state_costs(robot_, scene_, st->getJointStatesCuda(), n, (float*)c.p, nullptr, nullptr),
                 "mp_state_costs");
    mpcpp::check(mp_memcpy_d2h(st->total_costs_.data(), c.p, n * sizeof(float), nullptr), "d2h");
}

bool SingleArmSpace::checkMotions(const BaseStatesPtr& states1, const BaseStatesPtr& states2, std::vector<bool>& feas,
                                  std::vector<float>& costs) {
    auto a = as_states(states1), b = as_states(states2);
    int m = a->getNumOfStates();
    if (m != b->getNumOfStates()) throw std::runtime_error("Number of states in states1 and states2 are not equal");
    if (m == 0) throw std::runtime_error("No states to check");
    DevBuf ok(m), len((size_t)m * sizeof(float));
    mpcpp::check(mp_check_edges(robot_, scene_, a->getJointStatesCuda(), b->getJointStatesCuda(), m, resolution_,
                                (uint8_t*)ok.p, (float*)len.p, nullptr),
                 "mp_check_edges");
    std::vector<uint8_t> h(m);
    costs.assign(m, 0.0f);
    mpcpp::check(mp_memcpy_d2h(h.data(), ok.p, m, nullptr), "d2h");
    mpcpp::check(mp_memcpy_d2h(costs.data(), len.p, m * sizeof(float), nullptr), "d2h");
    feas.assign(m, false);
    for (int i = 0; i < m; i++) feas[i] = h[i] != 0;
    return true;
}

void SingleArmSpace::interpolate(const BaseStatesPtr& from_states, const BaseStatesPtr& to_states, float max_distance) {
    auto a = as_states(from_states), b = as_states(to_states);
    int n = a->getNumOfStates();
    if (n != b->getNumOfStates()) throw std::runtime_error("interpolate: size mismatch");
    if (n == 0) return;
    mpcpp::check(mp_steer(a->getJointStatesCuda(), b->getJointStatesCuda(), n, (int)dim_, max_distance, nullptr),
                 "mp_steer");
    std::vector<float> h((size_t)n * dim_);
    mpcpp::check(mp_memcpy_d2h(h.data(), b->getJointStatesCuda(), h.size() * sizeof(float), nullptr), "d2h");
    for (int i = 0; i < n; i++) b->q_[i].assign(h.begin() + i * dim_, h.begin() + (i + 1) * dim_);
}

BaseStatesPtr SingleArmSpace::getPathFromWaypoints(const BaseStatesPtr& waypoints) {
    // getUnconstrainedPathFromWaypoints (SingleArmSpace.cu:1428-1464)
    auto w = as_states(waypoints)->getJointStatesFullHost();
    std::vector<std::vector<float>> path;
    for (size_t i = 0; i + 1 < w.size(); i++) {
        auto seg = interpolateVectors(w[i], w[i + 1], resolution_);
        path.insert(path.end(), seg.begin(), seg.end());
    }
    return createStatesFromVectorFull(path);
}

// util.cpp:205-248
std::vector<std::vector<float>> interpolateVectors(const std::vector<float>& v1, const std::vector<float>& v2,
                                                   float resolution) {
    if (v1.size() != v2.size()) throw std::runtime_error("Vectors must be of the same size");
    if (v1 == v2) return {v1};
    std::vector<std::vector<float>> out{v1};
    float distance = 0.0f;
    for (size_t i = 0; i < v1.size(); ++i) distance += std::pow(v2[i] - v1[i], 2);
    distance = std::sqrt(distance);
    int num_steps = (int)std::floor(distance / resolution);
    std::vector<float> step(v1.size());
    for (size_t i = 0; i < v1.size(); ++i) step[i] = (v2[i] - v1[i]) / num_steps;
    for (int i = 1; i <= num_steps; ++i) {
        std::vector<float> x(v1.size());
        for (size_t j = 0; j < v1.size(); ++j) x[j] = v1[j] + i * step[j];
        out.push_back(x);
    }
    out.push_back(v2);
    return out;
}

// ---------------------------------------------------------------------------------- planners
namespace {
std::vector<std::vector<float>> densify(const std::vector<std::vector<float>>& path, float res) {
    std::vector<std::vector<float>> out;
    for (size_t i = 0; i + 1 < path.size(); i++) {
        auto seg = interpolateVectors(path[i], path[i + 1], res);
        out.insert(out.end(), seg.begin(), seg.end());
    }
    return out;
}
}  // namespace

void cRRTC::solve(BaseTerminationPtr termination) {
    auto space = std::dynamic_pointer_cast<SingleArmSpace>(space_);
    if (!space || !task_) throw std::runtime_error("cRRTC: space or task missing");
    termination->reset();
    auto starts = task_->getStartStatesVector(), goals = task_->getGoalStatesVector();
    if (starts.empty() || goals.empty()) {
        task_->setFailureReason("InvalidInput");
        return;
    }
    std::vector<float> s = starts[0], g = goals[0];
    mp_rrtc_params p;
    mp_rrtc_default_params(&p);
    p.resolution = space->getResolution();
    p.max_iterations = max_iterations_;
    std::vector<float> wp((size_t)p.max_path * s.size());
    for (unsigned long long attempt = 0; !termination->checkTerminationCondition(); attempt++) {
        p.seed = attempt;
        mp_path out{};
        out.waypoints = wp.data();
        out.capacity = p.max_path;
        mpcpp::check(mp_rrtc_solve(space->robot(), space->scene(), s.data(), g.data(), &p, &out), "mp_rrtc_solve");
        if (out.status == MP_PLAN_INVALID) {
            task_->setFailureReason("InvalidInput");
            return;
        }
        if (out.status == MP_PLAN_SOLVED) {
            std::vector<std::vector<float>> path;
            for (int k = 0; k < out.num_waypoints; k++) path.emplace_back(wp.begin() + k * s.size(), wp.begin() + (k + 1) * s.size());
            task_->setSolution(get_full_path_ ? densify(path, space->getResolution()) : path);
            return;
        }
    }
    task_->setFailureReason("MeetTerminationCondition");
}

void RRG::solve(BaseTerminationPtr termination) {
    // RRG.cpp:157-232: reject invalid start/goal, accept the direct motion when it is free
    auto space = std::dynamic_pointer_cast<SingleArmSpace>(space_);
    if (!space || !task_) throw std::runtime_error("RRG: space or task missing");
    auto starts = task_->getStartStatesVector(), goals = task_->getGoalStatesVector();
    if (starts.empty() || goals.empty()) {
        task_->setFailureReason("InvalidInput");
        return;
    }
    std::vector<bool> f1, f2;
    space->checkStates(space->createStatesFromVector({starts[0]}), f1);
    space->checkStates(space->createStatesFromVector({goals[0]}), f2);
    if (!f1[0] || !f2[0]) {
        task_->setFailureReason("InvalidInput");
        return;
    }
    std::vector<bool> mf;
    std::vector<float> mc;
    space->checkMotions(space->createStatesFromVector({starts[0]}), space->createStatesFromVector({goals[0]}), mf, mc);
    if (mf[0]) {
        std::vector<std::vector<float>> path{starts[0], goals[0]};
        task_->setSolution(get_full_path_ ? densify(path, space->getResolution()) : path);
        return;
    }
    cRRTC::solve(termination);
}

}  // namespace CUDAMPLib
