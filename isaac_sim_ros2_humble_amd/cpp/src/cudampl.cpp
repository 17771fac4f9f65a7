// CUDAMPLib API over the C ABI (reference: cudampl/CUDAMPLib/src).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <iostream>
#include <limits>
#include <queue>
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
    free_fk();
}

void SingleArmStates::free_fk() {
    if (d_poses_) mp_device_free(d_poses_);
    if (d_centres_) mp_device_free(d_centres_);
    d_poses_ = d_centres_ = nullptr;
    fk_current_ = false;
}

void SingleArmStates::update() {
    if (!robot_) throw std::runtime_error("update: states are not bound to a space");
    const int n = num_of_states_, L = (int)link_names_.size();
    if (!d_poses_) mpcpp::check(mp_device_alloc((void**)&d_poses_, std::max<size_t>((size_t)n * L * 64, 4)), "alloc");
    if (!d_centres_)
        mpcpp::check(mp_device_alloc((void**)&d_centres_, std::max<size_t>((size_t)n * num_spheres_ * 12, 4)), "alloc");
    if (n > 0) {
        mpcpp::check(mp_fk_link_poses(robot_, d_q_, n, d_poses_, nullptr), "mp_fk_link_poses");
        mpcpp::check(mp_fk_spheres(robot_, d_q_, n, d_centres_, nullptr), "mp_fk_spheres");
        mpcpp::check(mp_stream_sync(nullptr), "sync");
    }
    fk_current_ = true;
}

float* SingleArmStates::getLinkPosesInBaseLinkCuda() {
    if (!fk_current_) update();
    return d_poses_;
}

float* SingleArmStates::getSelfCollisionSpheresPosInBaseLinkCuda() {
    if (!fk_current_) update();
    return d_centres_;
}

std::vector<std::vector<std::vector<float>>> SingleArmStates::getSelfCollisionSpheresPosInBaseLinkHost() {
    const float* d = getSelfCollisionSpheresPosInBaseLinkCuda();
    const int n = num_of_states_, S = num_spheres_;
    std::vector<float> h((size_t)n * S * 3);
    if (!h.empty()) mpcpp::check(mp_memcpy_d2h(h.data(), d, h.size() * sizeof(float), nullptr), "d2h");
    std::vector<std::vector<std::vector<float>>> out(n, std::vector<std::vector<float>>(S));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < S; k++) out[i][k].assign(h.begin() + ((size_t)i * S + k) * 3, h.begin() + ((size_t)i * S + k + 1) * 3);
    return out;
}

std::vector<std::vector<mpx::Isometry3d>> SingleArmStates::getLinkPosesInBaseLinkHost() {
    const float* d = getLinkPosesInBaseLinkCuda();
    const int n = num_of_states_, L = (int)link_names_.size();
    std::vector<float> h((size_t)n * L * 16);
    if (!h.empty()) mpcpp::check(mp_memcpy_d2h(h.data(), d, h.size() * sizeof(float), nullptr), "d2h");
    std::vector<std::vector<mpx::Isometry3d>> out(n, std::vector<mpx::Isometry3d>(L));
    for (int i = 0; i < n; i++)
        for (int l = 0; l < L; l++)
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++) out[i][l].matrix()(r, c) = h[((size_t)i * L + l) * 16 + 4 * r + c];
    return out;
}

std::vector<mpx::Isometry3d> SingleArmStates::getLinkPoseInBaseLinkHost(const std::string& link_name) {
    // SingleArmStates.cu getLinkPoseInBaseLinkHost: the named link's pose in every state
    const auto it = std::find(link_names_.begin(), link_names_.end(), link_name);
    if (it == link_names_.end()) throw std::runtime_error("getLinkPoseInBaseLinkHost: unknown link " + link_name);
    const int l = (int)(it - link_names_.begin());
    auto all = getLinkPosesInBaseLinkHost();
    std::vector<mpx::Isometry3d> out;
    for (auto& p : all) out.push_back(p[l]);
    return out;
}

void SingleArmStates::upload() {
    free_fk();
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
    int rdof = 0, rpairs = 0, rlinks = 0;
    mpcpp::check(mp_robot_info(robot_, &rdof, &num_spheres_, &rpairs, &rlinks), "mp_robot_info");
    link_names_ = link_names;
    link_names_.resize(rlinks);
    mpcpp::check(mp_scene_create(s.data(), (int)(s.size() / 4), b.data(), (int)(b.size() / 18), c.data(),
                                 (int)(c.size() / 14), nullptr, 0, &scene_),
                 "mp_scene_create");
    // joint bounds are checked only through a BoundaryConstraint of the space, with that
    // constraint's own bounds (SingleArmSpace.cu:1572-1602, BoundaryConstraint.cu:56-92)
    auto active_bounds = [&](const BoundaryConstraint& bc, std::vector<float>& lo, std::vector<float>& hi) {
        for (size_t j = 0; j < active_joint_map.size(); j++) {
            if (!active_joint_map[j]) continue;
            const bool on = j < bc.active_.size() && bc.active_[j] && j < bc.lower_.size() && j < bc.upper_.size();
            lo.push_back(on ? bc.lower_[j] : -INFINITY);
            hi.push_back(on ? bc.upper_[j] : INFINITY);
        }
    };
    BoundaryConstraintPtr bc;
    for (const auto& con : constraints)
        if (auto p = std::dynamic_pointer_cast<BoundaryConstraint>(con)) bc = p;
    if (!bc) {
        mpcpp::check(mp_scene_set_joint_bounds(scene_, MP_BOUNDS_NONE, nullptr, nullptr, 0), "mp_scene_set_joint_bounds");
    } else {
        std::vector<float> lo, hi;
        active_bounds(*bc, lo, hi);
        mpcpp::check(mp_scene_set_joint_bounds(scene_, MP_BOUNDS_CUSTOM, lo.data(), hi.data(), (int)lo.size()),
                     "mp_scene_set_joint_bounds");
    }
    // one scene per constraint for its cost row (getCostsHost): the constraint's own obstacles
    // or bounds, nothing else
    for (const auto& con : constraints) {
        std::vector<float> cs, cb, cc;
        int row = -1;
        BoundaryConstraintPtr cbnd;
        if (auto e = std::dynamic_pointer_cast<EnvConstraintSphere>(con)) {
            row = MP_COST_ENV_SPHERE;
            for (size_t i = 0; i < e->pos_.size(); i++) cs.insert(cs.end(), {e->pos_[i][0], e->pos_[i][1], e->pos_[i][2], e->radius_[i]});
        } else if (auto e = std::dynamic_pointer_cast<EnvConstraintCuboid>(con)) {
            row = MP_COST_ENV_CUBOID;
            for (size_t i = 0; i < e->pos_.size(); i++) {
                cb.insert(cb.end(), e->pos_[i].begin(), e->pos_[i].end());
                cb.insert(cb.end(), e->rot_[i].begin(), e->rot_[i].end());
                cb.insert(cb.end(), e->min_[i].begin(), e->min_[i].end());
                cb.insert(cb.end(), e->max_[i].begin(), e->max_[i].end());
            }
        } else if (auto e = std::dynamic_pointer_cast<EnvConstraintCylinder>(con)) {
            row = MP_COST_ENV_CYLINDER;
            for (size_t i = 0; i < e->pos_.size(); i++) {
                cc.insert(cc.end(), e->pos_[i].begin(), e->pos_[i].end());
                cc.insert(cc.end(), e->rot_[i].begin(), e->rot_[i].end());
                cc.insert(cc.end(), {e->radius_[i], e->height_[i]});
            }
        } else if (std::dynamic_pointer_cast<SelfCollisionConstraint>(con)) {
            row = MP_COST_SELF;
        } else if (auto p = std::dynamic_pointer_cast<BoundaryConstraint>(con)) {
            row = MP_COST_BOUNDS;
            cbnd = p;
        }
        mp_scene* cscene = nullptr;
        mpcpp::check(mp_scene_create(cs.data(), (int)(cs.size() / 4), cb.data(), (int)(cb.size() / 18), cc.data(),
                                     (int)(cc.size() / 14), nullptr, 0, &cscene),
                     "mp_scene_create");
        con_scenes_.push_back(cscene);
        con_rows_.push_back(row);
        if (cbnd) {
            std::vector<float> lo, hi;
            active_bounds(*cbnd, lo, hi);
            mpcpp::check(mp_scene_set_joint_bounds(cscene, MP_BOUNDS_CUSTOM, lo.data(), hi.data(), (int)lo.size()),
                         "mp_scene_set_joint_bounds");
        } else {
            mpcpp::check(mp_scene_set_joint_bounds(cscene, MP_BOUNDS_NONE, nullptr, nullptr, 0), "mp_scene_set_joint_bounds");
        }
    }
}

SingleArmSpace::~SingleArmSpace() {
    for (mp_scene* cs : con_scenes_)
        if (cs) mp_scene_destroy(cs);
    if (scene_) mp_scene_destroy(scene_);
    if (robot_) mp_robot_destroy(robot_);
}

BaseStatesPtr SingleArmSpace::createStatesFromVector(const std::vector<std::vector<float>>& joint_values) {
    for (const auto& r : joint_values)
        if (r.size() != dim_) throw std::runtime_error("createStatesFromVector: expected active joint values");
    auto st = std::make_shared<SingleArmStates>(joint_values, active_, defaults_);
    st->robot_ = robot_;
    st->link_names_ = link_names_;
    st->num_spheres_ = num_spheres_;
    return st;
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

std::string SingleArmSpace::generateFKKernelSourceCode() const {
    const size_t n = mp_robot_source(robot_, nullptr, 0);
    std::string src(n + 1, '\0');
    mp_robot_source(robot_, &src[0], n + 1);
    src.resize(n);
    return src;
}

void SingleArmSpace::sampleInto(float* d_q, int n) {
    if (n < 0) throw std::runtime_error("sample: negative count");
    mpcpp::check(mp_sample(robot_, seed_, sample_offset_, n, d_q, nullptr), "mp_sample");
    sample_offset_ += (unsigned long long)n;
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
    st->costs_.assign(n, std::vector<float>(con_rows_.size(), 0.0f));
    if (n == 0) return;
    DevBuf c((size_t)n * sizeof(float));
    mpcpp::check(mp_state_costs(robot_, scene_, st->getJointStatesCuda(), n, (float*)c.p, nullptr, nullptr),
                 "mp_state_costs");
    mpcpp::check(mp_memcpy_d2h(st->total_costs_.data(), c.p, n * sizeof(float), nullptr), "d2h");
    // each constraint's cost: its row of the per-kind costs on its own scene
    DevBuf parts((size_t)MP_COST_KINDS * n * sizeof(float));
    std::vector<float> row(n);
    for (size_t k = 0; k < con_rows_.size(); k++) {
        if (con_rows_[k] < 0) continue;
        mpcpp::check(mp_constraint_costs(robot_, con_scenes_[k], st->getJointStatesCuda(), n, (float*)parts.p, nullptr,
                                         nullptr, nullptr),
                     "mp_constraint_costs");
        mpcpp::check(mp_memcpy_d2h(row.data(), (const float*)parts.p + (size_t)con_rows_[k] * n, n * sizeof(float), nullptr),
                     "d2h");
        for (int i = 0; i < n; i++) st->costs_[i][k] = row[i];
    }
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

BaseStateManagerPtr SingleArmSpace::createStateManager() {
    return std::make_shared<SingleArmStateManager>(this);
}

// ----------------------------------------------------------------------------- state manager
void SingleArmStateManager::clear() {
    q_.clear();
    num_of_states_ = 0;
}

std::vector<int> SingleArmStateManager::add_states(const BaseStatesPtr& states) {
    auto st = as_states(states);
    std::vector<int> idx;
    for (const auto& q : st->getJointStatesHost()) {
        idx.push_back((int)q_.size());
        q_.push_back(q);
    }
    num_of_states_ = (int)q_.size();
    return idx;
}

BaseStatesPtr SingleArmStateManager::get_states(const std::vector<int>& states_index) {
    std::vector<std::vector<float>> q;
    for (int i : states_index) {
        if (i < 0 || i >= (int)q_.size()) throw std::runtime_error("get_states: index out of range");
        q.push_back(q_[i]);
    }
    return space_->createStatesFromVector(q);
}

BaseStatesPtr SingleArmStateManager::concatinate_states(const std::vector<BaseStatesPtr>& states) {
    std::vector<std::vector<float>> q;
    for (const auto& s : states) {
        auto h = as_states(s)->getJointStatesHost();
        q.insert(q.end(), h.begin(), h.end());
    }
    return space_->createStatesFromVector(q);
}

void SingleArmStateManager::find_the_nearest_neighbors(const BaseStatesPtr& query_states,
                                                       const std::vector<std::vector<int>>& group_indexs,
                                                       std::vector<std::vector<int>>& neighbors_index) {
    // SingleArmStates.cu:1265-1431: per query and group, argmin of the squared L2 distance over
    // the active joints; the group's members are gathered into a SoA tree for mp_nn_argmin
    auto qs = as_states(query_states);
    const int nq = qs->getNumOfStates();
    const size_t dof = space_->getDim();
    neighbors_index.assign(nq, std::vector<int>(group_indexs.size(), -1));
    if (nq == 0) return;
    DevBuf d_idx((size_t)nq * sizeof(int));
    std::vector<int> pos(nq);
    for (size_t g = 0; g < group_indexs.size(); g++) {
        const auto& members = group_indexs[g];
        if (members.empty()) throw std::runtime_error("find_the_nearest_neighbors: empty group");
        const int n = (int)members.size();
        std::vector<float> soa(dof * (size_t)n);
        for (int i = 0; i < n; i++) {
            if (members[i] < 0 || members[i] >= (int)q_.size()) throw std::runtime_error("group index out of range");
            for (size_t j = 0; j < dof; j++) soa[j * n + i] = q_[members[i]][j];
        }
        DevBuf d_tree(soa.size() * sizeof(float));
        mpcpp::check(mp_memcpy_h2d(d_tree.p, soa.data(), soa.size() * sizeof(float), nullptr), "h2d");
        mpcpp::check(mp_nn_argmin((const float*)d_tree.p, n, n, (int)dof, qs->getJointStatesCuda(), nq, (int*)d_idx.p,
                                  nullptr),
                     "mp_nn_argmin");
        mpcpp::check(mp_memcpy_d2h(pos.data(), d_idx.p, nq * sizeof(int), nullptr), "d2h");
        for (int i = 0; i < nq; i++) neighbors_index[i][g] = members[pos[i]];
    }
}

int SingleArmStateManager::find_k_nearest_neighbors(int k, const BaseStatesPtr& query_states,
                                                    const std::vector<std::vector<int>>& group_indexs,
                                                    std::vector<std::vector<int>>& neighbors_index) {
    // unused by the planners (States.h:238-245); host selection by (distance, group order)
    auto qs = as_states(query_states)->getJointStatesHost();
    neighbors_index.assign(qs.size(), {});
    for (size_t i = 0; i < qs.size(); i++)
        for (const auto& members : group_indexs) {
            std::vector<std::pair<float, int>> d;
            for (size_t m = 0; m < members.size(); m++) {
                float acc = 0.0f;
                for (size_t j = 0; j < qs[i].size(); j++) {
                    const float df = qs[i][j] - q_[members[m]][j];
                    acc = std::fma(df, df, acc);
                }
                d.push_back({acc, (int)m});
            }
            const size_t kk = std::min(d.size(), (size_t)std::max(k, 0));
            std::partial_sort(d.begin(), d.begin() + kk, d.end());
            for (size_t r = 0; r < kk; r++) neighbors_index[i].push_back(members[d[r].second]);
        }
    return k;
}

void SingleArmStateManager::interpolateToStates(const BaseStatesPtr& states, std::vector<int>& nearest_neighbors_indexes,
                                                float max_distance) {
    // SingleArmStates.cu:1488-1557: pull each state to within max_distance of its neighbour
    auto from = get_states(nearest_neighbors_indexes);
    if (from->getNumOfStates() != states->getNumOfStates()) throw std::runtime_error("interpolateToStates: size mismatch");
    space_->interpolate(from, states, max_distance);
}

// The densified straight segment a -> b (the segment densifier of
// getUnconstrainedPathFromWaypoints, SingleArmSpace.cu:1428-1464 / util.cpp:205-248): a, then the
// k = floor(|b - a| / resolution) points a + i (b - a) / k for i = 1 .. k (the last one b up to
// rounding), then b itself, so both ends stay even where they repeat; just a when a == b.
std::vector<std::vector<float>> interpolateVectors(const std::vector<float>& a, const std::vector<float>& b,
                                                   float resolution) {
    if (a.size() != b.size()) throw std::runtime_error("interpolateVectors: the two states differ in dimension");
    if (a == b) return {a};
    const size_t n = a.size();
    std::vector<float> inc(n);
    float len2 = 0.0f;
    for (size_t j = 0; j < n; ++j) {
        const float d = b[j] - a[j];
        len2 += d * d;
    }
    const int k = (int)std::floor(std::sqrt(len2) / resolution);
    for (size_t j = 0; j < n; ++j) inc[j] = k > 0 ? (b[j] - a[j]) / (float)k : 0.0f;
    std::vector<std::vector<float>> pts;
    pts.reserve((size_t)k + 2);
    pts.push_back(a);
    for (int i = 1; i <= k; ++i) {
        pts.emplace_back(n);
        for (size_t j = 0; j < n; ++j) pts.back()[j] = a[j] + (float)i * inc[j];
    }
    pts.push_back(b);
    return pts;
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

std::vector<std::vector<float>> removeRedundantVectors(const std::vector<std::vector<float>>& vec, float tol) {
    std::vector<std::vector<float>> unique;
    for (const auto& v : vec) {
        bool dup = false;
        for (const auto& u : unique) {
            bool same = u.size() == v.size();
            for (size_t j = 0; same && j < v.size(); j++) same = std::fabs(v[j] - u[j]) <= tol;
            if (same) {
                dup = true;
                break;
            }
        }
        if (!dup) unique.push_back(v);
    }
    return unique;
}

SingleArmTask::SingleArmTask(const std::vector<std::vector<float>>& start_joint_values, const BaseSpacePtr& goal_region,
                             int goal_config_sampling_attempts)
    : start_(start_joint_values) {
    auto space = std::dynamic_pointer_cast<SingleArmSpace>(goal_region);
    if (!space) throw std::runtime_error("SingleArmTask: the goal region must be a SingleArmSpace");
    auto samples = space->sample(goal_config_sampling_attempts);
    std::vector<bool> feas;
    space->checkStates(samples, feas);
    auto st = std::dynamic_pointer_cast<SingleArmStates>(samples);
    st->filterStates(feas);
    goal_ = removeRedundantVectors(st->getJointStatesHost(), 1e-5f);
}


int RRG::add_vertex(const std::vector<float>& q) {
    node_.push_back(q);
    adj_.emplace_back();
    return (int)node_.size() - 1;
}

void RRG::add_edge(int a, int b, float weight) {
    adj_[a].push_back({b, weight});
    adj_[b].push_back({a, weight});
}

void RRG::setMotionTask(BaseTaskPtr task, bool get_full_path) {
    // RRG.cpp:30-94: one root per group, joined by zero-weight edges to each start / goal state
    task_ = task;
    get_full_path_ = get_full_path;
    node_.clear();
    adj_.clear();
    for (auto& m : members_) m.clear();
    start_root_ = add_vertex({});
    for (const auto& q : task->getStartStatesVector()) {
        const int v = add_vertex(q);
        add_edge(v, start_root_, 0.0f);
        members_[1].push_back(v);
    }
    goal_root_ = add_vertex({});
    for (const auto& q : task->getGoalStatesVector()) {
        const int v = add_vertex(q);
        add_edge(v, goal_root_, 0.0f);
        members_[2].push_back(v);
    }
}

std::vector<int> RRG::shortest_path() const {
    // Dijkstra from the start root (RRG.cpp:454-500); vertices from start root to goal root
    const int nv = (int)node_.size();
    std::vector<float> dist(nv, std::numeric_limits<float>::max());
    std::vector<int> pred(nv, -1);
    typedef std::pair<float, int> Item;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    dist[start_root_] = 0.0f;
    pq.push({0.0f, start_root_});
    while (!pq.empty()) {
        const Item it = pq.top();
        pq.pop();
        if (it.first > dist[it.second]) continue;
        if (it.second == goal_root_) break;
        for (const Edge& e : adj_[it.second]) {
            const float d = it.first + e.weight;
            if (d < dist[e.to]) {
                dist[e.to] = d;
                pred[e.to] = it.second;
                pq.push({d, e.to});
            }
        }
    }
    std::vector<int> path;
    if (pred[goal_root_] < 0) return path;
    for (int v = goal_root_; v != -1; v = pred[v]) path.push_back(v);
    std::reverse(path.begin(), path.end());
    return path;
}

BaseStateManagerPtr RRG::getStateManager() {
    auto space = std::dynamic_pointer_cast<SingleArmSpace>(space_);
    if (!space) throw std::runtime_error("RRG: space missing");
    auto m = space->createStateManager();
    std::vector<std::vector<float>> q;
    for (int v = 0; v < (int)node_.size(); v++)
        if (v != start_root_ && v != goal_root_) q.push_back(node_[v]);
    m->add_states(space->createStatesFromVector(q));
    return m;
}

void RRG::getStartAndGoalGroupStates(BaseStatesPtr& start_group_states, BaseStatesPtr& goal_group_states) {
    auto space = std::dynamic_pointer_cast<SingleArmSpace>(space_);
    if (!space) throw std::runtime_error("RRG: space missing");
    std::vector<std::vector<float>> a, b;
    for (int v : members_[1]) a.push_back(node_[v]);
    for (int v : members_[2]) b.push_back(node_[v]);
    start_group_states = space->createStatesFromVector(a);
    goal_group_states = space->createStatesFromVector(b);
}

void RRG::solve(BaseTerminationPtr termination) {
    auto space = std::dynamic_pointer_cast<SingleArmSpace>(space_);
    if (!space || !task_) throw std::runtime_error("RRG: space or task missing");
    const size_t dof = space->getDim();
    const float res = space->getResolution();
    auto states_of = [&](const std::vector<int>& vs) {
        std::vector<std::vector<float>> q;
        for (int v : vs) q.push_back(node_[v]);
        return q;
    };
    const std::vector<int> S0 = members_[1], G0 = members_[2];
    if (S0.empty() || G0.empty()) {
        task_->setFailureReason("InvalidInput");
        return;
    }
    // RRG.cpp:144-177: some start state and some goal state must be free
    std::vector<bool> fs, fg;
    space->checkStates(space->createStatesFromVector(states_of(S0)), fs);
    space->checkStates(space->createStatesFromVector(states_of(G0)), fg);
    if (std::find(fs.begin(), fs.end(), true) == fs.end() || std::find(fg.begin(), fg.end(), true) == fg.end()) {
        task_->setFailureReason("InvalidInput");
        return;
    }
    // RRG.cpp:179-232: the cheapest free direct motion between a start and a goal state wins
    {
        std::vector<std::vector<float>> a, b;
        for (int s : S0)
            for (int g : G0) {
                a.push_back(node_[s]);
                b.push_back(node_[g]);
            }
        std::vector<bool> mf;
        std::vector<float> mc;
        space->checkMotions(space->createStatesFromVector(a), space->createStatesFromVector(b), mf, mc);
        int best = -1;
        float best_cost = std::numeric_limits<float>::max();
        for (size_t i = 0; i < mf.size(); i++)
            if (mf[i] && mc[i] < best_cost) {
                best_cost = mc[i];
                best = (int)i;
            }
        if (best >= 0) {
            task_->setSolution(densify({a[best], b[best]}, res));
            return;
        }
    }
    termination->reset();
    // RRG.cpp:241-424 on the device: one mp_roadmap pass per loop pass, the roadmap in device
    // memory; the termination condition is checked once per pass, as in the reference's loop
    const int k = std::min(std::max(sample_attempts_, 1), 256);
    std::vector<float> sflat, gflat;
    for (int v : S0) sflat.insert(sflat.end(), node_[v].begin(), node_[v].end());
    for (int v : G0) gflat.insert(gflat.end(), node_[v].begin(), node_[v].end());
    const unsigned long long offset = space->sampleStreamOffset();
    mp_roadmap* rm = nullptr;
    mpcpp::check(mp_roadmap_create(space->robot(), space->scene(), sflat.data(), (int)S0.size(), gflat.data(),
                                   (int)G0.size(), k, max_travel_distance_, res, space->sampleStreamSeed(), offset, &rm),
                 "mp_roadmap_create");
    std::unique_ptr<mp_roadmap, int (*)(mp_roadmap*)> guard(rm, mp_roadmap_destroy);
    bool terminated = false;
    for (int t = 0; t < 1000000; t++) {
        if (termination->checkTerminationCondition()) {
            terminated = true;
            break;
        }
        int solved = 0;
        mpcpp::check(mp_roadmap_iterate(rm, &solved), "mp_roadmap_iterate");
        if (solved) break;
    }
    mp_roadmap_info info;
    mpcpp::check(mp_roadmap_info_get(rm, &info), "mp_roadmap_info_get");
    space->advanceSampleStream((unsigned long long)info.passes * (unsigned long long)k);
    // the device graph into the host one: device vertices 0..ns-1 are the start states, then the
    // goal states, then the added states in order
    std::vector<float> verts((size_t)info.vertices * dof);
    std::vector<int> groups(info.vertices), edges(2 * (size_t)info.edges), smem(info.start_group), gmem(info.goal_group);
    std::vector<float> weights(info.edges);
    mpcpp::check(mp_roadmap_read(rm, verts.data(), groups.data(), edges.data(), weights.data(), smem.data(), gmem.data()),
                 "mp_roadmap_read");
    std::vector<int> host_of(info.vertices);
    for (int v = 0; v < info.vertices; v++) {
        if (v < (int)S0.size()) host_of[v] = S0[v];
        else if (v < (int)(S0.size() + G0.size())) host_of[v] = G0[v - S0.size()];
        else host_of[v] = add_vertex(std::vector<float>(verts.begin() + (size_t)v * dof, verts.begin() + (size_t)(v + 1) * dof));
    }
    for (int e = 0; e < info.edges; e++) add_edge(host_of[edges[2 * e]], host_of[edges[2 * e + 1]], weights[e]);
    members_[1].clear();
    members_[2].clear();
    for (int v : smem) members_[1].push_back(host_of[v]);
    for (int v : gmem) members_[2].push_back(host_of[v]);
    if (!info.solved) {
        if (terminated) termination->printTerminationReason();
        task_->setFailureReason("MeetTerminationCondition");
        return;
    }
    std::vector<std::vector<float>> waypoints;
    for (int v : shortest_path())
        if (v != start_root_ && v != goal_root_) waypoints.push_back(node_[v]);
    task_->setSolution(get_full_path_ ? densify(waypoints, res) : waypoints);
}

}  // namespace CUDAMPLib
