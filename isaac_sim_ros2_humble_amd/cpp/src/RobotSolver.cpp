// CPRRTC::RobotSolver over the C ABI (reference: cudampl/CPRRTC/src/RobotSolver.cu).
#include "CPRRTC/RobotSolver.h"

#include <cmath>
#include <iostream>
#include <stdexcept>

#include "mp_abi.h"
#include "mp_cpp_util.h"

namespace CPRRTC {

RobotSolver::RobotSolver(std::string robot_name, size_t dim, const std::vector<int>& joint_types,
                         const std::vector<mpx::Isometry3d>& joint_poses, const std::vector<mpx::Vector3d>& joint_axes,
                         const std::vector<int>& link_parent_link_maps,
                         const std::vector<int>& self_collision_spheres_to_link_map,
                         const std::vector<mpx::Vector3d>& self_collision_spheres_pos_in_link,
                         const std::vector<float>& self_collision_spheres_radius,
                         const std::vector<bool>& active_joint_map, const std::vector<float>& lower,
                         const std::vector<float>& upper, const std::vector<float>& default_joint_values,
                         const std::vector<std::string>& link_names,
                         const std::vector<std::vector<bool>>& self_collision_enables_map, float resolution)
    : dim_(dim), resolution_(resolution) {
    mpcpp::ensure_init();
    mpcpp::RobotArrays a(robot_name, joint_types, joint_poses, joint_axes, link_parent_link_maps,
                         self_collision_spheres_to_link_map, self_collision_spheres_pos_in_link,
                         self_collision_spheres_radius, active_joint_map, lower, upper, default_joint_values,
                         link_names.size(), self_collision_enables_map);
    if (a.dof != (int)dim) throw std::runtime_error("RobotSolver: dim does not match the number of active joints");
    mpcpp::check(mp_robot_create(&a.desc, &robot_), "mp_robot_create");
    setEnvObstacleCache(0, 0, 0);
}

RobotSolver::~RobotSolver() {
    if (scene_) mp_scene_destroy(scene_);
    if (robot_) mp_robot_destroy(robot_);
}

void RobotSolver::setEnvObstacleCache(int num_of_spheres, int num_of_cuboids, int num_of_cylinders) {
    if (scene_) mp_scene_destroy(scene_);
    scene_ = nullptr;
    cap_s_ = num_of_spheres;
    cap_c_ = num_of_cuboids;
    cap_y_ = num_of_cylinders;
    std::vector<float> s(4 * (size_t)std::max(cap_s_, 1), 0.0f), b(18 * (size_t)std::max(cap_c_, 1), 0.0f),
        c(14 * (size_t)std::max(cap_y_, 1), 0.0f);
    mpcpp::check(mp_scene_create(s.data(), cap_s_, b.data(), cap_c_, c.data(), cap_y_, nullptr, 0, &scene_),
                 "mp_scene_create");
    mpcpp::check(mp_scene_update(scene_, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0), "mp_scene_update");
    // the CPRRTC kernel checks no joint limits (RobotSolver.cu:507-799): its samples are drawn
    // inside them (RobotSolver.cu:20-36) and its motions stay between checked states
    mpcpp::check(mp_scene_set_joint_bounds(scene_, MP_BOUNDS_NONE, nullptr, nullptr, 0), "mp_scene_set_joint_bounds");
}

void RobotSolver::updateEnvObstacle(std::vector<Sphere>& spheres, std::vector<Cuboid>& cuboids,
                                    std::vector<Cylinder>& cylinders) {
    if ((int)spheres.size() > cap_s_ || (int)cuboids.size() > cap_c_ || (int)cylinders.size() > cap_y_) {
        std::cerr << "The number of obstacles exceeds the maximum number of obstacles." << std::endl;
        return;  // reference behaviour (RobotSolver.cu:186-190)
    }
    std::vector<float> s, b, c;
    for (const Sphere& o : spheres) s.insert(s.end(), {o.x, o.y, o.z, o.radius});
    for (const Cuboid& o : cuboids) {
        float R[9];
        mpcpp::rpy_to_matrix(o.roll, o.pitch, o.yaw, R);
        b.insert(b.end(), {o.x, o.y, o.z});
        b.insert(b.end(), R, R + 9);
        b.insert(b.end(), {o.x_min, o.y_min, o.z_min, o.x_max, o.y_max, o.z_max});
    }
    for (const Cylinder& o : cylinders) {
        float R[9];
        mpcpp::rpy_to_matrix(o.roll, o.pitch, o.yaw, R);
        c.insert(c.end(), {o.x, o.y, o.z});
        c.insert(c.end(), R, R + 9);
        c.insert(c.end(), {o.radius, o.height});
    }
    mpcpp::check(mp_scene_update(scene_, s.data(), (int)spheres.size(), b.data(), (int)cuboids.size(), c.data(),
                                 (int)cylinders.size(), nullptr, 0),
                 "mp_scene_update");
}

std::vector<std::vector<float>> RobotSolver::solve(std::vector<float>& start, std::vector<float>& goal) {
    if (start.size() != dim_ || goal.size() != dim_) {
        std::cerr << "The size of start and goal configurations must be equal to the dimension of the robot."
                  << std::endl;
        return {};
    }
    mp_rrtc_params p;
    mp_rrtc_default_params(&p);
    p.resolution = resolution_;
    p.max_iterations = max_iterations_;
    p.seed = seed_;
    std::vector<float> wp((size_t)p.max_path * dim_);
    mp_path out{};
    out.waypoints = wp.data();
    out.capacity = p.max_path;
    mpcpp::check(mp_rrtc_solve(robot_, scene_, start.data(), goal.data(), &p, &out), "mp_rrtc_solve");
    last_status_ = out.status;
    std::vector<std::vector<float>> path;
    if (out.status != MP_PLAN_SOLVED) return path;
    for (int k = 0; k < out.num_waypoints; k++) path.emplace_back(wp.begin() + k * dim_, wp.begin() + (k + 1) * dim_);
    return path;
}

}  // namespace CPRRTC
