// CPRRTC::RobotSolver with the reference's public interface (cudampl/CPRRTC/include/RobotSolver.h:36-123),
// implemented over the MI355X C ABI (include/mp_abi.h): the robot-specialised module is
// generated and hipRTC-compiled in the constructor, obstacles stay resident on the device, and
// solve() runs one persistent planning launch.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../mp_eigen.h"
#include "Obstacle.h"

struct mp_robot;
struct mp_scene;

namespace CPRRTC {

class RobotSolver {
  public:
    RobotSolver(std::string robot_name, size_t dim, const std::vector<int>& joint_types,
                const std::vector<mpx::Isometry3d>& joint_poses, const std::vector<mpx::Vector3d>& joint_axes,
                const std::vector<int>& link_parent_link_maps, const std::vector<int>& self_collision_spheres_to_link_map,
                const std::vector<mpx::Vector3d>& self_collision_spheres_pos_in_link,
                const std::vector<float>& self_collision_spheres_radius, const std::vector<bool>& active_joint_map,
                const std::vector<float>& lower, const std::vector<float>& upper,
                const std::vector<float>& default_joint_values, const std::vector<std::string>& link_names,
                const std::vector<std::vector<bool>>& self_collision_enables_map, float resolution = 0.02f);
    ~RobotSolver();
    RobotSolver(const RobotSolver&) = delete;
    RobotSolver& operator=(const RobotSolver&) = delete;

    void setEnvObstacleCache(int num_of_spheres, int num_of_cuboids, int num_of_cylinders);
    void updateEnvObstacle(std::vector<Sphere>& spheres, std::vector<Cuboid>& cuboids, std::vector<Cylinder>& cylinders);
    // Path over active joints, start ... goal; empty on failure.
    std::vector<std::vector<float>> solve(std::vector<float>& start, std::vector<float>& goal);

    // Additions (no counterpart in the reference): planner settings and the last solve's status.
    void setMaxIterations(int n) { max_iterations_ = n; }
    void setSeed(unsigned long long seed) { seed_ = seed; }
    int lastStatus() const { return last_status_; }

  private:
    size_t dim_;
    float resolution_;
    int max_iterations_ = 10000;
    unsigned long long seed_ = 0;
    int last_status_ = -1;
    int cap_s_ = 0, cap_c_ = 0, cap_y_ = 0;
    mp_robot* robot_ = nullptr;
    mp_scene* scene_ = nullptr;
};

typedef std::shared_ptr<RobotSolver> RobotSolverPtr;

}  // namespace CPRRTC
