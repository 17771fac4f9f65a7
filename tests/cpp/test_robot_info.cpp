// Loads a robot with the ROS-free RobotInfo (URDF + SRDF + collision-sphere YAML) and prints its
// arrays as "key v0 v1 ..." lines for tests/test_robot_info.py.
#include <cstdio>
#include <string>

#include "foliation_planner/robot_info.hpp"

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    RobotInfo info(argv[1], argv[2], argv[3], argv[4]);
    auto ints = [](const char* k, const std::vector<int>& v) {
        std::printf("%s", k);
        for (int x : v) std::printf(" %d", x);
        std::printf("\n");
    };
    auto floats = [](const char* k, const std::vector<float>& v) {
        std::printf("%s", k);
        for (float x : v) std::printf(" %.9g", x);
        std::printf("\n");
    };
    std::printf("dimension %d\n", info.getDimension());
    std::printf("links");
    for (const auto& l : info.getLinkNames()) std::printf(" %s", l.c_str());
    std::printf("\njoint_names");
    for (const auto& j : info.getJointNames()) std::printf(" %s", j.empty() ? "-" : j.c_str());
    std::printf("\n");
    ints("joint_types", info.getJointTypes());
    ints("parents", info.getLinkMaps());
    std::vector<int> act;
    for (bool b : info.getActiveJointMap()) act.push_back(b);
    ints("active", act);
    floats("lower", info.getLowerBounds());
    floats("upper", info.getUpperBounds());
    floats("defaults", info.getDefaultJointValues());
    ints("sphere_link", info.getCollisionSpheresMap());
    std::vector<float> pos;
    for (const auto& p : info.getCollisionSpheresPos()) pos.insert(pos.end(), p.begin(), p.end());
    floats("sphere_pos", pos);
    floats("sphere_radius", info.getCollisionSpheresRadius());
    std::printf("joint_origins");
    for (const auto& T : info.getJointPoses())
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) std::printf(" %.17g", mpx::mat(T, r, c));
    std::printf("\njoint_axes");
    for (const auto& a : info.getJointAxes())
        for (int i = 0; i < 3; i++) std::printf(" %.17g", mpx::vec(a, i));
    std::printf("\n");
    std::vector<int> en;
    for (const auto& row : info.getSelfCollisionEnabledMap())
        for (bool b : row) en.push_back(b);
    ints("collision_enabled", en);
    return 0;
}
