// Helpers shared by the C++ API classes: C-ABI status checks and the RobotInfo -> mp_robot_desc
// conversion (robot_info.hpp:13-220 layout).
#pragma once
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "mp_abi.h"
#include "mp_eigen.h"

namespace mpcpp {

inline void check(int rc, const char* what) {
    if (rc != MP_OK) throw std::runtime_error(std::string(what) + ": " + mp_last_error());
}

inline void ensure_init() {
    static bool done = false;
    if (!done) {
        check(mp_init(0), "mp_init");
        done = true;
    }
}

// Rz(yaw) Ry(pitch) Rx(roll), row major, float (obstacle_generator.hpp:78-104)
inline void rpy_to_matrix(float roll, float pitch, float yaw, float* R) {
    float cx = std::cos(roll), sx = std::sin(roll), cy = std::cos(pitch), sy = std::sin(pitch), cz = std::cos(yaw),
          sz = std::sin(yaw);
    R[0] = cz * cy;
    R[1] = cz * sy * sx - sz * cx;
    R[2] = cz * sy * cx + sz * sx;
    R[3] = sz * cy;
    R[4] = sz * sy * sx + cz * cx;
    R[5] = sz * sy * cx - cz * sx;
    R[6] = -sy;
    R[7] = cy * sx;
    R[8] = cy * cx;
}

struct RobotArrays {
    std::string name;
    std::vector<int> jt, par, act, slink;
    std::vector<double> org, axes;
    std::vector<float> lo, hi, def, spos, srad;
    std::vector<unsigned char> en;
    mp_robot_desc desc{};
    int dof = 0;

    RobotArrays(const std::string& nm, const std::vector<int>& joint_types, const std::vector<mpx::Isometry3d>& poses,
                const std::vector<mpx::Vector3d>& axes_in, const std::vector<int>& parents,
                const std::vector<int>& sphere_map, const std::vector<mpx::Vector3d>& sphere_pos,
                const std::vector<float>& sphere_radius, const std::vector<bool>& active,
                const std::vector<float>& lower, const std::vector<float>& upper, const std::vector<float>& defaults,
                size_t num_links, const std::vector<std::vector<bool>>& enables)
        : name(nm), jt(joint_types), par(parents), slink(sphere_map), lo(lower), hi(upper), def(defaults),
          srad(sphere_radius) {
        const size_t L = joint_types.size();
        if (poses.size() != L || axes_in.size() != L || parents.size() != L || active.size() != L || lower.size() != L ||
            upper.size() != L || defaults.size() != L || (num_links && num_links != L))
            throw std::runtime_error("robot description: per-joint arrays must have one entry per link");
        if (sphere_map.size() != sphere_pos.size() || sphere_map.size() != sphere_radius.size())
            throw std::runtime_error("robot description: sphere arrays differ in length");
        for (size_t i = 0; i < L; i++) {
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++) org.push_back(mpx::mat(poses[i], r, c));
            for (int a = 0; a < 3; a++) axes.push_back(mpx::vec(axes_in[i], a));
            act.push_back(active[i] ? 1 : 0);
            dof += active[i] ? 1 : 0;
        }
        for (const auto& p : sphere_pos)
            for (int a = 0; a < 3; a++) spos.push_back((float)mpx::vec(p, a));
        en.assign(L * L, 1);
        if (!enables.empty()) {
            if (enables.size() != L) throw std::runtime_error("self-collision enable map must be L x L");
            for (size_t i = 0; i < L; i++)
                for (size_t j = 0; j < L; j++) en[i * L + j] = enables[i][j] ? 1 : 0;
        }
        desc.name = name.c_str();
        desc.num_links = (int)L;
        desc.joint_types = jt.data();
        desc.joint_origins = org.data();
        desc.joint_axes = axes.data();
        desc.parents = par.data();
        desc.active = act.data();
        desc.lower = lo.data();
        desc.upper = hi.data();
        desc.defaults = def.data();
        desc.num_spheres = (int)srad.size();
        desc.sphere_link = slink.data();
        desc.sphere_pos = spos.data();
        desc.sphere_radius = srad.data();
        desc.collision_enabled = en.data();
    }
};

}  // namespace mpcpp
