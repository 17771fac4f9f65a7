// Internal declarations of the motion-planning library (host side).
#pragma once

#include <atomic>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "mp_abi.h"

namespace mp {

// Error carrying an ABI status code; converted to a return code at the C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MP_HIP(expr)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            throw ::mp::Error(e_ == hipErrorOutOfMemory ? MP_E_NOMEM : MP_E_HIP,                          \
                              std::string(#expr) + ": " + hipGetErrorString(e_));                         \
    } while (0)

// Flat robot model in RobotInfo layout plus the derived FP32 constants the generated code uses.
struct RobotModel {
    std::string name;
    int L = 0, S = 0, dof = 0;
    std::vector<int> jtype, parent, active, active_index;
    std::vector<double> origin;  // [L][16]
    std::vector<double> axis;    // [L][3]
    std::vector<float> lower, upper, defaults;  // [L]
    std::vector<int> sphere_link;
    std::vector<float> sphere_pos, sphere_r;
    std::vector<unsigned char> enabled;  // [L][L]
    // derived
    std::vector<float> act_lo, act_hi;
    std::vector<int> pi, pj;       // self pairs
    std::vector<float> pthr, psr;  // fl((ri+rj)^2), sqrtf(thr)
    uint64_t hash = 0;             // robot_hash(*this), set once the model is complete
};

RobotModel model_from_desc(const mp_robot_desc* d);

// Code generation
std::string robot_struct_source(const RobotModel& m, const std::string& sname);
std::string module_source(const std::vector<const RobotModel*>& robots);
std::string kernel_template_prelude();
std::string kernel_template_body();

// hipRTC compile with an on-disk cache; returns the code object bytes.
std::vector<char> rtc_compile(const std::string& source, const std::string& tag);
std::string& cache_dir();

// A loaded module for a set of robots.
struct Module {
    hipModule_t mod = nullptr;
    std::vector<std::string> robot_keys;
    std::string source;
    hipFunction_t fn(const std::string& name);
    ~Module();

  private:
    std::mutex mu_;
    std::vector<std::pair<std::string, hipFunction_t>> fns_;
};

// A robot set maps to one code object whatever order callers list it in: the module's robot
// slots follow the canonical order (robot name, then the hash of its generated struct).
// canonical_order(...)[i] is the input index placed in module slot i.
std::vector<int> canonical_order(const std::vector<const RobotModel*>& robots);
// Loads (compiling on a cache miss) the module of a robot set; slot_of[i] receives the module
// slot of robots[i].
std::shared_ptr<Module> get_module(const std::vector<const RobotModel*>& robots, std::vector<int>* slot_of = nullptr);

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull);
uint64_t robot_hash(const RobotModel& m);  // fnv1a of the robot's generated struct (computed)

float sqrt_threshold(float x);  // smallest float y >= 0 with sqrtf(y) >= x

}  // namespace mp

struct mp_robot {
    mp::RobotModel model;
    uint64_t key = 0;  // fnv1a of the generated robot struct
    std::shared_ptr<mp::Module> module;
};

struct mp_scene {
    int ns = 0, nb = 0, nc = 0, ncap = 0;
    int cap_s = 0, cap_b = 0, cap_c = 0, cap_k = 0;
    // one allocation: spheres | box bounds | boxes | cylinder bounds | cylinders | capsule bounds
    // | capsules, each section sized by the capacity of its kind
    float* dev = nullptr;
    size_t offs[7] = {0, 0, 0, 0, 0, 0, 0};
    // joint bounds of every check on this scene: MP_BOUNDS_ROBOT (the robot's limits), or the
    // 32 floats lower[16] | upper[16] in bnd_dev
    int bounds_mode = 0;
    float* bnd_dev = nullptr;
    // queries of servers that are queued or active on this scene: while > 0 the scene's device
    // memory is read by a server's stream, so update and destroy are refused
    mutable std::atomic<int> server_refs{0};
};
