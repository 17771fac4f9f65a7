// Host mirrors of the planner's device structures (layouts must match the kernel template).
#pragma once

#include <functional>

#include "mp_internal.h"

#define MP_SPH_W_H 4
#define MP_BOX_W_H 24
#define MP_CYL_W_H 20
#define MP_CAP_W_H 12

namespace mp {

struct SceneDevHost {
    int ns, nb, nc, ncap;
    const float* sph;
    const float* box;
    const float* cyl;
    const float* cap;
    const float* bnd;  // null: the robot's joint limits; else lower[16], upper[16]
    const float* bbox;  // bounding spheres of the boxes, cylinders, capsules (4 floats each)
    const float* bcyl;
    const float* bcap;
};

struct QueryDevHost {
    SceneDevHost scene;
    float* tree[2];
    int* parent[2];
    float* path;
    uint64_t seed;
    int robot, dof, cap, max_iters;
    int max_path, dof_max, sampler, pad2;
    float start[16];
    float goal[16];
};
static_assert(sizeof(QueryDevHost) == 288, "QueryDev layout");

struct QueryStateHost {
    int alloc[2];
    int sample_ctr;
    int status;
    int conn[2];
    int path_len;
    int workers;
    uint64_t checked;
    uint64_t flops;
    uint64_t t_init, t_done;
};
static_assert(sizeof(QueryStateHost) == 64, "QueryState layout");

// control block in device memory: [0] ticket, [1] remaining, [2] n_active, [3] budget,
// [4] first new query (index into the active list), [5] number of new queries, [6] exclusive,
// [7] iterations per ticket
struct PlanCtlHost {
    QueryDevHost* q;
    QueryStateHost* st;
    const int* active;
    int* ctl;
    float res;
    int max_step;
    unsigned long long* prof;  // planner phase profile counters, null when off
    float dd_radius, dd_alpha, dd_min;  // dynamic-domain settings
    int dd;                             // 1: dynamic domain on
    int balance;                        // MP_BALANCE_*
    float tree_ratio;                   // the balance mode's ratio (resolved: never 0)
    int window;                         // iterations running at once on one query (pRRTC's W)
};

hipStream_t lib_stream();
int lib_num_cus();
unsigned long long* lib_plan_profile();  // device counters when the phase profile is on, else null
unsigned long long* lib_work_counter();  // the algorithmic-flop counter of the check kernels
void launch(hipFunction_t f, unsigned grid, unsigned block, hipStream_t s, void** args, const char* family);
SceneDevHost scene_view(const mp_scene* sc);
void plan_release_workspaces();

// Buffers for `nslots` query slots: two SoA trees and two parent arrays per slot in device
// memory, a path buffer per slot in pinned host memory that the kernel writes directly (a
// finished path needs no copy of its own; the kernel orders the path before the query's status
// at system scope and the host reads it after the stream has synchronised), the query
// descriptors / states, the active list and the control block.
struct SlotPool {
    int nslots = 0, cap = 0, dof_max = 0, max_path = 0;
    float* trees = nullptr;   // [nslots][2][dof_max][cap]
    int* parents = nullptr;   // [nslots][2][cap]
    float* hpaths = nullptr;  // pinned host [nslots][max_path * dof_max]; rows are the robot's dof wide
    float* paths = nullptr;   // device address of hpaths
    QueryDevHost* dq = nullptr;
    QueryStateHost* dst = nullptr;
    int* dactive = nullptr;
    int* dctl = nullptr;      // control block (8 ints)
    int* hpinned = nullptr;   // pinned staging: control block (8) + active list (nslots)
    QueryStateHost* hst_pinned = nullptr;
    QueryDevHost* hq = nullptr;  // pinned host staging of the descriptors [nslots]

    SlotPool(int nslots, int cap, int dof_max, int max_path);
    ~SlotPool();
    size_t bytes_per_slot() const;
    void fill_query(int slot, int robot_slot, int dof, const mp_scene* scene, const float* start, const float* goal,
                    uint64_t seed, int max_iters, int sampler);
    PlanCtlHost ctl(const mp_rrtc_params& p) const;
};

size_t slot_bytes(int cap, int dof_max);  // device bytes of one slot
void check_domain(const mp_rrtc_params& p);  // dynamic-domain settings (throws MP_E_INVALID)
PlanCtlHost make_ctl(QueryDevHost* dq, QueryStateHost* dst, const int* dactive, int* dctl, const mp_rrtc_params& p);
// Iterations running at once on one query for a launch shape (the balance modes' window).
int balance_window(int teams, int cap, int lanes, bool exclusive);

// Runs f, converting exceptions to an ABI status code (message in mp_last_error()).
int guard_call(const std::function<void()>& f);
void run_to_completion(Module& mod, SlotPool& pool, int nq, const mp_rrtc_params& p, int team, int teams,
                       int cap, int lanes, hipStream_t s);

// The buffers one persistent solve uses: descriptors, states, active list and control block in
// device memory, and their host staging (hctl holds 16 + nq ints; pinned or pageable).
struct RunBuffers {
    QueryDevHost* dq;
    QueryStateHost* dst;
    int* dactive;
    int* dctl;
    const QueryDevHost* hq;
    int* hctl;
    QueryStateHost* hst;
};
// Admission and planning launches for the nq queries of b.hq, then the states back into b.hst;
// returns after the stream has synchronised.  family: the kernel-timing family of the planning
// launch (null: untimed).
void run_queries(Module& mod, const RunBuffers& b, int nq, const mp_rrtc_params& p, int team, int teams, int cap,
                 int lanes, hipStream_t s, const char* family);

// Module, module robot slots and launch shape of a solve (explicit params, else the tuned
// shape of the robot set and class, else the defaults); validates the params.
struct LaunchPlan {
    std::shared_ptr<Module> mod;
    std::vector<int> slot_of;
    int dof_max = 1, team = 1, teams = 1, cap = 0, lanes = 64;
};
LaunchPlan plan_launch(int nq, mp_robot* const* robots, const mp_scene* const* scenes, const mp_rrtc_params* params,
                       int default_wave_count);

// The per-device address of the no-allocation entry point's scratch array, looked up once per
// device with that device current (called by mp_init).
void noalloc_setup(int device);

// Launch shape of the planning kernel: waves per team, total waves (0: resident teams) and
// teams per query (0: unbounded).
struct Shape {
    int team = 0;
    int waves = 0;
    int cap = 0;
    int lanes = 0;  // lanes per iteration (64 or 32; 0 = default)
};
std::string robot_set_key(const std::vector<mp_robot*>& robots);
Shape tuned_shape(const std::string& set_key, int cls);  // team 0 when nothing is tuned
void store_shape(const std::string& set_key, int cls, Shape sh);
// Fills `out` from a finished query's state; call after the stream that ran the query has synchronised.
void harvest(const SlotPool& pool, int slot, int dof, const QueryStateHost& st, mp_path* out);
// The status and counters part of harvest; returns the number of path waypoints to copy into
// out->waypoints (0: none).
int harvest_status(const QueryStateHost& st, int max_iters, int cap, mp_path* out);
constexpr int kDefaultTeamWaves = 4;
// Iterations a team runs on the query its ticket picked (solve paths; the server uses 1).
constexpr int kItersPerTicket = 4;
constexpr int kDefaultLanes = 64;  // lanes per iteration when neither given nor tuned
// admission workgroups: four waves whatever the planning team size (the start/goal check gives
// the same verdicts for any team size, and more threads empty a reused slot sooner)
constexpr unsigned kInitBlock = 256;
// control-block flag of the paired iteration: lanes per iteration 32 (max_step <= 32 checked)
inline int pair_flag(int lanes) { return lanes == 32 ? 1 : 0; }
// Resident teams of `team_waves` waves the planning kernel can keep on the device.
int default_teams(hipFunction_t plan_fn, int team_waves);
int team_size(const mp_rrtc_params& p);  // waves per team (the default when p.team_waves is 0)

}  // namespace mp
