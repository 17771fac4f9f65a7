// Planner host side: slot pools, the plain one-query-per-launch solve (the
// CPRRTC::RobotSolver::solve equivalent, cudampl/CPRRTC/src/RobotSolver.cu:348-473) and the
// batched persistent solve (many queries, one launch).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>

#include "mp_plan.h"

using mp::Error;

namespace mp {

size_t slot_bytes(int cap, int dof_max) {
    // per tree: coordinates [dof_max][cap], then parents int[cap] and domain radii float[cap]
    return (size_t)2 * dof_max * cap * sizeof(float) + (size_t)2 * cap * (sizeof(int) + sizeof(float));
}

SlotPool::SlotPool(int ns, int c, int d, int mp_) : nslots(ns), cap(c), dof_max(d), max_path(mp_) {
    try {
        MP_HIP(hipMalloc(&trees, (size_t)nslots * 2 * dof_max * cap * sizeof(float)));
        MP_HIP(hipMalloc(&parents, (size_t)nslots * 2 * 2 * cap * sizeof(int)));  // parents, radii
        MP_HIP(hipHostMalloc(&hpaths, (size_t)nslots * max_path * dof_max * sizeof(float),
                             hipHostMallocMapped | hipHostMallocCoherent));
        MP_HIP(hipHostGetDevicePointer((void**)&paths, hpaths, 0));
        MP_HIP(hipMalloc(&dq, (size_t)nslots * sizeof(QueryDevHost)));
        MP_HIP(hipMalloc(&dst, (size_t)nslots * sizeof(QueryStateHost)));
        MP_HIP(hipMalloc(&dactive, (size_t)nslots * sizeof(int)));
        MP_HIP(hipMalloc(&dctl, 16 * sizeof(int)));
        MP_HIP(hipHostMalloc(&hpinned, (size_t)(16 + nslots) * sizeof(int), hipHostMallocDefault));
        MP_HIP(hipHostMalloc(&hst_pinned, (size_t)nslots * sizeof(QueryStateHost), hipHostMallocDefault));
        MP_HIP(hipHostMalloc(&hq, (size_t)nslots * sizeof(QueryDevHost), hipHostMallocDefault));
        MP_HIP(hipMemset(dst, 0, (size_t)nslots * sizeof(QueryStateHost)));
        // every tree slot starts empty: NaN coordinates (0x7fc0dead) and parent -2 (the radii read
        // 0xfffffffe, a NaN: an unbounded domain)
        MP_HIP(hipMemsetD32((hipDeviceptr_t)trees, (int)0x7fc0deadu, (size_t)nslots * 2 * dof_max * cap));
        MP_HIP(hipMemsetD32((hipDeviceptr_t)parents, -2, (size_t)nslots * 2 * 2 * cap));
        MP_HIP(hipDeviceSynchronize());
    } catch (...) {
        this->~SlotPool();
        throw;
    }
    std::memset(hq, 0, (size_t)nslots * sizeof(QueryDevHost));
}

SlotPool::~SlotPool() {
    if (trees) hipFree(trees);
    if (parents) hipFree(parents);
    if (hpaths) hipHostFree(hpaths);
    if (dq) hipFree(dq);
    if (dst) hipFree(dst);
    if (dactive) hipFree(dactive);
    if (dctl) hipFree(dctl);
    if (hpinned) hipHostFree(hpinned);
    if (hst_pinned) hipHostFree(hst_pinned);
    if (hq) hipHostFree(hq);
    trees = nullptr;
    parents = nullptr;
    hpaths = nullptr;
    paths = nullptr;
    dq = nullptr;
    dst = nullptr;
    dactive = nullptr;
    dctl = nullptr;
    hpinned = nullptr;
    hst_pinned = nullptr;
    hq = nullptr;
}

size_t SlotPool::bytes_per_slot() const { return slot_bytes(cap, dof_max); }

void SlotPool::fill_query(int slot, int robot_slot, int dof, const mp_scene* scene, const float* start,
                          const float* goal, uint64_t seed, int max_iters, int sampler) {
    QueryDevHost& q = hq[slot];
    std::memset(&q, 0, sizeof(q));
    q.scene = scene_view(scene);
    for (int t = 0; t < 2; t++) {
        q.tree[t] = trees + ((size_t)slot * 2 + t) * dof_max * cap;
        q.parent[t] = parents + ((size_t)slot * 2 + t) * 2 * cap;  // radii follow at + cap
    }
    q.path = paths + (size_t)slot * max_path * dof_max;
    q.seed = seed;
    q.robot = robot_slot;
    q.dof = dof;
    q.cap = cap;
    q.max_iters = max_iters;
    q.max_path = max_path;
    q.dof_max = dof_max;
    q.sampler = sampler;
    for (int j = 0; j < dof; j++) {
        q.start[j] = start[j];
        q.goal[j] = goal[j];
    }
}

PlanCtlHost SlotPool::ctl(const mp_rrtc_params& p) const { return make_ctl(dq, dst, dactive, dctl, p); }

PlanCtlHost make_ctl(QueryDevHost* dq, QueryStateHost* dst, const int* dactive, int* dctl, const mp_rrtc_params& p) {
    PlanCtlHost c;
    c.q = dq;
    c.st = dst;
    c.active = dactive;
    c.ctl = dctl;
    c.res = p.resolution;
    c.max_step = p.max_step;
    c.prof = lib_plan_profile();
    c.dd_radius = p.dd_radius;
    c.dd_alpha = p.dd_alpha;
    c.dd_min = p.dd_min_radius;
    c.dd = p.dynamic_domain;
    c.balance = p.balance;
    c.tree_ratio = p.tree_ratio > 0.0f ? p.tree_ratio : (p.balance == MP_BALANCE_ALTERNATE ? 1.0f : 0.5f);
    c.window = 1;  // set by the launcher from its shape
    return c;
}

int balance_window(int teams, int cap, int lanes, bool exclusive) {
    const int per_query = exclusive ? 1 : (cap > 0 && cap < teams ? cap : teams);
    return std::max(1, per_query * (lanes == 32 ? 2 : 1));
}

int default_teams(hipFunction_t plan_fn, int team_waves) {
    int per_cu = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, plan_fn, 64 * team_waves, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return std::max(1, per_cu) * std::max(lib_num_cus(), 1);
}

int team_size(const mp_rrtc_params& p) { return p.team_waves > 0 ? p.team_waves : kDefaultTeamWaves; }

void check_domain(const mp_rrtc_params& p) {
    if (p.balance < MP_BALANCE_SMALLER || p.balance > MP_BALANCE_ALTERNATE)
        throw Error(MP_E_INVALID, "balance must be one of the MP_BALANCE_* modes");
    if (!(p.tree_ratio >= 0.0f) || p.tree_ratio > 1e6f) throw Error(MP_E_INVALID, "tree_ratio must be >= 0");
    if (p.dynamic_domain != 0 && p.dynamic_domain != 1) throw Error(MP_E_INVALID, "dynamic_domain must be 0 or 1");
    if (p.dynamic_domain && !(p.dd_radius > 0.0f && p.dd_min_radius > 0.0f && p.dd_alpha >= 0.0f && p.dd_alpha < 1.0f))
        throw Error(MP_E_INVALID, "dynamic domain needs dd_radius > 0, dd_min_radius > 0 and 0 <= dd_alpha < 1");
}

namespace {
std::mutex g_pool_mu;
std::map<std::string, std::unique_ptr<SlotPool>> g_pools;

SlotPool& pool_for(const char* tag, int nslots, int cap, int dof_max, int max_path) {
    std::string k = std::string(tag) + ":" + std::to_string(nslots) + ":" + std::to_string(cap) + ":" +
                    std::to_string(dof_max) + ":" + std::to_string(max_path);
    std::lock_guard<std::mutex> g(g_pool_mu);
    auto it = g_pools.find(k);
    if (it != g_pools.end()) return *it->second;
    // keep at most a few pools alive
    if (g_pools.size() >= 4) {
        hipStreamSynchronize(lib_stream());
        g_pools.clear();
    }
    auto p = std::make_unique<SlotPool>(nslots, cap, dof_max, max_path);
    auto& ref = *p;
    g_pools.emplace(k, std::move(p));
    return ref;
}

void check_params(const mp_rrtc_params* p) {
    if (!p) throw Error(MP_E_INVALID, "null params");
    if (!(p->resolution > 0.0f)) throw Error(MP_E_INVALID, "resolution must be > 0");
    if (p->max_step < 1 || p->max_step > 64) throw Error(MP_E_INVALID, "max_step must be in [1,64]");
    if (p->max_iterations < 1) throw Error(MP_E_INVALID, "max_iterations must be >= 1");
    if (p->tree_capacity < 2 || p->tree_capacity > (1 << 22)) throw Error(MP_E_INVALID, "tree_capacity out of range");
    if (p->max_path < 2) throw Error(MP_E_INVALID, "max_path must be >= 2");
    if (p->waves < 0) throw Error(MP_E_INVALID, "waves must be >= 0");
    if (p->team_waves != 0 && p->team_waves != 1 && p->team_waves != 2 && p->team_waves != 4)
        throw Error(MP_E_INVALID, "team_waves must be 0, 1, 2 or 4");
    if (p->teams_per_query < 0) throw Error(MP_E_INVALID, "teams_per_query must be >= 0");
    if (p->lanes_per_iteration != 0 && p->lanes_per_iteration != 32 && p->lanes_per_iteration != 64)
        throw Error(MP_E_INVALID, "lanes_per_iteration must be 0, 32 or 64");
    if (p->lanes_per_iteration == 32 && p->max_step > 32)
        throw Error(MP_E_INVALID, "lanes_per_iteration 32 needs max_step <= 32");
    if (p->sampler != MP_SAMPLER_UNIFORM && p->sampler != MP_SAMPLER_HALTON)
        throw Error(MP_E_INVALID, "sampler must be MP_SAMPLER_UNIFORM or MP_SAMPLER_HALTON");
    check_domain(*p);
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

void plan_release_workspaces() {
    std::lock_guard<std::mutex> g(g_pool_mu);
    g_pools.clear();
}

// Runs the queries already described in pool.hq[0..nq) to completion in one persistent launch
// of `teams` workgroups of `team` waves.
void run_to_completion(Module& mod, SlotPool& pool, int nq, const mp_rrtc_params& p, int team, int teams,
                       int cap, int lanes, hipStream_t s) {
    const RunBuffers b{pool.dq, pool.dst, pool.dactive, pool.dctl, pool.hq, pool.hpinned, pool.hst_pinned};
    run_queries(mod, b, nq, p, team, teams, cap, lanes, s, "plan");
}

void run_queries(Module& mod, const RunBuffers& b, int nq, const mp_rrtc_params& p, int team, int teams, int cap,
                 int lanes, hipStream_t s, const char* family) {
    const unsigned block = 64u * (unsigned)team;
    MP_HIP(hipMemcpyAsync(b.dq, b.hq, (size_t)nq * sizeof(QueryDevHost), hipMemcpyHostToDevice, s));
    int* h = b.hctl;
    h[0] = 0;       // ticket
    h[1] = nq;      // running
    h[2] = nq;      // active-list length
    h[3] = 0;       // no ticket budget
    h[4] = 0;       // first new
    h[5] = nq;      // new count
    h[6] = p.exclusive ? 1 : 0;
    h[7] = kItersPerTicket;
    h[8] = cap;
    h[9] = pair_flag(lanes);
    for (int k = 0; k < nq; k++) h[16 + k] = k;
    MP_HIP(hipMemcpyAsync(b.dctl, h, 16 * sizeof(int), hipMemcpyHostToDevice, s));
    MP_HIP(hipMemcpyAsync(b.dactive, h + 16, (size_t)nq * sizeof(int), hipMemcpyHostToDevice, s));
    PlanCtlHost ctl = make_ctl(b.dq, b.dst, b.dactive, b.dctl, p);
    ctl.window = balance_window(teams, cap, lanes, p.exclusive != 0);
    void* args[] = {&ctl};
    launch(mod.fn("mp_plan_init"), (unsigned)nq, kInitBlock, s, args, family ? "plan_init" : nullptr);
    launch(mod.fn("mp_plan"), (unsigned)teams, block, s, args, family);
    MP_HIP(hipMemcpyAsync(b.hst, b.dst, (size_t)nq * sizeof(QueryStateHost), hipMemcpyDeviceToHost, s));
    MP_HIP(hipStreamSynchronize(s));
}

void harvest(const SlotPool& pool, int slot, int dof, const QueryStateHost& st, mp_path* out) {
    const int n = harvest_status(st, pool.hq[slot].max_iters, pool.cap, out);
    if (n > 0)
        std::memcpy(out->waypoints, pool.hpaths + (size_t)slot * pool.max_path * pool.dof_max,
                    (size_t)n * dof * sizeof(float));
}

int harvest_status(const QueryStateHost& st, int max_iters, int cap, mp_path* out) {
    out->status = st.status < 0 ? MP_PLAN_RUNNING : st.status;
    out->iterations = std::min(st.sample_ctr, max_iters);
    out->start_tree_size = std::min(st.alloc[0], cap);
    out->goal_tree_size = std::min(st.alloc[1], cap);
    out->configs_checked = (int64_t)st.checked;
    out->work_flops = (double)st.flops;
    out->device_ms = st.t_done >= st.t_init ? (double)(st.t_done - st.t_init) * 1e-5 : 0.0;  // 100 MHz clock
    out->num_waypoints = 0;
    if (st.status == MP_PLAN_SOLVED) {
        if (st.path_len == -1) {  // a tree slot on the path never became readable: no path
            out->status = MP_PLAN_PATH_LOST;
            return 0;
        }
        // the path needs more waypoints than the pool's max_path (path_len = -length) or the
        // caller's buffer holds: report its length, copy nothing
        const int need = st.path_len < 0 ? -st.path_len : st.path_len;
        out->num_waypoints = need;
        if (st.path_len < 0 || need > out->capacity || !out->waypoints) {
            out->status = MP_PLAN_PATH_TOO_LONG;
            return 0;
        }
        return need;
    }
    return 0;
}

}  // namespace mp

namespace mp {

// ------------------------------------------------------------------------------ launch shapes
// The autotuner's choice of launch shape per robot set and class (0: one query per launch,
// 1: batches), kept in memory and in <cache dir>/tune_<hash>.txt, one line per class:
// "class team waves teams_per_query lanes_per_iteration".
namespace {
std::mutex g_tune_mu;
std::map<std::string, std::pair<Shape, bool>> g_tuned;  // key#class -> (shape, looked up)

std::string tune_file(const std::string& set_key) {
    char h[32];
    std::snprintf(h, sizeof(h), "%016llx", (unsigned long long)fnv1a(set_key));
    return cache_dir() + "/tune_" + h + ".txt";
}
}  // namespace

std::string robot_set_key(const std::vector<mp_robot*>& robots) {
    std::vector<std::string> parts;
    for (auto* r : robots) {
        char h[32];
        std::snprintf(h, sizeof(h), "%016llx", (unsigned long long)r->key);
        parts.push_back(r->model.name + "#" + h);
    }
    std::sort(parts.begin(), parts.end());
    parts.erase(std::unique(parts.begin(), parts.end()), parts.end());
    std::string k;
    for (auto& p : parts) k += p + ";";
    return k;
}

Shape tuned_shape(const std::string& set_key, int cls) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    const std::string k = set_key + "@" + std::to_string(cls);
    auto it = g_tuned.find(k);
    if (it != g_tuned.end()) return it->second.first;
    Shape sh;
    if (FILE* f = std::fopen(tune_file(set_key).c_str(), "r")) {
        char line[128];
        while (std::fgets(line, sizeof(line), f)) {
            int c, t, w, q, l = 0;
            const int n = std::sscanf(line, "%d %d %d %d %d", &c, &t, &w, &q, &l);
            if (n >= 4 && c == cls && (t == 1 || t == 2 || t == 4) && w >= 0 && q >= 0 && (l == 0 || l == 32 || l == 64))
                sh = Shape{t, w, q, l};
        }
        std::fclose(f);
    }
    g_tuned[k] = {sh, true};
    return sh;
}

void store_shape(const std::string& set_key, int cls, Shape sh) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tuned[set_key + "@" + std::to_string(cls)] = {sh, true};
    // rewrite the file with every class known for this set (best effort)
    std::string lines;
    for (int c = 0; c < 2; c++) {
        auto it = g_tuned.find(set_key + "@" + std::to_string(c));
        if (it != g_tuned.end() && it->second.first.team > 0)
            lines += std::to_string(c) + " " + std::to_string(it->second.first.team) + " " +
                     std::to_string(it->second.first.waves) + " " + std::to_string(it->second.first.cap) + " " +
                     std::to_string(it->second.first.lanes) + "\n";
    }
    const std::string path = tune_file(set_key), tmp = path + ".tmp";
    if (FILE* f = std::fopen(tmp.c_str(), "w")) {
        std::fputs(lines.c_str(), f);
        std::fclose(f);
        if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
    }
}

LaunchPlan plan_launch(int nq, mp_robot* const* robots, const mp_scene* const* scenes, const mp_rrtc_params* params,
                       int default_wave_count) {
    if (lib_stream() == nullptr) throw Error(MP_E_STATE, "mp_init() has not been called");
    check_params(params);
    if (nq < 1 || !robots || !scenes) throw Error(MP_E_INVALID, "bad batch arguments");
    LaunchPlan lp;
    // robot set -> module (distinct robots, mapped to the module's canonical slots)
    std::vector<mp_robot*> uniq;
    std::vector<int>& slot_of = lp.slot_of;
    slot_of.assign(nq, 0);
    int& dof_max = lp.dof_max;
    dof_max = 1;
    for (int k = 0; k < nq; k++) {
        if (!robots[k] || !scenes[k]) throw Error(MP_E_INVALID, "null robot or scene in batch");
        int s = -1;
        for (size_t u = 0; u < uniq.size(); u++)
            if (uniq[u] == robots[k] || uniq[u]->key == robots[k]->key) s = (int)u;
        if (s < 0) {
            s = (int)uniq.size();
            uniq.push_back(robots[k]);
        }
        slot_of[k] = s;
        dof_max = std::max(dof_max, robots[k]->model.dof);
    }
    std::shared_ptr<Module>& mod = lp.mod;
    if (uniq.size() == 1) {
        mod = uniq[0]->module;
    } else {
        std::vector<const RobotModel*> ms;
        for (auto* r : uniq) ms.push_back(&r->model);
        std::vector<int> module_slot;
        mod = get_module(ms, &module_slot);
        for (int k = 0; k < nq; k++) slot_of[k] = module_slot[slot_of[k]];
    }
    // launch shape: explicit parameters, else the tuned shape for this robot set and class, else
    // the built-in defaults
    int W = params->team_waves > 0 ? params->team_waves : kDefaultTeamWaves;
    int waves = params->waves > 0 ? params->waves : default_wave_count;
    int cap = params->teams_per_query;
    int lanes = params->lanes_per_iteration;
    if (params->team_waves == 0 && params->waves == 0 && params->teams_per_query == 0 &&
        params->lanes_per_iteration == 0) {
        const Shape tuned = tuned_shape(robot_set_key(uniq), nq == 1 ? 0 : 1);
        if (tuned.team > 0) {
            W = tuned.team;
            waves = tuned.waves;
            cap = tuned.cap;
            lanes = tuned.lanes;
        }
    }
    if (lanes == 0 || params->max_step > 32) lanes = params->max_step > 32 ? 64 : kDefaultLanes;
    int teams = waves > 0 ? std::max(1, waves / W) : default_teams(mod->fn("mp_plan"), W);
    if (params->exclusive) teams = nq;
    lp.team = W;
    lp.teams = teams;
    lp.cap = cap;
    lp.lanes = lanes;
    return lp;
}

namespace {
void solve_impl(const char* tag, int nq, mp_robot* const* robots, const mp_scene* const* scenes, const float* starts16,
                const float* goals16, const mp_rrtc_params* params, mp_path* out, int default_wave_count) {
    if (!starts16 || !goals16 || !out) throw Error(MP_E_INVALID, "bad batch arguments");
    double t0 = now_ms();
    const LaunchPlan lp = plan_launch(nq, robots, scenes, params, default_wave_count);
    hipStream_t s = lib_stream();
    SlotPool& pool = pool_for(tag, nq, params->tree_capacity, lp.dof_max, params->max_path);
    for (int k = 0; k < nq; k++)
        pool.fill_query(k, lp.slot_of[k], robots[k]->model.dof, scenes[k], starts16 + 16 * k, goals16 + 16 * k,
                        params->seed + (uint64_t)k, params->max_iterations, params->sampler);
    run_to_completion(*lp.mod, pool, nq, *params, lp.team, lp.teams, lp.cap, lp.lanes, s);
    for (int k = 0; k < nq; k++) harvest(pool, k, robots[k]->model.dof, pool.hst_pinned[k], &out[k]);
    double lat = now_ms() - t0;
    for (int k = 0; k < nq; k++) out[k].latency_ms = lat;
}
}  // namespace
}  // namespace mp

extern "C" {

void mp_rrtc_default_params(mp_rrtc_params* p) {
    if (!p) return;
    p->resolution = 0.02f;
    p->max_step = 32;
    p->max_iterations = 10000;
    p->tree_capacity = 65536;
    p->waves = 0;
    p->max_path = 4096;
    p->seed = 0;
    p->exclusive = 0;
    p->team_waves = 0;
    p->teams_per_query = 0;
    p->lanes_per_iteration = 0;
    p->sampler = MP_SAMPLER_UNIFORM;
    p->dynamic_domain = 0;  // pRRTC's settings (pRRTC_settings.hh:13-16) when turned on
    p->dd_radius = 4.0f;
    p->dd_alpha = 1e-4f;
    p->dd_min_radius = 1.0f;
    p->balance = MP_BALANCE_SMALLER;
    p->tree_ratio = 0.0f;
}

int mp_rrtc_solve(mp_robot* robot, const mp_scene* scene, const float* start, const float* goal,
                  const mp_rrtc_params* params, mp_path* out) {
    return mp::guard_call([&] {
        if (!robot || !start || !goal) throw Error(MP_E_INVALID, "null robot/start/goal");
        float s16[16] = {0}, g16[16] = {0};
        std::memcpy(s16, start, robot->model.dof * sizeof(float));
        std::memcpy(g16, goal, robot->model.dof * sizeof(float));
        // plain mode: one query per launch on a CPRRTC-sized grid (512 waves) unless given
        mp::solve_impl("plain", 1, &robot, &scene, s16, g16, params, out, 512);
    });
}

int mp_rrtc_solve_batch(int nq, mp_robot* const* robots, const mp_scene* const* scenes, const float* starts16,
                        const float* goals16, const mp_rrtc_params* params, mp_path* out) {
    return mp::guard_call([&] { mp::solve_impl("batch", nq, robots, scenes, starts16, goals16, params, out, 0); });
}

int mp_rrtc_autotune(int nq, mp_robot* const* robots, const mp_scene* const* scenes, const float* starts16,
                     const float* goals16, const mp_rrtc_params* params, int per_query, int reps,
                     mp_tune_result* results, int cap, int* n_results, mp_rrtc_params* best) {
    return mp::guard_call([&] {
        if (!params || !best || nq < 1 || reps < 1 || !robots || !scenes || !starts16 || !goals16)
            throw Error(MP_E_INVALID, "bad autotune arguments");
        if (params->exclusive) throw Error(MP_E_INVALID, "autotune the shared mode (exclusive = 0)");
        for (int k = 0; k < nq; k++)
            if (!robots[k] || !scenes[k]) throw Error(MP_E_INVALID, "null robot or scene in the sample");
        const std::string key = mp::robot_set_key(std::vector<mp_robot*>(robots, robots + nq));
        const int cls = per_query ? 0 : 1;
        // candidates: waves per team x total waves (0 = every team the device keeps resident) x
        // teams per query (0 = unbounded; single queries only use the grid)
        struct Cand {
            int team, waves, cap, lanes;
        };
        std::vector<Cand> cands;
        std::vector<int> lane_opts{64};
        if (params->max_step <= 32) lane_opts.push_back(32);
        // the resident grid of each team size; a 1024-wave candidate that is the resident grid
        // too (one wave per SIMD on 256 CUs) is the same launch and is not run twice
        std::shared_ptr<mp::Module> mod = mp::plan_launch(nq, robots, scenes, params, 0).mod;
        for (int l : lane_opts)
            for (int t : {1, 2, 4}) {
                if (per_query) {
                    for (int w : {256, 512, 1024, 2048}) cands.push_back({t, w, 0, l});
                } else {
                    const bool same = mp::default_teams(mod->fn("mp_plan"), t) * t == 1024;
                    for (int w : {0, 1024}) {
                        if (w == 1024 && same) continue;
                        for (int c : {0, 16, 32, 64, 128}) cands.push_back({t, w, c, l});
                    }
                }
            }
        std::vector<mp_path> out((size_t)nq);
        int n = 0;
        Cand best_c{0, 0, 0, 0};
        double bms = 1e300;
        // every candidate once per round, rounds = reps (one seed each), so that drifts of the
        // device over the tuning run fall on all candidates alike; the score is the mean time
        std::vector<double> total(cands.size(), 0.0);
        for (int r = 0; r < reps; r++) {
            for (size_t c = 0; c < cands.size(); c++) {
                const Cand& cd = cands[c];
                mp_rrtc_params p = *params;
                p.team_waves = cd.team;
                p.waves = cd.waves;
                p.teams_per_query = cd.cap;
                p.lanes_per_iteration = cd.lanes;
                p.seed = params->seed + 7919ull * (uint64_t)r;
                for (auto& o : out) std::memset(&o, 0, sizeof(o));
                if (per_query) {
                    for (int k = 0; k < nq; k++) {
                        mp::solve_impl("plain", 1, robots + k, scenes + k, starts16 + 16 * k, goals16 + 16 * k, &p,
                                       &out[k], 512);
                        total[c] += mp_last_kernel_ms("plan");
                    }
                } else {
                    mp::solve_impl("batch", nq, robots, scenes, starts16, goals16, &p, out.data(), 0);
                    total[c] += mp_last_kernel_ms("plan");
                }
            }
        }
        for (size_t c = 0; c < cands.size(); c++) {
            const Cand& cd = cands[c];
            const double ms = total[c] / reps;
            if (results && n < cap) results[n] = mp_tune_result{cd.team, cd.waves, cd.cap, cd.lanes, ms};
            n++;
            if (ms < bms) {
                bms = ms;
                best_c = cd;
            }
        }
        if (n_results) *n_results = n;
        mp::store_shape(key, cls, mp::Shape{best_c.team, best_c.waves, best_c.cap, best_c.lanes});
        *best = *params;
        best->team_waves = best_c.team;
        best->waves = best_c.waves;
        best->teams_per_query = best_c.cap;
        best->lanes_per_iteration = best_c.lanes;
    });
}

}  // extern "C"
