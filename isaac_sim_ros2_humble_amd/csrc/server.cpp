// Continuous-batching planning server.
//
// One worker thread owns the device.  Each step it admits queued queries into free tree slots
// (a slot = the fixed tree/path memory of one query, so admission is by free tree memory),
// launches the admission kernel for the new queries and one persistent planning launch with a
// ticket budget (extend-connect iterations) shared by every active query, copies the compact
// query states back, and retires the queries that finished.  Queries that arrive while others
// are on the device join at the next step; finished ones free their slot immediately.  The
// step (control-block copy, admission kernel, planning kernel, state copy) can be captured once
// in a hipGraph and replayed.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

#include "mp_plan.h"

using mp::Error;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Pending {
    int64_t ticket;
    int robot;
    const mp_scene* scene;
    float start[16];
    float goal[16];
    uint64_t seed;
    double t_submit;
};

struct Done {
    mp_path info;
    std::vector<float> waypoints;
};

struct ActiveQ {
    int64_t ticket;
    double t_submit;
    int dof;
    const mp_scene* scene;
};

// A scene stays pinned (mp_scene_destroy / mp_scene_update refuse it) from the submit of a query
// on it until that query is retired or the server is destroyed.
void scene_release(const mp_scene* sc) { sc->server_refs.fetch_sub(1, std::memory_order_acq_rel); }

}  // namespace

struct mp_server {
    mp_server_params prm;
    std::vector<mp_robot*> robots;
    std::vector<int> robot_slot;  // robot index -> the module's robot slot
    std::shared_ptr<mp::Module> mod;
    std::unique_ptr<mp::SlotPool> pool;
    int nslots = 0, teams = 0, team = 1, cap = 0, lanes = 64, budget = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;

    std::thread worker;
    std::mutex mu;
    std::condition_variable cv, idle_cv;
    bool stop = false;
    std::string error;

    int64_t next_ticket = 1;
    std::deque<Pending> queue;
    std::map<int, ActiveQ> active;  // slot -> query
    std::vector<int> free_slots;
    std::unordered_map<int64_t, Done> done;

    // stats
    int64_t submitted = 0, completed = 0, solved = 0, steps = 0;
    std::vector<double> latencies;
    double t_first = -1.0, t_last = 0.0;

    // worker-thread time in step() (enqueue and device) and in whole steps (admission, step,
    // retiring); written by the worker only
    std::atomic<double> busy_ms{0.0}, loop_ms{0.0};
    std::atomic<int64_t> backlog_steps{0};  // steps that ran with queries waiting for a slot

    void run();
    void step(const std::vector<int>& list, int n_new);
};

// A step's ticket budget (one team pass per ticket) grows with the number of active queries:
// kTicketsPerActive per query, at least the base budget.  With few queries active a step stays
// short (a new query waits at most one short step for admission); with many, each step gives
// every query a few passes, so the step's fixed costs (launch, copies, synchronisation, the
// wait for the last pass, retiring) are spread over more work.  Measured on the served
// workload at saturation: 4 / 8 / 16 / 32 / 64 tickets per active query gave 51.7 / 54.1 /
// 65.8 / 77.4 / 87.4 K queries/s with p50 2.1 / 1.8 / 1.9 / 3.0 / 22.6 ms; the latencies at
// 1000 and 5000 queries/s did not change.
constexpr int kTicketsPerActive = 32;

void mp_server::step(const std::vector<int>& list, int n_new) {
    int* h = pool->hpinned;
    h[0] = 0;
    h[1] = (int)list.size();
    h[2] = (int)list.size();
    // (exclusive mode counts the budget per team, one query each: it stays as configured)
    h[3] = prm.params.exclusive
               ? budget
               : (int)std::min<long long>(std::max<long long>(budget, (long long)kTicketsPerActive * (long long)list.size()),
                                          1 << 30);
    h[4] = (int)list.size() - n_new;
    h[5] = n_new;
    h[6] = prm.params.exclusive ? 1 : 0;
    h[7] = 1;  // one iteration per ticket: the step's budget stays a budget of iterations
    h[8] = cap;
    h[9] = mp::pair_flag(lanes);
    for (size_t k = 0; k < list.size(); k++) h[16 + k] = list[k];
    if (n_new > 0) {
        // the new queries' descriptors in one copy of the slot range that holds them (the
        // descriptors of running queries in that range are rewritten with the same bytes)
        int lo = nslots, hi = -1;
        for (int k = (int)list.size() - n_new; k < (int)list.size(); k++) {
            lo = std::min(lo, list[k]);
            hi = std::max(hi, list[k]);
        }
        MP_HIP(hipMemcpyAsync(pool->dq + lo, &pool->hq[lo], (size_t)(hi - lo + 1) * sizeof(mp::QueryDevHost),
                              hipMemcpyHostToDevice, stream));
    }
    if (prm.use_graph) {
        if (!graph) {
            mp::PlanCtlHost ctl = pool->ctl(prm.params);
            ctl.window = mp::balance_window(teams, cap, lanes, prm.params.exclusive != 0);
            void* args[] = {&ctl};
            MP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            MP_HIP(hipMemcpyAsync(pool->dctl, h, 16 * sizeof(int), hipMemcpyHostToDevice, stream));
            MP_HIP(hipMemcpyAsync(pool->dactive, h + 16, (size_t)nslots * sizeof(int), hipMemcpyHostToDevice, stream));
            MP_HIP(hipModuleLaunchKernel(mod->fn("mp_plan_init"), nslots, 1, 1, mp::kInitBlock, 1, 1, 0, stream, args,
                                         nullptr));
            MP_HIP(hipModuleLaunchKernel(mod->fn("mp_plan"), teams, 1, 1, 64 * team, 1, 1, 0, stream, args, nullptr));
            MP_HIP(hipMemcpyAsync(pool->hst_pinned, pool->dst, (size_t)nslots * sizeof(mp::QueryStateHost),
                                  hipMemcpyDeviceToHost, stream));
            MP_HIP(hipStreamEndCapture(stream, &graph_def));
            MP_HIP(hipGraphInstantiate(&graph, graph_def, nullptr, nullptr, 0));
        }
        MP_HIP(hipGraphLaunch(graph, stream));
    } else {
        mp::PlanCtlHost ctl = pool->ctl(prm.params);
        ctl.window = mp::balance_window(teams, cap, lanes, prm.params.exclusive != 0);
        void* args[] = {&ctl};
        MP_HIP(hipMemcpyAsync(pool->dctl, h, 16 * sizeof(int), hipMemcpyHostToDevice, stream));
        MP_HIP(hipMemcpyAsync(pool->dactive, h + 16, list.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        if (n_new > 0) mp::launch(mod->fn("mp_plan_init"), (unsigned)n_new, mp::kInitBlock, stream, args, "plan_init");
        mp::launch(mod->fn("mp_plan"), (unsigned)teams, 64u * team, stream, args, "plan");
        MP_HIP(hipMemcpyAsync(pool->hst_pinned, pool->dst, (size_t)nslots * sizeof(mp::QueryStateHost),
                              hipMemcpyDeviceToHost, stream));
    }
    MP_HIP(hipStreamSynchronize(stream));
}

void mp_server::run() {
    try {
        std::vector<int> list;
        std::vector<int> fresh;
        std::vector<char> is_fresh;
        double t_loop = 0.0;
        while (true) {
            bool backlog = false;
            {
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [&] { return stop || !queue.empty() || !active.empty(); });
                if (stop) break;
                t_loop = now_ms();
                fresh.clear();
                while (!queue.empty() && !free_slots.empty()) {
                    Pending p = queue.front();
                    queue.pop_front();
                    int slot = free_slots.back();
                    free_slots.pop_back();
                    mp_robot* r = robots[p.robot];
                    pool->fill_query(slot, robot_slot[p.robot], r->model.dof, p.scene, p.start, p.goal, p.seed, prm.params.max_iterations,
                                     prm.params.sampler);
                    active[slot] = ActiveQ{p.ticket, p.t_submit, r->model.dof, p.scene};
                    fresh.push_back(slot);
                }
                backlog = !queue.empty();
                list.clear();
                is_fresh.assign((size_t)nslots, 0);
                for (int s : fresh) is_fresh[s] = 1;
                for (auto& kv : active)
                    if (!is_fresh[kv.first]) list.push_back(kv.first);
                for (int s : fresh) list.push_back(s);
            }
            if (list.empty()) continue;
            const double t0 = now_ms();
            step(list, (int)fresh.size());
            if (backlog) backlog_steps.fetch_add(1, std::memory_order_relaxed);
            const double t1 = now_ms();
            busy_ms.store(busy_ms.load(std::memory_order_relaxed) + (t1 - t0), std::memory_order_relaxed);
            steps++;
            // retire finished queries
            std::vector<std::pair<int64_t, Done>> fin;
            std::vector<int> freed;
            for (int slot : list) {
                const mp::QueryStateHost& st = pool->hst_pinned[slot];
                if (st.status == MP_PLAN_RUNNING || st.status < 0) continue;
                ActiveQ aq;
                {
                    std::lock_guard<std::mutex> g(mu);
                    aq = active[slot];
                }
                Done d;
                std::memset(&d.info, 0, sizeof(d.info));
                const int need = st.status == MP_PLAN_SOLVED && st.path_len > 0 ? st.path_len : 0;
                d.waypoints.resize((size_t)need * aq.dof);
                d.info.waypoints = d.waypoints.data();
                d.info.capacity = need;
                mp::harvest(*pool, slot, aq.dof, st, &d.info);
                fin.emplace_back(aq.ticket, std::move(d));
                freed.push_back(slot);
            }
            double t = now_ms();
            {
                std::lock_guard<std::mutex> g(mu);
                for (size_t k = 0; k < fin.size(); k++) {
                    int slot = freed[k];
                    double lat = t - active[slot].t_submit;
                    fin[k].second.info.latency_ms = lat;
                    fin[k].second.waypoints.resize((size_t)std::max(fin[k].second.info.num_waypoints, 0) * active[slot].dof);
                    fin[k].second.info.waypoints = nullptr;
                    latencies.push_back(lat);
                    completed++;
                    if (fin[k].second.info.status == MP_PLAN_SOLVED) solved++;
                    t_last = t;
                    done.emplace(fin[k].first, std::move(fin[k].second));
                    scene_release(active[slot].scene);
                    active.erase(slot);
                    free_slots.push_back(slot);
                }
            }
            idle_cv.notify_all();
            loop_ms.store(loop_ms.load(std::memory_order_relaxed) + (now_ms() - t_loop), std::memory_order_relaxed);
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> g(mu);
        error = e.what();
        stop = true;
    }
    idle_cv.notify_all();
}

extern "C" {

int mp_server_create(const mp_server_params* params, mp_server** out) {
    return mp::guard_call([&] {
        if (mp::lib_stream() == nullptr) throw Error(MP_E_STATE, "mp_init() has not been called");
        if (!params || !out || params->num_robots < 1 || !params->robots) throw Error(MP_E_INVALID, "bad server params");
        if (params->tree_capacity < 2 || params->max_active < 1) throw Error(MP_E_INVALID, "bad server capacities");
        const mp_rrtc_params& p = params->params;
        if (!(p.resolution > 0.0f) || p.max_step < 1 || p.max_step > 64 || p.max_iterations < 1 || p.max_path < 2)
            throw Error(MP_E_INVALID, "bad planner params");
        auto s = std::make_unique<mp_server>();
        s->prm = *params;
        s->robots.assign(params->robots, params->robots + params->num_robots);
        int dof_max = 1;
        std::vector<const mp::RobotModel*> ms;
        for (auto* r : s->robots) {
            if (!r) throw Error(MP_E_INVALID, "null robot");
            ms.push_back(&r->model);
            dof_max = std::max(dof_max, r->model.dof);
        }
        if (ms.size() == 1) {
            s->mod = s->robots[0]->module;
            s->robot_slot.assign(1, 0);
        } else {
            s->mod = mp::get_module(ms, &s->robot_slot);
        }
        size_t per = mp::slot_bytes(params->tree_capacity, dof_max);
        size_t fit = params->tree_memory_bytes / per;
        s->nslots = (int)std::min<size_t>((size_t)params->max_active, fit);
        if (s->nslots < 1) throw Error(MP_E_CAPACITY, "tree memory too small for one query");
        s->pool = std::make_unique<mp::SlotPool>(s->nslots, params->tree_capacity, dof_max, p.max_path);
        if (p.team_waves != 0 && p.team_waves != 1 && p.team_waves != 2 && p.team_waves != 4)
            throw Error(MP_E_INVALID, "team_waves must be 0, 1, 2 or 4");
        if (p.teams_per_query < 0) throw Error(MP_E_INVALID, "teams_per_query must be >= 0");
        if (p.lanes_per_iteration != 0 && p.lanes_per_iteration != 32 && p.lanes_per_iteration != 64)
            throw Error(MP_E_INVALID, "lanes_per_iteration must be 0, 32 or 64");
        if (p.lanes_per_iteration == 32 && p.max_step > 32)
            throw Error(MP_E_INVALID, "lanes_per_iteration 32 needs max_step <= 32");
        if (p.sampler != MP_SAMPLER_UNIFORM && p.sampler != MP_SAMPLER_HALTON)
            throw Error(MP_E_INVALID, "sampler must be MP_SAMPLER_UNIFORM or MP_SAMPLER_HALTON");
        mp::check_domain(p);
        s->team = mp::team_size(p);
        s->cap = p.teams_per_query;
        s->lanes = p.lanes_per_iteration;
        int waves = params->waves;
        if (p.team_waves == 0 && params->waves == 0 && p.teams_per_query == 0 &&
            p.lanes_per_iteration == 0) {  // the tuned batch shape
            const mp::Shape tuned = mp::tuned_shape(mp::robot_set_key(s->robots), 1);
            if (tuned.team > 0) {
                s->team = tuned.team;
                waves = tuned.waves;
                s->cap = tuned.cap;
                s->lanes = tuned.lanes;
            }
        }
        if (s->lanes == 0 || p.max_step > 32) s->lanes = p.max_step > 32 ? 64 : mp::kDefaultLanes;
        s->teams = waves > 0 ? std::max(1, waves / s->team) : mp::default_teams(s->mod->fn("mp_plan"), s->team);
        if (p.exclusive) s->teams = s->nslots;
        s->budget = params->step_iterations > 0 ? params->step_iterations : 2 * s->teams;
        MP_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        for (int k = s->nslots - 1; k >= 0; k--) s->free_slots.push_back(k);
        mp_server* raw = s.release();
        raw->worker = std::thread([raw] { raw->run(); });
        *out = raw;
    });
}

int mp_server_submit(mp_server* s, int robot_index, const mp_scene* scene, const float* start, const float* goal,
                     uint64_t seed, int64_t* ticket) {
    return mp::guard_call([&] {
        if (!s || !scene || !start || !goal || !ticket) throw Error(MP_E_INVALID, "bad submit arguments");
        if (robot_index < 0 || robot_index >= (int)s->robots.size()) throw Error(MP_E_INVALID, "robot index out of range");
        Pending p;
        std::memset(&p, 0, sizeof(p));
        p.robot = robot_index;
        p.scene = scene;
        int dof = s->robots[robot_index]->model.dof;
        std::memcpy(p.start, start, dof * sizeof(float));
        std::memcpy(p.goal, goal, dof * sizeof(float));
        p.seed = seed;
        p.t_submit = now_ms();
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (!s->error.empty()) throw Error(MP_E_STATE, "server stopped: " + s->error);
            scene->server_refs.fetch_add(1, std::memory_order_acq_rel);
            p.ticket = s->next_ticket++;
            s->queue.push_back(p);
            s->submitted++;
            if (s->t_first < 0) s->t_first = p.t_submit;
        }
        s->cv.notify_one();
        *ticket = p.ticket;
    });
}

int mp_server_poll(mp_server* s, int64_t ticket, mp_path* out) {
    int ready = 0;
    int rc = mp::guard_call([&] {
        if (!s || !out) throw Error(MP_E_INVALID, "bad poll arguments");
        std::lock_guard<std::mutex> g(s->mu);
        if (!s->error.empty()) throw Error(MP_E_STATE, "server stopped: " + s->error);
        auto it = s->done.find(ticket);
        if (it == s->done.end()) return;
        const Done& d = it->second;
        float* wp = out->waypoints;
        int cap = out->capacity;
        *out = d.info;
        out->waypoints = wp;
        out->capacity = cap;
        // the buffer holds `capacity` waypoints of the ticket's robot; a longer path is reported
        // with its length and not copied
        const int n = d.info.num_waypoints;
        if (d.info.status == MP_PLAN_SOLVED && (n > cap || !wp)) {
            out->status = MP_PLAN_PATH_TOO_LONG;
        } else if (d.info.status == MP_PLAN_SOLVED && n > 0) {
            std::memcpy(wp, d.waypoints.data(), d.waypoints.size() * sizeof(float));
        }
        s->done.erase(it);
        ready = 1;
    });
    return rc == MP_OK ? ready : rc;
}

int mp_server_wait_idle(mp_server* s, double timeout_s) {
    return mp::guard_call([&] {
        if (!s) throw Error(MP_E_INVALID, "null server");
        std::unique_lock<std::mutex> g(s->mu);
        bool ok = s->idle_cv.wait_for(g, std::chrono::duration<double>(timeout_s), [&] {
            return !s->error.empty() || (s->queue.empty() && s->active.empty());
        });
        if (!s->error.empty()) throw Error(MP_E_STATE, "server stopped: " + s->error);
        if (!ok) throw Error(MP_E_STATE, "timeout waiting for the server to drain");
    });
}

int mp_server_stats_get(mp_server* s, mp_server_stats* out) {
    return mp::guard_call([&] {
        if (!s || !out) throw Error(MP_E_INVALID, "bad arguments");
        std::lock_guard<std::mutex> g(s->mu);
        std::memset(out, 0, sizeof(*out));
        out->submitted = s->submitted;
        out->completed = s->completed;
        out->solved = s->solved;
        out->active = (int)s->active.size();
        out->queued = (int)s->queue.size();
        out->steps = s->steps;
        out->step_ms = s->busy_ms.load(std::memory_order_relaxed);
        out->loop_ms = s->loop_ms.load(std::memory_order_relaxed);
        out->backlog_steps = s->backlog_steps.load(std::memory_order_relaxed);
        if (!s->latencies.empty()) {
            std::vector<double> v = s->latencies;
            std::sort(v.begin(), v.end());
            auto pct = [&](double q) { return v[std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5))]; };
            out->p50_ms = pct(0.50);
            out->p99_ms = pct(0.99);
            double sum = 0;
            for (double x : v) sum += x;
            out->mean_ms = sum / v.size();
            double span = (s->t_last - s->t_first) / 1000.0;
            out->throughput_qps = span > 0 ? s->completed / span : 0.0;
        }
    });
}

int mp_server_destroy(mp_server* s) {
    return mp::guard_call([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv.notify_all();
        if (s->worker.joinable()) s->worker.join();
        if (s->stream) hipStreamSynchronize(s->stream);
        for (const Pending& p : s->queue) scene_release(p.scene);
        for (auto& kv : s->active) scene_release(kv.second.scene);
        if (s->graph) hipGraphExecDestroy(s->graph);
        if (s->graph_def) hipGraphDestroy(s->graph_def);
        s->pool.reset();
        if (s->stream) hipStreamDestroy(s->stream);
        delete s;
    });
}

}  // extern "C"
