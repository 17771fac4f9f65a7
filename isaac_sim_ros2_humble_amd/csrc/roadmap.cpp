// Device-resident roadmap planner (the extend-connect loop of CUDAMPLib::RRG::solve,
// cudampl/CUDAMPLib/src/planners/RRG.cpp:241-424): the roadmap's vertices, group members and
// edges live in device memory; every loop pass is three launches on the library stream (sample,
// motions, commit) with no host round trip.  The host keeps a few passes in flight, watches a
// host-mapped copy of the loop state for the solution, and grows the buffers before a pass could
// run out of room.
#include <algorithm>
#include <cstring>

#include "mp_plan.h"

using mp::Error;

namespace {

constexpr int kNctl = 16;      // MP_RM_NCTL
constexpr int kMaxK = 256;     // MP_RM_MAX_K
constexpr int kDepth = 3;      // passes in flight
enum { T = 0, SOLVED = 1, NV = 2, NE = 3, G1 = 4, G2 = 5, PASS = 6, FULL = 7, TRAVEL1 = 8, TRAVEL2 = 9 };

// Host mirror of the kernels' RoadmapDev (layout must match the kernel template).
struct RoadmapDevHost {
    float* node;
    int* vgroup;
    float* gq[2];
    int* gv[2];
    int* ea;
    int* eb;
    float* ew;
    float* smp;
    int* nn;
    unsigned char* ok;
    float* len;
    unsigned char* freed;
    int* ctl;
    int* mirror;
    uint64_t seed, offset;
    int k, vcap, ecap, dof;
    float res, max_travel;
};

int f2i(float f) {
    int i;
    std::memcpy(&i, &f, 4);
    return i;
}
float i2f(int i) {
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

}  // namespace

struct mp_roadmap {
    mp_robot* robot = nullptr;
    const mp_scene* scene = nullptr;
    RoadmapDevHost g{};
    int* mirror_h = nullptr;   // host-mapped loop state
    hipEvent_t ev[kDepth] = {};
    long long enqueued = 0, completed = 0;  // passes enqueued / known finished
    hipFunction_t f_sample = nullptr, f_edges = nullptr, f_commit = nullptr;

    void alloc(int vcap, int ecap);
    void release();
    void grow(int need_v, int need_e);
    void wait_all();
};

void mp_roadmap::alloc(int vcap, int ecap) {
    const int dof = g.dof;
    g.vcap = vcap;
    g.ecap = ecap;
    MP_HIP(hipMalloc(&g.node, (size_t)dof * vcap * sizeof(float)));
    MP_HIP(hipMalloc(&g.vgroup, (size_t)vcap * sizeof(int)));
    for (int s = 0; s < 2; s++) {
        MP_HIP(hipMalloc(&g.gq[s], (size_t)dof * vcap * sizeof(float)));
        MP_HIP(hipMalloc(&g.gv[s], (size_t)vcap * sizeof(int)));
    }
    MP_HIP(hipMalloc(&g.ea, (size_t)ecap * sizeof(int)));
    MP_HIP(hipMalloc(&g.eb, (size_t)ecap * sizeof(int)));
    MP_HIP(hipMalloc(&g.ew, (size_t)ecap * sizeof(float)));
}

void mp_roadmap::release() {
    void* ps[] = {g.node, g.vgroup, g.gq[0], g.gq[1], g.gv[0], g.gv[1], g.ea, g.eb, g.ew};
    for (void* p : ps)
        if (p) hipFree(p);
    g.node = nullptr;
    g.vgroup = g.gv[0] = g.gv[1] = g.ea = g.eb = nullptr;
    g.gq[0] = g.gq[1] = g.ew = nullptr;
}

void mp_roadmap::wait_all() {
    MP_HIP(hipStreamSynchronize(mp::lib_stream()));
    completed = enqueued;
}

// Doubles the capacities until need_v vertices and need_e edges fit, copying the roadmap (the
// structure-of-arrays rows move to the new stride).
void mp_roadmap::grow(int need_v, int need_e) {
    int vcap = g.vcap, ecap = g.ecap;
    while (vcap < need_v) vcap *= 2;
    while (ecap < need_e) ecap *= 2;
    if (vcap == g.vcap && ecap == g.ecap) return;
    wait_all();
    hipStream_t s = mp::lib_stream();
    int ctl[kNctl];
    MP_HIP(hipMemcpy(ctl, g.ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    RoadmapDevHost old = g;
    g.node = nullptr;
    g.vgroup = g.gv[0] = g.gv[1] = g.ea = g.eb = nullptr;
    g.gq[0] = g.gq[1] = g.ew = nullptr;
    alloc(vcap, ecap);
    const int dof = g.dof, nv = ctl[NV], ne = ctl[NE], n1 = ctl[G1], n2 = ctl[G2];
    for (int j = 0; j < dof; j++) {
        MP_HIP(hipMemcpyAsync(g.node + (size_t)j * vcap, old.node + (size_t)j * old.vcap, (size_t)nv * 4,
                              hipMemcpyDeviceToDevice, s));
        MP_HIP(hipMemcpyAsync(g.gq[0] + (size_t)j * vcap, old.gq[0] + (size_t)j * old.vcap, (size_t)n1 * 4,
                              hipMemcpyDeviceToDevice, s));
        MP_HIP(hipMemcpyAsync(g.gq[1] + (size_t)j * vcap, old.gq[1] + (size_t)j * old.vcap, (size_t)n2 * 4,
                              hipMemcpyDeviceToDevice, s));
    }
    MP_HIP(hipMemcpyAsync(g.vgroup, old.vgroup, (size_t)nv * 4, hipMemcpyDeviceToDevice, s));
    MP_HIP(hipMemcpyAsync(g.gv[0], old.gv[0], (size_t)n1 * 4, hipMemcpyDeviceToDevice, s));
    MP_HIP(hipMemcpyAsync(g.gv[1], old.gv[1], (size_t)n2 * 4, hipMemcpyDeviceToDevice, s));
    MP_HIP(hipMemcpyAsync(g.ea, old.ea, (size_t)ne * 4, hipMemcpyDeviceToDevice, s));
    MP_HIP(hipMemcpyAsync(g.eb, old.eb, (size_t)ne * 4, hipMemcpyDeviceToDevice, s));
    MP_HIP(hipMemcpyAsync(g.ew, old.ew, (size_t)ne * 4, hipMemcpyDeviceToDevice, s));
    MP_HIP(hipStreamSynchronize(s));
    void* ps[] = {old.node, old.vgroup, old.gq[0], old.gq[1], old.gv[0], old.gv[1], old.ea, old.eb, old.ew};
    for (void* p : ps) hipFree(p);
}

extern "C" {

int mp_roadmap_create(mp_robot* robot, const mp_scene* scene, const float* starts, int ns, const float* goals, int ng,
                      int samples_per_iteration, float max_travel_distance, float resolution, uint64_t seed,
                      uint64_t sample_offset, mp_roadmap** out) {
    return mp::guard_call([&] {
        if (mp::lib_stream() == nullptr) throw Error(MP_E_STATE, "mp_init() has not been called");
        if (!robot || !scene || !out || ns < 1 || ng < 1 || !starts || !goals) throw Error(MP_E_INVALID, "bad roadmap arguments");
        if (samples_per_iteration < 1 || samples_per_iteration > kMaxK)
            throw Error(MP_E_INVALID, "samples per iteration must be in [1, 256]");
        if (!(max_travel_distance > 0.0f) || !(resolution > 0.0f)) throw Error(MP_E_INVALID, "bad travel distance or resolution");
        auto rm = std::make_unique<mp_roadmap>();
        rm->robot = robot;
        rm->scene = scene;
        const int dof = robot->model.dof;
        RoadmapDevHost& g = rm->g;
        g.dof = dof;
        g.k = samples_per_iteration;
        g.seed = seed;
        g.offset = sample_offset;
        g.res = resolution;
        g.max_travel = max_travel_distance;
        const int nv0 = ns + ng;
        int vcap = 4096;
        while (vcap < nv0 + kDepth * 2 * g.k) vcap *= 2;
        try {
            rm->alloc(vcap, 2 * vcap);
            MP_HIP(hipMalloc(&g.smp, (size_t)g.k * dof * sizeof(float)));
            MP_HIP(hipMalloc(&g.nn, (size_t)3 * g.k * sizeof(int)));
            MP_HIP(hipMalloc(&g.ok, (size_t)2 * g.k));
            MP_HIP(hipMalloc(&g.len, (size_t)2 * g.k * sizeof(float)));
            MP_HIP(hipMalloc(&g.freed, (size_t)g.k));
            MP_HIP(hipMalloc(&g.ctl, kNctl * sizeof(int)));
            MP_HIP(hipHostMalloc(&rm->mirror_h, kNctl * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
            MP_HIP(hipHostGetDevicePointer((void**)&g.mirror, rm->mirror_h, 0));
            for (auto& e : rm->ev) MP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            // the start and goal states: vertices, and the first members of their groups
            std::vector<float> soa((size_t)dof * nv0);
            std::vector<int> grp(nv0), ids(nv0);
            for (int v = 0; v < nv0; v++) {
                const float* q = v < ns ? starts + (size_t)v * dof : goals + (size_t)(v - ns) * dof;
                for (int j = 0; j < dof; j++) soa[(size_t)j * nv0 + v] = q[j];
                grp[v] = v < ns ? 1 : 2;
                ids[v] = v;
            }
            hipStream_t s = mp::lib_stream();
            for (int j = 0; j < dof; j++) {
                MP_HIP(hipMemcpyAsync(g.node + (size_t)j * vcap, &soa[(size_t)j * nv0], (size_t)nv0 * 4, hipMemcpyHostToDevice, s));
                MP_HIP(hipMemcpyAsync(g.gq[0] + (size_t)j * vcap, &soa[(size_t)j * nv0], (size_t)ns * 4, hipMemcpyHostToDevice, s));
                MP_HIP(hipMemcpyAsync(g.gq[1] + (size_t)j * vcap, &soa[(size_t)j * nv0 + ns], (size_t)ng * 4,
                                      hipMemcpyHostToDevice, s));
            }
            MP_HIP(hipMemcpyAsync(g.vgroup, grp.data(), (size_t)nv0 * 4, hipMemcpyHostToDevice, s));
            MP_HIP(hipMemcpyAsync(g.gv[0], ids.data(), (size_t)ns * 4, hipMemcpyHostToDevice, s));
            MP_HIP(hipMemcpyAsync(g.gv[1], ids.data() + ns, (size_t)ng * 4, hipMemcpyHostToDevice, s));
            int ctl[kNctl] = {0};
            ctl[NV] = nv0;
            ctl[G1] = ns;
            ctl[G2] = ng;
            ctl[TRAVEL1] = ctl[TRAVEL2] = f2i(max_travel_distance);
            MP_HIP(hipMemcpyAsync(g.ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice, s));
            std::memcpy(rm->mirror_h, ctl, sizeof(ctl));
            MP_HIP(hipStreamSynchronize(s));
            mp::Module& mod = *robot->module;
            rm->f_sample = mod.fn("mp_rm_sample_0");
            rm->f_edges = mod.fn("mp_rm_edges_0");
            rm->f_commit = mod.fn("mp_rm_commit_0");
        } catch (...) {
            mp_roadmap_destroy(rm.release());
            throw;
        }
        *out = rm.release();
    });
}

int mp_roadmap_iterate(mp_roadmap* rm, int* solved) {
    return mp::guard_call([&] {
        if (!rm || !solved) throw Error(MP_E_INVALID, "bad arguments");
        RoadmapDevHost& g = rm->g;
        hipStream_t s = mp::lib_stream();
        if (rm->enqueued - rm->completed >= kDepth) {  // keep at most kDepth passes in flight
            MP_HIP(hipEventSynchronize(rm->ev[rm->completed % kDepth]));
            rm->completed++;
        }
        volatile int* m = rm->mirror_h;
        if (m[SOLVED]) {
            *solved = 1;
            return;
        }
        if (m[FULL]) {  // a pass found no room: grow, then run it again
            rm->wait_all();
            rm->grow(2 * g.vcap, 2 * g.ecap);
            const int zero = 0;
            MP_HIP(hipMemcpy(g.ctl + FULL, &zero, sizeof(int), hipMemcpyHostToDevice));
            m[FULL] = 0;
        }
        // the mirror holds at least every finished pass; the passes in flight and this one add
        // at most k vertices, 2k edges and k members each
        const int pending = (int)(rm->enqueued - rm->completed) + 1;
        const int need_v = std::max(m[NV], std::max(m[G1], m[G2])) + pending * g.k;
        const int need_e = m[NE] + pending * 2 * g.k;
        if (need_v > g.vcap || need_e > g.ecap) rm->grow(need_v, need_e);
        mp::SceneDevHost sv = mp::scene_view(rm->scene);
        unsigned long long* work = mp::lib_work_counter();
        void* a1[] = {&g, &sv};
        void* a2[] = {&g, &sv, &work};
        void* a3[] = {&g};
        mp::launch(rm->f_sample, 1, kMaxK, s, a1, nullptr);
        mp::launch(rm->f_edges, (unsigned)((2 * g.k + 3) / 4), 256, s, a2, "rrg");
        mp::launch(rm->f_commit, 1, 64, s, a3, nullptr);
        MP_HIP(hipEventRecord(rm->ev[rm->enqueued % kDepth], s));
        rm->enqueued++;
        *solved = 0;
    });
}

int mp_roadmap_info_get(mp_roadmap* rm, mp_roadmap_info* info) {
    return mp::guard_call([&] {
        if (!rm || !info) throw Error(MP_E_INVALID, "bad arguments");
        rm->wait_all();
        int c[kNctl];
        MP_HIP(hipMemcpy(c, rm->g.ctl, sizeof(c), hipMemcpyDeviceToHost));
        info->solved = c[SOLVED];
        info->passes = c[PASS];
        info->loop_index = c[T];
        info->vertices = c[NV];
        info->edges = c[NE];
        info->start_group = c[G1];
        info->goal_group = c[G2];
        info->travel_start = i2f(c[TRAVEL1]);
        info->travel_goal = i2f(c[TRAVEL2]);
    });
}

int mp_roadmap_read(mp_roadmap* rm, float* vertices, int* groups, int* edges, float* weights, int* start_members,
                    int* goal_members) {
    return mp::guard_call([&] {
        if (!rm) throw Error(MP_E_INVALID, "null roadmap");
        rm->wait_all();
        const RoadmapDevHost& g = rm->g;
        int c[kNctl];
        MP_HIP(hipMemcpy(c, g.ctl, sizeof(c), hipMemcpyDeviceToHost));
        const int nv = c[NV], ne = c[NE], dof = g.dof;
        if (vertices && nv > 0) {
            std::vector<float> soa((size_t)dof * nv);
            for (int j = 0; j < dof; j++)
                MP_HIP(hipMemcpy(&soa[(size_t)j * nv], g.node + (size_t)j * g.vcap, (size_t)nv * 4, hipMemcpyDeviceToHost));
            for (int v = 0; v < nv; v++)
                for (int j = 0; j < dof; j++) vertices[(size_t)v * dof + j] = soa[(size_t)j * nv + v];
        }
        if (groups && nv > 0) MP_HIP(hipMemcpy(groups, g.vgroup, (size_t)nv * 4, hipMemcpyDeviceToHost));
        if (edges && ne > 0) {
            std::vector<int> a(ne), b(ne);
            MP_HIP(hipMemcpy(a.data(), g.ea, (size_t)ne * 4, hipMemcpyDeviceToHost));
            MP_HIP(hipMemcpy(b.data(), g.eb, (size_t)ne * 4, hipMemcpyDeviceToHost));
            for (int e = 0; e < ne; e++) {
                edges[2 * e] = a[e];
                edges[2 * e + 1] = b[e];
            }
        }
        if (weights && ne > 0) MP_HIP(hipMemcpy(weights, g.ew, (size_t)ne * 4, hipMemcpyDeviceToHost));
        if (start_members && c[G1] > 0) MP_HIP(hipMemcpy(start_members, g.gv[0], (size_t)c[G1] * 4, hipMemcpyDeviceToHost));
        if (goal_members && c[G2] > 0) MP_HIP(hipMemcpy(goal_members, g.gv[1], (size_t)c[G2] * 4, hipMemcpyDeviceToHost));
    });
}

int mp_roadmap_destroy(mp_roadmap* rm) {
    return mp::guard_call([&] {
        if (!rm) return;
        if (mp::lib_stream()) hipStreamSynchronize(mp::lib_stream());
        rm->release();
        void* ps[] = {rm->g.smp, rm->g.nn, rm->g.ok, rm->g.len, rm->g.freed, rm->g.ctl};
        for (void* p : ps)
            if (p) hipFree(p);
        if (rm->mirror_h) hipHostFree(rm->mirror_h);
        for (auto& e : rm->ev)
            if (e) hipEventDestroy(e);
        delete rm;
    });
}

}  // extern "C"
