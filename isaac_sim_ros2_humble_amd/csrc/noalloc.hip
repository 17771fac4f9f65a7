// mp_rrtc_solve_noalloc: the plain one-query solve (mp_rrtc_solve) with every buffer its kernels
// use beyond the caller's inputs and outputs -- the two trees with their parents and domain
// radii, the query descriptor and state, the active list, the control block and the path -- in a
// file-scope device array instead of library allocations.  For callers that cannot allocate
// device memory (code under stream capture, tight serving loops, harnesses that forbid
// allocation).  The planning kernels are the same hipRTC-compiled ones; they take addresses in
// the array as arguments, as they take the slot pool's buffers.
//
// Built with hipcc for gfx950 and linked into the host library: the array lives in this source's
// code object, one copy per device, held until the process ends.  Its per-device address is
// looked up once per device at mp_init with that device current and used for every kernel
// argument, memset and copy.  Every call on a device shares the device's array: a call clears the
// parts that must start empty and the library serialises calls; a caller that ran the planning
// kernels on the same array from elsewhere concurrently would race.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>

#include "mp_plan.h"

namespace {

constexpr size_t kScratchBytes = MP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;                           // serialises calls and the address table
unsigned char* g_ptr[kMaxDevices] = {};    // per-device address of `scratch`

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Offsets of one query's buffers in the array.
struct Layout {
    size_t ctl = 0, active = 256, q = 512, st = 1024, trees = 1280, parents = 0, path = 0, end = 0;
};
static_assert(sizeof(mp::QueryDevHost) <= 1024 - 512 && sizeof(mp::QueryStateHost) <= 1280 - 1024,
              "descriptor slots of the scratch layout");
Layout layout(int cap, int dof, int max_path) {
    Layout L;
    L.parents = up256(L.trees + (size_t)2 * dof * cap * sizeof(float));
    L.path = up256(L.parents + (size_t)2 * 2 * cap * sizeof(int));  // parents, then radii, per tree
    L.end = L.path + (size_t)max_path * dof * sizeof(float);
    return L;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

namespace mp {

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(MP_E_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ptr[device]) return;
    void* p = nullptr;
    MP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_ptr[device] = static_cast<unsigned char*>(p);
}

}  // namespace mp

extern "C" int mp_rrtc_solve_noalloc(mp_robot* robot, const mp_scene* scene, const float* start, const float* goal,
                                     const mp_rrtc_params* params, mp_path* out) {
    using mp::Error;
    return mp::guard_call([&] {
        if (!robot || !scene || !start || !goal || !params || !out) throw Error(MP_E_INVALID, "null argument");
        const double t0 = now_ms();
        const mp::LaunchPlan lp = mp::plan_launch(1, &robot, &scene, params, 512);  // validates params
        const int dof = robot->model.dof, cap = params->tree_capacity;
        const Layout L = layout(cap, dof, params->max_path);
        if (L.end > kScratchBytes)
            throw Error(MP_E_CAPACITY, "mp_rrtc_solve_noalloc: tree_capacity " + std::to_string(cap) + " x " +
                                           std::to_string(dof) + " joints and max_path " +
                                           std::to_string(params->max_path) + " need " + std::to_string(L.end) +
                                           " bytes of scratch; the maximum is " + std::to_string(kScratchBytes));
        int dev = 0;
        MP_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> g(g_mu);
        unsigned char* base = dev >= 0 && dev < kMaxDevices ? g_ptr[dev] : nullptr;
        if (!base) throw Error(MP_E_STATE, "mp_init() has not set up this device");
        hipStream_t s = mp::lib_stream();
        // empty trees (the sentinel coordinates), parents and radii (-2), and a zero state, so the
        // admission kernel finds nothing of an earlier call to clear
        MP_HIP(hipMemsetD32Async((hipDeviceptr_t)(base + L.trees), (int)0x7fc0deadu, (size_t)2 * dof * cap, s));
        MP_HIP(hipMemsetD32Async((hipDeviceptr_t)(base + L.parents), -2, (size_t)2 * 2 * cap, s));
        MP_HIP(hipMemsetAsync(base + L.st, 0, sizeof(mp::QueryStateHost), s));
        mp::QueryDevHost q;
        std::memset(&q, 0, sizeof(q));
        q.scene = mp::scene_view(scene);
        for (int t = 0; t < 2; t++) {
            q.tree[t] = reinterpret_cast<float*>(base + L.trees) + (size_t)t * dof * cap;
            q.parent[t] = reinterpret_cast<int*>(base + L.parents) + (size_t)t * 2 * cap;  // radii at + cap
        }
        q.path = reinterpret_cast<float*>(base + L.path);
        q.seed = params->seed;
        q.robot = lp.slot_of[0];
        q.dof = dof;
        q.cap = cap;
        q.max_iters = params->max_iterations;
        q.max_path = params->max_path;
        q.dof_max = dof;
        q.sampler = params->sampler;
        std::memcpy(q.start, start, dof * sizeof(float));
        std::memcpy(q.goal, goal, dof * sizeof(float));
        int hctl[17];
        mp::QueryStateHost hst;
        const mp::RunBuffers b{reinterpret_cast<mp::QueryDevHost*>(base + L.q),
                               reinterpret_cast<mp::QueryStateHost*>(base + L.st), reinterpret_cast<int*>(base + L.active),
                               reinterpret_cast<int*>(base + L.ctl), &q, hctl, &hst};
        mp::run_queries(*lp.mod, b, 1, *params, lp.team, lp.teams, lp.cap, lp.lanes, s, nullptr);
        const int n = mp::harvest_status(hst, params->max_iterations, cap, out);
        if (n > 0) {
            MP_HIP(hipMemcpyAsync(out->waypoints, base + L.path, (size_t)n * dof * sizeof(float), hipMemcpyDeviceToHost, s));
            MP_HIP(hipStreamSynchronize(s));
        }
        out->latency_ms = now_ms() - t0;
    });
}
