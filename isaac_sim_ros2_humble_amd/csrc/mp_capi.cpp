// C ABI of the motion-planning hot path (declarations and reference citations in include/mp_abi.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "mp_internal.h"
#include "mp_plan.h"

using mp::Error;

namespace {

thread_local std::string g_err;

struct Global {
    int device = -1;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    unsigned long long* work = nullptr;  // device counter of algorithmic flops (check/cost/edges)
    unsigned long long* prof = nullptr;  // planner phase profile counters [MP_PLAN_PROFILE_N]
    bool prof_on = false;
    std::mutex mu;
    std::map<std::string, std::pair<hipEvent_t, hipEvent_t>> events;
} G;

template <class F>
int guard(F&& f) {
    try {
        f();
        return MP_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return MP_E_INVALID;
    }
}

void need_init() {
    if (G.device < 0) throw Error(MP_E_STATE, "mp_init() has not been called");
}

hipStream_t stream_of(void* s) { return s ? (hipStream_t)s : G.stream; }

}  // namespace

namespace mp {

hipStream_t lib_stream() { return G.stream; }
int guard_call(const std::function<void()>& f) { return guard(f); }
int lib_num_cus() { return G.num_cus; }
unsigned long long* lib_plan_profile() { return G.prof_on ? G.prof : nullptr; }
unsigned long long* lib_work_counter() { return G.work; }

void launch(hipFunction_t f, unsigned grid, unsigned block, hipStream_t s, void** args, const char* family) {
    if (grid == 0) return;
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (family) {
        std::lock_guard<std::mutex> g(G.mu);
        auto it = G.events.find(family);
        if (it == G.events.end()) {
            hipEvent_t a, b;
            MP_HIP(hipEventCreate(&a));
            MP_HIP(hipEventCreate(&b));
            it = G.events.emplace(family, std::make_pair(a, b)).first;
        }
        ev = it->second;
    }
    if (ev.first) MP_HIP(hipEventRecord(ev.first, s));
    MP_HIP(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, s, args, nullptr));
    if (ev.second) MP_HIP(hipEventRecord(ev.second, s));
}

SceneDevHost scene_view(const mp_scene* sc) {
    SceneDevHost v;
    v.ns = sc->ns;
    v.nb = sc->nb;
    v.nc = sc->nc;
    v.ncap = sc->ncap;
    v.sph = sc->dev + sc->offs[0];
    v.bbox = sc->dev + sc->offs[1];
    v.box = sc->dev + sc->offs[2];
    v.bcyl = sc->dev + sc->offs[3];
    v.cyl = sc->dev + sc->offs[4];
    v.bcap = sc->dev + sc->offs[5];
    v.cap = sc->dev + sc->offs[6];
    v.bnd = sc->bounds_mode == MP_BOUNDS_ROBOT ? nullptr : sc->bnd_dev;
    return v;
}

RobotModel model_from_desc(const mp_robot_desc* d) {
    if (!d || d->num_links <= 0 || d->num_links > 64 || d->num_spheres < 0 || d->num_spheres > 1024)
        throw Error(MP_E_INVALID, "robot description: bad link/sphere counts");
    if (!d->joint_types || !d->parents || !d->active || !d->joint_origins || !d->joint_axes || !d->lower || !d->upper ||
        !d->defaults || !d->collision_enabled)
        throw Error(MP_E_INVALID, "robot description: a per-link array is null");
    if (d->num_spheres > 0 && (!d->sphere_link || !d->sphere_pos || !d->sphere_radius))
        throw Error(MP_E_INVALID, "robot description: a per-sphere array is null");
    RobotModel m;
    m.name = d->name ? d->name : "robot";
    for (char& c : m.name)
        if (!isalnum((unsigned char)c)) c = '_';
    m.L = d->num_links;
    m.S = d->num_spheres;
    m.jtype.assign(d->joint_types, d->joint_types + m.L);
    m.parent.assign(d->parents, d->parents + m.L);
    m.active.assign(d->active, d->active + m.L);
    m.origin.assign(d->joint_origins, d->joint_origins + 16 * m.L);
    m.axis.assign(d->joint_axes, d->joint_axes + 3 * m.L);
    m.lower.assign(d->lower, d->lower + m.L);
    m.upper.assign(d->upper, d->upper + m.L);
    m.defaults.assign(d->defaults, d->defaults + m.L);
    m.sphere_link.assign(d->sphere_link, d->sphere_link + m.S);
    m.sphere_pos.assign(d->sphere_pos, d->sphere_pos + 3 * m.S);
    m.sphere_r.assign(d->sphere_radius, d->sphere_radius + m.S);
    m.enabled.assign(d->collision_enabled, d->collision_enabled + m.L * m.L);
    // topological order check (SingleArmSpace.cu:62-69)
    if (m.parent[0] != -1) throw Error(MP_E_INVALID, "link 0 must be the root (parent -1)");
    for (int i = 1; i < m.L; i++)
        if (m.parent[i] < 0 || m.parent[i] >= i) throw Error(MP_E_INVALID, "links must be in topological order");
    for (int s = 0; s < m.S; s++)
        if (m.sphere_link[s] < 0 || m.sphere_link[s] >= m.L) throw Error(MP_E_INVALID, "sphere link index out of range");
    m.active_index.assign(m.L, -1);
    for (int i = 0; i < m.L; i++) {
        int t = m.jtype[i];
        if (t != 0 && t != 1 && t != 2 && t != 5) throw Error(MP_E_INVALID, "unsupported joint type");
        if (m.active[i]) {
            if (t != 1 && t != 2) throw Error(MP_E_INVALID, "active joints must be revolute or prismatic");
            m.active_index[i] = m.dof++;
            m.act_lo.push_back(m.lower[i]);
            m.act_hi.push_back(m.upper[i]);
        }
    }
    if (m.dof < 1 || m.dof > 16) throw Error(MP_E_INVALID, "dof must be in [1,16]");
    // self-collision pairs (SelfCollisionConstraint.cu:29-39, RobotSolver.cu:86-96)
    for (int i = 0; i < m.S; i++)
        for (int j = i + 1; j < m.S; j++) {
            int li = m.sphere_link[i], lj = m.sphere_link[j];
            if (li != lj && m.enabled[li * m.L + lj]) {
                float sr = m.sphere_r[i] + m.sphere_r[j];
                float thr = sr * sr;
                m.pi.push_back(i);
                m.pj.push_back(j);
                m.pthr.push_back(thr);
                m.psr.push_back(std::sqrt(thr));
            }
        }
    m.hash = robot_hash(m);
    return m;
}

}  // namespace mp

// =========================================================================================
extern "C" {

const char* mp_last_error(void) { return g_err.c_str(); }

int mp_init(int device) {
    return guard([&] {
        if (G.device == device && G.stream) return;
        int n = 0;
        MP_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw Error(MP_E_INVALID, "no such HIP device");
        MP_HIP(hipSetDevice(device));
        mp::noalloc_setup(device);  // the no-allocation entry point's scratch address on this device
        MP_HIP(hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking));
        hipDeviceProp_t p;
        MP_HIP(hipGetDeviceProperties(&p, device));
        G.num_cus = p.multiProcessorCount;
        MP_HIP(hipMalloc(&G.work, sizeof(unsigned long long)));
        MP_HIP(hipMemset(G.work, 0, sizeof(unsigned long long)));
        MP_HIP(hipMalloc(&G.prof, MP_PLAN_PROFILE_N * sizeof(unsigned long long)));
        MP_HIP(hipMemset(G.prof, 0, MP_PLAN_PROFILE_N * sizeof(unsigned long long)));
        G.device = device;
    });
}

int mp_shutdown(void) {
    return guard([&] {
        if (G.device < 0) return;
        hipStreamSynchronize(G.stream);
        for (auto& e : G.events) {
            hipEventDestroy(e.second.first);
            hipEventDestroy(e.second.second);
        }
        G.events.clear();
        mp::plan_release_workspaces();
        if (G.work) hipFree(G.work);
        G.work = nullptr;
        if (G.prof) hipFree(G.prof);
        G.prof = nullptr;
        G.prof_on = false;
        hipStreamDestroy(G.stream);
        G.stream = nullptr;
        G.device = -1;
    });
}

int mp_set_cache_dir(const char* dir) {
    return guard([&] {
        if (!dir) throw Error(MP_E_INVALID, "null dir");
        mp::cache_dir() = dir;
    });
}

int mp_device_alloc(void** dptr, size_t bytes) {
    return guard([&] {
        need_init();
        MP_HIP(hipMalloc(dptr, bytes ? bytes : 1));
    });
}
int mp_device_free(void* dptr) {
    return guard([&] {
        need_init();
        if (dptr) MP_HIP(hipFree(dptr));
    });
}
int mp_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    return guard([&] {
        need_init();
        MP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(stream)));
        MP_HIP(hipStreamSynchronize(stream_of(stream)));
    });
}
int mp_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    return guard([&] {
        need_init();
        MP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(stream)));
        MP_HIP(hipStreamSynchronize(stream_of(stream)));
    });
}
int mp_memset_d8(void* dst, int value, size_t bytes, void* stream) {
    return guard([&] {
        need_init();
        MP_HIP(hipMemsetAsync(dst, value, bytes, stream_of(stream)));
    });
}
int mp_stream_sync(void* stream) {
    return guard([&] {
        need_init();
        MP_HIP(hipStreamSynchronize(stream_of(stream)));
    });
}
void* mp_default_stream(void) { return (void*)G.stream; }

int mp_plan_profile(int enable, uint64_t* counters, int n) {
    return guard([&] {
        need_init();
        if (counters && (n < 0 || n > MP_PLAN_PROFILE_N)) throw Error(MP_E_INVALID, "bad profile counter count");
        MP_HIP(hipStreamSynchronize(G.stream));
        if (counters) {
            unsigned long long v[MP_PLAN_PROFILE_N];
            MP_HIP(hipMemcpy(v, G.prof, sizeof(v), hipMemcpyDeviceToHost));
            for (int k = 0; k < n; k++) counters[k] = v[k];
            MP_HIP(hipMemset(G.prof, 0, sizeof(v)));
        }
        if (enable >= 0) G.prof_on = enable != 0;
    });
}

int mp_work_flops(uint64_t* out, int reset) {
    return guard([&] {
        need_init();
        unsigned long long v = 0;
        MP_HIP(hipMemcpyAsync(&v, G.work, sizeof(v), hipMemcpyDeviceToHost, G.stream));
        MP_HIP(hipStreamSynchronize(G.stream));
        if (out) *out = v;
        if (reset) {
            MP_HIP(hipMemsetAsync(G.work, 0, sizeof(v), G.stream));
            MP_HIP(hipStreamSynchronize(G.stream));
        }
    });
}

double mp_last_kernel_ms(const char* family) {
    std::lock_guard<std::mutex> g(G.mu);
    auto it = G.events.find(family ? family : "");
    if (it == G.events.end()) return -1.0;
    if (hipEventSynchronize(it->second.second) != hipSuccess) return -1.0;
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, it->second.first, it->second.second) != hipSuccess) return -1.0;
    return ms;
}

// ---------------------------------------------------------------------------------------- robot
int mp_robot_create(const mp_robot_desc* desc, mp_robot** out) {
    return guard([&] {
        need_init();
        if (!out) throw Error(MP_E_INVALID, "null out");
        auto* r = new mp_robot();
        try {
            r->model = mp::model_from_desc(desc);
            r->key = r->model.hash;
            r->module = mp::get_module({&r->model});
        } catch (...) {
            delete r;
            throw;
        }
        *out = r;
    });
}

int mp_precompile(const mp_robot_desc* const* descs, int n) {
    return guard([&] {
        if (n < 0 || (n > 0 && !descs)) throw Error(MP_E_INVALID, "bad robot list");
        std::vector<mp::RobotModel> ms;
        ms.reserve(n);
        for (int i = 0; i < n; i++) ms.push_back(mp::model_from_desc(descs[i]));
        std::vector<const mp::RobotModel*> ptrs;
        for (auto& m : ms) ptrs.push_back(&m);
        std::vector<const mp::RobotModel*> canon;
        std::string tag = "mp";
        for (int i : mp::canonical_order(ptrs)) {
            canon.push_back(ptrs[i]);
            tag += "_" + ptrs[i]->name;
        }
        mp::rtc_compile(mp::module_source(canon), tag);
    });
}

int mp_robot_precompile(const mp_robot_desc* desc) { return mp_precompile(&desc, 1); }

size_t mp_generate_source(const mp_robot_desc* const* descs, int n, char* buf, size_t cap) {
    std::string src;
    int rc = guard([&] {
        if (n < 0 || (n > 0 && !descs)) throw Error(MP_E_INVALID, "bad robot list");
        std::vector<mp::RobotModel> ms;
        ms.reserve(n);
        for (int i = 0; i < n; i++) ms.push_back(mp::model_from_desc(descs[i]));
        std::vector<const mp::RobotModel*> ptrs;
        for (auto& m : ms) ptrs.push_back(&m);
        src = mp::module_source(ptrs);
    });
    if (rc != MP_OK) return 0;
    if (buf && cap) {
        size_t k = std::min(cap - 1, src.size());
        std::memcpy(buf, src.data(), k);
        buf[k] = 0;
    }
    return src.size();
}

int mp_robot_destroy(mp_robot* robot) {
    return guard([&] { delete robot; });
}

int mp_robot_info(const mp_robot* r, int* dof, int* ns, int* np, int* nl) {
    return guard([&] {
        if (!r) throw Error(MP_E_INVALID, "null robot");
        if (dof) *dof = r->model.dof;
        if (ns) *ns = r->model.S;
        if (np) *np = (int)r->model.pi.size();
        if (nl) *nl = r->model.L;
    });
}

size_t mp_robot_source(const mp_robot* r, char* buf, size_t cap) {
    if (!r || !r->module) return 0;
    const std::string& s = r->module->source;
    if (buf && cap) {
        size_t k = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), k);
        buf[k] = 0;
    }
    return s.size();
}

// ---------------------------------------------------------------------------------------- scene
namespace {
float dot3f(float a0, float b0, float a1, float b1, float a2, float b2) {
    return std::fmaf(a0, b0, std::fmaf(a1, b1, a2 * b2));
}
// computeInversePoseMatrix (EnvConstraintCuboid.cu:13-34): [R^T | -R^T p], float
void inverse_pose(const float* p, const float* R, float* m) {
    float t00 = R[0], t01 = R[3], t02 = R[6];
    float t10 = R[1], t11 = R[4], t12 = R[7];
    float t20 = R[2], t21 = R[5], t22 = R[8];
    m[0] = t00; m[1] = t01; m[2] = t02;
    m[4] = t10; m[5] = t11; m[6] = t12;
    m[8] = t20; m[9] = t21; m[10] = t22;
    m[3] = -dot3f(t00, p[0], t01, p[1], t02, p[2]);
    m[7] = -dot3f(t10, p[0], t11, p[1], t12, p[2]);
    m[11] = -dot3f(t20, p[0], t21, p[1], t22, p[2]);
}

// World bounding sphere of the region {p : m p + t in [lo, hi]} (m, t = the inverse pose a kernel
// applies, o[0..11]).  Rotations that are not orthonormal to 1e-4 get an infinite radius (never
// culled): the link-level tests measure distances in the obstacle frame, which equal world
// distances only for rotations.
void local_box_bound(const float* o, const float* lo, const float* hi, double wc[3], double& rho) {
    double M[3][3], t[3];
    for (int a = 0; a < 3; a++) {
        for (int c = 0; c < 3; c++) M[a][c] = o[4 * a + c];
        t[a] = o[4 * a + 3];
    }
    double dev = 0.0;
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) {
            double d = M[a][0] * M[c][0] + M[a][1] * M[c][1] + M[a][2] * M[c][2] - (a == c ? 1.0 : 0.0);
            dev = std::max(dev, std::fabs(d));
        }
    // inverse of M by cofactors
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    if (dev > 1e-4 || !(std::fabs(det) > 0.5)) {
        wc[0] = wc[1] = wc[2] = 0.0;
        rho = 1e30;
        return;
    }
    double I[3][3];
    I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    auto to_world = [&](const double l[3], double w[3]) {
        for (int a = 0; a < 3; a++)
            w[a] = I[a][0] * (l[0] - t[0]) + I[a][1] * (l[1] - t[1]) + I[a][2] * (l[2] - t[2]);
    };
    const double lc[3] = {0.5 * ((double)lo[0] + hi[0]), 0.5 * ((double)lo[1] + hi[1]), 0.5 * ((double)lo[2] + hi[2])};
    to_world(lc, wc);
    rho = 0.0;
    for (int corner = 0; corner < 8; corner++) {
        const double l[3] = {(corner & 1) ? hi[0] : lo[0], (corner & 2) ? hi[1] : lo[1], (corner & 4) ? hi[2] : lo[2]};
        double w[3];
        to_world(l, w);
        rho = std::max(rho, std::sqrt((w[0] - wc[0]) * (w[0] - wc[0]) + (w[1] - wc[1]) * (w[1] - wc[1]) +
                                      (w[2] - wc[2]) * (w[2] - wc[2])));
    }
}

// Bounding sphere of an obstacle for the obstacle-level cull: world centre, radius computed in
// double and rounded up so that the float sphere contains the obstacle.
void bound_sphere(float* o, double cx, double cy, double cz, double r) {
    o[0] = (float)cx;
    o[1] = (float)cy;
    o[2] = (float)cz;
    o[3] = (float)(r * (1.0 + 1e-6) + 1e-6);
}

void scene_upload(mp_scene* sc, const float* spheres, int ns, const float* boxes, int nb, const float* cyl, int nc,
                  const float* caps, int ncap) {
    std::vector<float> h(sc->offs[6] + (size_t)MP_CAP_W_H * std::max(sc->cap_k, 1), 0.0f);
    for (int i = 0; i < ns; i++) std::memcpy(&h[sc->offs[0] + 4 * i], spheres + 4 * i, 4 * sizeof(float));
    for (int i = 0; i < nb; i++) {
        float* o = &h[sc->offs[2] + MP_BOX_W_H * i];
        const float* b = boxes + 18 * i;
        inverse_pose(b, b + 3, o);
        std::memcpy(o + 12, b + 12, 6 * sizeof(float));
        // the kernel tests local = M p + t (o[0..11]); the box is the preimage of [min, max]:
        // map its corners back exactly (double) and bound them
        double wc[3], rho;
        local_box_bound(o, b + 12, b + 15, wc, rho);
        bound_sphere(o + 18, wc[0], wc[1], wc[2], rho);
        std::memcpy(&h[sc->offs[1] + 4 * i], o + 18, 4 * sizeof(float));
    }
    for (int i = 0; i < nc; i++) {
        float* o = &h[sc->offs[4] + MP_CYL_W_H * i];
        inverse_pose(cyl + 14 * i, cyl + 14 * i + 3, o);
        o[12] = cyl[14 * i + 12];
        o[13] = cyl[14 * i + 13] * 0.5f;
        // a finite cylinder lies in its local box [-r, r] x [-r, r] x [-h/2, h/2]
        const float lo[3] = {-o[12], -o[12], -o[13]}, hi[3] = {o[12], o[12], o[13]};
        double wc[3], rho;
        local_box_bound(o, lo, hi, wc, rho);
        bound_sphere(o + 14, wc[0], wc[1], wc[2], rho);
        std::memcpy(&h[sc->offs[3] + 4 * i], o + 14, 4 * sizeof(float));
    }
    for (int i = 0; i < ncap; i++) {
        const float* c = caps + 7 * i;
        float* o = &h[sc->offs[6] + MP_CAP_W_H * i];
        o[0] = c[0]; o[1] = c[1]; o[2] = c[2];
        o[3] = c[3] - c[0]; o[4] = c[4] - c[1]; o[5] = c[5] - c[2];
        o[6] = 1.0f / dot3f(o[3], o[3], o[4], o[4], o[5], o[5]);
        o[7] = c[6];
        const double vx = (double)c[3] - c[0], vy = (double)c[4] - c[1], vz = (double)c[5] - c[2];
        bound_sphere(o + 8, 0.5 * ((double)c[0] + c[3]), 0.5 * ((double)c[1] + c[4]), 0.5 * ((double)c[2] + c[5]),
                     0.5 * std::sqrt(vx * vx + vy * vy + vz * vz) + c[6]);
        std::memcpy(&h[sc->offs[5] + 4 * i], o + 8, 4 * sizeof(float));
    }
    MP_HIP(hipMemcpyAsync(sc->dev, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, G.stream));
    MP_HIP(hipStreamSynchronize(G.stream));
    sc->ns = ns;
    sc->nb = nb;
    sc->nc = nc;
    sc->ncap = ncap;
}
}  // namespace

int mp_scene_create(const float* spheres, int ns, const float* boxes, int nb, const float* cylinders, int nc,
                    const float* capsules, int ncap, mp_scene** out) {
    return guard([&] {
        need_init();
        if (!out || ns < 0 || nb < 0 || nc < 0 || ncap < 0) throw Error(MP_E_INVALID, "bad scene arguments");
        if ((ns && !spheres) || (nb && !boxes) || (nc && !cylinders) || (ncap && !capsules))
            throw Error(MP_E_INVALID, "null obstacle array with nonzero count");
        auto* sc = new mp_scene();
        sc->cap_s = ns;
        sc->cap_b = nb;
        sc->cap_c = nc;
        sc->cap_k = ncap;
        sc->offs[0] = 0;
        sc->offs[1] = sc->offs[0] + (size_t)4 * std::max(ns, 1);
        sc->offs[2] = sc->offs[1] + (size_t)4 * std::max(nb, 1);
        sc->offs[3] = sc->offs[2] + (size_t)MP_BOX_W_H * std::max(nb, 1);
        sc->offs[4] = sc->offs[3] + (size_t)4 * std::max(nc, 1);
        sc->offs[5] = sc->offs[4] + (size_t)MP_CYL_W_H * std::max(nc, 1);
        sc->offs[6] = sc->offs[5] + (size_t)4 * std::max(ncap, 1);
        size_t total = sc->offs[6] + (size_t)MP_CAP_W_H * std::max(ncap, 1);
        try {
            MP_HIP(hipMalloc(&sc->dev, total * sizeof(float)));
            scene_upload(sc, spheres, ns, boxes, nb, cylinders, nc, capsules, ncap);
        } catch (...) {
            if (sc->dev) hipFree(sc->dev);
            delete sc;
            throw;
        }
        *out = sc;
    });
}

int mp_scene_update(mp_scene* sc, const float* spheres, int ns, const float* boxes, int nb, const float* cylinders,
                    int nc, const float* capsules, int ncap) {
    return guard([&] {
        need_init();
        if (!sc) throw Error(MP_E_INVALID, "null scene");
        if (ns < 0 || nb < 0 || nc < 0 || ncap < 0) throw Error(MP_E_INVALID, "negative obstacle count");
        if (ns > sc->cap_s || nb > sc->cap_b || nc > sc->cap_c || ncap > sc->cap_k)
            throw Error(MP_E_CAPACITY, "the number of obstacles exceeds the scene's capacity");
        if (sc->server_refs.load() > 0)
            throw Error(MP_E_STATE, "the scene has server queries queued or running; update it after they finish");
        MP_HIP(hipStreamSynchronize(G.stream));
        scene_upload(sc, spheres, ns, boxes, nb, cylinders, nc, capsules, ncap);
    });
}

int mp_scene_set_joint_bounds(mp_scene* sc, int mode, const float* lower, const float* upper, int n) {
    return guard([&] {
        need_init();
        if (!sc) throw Error(MP_E_INVALID, "null scene");
        if (mode != MP_BOUNDS_ROBOT && mode != MP_BOUNDS_NONE && mode != MP_BOUNDS_CUSTOM)
            throw Error(MP_E_INVALID, "joint-bound mode must be MP_BOUNDS_ROBOT, MP_BOUNDS_NONE or MP_BOUNDS_CUSTOM");
        if (mode == MP_BOUNDS_CUSTOM && (n < 0 || n > 16 || (n > 0 && (!lower || !upper))))
            throw Error(MP_E_INVALID, "custom joint bounds: n in [0, 16] with lower and upper arrays");
        if (sc->server_refs.load() > 0)
            throw Error(MP_E_STATE, "the scene has server queries queued or running; change it after they finish");
        float h[32];
        for (int j = 0; j < 16; j++) {
            const bool given = mode == MP_BOUNDS_CUSTOM && j < n;
            h[j] = given ? lower[j] : -INFINITY;
            h[16 + j] = given ? upper[j] : INFINITY;
        }
        MP_HIP(hipStreamSynchronize(G.stream));
        if (mode != MP_BOUNDS_ROBOT) {
            if (!sc->bnd_dev) MP_HIP(hipMalloc((void**)&sc->bnd_dev, sizeof(h)));
            MP_HIP(hipMemcpy(sc->bnd_dev, h, sizeof(h), hipMemcpyHostToDevice));
        }
        sc->bounds_mode = mode;
    });
}

int mp_scene_destroy(mp_scene* sc) {
    return guard([&] {
        if (!sc) return;
        if (sc->server_refs.load() > 0)
            throw Error(MP_E_STATE, "the scene has server queries queued or running; destroy it after they finish");
        if (sc->dev || sc->bnd_dev) hipStreamSynchronize(G.stream);
        if (sc->dev) hipFree(sc->dev);
        if (sc->bnd_dev) hipFree(sc->bnd_dev);
        delete sc;
    });
}

// ---------------------------------------------------------------------------------------- kernels
int mp_fk_spheres(const mp_robot* robot, const float* d_q, int64_t n, float* d_centres, void* stream) {
    return guard([&] {
        need_init();
        if (!robot || n < 0 || (n && (!d_q || !d_centres))) throw Error(MP_E_INVALID, "bad arguments");
        if (n == 0) return;
        long long nn = n;
        void* args[] = {&d_q, &nn, &d_centres};
        mp::launch(robot->module->fn("mp_fk_0"), (unsigned)((n + 255) / 256), 256, stream_of(stream), args, "fk");
    });
}

int mp_fk_link_poses(const mp_robot* robot, const float* d_q, int64_t n, float* d_poses, void* stream) {
    return guard([&] {
        need_init();
        if (!robot || n < 0 || (n && (!d_q || !d_poses))) throw Error(MP_E_INVALID, "bad arguments");
        if (n == 0) return;
        long long nn = n;
        void* args[] = {&d_q, &nn, &d_poses};
        mp::launch(robot->module->fn("mp_poses_0"), (unsigned)((n + 255) / 256), 256, stream_of(stream), args, "fk");
    });
}

int mp_check_states(const mp_robot* robot, const mp_scene* scene, const float* d_q, int64_t n, uint8_t* d_verdict,
                    float* d_centres, void* stream) {
    return guard([&] {
        need_init();
        if (!robot || !scene || n < 0 || (n && (!d_q || !d_verdict))) throw Error(MP_E_INVALID, "bad arguments");
        if (n == 0) return;
        long long nn = n;
        mp::SceneDevHost sv = mp::scene_view(scene);
        void* args[] = {&d_q, &nn, &sv, &d_verdict, &d_centres, &G.work};
        mp::launch(robot->module->fn("mp_check_0"), (unsigned)((n + 255) / 256), 256, stream_of(stream), args, "check");
    });
}

int mp_state_costs(const mp_robot* robot, const mp_scene* scene, const float* d_q, int64_t n, float* d_cost,
                   uint8_t* d_verdict, void* stream) {
    return guard([&] {
        need_init();
        if (!robot || !scene || n < 0 || (n && !d_q)) throw Error(MP_E_INVALID, "bad arguments");
        if (n == 0) return;
        long long nn = n;
        mp::SceneDevHost sv = mp::scene_view(scene);
        float* d_parts = nullptr;
        void* args[] = {&d_q, &nn, &sv, &d_cost, &d_verdict, &d_parts, &G.work};
        mp::launch(robot->module->fn("mp_cost_0"), (unsigned)((n + 255) / 256), 256, stream_of(stream), args, "cost");
    });
}

int mp_constraint_costs(const mp_robot* robot, const mp_scene* scene, const float* d_q, int64_t n, float* d_parts,
                        float* d_cost, uint8_t* d_verdict, void* stream) {
    return guard([&] {
        need_init();
        if (!robot || !scene || n < 0 || (n && (!d_q || !d_parts))) throw Error(MP_E_INVALID, "bad arguments");
        if (n == 0) return;
        long long nn = n;
        mp::SceneDevHost sv = mp::scene_view(scene);
        void* args[] = {&d_q, &nn, &sv, &d_cost, &d_verdict, &d_parts, &G.work};
        mp::launch(robot->module->fn("mp_cost_0"), (unsigned)((n + 255) / 256), 256, stream_of(stream), args, "cost");
    });
}

int mp_check_edges(const mp_robot* robot, const mp_scene* scene, const float* d_q0, const float* d_q1, int m,
                   float resolution, uint8_t* d_ok, float* d_len, void* stream) {
    return guard([&] {
        need_init();
        if (!robot || !scene || m < 0 || (m && (!d_q0 || !d_q1 || !d_ok || !d_len)) || !(resolution > 0.0f))
            throw Error(MP_E_INVALID, "bad arguments");
        if (m == 0) return;
        mp::SceneDevHost sv = mp::scene_view(scene);
        int waves = std::min(m, std::max(G.num_cus, 1) * 16);
        void* args[] = {&d_q0, &d_q1, &m, &resolution, &sv, &d_ok, &d_len, &G.work};
        mp::launch(robot->module->fn("mp_edges_0"), (unsigned)((waves + 3) / 4), 256, stream_of(stream), args, "edges");
    });
}

int mp_nn_argmin(const float* d_tree, int n_nodes, int capacity, int dof, const float* d_q, int nq, int* d_idx,
                 void* stream) {
    return guard([&] {
        need_init();
        if (n_nodes < 0 || capacity < n_nodes || dof < 1 || nq < 0 || (nq && (!d_q || !d_idx)) || (n_nodes && !d_tree))
            throw Error(MP_E_INVALID, "bad arguments");
        if (nq == 0) return;
        auto mod = mp::get_module({});
        void* args[] = {&d_tree, &n_nodes, &capacity, &dof, &d_q, &nq, &d_idx};
        mp::launch(mod->fn("mp_nn_generic"), (unsigned)((nq + 3) / 4), 256, stream_of(stream), args, "nn");
    });
}

int mp_steer(const float* d_from, float* d_to, int n, int dof, float max_distance, void* stream) {
    return guard([&] {
        need_init();
        if (n < 0 || dof < 1 || (n && (!d_from || !d_to))) throw Error(MP_E_INVALID, "bad arguments");
        if (n == 0) return;
        auto mod = mp::get_module({});
        void* args[] = {&d_from, &d_to, &n, &dof, &max_distance};
        mp::launch(mod->fn("mp_steer_generic"), (unsigned)((n + 255) / 256), 256, stream_of(stream), args, "steer");
    });
}

int mp_sample(const mp_robot* robot, uint64_t seed, uint64_t offset, int64_t n, float* d_q, void* stream) {
    return guard([&] {
        need_init();
        if (!robot || n < 0 || (n && !d_q)) throw Error(MP_E_INVALID, "bad arguments");
        if (n == 0) return;
        long long nn = n;
        void* args[] = {&seed, &offset, &nn, &d_q};
        mp::launch(robot->module->fn("mp_sample_0"), (unsigned)((n + 255) / 256), 256, stream_of(stream), args, "sample");
    });
}

int mp_sample_halton(const mp_robot* robot, uint64_t seed, uint64_t offset, int64_t n, float* d_q, void* stream) {
    return guard([&] {
        need_init();
        if (!robot || n < 0 || (n && !d_q)) throw Error(MP_E_INVALID, "bad arguments");
        if (n == 0) return;
        long long nn = n;
        void* args[] = {&seed, &offset, &nn, &d_q};
        mp::launch(robot->module->fn("mp_halton_0"), (unsigned)((n + 255) / 256), 256, stream_of(stream), args, "halton");
    });
}

}  // extern "C"
