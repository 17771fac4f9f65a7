// hipRTC compilation of the generated module source for gfx950, with an on-disk code-object
// cache keyed by a hash of (source, options, hipRTC version).  Replaces KernelFunction::create
// (cudampl/CUDAMPLib/include/NvrtcUtil.h:58-148, cudampl/CPRRTC/include/Nvrtc.h:58-150), whose
// PTX cache was keyed by kernel name only, so a second robot silently reused the first robot's
// kernel.
#include <dlfcn.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>
#include <utime.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <algorithm>
#include <map>
#include <sstream>

#include "mp_internal.h"

namespace mp {

namespace {
const char* kKernelTemplate =
#include "_kernels_src.inc"
    ;

const char* kOptions[] = {"--offload-arch=gfx950",
                          "-O3",
                          "-std=c++17",
                          "-ffp-contract=off",
                          "-fhip-fp32-correctly-rounded-divide-sqrt",
                          "-Xclang",
                          "-target-feature",
                          "-Xclang",
                          "-packed-fp32-ops"};
const int kNumOptions = sizeof(kOptions) / sizeof(kOptions[0]);

std::string section(const char* begin, const char* end) {
    std::string t(kKernelTemplate);
    size_t a = t.find(begin), b = t.find(end);
    if (a == std::string::npos || b == std::string::npos) throw Error(MP_E_RTC, "kernel template markers missing");
    a += std::string(begin).size();
    return t.substr(a, b - a);
}

std::string lib_dir() {
    Dl_info info;
    if (dladdr((void*)&lib_dir, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t k = p.rfind('/');
        if (k != std::string::npos) return p.substr(0, k);
    }
    return ".";
}
}  // namespace

uint64_t fnv1a(const std::string& s, uint64_t h) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

uint64_t robot_hash(const RobotModel& m) { return fnv1a(robot_struct_source(m, "Robot0")); }

std::string kernel_template_prelude() { return section("//@@PRELUDE_BEGIN@@", "//@@PRELUDE_END@@"); }
std::string kernel_template_body() { return section("//@@TEMPLATES_BEGIN@@", "//@@TEMPLATES_END@@"); }

std::string& cache_dir() {
    static std::string dir = [] {
        const char* e = std::getenv("MP_KERNEL_CACHE");
        return e && *e ? std::string(e) : lib_dir() + "/_kcache";
    }();
    return dir;
}

std::vector<char> rtc_compile(const std::string& source, const std::string& tag) {
    int major = 0, minor = 0;
    hiprtcVersion(&major, &minor);
    std::string opts;
    for (int i = 0; i < kNumOptions; i++) opts += std::string(kOptions[i]) + " ";
    char key[64];
    std::snprintf(key, sizeof(key), "%016llx",
                  (unsigned long long)fnv1a(source + "\n" + opts + std::to_string(major) + "." + std::to_string(minor)));
    std::string path = cache_dir() + "/" + tag + "_" + key + ".hsaco";
    {
        std::ifstream f(path, std::ios::binary);
        if (f) {
            std::vector<char> code((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
            if (!code.empty()) {
                utime(path.c_str(), nullptr);  // mark as in use (the build prunes entries it did not touch)
                return code;
            }
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, source.c_str(), (tag + ".hip").c_str(), 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        throw Error(MP_E_RTC, "hiprtcCreateProgram failed");
    hiprtcResult r = hiprtcCompileProgram(prog, kNumOptions, kOptions);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        if (log.size() > 4000) log = log.substr(0, 4000);
        throw Error(MP_E_RTC, "hipRTC compile of " + tag + " failed: " + log);
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    // best effort: store atomically (write temp, rename)
    mkdir(cache_dir().c_str(), 0775);
    std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (f) f.write(code.data(), (std::streamsize)code.size());
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
    return code;
}

Module::~Module() {
    if (mod) hipModuleUnload(mod);
}

hipFunction_t Module::fn(const std::string& name) {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& p : fns_)
        if (p.first == name) return p.second;
    hipFunction_t f = nullptr;
    MP_HIP(hipModuleGetFunction(&f, mod, name.c_str()));
    fns_.emplace_back(name, f);
    return f;
}

std::vector<int> canonical_order(const std::vector<const RobotModel*>& robots) {
    std::vector<std::pair<std::string, uint64_t>> key(robots.size());
    for (size_t i = 0; i < robots.size(); i++)
        key[i] = {robots[i]->name, robots[i]->hash ? robots[i]->hash : robot_hash(*robots[i])};
    std::vector<int> order(robots.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    return order;
}

std::shared_ptr<Module> get_module(const std::vector<const RobotModel*>& robots, std::vector<int>* slot_of) {
    static std::mutex mu;
    // canonical robot hashes -> module.  Modules stay loaded (a multi-robot module is otherwise
    // released after every batched call and reloaded by the next); the map is never destroyed,
    // so no module is unloaded after the HIP runtime has shut down at process exit.
    static auto* cache = new std::map<std::string, std::shared_ptr<Module>>();
    static auto* order_lru = new std::vector<std::string>();
    constexpr size_t kMaxModules = 16;
    std::vector<int> order = canonical_order(robots);
    if (slot_of) {
        slot_of->assign(robots.size(), 0);
        for (size_t s = 0; s < order.size(); s++) (*slot_of)[order[s]] = (int)s;
    }
    std::vector<const RobotModel*> canon;
    std::string k;
    for (int i : order) {
        canon.push_back(robots[i]);
        k += robots[i]->name + "#" + std::to_string(robots[i]->hash ? robots[i]->hash : robot_hash(*robots[i])) + ":";
    }
    std::lock_guard<std::mutex> g(mu);
    auto it = cache->find(k);
    if (it != cache->end()) {
        order_lru->erase(std::find(order_lru->begin(), order_lru->end(), k));
        order_lru->push_back(k);
        return it->second;
    }
    std::string src = module_source(canon);
    std::string tag = "mp";
    for (auto* r : canon) tag += "_" + r->name;
    std::vector<char> code = rtc_compile(src, tag);
    auto m = std::make_shared<Module>();
    m->source = src;
    MP_HIP(hipModuleLoadData(&m->mod, code.data()));
    if (cache->size() >= kMaxModules) {  // drop the least recently used (users keep their shared_ptr)
        cache->erase(order_lru->front());
        order_lru->erase(order_lru->begin());
    }
    (*cache)[k] = m;
    order_lru->push_back(k);
    return m;
}

}  // namespace mp
