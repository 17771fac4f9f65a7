"""Build the native library in-tree: isaac_sim_ros2_humble_amd/libmpamd.so.

Steps: embed the kernel template (csrc/mp_kernels.hip.tmpl) into the host library as a string,
compile the C++ host sources against the HIP runtime + hipRTC, then (optionally) pre-populate
the code-object cache by hipRTC-compiling the gfx950 modules for the bundled robots, so a fresh
process on the GPU finds them on disk.  No GPU is needed for any step.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmpamd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SOURCES = ["mp_capi.cpp", "codegen.cpp", "rtc.cpp", "plan.cpp", "server.cpp", "roadmap.cpp"]


def embed_template():
    src = open(os.path.join(CSRC, "mp_kernels.hip.tmpl")).read()
    out = os.path.join(CSRC, "_kernels_src.inc")
    # split into chunks: some compilers cap single string literal length
    chunks = [src[i:i + 12000] for i in range(0, len(src), 12000)]
    body = "\n".join('R"MPKSRC(' + c + ')MPKSRC"' for c in chunks)
    new = body + "\n"
    if not os.path.exists(out) or open(out).read() != new:
        with open(out, "w") as f:
            f.write(new)


HIP_SOURCES = ["noalloc.hip"]  # sources with device code of their own, compiled by hipcc
OBJ_DIR = os.path.join(HERE, "_build")


def compile_hip_objects(verbose=False):
    """hipcc -c for gfx950 of the library's own device-code sources (the no-allocation entry
    point's scratch array); returns the object paths."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [os.path.join(ROCM, "bin", "hipcc"), "-c", "-fPIC", "-O2", "-std=c++17", "--offload-arch=gfx950",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
               "-I" + os.path.join(REPO, "include"), "-I" + CSRC, os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.check_call(cmd)
        objs.append(obj)
    return objs


def compile_lib(verbose=False):
    embed_template()
    objs = compile_hip_objects(verbose)
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wall",
           "-Wno-unused-function", "-Wno-unused-result", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
           "-I" + os.path.join(ROCM, "include")]
    cmd += [os.path.join(CSRC, s) for s in SOURCES] + objs
    cmd += ["-o", LIB + ".tmp", "-L" + os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib"),
            "-lamdhip64", "-lhiprtc", "-ldl", "-lpthread"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)


def _precompile_one(names):
    sys.path.insert(0, REPO)
    from isaac_sim_ros2_humble_amd import mbm, native
    from isaac_sim_ros2_humble_amd.model import load_robot
    models = [mbm.mbm_model(load_robot(n[:-4])) if n.endswith("_mbm") else load_robot(n) for n in names]
    try:
        native.precompile(models)
    except Exception as e:  # the parent may not import this package: hand back a plain error
        raise RuntimeError("precompiling %s: %s" % (names, e)) from None
    return names


CPP = os.path.join(HERE, "cpp")
CPP_LIB = os.path.join(HERE, "libcudampl_amd.so")
CPP_TEST_SRC = os.path.join(REPO, "tests", "cpp", "test_cpp_api.cpp")
CPP_TEST_BIN = os.path.join(REPO, "tests", "cpp", "test_cpp_api")


def compile_cpp_api():
    """C++ planner API (CPRRTC::RobotSolver, CUDAMPLib classes) over the C ABI, plus its test."""
    inc = ["-I" + os.path.join(REPO, "include"), "-I" + os.path.join(CPP, "include")]
    srcs = [os.path.join(CPP, "src", f) for f in ("RobotSolver.cpp", "cudampl.cpp")]
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wno-unused-result"] + inc + srcs +
                          ["-o", CPP_LIB, "-L" + HERE, "-Wl,-rpath," + HERE, "-lmpamd"])
    if os.path.exists(CPP_TEST_SRC):
        subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall"] + inc + [CPP_TEST_SRC, "-o", CPP_TEST_BIN, "-L" + HERE,
                              "-Wl,-rpath," + HERE, "-lcudampl_amd", "-lmpamd"])


def precompile_bundled():
    """hipRTC-compile the modules for the bundled robots into the on-disk cache (one process per
    robot set, in parallel)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    import glob
    import time
    sets = [(), ("fr3",), ("fetch",), ("fr3", "fetch"), ("fer_mbm",), ("fetch_mbm",), ("fer_mbm", "fetch_mbm")]
    t0 = time.time() - 1.0
    with cf.ProcessPoolExecutor(len(sets), mp_context=mp.get_context("spawn")) as ex:  # noqa: SIM117
        for names in ex.map(_precompile_one, sets):
            pass
    # drop code objects of earlier sources (every bundled set was written or touched above)
    for f in glob.glob(os.path.join(HERE, "_kcache", "*.hsaco")):
        if os.path.getmtime(f) < t0:
            os.remove(f)


def build(precompile=True, verbose=False):
    compile_lib(verbose)
    compile_cpp_api()
    if precompile:
        precompile_bundled()


if __name__ == "__main__":
    build(precompile="--no-precompile" not in sys.argv, verbose=True)
