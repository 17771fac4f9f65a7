"""ctypes binding of the native library (libmpamd.so, C ABI in include/mp_abi.h).

Loading fails loudly when the library has not been built: there is no Python fallback for any
operation of the hot path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpamd.so")

MP_OK = 0
MP_E_INVALID, MP_E_HIP, MP_E_RTC, MP_E_NOMEM, MP_E_CAPACITY, MP_E_STATE = -1, -2, -3, -4, -5, -6
PLAN_STATUS = {0: "running", 1: "solved", 2: "exhausted", 3: "tree_full", 4: "invalid", 5: "path_lost",
               6: "path_too_long"}


class MPError(RuntimeError):
    pass


class RobotDesc(C.Structure):
    _fields_ = [("name", C.c_char_p), ("num_links", C.c_int), ("joint_types", C.c_void_p),
                ("joint_origins", C.c_void_p), ("joint_axes", C.c_void_p), ("parents", C.c_void_p),
                ("active", C.c_void_p), ("lower", C.c_void_p), ("upper", C.c_void_p), ("defaults", C.c_void_p),
                ("num_spheres", C.c_int), ("sphere_link", C.c_void_p), ("sphere_pos", C.c_void_p),
                ("sphere_radius", C.c_void_p), ("collision_enabled", C.c_void_p)]


class RRTCParams(C.Structure):
    _fields_ = [("resolution", C.c_float), ("max_step", C.c_int), ("max_iterations", C.c_int),
                ("tree_capacity", C.c_int), ("waves", C.c_int), ("max_path", C.c_int), ("seed", C.c_uint64),
                ("exclusive", C.c_int), ("team_waves", C.c_int), ("teams_per_query", C.c_int),
                ("lanes_per_iteration", C.c_int), ("sampler", C.c_int), ("dynamic_domain", C.c_int),
                ("dd_radius", C.c_float), ("dd_alpha", C.c_float), ("dd_min_radius", C.c_float),
                ("balance", C.c_int), ("tree_ratio", C.c_float)]


BOUNDS_ROBOT, BOUNDS_NONE, BOUNDS_CUSTOM = 0, 1, 2  # mp_scene_set_joint_bounds modes
BALANCE_SMALLER, BALANCE_SPLIT, BALANCE_DYNAMIC, BALANCE_ALTERNATE = 0, 1, 2, 3


SAMPLER_UNIFORM, SAMPLER_HALTON = 0, 1


class Path(C.Structure):
    _fields_ = [("status", C.c_int), ("num_waypoints", C.c_int), ("waypoints", C.c_void_p), ("capacity", C.c_int),
                ("iterations", C.c_int), ("start_tree_size", C.c_int), ("goal_tree_size", C.c_int),
                ("configs_checked", C.c_int64), ("latency_ms", C.c_double), ("work_flops", C.c_double),
                ("device_ms", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("team_waves", C.c_int), ("waves", C.c_int), ("teams_per_query", C.c_int),
                ("lanes_per_iteration", C.c_int), ("kernel_ms", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("num_robots", C.c_int), ("robots", C.c_void_p), ("tree_memory_bytes", C.c_size_t),
                ("tree_capacity", C.c_int), ("max_active", C.c_int), ("step_iterations", C.c_int), ("waves", C.c_int),
                ("use_graph", C.c_int), ("params", RRTCParams)]


class RoadmapInfo(C.Structure):
    _fields_ = [("solved", C.c_int), ("passes", C.c_int), ("loop_index", C.c_int), ("vertices", C.c_int),
                ("edges", C.c_int), ("start_group", C.c_int), ("goal_group", C.c_int), ("travel_start", C.c_float),
                ("travel_goal", C.c_float)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("solved", C.c_int64), ("active", C.c_int),
                ("queued", C.c_int), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("throughput_qps", C.c_double), ("steps", C.c_int64), ("step_ms", C.c_double),
                ("loop_ms", C.c_double), ("backlog_steps", C.c_int64)]


# name -> (restype, argtypes)
P = C.c_void_p
_SIGS = {
    "mp_init": (C.c_int, [C.c_int]),
    "mp_shutdown": (C.c_int, []),
    "mp_last_error": (C.c_char_p, []),
    "mp_set_cache_dir": (C.c_int, [C.c_char_p]),
    "mp_device_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "mp_device_free": (C.c_int, [P]),
    "mp_memcpy_h2d": (C.c_int, [P, P, C.c_size_t, P]),
    "mp_memcpy_d2h": (C.c_int, [P, P, C.c_size_t, P]),
    "mp_memset_d8": (C.c_int, [P, C.c_int, C.c_size_t, P]),
    "mp_stream_sync": (C.c_int, [P]),
    "mp_default_stream": (P, []),
    "mp_last_kernel_ms": (C.c_double, [C.c_char_p]),
    "mp_work_flops": (C.c_int, [C.POINTER(C.c_uint64), C.c_int]),
    "mp_plan_profile": (C.c_int, [C.c_int, P, C.c_int]),
    "mp_robot_create": (C.c_int, [C.POINTER(RobotDesc), C.POINTER(C.c_void_p)]),
    "mp_robot_destroy": (C.c_int, [P]),
    "mp_robot_info": (C.c_int, [P, P, P, P, P]),
    "mp_robot_source": (C.c_size_t, [P, P, C.c_size_t]),
    "mp_robot_precompile": (C.c_int, [C.POINTER(RobotDesc)]),
    "mp_precompile": (C.c_int, [P, C.c_int]),
    "mp_generate_source": (C.c_size_t, [P, C.c_int, P, C.c_size_t]),
    "mp_scene_create": (C.c_int, [P, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int, C.POINTER(C.c_void_p)]),
    "mp_scene_update": (C.c_int, [P, P, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int]),
    "mp_scene_destroy": (C.c_int, [P]),
    "mp_scene_set_joint_bounds": (C.c_int, [P, C.c_int, P, P, C.c_int]),
    "mp_fk_spheres": (C.c_int, [P, P, C.c_int64, P, P]),
    "mp_fk_link_poses": (C.c_int, [P, P, C.c_int64, P, P]),
    "mp_check_states": (C.c_int, [P, P, P, C.c_int64, P, P, P]),
    "mp_state_costs": (C.c_int, [P, P, P, C.c_int64, P, P, P]),
    "mp_constraint_costs": (C.c_int, [P, P, P, C.c_int64, P, P, P, P]),
    "mp_check_edges": (C.c_int, [P, P, P, P, C.c_int, C.c_float, P, P, P]),
    "mp_nn_argmin": (C.c_int, [P, C.c_int, C.c_int, C.c_int, P, C.c_int, P, P]),
    "mp_sample": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_int64, P, P]),
    "mp_sample_halton": (C.c_int, [P, C.c_uint64, C.c_uint64, C.c_int64, P, P]),
    "mp_steer": (C.c_int, [P, P, C.c_int, C.c_int, C.c_float, P]),
    "mp_rrtc_default_params": (None, [C.POINTER(RRTCParams)]),
    "mp_rrtc_solve": (C.c_int, [P, P, P, P, C.POINTER(RRTCParams), C.POINTER(Path)]),
    "mp_rrtc_solve_batch": (C.c_int, [C.c_int, P, P, P, P, C.POINTER(RRTCParams), P]),
    "mp_rrtc_solve_noalloc": (C.c_int, [P, P, P, P, C.POINTER(RRTCParams), C.POINTER(Path)]),
    "mp_rrtc_autotune": (C.c_int, [C.c_int, P, P, P, P, C.POINTER(RRTCParams), C.c_int, C.c_int, P, C.c_int,
                                   C.POINTER(C.c_int), C.POINTER(RRTCParams)]),
    "mp_roadmap_create": (C.c_int, [P, P, P, C.c_int, P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, C.c_uint64,
                                    C.POINTER(C.c_void_p)]),
    "mp_roadmap_iterate": (C.c_int, [P, C.POINTER(C.c_int)]),
    "mp_roadmap_info_get": (C.c_int, [P, C.POINTER(RoadmapInfo)]),
    "mp_roadmap_read": (C.c_int, [P, P, P, P, P, P, P]),
    "mp_roadmap_destroy": (C.c_int, [P]),
    "mp_server_create": (C.c_int, [C.POINTER(ServerParams), C.POINTER(C.c_void_p)]),
    "mp_server_submit": (C.c_int, [P, C.c_int, P, P, P, C.c_uint64, C.POINTER(C.c_int64)]),
    "mp_server_poll": (C.c_int, [P, C.c_int64, C.POINTER(Path)]),
    "mp_server_wait_idle": (C.c_int, [P, C.c_double]),
    "mp_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
    "mp_server_destroy": (C.c_int, [P]),
}
EXPORTED = sorted(_SIGS)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MPError("native library %s is not built (run isaac_sim_ros2_humble_amd/build.py)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != MP_OK:
        msg = lib().mp_last_error()
        raise MPError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def robot_desc(m):
    """RobotDesc for a model.RobotModel; the returned tuple keeps the arrays alive."""
    keep = [np.ascontiguousarray(m.joint_types, np.int32), np.ascontiguousarray(m.joint_origins, np.float64),
            np.ascontiguousarray(m.joint_axes, np.float64), np.ascontiguousarray(m.parents, np.int32),
            np.ascontiguousarray(m.active, np.int32), np.ascontiguousarray(m.lower, np.float32),
            np.ascontiguousarray(m.upper, np.float32), np.ascontiguousarray(m.defaults, np.float32),
            np.ascontiguousarray(m.sphere_link, np.int32), np.ascontiguousarray(m.sphere_pos, np.float32),
            np.ascontiguousarray(m.sphere_radius, np.float32), np.ascontiguousarray(m.enabled, np.uint8)]
    d = RobotDesc(m.name.encode(), m.L, _p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]), _p(keep[5]),
                  _p(keep[6]), _p(keep[7]), m.S, _p(keep[8]), _p(keep[9]), _p(keep[10]), _p(keep[11]))
    return d, keep


def precompile(models):
    descs = [robot_desc(m) for m in models]
    arr = (C.POINTER(RobotDesc) * len(descs))(*[C.pointer(d[0]) for d in descs])
    check(lib().mp_precompile(C.cast(arr, C.c_void_p), len(descs)), "mp_precompile")


def generate_source(models):
    descs = [robot_desc(m) for m in models]
    arr = (C.POINTER(RobotDesc) * len(descs))(*[C.pointer(d[0]) for d in descs])
    n = lib().mp_generate_source(C.cast(arr, C.c_void_p), len(descs), None, 0)
    if n == 0:
        raise MPError("mp_generate_source failed: %s" % lib().mp_last_error().decode())
    buf = C.create_string_buffer(n + 1)
    lib().mp_generate_source(C.cast(arr, C.c_void_p), len(descs), buf, n + 1)
    return buf.value.decode()


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        check(lib().mp_init(device), "mp_init")
        _initialised = True


class DeviceBuffer:
    """Device allocation owned by the library (hipMalloc through the C ABI)."""

    def __init__(self, nbytes):
        init()
        self.nbytes = int(nbytes)
        ptr = C.c_void_p()
        check(lib().mp_device_alloc(C.byref(ptr), max(self.nbytes, 1)), "mp_device_alloc")
        self.ptr = ptr.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().mp_memcpy_h2d(b.ptr, _p(a), a.nbytes, None), "mp_memcpy_h2d")
        return b

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype)
        if out.nbytes:
            check(lib().mp_memcpy_d2h(_p(out), self.ptr, out.nbytes, None), "mp_memcpy_d2h")
        return out

    def free(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.mp_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class Robot:
    def __init__(self, model):
        init()
        self.model = model
        d, self._keep = robot_desc(model)
        h = C.c_void_p()
        check(lib().mp_robot_create(C.byref(d), C.byref(h)), "mp_robot_create")
        self.h = h.value
        dof, ns, npairs, nl = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().mp_robot_info(self.h, C.byref(dof), C.byref(ns), C.byref(npairs), C.byref(nl)), "mp_robot_info")
        self.dof, self.S, self.P, self.L = dof.value, ns.value, npairs.value, nl.value

    def source(self):
        n = lib().mp_robot_source(self.h, None, 0)
        buf = C.create_string_buffer(n + 1)
        lib().mp_robot_source(self.h, buf, n + 1)
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mp_robot_destroy(self.h)


class Scene:
    def __init__(self, scene):
        init()
        self.scene = scene
        self._keep = [scene.spheres, scene.boxes, scene.cylinders, scene.capsules]
        ns, nb, nc, ncap = scene.counts
        h = C.c_void_p()
        check(lib().mp_scene_create(_p(scene.spheres), ns, _p(scene.boxes), nb, _p(scene.cylinders), nc,
                                    _p(scene.capsules), ncap, C.byref(h)), "mp_scene_create")
        self.h = h.value
        if getattr(scene, "joint_bounds", None) is not None:
            self.set_joint_bounds(scene.joint_bounds)

    def update(self, scene):
        ns, nb, nc, ncap = scene.counts
        check(lib().mp_scene_update(self.h, _p(scene.spheres), ns, _p(scene.boxes), nb, _p(scene.cylinders), nc,
                                    _p(scene.capsules), ncap), "mp_scene_update")
        self.set_joint_bounds(getattr(scene, "joint_bounds", None))
        self.scene = scene
        self._keep = [scene.spheres, scene.boxes, scene.cylinders, scene.capsules]

    def set_joint_bounds(self, bounds):
        """None: the robot's joint limits; "none": no joint-bound check; (lower, upper): bounds
        over the active joints."""
        if bounds is None:
            check(lib().mp_scene_set_joint_bounds(self.h, BOUNDS_ROBOT, None, None, 0), "mp_scene_set_joint_bounds")
        elif isinstance(bounds, str):
            if bounds != "none":
                raise ValueError("joint bounds: None, 'none' or (lower, upper)")
            check(lib().mp_scene_set_joint_bounds(self.h, BOUNDS_NONE, None, None, 0), "mp_scene_set_joint_bounds")
        else:
            lo = np.ascontiguousarray(bounds[0], np.float32).reshape(-1)
            hi = np.ascontiguousarray(bounds[1], np.float32).reshape(-1)
            if lo.shape != hi.shape:
                raise ValueError("joint bounds: lower and upper differ in length")
            check(lib().mp_scene_set_joint_bounds(self.h, BOUNDS_CUSTOM, _p(lo), _p(hi), len(lo)),
                  "mp_scene_set_joint_bounds")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mp_scene_destroy(self.h)


def work_flops(reset=True):
    v = C.c_uint64()
    check(lib().mp_work_flops(C.byref(v), 1 if reset else 0), "mp_work_flops")
    return v.value


PROFILE_PHASES = ["sched", "nn_extend", "check_extend", "insert", "nn_connect", "check_connect", "path", "iteration",
                  "wave", "n_iterations", "n_inserts", "n_sleeps"]


def plan_profile(enable=None):
    """Reads and resets the planner phase counters (dict, ticks of 10 ns); enable: True/False/None."""
    v = (C.c_uint64 * len(PROFILE_PHASES))()
    e = -1 if enable is None else (1 if enable else 0)
    check(lib().mp_plan_profile(e, C.cast(v, C.c_void_p), len(PROFILE_PHASES)), "mp_plan_profile")
    return dict(zip(PROFILE_PHASES, [int(x) for x in v]))


def kernel_ms(family):
    return lib().mp_last_kernel_ms(family.encode())


def sync():
    check(lib().mp_stream_sync(None), "mp_stream_sync")


# ---------------------------------------------------------------------------------------------
# host-array conveniences (copy in, run, copy out); the timed benchmarks use device buffers
# ---------------------------------------------------------------------------------------------
def fk_spheres(robot, q):
    q = np.ascontiguousarray(q, np.float32).reshape(-1, robot.dof)
    n = len(q)
    dq = DeviceBuffer.from_array(q)
    dc = DeviceBuffer(n * robot.S * 12)
    check(lib().mp_fk_spheres(robot.h, dq.ptr, n, dc.ptr, None), "mp_fk_spheres")
    return dc.to_array(np.float32, (n, robot.S, 3))


def fk_link_poses(robot, q):
    """Pose of every link in the base frame, [n][L][4][4] (row-major, last row 0 0 0 1)."""
    q = np.ascontiguousarray(q, np.float32).reshape(-1, robot.dof)
    n = len(q)
    dq = DeviceBuffer.from_array(q)
    dp = DeviceBuffer(n * robot.L * 64)
    check(lib().mp_fk_link_poses(robot.h, dq.ptr, n, dp.ptr, None), "mp_fk_link_poses")
    return dp.to_array(np.float32, (n, robot.L, 4, 4))


def check_states(robot, scene, q, want_centres=False):
    q = np.ascontiguousarray(q, np.float32).reshape(-1, robot.dof)
    n = len(q)
    dq = DeviceBuffer.from_array(q)
    dv = DeviceBuffer(n)
    dc = DeviceBuffer(n * robot.S * 12) if want_centres else None
    check(lib().mp_check_states(robot.h, scene.h, dq.ptr, n, dv.ptr, dc.ptr if dc else None, None), "mp_check_states")
    v = dv.to_array(np.uint8, (n,))
    return (v, dc.to_array(np.float32, (n, robot.S, 3))) if want_centres else v


def state_costs(robot, scene, q):
    q = np.ascontiguousarray(q, np.float32).reshape(-1, robot.dof)
    n = len(q)
    dq = DeviceBuffer.from_array(q)
    dcost = DeviceBuffer(4 * n)
    dv = DeviceBuffer(n)
    check(lib().mp_state_costs(robot.h, scene.h, dq.ptr, n, dcost.ptr, dv.ptr, None), "mp_state_costs")
    return dcost.to_array(np.float32, (n,)), dv.to_array(np.uint8, (n,))


COST_KINDS = ("env_sphere", "env_cuboid", "env_cylinder", "env_capsule", "self", "bounds")


def constraint_costs(robot, scene, q):
    """Per-constraint costs [6][n] (rows in COST_KINDS order), total cost [n], verdict [n]."""
    q = np.ascontiguousarray(q, np.float32).reshape(-1, robot.dof)
    n = len(q)
    dq = DeviceBuffer.from_array(q)
    dparts = DeviceBuffer(4 * n * len(COST_KINDS))
    dcost = DeviceBuffer(4 * n)
    dv = DeviceBuffer(n)
    check(lib().mp_constraint_costs(robot.h, scene.h, dq.ptr, n, dparts.ptr, dcost.ptr, dv.ptr, None),
          "mp_constraint_costs")
    return (dparts.to_array(np.float32, (len(COST_KINDS), n)), dcost.to_array(np.float32, (n,)),
            dv.to_array(np.uint8, (n,)))


def check_edges(robot, scene, q0, q1, resolution):
    q0 = np.ascontiguousarray(q0, np.float32).reshape(-1, robot.dof)
    q1 = np.ascontiguousarray(q1, np.float32).reshape(-1, robot.dof)
    m = len(q0)
    a, b = DeviceBuffer.from_array(q0), DeviceBuffer.from_array(q1)
    ok, ln = DeviceBuffer(m), DeviceBuffer(4 * m)
    check(lib().mp_check_edges(robot.h, scene.h, a.ptr, b.ptr, m, resolution, ok.ptr, ln.ptr, None), "mp_check_edges")
    return ok.to_array(np.uint8, (m,)), ln.to_array(np.float32, (m,))


def nn_argmin(tree, queries, capacity=None):
    """tree: [n, dof] host array (stored SoA on device); queries [nq, dof]."""
    tree = np.ascontiguousarray(tree, np.float32)
    n, dof = tree.shape
    cap = capacity or max(n, 1)
    soa = np.zeros((dof, cap), np.float32)
    soa[:, :n] = tree.T
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, dof)
    dt, dq = DeviceBuffer.from_array(soa), DeviceBuffer.from_array(q)
    di = DeviceBuffer(4 * len(q))
    check(lib().mp_nn_argmin(dt.ptr, n, cap, dof, dq.ptr, len(q), di.ptr, None), "mp_nn_argmin")
    return di.to_array(np.int32, (len(q),))


def sample(robot, seed, offset, n):
    dq = DeviceBuffer(4 * n * robot.dof)
    check(lib().mp_sample(robot.h, seed, offset, n, dq.ptr, None), "mp_sample")
    return dq.to_array(np.float32, (n, robot.dof))


def sample_halton(robot, seed, offset, n):
    dq = DeviceBuffer(4 * n * robot.dof)
    check(lib().mp_sample_halton(robot.h, seed, offset, n, dq.ptr, None), "mp_sample_halton")
    return dq.to_array(np.float32, (n, robot.dof))


def steer(frm, to, max_distance):
    frm = np.ascontiguousarray(frm, np.float32)
    to = np.ascontiguousarray(to, np.float32)
    n, dof = frm.shape
    a, b = DeviceBuffer.from_array(frm), DeviceBuffer.from_array(to)
    check(lib().mp_steer(a.ptr, b.ptr, n, dof, max_distance, None), "mp_steer")
    return b.to_array(np.float32, (n, dof))


def default_params(**kw):
    p = RRTCParams()
    lib().mp_rrtc_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _path_result(pth, wp, dof):
    n = max(pth.num_waypoints, 0) if pth.status == 1 else 0
    return {"status": PLAN_STATUS.get(pth.status, str(pth.status)), "path": wp[:n].reshape(-1, dof).copy(),
            "iterations": pth.iterations, "start_tree": pth.start_tree_size, "goal_tree": pth.goal_tree_size,
            "configs_checked": pth.configs_checked, "latency_ms": pth.latency_ms, "work_flops": pth.work_flops,
            "device_ms": pth.device_ms}


def rrtc_solve(robot, scene, start, goal, params=None, max_path=4096):
    params = params or default_params()
    wp = np.empty((params.max_path, robot.dof), np.float32)
    pth = Path(0, 0, _p(wp), params.max_path, 0, 0, 0, 0, 0.0, 0.0, 0.0)
    s = np.ascontiguousarray(start, np.float32)
    g = np.ascontiguousarray(goal, np.float32)
    check(lib().mp_rrtc_solve(robot.h, scene.h, _p(s), _p(g), C.byref(params), C.byref(pth)), "mp_rrtc_solve")
    return _path_result(pth, wp, robot.dof)


# numpy view of an mp_path array (layout checked against the ctypes structure at import)
PATH_DTYPE = np.dtype([("status", "<i4"), ("num_waypoints", "<i4"), ("waypoints", "<u8"), ("capacity", "<i4"),
                       ("iterations", "<i4"), ("start_tree_size", "<i4"), ("goal_tree_size", "<i4"),
                       ("configs_checked", "<i8"), ("latency_ms", "<f8"), ("work_flops", "<f8"), ("device_ms", "<f8")])
assert PATH_DTYPE.itemsize == C.sizeof(Path) and all(
    PATH_DTYPE.fields[n][1] == getattr(Path, n).offset for n, _ in Path._fields_)

class BatchRequest:
    """A set of planning queries packed once for repeated solves: start/goal vectors (16-float
    rows), robot and scene handles, an mp_path array and its waypoint storage.

    solve(params): all queries in one persistent launch (mp_rrtc_solve_batch).
    solve_each(params): the same queries one launch each (mp_rrtc_solve), query k with seed
    params.seed + k, as in the batch.
    Both return a BatchResult; a result's paths stay valid until the request is solved again."""

    def __init__(self, robots, scenes, starts, goals, max_path=4096):
        self.nq = len(robots)
        self.dof = np.array([r.dof for r in robots], np.int64)
        self.s16 = np.zeros((self.nq, 16), np.float32)
        self.g16 = np.zeros((self.nq, 16), np.float32)
        for k in range(self.nq):
            self.s16[k, :self.dof[k]] = starts[k]
            self.g16[k, :self.dof[k]] = goals[k]
        self.rh = (C.c_void_p * self.nq)(*[r.h for r in robots])
        self.sh = (C.c_void_p * self.nq)(*[s.h for s in scenes])
        self.max_path = max_path
        self.stride = max_path * 16
        self.block = np.zeros(self.nq * self.stride, np.float32)
        self.paths = np.zeros(self.nq, PATH_DTYPE)
        self.paths["waypoints"] = self.block.ctypes.data + np.arange(self.nq, dtype=np.uint64) * np.uint64(self.stride * 4)
        self.paths["capacity"] = max_path
        self.generation = 0
        self._keep = (list(robots), list(scenes))

    def _check_params(self, params):
        if params.max_path > self.max_path:
            raise MPError("params.max_path %d exceeds the request's %d" % (params.max_path, self.max_path))

    def solve(self, params=None):
        params = params or default_params(max_path=self.max_path)
        self._check_params(params)
        check(lib().mp_rrtc_solve_batch(self.nq, C.cast(self.rh, C.c_void_p), C.cast(self.sh, C.c_void_p),
                                        _p(self.s16), _p(self.g16), C.byref(params),
                                        self.paths.ctypes.data_as(C.c_void_p)), "mp_rrtc_solve_batch")
        self.generation += 1
        return BatchResult(self)

    def solve_each(self, params=None):
        params = params or default_params(max_path=self.max_path)
        self._check_params(params)
        p = RRTCParams.from_buffer_copy(params)
        base = params.seed
        L = lib()
        s0, g0, p0 = self.s16.ctypes.data, self.g16.ctypes.data, self.paths.ctypes.data
        for k in range(self.nq):
            p.seed = base + k
            check(L.mp_rrtc_solve(self.rh[k], self.sh[k], s0 + 64 * k, g0 + 64 * k, C.byref(p),
                                  C.cast(p0 + PATH_DTYPE.itemsize * k, C.POINTER(Path))), "mp_rrtc_solve")
        self.generation += 1
        return BatchResult(self)


class BatchResult:
    """Per-query outcome columns of a solved BatchRequest (numpy arrays) and path access."""

    def __init__(self, req):
        self._req = req
        self._gen = req.generation
        r = req.paths.copy()
        self.code = r["status"]
        self._status = None
        self.solved = int((self.code == 1).sum())
        self.num_waypoints = r["num_waypoints"]
        self.iterations = r["iterations"]
        self.start_tree = r["start_tree_size"]
        self.goal_tree = r["goal_tree_size"]
        self.configs_checked = r["configs_checked"]
        self.latency_ms = r["latency_ms"]
        self.work_flops = r["work_flops"]
        self.device_ms = r["device_ms"]

    @property
    def status(self):
        """Status names per query (from the codes, built on first use)."""
        if self._status is None:
            self._status = [PLAN_STATUS.get(c, str(c)) for c in self.code.tolist()]
        return self._status

    def path(self, k):
        if self._req.generation != self._gen:
            raise MPError("the request has been solved again: its paths were overwritten")
        n = max(int(self.num_waypoints[k]), 0) if self.code[k] == 1 else 0
        d = int(self._req.dof[k])
        o = k * self._req.stride
        return self._req.block[o:o + n * d].reshape(n, d).copy()

    def to_dicts(self):
        return [{"status": self.status[k], "path": self.path(k), "iterations": int(self.iterations[k]),
                 "start_tree": int(self.start_tree[k]), "goal_tree": int(self.goal_tree[k]),
                 "configs_checked": int(self.configs_checked[k]), "latency_ms": float(self.latency_ms[k]),
                 "work_flops": float(self.work_flops[k]), "device_ms": float(self.device_ms[k])}
                for k in range(len(self.code))]


def rrtc_solve_batch(robots, scenes, starts, goals, params=None):
    params = params or default_params()
    return BatchRequest(robots, scenes, starts, goals, max_path=params.max_path).solve(params).to_dicts()


def rrtc_autotune(robots, scenes, starts, goals, params=None, per_query=False, reps=2):
    """Tunes the planning launch shape on the sample (mp_rrtc_autotune); returns the best params
    and the candidates as dicts (team_waves, waves, teams_per_query, lanes_per_iteration, kernel_ms)."""
    params = params or default_params()
    nq = len(robots)
    s16 = np.zeros((nq, 16), np.float32)
    g16 = np.zeros((nq, 16), np.float32)
    for k in range(nq):
        s16[k, :robots[k].dof] = starts[k]
        g16[k, :robots[k].dof] = goals[k]
    rh = (C.c_void_p * nq)(*[r.h for r in robots])
    sh = (C.c_void_p * nq)(*[s.h for s in scenes])
    res = (TuneResult * 128)()
    n = C.c_int()
    best = RRTCParams()
    check(lib().mp_rrtc_autotune(nq, C.cast(rh, C.c_void_p), C.cast(sh, C.c_void_p), _p(s16), _p(g16), C.byref(params),
                                 1 if per_query else 0, reps, C.cast(res, C.c_void_p), 128, C.byref(n), C.byref(best)),
          "mp_rrtc_autotune")
    return best, [{"team_waves": r.team_waves, "waves": r.waves, "teams_per_query": r.teams_per_query,
                   "lanes_per_iteration": r.lanes_per_iteration, "kernel_ms": r.kernel_ms}
                  for r in res[:min(n.value, 128)]]


class Roadmap:
    """Device-resident roadmap planner loop (mp_roadmap_*): the robot and scene outlive it."""

    def __init__(self, robot, scene, starts, goals, samples_per_iteration=30, max_travel_distance=0.5,
                 resolution=0.02, seed=0x5eed, sample_offset=0):
        init()
        self.robot, self.scene = robot, scene
        s = np.ascontiguousarray(starts, np.float32).reshape(-1, robot.dof)
        g = np.ascontiguousarray(goals, np.float32).reshape(-1, robot.dof)
        h = C.c_void_p()
        check(lib().mp_roadmap_create(robot.h, scene.h, _p(s), len(s), _p(g), len(g), samples_per_iteration,
                                      max_travel_distance, resolution, seed, sample_offset, C.byref(h)),
              "mp_roadmap_create")
        self.h = h.value

    def iterate(self):
        """Enqueues one loop pass; True once a finished pass has joined the groups."""
        s = C.c_int()
        check(lib().mp_roadmap_iterate(self.h, C.byref(s)), "mp_roadmap_iterate")
        return bool(s.value)

    def info(self):
        i = RoadmapInfo()
        check(lib().mp_roadmap_info_get(self.h, C.byref(i)), "mp_roadmap_info_get")
        return {f: getattr(i, f) for f, _ in RoadmapInfo._fields_}

    def read(self):
        inf = self.info()
        nv, ne = inf["vertices"], inf["edges"]
        inf["nv"], inf["ne"] = nv, ne  # counts; "vertices" and "edges" below are the arrays
        v = np.zeros((nv, self.robot.dof), np.float32)
        grp = np.zeros(nv, np.int32)
        e = np.zeros((ne, 2), np.int32)
        w = np.zeros(ne, np.float32)
        sm = np.zeros(inf["start_group"], np.int32)
        gm = np.zeros(inf["goal_group"], np.int32)
        check(lib().mp_roadmap_read(self.h, _p(v), _p(grp), _p(e), _p(w), _p(sm), _p(gm)), "mp_roadmap_read")
        inf.update(vertices=v, groups=grp, edges=e, weights=w, start_members=sm, goal_members=gm)
        return inf

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mp_roadmap_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Server:
    """Continuous-batching planning server (mp_server_*)."""

    def __init__(self, robots, tree_memory_bytes=8 << 30, tree_capacity=65536, max_active=512, step_iterations=0,
                 waves=0, use_graph=False, params=None):
        init()
        self.robots = list(robots)
        self._rh = (C.c_void_p * len(self.robots))(*[r.h for r in self.robots])
        p = params or default_params(tree_capacity=tree_capacity)
        p.tree_capacity = tree_capacity
        sp = ServerParams(len(self.robots), C.cast(self._rh, C.c_void_p), tree_memory_bytes, tree_capacity, max_active,
                          step_iterations, waves, 1 if use_graph else 0, p)
        h = C.c_void_p()
        check(lib().mp_server_create(C.byref(sp), C.byref(h)), "mp_server_create")
        self.h = h.value
        self.max_path = p.max_path
        # ticket -> (scene, dof): the Scene object stays alive until its result is handed out
        self._tickets = {}

    def submit(self, robot_index, scene, start, goal, seed=0):
        t = C.c_int64()
        dof = self.robots[robot_index].dof
        s = np.ascontiguousarray(start, np.float32)
        g = np.ascontiguousarray(goal, np.float32)
        if s.shape != (dof,) or g.shape != (dof,):
            raise MPError("start/goal must have the robot's %d joints" % dof)
        check(lib().mp_server_submit(self.h, robot_index, scene.h, _p(s), _p(g), seed, C.byref(t)), "mp_server_submit")
        self._tickets[t.value] = (scene, dof)
        return t.value

    def poll(self, ticket, dof=None):
        """The finished query's result (once), or None while it runs.  The waypoint buffer is
        sized by the ticket's robot; `dof`, if given, must match it."""
        if ticket not in self._tickets:
            return None
        tdof = self._tickets[ticket][1]
        if dof is not None and dof != tdof:
            raise MPError("ticket %d was submitted for a %d-joint robot, not %d" % (ticket, tdof, dof))
        wp = np.zeros((self.max_path, tdof), np.float32)
        pth = Path(0, 0, _p(wp), self.max_path, 0, 0, 0, 0, 0.0, 0.0, 0.0)
        rc = lib().mp_server_poll(self.h, ticket, C.byref(pth))
        if rc < 0:
            check(rc, "mp_server_poll")
        if rc != 1:
            return None
        del self._tickets[ticket]
        return _path_result(pth, wp, tdof)

    def wait_idle(self, timeout_s=60.0):
        check(lib().mp_server_wait_idle(self.h, timeout_s), "mp_server_wait_idle")

    def stats(self):
        st = ServerStats()
        check(lib().mp_server_stats_get(self.h, C.byref(st)), "mp_server_stats_get")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            check(_lib.mp_server_destroy(self.h), "mp_server_destroy")
            self.h = None
            self._tickets = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
