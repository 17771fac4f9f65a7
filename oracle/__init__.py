"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker
or the timed CPU baseline.  The product (isaac_sim_ros2_humble_amd) never imports this package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        fp = np.ctypeslib.ndpointer
        L.orc_robot_create.restype = P
        L.orc_robot_create.argtypes = [C.c_int, P, P, P, P, P, P, P, P, C.c_int, P, P, P, P]
        L.orc_robot_destroy.argtypes = [P]
        L.orc_robot_dof.argtypes = [P]
        L.orc_robot_num_pairs.argtypes = [P]
        L.orc_robot_pairs.argtypes = [P, P, P, P]
        L.orc_full_flops.argtypes = [P, P]
        L.orc_full_flops.restype = C.c_double
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [P, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_set_joint_bounds.argtypes = [P, C.c_int, P, P, C.c_int]
        L.orc_sincosf.argtypes = [C.c_float, P, P]
        L.orc_fk64.argtypes = [P, P, P, P]
        L.orc_fk32.argtypes = [P, P, P]
        L.orc_link_poses32.argtypes = [P, P, P]
        L.orc_check32.argtypes = [P, P, P, C.c_int64, P, P, P, P, C.c_int]
        L.orc_check_centres32.argtypes = [P, P, P, P, C.c_int64, P]
        L.orc_edges32.argtypes = [P, P, P, P, C.c_int, C.c_float, P, P]
        L.orc_nn.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_steer.argtypes = [P, P, C.c_int, C.c_int, C.c_float]
        L.orc_constraint_costs.argtypes = [P, P, P, C.c_int64, P]
        L.orc_rrg.argtypes = [P, P, P, C.c_int, P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, C.c_uint64,
                              C.c_int, P, P, C.c_int, P, P, C.c_int, P, P, P]
        L.orc_validate_path.argtypes = [P, P, P, C.c_int, C.c_float, P]
        L.orc_sample.argtypes = [P, C.c_uint64, C.c_uint64, P]
        L.orc_sample_halton.argtypes = [P, C.c_uint64, C.c_uint64, P]
        L.orc_radical_inverse.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_radical_inverse.restype = C.c_float
        L.orc_philox4x32.argtypes = [C.c_uint32] * 6 + [P]
        L.orc_rrtconnect.argtypes = [P, P, P, P, C.c_float, C.c_float, C.c_uint64, C.c_int, C.c_double, P, C.c_int, P]
        L.orc_cprrtc.argtypes = [P, P, P, P, C.c_float, C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int, P, C.c_int, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class PlanStats(C.Structure):
    _fields_ = [("solved", C.c_int), ("iterations", C.c_int), ("start_tree", C.c_int), ("goal_tree", C.c_int),
                ("states_checked", C.c_int64), ("seconds", C.c_double)]


class OracleRobot:
    def __init__(self, m):
        self.m = m
        self._keep = [np.ascontiguousarray(m.joint_types, np.int32), np.ascontiguousarray(m.joint_origins, np.float64),
                      np.ascontiguousarray(m.joint_axes, np.float64), np.ascontiguousarray(m.parents, np.int32),
                      np.ascontiguousarray(m.active, np.int32), np.ascontiguousarray(m.lower, np.float32),
                      np.ascontiguousarray(m.upper, np.float32), np.ascontiguousarray(m.defaults, np.float32),
                      np.ascontiguousarray(m.sphere_link, np.int32), np.ascontiguousarray(m.sphere_pos, np.float32),
                      np.ascontiguousarray(m.sphere_radius, np.float32), np.ascontiguousarray(m.enabled, np.uint8)]
        k = self._keep
        self.h = lib().orc_robot_create(m.L, _p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]), _p(k[4]), _p(k[5]), _p(k[6]),
                                        _p(k[7]), m.S, _p(k[8]), _p(k[9]), _p(k[10]), _p(k[11]))
        assert self.h, "orc_robot_create failed"
        self.dof = lib().orc_robot_dof(self.h)
        self.S = m.S
        self.P = lib().orc_robot_num_pairs(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_robot_destroy(self.h)

    def pairs(self):
        i = np.zeros(self.P, np.int32)
        j = np.zeros(self.P, np.int32)
        t = np.zeros(self.P, np.float32)
        lib().orc_robot_pairs(self.h, _p(i), _p(j), _p(t))
        return i, j, t

    def fk64(self, q):
        q = np.ascontiguousarray(q, np.float64).reshape(-1, self.dof)
        out = np.zeros((len(q), self.S, 3), np.float64)
        for n in range(len(q)):
            lib().orc_fk64(self.h, _p(q[n]), _p(out[n]), None)
        return out

    def link_poses64(self, q):
        q = np.ascontiguousarray(q, np.float64).reshape(self.dof)
        c = np.zeros((self.S, 3))
        T = np.zeros((self.m.L, 4, 4))
        lib().orc_fk64(self.h, _p(q), _p(c), _p(T))
        return T

    def fk32(self, q):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, self.dof)
        out = np.zeros((len(q), self.S, 3), np.float32)
        for n in range(len(q)):
            lib().orc_fk32(self.h, _p(q[n]), _p(out[n]))
        return out

    def link_poses32(self, q):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, self.dof)
        out = np.zeros((len(q), self.m.L, 4, 4), np.float32)
        for n in range(len(q)):
            lib().orc_link_poses32(self.h, _p(q[n]), _p(out[n]))
        return out

    def sample(self, seed, idx0, n):
        out = np.zeros((n, self.dof), np.float32)
        for k in range(n):
            lib().orc_sample(self.h, seed, idx0 + k, _p(out[k]))
        return out

    def sample_halton(self, seed, idx0, n):
        out = np.zeros((n, self.dof), np.float32)
        for k in range(n):
            lib().orc_sample_halton(self.h, seed, idx0 + k, _p(out[k]))
        return out


class OracleScene:
    def __init__(self, scene):
        self.scene = scene
        self._keep = [scene.spheres, scene.boxes, scene.cylinders, scene.capsules]
        ns, nb, nc, ncap = scene.counts
        self.h = lib().orc_scene_create(_p(scene.spheres), ns, _p(scene.boxes), nb, _p(scene.cylinders), nc,
                                        _p(scene.capsules), ncap)
        jb = getattr(scene, "joint_bounds", None)
        if isinstance(jb, str):
            if jb != "none":
                raise ValueError("joint bounds: None, 'none' or (lower, upper)")
            lib().orc_scene_set_joint_bounds(self.h, 1, None, None, 0)
        elif jb is not None:
            lo = np.ascontiguousarray(jb[0], np.float32).reshape(-1)
            hi = np.ascontiguousarray(jb[1], np.float32).reshape(-1)
            self._keep += [lo, hi]
            lib().orc_scene_set_joint_bounds(self.h, 2, _p(lo), _p(hi), len(lo))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_destroy(self.h)


def check(robot, scene, q, want_cost=False, want_centres=False, want_flops=False, nthreads=1):
    q = np.ascontiguousarray(q, np.float32).reshape(-1, robot.dof)
    n = len(q)
    v = np.zeros(n, np.uint8)
    cost = np.zeros(n, np.float32) if want_cost else None
    cen = np.zeros((n, robot.S, 3), np.float32) if want_centres else None
    fl = np.zeros(n, np.float64) if want_flops else None
    lib().orc_check32(robot.h, scene.h, _p(q), n, _p(v), _p(cost), _p(cen), _p(fl), nthreads)
    return v, cost, cen, fl


COST_KINDS = ("env_sphere", "env_cuboid", "env_cylinder", "env_capsule", "self", "bounds")


def constraint_costs(robot, scene, q):
    """Per-constraint costs [6][n], rows in COST_KINDS order."""
    q = np.ascontiguousarray(q, np.float32).reshape(-1, robot.dof)
    parts = np.zeros((len(COST_KINDS), len(q)), np.float32)
    lib().orc_constraint_costs(robot.h, scene.h, _p(q), len(q), _p(parts))
    return parts


def check_centres(robot, scene, q, centres):
    q = np.ascontiguousarray(q, np.float32).reshape(-1, robot.dof)
    c = np.ascontiguousarray(centres, np.float32)
    v = np.zeros(len(q), np.uint8)
    lib().orc_check_centres32(robot.h, scene.h, _p(q), _p(c), len(q), _p(v))
    return v


def edges(robot, scene, q0, q1, resolution):
    q0 = np.ascontiguousarray(q0, np.float32).reshape(-1, robot.dof)
    q1 = np.ascontiguousarray(q1, np.float32).reshape(-1, robot.dof)
    ok = np.zeros(len(q0), np.uint8)
    ln = np.zeros(len(q0), np.float32)
    lib().orc_edges32(robot.h, scene.h, _p(q0), _p(q1), len(q0), resolution, _p(ok), _p(ln))
    return ok, ln


def nn(tree, q):
    tree = np.ascontiguousarray(tree, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    return lib().orc_nn(_p(tree), tree.shape[0], tree.shape[1], _p(q))


def steer(frm, to, max_distance):
    frm = np.ascontiguousarray(frm, np.float32)
    out = np.array(to, np.float32, copy=True, order="C").reshape(frm.shape)
    lib().orc_steer(_p(frm), _p(out), frm.shape[0], frm.shape[1], max_distance)
    return out


def validate_path(robot, scene, path, resolution=0.005):
    path = np.ascontiguousarray(path, np.float32).reshape(-1, robot.dof)
    bad = C.c_int(0)
    ok = lib().orc_validate_path(robot.h, scene.h, _p(path), len(path), resolution, C.byref(bad))
    return bool(ok), bad.value


def sincosf(x):
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    lib().orc_sincosf(float(x), _p(s), _p(c))
    return float(s[0]), float(c[0])


def philox(c, k):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32(c[0], c[1], c[2], c[3], k[0], k[1], _p(out))
    return out


def radical_inverse(i, b):
    return float(lib().orc_radical_inverse(i, b))


def full_flops(robot, scene):
    return lib().orc_full_flops(robot.h, scene.h)


class RrgInfo(C.Structure):
    _fields_ = [("solved", C.c_int), ("passes", C.c_int), ("loop_index", C.c_int), ("vertices", C.c_int),
                ("edges", C.c_int), ("start_group", C.c_int), ("goal_group", C.c_int), ("travel_start", C.c_float),
                ("travel_goal", C.c_float)]


def rrg(robot, scene, starts, goals, k, max_travel, resolution, seed, offset, max_passes, vcap=1 << 16):
    """The roadmap loop for up to max_passes passes: a dict with the info fields and the graph."""
    starts = np.ascontiguousarray(starts, np.float32).reshape(-1, robot.dof)
    goals = np.ascontiguousarray(goals, np.float32).reshape(-1, robot.dof)
    ecap = 2 * vcap
    v = np.zeros((vcap, robot.dof), np.float32)
    g = np.zeros(vcap, np.int32)
    e = np.zeros((ecap, 2), np.int32)
    w = np.zeros(ecap, np.float32)
    sm = np.zeros(vcap, np.int32)
    gm = np.zeros(vcap, np.int32)
    info = RrgInfo()
    rc = lib().orc_rrg(robot.h, scene.h, _p(starts), len(starts), _p(goals), len(goals), k, max_travel, resolution,
                       seed, offset, max_passes, _p(v), _p(g), vcap, _p(e), _p(w), ecap, _p(sm), _p(gm), C.byref(info))
    assert rc == 0, "orc_rrg: capacity"
    out = {f: getattr(info, f) for f, _ in RrgInfo._fields_}
    nv, ne = info.vertices, info.edges
    out["nv"], out["ne"] = nv, ne  # counts; "vertices" and "edges" below are the arrays
    out.update(vertices=v[:nv].copy(), groups=g[:nv].copy(), edges=e[:ne].copy(), weights=w[:ne].copy(),
               start_members=sm[:info.start_group].copy(), goal_members=gm[:info.goal_group].copy())
    return out


class CprrtcInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("start_tree", C.c_int), ("goal_tree", C.c_int),
                ("path_len", C.c_int)]


CPRRTC_STATUS = {1: "solved", 2: "exhausted", 3: "tree_full", 4: "invalid"}


def cprrtc(robot, scene, start, goal, resolution=0.02, max_step=32, seed=0, max_iterations=10000,
           tree_capacity=65536, lanes=64, out_cap=4096):
    """CPRRTC's planner run sequentially, as the GPU planner runs one query with one team
    (exclusive mode; lanes 64: one iteration per pass, 32: two); see orc_cprrtc.  Returns a dict
    with the keys of the GPU planner's result that the restatement fixes: status, path,
    iterations, start_tree, goal_tree (and path_len, the length of a path too long for out_cap)."""
    start = np.ascontiguousarray(start, np.float32)
    goal = np.ascontiguousarray(goal, np.float32)
    out = np.zeros((out_cap, robot.dof), np.float32)
    inf = CprrtcInfo()
    k = lib().orc_cprrtc(robot.h, scene.h, _p(start), _p(goal), resolution, max_step, seed, max_iterations,
                         tree_capacity, lanes, _p(out), out_cap, C.byref(inf))
    return {"status": CPRRTC_STATUS[inf.status], "path": out[:k].copy(), "iterations": inf.iterations,
            "start_tree": inf.start_tree, "goal_tree": inf.goal_tree, "path_len": inf.path_len}


def rrtconnect(robot, scene, start, goal, range_=0.3, resolution=0.02, seed=0, max_iterations=1_000_000,
               time_limit=10.0, out_cap=100000):
    start = np.ascontiguousarray(start, np.float32)
    goal = np.ascontiguousarray(goal, np.float32)
    out = np.zeros((out_cap, robot.dof), np.float32)
    st = PlanStats()
    k = lib().orc_rrtconnect(robot.h, scene.h, _p(start), _p(goal), range_, resolution, seed, max_iterations,
                             time_limit, _p(out), out_cap, C.byref(st))
    return out[:k].copy(), st
