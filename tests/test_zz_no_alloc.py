"""The no-allocation planning entry point (mp_rrtc_solve_noalloc) against mp_rrtc_solve, through
the C ABI: in exclusive mode (one team, deterministic per seed) both return the same status,
counters and path bit for bit, on two calls in a row, the second with a smaller robot and tree;
a workspace beyond the scratch array is refused with MP_E_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

from isaac_sim_ros2_humble_amd import native, workloads

pytestmark = pytest.mark.gpu


def _call(fn, robot, scene, start, goal, params):
    wp = np.zeros((params.max_path, robot.dof), np.float32)
    pth = native.Path(0, 0, wp.ctypes.data_as(C.c_void_p), params.max_path, 0, 0, 0, 0, 0.0, 0.0, 0.0)
    s = np.ascontiguousarray(start, np.float32)
    g = np.ascontiguousarray(goal, np.float32)
    rc = fn(robot.h, scene.h, s.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p), C.byref(params),
            C.byref(pth))
    return rc, pth, wp


def _query(model, robot, seed):
    rng = np.random.default_rng(seed)
    scene = workloads.random_scene(rng, 40, *workloads.workspace(model))
    ns = native.Scene(scene)

    def check(m, sc, q):
        return native.check_states(robot, ns, q)

    def edge(m, sc, a, b):
        return native.check_edges(robot, ns, a, b, 0.02)[0]
    pair = workloads.sample_pair(model, scene, rng, check, edge, want_free_line=False)
    assert pair is not None
    return ns, pair


def test_noalloc_solve_equals_solve(fr3, fetch):
    native.init(0)
    L = native.lib()
    # first call: Fetch (8 joints) with 32768-node trees; second: FR3 (7 joints), 8192 nodes
    for model, cap, seed in ((fetch, 32768, 3), (fr3, 8192, 4)):
        robot = native.Robot(model)
        scene, (s, g) = _query(model, robot, seed)
        p = native.default_params(exclusive=1, seed=seed, tree_capacity=cap, max_path=1024)
        rc0, ref, wref = _call(L.mp_rrtc_solve, robot, scene, s, g, p)
        rc1, got, wgot = _call(L.mp_rrtc_solve_noalloc, robot, scene, s, g, p)
        assert rc0 == 0 and rc1 == 0, native.lib().mp_last_error()
        assert got.status == ref.status and got.num_waypoints == ref.num_waypoints
        assert (got.iterations, got.start_tree_size, got.goal_tree_size, got.configs_checked) == (
            ref.iterations, ref.start_tree_size, ref.goal_tree_size, ref.configs_checked)
        assert ref.status == 1, "the query should solve"
        n = ref.num_waypoints
        assert np.array_equal(wgot[:n], wref[:n])
        assert np.array_equal(wgot[0], s) and np.array_equal(wgot[n - 1], g)


def test_noalloc_rejects_oversized_workspace(fr3):
    native.init(0)
    robot = native.Robot(fr3)
    scene, (s, g) = _query(fr3, robot, 5)
    p = native.default_params(exclusive=1, tree_capacity=65536)  # 2 x 65536 x 7 floats of trees > 4 MiB
    rc, _, _ = _call(native.lib().mp_rrtc_solve_noalloc, robot, scene, s, g, p)
    assert rc == native.MP_E_CAPACITY
    assert b"scratch" in native.lib().mp_last_error()
