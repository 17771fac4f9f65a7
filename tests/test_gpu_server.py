"""Serving (continuous batching): queries admitted while others run, results identical alone and
in the mix (exclusive mode: one wave per query, deterministic per seed), graph replays equal
eager steps, statistics sane."""
import time

import numpy as np
import pytest

from isaac_sim_ros2_humble_amd import native, workloads

pytestmark = pytest.mark.gpu

_R = {}


def robot(m):
    if m.name not in _R:
        native.init(0)
        _R[m.name] = native.Robot(m)
    return _R[m.name]


def make_queries(models, n, seed):
    def check(m, s, q):
        return native.check_states(robot(m), native.Scene(s), q)

    def edge(m, s, a, b):
        return native.check_edges(robot(m), native.Scene(s), a, b, 0.02)[0]
    return workloads.c4_batch(models, n, check, edge, seed=seed)


@pytest.mark.parametrize("use_graph,lanes", [(False, 64), (True, 64), (True, 32)])
def test_served_equals_solo_exclusive(fr3, fetch, use_graph, lanes):
    models = [fr3, fetch]
    qs = make_queries(models, 12, seed=21)
    scenes = [native.Scene(q[1]) for q in qs]
    solo = [native.rrtc_solve(robot(models[k]), scenes[i], s, g,
                              native.default_params(exclusive=1, seed=100 + i, lanes_per_iteration=lanes))
            for i, (k, _, s, g, _) in enumerate(qs)]
    srv = native.Server([robot(fr3), robot(fetch)], tree_memory_bytes=1 << 30, max_active=5, step_iterations=7,
                        use_graph=use_graph, params=native.default_params(exclusive=1, lanes_per_iteration=lanes))
    tickets = []
    for i, (k, _, s, g, _) in enumerate(qs):
        tickets.append(srv.submit(k, scenes[i], s, g, seed=100 + i))
        time.sleep(0.002)  # arrivals while earlier queries are on the device
    srv.wait_idle(120.0)
    for i, (k, _, s, g, _) in enumerate(qs):
        r = srv.poll(tickets[i], models[k].dof)
        assert r is not None
        assert r["status"] == solo[i]["status"]
        assert np.array_equal(r["path"], solo[i]["path"])
    st = srv.stats()
    assert st["completed"] == len(qs) and st["steps"] > 1
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    srv.close()


@pytest.mark.parametrize("lanes,options", [(0, ""), (32, ""), (32, "halton+domain")])
def test_served_shared_mode_results_valid(fr3, fetch, lanes, options):
    import oracle
    opt = {"sampler": native.SAMPLER_HALTON, "dynamic_domain": 1} if options else {}
    models = [fr3, fetch]
    qs = make_queries(models, 24, seed=5)
    scenes = [native.Scene(q[1]) for q in qs]
    srv = native.Server([robot(fr3), robot(fetch)], tree_memory_bytes=2 << 30, max_active=16, use_graph=True,
                        params=native.default_params(lanes_per_iteration=lanes, **opt))
    tickets = [srv.submit(k, scenes[i], s, g, seed=i) for i, (k, _, s, g, _) in enumerate(qs)]
    srv.wait_idle(120.0)
    nsol, nvalid = 0, 0
    for i, (k, sc, s, g, _) in enumerate(qs):
        r = srv.poll(tickets[i], models[k].dof)
        if r["status"] == "solved":
            nsol += 1
            assert np.array_equal(r["path"][0], s) and np.array_equal(r["path"][-1], g)
            ok, _ = oracle.validate_path(oracle.OracleRobot(models[k]), oracle.OracleScene(sc), r["path"], 0.005)
            nvalid += ok
    assert nsol >= 0.8 * len(qs)
    assert nvalid >= 0.9 * nsol
    assert srv.poll(tickets[0], 7) is None  # each result is handed out once
    srv.close()


def test_scene_pinned_while_queries_use_it(fr3):
    """A scene with queued or running server queries cannot be updated or destroyed (MP_E_STATE);
    the Python server keeps the Scene of each pending ticket alive, so a scene handle the caller
    drops at once stays valid until its result is polled."""
    import oracle
    from helpers import point_robot
    from isaac_sim_ros2_humble_amd.scene import Scene, box_record
    m = point_robot(0.02, lim=1.0)
    walls = [box_record([0.5 + dx, dy, dz], np.eye(3), h) for dx, dy, dz, h in [
        (-0.15, 0, 0, [0.01, 0.2, 0.2]), (0.15, 0, 0, [0.01, 0.2, 0.2]), (0, -0.15, 0, [0.2, 0.01, 0.2]),
        (0, 0.15, 0, [0.2, 0.01, 0.2]), (0, 0, -0.15, [0.2, 0.2, 0.01]), (0, 0, 0.15, [0.2, 0.2, 0.01])]]
    shell = Scene(boxes=np.array(walls))
    srv = native.Server([robot(m)], tree_memory_bytes=256 << 20, tree_capacity=65536, max_active=1,
                        params=native.default_params(max_iterations=400000))
    s, g = np.array([-0.5, 0, 0], np.float32), np.array([0.5, 0, 0], np.float32)
    blocker = native.Scene(shell)
    t1 = srv.submit(0, blocker, s, g, seed=1)          # unreachable goal: holds the only slot
    free_scene = Scene(boxes=box_record([0, 0.8, 0], np.eye(3), [0.05, 0.05, 0.05])[None])
    pinned = native.Scene(free_scene)
    t2 = srv.submit(0, pinned, s, g, seed=2)           # queued behind it
    L = native.lib()
    assert L.mp_scene_destroy(pinned.h) == native.MP_E_STATE
    assert L.mp_scene_update(pinned.h, None, 0, None, 0, None, 0, None, 0) == native.MP_E_STATE
    t3 = srv.submit(0, native.Scene(free_scene), s, g, seed=3)  # the caller keeps no handle
    srv.wait_idle(120.0)
    r1, r2, r3 = srv.poll(t1), srv.poll(t2), srv.poll(t3)
    assert r1["status"] in ("exhausted", "tree_full")
    for r in (r2, r3):
        assert r["status"] == "solved"
        ok, _ = oracle.validate_path(oracle.OracleRobot(m), oracle.OracleScene(free_scene), r["path"], 0.005)
        assert ok
    pinned.update(free_scene)                          # released once its query finished
    srv.close()
