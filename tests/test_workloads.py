"""Workload generators (CPU, with the oracle as the checker)."""
import numpy as np

import oracle
from isaac_sim_ros2_humble_amd import workloads


def _orc(models):
    cache = {m.name: oracle.OracleRobot(m) for m in models}

    def check(m, scene, q):
        return oracle.check(cache[m.name], oracle.OracleScene(scene), q, nthreads=8)[0]

    def edge(m, scene, a, b):
        return oracle.edges(cache[m.name], oracle.OracleScene(scene), a, b, 0.02)[0]
    return check, edge


def test_c2_scene_counts_and_free_space(fr3, orc_fr3):
    sc = workloads.c2_scene(seed=1)
    assert sc.counts == (12, 40, 6, 6)
    q = workloads.c2_configs(fr3, 4000, seed=1)
    v = oracle.check(orc_fr3, oracle.OracleScene(sc), q, nthreads=8)[0]
    assert 0.05 < v.mean() < 0.9


def test_c1_scene_yaw_only_boxes():
    sc = workloads.c1_scene(seed=0)
    assert sc.counts == (0, 20, 0, 0)
    for b in sc.boxes:
        R = b[3:12].reshape(3, 3)
        assert abs(R[2, 2] - 1) < 1e-6 and np.allclose(R[2, :2], 0) and np.allclose(R[:2, 2], 0)
        d = np.linalg.norm(b[:3])
        assert 0.3 - 1e-6 <= d <= 0.9 + 1e-6


def test_c4_batch_uneven_difficulty(fr3, fetch):
    check, edge = _orc([fr3, fetch])
    qs = workloads.c4_batch([fr3, fetch], 6, check, edge, seed=2, batch=128)
    assert len(qs) == 6
    for i, (k, scene, s, g, easy) in enumerate(qs):
        m = [fr3, fetch][k]
        assert 5 <= scene.num_obstacles <= 200
        assert check(m, scene, np.stack([s, g])).all()
        free = edge(m, scene, s[None], g[None])[0]
        assert bool(free) == easy == (i % 2 == 0)


def test_c3_shelf_and_queries(fetch):
    scene, shelf = workloads.c3_scene(seed=0)
    assert scene.counts[1] == 4 + 3 + 30  # boards, side and back walls, clutter
    r = oracle.OracleRobot(fetch)
    check, _ = _orc([fetch])
    s, g = workloads.c3_queries(fetch, scene, shelf, 2, check, lambda m, q: r.fk32(q), seed=0, batch=4096)
    gc = workloads.gripper_centres(fetch, r.fk32(g))
    assert np.all(gc[:, 0] > shelf["x0"])


def test_poisson_arrivals_rate():
    t = workloads.poisson_arrivals(1000.0, 5.0, seed=0)
    assert abs(len(t) / 5.0 - 1000.0) < 100 and np.all(np.diff(t) > 0)
