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


def test_sample_pairs_alternate_difficulty(fr3):
    check, edge = _orc([fr3])
    rng = np.random.default_rng(3)
    scene = workloads.random_scene(rng, 40, *workloads.workspace(fr3))
    pairs = workloads.sample_pairs(fr3, scene, rng, check, edge, 8, batch=512)
    assert len(pairs) == 8
    a = np.array([p[0] for p in pairs])
    b = np.array([p[1] for p in pairs])
    assert check(fr3, scene, a).all() and check(fr3, scene, b).all()
    free = edge(fr3, scene, a, b)
    assert list(free[:6]) == [1, 0, 1, 0, 1, 0]


def test_c5_library_pairs_are_free(fr3):
    check, edge = _orc([fr3])
    lib = workloads.c5_library([fr3], 3, 4, check, edge, seed=1, batch=256)
    assert len(lib) == 3
    for k, scene, pairs in lib:
        assert k == 0 and 5 <= scene.num_obstacles <= 200 and len(pairs) >= 2
        q = np.array([x for p in pairs for x in p])
        assert check(fr3, scene, q).all()


def test_c5_stream_reuse_and_uploads():
    lib = [(0, None, [None] * (2 + i % 3)) for i in range(12)]
    st = workloads.c5_stream(lib, 1000.0, 2.0, p_reuse=0.8, resident_max=8, seed=2)
    assert abs(len(st) - 2000) < 200
    up = np.array([a[3] for a in st])
    assert abs(up.mean() - 0.2) < 0.04          # about one query in five brings a new scene
    assert st[0][3]                              # the first query has nothing to reuse
    # a reused scene was uploaded before and is among the last resident_max uploads
    uploads = []
    for t, s, p, u in st:
        if u:
            uploads.append(s)
        else:
            assert s in uploads[-8:]
        assert 0 <= p < len(lib[s][2])
