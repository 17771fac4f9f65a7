"""GPU parity of the data-parallel kernels against the CPU oracle, through the C ABI.

Bar: bit-identical FK sphere centres, verdicts, costs, edge results, samples and NN indices
(the kernels evaluate the same FP32 expressions as the oracle); FK also within 1e-5 m of the
FP64 FK.
"""
import numpy as np
import pytest

import oracle
from helpers import point_robot
from isaac_sim_ros2_humble_amd import native, workloads
from isaac_sim_ros2_humble_amd.scene import Scene, box_record

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_robots(fr3, fetch):
    native.init(0)
    return {"fr3": native.Robot(fr3), "fetch": native.Robot(fetch)}


@pytest.fixture(scope="module")
def c2():
    return workloads.c2_scene(seed=1)


@pytest.fixture(scope="module")
def scenes(c2, fetch):
    # the c2 scene crowds Fetch's base column (every Fetch configuration collides there);
    # Fetch gets 40 mixed obstacles in its own workspace (about 75 % of uniform configs valid)
    return {"fr3": c2, "fetch": workloads.random_scene(np.random.default_rng(1), 40, *workloads.workspace(fetch))}


@pytest.mark.parametrize("name", ["fr3", "fetch"])
def test_fk_bit_exact(name, gpu_robots, request, rng):
    m = request.getfixturevalue(name)
    orc = request.getfixturevalue("orc_" + name)
    q = workloads.uniform_configs(m, 4099, rng)
    g = native.fk_spheres(gpu_robots[name], q)
    c32 = orc.fk32(q)
    assert np.array_equal(g, c32), np.abs(g - c32).max()
    assert np.abs(g.astype(np.float64) - orc.fk64(q.astype(np.float64))).max() < 1e-5


@pytest.mark.parametrize("name", ["fr3", "fetch"])
def test_link_poses_bit_exact(name, gpu_robots, request, rng):
    # every link's pose (ragged last block), equal to the FP32 oracle, within 1e-5 of FP64
    m = request.getfixturevalue(name)
    orc = request.getfixturevalue("orc_" + name)
    q = workloads.uniform_configs(m, 1027, rng)
    g = native.fk_link_poses(gpu_robots[name], q)
    assert g.shape == (1027, m.L, 4, 4)
    p32 = orc.link_poses32(q)
    assert np.array_equal(g, p32), np.abs(g - p32).max()
    for k in (0, 513, 1026):
        assert np.abs(g[k].astype(np.float64) - orc.link_poses64(q[k].astype(np.float64))).max() < 1e-5


@pytest.mark.parametrize("name", ["fr3", "fetch"])
def test_check_states_verdicts_identical(name, gpu_robots, request, scenes, rng):
    m = request.getfixturevalue(name)
    orc = request.getfixturevalue("orc_" + name)
    q = workloads.uniform_configs(m, 16387, rng)
    v, cen = native.check_states(gpu_robots[name], native.Scene(scenes[name]), q, want_centres=True)
    vo, _, _, _ = oracle.check(orc, oracle.OracleScene(scenes[name]), q, nthreads=8)
    mism = np.nonzero(v != vo)[0]
    assert len(mism) == 0, "verdict mismatches at %s" % mism[:10]
    assert np.array_equal(cen, orc.fk32(q))
    assert 0.01 < v.mean() < 0.99


@pytest.mark.parametrize("name", ["fr3", "fetch"])
def test_costs_identical(name, gpu_robots, request, scenes, rng):
    m = request.getfixturevalue(name)
    orc = request.getfixturevalue("orc_" + name)
    q = workloads.uniform_configs(m, 2051, rng)
    sc = scenes[name]
    cost, v = native.state_costs(gpu_robots[name], native.Scene(sc), q)
    vo, co, _, _ = oracle.check(orc, oracle.OracleScene(sc), q, want_cost=True, nthreads=8)
    assert np.array_equal(v, vo)
    assert np.array_equal(cost, co)
    assert 0 < v.sum() < len(v)
    # the early-out verdict kernel agrees with the full evaluation
    assert np.array_equal(native.check_states(gpu_robots[name], native.Scene(sc), q), v)


def test_contact_band_cases():
    m = point_robot(0.05)
    r = native.Robot(m)
    orc = oracle.OracleRobot(m)
    R = 0.25
    scene = Scene(spheres=np.array([[0.0, 0.0, 0.0, R]], np.float32),
                  boxes=box_record([1.0, 0, 0], np.eye(3), [0.1, 0.1, 0.1])[None])
    d = np.float32(0.3)
    pts = []
    for k in range(-40, 41):
        x = np.nextafter(d, np.float32(1), dtype=np.float32) if k == 0 else np.float32(d * (1 + k * 1e-8))
        pts.append([x, 0, 0])
        pts.append([1.0 - 0.15 * (1 + k * 1e-8), 0, 0])
    pts = np.array(pts, np.float32)
    v = native.check_states(r, native.Scene(scene), pts)
    vo, _, _, _ = oracle.check(orc, oracle.OracleScene(scene), pts)
    assert np.array_equal(v, vo)
    assert 0 < v.sum() < len(v)


def test_edges_identical(fr3, orc_fr3, gpu_robots, c2, rng):
    q = workloads.uniform_configs(fr3, 1024, rng)
    q0, q1 = q[:512], q[512:]
    q1 = q0 + (q1 - q0) * rng.uniform(0.0, 0.6, size=(512, 1)).astype(np.float32)  # mixed lengths
    ok, ln = native.check_edges(gpu_robots["fr3"], native.Scene(c2), q0, q1, 0.02)
    oko, lno = oracle.edges(orc_fr3, oracle.OracleScene(c2), q0, q1, 0.02)
    assert np.array_equal(ln, lno)
    assert np.array_equal(ok, oko)
    assert 0 < ok.sum() < len(ok)


def test_edge_skipped_state_semantics():
    m = point_robot(0.001)
    r = native.Robot(m)
    probe = native.Scene(Scene(spheres=np.array([[0.1 * 5 / 6, 0.0, 0.0, 0.004]], np.float32)))
    ok, ln = native.check_edges(r, probe, np.zeros((1, 3), np.float32), np.array([[0.1, 0, 0]], np.float32), 0.02)
    assert ok[0] == 1 and ln[0] == np.float32(0.1)


@pytest.mark.parametrize("name", ["fr3", "fetch"])
def test_sample_bit_exact(name, gpu_robots, request):
    orc = request.getfixturevalue("orc_" + name)
    g = native.sample(gpu_robots[name], seed=0x1234567890, offset=77, n=3000)
    assert np.array_equal(g, orc.sample(0x1234567890, 77, 3000))


@pytest.mark.parametrize("name", ["fr3", "fetch"])
def test_sample_halton_bit_exact(name, gpu_robots, request):
    orc = request.getfixturevalue("orc_" + name)
    for seed, off, n in ((0x1234567890, 0, 3000), (5, 123456, 500), (9, 2**32 - 4, 8)):  # last: index wraps to 0
        g = native.sample_halton(gpu_robots[name], seed=seed, offset=off, n=n)
        assert np.array_equal(g, orc.sample_halton(seed, off, n)), (seed, off)


def test_nn_argmin(rng):
    tree = rng.normal(size=(5000, 8)).astype(np.float32)
    tree[4000] = tree[17]  # duplicate node: ties -> lowest index
    qs = np.concatenate([rng.normal(size=(60, 8)).astype(np.float32), tree[17:18] + 1e-3])
    got = native.nn_argmin(tree, qs, capacity=8192)
    ref = [oracle.nn(tree, q) for q in qs]
    assert list(got) == ref
    assert got[-1] == 17


def test_steer(rng):
    a = rng.normal(size=(100, 7)).astype(np.float32)
    b = rng.normal(size=(100, 7)).astype(np.float32)
    out = native.steer(a, b, 0.5)
    d = np.linalg.norm(out - a, axis=1)
    db = np.linalg.norm(b - a, axis=1)
    assert np.all(d <= 0.5 + 1e-5)
    assert np.allclose(out[db <= 0.5], b[db <= 0.5])


def test_c2_full_size_properties(fr3, orc_fr3, gpu_robots, c2):
    q = workloads.c2_configs(fr3, 1 << 20, seed=1)
    sc = native.Scene(c2)
    v1 = native.check_states(gpu_robots["fr3"], sc, q)
    v2 = native.check_states(gpu_robots["fr3"], sc, q)
    assert np.array_equal(v1, v2)  # idempotent
    # chunked evaluation equals the whole-batch evaluation
    parts = np.concatenate([native.check_states(gpu_robots["fr3"], sc, q[i:i + 300001]) for i in range(0, len(q), 300001)])
    assert np.array_equal(parts, v1)
    # a strided sample against the oracle
    idx = np.arange(0, len(q), 97)
    vo, _, _, _ = oracle.check(orc_fr3, oracle.OracleScene(c2), q[idx], nthreads=8)
    assert np.array_equal(v1[idx], vo)
