"""GPU parity of the data-parallel kernels against the CPU oracle, through the C ABI.

Bar: bit-identical FK sphere centres, verdicts, costs, edge results, samples and NN indices
(the kernels evaluate the same FP32 expressions as the oracle); FK also within 1e-5 m of the
FP64 FK.
"""
import os

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


def test_steer_bit_exact_and_geometry(rng):
    a = rng.normal(size=(1000, 7)).astype(np.float32)
    b = (a + rng.normal(size=(1000, 7)) * rng.uniform(0.01, 0.5, size=(1000, 1))).astype(np.float32)
    out = native.steer(a, b, 0.3)
    assert np.array_equal(out, oracle.steer(a, b, 0.3))
    # the oracle's geometry pins (length min(|b - a|, max), toward b) hold for the GPU result too
    d = np.linalg.norm(b.astype(np.float64) - a, axis=1)
    step = out.astype(np.float64) - a
    far = d > 0.3 * (1 + 1e-6)
    assert far.sum() > 100 and (~far).sum() > 100
    assert np.allclose(np.linalg.norm(step[far], axis=1), 0.3, rtol=1e-5)
    assert np.all((step[far] * (b[far] - a[far])).sum(1) > 0)
    assert np.array_equal(out[d <= 0.3 * (1 - 1e-6)], b[d <= 0.3 * (1 - 1e-6)])


@pytest.mark.parametrize("name", ["fr3", "fetch"])
def test_out_of_bounds_states(name, gpu_robots, request, scenes, rng):
    """Configurations outside the joint limits, inside and outside the 1e-4 tolerance on the
    summed squared violation (BoundaryConstraint.cu:46-93): verdicts, total and per-constraint
    costs identical to the oracle's."""
    m = request.getfixturevalue(name)
    orc = request.getfixturevalue("orc_" + name)
    lo, hi = m.active_lower.astype(np.float64), m.active_upper.astype(np.float64)
    q = workloads.uniform_configs(m, 3000, rng).astype(np.float64)
    n = len(q)
    j = rng.integers(0, m.dof, size=n)
    viol = np.choose(rng.integers(0, 5, size=n), [np.full(n, 0.0099), np.full(n, 0.0101), rng.uniform(0, 0.02, n),
                                                  rng.uniform(0.02, 1.0, n), np.full(n, 0.01)])
    side = rng.random(n) < 0.5
    q[np.arange(n), j] = np.where(side, lo[j] - viol, hi[j] + viol)
    two = rng.random(n) < 0.3  # a second joint out as well
    j2 = (j + 1) % m.dof
    q[two, j2[two]] = hi[j2[two]] + 0.007
    q = q.astype(np.float32)
    sc = scenes[name]
    v = native.check_states(gpu_robots[name], native.Scene(sc), q)
    cost, vc = native.state_costs(gpu_robots[name], native.Scene(sc), q)
    parts, ptot, pv = native.constraint_costs(gpu_robots[name], native.Scene(sc), q)
    vo, co, _, _ = oracle.check(orc, oracle.OracleScene(sc), q, want_cost=True)
    po = oracle.constraint_costs(orc, oracle.OracleScene(sc), q)
    assert np.array_equal(v, vo) and np.array_equal(vc, vo) and np.array_equal(pv, vo)
    assert np.array_equal(cost, co) and np.array_equal(ptot, co)
    assert np.array_equal(parts, po)
    out = parts[5] > 0
    assert 0.2 < out.mean() < 0.95 and (parts[5] == 0).sum() > 100  # both sides of the tolerance


@pytest.mark.parametrize("name", ["fr3", "fetch"])
def test_constraint_costs_identical(name, gpu_robots, request, scenes, c2, rng):
    """Per-constraint rows [6][n] bit-identical to the oracle's on scenes with every obstacle
    kind (the C2 scene has spheres, boxes, cylinders and capsules)."""
    m = request.getfixturevalue(name)
    orc = request.getfixturevalue("orc_" + name)
    q = workloads.uniform_configs(m, 2053, rng)
    for sc in (scenes[name], c2):
        parts, tot, v = native.constraint_costs(gpu_robots[name], native.Scene(sc), q)
        po = oracle.constraint_costs(orc, oracle.OracleScene(sc), q)
        assert np.array_equal(parts, po)
        cost, _ = native.state_costs(gpu_robots[name], native.Scene(sc), q)
        assert np.array_equal(tot, cost)
        assert np.array_equal(v, (cost == 0).astype(np.uint8))
    assert (po[4] > 0).any() and (po[1] > 0).any()


def test_capsule_contact_band():
    """A point robot at distances r + R (1 + k 1e-8) from a capsule's segment: verdicts (the
    fast test plus the exact re-check in the band) and costs equal the oracle's."""
    m = point_robot(0.05)
    r = native.Robot(m)
    orc = oracle.OracleRobot(m)
    from isaac_sim_ros2_humble_amd.scene import capsule_record
    R = 0.1
    scene = Scene(capsules=capsule_record([0.0, 0.0, 0.0], np.eye(3), R, 0.4)[None])
    pts = []
    for k in range(-60, 61):
        d = np.float32(0.15 * (1 + k * 1e-8))
        pts += [[d, 0, 0.05], [0, -d, -0.1], [0, 0, 0.2 + d]]
    pts = np.array(pts, np.float32)
    v = native.check_states(r, native.Scene(scene), pts)
    cost, vc = native.state_costs(r, native.Scene(scene), pts)
    vo, co, _, _ = oracle.check(orc, oracle.OracleScene(scene), pts, want_cost=True)
    assert np.array_equal(v, vo) and np.array_equal(vc, vo) and np.array_equal(cost, co)
    assert 0 < v.sum() < len(v)


def test_self_and_env_rows_point_robot():
    """The two-sphere point robot touching a sphere, a box and its own base sphere: each row
    equal to the oracle's (whose values are pinned by closed forms)."""
    from helpers import point_robot as pr
    m = pr(0.05, base_sphere=(0.0, 0.0, 0.0, 0.16))
    r, orc = native.Robot(m), oracle.OracleRobot(m)
    lo, hi = np.array([0.1, 0.13, -0.05]), np.array([0.2, 0.37, 0.05])
    scene = Scene(spheres=np.array([[0.25, 0.1, 0.0, 0.08]], np.float32),
                  boxes=box_record((lo + hi) / 2, np.eye(3), (hi - lo) / 2)[None])
    q = np.array([[0.15, 0.1, 0.0], [0.5, 0.5, 0.5], [0.12, 0.0, 0.0]], np.float32)
    parts, tot, v = native.constraint_costs(r, native.Scene(scene), q)
    assert np.array_equal(parts, oracle.constraint_costs(orc, oracle.OracleScene(scene), q))
    assert list(v) == [0, 1, 0] and (parts[:, 0] > 0).sum() == 3


@pytest.mark.parametrize("mode", ["robot", "none", "custom"])
def test_scene_joint_bounds_identical(mode, fr3, orc_fr3, gpu_robots, scenes, rng):
    """Joint bounds carried by the scene (the robot's limits, none, or a BoundaryConstraint's
    own): configurations drawn 30% past the limits on every side; verdicts of the early-out and
    full kernels, total costs, the per-constraint rows and edge checks equal the oracle's."""
    lo, hi = fr3.active_lower.astype(np.float64), fr3.active_upper.astype(np.float64)
    span = hi - lo
    q = rng.uniform(lo - 0.3 * span, hi + 0.3 * span, size=(4099, fr3.dof)).astype(np.float32)
    mid = (lo + hi) / 2
    jb = {"robot": None, "none": "none",
          "custom": ((mid - 0.25 * span).astype(np.float32), (mid + 0.25 * span).astype(np.float32))}[mode]
    base = scenes["fr3"]
    sc = Scene(spheres=base.spheres, boxes=base.boxes, cylinders=base.cylinders, capsules=base.capsules,
               joint_bounds=jb)
    ns, so = native.Scene(sc), oracle.OracleScene(sc)
    r = gpu_robots["fr3"]
    v = native.check_states(r, ns, q)
    cost, vc = native.state_costs(r, ns, q)
    parts, tot, vp = native.constraint_costs(r, ns, q)
    vo, co, _, _ = oracle.check(orc_fr3, so, q, want_cost=True, nthreads=8)
    assert np.array_equal(v, vo) and np.array_equal(vc, vo) and np.array_equal(vp, vo)
    assert np.array_equal(cost, co) and np.array_equal(tot, co)
    assert np.array_equal(parts, oracle.constraint_costs(orc_fr3, so, q))
    out = (q < lo) | (q > hi)
    if mode == "none":
        assert not parts[5].any()
    else:
        assert (parts[5] > 0).sum() > 1000
    # edges between consecutive configurations
    ge, gl = native.check_edges(r, ns, q[:-1:4], q[1::4], 0.02)
    oe, ol = oracle.edges(orc_fr3, so, q[:-1:4], q[1::4], 0.02)
    assert np.array_equal(ge, oe) and np.array_equal(gl, ol)
    assert out.any()


def test_planner_follows_scene_joint_bounds(fr3, gpu_robots):
    """A start past the robot's joint limit is invalid under the robot's limits and plannable
    without a joint-bound check (CPRRTC's kernel checks none)."""
    r = gpu_robots["fr3"]
    orc = oracle.OracleRobot(fr3)
    free = oracle.OracleScene(Scene(joint_bounds="none"))
    rng = np.random.default_rng(5)
    for _ in range(1000):  # a start 0.2 past joint 4's upper limit and a goal inside, both free
        start = orc.sample(seed=int(rng.integers(1 << 30)), idx0=0, n=2)
        start, goal = start[0].copy(), start[1].copy()
        start[3] = fr3.active_upper[3] + np.float32(0.2)
        if oracle.check(orc, free, np.stack([start, goal]))[0].all():
            break
    res = {}
    for jb in (None, "none"):
        res[jb] = native.rrtc_solve(r, native.Scene(Scene(joint_bounds=jb)), start, goal, native.default_params(seed=1))
    assert res[None]["status"] == "invalid"
    assert res["none"]["status"] == "solved"
    p = res["none"]["path"]
    assert np.array_equal(p[0], start) and np.array_equal(p[-1], goal)
    assert oracle.validate_path(orc, free, p, 0.005)[0]


def test_c2_full_size_against_oracle(fr3, orc_fr3, gpu_robots, c2):
    """All 2^20 C2 verdicts and sphere centres against the oracle, plus size-independent
    properties: idempotence and chunked == whole-batch evaluation."""
    q = workloads.c2_configs(fr3, 1 << 20, seed=1)
    sc = native.Scene(c2)
    v1, cen = native.check_states(gpu_robots["fr3"], sc, q, want_centres=True)
    vo, _, co, _ = oracle.check(orc_fr3, oracle.OracleScene(c2), q, want_centres=True,
                                nthreads=min(16, os.cpu_count() or 8))
    mism = np.nonzero(v1 != vo)[0]
    assert len(mism) == 0, "verdict mismatches at %s" % mism[:10]
    assert np.array_equal(cen, co)
    del cen, co
    v2 = native.check_states(gpu_robots["fr3"], sc, q)
    assert np.array_equal(v1, v2)  # idempotent
    parts = np.concatenate([native.check_states(gpu_robots["fr3"], sc, q[i:i + 300001]) for i in range(0, len(q), 300001)])
    assert np.array_equal(parts, v1)
    assert 0.05 < v1.mean() < 0.95
