"""GPU RRT-Connect (plain and batched): statuses, path validity re-checked by the CPU oracle at
0.005 rad, endpoint identity, determinism per seed in exclusive mode, batch == solo."""
import numpy as np
import pytest

import oracle
from helpers import assert_valid as _assert_valid, point_robot
from isaac_sim_ros2_humble_amd import native, workloads
from isaac_sim_ros2_humble_amd.scene import Scene, box_record

pytestmark = pytest.mark.gpu


def gpu_check(model, scene, q):
    return native.check_states(_robot(model), native.Scene(scene), q)


_ROBOTS = {}


def _robot(model):
    if model.name not in _ROBOTS:
        native.init(0)
        _ROBOTS[model.name] = native.Robot(model)
    return _ROBOTS[model.name]


def gpu_edge(model, scene, a, b):
    ok, _ = native.check_edges(_robot(model), native.Scene(scene), a, b, 0.02)
    return ok


def test_point_robot_wall():
    m = point_robot(0.02, lim=1.0)
    r = _robot(m)
    scene = Scene(boxes=box_record([0, 0, 0], np.eye(3), [0.05, 0.6, 0.6])[None])
    s, g = np.array([-0.5, 0, 0], np.float32), np.array([0.5, 0, 0], np.float32)
    res = native.rrtc_solve(r, native.Scene(scene), s, g, native.default_params(seed=1))
    _assert_valid(m, scene, res, s, g)
    # the path goes around the wall (|y| or |z| beyond the wall's half extent + radius at x = 0)
    path = res["path"]
    cross = np.nonzero(np.diff(np.sign(path[:, 0])))[0]
    assert len(cross) >= 1
    assert res["start_tree"] >= 1 and res["goal_tree"] >= 1


def test_fr3_c1_query(fr3):
    scene = workloads.c1_scene(seed=0)
    rng = np.random.default_rng(0)
    pair = workloads.sample_pair(fr3, scene, rng, gpu_check, gpu_edge, want_free_line=False)
    assert pair is not None
    s, g = pair
    res = native.rrtc_solve(_robot(fr3), native.Scene(scene), s, g, native.default_params(seed=0))
    _assert_valid(fr3, scene, res, s, g)  # status + endpoints; validity rate is asserted over batches


def test_invalid_start_reported(fr3):
    scene = Scene(boxes=box_record([0, 0, 0.3], np.eye(3), [0.5, 0.5, 0.5])[None])  # swallows the arm
    q = workloads.uniform_configs(fr3, 2, np.random.default_rng(0))
    res = native.rrtc_solve(_robot(fr3), native.Scene(scene), q[0], q[1])
    assert res["status"] == "invalid" and len(res["path"]) == 0


@pytest.mark.parametrize("lanes", [64, 32])
def test_exhausted_and_tree_full(lanes):
    m = point_robot(0.02, lim=1.0)
    r = _robot(m)
    # goal enclosed by a box shell: unreachable
    walls = [box_record([0.5 + dx, dy, dz], np.eye(3), h) for dx, dy, dz, h in [
        (-0.15, 0, 0, [0.01, 0.2, 0.2]), (0.15, 0, 0, [0.01, 0.2, 0.2]), (0, -0.15, 0, [0.2, 0.01, 0.2]),
        (0, 0.15, 0, [0.2, 0.01, 0.2]), (0, 0, -0.15, [0.2, 0.2, 0.01]), (0, 0, 0.15, [0.2, 0.2, 0.01])]]
    scene = native.Scene(Scene(boxes=np.array(walls)))
    s, g = np.array([-0.5, 0, 0], np.float32), np.array([0.5, 0, 0], np.float32)
    for budget in (300, 301):  # an odd budget leaves the paired iteration's second half idle
        res = native.rrtc_solve(r, scene, s, g, native.default_params(max_iterations=budget, seed=2,
                                                                      lanes_per_iteration=lanes))
        assert res["status"] == "exhausted" and res["iterations"] == budget
    res = native.rrtc_solve(r, scene, s, g, native.default_params(max_iterations=100000, tree_capacity=64, seed=2,
                                                                  lanes_per_iteration=lanes))
    assert res["status"] == "tree_full"


@pytest.mark.parametrize("lanes", [64, 32])
def test_exclusive_mode_deterministic_and_batch_equals_solo(fr3, fetch, lanes):
    rng = np.random.default_rng(5)
    queries = []
    for k in range(6):
        m = fr3 if k % 2 == 0 else fetch
        scene = workloads.random_scene(rng, 20, *workloads.workspace(m))
        pair = workloads.sample_pair(m, scene, rng, gpu_check, gpu_edge, want_free_line=(k % 3 == 0))
        assert pair is not None
        queries.append((m, scene, pair[0], pair[1]))
    p = native.default_params(exclusive=1, seed=11, lanes_per_iteration=lanes)
    solo = [native.rrtc_solve(_robot(m), native.Scene(sc), s, g,
                              native.default_params(exclusive=1, seed=11 + k, lanes_per_iteration=lanes))
            for k, (m, sc, s, g) in enumerate(queries)]
    again = native.rrtc_solve(_robot(queries[0][0]), native.Scene(queries[0][1]), queries[0][2], queries[0][3],
                              native.default_params(exclusive=1, seed=11, lanes_per_iteration=lanes))
    assert np.array_equal(again["path"], solo[0]["path"])
    batch = native.rrtc_solve_batch([_robot(m) for m, _, _, _ in queries], [native.Scene(sc) for _, sc, _, _ in queries],
                                    [s for _, _, s, _ in queries], [g for _, _, _, g in queries], p)
    for k, (m, sc, s, g) in enumerate(queries):
        assert batch[k]["status"] == solo[k]["status"]
        assert np.array_equal(batch[k]["path"], solo[k]["path"])
        if solo[k]["status"] == "solved":
            _assert_valid(m, sc, solo[k], s, g)


@pytest.mark.parametrize("lanes", [64, 32])
def test_exclusive_mode_team_size_invariant(fr3, fetch, lanes):
    """One team per query: the nearest-neighbour ties go to the lowest index and a check's
    verdict on its deciding lanes is exact for any split of the work across the team's waves,
    so the planner's decisions, and the paths, are the same for 1, 2 and 4 waves per team."""
    rng = np.random.default_rng(21)
    queries = []
    for k in range(6):
        m = fr3 if k % 2 == 0 else fetch
        scene = workloads.random_scene(rng, workloads.log_uniform_count(rng, 5, 120), *workloads.workspace(m))
        pair = workloads.sample_pair(m, scene, rng, gpu_check, gpu_edge, want_free_line=(k % 3 == 0))
        assert pair is not None
        queries.append((m, scene, pair[0], pair[1]))
    robots = [_robot(m) for m, _, _, _ in queries]
    scenes = [native.Scene(sc) for _, sc, _, _ in queries]
    starts, goals = [s for _, _, s, _ in queries], [g for _, _, _, g in queries]
    ref = None
    for team in (1, 2, 4):
        res = native.rrtc_solve_batch(robots, scenes, starts, goals,
                                      native.default_params(exclusive=1, seed=5, team_waves=team,
                                                            lanes_per_iteration=lanes))
        got = [(r["status"], r["path"], r["iterations"], r["start_tree"], r["goal_tree"], r["configs_checked"])
               for r in res]
        if ref is None:
            ref = got
            assert sum(r[0] == "solved" for r in got) >= 4
            continue
        for a, b in zip(ref, got):
            assert a[0] == b[0] and np.array_equal(a[1], b[1]) and a[2:] == b[2:], (team, a[0], b[0], a[2:], b[2:])


@pytest.mark.parametrize("lanes", [64, 32])
def test_batch_shared_mixed_robots(fr3, fetch, lanes):
    rng = np.random.default_rng(9)
    robots, scenes, starts, goals, models = [], [], [], [], []
    for k in range(24):
        m = fr3 if k % 2 == 0 else fetch
        scene = workloads.random_scene(rng, workloads.log_uniform_count(rng, 5, 60), *workloads.workspace(m))
        pair = workloads.sample_pair(m, scene, rng, gpu_check, gpu_edge, want_free_line=(k % 2 == 0))
        if pair is None:
            continue
        robots.append(_robot(m)); scenes.append(native.Scene(scene)); starts.append(pair[0]); goals.append(pair[1])
        models.append((m, scene))
    res = native.rrtc_solve_batch(robots, scenes, starts, goals, native.default_params(seed=3, lanes_per_iteration=lanes))
    solved = [r for r in res if r["status"] == "solved"]
    assert len(solved) >= 0.8 * len(res)
    valid = 0
    for (m, sc), r, s, g in zip(models, res, starts, goals):
        if r["status"] == "solved":
            valid += _assert_valid(m, sc, r, s, g)
    assert valid >= 0.9 * len(solved)


@pytest.mark.parametrize("lanes", [64, 32])
@pytest.mark.parametrize("option", ["halton", "dynamic_domain", "balance_split", "balance_dynamic",
                                    "balance_alternate"])
def test_sampling_options(fr3, fetch, lanes, option):
    """Halton sample streams, dynamic-domain sampling (pRRTC's settings) and pRRTC's tree-balance
    modes: shared-mode batches solve and keep the planner's contract; exclusive mode stays
    deterministic per seed
    with batch == solo; the runs differ from the plain ones (the setting reaches the kernel)."""
    rng = np.random.default_rng(31)
    queries = []
    for k in range(12):
        m = fr3 if k % 2 == 0 else fetch
        scene = workloads.random_scene(rng, workloads.log_uniform_count(rng, 5, 60), *workloads.workspace(m))
        pair = workloads.sample_pair(m, scene, rng, gpu_check, gpu_edge, want_free_line=False)
        if pair is not None:
            queries.append((m, scene, pair[0], pair[1]))
    assert len(queries) >= 8
    robots = [_robot(m) for m, _, _, _ in queries]
    scenes = [native.Scene(sc) for _, sc, _, _ in queries]
    starts, goals = [s for _, _, s, _ in queries], [g for _, _, _, g in queries]
    opt = {"halton": {"sampler": native.SAMPLER_HALTON}, "dynamic_domain": {"dynamic_domain": 1},
           "balance_split": {"balance": native.BALANCE_SPLIT}, "balance_dynamic": {"balance": native.BALANCE_DYNAMIC},
           "balance_alternate": {"balance": native.BALANCE_ALTERNATE}}[option]
    res = native.rrtc_solve_batch(robots, scenes, starts, goals,
                                  native.default_params(seed=3, lanes_per_iteration=lanes, **opt))
    solved = [k for k, r in enumerate(res) if r["status"] == "solved"]
    assert len(solved) >= 0.8 * len(res)
    valid = sum(_assert_valid(queries[k][0], queries[k][1], res[k], starts[k], goals[k]) for k in solved)
    assert valid >= 0.9 * len(solved)
    ex = native.default_params(exclusive=1, seed=7, lanes_per_iteration=lanes, **opt)
    batch = native.rrtc_solve_batch(robots[:4], scenes[:4], starts[:4], goals[:4], ex)
    uni = native.rrtc_solve_batch(robots[:4], scenes[:4], starts[:4], goals[:4],
                                  native.default_params(exclusive=1, seed=7, lanes_per_iteration=lanes))
    differs = False
    for k in range(4):
        solo = native.rrtc_solve(robots[k], scenes[k], starts[k], goals[k],
                                 native.default_params(exclusive=1, seed=7 + k, lanes_per_iteration=lanes, **opt))
        assert solo["status"] == batch[k]["status"] and np.array_equal(solo["path"], batch[k]["path"])
        if solo["status"] == "solved":
            _assert_valid(queries[k][0], queries[k][1], solo, starts[k], goals[k])
        differs |= solo["iterations"] != uni[k]["iterations"] or not np.array_equal(solo["path"], uni[k]["path"])
    assert differs


@pytest.mark.parametrize("lanes", [64, 32])
def test_reused_slots_hold_no_earlier_nodes(fr3, fetch, lanes):
    """Admission empties every coordinate row of a reused slot: batches of the same size run in
    the same slot pool, here with the robots swapped between slots from one batch to the next
    and many teams per query, and every solved path keeps the planner's contract."""
    rng = np.random.default_rng(13)
    sets = []
    for swap in (0, 1):
        qs = []
        for k in range(16):
            m = fr3 if (k + swap) % 2 == 0 else fetch
            scene = workloads.random_scene(rng, workloads.log_uniform_count(rng, 5, 60), *workloads.workspace(m))
            pair = workloads.sample_pair(m, scene, rng, gpu_check, gpu_edge, want_free_line=False)
            if pair is not None:
                qs.append((m, scene, pair[0], pair[1]))
        sets.append(qs[:12])
    solved = 0
    for rep in range(6):
        qs = sets[rep % 2]
        res = native.rrtc_solve_batch([_robot(m) for m, _, _, _ in qs], [native.Scene(sc) for _, sc, _, _ in qs],
                                      [s for _, _, s, _ in qs], [g for _, _, _, g in qs],
                                      native.default_params(seed=40 + rep, lanes_per_iteration=lanes))
        for (m, sc, s, g), r in zip(qs, res):
            if r["status"] == "solved":
                _assert_valid(m, sc, r, s, g)
                solved += 1
    assert solved >= 0.8 * 6 * 12


def test_autotune_picks_fastest_candidate_and_is_used(fr3, fetch):
    rng = np.random.default_rng(31)
    robots, scenes, starts, goals = [], [], [], []
    for k in range(12):
        m = fr3 if k % 2 == 0 else fetch
        scene = workloads.random_scene(rng, workloads.log_uniform_count(rng, 5, 60), *workloads.workspace(m))
        pair = workloads.sample_pair(m, scene, rng, gpu_check, gpu_edge, want_free_line=(k % 2 == 0))
        if pair is None:
            continue
        robots.append(_robot(m)); scenes.append(native.Scene(scene)); starts.append(pair[0]); goals.append(pair[1])
    best, cands = native.rrtc_autotune(robots, scenes, starts, goals, native.default_params(seed=4), reps=1)
    # 3 team sizes x {resident, 1024 waves} x 5 caps x 2 lane widths, less the 1024-wave
    # candidates that are the resident grid (the same launch)
    assert 30 <= len(cands) <= 60
    assert len({(c["team_waves"], c["waves"], c["teams_per_query"], c["lanes_per_iteration"]) for c in cands}) == len(cands)
    assert {c["team_waves"] for c in cands} == {1, 2, 4}
    assert {c["teams_per_query"] for c in cands} == {0, 16, 32, 64, 128}
    assert {c["lanes_per_iteration"] for c in cands} == {32, 64}
    fastest = min(cands, key=lambda c: c["kernel_ms"])
    assert (best.team_waves, best.waves, best.teams_per_query, best.lanes_per_iteration) == (
        fastest["team_waves"], fastest["waves"], fastest["teams_per_query"], fastest["lanes_per_iteration"])
    assert all(c["kernel_ms"] > 0 for c in cands)
    # later batches with team_waves = waves = 0 run with the stored shape
    res = native.rrtc_solve_batch(robots, scenes, starts, goals, native.default_params(seed=9))
    assert sum(r["status"] == "solved" for r in res) >= 0.8 * len(res)


def test_batch_request_batch_and_each_agree_in_exclusive_mode(fr3, fetch):
    """A packed request solved in one launch and one launch per query (query k with seed + k)
    gives the same results in exclusive mode; paths are readable until the next solve."""
    rng = np.random.default_rng(41)
    robots, scenes, starts, goals = [], [], [], []
    for k in range(8):
        m = fr3 if k % 2 == 0 else fetch
        scene = workloads.random_scene(rng, workloads.log_uniform_count(rng, 5, 80), *workloads.workspace(m))
        pair = workloads.sample_pair(m, scene, rng, gpu_check, gpu_edge, want_free_line=(k % 4 == 0))
        assert pair is not None
        robots.append(_robot(m)); scenes.append(native.Scene(scene)); starts.append(pair[0]); goals.append(pair[1])
    req = native.BatchRequest(robots, scenes, starts, goals)
    p = native.default_params(exclusive=1, seed=17)
    a = req.solve(p)
    pa = [a.path(k) for k in range(req.nq)]
    b = req.solve_each(p)
    with pytest.raises(native.MPError):
        a.path(0)
    assert list(a.status) == list(b.status) and a.solved >= 6
    assert np.array_equal(a.iterations, b.iterations)
    for k in range(req.nq):
        assert np.array_equal(pa[k], b.path(k))
