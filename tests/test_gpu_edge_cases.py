"""Edge cases of the planning entry points through the C ABI: start == goal, an empty scene, a
scene updated between solves, single-query batches and rejected parameters."""
import numpy as np
import pytest

import oracle
from isaac_sim_ros2_humble_amd import native, workloads
from isaac_sim_ros2_humble_amd.scene import Scene, box_record
from helpers import assert_valid as _assert_valid, planner_states
from test_gpu_planner import _robot

pytestmark = pytest.mark.gpu


def _free_config(model, scene, rng):
    orc, osc = oracle.OracleRobot(model), oracle.OracleScene(scene)
    for _ in range(64):
        q = workloads.uniform_configs(model, 256, rng)
        v, _, _, _ = oracle.check(orc, osc, q)
        if v.any():
            return q[np.nonzero(v)[0][0]]
    raise AssertionError("no free configuration")


def test_start_equals_goal(fr3):
    rng = np.random.default_rng(1)
    scene = Scene()
    q = _free_config(fr3, scene, rng)
    res = native.rrtc_solve(_robot(fr3), native.Scene(scene), q, q, native.default_params(seed=3))
    _assert_valid(fr3, scene, res, q, q)
    assert len(res["path"]) <= 4


def test_empty_scene(fetch):
    rng = np.random.default_rng(2)
    scene = Scene()
    a, b = _free_config(fetch, scene, rng), _free_config(fetch, scene, rng)
    res = native.rrtc_solve(_robot(fetch), native.Scene(scene), a, b, native.default_params(seed=4))
    if res["status"] == "solved":
        _assert_valid(fetch, scene, res, a, b)
    else:  # self-collision can block an empty-scene query; it must then be reported, not invented
        assert res["status"] in ("exhausted", "invalid") and len(res["path"]) == 0


def test_scene_update_between_solves(fr3):
    rng = np.random.default_rng(7)
    # a scene's capacity is fixed at creation (setEnvObstacleCache): one box, far away at first
    empty = Scene(boxes=box_record([5.0, 5.0, 5.0], np.eye(3), [0.04, 0.04, 0.04])[None])
    orc = oracle.OracleRobot(fr3)
    # a pair whose straight line is free in the empty scene
    for _ in range(50):
        a, b = _free_config(fr3, empty, rng), _free_config(fr3, empty, rng)
        line = planner_states(np.stack([a, b]), 0.02)
        if oracle.check(orc, oracle.OracleScene(empty), line)[0].all():
            break
    nscene = native.Scene(empty)
    r1 = native.rrtc_solve(_robot(fr3), nscene, a, b, native.default_params(seed=5))
    _assert_valid(fr3, empty, r1, a, b)
    # put a box on the straight line's midpoint (hand position there), clear of both ends
    mid = (a + b) / 2
    c = orc.fk32(mid[None])[0].reshape(-1, 3)[-1]
    blocked = Scene(boxes=box_record(c, np.eye(3), [0.04, 0.04, 0.04])[None])
    v, _, _, _ = oracle.check(orc, oracle.OracleScene(blocked), np.stack([a, b]))
    if not v.all():
        pytest.skip("the box reaches an endpoint for this draw")
    nscene.update(blocked)
    assert np.array_equal(native.check_states(_robot(fr3), nscene, np.stack([a, mid, b])),
                          oracle.check(orc, oracle.OracleScene(blocked), np.stack([a, mid, b]))[0])
    r2 = native.rrtc_solve(_robot(fr3), nscene, a, b, native.default_params(seed=6))
    _assert_valid(fr3, blocked, r2, a, b)


def test_single_query_batch_and_empty_batch(fr3):
    rng = np.random.default_rng(9)
    scene = workloads.c1_scene(seed=0)
    a, b = _free_config(fr3, scene, rng), _free_config(fr3, scene, rng)
    res = native.rrtc_solve_batch([_robot(fr3)], [native.Scene(scene)], [a], [b], native.default_params(seed=1))
    assert res[0]["status"] in ("solved", "exhausted")
    if res[0]["status"] == "solved":
        _assert_valid(fr3, scene, res[0], a, b)
    with pytest.raises(native.MPError):
        native.rrtc_solve_batch([], [], [], [], native.default_params())


@pytest.mark.parametrize("field,value", [("team_waves", 3), ("team_waves", 8), ("max_step", 65), ("max_step", 0),
                                         ("teams_per_query", -1), ("resolution", 0.0), ("tree_capacity", 1),
                                         ("lanes_per_iteration", 16), ("lanes_per_iteration", 48),
                                         ("sampler", 2), ("sampler", -1), ("dynamic_domain", 2)])
def test_bad_params_rejected(fr3, field, value):
    q = np.zeros(7, np.float32)
    p = native.default_params()
    setattr(p, field, value)
    with pytest.raises(native.MPError) as e:
        native.rrtc_solve(_robot(fr3), native.Scene(Scene()), q, q, p)
    assert "(-1)" in str(e.value)  # MP_E_INVALID, reported as a status code


@pytest.mark.parametrize("field,value", [("dd_radius", 0.0), ("dd_min_radius", -1.0), ("dd_alpha", 1.0),
                                         ("dd_alpha", -0.5)])
def test_bad_domain_settings_rejected(fr3, field, value):
    q = np.zeros(7, np.float32)
    p = native.default_params(dynamic_domain=1)
    setattr(p, field, value)
    with pytest.raises(native.MPError) as e:
        native.rrtc_solve(_robot(fr3), native.Scene(Scene()), q, q, p)
    assert "(-1)" in str(e.value)
