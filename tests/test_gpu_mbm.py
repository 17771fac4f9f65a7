"""MotionBenchMaker problems on the GPU (the 70 committed fixtures: 14 sets x 5, Panda problems on
the FER model, Fetch problems on Fetch), planned as one batch per seed for seeds 0..2 at CPRRTC's
budget of 10,000 samples:

  * a problem whose start or goal collides under the sphere model (the oracle's verdict) is
    reported invalid, the others solve or exhaust;
  * every solved path keeps the planner's contract (oracle check of the states it checked) and at
    least VALIDITY_MIN of them pass the 0.005 rad re-check;
  * per problem set the GPU solves at least as many (problem, seed) runs as the CPU RRT-Connect
    port did at the same budget (tests/golden/mbm/cpu_port_status.json, made by
    make_cpu_port_status.py next to it), less one run of slack for the seeds' luck.
"""
import json
import os

import numpy as np
import pytest

import oracle
from helpers import assert_valid_many
from isaac_sim_ros2_humble_amd import mbm, native

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "mbm")
VALIDITY_MIN = 0.9


def test_mbm_problems_against_cpu_port():
    native.init(0)
    models = mbm.problem_models()
    robots = [native.Robot(m) for m in models]
    probs = mbm.load_problems(GOLD, models)
    assert len(probs) == 70
    with open(os.path.join(GOLD, "cpu_port_status.json")) as f:
        cpu = json.load(f)
    orc = [oracle.OracleRobot(m) for m in models]
    invalid = [not oracle.check(orc[k], oracle.OracleScene(sc), np.stack([s, g]))[0].all()
               for _, k, sc, s, g in probs]
    scenes = [native.Scene(p[2]) for p in probs]
    gpu_solved, items = {}, []
    for seed in cpu["seeds"]:
        res = native.rrtc_solve_batch([robots[p[1]] for p in probs], scenes, [p[3] for p in probs],
                                      [p[4] for p in probs], native.default_params(seed=1000 * seed))
        for i, (p, r) in enumerate(zip(probs, res)):
            if invalid[i]:
                assert r["status"] == "invalid", (p[0], i)
                continue
            assert r["status"] in ("solved", "exhausted"), (p[0], i, r["status"])
            if r["status"] == "solved":
                gpu_solved[p[0]] = gpu_solved.get(p[0], 0) + 1
                items.append((models[p[1]], p[2], r, p[3], p[4]))
    cpu_solved = {}
    for p in cpu["problems"]:
        cpu_solved[p["set"]] = cpu_solved.get(p["set"], 0) + sum(x == 1 for x in p["solved"])
    for name in sorted(cpu_solved):
        print("%-24s gpu %2d  cpu port %2d" % (name, gpu_solved.get(name, 0), cpu_solved[name]))
        assert gpu_solved.get(name, 0) >= cpu_solved[name] - 1, name
    valid = assert_valid_many(items, min(16, os.cpu_count() or 1))
    print("solved %d, valid at 0.005 rad %d" % (len(items), valid))
    assert valid >= VALIDITY_MIN * len(items)
    assert sum(gpu_solved.values()) >= sum(cpu_solved.values())
