"""Writes cpu_port_status.json: the CPU RRT-Connect port's outcome (the oracle, OMPL RRTConnect
semantics, range 0.3 rad, resolution 0.02 rad) on every committed MotionBenchMaker problem, for
seeds 0..2 at the GPU planner's budget of 10,000 samples, with each solved path's re-check at
0.005 rad.  The GPU MBM test compares its solve counts per problem set with these.

usage: python tests/golden/mbm/make_cpu_port_status.py
"""
import concurrent.futures as cf
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from isaac_sim_ros2_humble_amd import mbm  # noqa: E402

SEEDS = (0, 1, 2)
BUDGET = 10000


def main():
    models = mbm.problem_models()
    orc = [oracle.OracleRobot(m) for m in models]
    probs = mbm.load_problems(HERE, models)

    def one(item):
        i, seed = item
        name, k, scene, s, g = probs[i]
        osc = oracle.OracleScene(scene)
        path, st = oracle.rrtconnect(orc[k], osc, s, g, range_=0.3, resolution=0.02, seed=seed,
                                     max_iterations=BUDGET, time_limit=600.0)
        valid = bool(st.solved == 1 and oracle.validate_path(orc[k], osc, path, 0.005)[0])
        return i, seed, int(st.solved), int(st.iterations), valid

    items = [(i, s) for i in range(len(probs)) for s in SEEDS]
    with cf.ThreadPoolExecutor(os.cpu_count() or 4) as ex:
        res = list(ex.map(one, items))
    out = {"budget_samples": BUDGET, "seeds": list(SEEDS), "problems": []}
    for i, (name, k, scene, s, g) in enumerate(probs):
        rows = [r for r in res if r[0] == i]
        out["problems"].append({"set": name, "index": i, "solved": [r[2] for r in rows],
                                "iterations": [r[3] for r in rows], "valid": [r[4] for r in rows]})
    with open(os.path.join(HERE, "cpu_port_status.json"), "w") as f:
        json.dump(out, f, indent=1)
    by = {}
    for p in out["problems"]:
        d = by.setdefault(p["set"], [0, 0])
        d[0] += sum(x == 1 for x in p["solved"])
        d[1] += len(p["solved"])
    for k, v in sorted(by.items()):
        print("%-24s %d/%d" % (k, v[0], v[1]))


if __name__ == "__main__":
    np.seterr(all="ignore")
    main()
