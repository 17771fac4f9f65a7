#!/usr/bin/env python3
"""Benchmark of the planning hot path on one MI355X.

Headline: solved plan queries/s of the batched planner on the 256-query mixed FR3/Fetch workload
(one persistent launch per step), next to the plain one-query-per-launch planner on the same
queries, the served planner's p50/p99 latency under Poisson arrivals, the 2^20-configuration
FK + collision microbenchmark, and the CPU RRT-Connect baseline (oracle, on the host's cores).

Timed region: K steps of the batched solve, each bracketed by a device synchronisation; the
inputs (scenes, robot modules, query descriptors) are resident in HBM before timing starts.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--serve-seconds S] [--rate L]
                       [--team-waves W --waves N --teams-per-query C --lanes 32|64]
                       [--sampler uniform|halton] [--dynamic-domain] [--skip PART ...]
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from isaac_sim_ros2_humble_amd import native, workloads  # noqa: E402
from isaac_sim_ros2_humble_amd.model import load_robot  # noqa: E402

METRIC = "solved plan queries/sec + p50/p99 latency, batched/served vs plain, 1 MI355X"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel, path=os.path.join(REPO, "profiles", "traffic.txt")):
    """HBM bytes per launch (fetch + write) of `kernel` from the last rocprofv3 --pmc runs
    (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel factors), or None."""
    try:
        with open(path) as f:
            for line in f:
                w = line.split()
                if len(w) >= 3 and w[0] == kernel:
                    return float(w[1]) + float(w[2])
    except OSError:
        pass
    return None


def build_c4(robots, models, n, seed):
    def check(m, scene, q):
        return native.check_states(robots[m.name], native.Scene(scene), q)

    def edge(m, scene, a, b):
        ok, _ = native.check_edges(robots[m.name], native.Scene(scene), a, b, 0.02)
        return ok
    return workloads.c4_batch(models, n, check, edge, seed=seed)


def run_batch(rob_list, scenes, starts, goals, params):
    native.sync()
    t0 = time.perf_counter()
    res = native.rrtc_solve_batch(rob_list, scenes, starts, goals, params)
    native.sync()
    return res, (time.perf_counter() - t0) * 1e3, native.kernel_ms("plan")


def full_flops_per_config(m, counts):
    """Cost model per configuration, full evaluation: FK + 18 S + 9 P + S (11 ns + 34 nb + 35 nc + 28 ncap)."""
    fk = 0
    for i in range(1, m.L):
        t = int(m.joint_types[i])
        fk += 130 if t == 1 else 87 if t == 2 else 63
    S = m.S
    P = 0
    for i in range(S):
        for j in range(i + 1, S):
            li, lj = m.sphere_link[i], m.sphere_link[j]
            if li != lj and m.enabled[li, lj]:
                P += 1
    ns, nb, nc, ncap = counts
    return fk + 18 * S + 9 * P + S * (11 * ns + 34 * nb + 35 * nc + 28 * ncap)


def bench_c2(robot, model, steps, warmup):
    scene = native.Scene(workloads.c2_scene(seed=1))
    q = workloads.c2_configs(model, 1 << 20, seed=1)
    n = len(q)
    dq = native.DeviceBuffer.from_array(q)
    dv = native.DeviceBuffer(n)
    dc = native.DeviceBuffer(n * model.S * 12)
    dcost = native.DeviceBuffer(4 * n)
    L = native.lib()
    out = {}
    for mode in ("verdict", "full"):
        times, flops = [], []
        for it in range(warmup + steps):
            native.work_flops(reset=True)
            if mode == "verdict":
                native.check(L.mp_check_states(robot.h, scene.h, dq.ptr, n, dv.ptr, dc.ptr, None), "check")
                fam = "check"
            else:
                native.check(L.mp_state_costs(robot.h, scene.h, dq.ptr, n, dcost.ptr, dv.ptr, None), "cost")
                fam = "cost"
            native.sync()
            if it >= warmup:
                times.append(native.kernel_ms(fam))
                flops.append(native.work_flops(reset=True))
        ms = float(np.mean(times))
        fl = float(np.mean(flops))
        v = dv.to_array(np.uint8, (n,))
        out[mode] = {"ms": ms, "configs_per_s": n / (ms / 1e3), "valid_fraction": float(v.mean()),
                     "executed_flops_per_config": fl / n, "achieved_tflops": fl / (ms / 1e3) / 1e12}
    # bytes: q 28 B + centres 12 S B + verdict 1 B per configuration
    bpc = model.dof * 4 + 12 * model.S + 1
    vms = out["verdict"]["ms"]
    out["verdict"]["hbm_gbs"] = n * bpc / (vms / 1e3) / 1e9
    ff = full_flops_per_config(model, scene.scene.counts)
    out["full"]["model_flops_per_config"] = ff
    out["full"]["model_tflops"] = n * ff / (out["full"]["ms"] / 1e3) / 1e12
    return out


def cpu_baseline(models, queries, budget_s, threads):
    import oracle  # the CPU baseline leg: the oracle's RRT-Connect (OMPL semantics) on the host
    orc = {m.name: oracle.OracleRobot(m) for m in models}
    per_query_limit = 2.0
    sample = queries[:max(1, min(len(queries), int(budget_s * threads / per_query_limit)))]

    def one(q):
        k, scene, s, g, _ = q
        m = models[k]
        path, st = oracle.rrtconnect(orc[m.name], oracle.OracleScene(scene), s, g, range_=0.3, resolution=0.02,
                                     seed=0, time_limit=per_query_limit)
        return st.solved == 1
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        solved = sum(ex.map(one, sample))
    dt = time.perf_counter() - t0
    return {"value": solved / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": "first %d of the %d batched queries, RRT-Connect range 0.3 rad, resolution 0.02 rad, "
                      "%.0f s limit per query, %d solved in %.1f s" % (len(sample), len(queries), per_query_limit,
                                                                      solved, dt)}


def bench_c1(robot, model, seeds, threads):
    """C1: FR3 with 20 yaw-only boxes, one query per seed.  CPU: the oracle's RRT-Connect (OMPL
    semantics, range 0.3 rad, resolution 0.02 rad), one thread per query, `threads` queries at a
    time; GPU: the plain one-query-per-launch planner on the same queries."""
    import oracle
    scene = workloads.c1_scene(seed=0)
    nscene = native.Scene(scene)

    def check(m, sc, q):
        return native.check_states(robot, nscene, q)

    def edge(m, sc, a, b):
        ok, _ = native.check_edges(robot, nscene, a, b, 0.02)
        return ok
    pairs = [workloads.sample_pair(model, scene, np.random.default_rng(k), check, edge, want_free_line=False)
             for k in range(seeds)]
    orc, osc = oracle.OracleRobot(model), oracle.OracleScene(scene)

    def cpu_one(k):
        t0 = time.perf_counter()
        _, st = oracle.rrtconnect(orc, osc, pairs[k][0], pairs[k][1], range_=0.3, resolution=0.02, seed=k,
                                  time_limit=10.0)
        return time.perf_counter() - t0, st.solved == 1
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        cpu = list(ex.map(cpu_one, range(seeds)))
    wall = time.perf_counter() - t0
    gpu = []
    for k in range(seeds):
        r = native.rrtc_solve(robot, nscene, pairs[k][0], pairs[k][1], native.default_params(seed=k))
        gpu.append((r["latency_ms"] / 1e3, r["status"] == "solved"))
    ct = np.array([c[0] for c in cpu]) * 1e3
    gt = np.array([g[0] for g in gpu]) * 1e3
    return {"seeds": seeds, "cpu_seed0_ms": float(ct[0]), "cpu_median_ms": float(np.median(ct)),
            "cpu_solved": int(sum(c[1] for c in cpu)), "cpu_threads": threads,
            "cpu_qps": float(sum(c[1] for c in cpu) / wall), "gpu_seed0_ms": float(gt[0]),
            "gpu_median_ms": float(np.median(gt)), "gpu_solved": int(sum(g[1] for g in gpu)),
            "median_speedup": float(np.median(ct) / np.median(gt))}


def cpu_checker_c2(model, threads, n=1 << 20):
    """The CPU FP32 FK + collision checker (oracle) on the C2 configurations."""
    import oracle
    scene = workloads.c2_scene(seed=1)
    q = workloads.c2_configs(model, 1 << 20, seed=1)[:n]
    orc, osc = oracle.OracleRobot(model), oracle.OracleScene(scene)
    t0 = time.perf_counter()
    oracle.check(orc, osc, q, nthreads=threads)
    dt = time.perf_counter() - t0
    # untimed: the canonical-order evaluated work (self pairs, then spheres x obstacles in index
    # order, stopping at the first hit) per configuration, on a sample of the same configurations
    ns = min(n, 65536)
    _, _, _, fl = oracle.check(orc, osc, q[:ns], want_flops=True, nthreads=threads)
    return {"value": n / dt, "unit": "configs/s", "cores": threads, "kind": "port",
            "sample": "%d C2 configurations (FR3, 64 obstacles), verdict mode" % n,
            "canonical_flops_per_config": float(fl.mean()), "canonical_sample": ns}


def bench_c3(robot, model, seeds):
    """Fetch 8-DOF in a 3-shelf unit: start outside, goal between shelves, one query per seed."""
    scene, shelf = workloads.c3_scene(seed=0)
    nscene = native.Scene(scene)

    def check(m, sc, q):
        return native.check_states(robot, nscene, q)

    def fk(m, q):
        return native.fk_spheres(robot, q)
    starts, goals = workloads.c3_queries(model, scene, shelf, seeds, check, fk, seed=0)
    ok, lens, times = 0, [], []
    for k in range(seeds):
        r = native.rrtc_solve(robot, nscene, starts[k], goals[k], native.default_params(seed=k))
        times.append(r["latency_ms"])
        if r["status"] == "solved":
            ok += 1
            lens.append(float(np.linalg.norm(np.diff(r["path"], axis=0), axis=1).sum()))
    return {"seeds": seeds, "success_rate": ok / seeds, "mean_path_length_rad": float(np.mean(lens)) if lens else None,
            "mean_ms": float(np.mean(times)), "p50_ms": float(np.percentile(times, 50)),
            "p99_ms": float(np.percentile(times, 99)), "obstacles": scene.num_obstacles}


def bench_mbm(models, sampler=0, dynamic_domain=0):
    """MotionBenchMaker problems (committed fixtures: 5 per set, 14 sets) in one batched launch."""
    from isaac_sim_ros2_humble_amd import mbm
    mm = mbm.problem_models()
    rob = [native.Robot(m) for m in mm]
    probs = mbm.load_problems(os.path.join(REPO, "tests", "golden", "mbm"), mm)
    scenes = [native.Scene(p[2]) for p in probs]
    res, ms, kms = run_batch([rob[p[1]] for p in probs], scenes, [p[3] for p in probs], [p[4] for p in probs],
                             native.default_params(seed=0, sampler=sampler, dynamic_domain=dynamic_domain))
    by = {}
    for p, r in zip(probs, res):
        d = by.setdefault(p[0], [0, 0])
        d[0] += r["status"] == "solved"
        d[1] += 1
    return {"problems": len(probs), "solved": sum(v[0] for v in by.values()), "batch_ms": ms,
            "statuses": {s: sum(r["status"] == s for r in res) for s in set(r["status"] for r in res)},
            "per_set_solved": {k: "%d/%d" % tuple(v) for k, v in by.items()}}


def serve(robots, rob_names, queries, rate, seconds, exclusive=False):
    srv = native.Server([robots[n] for n in rob_names], tree_memory_bytes=16 << 30, tree_capacity=65536,
                        max_active=1024, use_graph=True,
                        params=native.default_params(exclusive=1 if exclusive else 0))
    arrivals = workloads.poisson_arrivals(rate, seconds, seed=7)
    scenes = [native.Scene(q[1]) for q in queries]
    rng = np.random.default_rng(3)
    tickets = []
    t0 = time.perf_counter()
    for k, ta in enumerate(arrivals):
        while time.perf_counter() - t0 < ta:
            time.sleep(0.0002)
        # scene reuse: with probability 0.8 a resident scene of an earlier query, else the next one
        j = int(rng.integers(len(queries))) if rng.random() < 0.8 else k % len(queries)
        mi, _, s, g, _ = queries[j]
        tickets.append((srv.submit(mi, scenes[j], s, g, seed=k), len(s)))
    srv.wait_idle(120.0)
    st = srv.stats()
    statuses = {}
    for t, dof in tickets:
        r = srv.poll(t, dof)
        statuses[r["status"]] = statuses.get(r["status"], 0) + 1
    srv.close()
    return {"rate_qps": rate, "seconds": seconds, "submitted": st["submitted"], "completed": st["completed"],
            "solved": st["solved"], "throughput_qps": st["throughput_qps"], "p50_ms": st["p50_ms"],
            "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"], "steps": st["steps"], "statuses": statuses}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--rate", type=float, nargs="*", default=[100.0, 1000.0, 5000.0])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--c3-seeds", type=int, default=1000)
    ap.add_argument("--c1-seeds", type=int, default=100)
    ap.add_argument("--tune-queries", type=int, default=256)
    ap.add_argument("--team-waves", type=int, default=0, help="planning launch shape (0: tuned / default)")
    ap.add_argument("--waves", type=int, default=0, help="planning grid in waves (0: tuned / resident)")
    ap.add_argument("--teams-per-query", type=int, default=0, help="teams per query (0: tuned / unbounded)")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per iteration, 32 or 64 (0: tuned / default)")
    ap.add_argument("--sampler", choices=["uniform", "halton"], default="uniform",
                    help="sample streams of the batched planner: Philox uniform or rotated Halton")
    ap.add_argument("--dynamic-domain", action="store_true",
                    help="dynamic-domain sampling in the batched planner (pRRTC's settings)")
    ap.add_argument("--skip", nargs="*", default=[], help="parts to skip: tune plain serve c2 c3 mbm cpu (cpu also skips c1)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: queries shard across GPUs with no exchange, so the "
              "single-GPU number is the per-GPU share of an N-GPU job (one process per GPU, each "
              "with its own queries, scenes and trees)", file=sys.stderr)
        sys.exit(2)

    native.init(0)
    models = [load_robot("fr3"), load_robot("fetch")]
    robots = {m.name: native.Robot(m) for m in models}
    log("robots ready")
    queries = build_c4(robots, models, a.queries, seed=0)
    log("c4 workload: %d queries" % len(queries))
    rob_list = [robots[models[k].name] for k, _, _, _, _ in queries]
    scenes = [native.Scene(sc) for _, sc, _, _, _ in queries]
    starts = [s for _, _, s, _, _ in queries]
    goals = [g for _, _, _, g, _ in queries]

    # autotune the launch shapes (team size x grid) on a separate calibration workload; the
    # choice is stored per robot set and class and used by the batched, plain and served runs
    tune = {}
    if "tune" not in a.skip:
        cal = build_c4(robots, models, a.tune_queries, seed=1)
        c_rob = [robots[models[k].name] for k, _, _, _, _ in cal]
        c_sc = [native.Scene(sc) for _, sc, _, _, _ in cal]
        c_s = [q[2] for q in cal]
        c_g = [q[3] for q in cal]
        for name, per_query, n in (("batch", False, len(cal)), ("plain", True, min(32, len(cal)))):
            best, cands = native.rrtc_autotune(c_rob[:n], c_sc[:n], c_s[:n], c_g[:n], native.default_params(seed=77),
                                               per_query=per_query, reps=2 if per_query else 4)
            tune[name] = {"team_waves": best.team_waves, "waves": best.waves, "teams_per_query": best.teams_per_query,
                          "lanes_per_iteration": best.lanes_per_iteration, "calibration_queries": n,
                          "candidates": cands}
            log("autotune %s: team_waves %d waves %d teams_per_query %d lanes_per_iteration %d" % (
                name, best.team_waves, best.waves, best.teams_per_query, best.lanes_per_iteration))
    params = native.default_params(seed=0, team_waves=a.team_waves, waves=a.waves, teams_per_query=a.teams_per_query,
                                   lanes_per_iteration=a.lanes,
                                   sampler=native.SAMPLER_HALTON if a.sampler == "halton" else native.SAMPLER_UNIFORM,
                                   dynamic_domain=1 if a.dynamic_domain else 0)

    # the 256 queries are packed once (host arrays, handles of resident scenes and robots); each
    # step submits them, runs the persistent launch and has every path on the host
    req = native.BatchRequest(rob_list, scenes, starts, goals, max_path=params.max_path)

    def step(p):
        native.sync()
        t0 = time.perf_counter()
        r = req.solve(p)
        native.sync()
        return r, (time.perf_counter() - t0) * 1e3, native.kernel_ms("plan")

    # batched: warmup, then exactly K timed steps
    for _ in range(a.warmup):
        step(params)
    step_ms, kern_ms, solved, flops, checked, host_ms = [], [], [], [], [], []
    for k in range(a.steps):
        params.seed = 1000 * (k + 1)
        res, ms, kms = step(params)
        step_ms.append(ms)
        kern_ms.append(kms)
        solved.append(res.solved)
        flops.append(float(res.work_flops.sum()))
        checked.append(float(res.configs_checked.sum()))
        host_ms.append(float(res.latency_ms[0]))  # the library's own submit-to-results time
        log("batch step %d: %.1f ms (kernel %.1f ms), solved %d/%d" % (k, ms, kms, solved[-1], req.nq))
    ms_per_step = float(np.mean(step_ms))
    batched_qps = float(np.sum(solved)) / (np.sum(step_ms) / 1e3)
    plan_kernel_ms = float(np.mean(kern_ms))
    achieved = float(np.mean(flops)) / (plan_kernel_ms / 1e3) / 1e12
    iters = res.iterations
    dev_ms = res.device_ms
    # one more (untimed) step with the planner's phase profile on: where the waves spend their time
    native.plan_profile(enable=True)
    params.seed = 999
    shape = ({"team_waves": a.team_waves, "waves": a.waves, "teams_per_query": a.teams_per_query,
              "lanes_per_iteration": a.lanes}
             if a.team_waves or a.waves or a.teams_per_query or a.lanes else tune.get("batch"))
    step(params)
    prof = native.plan_profile(enable=False)
    wave = max(prof["wave"], 1)
    phase = {k: prof[k] / wave for k in native.PROFILE_PHASES[:8]}
    phase.update({k: prof[k] for k in native.PROFILE_PHASES[9:]})

    out = {"metric": METRIC, "value": batched_qps, "unit": "queries/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "batched planning: %d independent RRT-Connect queries, FR3 7-DOF / Fetch 8-DOF "
                                  "mixed, 5-200 obstacles (box/cylinder/sphere), half straight-line-free, tree "
                                  "capacity 65536, resolution 0.02, one persistent launch per step" % len(queries),
                      "queries_per_step": len(queries), "tree_capacity": 65536, "resolution": 0.02,
                      "sampler": a.sampler, "dynamic_domain": bool(a.dynamic_domain)},
           "batched": {"solved_per_step": solved, "success_rate": float(np.mean(solved)) / len(queries),
                       "plan_kernel_ms": plan_kernel_ms, "configs_checked_per_step": float(np.mean(checked)),
                       "host_ms": float(np.mean(host_ms)), "iterations_p50_p99_max": [
                           float(np.percentile(iters, 50)), float(np.percentile(iters, 99)), int(iters.max())],
                       "query_device_ms_p50_p99_max": [float(np.percentile(dev_ms, 50)),
                                                       float(np.percentile(dev_ms, 99)), float(dev_ms.max())],
                       "phase_fraction_of_wave_time": phase},
           "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": measured_traffic("mp_plan"),
                        "kernel": "mp_plan (batched)", "flops_per_launch": float(np.mean(flops)),
                        "launch_shape": {k: shape[k] for k in ("team_waves", "waves", "teams_per_query",
                                                               "lanes_per_iteration")} if shape else None},
           "autotune": tune}

    if "plain" not in a.skip:
        # the same packed queries, one launch each (query k with seed k), with the plain shape
        pp = native.default_params(seed=0)
        native.sync()
        t0 = time.perf_counter()
        pr = req.solve_each(pp)
        native.sync()
        dt = time.perf_counter() - t0
        nsol = pr.solved
        lat = pr.latency_ms
        plain_qps = nsol / dt
        out["plain"] = {"qps": plain_qps, "solved": nsol, "p50_ms": float(np.percentile(lat, 50)),
                        "p99_ms": float(np.percentile(lat, 99)), "seconds": dt}
        out["batched_vs_plain"] = batched_qps / plain_qps if plain_qps > 0 else None
        log("plain: %.1f q/s" % plain_qps)

    if "serve" not in a.skip:
        out["served"] = [serve(robots, [m.name for m in models], queries, rate, a.serve_seconds) for rate in a.rate]
        log("served:", out["served"])

    if "c2" not in a.skip:
        out["c2"] = bench_c2(robots["fr3"], models[0], max(a.steps, 3), 1)
        c2v = out["c2"]["verdict"]
        out["roofline_c2"] = {"bound": "valu", "kernel": "mp_check_0 (2^20 FR3 configs, 64 obstacles)",
                              "achieved": c2v["achieved_tflops"], "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                              "frac": c2v["achieved_tflops"] / PEAK_FP32_TFLOPS, "hbm_gbs": c2v["hbm_gbs"],
                              "hbm_frac": c2v["hbm_gbs"] / PEAK_HBM_GBS}
        log("c2:", out["c2"])

    if "c3" not in a.skip and a.c3_seeds > 0:
        out["c3"] = bench_c3(robots["fetch"], models[1], a.c3_seeds)
        log("c3:", out["c3"])

    if "mbm" not in a.skip:
        out["mbm"] = bench_mbm(models, params.sampler, params.dynamic_domain)
        log("mbm:", out["mbm"])

    if "cpu" not in a.skip:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(models, queries, a.cpu_seconds, threads)
        log("cpu:", out["cpu_baseline"])
        out["cpu_checker_c2"] = cpu_checker_c2(models[0], threads)
        log("cpu c2:", out["cpu_checker_c2"])
        if "c2" in out:  # C2's early-out throughput at the canonical-order work count
            cf_ = out["cpu_checker_c2"]["canonical_flops_per_config"]
            v = out["c2"]["verdict"]
            v["canonical_flops_per_config"] = cf_
            v["canonical_tflops"] = cf_ * (1 << 20) / (v["ms"] / 1e3) / 1e12
            if "roofline_c2" in out:
                out["roofline_c2"]["achieved_canonical"] = v["canonical_tflops"]
                out["roofline_c2"]["frac_canonical"] = v["canonical_tflops"] / PEAK_FP32_TFLOPS
        if a.c1_seeds > 0:
            out["c1"] = bench_c1(robots["fr3"], models[0], a.c1_seeds, threads)
            log("c1:", out["c1"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
