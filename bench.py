#!/usr/bin/env python3
"""Benchmark of the planning hot path on one MI355X.

Headline: solved plan queries/s of the batched planner on the 256-query mixed FR3/Fetch workload
(one persistent launch per step), next to the plain one-query-per-launch planner on the same
queries, the served planner under Poisson arrivals (p50/p99 latency, scenes uploaded while
serving) and at saturation, the 2^20-configuration FK + collision microbenchmark, the Fetch shelf
workload, the MotionBenchMaker fixtures, and the CPU leg: the CPU RRT-Connect port and checker
timed on the host's cores, and the independent CPU re-check of the GPU paths at 0.005 rad.

Timed region: K steps of the batched solve, each bracketed by a device synchronisation; the
inputs (scenes, robot modules, query descriptors) are resident in HBM before timing starts.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--serve-seconds S] [--rate L ...]
                       [--saturation-queries N] [--team-waves W --waves N --teams-per-query C
                       --lanes 32|64] [--sampler uniform|halton] [--dynamic-domain]
                       [--balance smaller|split|dynamic|alternate] [--skip PART ...]
"""
import argparse
import concurrent.futures as cf
import json
import os
import subprocess
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
VALIDATION_RES = 0.005    # rad: the independent CPU re-check of returned paths
CONFIGURED_SERVE_SECONDS = 60.0


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


def host_info():
    """The host CPUs this process may use: affinity mask, cgroup quota, the per-GPU share the
    box grants (OMP_NUM_THREADS), and lscpu's description of the processor."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    usable = aff if quota is None else max(1, min(aff, int(quota)))
    share = os.environ.get("OMP_NUM_THREADS")
    threads = max(1, min(usable, int(share))) if share and share.isdigit() else usable
    lscpu = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)",
                             "NUMA node(s)", "CPU max MHz"):
                lscpu[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"os_cpu_count": os.cpu_count(), "affinity_cpus": aff, "cgroup_cpu_quota": quota, "usable_cpus": usable,
            "omp_num_threads": share, "threads_used": threads, "lscpu": lscpu}


def build_c4(robots, models, n, seed):
    def check(m, scene, q):
        return native.check_states(robots[m.name], native.Scene(scene), q)

    def edge(m, scene, a, b):
        ok, _ = native.check_edges(robots[m.name], native.Scene(scene), a, b, 0.02)
        return ok
    return workloads.c4_batch(models, n, check, edge, seed=seed)


def full_flops_per_config(m, pairs, counts):
    """Full-evaluation cost model per configuration: FK (revolute 130, prismatic 87, fixed 63)
    + 18 S (sphere transforms) + 9 P (self pairs) + S (11 ns + 34 nb + 35 nc + 28 ncap)."""
    fk = 0
    for i in range(1, m.L):
        t = int(m.joint_types[i])
        fk += 130 if t == 1 else 87 if t == 2 else 63
    ns, nb, nc, ncap = counts
    return fk + 18 * m.S + 9 * pairs + m.S * (11 * ns + 34 * nb + 35 * nc + 28 * ncap)


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
    ff = full_flops_per_config(model, robot.P, scene.scene.counts)
    out["full"]["model_flops_per_config"] = ff
    out["full"]["model_tflops"] = n * ff / (out["full"]["ms"] / 1e3) / 1e12
    return out


def bench_c3(robot, model, seeds):
    """Fetch 8-DOF in a 3-shelf unit: start outside, goal between shelves, one query per seed."""
    scene, shelf = workloads.c3_scene(seed=0)
    nscene = native.Scene(scene)

    def check(m, sc, q):
        return native.check_states(robot, nscene, q)

    def fk(m, q):
        return native.fk_spheres(robot, q)
    starts, goals = workloads.c3_queries(model, scene, shelf, seeds, check, fk, seed=0)
    ok, lens, times, paths = 0, [], [], []
    for k in range(seeds):
        r = native.rrtc_solve(robot, nscene, starts[k], goals[k], native.default_params(seed=k))
        times.append(r["latency_ms"])
        if r["status"] == "solved":
            ok += 1
            lens.append(float(np.linalg.norm(np.diff(r["path"], axis=0), axis=1).sum()))
            paths.append((model, scene, r["path"], starts[k], goals[k]))
    return ({"seeds": seeds, "success_rate": ok / seeds, "mean_path_length_rad": float(np.mean(lens)) if lens else None,
             "mean_ms": float(np.mean(times)), "p50_ms": float(np.percentile(times, 50)),
             "p99_ms": float(np.percentile(times, 99)), "obstacles": scene.num_obstacles}, paths)


def bench_mbm(models, sampler=0, dynamic_domain=0, balance=0):
    """MotionBenchMaker problems (committed fixtures: 5 per set, 14 sets) in one batched launch per
    seed (0, 1, 2), next to the CPU RRT-Connect port's outcome at the same 10,000-sample budget
    (tests/golden/mbm/cpu_port_status.json, written by make_cpu_port_status.py beside it)."""
    from isaac_sim_ros2_humble_amd import mbm
    gold = os.path.join(REPO, "tests", "golden", "mbm")
    mm = mbm.problem_models()
    rob = [native.Robot(m) for m in mm]
    probs = mbm.load_problems(gold, mm)
    scenes = [native.Scene(p[2]) for p in probs]
    by, statuses, ms, paths = {}, {}, [], []
    for seed in (0, 1, 2):
        native.sync()
        t0 = time.perf_counter()
        res = native.rrtc_solve_batch([rob[p[1]] for p in probs], scenes, [p[3] for p in probs], [p[4] for p in probs],
                                      native.default_params(seed=1000 * seed, sampler=sampler,
                                                            dynamic_domain=dynamic_domain, balance=balance))
        native.sync()
        ms.append((time.perf_counter() - t0) * 1e3)
        for p, r in zip(probs, res):
            d = by.setdefault(p[0], [0, 0])
            d[0] += r["status"] == "solved"
            d[1] += 1
            statuses[r["status"]] = statuses.get(r["status"], 0) + 1
            if r["status"] == "solved":
                paths.append((mm[p[1]], p[2], r["path"], p[3], p[4]))
    cpu = {}
    try:
        with open(os.path.join(gold, "cpu_port_status.json")) as f:
            for p in json.load(f)["problems"]:
                d = cpu.setdefault(p["set"], [0, 0])
                d[0] += sum(x == 1 for x in p["solved"])
                d[1] += len(p["solved"])
    except OSError:
        pass
    out = {"problems": len(probs), "seeds": 3, "solved": sum(v[0] for v in by.values()), "runs": 3 * len(probs),
           "batch_ms": float(np.mean(ms)), "statuses": statuses,
           "per_set_solved": {k: "%d/%d" % tuple(v) for k, v in sorted(by.items())},
           "cpu_port_per_set_solved": {k: "%d/%d" % tuple(v) for k, v in sorted(cpu.items())},
           "cpu_port_solved": sum(v[0] for v in cpu.values()) if cpu else None}
    return out, paths


RRG_SETTINGS = {"samples_per_iteration": 60, "max_travel_distance": 6.0, "resolution": 0.02}


def rrg_tasks(robot, model, n, seed):
    """evaluate_cudampl.cpp's RRG workload: FR3 with self collision and 20 sphere obstacles
    (radius U[0.03, 0.1], centres in [-0.8, 0.8]^2 x [0, 1.2] clear of the base column) per task,
    a collision-free start and goal whose straight line is blocked."""
    rng = np.random.default_rng(seed)
    tasks = []
    while len(tasks) < n:
        c = rng.uniform([-0.8, -0.8, 0.0], [0.8, 0.8, 1.2], size=(20, 3))
        c = c[np.hypot(c[:, 0], c[:, 1]) > 0.3]
        r = rng.uniform(0.03, 0.1, size=len(c))
        scene = workloads.Scene(spheres=np.concatenate([c, r[:, None]], axis=1).astype(np.float32))
        ns = native.Scene(scene)
        pair = workloads.sample_pair(model, scene, rng, lambda m, sc, q: native.check_states(robot, ns, q),
                                     lambda m, sc, a, b: native.check_edges(robot, ns, a, b, 0.02)[0],
                                     want_free_line=False)
        if pair is not None:
            tasks.append((scene, ns, pair))
    return tasks


def roadmap_path(g):
    """Shortest start-goal path over the roadmap's edge lengths (RRG.cpp:433-491), or None."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import dijkstra
    nv = g["nv"]
    if not g["solved"] or g["ne"] == 0:
        return None
    e, w = g["edges"], g["weights"].astype(np.float64)
    m = coo_matrix((np.concatenate([w, w]), (np.concatenate([e[:, 0], e[:, 1]]), np.concatenate([e[:, 1], e[:, 0]]))),
                   shape=(nv, nv)).tocsr()
    dist, pred = dijkstra(m, indices=0, return_predecessors=True)
    if not np.isfinite(dist[1]):
        return None
    path, v = [1], 1
    while v != 0:
        v = int(pred[v])
        path.append(v)
    return g["vertices"][path[::-1]]


def bench_rrg(robot, model, n_tasks, max_passes=2000):
    """The roadmap planner (RRG::solve's loop, RRG.cpp:241-424) with the roadmap in device memory,
    evaluate_cudampl.cpp's settings (travel distance 6.0, 60 samples per pass, resolution 0.02):
    per task, the time from the roadmap's creation to the graph read back, and Dijkstra's path."""
    tasks = rrg_tasks(robot, model, n_tasks, seed=11)
    times, passes, verts, paths, solved = [], [], [], [], 0
    for k, (scene, ns, (s, g)) in enumerate(tasks):
        native.sync()
        t0 = time.perf_counter()
        rm = native.Roadmap(robot, ns, s[None], g[None], seed=k, **RRG_SETTINGS)
        for _ in range(max_passes):
            if rm.iterate():
                break
        graph = rm.read()
        times.append((time.perf_counter() - t0) * 1e3)
        rm.close()
        passes.append(graph["passes"])
        verts.append(graph["nv"])
        path = roadmap_path(graph)
        if path is not None:
            solved += 1
            paths.append((model, scene, path, s, g))
    return ({"tasks": n_tasks, "solved": solved, "success_rate": solved / n_tasks, "settings": RRG_SETTINGS,
             "p50_ms": float(np.percentile(times, 50)), "p99_ms": float(np.percentile(times, 99)),
             "mean_ms": float(np.mean(times)), "passes_p50_max": [float(np.median(passes)), int(max(passes))],
             "vertices_p50_max": [float(np.median(verts)), int(max(verts))]}, paths, tasks)


def cpu_rrg(model, tasks, n, max_passes=2000):
    """The oracle's roadmap loop (the same passes, bit-identical graphs) on the first n tasks."""
    import oracle
    orc = oracle.OracleRobot(model)
    ms = []
    for k, (scene, _, (s, g)) in enumerate(tasks[:n]):
        t0 = time.perf_counter()
        oracle.rrg(orc, oracle.OracleScene(scene), s[None], g[None], RRG_SETTINGS["samples_per_iteration"],
                   RRG_SETTINGS["max_travel_distance"], RRG_SETTINGS["resolution"], k, 0, max_passes)
        ms.append((time.perf_counter() - t0) * 1e3)
    return {"tasks": n, "p50_ms": float(np.percentile(ms, 50)), "mean_ms": float(np.mean(ms)), "cores": 1,
            "kind": "port"}


def serve(robots, lib, rate, seconds, seed, sample_paths=1000, saturation_queries=0):
    """Served planning over a scene library (workloads.c5_stream): Poisson arrivals at `rate`
    (or, with saturation_queries, that many queries submitted at once), scenes reused with
    probability 0.8 and otherwise uploaded at arrival while the server plans, a fresh start/goal
    per query.  Returns the statistics and a sample of solved paths for the CPU re-check."""
    srv = native.Server(robots, tree_memory_bytes=16 << 30, tree_capacity=65536, max_active=1024, use_graph=True,
                        params=native.default_params())
    if saturation_queries:
        stream = workloads.c5_stream(lib, 20000.0, saturation_queries / 20000.0, p_reuse=0.8, seed=seed)
    else:
        stream = workloads.c5_stream(lib, rate, seconds, p_reuse=0.8, seed=seed)
    dev, tickets, uploads = {}, [], 0
    t0 = time.perf_counter()
    for t, s, p, up in stream:
        if not saturation_queries:
            while time.perf_counter() - t0 < t:
                time.sleep(0.0001)
        if up or s not in dev:
            # the server holds the scenes of its pending tickets, so nothing is freed (and no
            # device-wide synchronisation happens) while the stream runs
            dev.pop(s, None)
            dev[s] = native.Scene(lib[s][1])
            uploads += 1
            if len(dev) > 512:
                dev.pop(next(iter(dev)))
        a, b = lib[s][2][p]
        tickets.append((srv.submit(lib[s][0], dev[s], a, b, seed=len(tickets)), s, p))
    submit_s = time.perf_counter() - t0
    at_submit = srv.stats()
    t_sub = time.perf_counter()
    srv.wait_idle(600.0)
    t_idle = time.perf_counter()
    st = srv.stats()
    statuses, paths = {}, []
    every = max(1, len(tickets) // sample_paths)
    for i, (tk, s, p) in enumerate(tickets):
        r = srv.poll(tk)
        statuses[r["status"]] = statuses.get(r["status"], 0) + 1
        if r["status"] == "solved" and i % every == 0:
            paths.append((s, p, r["path"]))
    srv.close()
    out = {"rate_qps": rate if not saturation_queries else None, "seconds": seconds if not saturation_queries else None,
           "submitted": st["submitted"], "completed": st["completed"], "solved": st["solved"],
           "success_rate": st["solved"] / max(st["completed"], 1), "scene_uploads": uploads,
           "throughput_qps": st["throughput_qps"], "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
           "mean_ms": st["mean_ms"], "steps": st["steps"], "worker_step_ms": st["step_ms"],
           "worker_loop_ms": st["loop_ms"], "backlog_steps": st["backlog_steps"], "statuses": statuses,
           "submit_seconds": submit_s}
    if saturation_queries:
        # after the last submit every query is queued or running: the completions from there to
        # idle are the server's rate with a full backlog (the submit loop itself, ~13 us per
        # call from Python, caps the arrival rate before that)
        drained = st["completed"] - at_submit["completed"]
        out["backlog_at_last_submit"] = at_submit["queued"] + at_submit["active"]
        out["drain_qps"] = drained / max(t_idle - t_sub, 1e-9)
        out["drain_queries"] = drained
    return out, paths


def cpu_leg(models, queries, gpu_results, budget_s, threads, checks):
    """The CPU leg.  (1) The baseline: the oracle's RRT-Connect (OMPL RRTConnect semantics, range
    0.3 rad, resolution 0.02 rad) on the first batched queries, `threads` queries at a time.  (2)
    The CPU checker over the C2 configurations.  (3) The independent re-check, at 0.005 rad, of
    the paths the GPU legs returned (the oracle as the checker)."""
    import oracle
    orc = {m.name: oracle.OracleRobot(m) for m in models}
    for items in checks.values():  # the MBM legs plan with their own robot variants
        for it in items:
            if it[0].name not in orc:
                orc[it[0].name] = oracle.OracleRobot(it[0])
    per_query_limit = 2.0
    nsample = max(1, min(len(queries), int(budget_s * threads / per_query_limit)))
    sample = queries[:nsample]

    def one(q):
        k, scene, s, g, _ = q
        m = models[k]
        osc = oracle.OracleScene(scene)
        path, st = oracle.rrtconnect(orc[m.name], osc, s, g, range_=0.3, resolution=0.02, seed=0,
                                     time_limit=per_query_limit)
        ok = st.solved == 1
        valid = ok and oracle.validate_path(orc[m.name], osc, path, VALIDATION_RES)[0]
        return ok, valid
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(one, sample))
    dt = time.perf_counter() - t0
    solved = sum(r[0] for r in res)
    valid = sum(r[1] for r in res)
    # the GPU planners on the same queries (first step of the batched run, and the plain run)
    same = {}
    for name, (status, paths) in gpu_results.items():
        sol = [i for i in range(nsample) if status[i] == "solved"]
        same[name] = {"solved": len(sol), "valid": sum(
            oracle.validate_path(orc[models[sample[i][0]].name], oracle.OracleScene(sample[i][1]), paths[i],
                                 VALIDATION_RES)[0] for i in sol)}
    baseline = {"value": solved / dt, "unit": "queries/s", "cores": threads, "kind": "port",
                "sample": "first %d of the %d batched queries, RRT-Connect range 0.3 rad, resolution 0.02 rad, "
                          "%.0f s limit per query, %d solved in %.1f s on %d threads" % (
                              nsample, len(queries), per_query_limit, solved, dt, threads),
                "queries": nsample, "solved": solved, "success_rate": solved / nsample,
                "validity_rate": valid / max(solved, 1), "gpu_same_queries": same}

    def recheck(item):
        m, scene, path, s, g = item
        ends = np.array_equal(path[0], s) and np.array_equal(path[-1], g)
        return ends and oracle.validate_path(orc[m.name], oracle.OracleScene(scene), path, VALIDATION_RES)[0]
    validity = {}
    with cf.ThreadPoolExecutor(threads) as ex:
        for name, items in checks.items():
            ok = sum(ex.map(recheck, items))
            validity[name] = {"paths": len(items), "valid": ok, "rate": ok / max(len(items), 1)}
    return baseline, validity


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
        path, st = oracle.rrtconnect(orc, osc, pairs[k][0], pairs[k][1], range_=0.3, resolution=0.02, seed=k,
                                     time_limit=10.0)
        dt = time.perf_counter() - t0
        return dt, st.solved == 1, st.solved == 1 and oracle.validate_path(orc, osc, path, VALIDATION_RES)[0]
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        cpu = list(ex.map(cpu_one, range(seeds)))
    wall = time.perf_counter() - t0
    gpu, gpaths = [], []
    for k in range(seeds):
        r = native.rrtc_solve(robot, nscene, pairs[k][0], pairs[k][1], native.default_params(seed=k))
        gpu.append((r["latency_ms"] / 1e3, r["status"] == "solved"))
        if r["status"] == "solved":
            gpaths.append((model, scene, r["path"], pairs[k][0], pairs[k][1]))
    ct = np.array([c[0] for c in cpu]) * 1e3
    gt = np.array([g[0] for g in gpu]) * 1e3
    csol = int(sum(c[1] for c in cpu))
    return ({"seeds": seeds, "cpu_seed0_ms": float(ct[0]), "cpu_median_ms": float(np.median(ct)),
             "cpu_solved": csol, "cpu_validity_rate": sum(c[2] for c in cpu) / max(csol, 1), "cpu_threads": threads,
             "cpu_qps": float(csol / wall), "gpu_seed0_ms": float(gt[0]),
             "gpu_median_ms": float(np.median(gt)), "gpu_solved": int(sum(g[1] for g in gpu)),
             "median_speedup": float(np.median(ct) / np.median(gt))}, gpaths)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="seconds per served rate (the configured workload is 60 s)")
    ap.add_argument("--rate", type=float, nargs="*", default=[100.0, 1000.0, 5000.0])
    ap.add_argument("--saturation-queries", type=int, default=20000,
                    help="queries submitted at once for the served throughput at saturation (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--c3-seeds", type=int, default=1000)
    ap.add_argument("--c1-seeds", type=int, default=100)
    ap.add_argument("--rrg-tasks", type=int, default=32)
    ap.add_argument("--tune-queries", type=int, default=256)
    ap.add_argument("--team-waves", type=int, default=0, help="planning launch shape (0: tuned / default)")
    ap.add_argument("--waves", type=int, default=0, help="planning grid in waves (0: tuned / resident)")
    ap.add_argument("--teams-per-query", type=int, default=0, help="teams per query (0: tuned / unbounded)")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per iteration, 32 or 64 (0: tuned / default)")
    ap.add_argument("--sampler", choices=["uniform", "halton"], default="uniform",
                    help="sample streams of the batched planner: Philox uniform or rotated Halton")
    ap.add_argument("--dynamic-domain", action="store_true",
                    help="dynamic-domain sampling in the batched planner (pRRTC's settings)")
    ap.add_argument("--balance", choices=["smaller", "split", "dynamic", "alternate"], default="smaller",
                    help="tree balance of the batched planner (CPRRTC's smaller tree, or pRRTC's modes)")
    ap.add_argument("--skip", nargs="*", default=[],
                    help="parts to skip: tune plain serve c2 c3 mbm rrg cpu (cpu also skips c1 and the CPU re-check)")
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
    # full-evaluation flops of one checked configuration, per query (its robot and scene)
    f_query = np.array([full_flops_per_config(models[k], robots[models[k].name].P, sc.counts)
                        for k, sc, _, _, _ in queries], np.float64)

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
                                               per_query=per_query, reps=2 if per_query else 12)
            tune[name] = {"team_waves": best.team_waves, "waves": best.waves, "teams_per_query": best.teams_per_query,
                          "lanes_per_iteration": best.lanes_per_iteration, "calibration_queries": n,
                          "candidates": cands}
            log("autotune %s: team_waves %d waves %d teams_per_query %d lanes_per_iteration %d" % (
                name, best.team_waves, best.waves, best.teams_per_query, best.lanes_per_iteration))
    params = native.default_params(seed=0, team_waves=a.team_waves, waves=a.waves, teams_per_query=a.teams_per_query,
                                   lanes_per_iteration=a.lanes,
                                   sampler=native.SAMPLER_HALTON if a.sampler == "halton" else native.SAMPLER_UNIFORM,
                                   dynamic_domain=1 if a.dynamic_domain else 0,
                                   balance=["smaller", "split", "dynamic", "alternate"].index(a.balance))

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
    step_ms, kern_ms, solved, flops, model_flops, checked, host_ms = [], [], [], [], [], [], []
    batch_checks, first_step = [], None
    for k in range(a.steps):
        params.seed = 1000 * (k + 1)
        res, ms, kms = step(params)
        # untimed from here: keep the step's paths for the CPU re-check
        step_ms.append(ms)
        kern_ms.append(kms)
        solved.append(res.solved)
        flops.append(float(res.work_flops.sum()))
        model_flops.append(float((res.configs_checked.astype(np.float64) * f_query).sum()))
        checked.append(float(res.configs_checked.sum()))
        host_ms.append(float(res.latency_ms[0]))  # the library's own submit-to-results time
        paths = [res.path(i) for i in range(req.nq)]
        if first_step is None:
            first_step = (list(res.status), paths)
        batch_checks += [(models[queries[i][0]], queries[i][1], paths[i], starts[i], goals[i])
                         for i in range(req.nq) if res.status[i] == "solved"]
        log("batch step %d: %.1f ms (kernel %.1f ms), solved %d/%d" % (k, ms, kms, solved[-1], req.nq))
    ms_per_step = float(np.mean(step_ms))
    batched_qps = float(np.sum(solved)) / (np.sum(step_ms) / 1e3)
    plan_kernel_ms = float(np.mean(kern_ms))
    achieved_model = float(np.mean(model_flops)) / (plan_kernel_ms / 1e3) / 1e12
    achieved_exec = float(np.mean(flops)) / (plan_kernel_ms / 1e3) / 1e12
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
                      "sampler": a.sampler, "dynamic_domain": bool(a.dynamic_domain), "balance": a.balance},
           "batched": {"solved_per_step": solved, "success_rate": float(np.mean(solved)) / len(queries),
                       "plan_kernel_ms": plan_kernel_ms, "configs_checked_per_step": float(np.mean(checked)),
                       "host_ms": float(np.mean(host_ms)), "iterations_p50_p99_max": [
                           float(np.percentile(iters, 50)), float(np.percentile(iters, 99)), int(iters.max())],
                       "query_device_ms_p50_p99_max": [float(np.percentile(dev_ms, 50)),
                                                       float(np.percentile(dev_ms, 99)), float(dev_ms.max())],
                       "phase_fraction_of_wave_time": phase},
           "roofline": {"bound": "valu", "achieved": achieved_model, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved_model / PEAK_FP32_TFLOPS, "traffic": measured_traffic("mp_plan"),
                        "kernel": "mp_plan (batched)",
                        "definition": "planner work = configurations checked x full-evaluation flops per "
                                      "configuration of the query's robot and scene (FK + sphere transforms + "
                                      "self pairs + every sphere-obstacle pair), per launch, over the mean "
                                      "HIP-event time of the planning kernel",
                        "flops_per_launch": float(np.mean(model_flops)),
                        "achieved_executed": achieved_exec, "frac_executed": achieved_exec / PEAK_FP32_TFLOPS,
                        "executed_flops_per_launch": float(np.mean(flops)),
                        "executed_definition": "the flops the kernel actually evaluated (bounding-sphere culls "
                                               "and early exits skip the rest), counted in-kernel",
                        "launch_shape": {k: shape[k] for k in ("team_waves", "waves", "teams_per_query",
                                                               "lanes_per_iteration")} if shape else None},
           "autotune": tune}
    checks = {"c4_batched": batch_checks}
    gpu_same = {"batched": first_step}

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
        ppaths = [pr.path(i) for i in range(req.nq)]
        checks["c4_plain"] = [(models[queries[i][0]], queries[i][1], ppaths[i], starts[i], goals[i])
                              for i in range(req.nq) if pr.status[i] == "solved"]
        gpu_same["plain"] = (list(pr.status), ppaths)
        out["plain"] = {"qps": plain_qps, "solved": nsol, "success_rate": nsol / req.nq,
                        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                        "seconds": dt}
        out["batched_vs_plain"] = batched_qps / plain_qps if plain_qps > 0 else None
        log("plain: %.1f q/s" % plain_qps)

    if "serve" not in a.skip:
        def check(m, scene, q):
            return native.check_states(robots[m.name], native.Scene(scene), q)

        def edge(m, scene, q0, q1):
            return native.check_edges(robots[m.name], native.Scene(scene), q0, q1, 0.02)[0]
        t0 = time.perf_counter()
        lib = workloads.c5_library(models, 256, 64, check, edge, seed=0)
        log("c5 scene library: %d scenes in %.1f s" % (len(lib), time.perf_counter() - t0))
        srobots = [robots[m.name] for m in models]
        served, served_checks = [], []
        for i, rate in enumerate(a.rate):
            r, sp = serve(srobots, lib, rate, a.serve_seconds, seed=10 + i)
            r["configured_seconds"] = CONFIGURED_SERVE_SECONDS
            served.append(r)
            served_checks += [(models[lib[s][0]], lib[s][1], path, *lib[s][2][p]) for s, p, path in sp]
            log("served %.0f q/s: %s" % (rate, {k: v for k, v in r.items() if k != "statuses"}))
        out["served"] = served
        out["served_note"] = ("%.0f s per rate (the configured workload runs 60 s per rate: --serve-seconds 60); "
                              "scenes reused with probability 0.8, otherwise uploaded at arrival while the server "
                              "plans; each query a fresh start/goal pair from its scene's pool of 64" % a.serve_seconds)
        if a.saturation_queries > 0:
            r, sp = serve(srobots, lib, None, None, seed=99, saturation_queries=a.saturation_queries)
            out["served_saturation"] = r
            served_checks += [(models[lib[s][0]], lib[s][1], path, *lib[s][2][p]) for s, p, path in sp]
            if "plain" in out:
                out["served_vs_plain"] = r["throughput_qps"] / out["plain"]["qps"]
            log("served saturation: %s" % {k: v for k, v in r.items() if k != "statuses"})
            if "plain" not in a.skip:
                # the plain planner on the served workload itself: the first 256 queries of the
                # saturation stream, one launch each, scenes resident beforehand
                first = [(s_, p_) for _, s_, p_, _ in workloads.c5_stream(lib, 20000.0, a.saturation_queries / 20000.0,
                                                                          p_reuse=0.8, seed=99)][:256]
                dscenes = {s_: native.Scene(lib[s_][1]) for s_, _ in first}
                native.sync()
                t0 = time.perf_counter()
                nsol = 0
                for k, (s_, p_) in enumerate(first):
                    st_, gl_ = lib[s_][2][p_]
                    res_ = native.rrtc_solve(robots[models[lib[s_][0]].name], dscenes[s_], st_, gl_,
                                             native.default_params(seed=k))
                    nsol += res_["status"] == "solved"
                native.sync()
                pq = nsol / (time.perf_counter() - t0)
                out["plain_served_workload"] = {"qps": pq, "queries": len(first), "solved": nsol}
                out["served_vs_plain_same_workload"] = r["throughput_qps"] / pq if pq > 0 else None
                out["served_drain_vs_plain_same_workload"] = r["drain_qps"] / pq if pq > 0 else None
                log("plain on the served workload: %.1f q/s" % pq)
        checks["served"] = served_checks

    if "c2" not in a.skip:
        out["c2"] = bench_c2(robots["fr3"], models[0], max(a.steps, 3), 1)
        c2v = out["c2"]["verdict"]
        out["roofline_c2"] = {"bound": "valu", "kernel": "mp_check_0 (2^20 FR3 configs, 64 obstacles)",
                              "achieved": c2v["achieved_tflops"], "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                              "frac": c2v["achieved_tflops"] / PEAK_FP32_TFLOPS, "hbm_gbs": c2v["hbm_gbs"],
                              "hbm_frac": c2v["hbm_gbs"] / PEAK_HBM_GBS}
        log("c2:", out["c2"])

    if "c3" not in a.skip and a.c3_seeds > 0:
        out["c3"], checks["c3"] = bench_c3(robots["fetch"], models[1], a.c3_seeds)
        log("c3:", out["c3"])

    if "mbm" not in a.skip:
        out["mbm"], checks["mbm"] = bench_mbm(models, params.sampler, params.dynamic_domain, params.balance)
        log("mbm:", out["mbm"])

    rrg_tasks_ = None
    if "rrg" not in a.skip and a.rrg_tasks > 0:
        out["rrg"], checks["rrg"], rrg_tasks_ = bench_rrg(robots["fr3"], models[0], a.rrg_tasks)
        log("rrg:", out["rrg"])

    if "cpu" not in a.skip:
        host = host_info()
        out["host"] = host
        threads = host["threads_used"]
        log("host:", host)
        if a.c1_seeds > 0:
            out["c1"], checks["c1_gpu"] = bench_c1(robots["fr3"], models[0], a.c1_seeds, threads)
            log("c1:", out["c1"])
        out["cpu_baseline"], out["validity_0p005"] = cpu_leg(models, queries, gpu_same, a.cpu_seconds, threads, checks)
        log("cpu:", out["cpu_baseline"])
        log("validity:", out["validity_0p005"])
        if "c4_batched" in out["validity_0p005"]:
            v = out["validity_0p005"]["c4_batched"]
            out["batched"]["validity_rate_0p005"] = v["rate"]
            out["batched"]["valid_solved_qps"] = batched_qps * v["rate"]
        out["cpu_checker_c2"] = cpu_checker_c2(models[0], threads)
        log("cpu c2:", out["cpu_checker_c2"])
        if rrg_tasks_:
            out["rrg"]["cpu_port"] = cpu_rrg(models[0], rrg_tasks_, min(8, len(rrg_tasks_)))
            out["rrg"]["speedup_p50"] = out["rrg"]["cpu_port"]["p50_ms"] / out["rrg"]["p50_ms"]
            log("cpu rrg:", out["rrg"]["cpu_port"])
        if "c2" in out:  # C2's early-out throughput at the canonical-order work count
            cf_ = out["cpu_checker_c2"]["canonical_flops_per_config"]
            v = out["c2"]["verdict"]
            v["canonical_flops_per_config"] = cf_
            v["canonical_tflops"] = cf_ * (1 << 20) / (v["ms"] / 1e3) / 1e12
            if "roofline_c2" in out:
                out["roofline_c2"]["achieved_canonical"] = v["canonical_tflops"]
                out["roofline_c2"]["frac_canonical"] = v["canonical_tflops"] / PEAK_FP32_TFLOPS
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
