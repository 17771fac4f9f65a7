"""Measurement helpers for rocprofv3 runs of the planning hot path.

    python -m isaac_sim_ros2_humble_amd.profiling copy --mib 512 --reps 3
        launches a plain device-to-device copy (torch's copy_ of one resident fp32 buffer into
        another) of a known size, so that FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc
        run can be scaled to bytes: factor = bytes the copy moves / bytes the counter reports,
        per dispatch.

    python -m isaac_sim_ros2_humble_amd.profiling summarize|traffic <dir> [--command "..."]
        reads the CSV output of the runs below (one sub-directory each) and prints a markdown
        summary: kernel-trace statistics, HBM bytes per dispatch (counter x copy factor) and the
        SQ instruction counters per kernel.

            <dir>/kt            rocprofv3 --kernel-trace --stats -f csv
            <dir>/copy_FETCH_SIZE, <dir>/copy_WRITE_SIZE
                                rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE -f csv (copy above)
            <dir>/pmc_FETCH_SIZE, <dir>/pmc_WRITE_SIZE, <dir>/pmc_sq
                                rocprofv3 --pmc ... -f csv (the workload)
"""
import argparse
import collections
import csv
import glob
import os
import sys


def copy_kernel(mib, reps):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copy kernel: %d launches of %d bytes read + %d bytes written" % (reps, n * 4, n * 4), flush=True)


def _csv(path_glob):
    files = sorted(glob.glob(path_glob))
    rows = []
    for f in files:
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def _group(kernel, grid):
    # the planning kernels and the 2^20-configuration checks by their grid; the many small
    # dispatches of workload generation (edge and state checks) as one row per kernel
    return (kernel, grid) if kernel.startswith("mp_plan") or grid >= (1 << 20) else (kernel, 0)


def counters(d):
    """{(kernel, grid or 0 for small grids): {counter: [values per dispatch]}} from a
    counter_collection.csv."""
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in _csv(os.path.join(d, "*counter_collection.csv")):
        out[_group(r["Kernel_Name"], int(r["Grid_Size"]))][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return out


def copy_factors(root, copy_bytes):
    """Bytes per counter unit for FETCH_SIZE and WRITE_SIZE, from the calibration copy."""
    f = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = []
        for (k, g), cs in counters(os.path.join(root, "copy_" + c)).items():
            if "copyBuffer" in k or "copy" in k.lower():
                vals += [v for v in cs.get(c, []) if v > 0]
        big = max(vals) if vals else None
        f[c] = copy_bytes / big if big else None
    return f


def summarize(root, command, copy_bytes):
    lines = []
    lines.append("Command: `%s`" % command if command else "")
    kt = _csv(os.path.join(root, "kt", "*kernel_stats.csv"))
    if kt:
        lines += ["", "## Kernel trace (rocprofv3 --kernel-trace --stats)", "",
                  "| kernel | calls | avg us | min us | max us | total % |", "|---|---|---|---|---|---|"]
        for r in sorted(kt, key=lambda r: -float(r["TotalDurationNs"])):
            lines.append("| %s | %s | %.1f | %.1f | %.1f | %.1f |" % (
                r["Name"][:60], r["Calls"], float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    fac = copy_factors(root, copy_bytes)
    lines += ["", "## HBM traffic per dispatch (FETCH_SIZE / WRITE_SIZE x copy-kernel factor)", "",
              "Copy calibration: %d bytes read and written per dispatch; factor FETCH_SIZE %s, WRITE_SIZE %s "
              "bytes per counter unit." % (copy_bytes, "%.1f" % fac["FETCH_SIZE"] if fac["FETCH_SIZE"] else "n/a",
                                            "%.1f" % fac["WRITE_SIZE"] if fac["WRITE_SIZE"] else "n/a"), "",
              "| kernel | grid | dispatches | fetch bytes | write bytes |", "|---|---|---|---|---|"]
    fe = counters(os.path.join(root, "pmc_FETCH_SIZE"))
    wr = counters(os.path.join(root, "pmc_WRITE_SIZE"))
    for key in sorted(set(fe) | set(wr), key=lambda k: (k[0], k[1])):
        if key[0].startswith("__amd"):
            continue
        fv = fe.get(key, {}).get("FETCH_SIZE", [])
        wv = wr.get(key, {}).get("WRITE_SIZE", [])
        mf = sum(fv) / len(fv) * fac["FETCH_SIZE"] if fv and fac["FETCH_SIZE"] else float("nan")
        mw = sum(wv) / len(wv) * fac["WRITE_SIZE"] if wv and fac["WRITE_SIZE"] else float("nan")
        lines.append("| %s | %s | %d | %.4g | %.4g |" % (key[0][:40], key[1] or "small", max(len(fv), len(wv)), mf, mw))
    sq = counters(os.path.join(root, "pmc_sq"))
    if sq:
        names = sorted({c for v in sq.values() for c in v})
        lines += ["", "## SQ counters per dispatch (mean)", "",
                  "| kernel | grid | " + " | ".join(names) + " |", "|---|---|" + "---|" * len(names)]
        for key in sorted(sq, key=lambda k: (k[0], k[1])):
            if key[0].startswith("__amd"):
                continue
            v = sq[key]
            lines.append("| %s | %s | " % (key[0][:40], key[1] or "small") +
                         " | ".join("%.4g" % (sum(v[n]) / len(v[n])) if v.get(n) else "" for n in names) + " |")
    return "\n".join(lines) + "\n"


def traffic_table(root, copy_bytes):
    """`kernel fetch_bytes write_bytes` per planning / C2 kernel (largest grid of each), the
    lines bench.py reads for its roofline's traffic."""
    fac = copy_factors(root, copy_bytes)
    fe = counters(os.path.join(root, "pmc_FETCH_SIZE"))
    wr = counters(os.path.join(root, "pmc_WRITE_SIZE"))
    best = {}
    for key in set(fe) & set(wr):
        if key[0].startswith("__amd") or not key[1]:
            continue
        if key[0] not in best or key[1] > best[key[0]][1]:
            best[key[0]] = key
    out = []
    for k, key in sorted(best.items()):
        fv, wv = fe[key]["FETCH_SIZE"], wr[key]["WRITE_SIZE"]
        out.append("%s %.6g %.6g" % (k, sum(fv) / len(fv) * fac["FETCH_SIZE"], sum(wv) / len(wv) * fac["WRITE_SIZE"]))
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "summarize", "traffic"])
    ap.add_argument("root", nargs="?")
    ap.add_argument("--mib", type=int, default=512)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--command", default="")
    a = ap.parse_args(argv)
    if a.what == "copy":
        copy_kernel(a.mib, a.reps)
    elif a.what == "traffic":
        sys.stdout.write(traffic_table(a.root, a.mib * (1 << 20)))
    else:
        sys.stdout.write(summarize(a.root, a.command, a.mib * (1 << 20)))


if __name__ == "__main__":
    sys.exit(main())
