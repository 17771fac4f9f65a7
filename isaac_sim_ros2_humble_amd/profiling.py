"""Measurement helpers for rocprofv3 runs of the planning hot path.

    python -m isaac_sim_ros2_humble_amd.profiling copy --mib 512 --reps 5

launches a plain device-to-device copy kernel (torch's copy_ on two resident fp32 buffers) of a
known size, so that the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run can be scaled
to bytes: factor = known bytes / counter bytes (both per dispatch).
"""
import argparse
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy"])
    ap.add_argument("--mib", type=int, default=512)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args(argv)
    if a.what == "copy":
        copy_kernel(a.mib, a.reps)


if __name__ == "__main__":
    sys.exit(main())
