"""Autotuner for the split of source ranges across waves, per size bucket.

For every bucket b (systems of n with floor(log2 n) == b, sampled at
n = 2^b and n = 1.5 * 2^b), each candidate part count k in {1, 2, 4, 8, 16, 32} (with
at least 64 sources per part) is timed on a Plummer sphere stepped with
velocity Verlet in both f32-compensated forms (f32r, f32c) and the f64
arithmetic.  The clocks are warmed first, and every candidate is timed twice
(candidates in increasing, then decreasing order; the faster time counts),
so a clock ramp cannot favour the candidates timed last.  The k with the
smallest summed time wins.  The table is written as
`split <bucket> <parts>` lines (the format the library loads at start-up
from tuning/gfx950.txt next to libnbody_amd.so).

Run on the GPU from the repository root:
    python n-body-sim_amd/nbody_amd/autotune.py [--out PATH] [--max-bucket 20]
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import nbody_amd as nb  # noqa: E402

CANDIDATES = (1, 2, 4, 8, 16, 32)
TUNED = (nb.Precision.F32R, nb.Precision.F32C, nb.Precision.F64)


def time_case(gpu, bodies, prec, budget_ms=120.0):
    """ms per velocity-Verlet step (one force evaluation per step)."""
    lib = nb.load_library()
    db = nb.DeviceBodies(gpu, bodies)
    try:
        args = (gpu.handle, db.handle, 1, 1, int(prec), 5e-4, 1.0, 0.005 ** 2)
        nb._check(lib.nbody_step(*args, 2))          # prime + warm up
        gpu.synchronize()
        gpu.record(0)
        nb._check(lib.nbody_step(*args, 1))
        gpu.record(1)
        one = max(gpu.elapsed_ms(0, 1), 1e-3)
        steps = int(max(2, min(200, budget_ms / one)))
        gpu.record(0)
        nb._check(lib.nbody_step(*args, steps))
        gpu.record(1)
        return gpu.elapsed_ms(0, 1) / steps
    finally:
        db.close()


def tune(max_bucket=20, min_bucket=6, log=print):
    gpu = nb.GpuContext(0)
    table = nb.tuning_split()
    report = []
    warm = nb.plummer(65536, 1)
    t_end = time.time() + 1.0
    while time.time() < t_end:           # bring the clocks up
        time_case(gpu, warm, nb.Precision.F64, budget_ms=200.0)
    for b in range(min_bucket, max_bucket + 1):
        sizes = [1 << b] + ([3 << (b - 1)] if b < max_bucket else [])
        systems = [nb.plummer(n, 7 + n) for n in sizes]
        cands = [k for k in CANDIDATES if min(sizes) // k >= 64]
        times = {}
        for order in (cands, cands[::-1]):
            for k in order:
                trial = list(table)
                trial[b] = k
                nb.set_tuning_split(trial)
                t = 0.0
                for bodies in systems:
                    for prec in TUNED:
                        t += time_case(gpu, bodies, prec)
                times[k] = min(times.get(k, t), t)
        row = [(k, times[k]) for k in cands]
        best = min(cands, key=lambda k: times[k])
        table[b] = best
        nb.set_tuning_split(table)
        log(f"bucket {b:2d} (n={sizes}): " +
            " ".join(f"k={k}:{t:.3f}ms" for k, t in row) + f" -> {best}")
        report.append((b, sizes, row, best))
    gpu.close()
    return table, report


def write_table(path, table, report, max_bucket):
    lines = ["# source-range parts per target tile by bucket floor(log2 n)",
             "# written by nbody_amd/autotune.py on " + time.strftime("%Y-%m-%d"),
             "# sum of ms/step (f32r + f32c + f64, velocity Verlet, Plummer) per candidate:"]
    for b, sizes, row, best in report:
        lines.append("#   bucket %d n=%s: %s -> %d" % (
            b, sizes, " ".join(f"{k}:{t:.3f}" for k, t in row), best))
    for b in range(len(table)):
        lines.append(f"split {b} {table[b]}")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(PKG, "tuning", "gfx950.txt"))
    ap.add_argument("--max-bucket", type=int, default=20)
    ap.add_argument("--min-bucket", type=int, default=6)
    a = ap.parse_args()
    table, report = tune(a.max_bucket, a.min_bucket,
                         log=lambda m: print(m, flush=True))
    write_table(a.out, table, report, a.max_bucket)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
