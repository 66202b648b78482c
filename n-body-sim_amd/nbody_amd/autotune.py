"""Autotuner of the force kernel per size bucket: the split of source ranges
across waves, then the kernel variant (waves per workgroup and loop shape)
per arithmetic mode.

Stage "split":

For every bucket b (systems of n with floor(log2 n) == b, sampled at
n = 2^b and n = 1.5 * 2^b), each candidate part count k in {1, 2, 4, 8, 16, 32} (with
at least 64 sources per part) is timed on a Plummer sphere stepped with
velocity Verlet in both f32-compensated forms (f32r, f32c) and the f64
arithmetic.  The clocks are warmed first, and every candidate is timed twice
(candidates in increasing, then decreasing order; the faster time counts),
so a clock ramp cannot favour the candidates timed last.  The k with the
smallest summed time wins.

Stage "variant": with the split table in place, every force-kernel variant
v = 3 * w + shape (w: 1, 2 or 4 waves per workgroup; shape: sources per
chunk and chunks per compensated group in f64 and f32c, sources per LDS
batch in f32r) is timed per mode and bucket, twice in opposite orders (the
faster time counts).  Buckets below 2^14 bodies are timed as a batch of
equal systems (the batched and served regime, graph-stepped velocity
Verlet); larger ones as one system.  Systems of different variants in one
batch need one launch per variant, so each mode takes one uniform variant
(best summed over the buckets) and a bucket keeps its own best only where
that is at least 8 % faster (choose_variants).  Buckets outside the timed
range (below 2^6 and above the largest timed one) take the uniform variant.

The table is written as `split <bucket> <parts>` and `variant <mode>
<bucket> <v>` lines (the format the library loads at start-up from
tuning/gfx950.txt next to libnbody_amd.so); the stage not run is copied from
the current table.

Run on the GPU from the repository root:
    python n-body-sim_amd/nbody_amd/autotune.py [--stage split|variant|all]
        [--out PATH] [--max-bucket 20]
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


VARIANT_MODES = (nb.Precision.F32R, nb.Precision.F32C, nb.Precision.F64)
BATCH_BELOW = 14          # buckets below 2^14 bodies: timed as a batch


def time_variant(gpu, bodies, prec, budget_ms=150.0):
    """ms per step: a batch of copies of `bodies` (enough for about 2^31
    pair interactions per step, at most 1024 systems) stepped by graph, or
    one system for the large buckets."""
    n = len(bodies)
    if n >= 1 << BATCH_BELOW:
        return time_case(gpu, bodies, prec, budget_ms)
    count = max(1, min(1024, (1 << 31) // (n * n)))
    bt = nb.Batch(gpu, [bodies] * count, [1] * count, 5e-4, 1.0, 0.005 ** 2, prec)
    try:
        bt.step(2)
        gpu.synchronize()
        gpu.record(0)
        bt.step(1)
        gpu.record(1)
        one = max(gpu.elapsed_ms(0, 1), 1e-3)
        steps = int(max(3, min(100, budget_ms / one)))
        gpu.record(0)
        bt.step(steps)
        gpu.record(1)
        return gpu.elapsed_ms(0, 1) / steps
    finally:
        bt.close()


OVERRIDE_GAIN = 0.08      # a bucket leaves the uniform variant only if 8 % faster


def choose_variants(times):
    """Variant per bucket from times[b][v] (ms per step).  Systems of
    different variants in one batch run in separate launches (one per
    variant and round), so each launch drains before the next and the
    longest-first order of the round is broken.  Hence one uniform variant
    per mode (the smallest sum of times normalised by each bucket's best),
    and a bucket takes its own best variant only where that is at least
    OVERRIDE_GAIN faster than the uniform one.  Returns (uniform, {b: v})."""
    nv = len(next(iter(times.values())))
    score = [sum(t[v] / min(t) for t in times.values()) for v in range(nv)]
    uniform = min(range(nv), key=lambda v: score[v])
    pick = {}
    for b, t in times.items():
        best = min(range(nv), key=lambda v: t[v])
        pick[b] = best if t[uniform] > (1.0 + OVERRIDE_GAIN) * t[best] else uniform
    return uniform, pick


def variant_table(uniform, pick, nbuckets):
    """The full table of one mode: the chosen variant in the timed buckets
    (pick: {bucket: v}), the uniform variant everywhere else, so a round
    mixing very small (or very large) systems with the rest stays one
    launch."""
    table = [uniform] * nbuckets
    for b, v in pick.items():
        table[b] = v
    return table


def tune_variants(max_bucket=20, min_bucket=6, log=print):
    gpu = nb.GpuContext(0)
    report = []
    tables = {}
    warm = nb.plummer(65536, 1)
    t_end = time.time() + 1.0
    while time.time() < t_end:           # bring the clocks up
        time_case(gpu, warm, nb.Precision.F64, budget_ms=200.0)
    for prec in VARIANT_MODES:
        table = nb.tuning_variant(prec)
        measured = {}
        for b in range(min_bucket, max_bucket + 1):
            bodies = nb.plummer(1 << b, 11 + b)
            cands = list(range(9))
            times = {}
            for order in (cands, cands[::-1]):
                for v in order:
                    trial = list(table)
                    trial[b] = v
                    nb.set_tuning_variant(prec, trial)
                    t = time_variant(gpu, bodies, prec)
                    times[v] = min(times.get(v, t), t)
            nb.set_tuning_variant(prec, table)
            measured[b] = [times[v] for v in cands]
            log(f"{prec.name} bucket {b:2d}: " +
                " ".join(f"v{v}:{times[v]:.4f}" for v in cands))
        uniform, pick = choose_variants(measured)
        table = variant_table(uniform, pick, len(table))
        for b in range(min_bucket, max_bucket + 1):
            report.append((prec, b, list(enumerate(measured[b])), pick[b]))
        log(f"{prec.name}: uniform variant {uniform}; per bucket {pick}")
        nb.set_tuning_variant(prec, table)
        tables[prec] = table
    gpu.close()
    return tables, report


def read_comments(path, prefix):
    """The comment block of one stage from an existing table (kept when the
    other stage is re-run)."""
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return []
    out, on = [], False
    for line in lines:
        if line.startswith("# " + prefix):
            on = True
        elif line.startswith("# ") and not line.startswith("#   "):
            on = False
        if on and line.startswith("#"):
            out.append(line)
    return out


def current_table_path():
    """The table the library loaded (NBODY_AMD_TUNING, else the one next to
    the library)."""
    return os.environ.get("NBODY_AMD_TUNING") or os.path.join(PKG, "tuning", "gfx950.txt")


def write_table(path, split, split_report, variants, variant_report):
    day = time.strftime("%Y-%m-%d")
    lines = []
    current = current_table_path()
    if split_report is not None:
        lines += ["# source-range parts per target tile by bucket floor(log2 n)",
                  "#   written by nbody_amd/autotune.py on " + day,
                  "#   sum of ms/step (f32r + f32c + f64, velocity Verlet, Plummer) "
                  "per candidate:"]
        for b, sizes, row, best in split_report:
            lines.append("#   bucket %d n=%s: %s -> %d" % (
                b, sizes, " ".join(f"{k}:{t:.3f}" for k, t in row), best))
    else:
        lines += read_comments(current, "source-range parts")
    if variant_report is not None:
        lines += ["# force-kernel variant (3 * w + shape) by precision and bucket",
                  "#   written by nbody_amd/autotune.py on " + day,
                  "#   ms/step per variant (velocity Verlet, Plummer; buckets below "
                  f"{BATCH_BELOW} as a batch);",
                  "#   one uniform variant per precision, a bucket's own best where "
                  f">= {round(100 * OVERRIDE_GAIN)} % faster;",
                  "#   buckets outside the timed range take the uniform variant:"]
        for prec, b, row, best in variant_report:
            lines.append("#   %s bucket %d: %s -> %d" % (
                prec.name, b, " ".join(f"{v}:{t:.4f}" for v, t in row), best))
    else:
        lines += read_comments(current, "force-kernel variant")
    for b in range(len(split)):
        lines.append(f"split {b} {split[b]}")
    for prec in VARIANT_MODES:
        for b, v in enumerate(variants[prec]):
            if v:
                lines.append(f"variant {int(prec)} {b} {v}")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(PKG, "tuning", "gfx950.txt"))
    ap.add_argument("--stage", choices=["split", "variant", "all"], default="all")
    ap.add_argument("--max-bucket", type=int, default=20)
    ap.add_argument("--min-bucket", type=int, default=6)
    a = ap.parse_args()
    say = lambda m: print(m, flush=True)  # noqa: E731
    split, split_report = nb.tuning_split(), None
    if a.stage in ("split", "all"):
        split, split_report = tune(a.max_bucket, a.min_bucket, log=say)
    variants = {p: nb.tuning_variant(p) for p in VARIANT_MODES}
    variant_report = None
    if a.stage in ("variant", "all"):
        variants, variant_report = tune_variants(a.max_bucket, a.min_bucket, log=say)
    write_table(a.out, split, split_report, variants, variant_report)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
