"""Benchmark: pair-interactions/s of the direct-sum gravity + integrator step.

Default workload (one GPU): a single N=65536 Plummer sphere (seed 2, scale
radius 1, equal masses 1/N, G=1), softening 0.005, dt=5e-4, velocity Verlet
with the acceleration carried between steps (one force evaluation per step).
The headline `value` is an f32-compensated form (--f32: f32r, per-wave
relative coordinates, the default for one large system; or f32c, one f64
difference per pair, the default for the batched and served mixes of small
systems); the other f32 form and the plain f64 form are timed in the same run
and reported beside it.  A "step" is one velocity-Verlet step
of the whole system; interactions counted as N(N-1) per force evaluation and
20 flops per interaction.  Inputs are resident in HBM before timing starts;
the timed region is bracketed by HIP events on the library's stream.

Other workloads: --workload batched (512 star systems, N log-uniform in
[64, 8192], Euler/VV/RK4 1:1:1, dt=1e-3), --workload large (two colliding
Plummer spheres, N=1,048,576, RK4), --workload serve (the request server
under Poisson arrivals) and --workload bh (Barnes-Hut, theta 0.5, on the
N=1,048,576 system).

cpu_baseline: the oracle's threaded direct sum (the reference's
newton-parallel restated in C) on this host, full N, repeated for ~10 s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "n-body-sim_amd")]

import nbody_amd as nb  # noqa: E402

PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}  # MI355X vector peaks (datasheet)
FLOPS_PER_INTERACTION = 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_err(a, ref):
    a, ref = np.stack(a, 1), np.stack(ref, 1)
    e = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(e.max()), float(np.median(e))


def cpu_baseline(b, g, eps, seconds):
    """Oracle newton-parallel on the host: full-N evaluations for ~seconds."""
    import oracle
    threads = min(16, oracle.default_threads())
    n = len(b)
    t0 = time.perf_counter()
    reps = 0
    acc = None
    while True:
        acc = oracle.newton_parallel(b.masses, b.pos_x, b.pos_y, b.pos_z, g, eps, threads)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    rate = n * (n - 1) * reps / dt
    # ms per force evaluation at other sizes: median of 5 after one warm-up
    # (seeded Plummer spheres); N = 1M extrapolated from the N above
    per_eval = {}
    for m in (1024, 8192):
        p = nb.plummer(m, 2)
        ts = []
        for _ in range(6):
            t1 = time.perf_counter()
            oracle.newton_parallel(p.masses, p.pos_x, p.pos_y, p.pos_z, g, eps, threads)
            ts.append(time.perf_counter() - t1)
        per_eval[str(m)] = 1e3 * float(np.median(ts[1:]))
    per_eval[str(n)] = 1e3 * dt / reps
    per_eval["1048576 (extrapolated)"] = 1e3 * (1 << 20) * ((1 << 20) - 1) / rate
    return {"value": rate, "unit": "pair-interactions/s", "cores": threads,
            "kind": "port",
            "sample": f"{reps} full force evaluation(s) of the same N={n} system "
                      f"(oracle newton-parallel, {threads} threads, {dt:.1f} s)",
            "ms_per_eval": per_eval}, acc


def pair_scale(b, g, eps, idx):
    """S_i = sum_j |a_ij| for the listed targets (numpy f64): the size of the
    terms summed into a_i, against which an error shows the conditioning."""
    pos = np.stack([b.pos_x, b.pos_y, b.pos_z], 1)
    out = np.empty(len(idx))
    for k in range(0, len(idx), 32):
        sel = idx[k:k + 32]
        d = pos[None, :, :] - pos[sel][:, None, :]
        r2 = (d * d).sum(2)
        w = g * b.masses[None, :] * np.sqrt(r2) / (r2 + eps * eps) ** 1.5
        w[np.arange(len(sel)), sel] = 0.0
        out[k:k + 32] = w.sum(1)
    return out


def time_single(gpu, b, prec, g, eps, dt, steps, warmup):
    grav = nb.GpuNewtonParallelGravity(g, eps * eps, prec)
    integ = nb.GpuVelocityVerletIntegrator(dt)
    db = nb.DeviceBodies(gpu, b)
    acc0 = grav.calculate_accelerations(gpu, db)
    e0 = energy(gpu, db, g, eps)
    integ.step(gpu, db, grav, warmup)
    gpu.synchronize()
    gpu.record(0)
    integ.step(gpu, db, grav, steps)
    gpu.record(1)
    ms = gpu.elapsed_ms(0, 1)
    e1 = energy(gpu, db, g, eps)
    final = db.download()
    db.close()
    return ms, acc0, e0, e1, final


def energy(gpu, db, g, eps):
    out = np.zeros(10)
    nb._check(nb.load_library().nbody_energy_momentum(gpu.handle, db.handle, g,
                                                      eps * eps, nb._p(out)))
    return out


def measured_traffic():
    """HBM bytes per force launch of the N=65536 workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (counters scaled by a copy-kernel
    calibration).  None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "force_single_n65536_pmc.md")
    try:
        for line in open(path):
            if line.startswith("machine-readable: traffic_bytes_per_launch"):
                kv = dict(t.split("=") for t in line.split()[2:])
                return {k: int(v) for k, v in kv.items()}
    except OSError:
        pass
    return None


def roofline(rate, dtype, kernel_ms, traffic=None):
    """VALU roofline of the force kernel: achieved = interactions/s x 20 flops
    (the counting convention of the reference's published figures) against
    MI355X's vector peak for the arithmetic type."""
    ach = rate * FLOPS_PER_INTERACTION / 1e12
    return {"bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS[dtype],
            "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS[dtype], "traffic": traffic,
            "kernel_avg_ms": kernel_ms}


F32_FORMS = {"f32r": nb.Precision.F32R, "f32c": nb.Precision.F32C}
DTYPE = {
    "f32r": "f32 pair terms (per-wave relative coordinates; f64 differences "
            "split in two f32 words for near chunks), f64 state",
    "f32c": "f32 pair terms (f64 differences), f64 state",
}


def default_f32(workload):
    """The faster f32 form for a workload: many small systems (a few chunks
    each, most of them near their targets) favour the per-pair f64 difference
    of f32c; one large system favours f32r."""
    return "f32c" if workload in ("batched", "serve") else "f32r"


def f32_forms(args):
    """The headline f32 form first, then the other one."""
    other = "f32c" if args.f32 == "f32r" else "f32r"
    return [args.f32, other]


def run_single(args, gpu):
    n, g, eps, dt = 65536, 1.0, 0.005, 5e-4
    b = nb.plummer(n, 2)
    inter = n * (n - 1) * args.steps
    forms = f32_forms(args) + ["f64"]
    precs = dict(F32_FORMS, f64=nb.Precision.F64)
    res = {}
    for name in forms:
        prec = precs[name]
        log(f"timing {prec.name}: N={n} VV {args.steps} steps (+{args.warmup} warmup)")
        res[name] = time_single(gpu, b, prec, g, eps, dt, args.steps, args.warmup)
        log(f"  {res[name][0]:.1f} ms -> {inter / (res[name][0] / 1e3):.4e} int/s")
    traffic = measured_traffic() or {}
    head, alt = forms[0], forms[1]
    rate = {k: inter / (res[k][0] / 1e3) for k in forms}
    ms = {k: res[k][0] for k in forms}
    cpu = None
    acc_ref = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle newton-parallel at N={n} for ~{args.cpu_seconds} s")
        cpu, acc_ref = cpu_baseline(b, g, eps, args.cpu_seconds)
    acc_rep = {}
    drift = lambda e0, e1: abs(e1[2] - e0[2]) / abs(e0[2])  # noqa: E731
    # momentum drift |P(t) - P(0)| / sum_i m_i |v_i(0)|
    pscale = float(np.sum(b.masses * np.sqrt(b.vel_x ** 2 + b.vel_y ** 2 + b.vel_z ** 2)))
    pdrift = lambda e0, e1: float(np.linalg.norm(e1[3:6] - e0[3:6])) / pscale  # noqa: E731
    fin64 = res["f64"][4]
    if acc_ref is not None:
        # normalised error |da_i| / sum_j |a_ij| on a seeded target sample
        sample = np.sort(np.random.default_rng(5).choice(n, 2048, replace=False))
        scale = pair_scale(b, g, eps, sample)
        ref_s = np.stack(acc_ref, 1)[sample]
    for k in forms:
        _, acc0, e0, e1, fin = res[k]
        if acc_ref is not None:
            acc_rep[f"{k}_max_rel"], acc_rep[f"{k}_median_rel"] = rel_err(acc0, acc_ref)
            dn = np.linalg.norm(np.stack(acc0, 1)[sample] - ref_s, axis=1) / scale
            acc_rep[f"{k}_max_norm_sampled"] = float(dn.max())
            acc_rep[f"{k}_median_norm_sampled"] = float(np.median(dn))
        acc_rep[f"{k}_energy_drift"] = drift(e0, e1)
        acc_rep[f"{k}_momentum_drift"] = pdrift(e0, e1)
        if k != "f64":
            # trajectories after warmup + steps against f64, in units of the
            # Plummer scale radius (1)
            dev = np.sqrt((fin.pos_x - fin64.pos_x) ** 2 + (fin.pos_y - fin64.pos_y) ** 2 +
                          (fin.pos_z - fin64.pos_z) ** 2)
            acc_rep[f"trajectory_max_dev_{k}_vs_f64"] = float(dev.max())
            acc_rep[f"trajectory_rms_dev_{k}_vs_f64"] = float(np.sqrt(np.mean(dev ** 2)))
    acc_rep["trajectory_steps"] = args.warmup + args.steps
    return {
        "metric": "pair-interactions/s (f32-compensated forces, f64 state)",
        "value": rate[head], "unit": "pair-interactions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms[head] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[head],
        "data": "synthetic (seeded Plummer sphere, generated on the host)",
        "config": {"workload": "single Plummer sphere N=65536, softening 0.005, "
                               "dt 5e-4, velocity-verlet", "n": n, "g": g,
                   "softening": eps, "dt": dt, "integrator": "velocity-verlet",
                   "f32_form": head, "global_batch": 1,
                   "parallelism": "none (1 GPU)"},
        "roofline": roofline(rate[head], "f32", ms[head] / args.steps,
                             traffic.get(head)),
        "cpu_baseline": cpu,
        "f32_alt": {"form": alt, "value": rate[alt], "dtype": DTYPE[alt],
                    "ms_per_step": ms[alt] / args.steps,
                    "roofline": roofline(rate[alt], "f32", ms[alt] / args.steps,
                                         traffic.get(alt))},
        "fp64_plain": {"value": rate["f64"], "ms_per_step": ms["f64"] / args.steps,
                       "roofline": roofline(rate["f64"], "f64", ms["f64"] / args.steps,
                                            traffic.get("f64")),
                       "f32_speedup": rate[head] / rate["f64"]},
        "accuracy": acc_rep,
    }


def run_batched(args, gpu):
    rng = np.random.default_rng(3)
    nsys = 512
    ns = np.exp(rng.uniform(np.log(64), np.log(8192), nsys)).astype(np.int64)
    systems = [nb.star_system(int(n), 10_000 + s) for s, n in enumerate(ns)]
    integ = [s % 3 for s in range(nsys)]
    evals = np.array([1, 1, 4])[integ]
    inter = float(np.sum(ns * (ns - 1) * evals)) * args.steps
    forms = f32_forms(args) + ["f64"]
    precs = dict(F32_FORMS, f64=nb.Precision.F64)
    rate = {}
    for name in forms:
        prec = precs[name]
        bt = nb.Batch(gpu, systems, integ, 1e-3, 1.0, 1e-6, prec)
        bt.step(args.warmup)
        gpu.synchronize()
        gpu.record(0)
        bt.step(args.steps)
        gpu.record(1)
        ms = gpu.elapsed_ms(0, 1)
        rate[name] = inter / (ms / 1e3)
        log(f"batched {prec.name}: {ms:.1f} ms -> {rate[name]:.4e} int/s")
        bt.close()
    head, alt = forms[0], forms[1]
    ms_head = inter / rate[head] * 1e3
    return {
        "metric": "pair-interactions/s (batched, f32-compensated forces)",
        "value": rate[head], "unit": "pair-interactions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_head / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[head],
        "data": "synthetic (seeded star systems)",
        "config": {"workload": "512 star systems, N log-uniform [64, 8192], "
                               "euler/vv/rk4 1:1:1, dt 1e-3", "systems": nsys,
                   "bodies": int(ns.sum()), "f32_form": head, "global_batch": nsys,
                   "parallelism": "none (1 GPU)"},
        "roofline": roofline(rate[head], "f32", ms_head / args.steps),
        "cpu_baseline": None,
        "f32_alt": {"form": alt, "value": rate[alt]},
        "fp64_plain": {"value": rate["f64"], "f32_speedup": rate[head] / rate["f64"]},
    }


def run_serve(args, gpu):
    """Poisson request stream against the continuous-batching server."""
    import threading
    rng = np.random.default_rng(11)
    rates = [float(r) for r in args.rates.split(",")]
    nreq = args.requests
    # pre-generate the request stream (host side, not timed)
    kinds = args.mix.split(",")
    make = {"plummer": nb.plummer, "star": nb.star_system, "disc": nb.disc_galaxy}
    reqs = []
    for k in range(nreq):
        n = int(np.exp(rng.uniform(np.log(32), np.log(16384))))
        kind = kinds[int(rng.random() * len(kinds))]
        b = make[kind](n, 50_000 + k)
        reqs.append(dict(b=b, integ=nb.IntegratorMethod(int(rng.integers(0, 3))),
                         dt=float(rng.choice([1e-3, 5e-4])),
                         steps=int(rng.integers(10, 1001))))
    results = []
    for rate in rates:
        srv = nb.Server(capacity_bodies=4 << 20, max_systems=4096, snapshot_every=10,
                        precision=F32_FORMS[args.f32])
        gaps = rng.exponential(1.0 / rate, nreq)
        lat, done, interactions = [], [0], 0.0
        t_first = [None]

        def consume():
            while done[0] < nreq:
                r = srv.poll(timeout_ms=2000, capacity=16384)
                if r is None:
                    continue
                ev, _ = r
                if ev["kind"] in (nb.EVENT_DONE, nb.EVENT_FAILED, nb.EVENT_CANCELLED):
                    if ev["kind"] != nb.EVENT_DONE:
                        raise RuntimeError(ev["message"])
                    lat.append(ev["t_event"] - ev["t_submit"])
                    done[0] += 1

        th = threading.Thread(target=consume)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        for k, q in enumerate(reqs):
            t_next += gaps[k]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(q["b"], q["integ"], q["dt"], q["steps"], eps2=1e-4,
                       want_snapshots=True)
        th.join(timeout=600)
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        lat_ms = np.array(lat) * 1e3
        res = {"rate_req_s": rate, "requests": nreq, "wall_s": wall,
               "solves_s": nreq / wall, "p50_ms": float(np.percentile(lat_ms, 50)),
               "p99_ms": float(np.percentile(lat_ms, 99)),
               "interactions_s": st["interactions"] / wall,
               "shared_steps": st["shared_steps"]}
        log(f"serve rate {rate}/s: {res}")
        results.append(res)
    best = max(results, key=lambda r: r["interactions_s"])
    return {
        "metric": "served solves/s (p50/p99 latency per rate in `serving`)",
        "value": best["solves_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.f32],
        "data": f"synthetic Poisson request stream ({args.mix.replace(',', ' / ')})",
        "config": {"workload": "served requests, N log-uniform [32, 16384], steps "
                               "uniform [10, 1000], random integrator, dt in "
                               "{1e-3, 5e-4}, snapshots every 10 steps",
                   "requests_per_rate": nreq, "rates": rates,
                   "parallelism": "none (1 GPU)"},
        "roofline": None, "cpu_baseline": None, "serving": results,
    }


def run_large(args, gpu):
    """Two colliding Plummer spheres, N=1,048,576, RK4 (4 evaluations/step)."""
    n, g, eps, dt = 1 << 20, 1.0, 0.002, 2.5e-4
    b = nb.colliding(n, 5, separation=10.0, rel_speed=0.5)
    forms = f32_forms(args) + ["f64"]
    precs = dict(F32_FORMS, f64=nb.Precision.F64)
    # sampled-target oracle: 4096 random targets against all sources (f64 CPU)
    acc_rep = {}
    if not args.no_cpu_baseline:
        import oracle
        idx = np.sort(np.random.default_rng(9).choice(n, 4096, replace=False))
        t0 = time.perf_counter()
        ref = oracle.newton_targets(b.masses, b.pos_x, b.pos_y, b.pos_z, g, eps, idx,
                                    min(16, oracle.default_threads()))
        log(f"sampled oracle: 4096 targets in {time.perf_counter() - t0:.1f} s")
        for name in forms:
            db = nb.DeviceBodies(gpu, b)
            got = nb.GpuNewtonParallelGravity(g, eps * eps, precs[name]) \
                .calculate_accelerations(gpu, db)
            db.close()
            mx, med = rel_err(tuple(a[idx] for a in got), ref)
            acc_rep[f"{name}_max_rel_sampled"] = mx
            acc_rep[f"{name}_median_rel_sampled"] = med
    inter = n * (n - 1) * 4 * args.steps
    rate, ms = {}, {}
    for name in forms:
        grav = nb.GpuNewtonParallelGravity(g, eps * eps, precs[name])
        integ = nb.GpuRungeKuttaIntegrator(dt)
        db = nb.DeviceBodies(gpu, b)
        e0 = energy(gpu, db, g, eps)
        gpu.synchronize()
        gpu.record(0)
        integ.step(gpu, db, grav, args.steps)
        gpu.record(1)
        ms[name] = gpu.elapsed_ms(0, 1)
        e1 = energy(gpu, db, g, eps)
        rate[name] = inter / (ms[name] / 1e3)
        acc_rep[f"{name}_energy_drift"] = abs(e1[2] - e0[2]) / abs(e0[2])
        log(f"large {name}: {ms[name]:.0f} ms for {args.steps} RK4 steps, "
            f"energy drift {acc_rep[f'{name}_energy_drift']:.3e}")
        db.close()
    head, alt = forms[0], forms[1]
    return {
        "metric": "pair-interactions/s (f32-compensated forces, f64 state)",
        "value": rate[head], "unit": "pair-interactions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 0, "ms_per_step": ms[head] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[head],
        "data": "synthetic (two seeded Plummer spheres)",
        "config": {"workload": "two colliding Plummer spheres N=1048576, softening "
                               "0.002, dt 2.5e-4, runge-kutta", "n": n,
                   "f32_form": head, "parallelism": "none (1 GPU)"},
        "roofline": roofline(rate[head], "f32", ms[head] / args.steps / 4),
        "cpu_baseline": None,
        "f32_alt": {"form": alt, "value": rate[alt],
                    "roofline": roofline(rate[alt], "f32", ms[alt] / args.steps / 4)},
        "fp64_plain": {"value": rate["f64"], "f32_speedup": rate[head] / rate["f64"],
                       "roofline": roofline(rate["f64"], "f64", ms["f64"] / args.steps / 4)},
        "accuracy": acc_rep,
    }


def run_bh(args, gpu):
    """Barnes-Hut, theta 0.5, on the N=1,048,576 colliding spheres: the tree
    is rebuilt for every evaluation, as the reference does.  Timed: fast-f64
    RK4 steps (4 evaluations each) and single strict evaluations; beside
    them one f64 direct-sum evaluation of the same system for the speedup.
    Parity at full size: strict forces against the oracle's Barnes-Hut on
    4096 sampled targets, bit for bit."""
    n, g, eps, dt, theta = 1 << 20, 1.0, 0.002, 2.5e-4, 0.5
    b = nb.colliding(n, 5, separation=10.0, rel_speed=0.5)
    strict = nb.GpuBarnesHutGravity(g, eps * eps, theta, nb.Precision.F64_STRICT)
    fast = nb.GpuBarnesHutGravity(g, eps * eps, theta, nb.Precision.F64)
    db = nb.DeviceBodies(gpu, b)
    acc_strict = strict.calculate_accelerations(gpu, db)      # warm-up + parity
    nodes = strict.tree_nodes(db)
    times = {}
    for name, grav in (("strict", strict), ("fast", fast)):
        grav.calculate_accelerations(gpu, db, download=False)
        gpu.synchronize()
        gpu.record(0)
        for _ in range(3):
            grav.calculate_accelerations(gpu, db, download=False)
        gpu.record(1)
        times[name] = gpu.elapsed_ms(0, 1) / 3
        log(f"bh {name}: {times[name]:.2f} ms per evaluation")
    acc_fast = fast.calculate_accelerations(gpu, db)
    direct = nb.GpuNewtonParallelGravity(g, eps * eps, nb.Precision.F64)
    gpu.record(0)
    acc_direct = direct.calculate_accelerations(gpu, db)
    gpu.record(1)
    direct_ms = gpu.elapsed_ms(0, 1)
    db.close()
    # RK4 steps with fast Barnes-Hut forces
    db = nb.DeviceBodies(gpu, b)
    rk = nb.GpuRungeKuttaIntegrator(dt)
    e0 = energy(gpu, db, g, eps)
    rk.step(gpu, db, fast, 1)
    gpu.synchronize()
    gpu.record(0)
    rk.step(gpu, db, fast, args.steps)
    gpu.record(1)
    step_ms = gpu.elapsed_ms(0, 1) / args.steps
    e1 = energy(gpu, db, g, eps)
    db.close()
    acc = {"fast_vs_direct_max_rel": None, "fast_vs_direct_median_rel": None}
    mx, med = rel_err(acc_fast, acc_direct)
    acc.update(fast_vs_direct_max_rel=mx, fast_vs_direct_median_rel=med,
               strict_vs_direct_median_rel=rel_err(acc_strict, acc_direct)[1],
               energy_drift_rk4=abs(e1[2] - e0[2]) / abs(e0[2]))
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        idx = np.sort(np.random.default_rng(9).choice(n, 4096, replace=False))
        threads = min(16, oracle.default_threads())
        t0 = time.perf_counter()
        ref = oracle.barnes_hut_targets(b.masses, b.pos_x, b.pos_y, b.pos_z, g, eps,
                                        theta, idx, threads)
        t_sample = time.perf_counter() - t0
        same = int(sum(np.array_equal(np.array([a[i] for a in acc_strict]),
                                      np.array([r[q] for r in ref]))
                       for q, i in enumerate(idx)))
        acc["strict_bitwise_sampled"] = f"{same}/4096"
        mx_f, _ = rel_err(tuple(a[idx] for a in acc_fast), ref)
        acc["fast_vs_oracle_max_rel_sampled"] = mx_f
        # CPU baseline: the oracle's full evaluation (tree + every target)
        t0 = time.perf_counter()
        oracle.barnes_hut(b.masses, b.pos_x, b.pos_y, b.pos_z, g, eps, theta, threads)
        t_full = time.perf_counter() - t0
        log(f"oracle barnes-hut: sample {t_sample:.1f} s, full {t_full:.1f} s")
        cpu = {"value": 1.0 / t_full, "unit": "evaluations/s", "cores": threads,
               "kind": "port",
               "sample": f"one full evaluation of the same N={n} system (oracle "
                         f"barnes-hut: tree build + all targets, {threads} threads, "
                         f"{t_full:.1f} s)"}
    evals_per_s = 1e3 / times["fast"]
    return {
        "metric": "barnes-hut force evaluations/s (theta 0.5, tree rebuilt per evaluation)",
        "value": evals_per_s, "unit": "evaluations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": step_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (two seeded Plummer spheres)",
        "config": {"workload": "barnes-hut theta 0.5 on two colliding Plummer spheres "
                               "N=1048576, softening 0.002; RK4 dt 2.5e-4", "n": n,
                   "parallelism": "none (1 GPU)"},
        "roofline": None,
        "roofline_note": "tree walk: latency-bound pointer chasing through "
                         "scalar-cached node data; no VALU or HBM roofline applies",
        "cpu_baseline": cpu,
        "bh": {"fast_ms_per_eval": times["fast"], "strict_ms_per_eval": times["strict"],
               "direct_f64_ms_per_eval": direct_ms,
               "speedup_vs_direct_f64": direct_ms / times["fast"],
               "direct_equivalent_interactions_per_s": n * (n - 1) / (times["fast"] / 1e3),
               "tree_nodes": nodes},
        "accuracy": acc,
    }


def run_copy(args, gpu):
    """Counter calibration: one device-to-device copy of exactly 256 MiB
    (HIP runtime blit).  Under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE its
    kernel must read and write 262144 KiB; the ratio measured/expected scales
    the counters of the force kernels into HBM bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # device to device
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return {"metric": "copy calibration", "value": nbytes, "unit": "bytes",
            "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u8", "data": "synthetic", "config": {"workload": "copy 256 MiB"},
            "roofline": None, "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload",
                    choices=["single", "batched", "serve", "large", "bh", "copy"],
                    default="single")
    ap.add_argument("--rates", default="20,50,80,160,640",
                    help="serve: request rates (1/s); the last ones saturate the GPU")
    ap.add_argument("--requests", type=int, default=200, help="serve: requests per rate")
    ap.add_argument("--mix", default="plummer,star",
                    help="serve: request scenarios, from plummer, star, disc")
    ap.add_argument("--f32", choices=sorted(F32_FORMS), default=None,
                    help="the f32 form of the headline value (the other one is "
                         "timed beside it); default: the faster one for the "
                         "workload (f32r for one large system, f32c for many "
                         "small ones)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(independent systems shard by system with no communication; a single "
              "system would shard by target range with one all-gather of source "
              "positions per force evaluation -- not built here)", file=sys.stderr)
        sys.exit(2)
    if args.f32 is None:
        args.f32 = default_f32(args.workload)
    if args.steps is None:
        args.steps = {"single": 200, "batched": 50, "serve": 0, "large": 2,
                      "bh": 5, "copy": 1}[args.workload]
    gpu = nb.GpuContext(0)
    name, cus = gpu.device_info()
    log(f"device {name}, {cus} CUs")
    rec = {"single": run_single, "batched": run_batched, "serve": run_serve,
           "large": run_large, "bh": run_bh, "copy": run_copy}[args.workload](args, gpu)
    rec["device"] = {"arch": name, "compute_units": cus}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
