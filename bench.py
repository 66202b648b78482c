"""Benchmark: pair-interactions/s of the direct-sum gravity + integrator step.

The default run (`python bench.py`, one GPU) prints ONE JSON line: the
single-system workload below as the headline record, with the batched, served
and large workloads as sub-records (`batched`, `serving`, `large`) of the same
line.  Every workload is timed in an f32-compensated form (the headline) and
in the plain f64 form (`fp64_plain`, with the f32/f64 ratio), and carries its
accuracy against the f64 oracle.

* single (headline `value`): one N=65536 Plummer sphere (seed 2, scale radius
  1, equal masses 1/N, G=1), softening 0.005, dt=5e-4, velocity Verlet with
  the acceleration carried between steps (one force evaluation per step),
  `--steps` timed steps after `--warmup`.  Form f32r (per-wave relative
  coordinates); f32c is timed beside it.
* batched: 512 star systems, N log-uniform in [64, 8192], Euler/VV/RK4
  1:1:1, dt=1e-3, softening 1e-3, 500 steps (`--batched-steps`), graph
  stepping.
* serving: the request server under Poisson arrivals at several rates
  (`--rates`, 200 requests each): N log-uniform in [32, 16384], 10-1000
  steps, random integrator, dt in {1e-3, 5e-4}, snapshots every 10 steps;
  solves/s, p50/p99 latency and interactions/s per rate, and the saturated
  rate again with the f64 server.
* large: two colliding Plummer spheres, N=1,048,576, softening 0.002, RK4,
  dt=2.5e-4, 20 steps (`--large-steps`; the second f32 form runs 5).

A "step" of the headline is one velocity-Verlet step of the whole system;
interactions are N(N-1) per force evaluation, 20 flops each.  Inputs are
resident in HBM before timing starts; each timed region is bracketed by HIP
events on the library's stream.  Before each single-system form (and before
any other workload run on its own) the GPU runs `--clock-warm` seconds (0.5)
of untimed force evaluations of the same kind on a throwaway copy of the
system: the clock drops while the host works between measurements and
ramps up again over tens of milliseconds, and a 20-step timed region would
otherwise average over that ramp.  `roofline` is the force kernel's: its
launches inside the timed region are bracketed by HIP events, and achieved
FLOP/s = 20 x interactions per launch / mean launch time.

cpu_baseline: the oracle's threaded direct sum (the reference's
newton-parallel restated in C, bit-identical per target) on this host, full
N, repeated for about --cpu-seconds.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "n-body-sim_amd")]

import nbody_amd as nb  # noqa: E402

PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}  # MI355X vector peaks (datasheet)
HBM_PEAK_GBS = 8000.0                        # MI355X HBM3E (datasheet)
FLOPS_PER_INTERACTION = 20
# one GPU's share of a box's host cores: worker pools are capped at this
HOST_CPU_SHARE = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_err_vec(a, ref):
    a, ref = np.stack(a, 1), np.stack(ref, 1)
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


def rel_err(a, ref):
    e = rel_err_vec(a, ref)
    return float(e.max()), float(np.median(e))


def host_cpus():
    """Threads for the CPU baseline (the reference's rayon pool spans every
    logical CPU the process may run on, src/gravity/cpu/newton_parallel.rs:
    44-51) capped at one GPU's share of the box, and the counts seen."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        aff = os.cpu_count() or 1
    return min(aff, HOST_CPU_SHARE), aff, os.cpu_count()


def cpu_baseline(b, g, eps, seconds):
    """Oracle newton-parallel on the host: full-N evaluations for ~seconds."""
    import oracle
    threads, aff, logical = host_cpus()
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
            "host_logical_cpus": logical, "affinity_cpus": aff,
            "sample": f"{reps} full force evaluation(s) of the same N={n} system "
                      f"(oracle newton-parallel, {threads} threads = the "
                      f"{aff} CPUs this process may use, capped at one GPU's "
                      f"share of {HOST_CPU_SHARE}; {dt:.1f} s)",
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


def roofline(dtype, interactions, kt, step_rate=None, traffic=None):
    """VALU roofline of the force kernel.  kt: launch times of the timed
    region (GpuContext.kernel_times()); interactions: pair interactions those
    force launches evaluated.  achieved = 20 x interactions / force-launch
    time (the counting convention of the reference's published figures)
    against MI355X's vector peak for the arithmetic type."""
    force_s = kt["force_ms"] / 1e3
    ach = interactions * FLOPS_PER_INTERACTION / force_s / 1e12
    r = {"bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS[dtype],
         "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS[dtype], "traffic": traffic,
         "basis": "force-kernel launches bracketed by HIP events in the timed region",
         "force_kernel_avg_ms": kt["force_ms"] / max(kt["force_launches"], 1),
         "force_launches": kt["force_launches"],
         "combine_kernel_avg_ms": (kt["combine_ms"] / kt["combine_launches"]
                                   if kt["combine_launches"] else 0.0),
         "combine_launches": kt["combine_launches"]}
    if step_rate is not None:
        r["frac_of_whole_step"] = step_rate * FLOPS_PER_INTERACTION / 1e12 / PEAK_TFLOPS[dtype]
    if traffic:
        # measured HBM bytes per launch over the live launch time, against
        # MI355X's 8 TB/s
        gbs = traffic / (r["force_kernel_avg_ms"] / 1e3) / 1e9
        r["hbm_achieved_gbs"] = gbs
        r["hbm_frac"] = gbs / HBM_PEAK_GBS
    return r


F32_FORMS = {"f32r": nb.Precision.F32R, "f32c": nb.Precision.F32C}
PRECS = dict(F32_FORMS, f64=nb.Precision.F64)
DTYPE = {
    "f32r": "f32 pair terms (per-wave relative coordinates; r^2 of far chunks "
            "from bf16-split Gram products on the matrix cores; f64 differences "
            "split in two f32 words for near chunks), f64 state",
    "f32c": "f32 pair terms (f64 differences), f64 state",
    "f64": "f64",
}


def other_f32(form):
    return "f32c" if form == "f32r" else "f32r"


def warm_clocks(gpu, seconds, prec=None, bodies=None):
    """Untimed force evaluations for about `seconds` right before a timed
    region, on a throwaway copy of the system (`bodies`, else a seeded
    N=65536 sphere) in the precision about to be timed.  The GPU clock drops
    while the host works between measurements (allocation, energies,
    accuracy against the oracle) and ramps up again over tens of
    milliseconds of sustained load; without this a short timed region would
    average over that ramp rather than measure the sustained rate."""
    if seconds <= 0:
        return
    grav = nb.GpuNewtonParallelGravity(1.0, 0.005 ** 2,
                                       prec if prec is not None else nb.Precision.F32C)
    db = nb.DeviceBodies(gpu, bodies if bodies is not None else nb.plummer(65536, 99))
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end:
        grav.calculate_accelerations(gpu, db, download=False)
        gpu.synchronize()
    db.close()


# ---- config 2: one N=65536 Plummer sphere, velocity Verlet -------------------
def time_single(gpu, b, prec, g, eps, dt, steps, warmup, clock_warm):
    grav = nb.GpuNewtonParallelGravity(g, eps * eps, prec)
    integ = nb.GpuVelocityVerletIntegrator(dt)
    db = nb.DeviceBodies(gpu, b)
    acc0 = grav.calculate_accelerations(gpu, db)
    e0 = energy(gpu, db, g, eps)
    warm_clocks(gpu, clock_warm, prec, b)
    integ.step(gpu, db, grav, warmup)
    gpu.synchronize()
    gpu.kernel_timing(True)
    gpu.record(0)
    integ.step(gpu, db, grav, steps)
    gpu.record(1)
    ms = gpu.elapsed_ms(0, 1)
    kt = gpu.kernel_times()
    gpu.kernel_timing(False)
    e1 = energy(gpu, db, g, eps)
    final = db.download()
    db.close()
    return ms, kt, acc0, e0, e1, final


def run_single(args, gpu):
    n, g, eps, dt = 65536, 1.0, 0.005, 5e-4
    b = nb.plummer(n, 2)
    inter_eval = n * (n - 1)
    inter = inter_eval * args.steps
    head = args.f32 or "f32r"
    alt = other_f32(head)
    forms = [head, alt, "f64"]
    res = {}
    for name in forms:
        log(f"single {name}: N={n} VV {args.steps} steps (+{args.warmup} warmup)")
        res[name] = time_single(gpu, b, PRECS[name], g, eps, dt, args.steps, args.warmup,
                                args.clock_warm)
        log(f"  {res[name][0]:.1f} ms -> {inter / (res[name][0] / 1e3):.4e} int/s; "
            f"force kernel {res[name][1]['force_ms'] / max(res[name][1]['force_launches'], 1):.3f} ms")
    traffic = measured_traffic() or {}
    rate = {k: inter / (res[k][0] / 1e3) for k in forms}
    ms = {k: res[k][0] for k in forms}

    def roof(k):
        kt = res[k][1]
        return roofline("f64" if k == "f64" else "f32", inter_eval * kt["force_launches"],
                        kt, rate[k], traffic.get(k))

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
    fin64 = res["f64"][5]
    if acc_ref is not None:
        # normalised error |da_i| / sum_j |a_ij| on a seeded target sample
        sample = np.sort(np.random.default_rng(5).choice(n, 2048, replace=False))
        scale = pair_scale(b, g, eps, sample)
        ref_s = np.stack(acc_ref, 1)[sample]
    for k in forms:
        _, _, acc0, e0, e1, fin = res[k]
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
    acc_rep["reference"] = "oracle newton-parallel (f64, bit-identical to the reference's newton)"
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
        "roofline": roof(head),
        "cpu_baseline": cpu,
        "f32_alt": {"form": alt, "value": rate[alt], "dtype": DTYPE[alt],
                    "ms_per_step": ms[alt] / args.steps, "roofline": roof(alt)},
        "fp64_plain": {"value": rate["f64"], "ms_per_step": ms["f64"] / args.steps,
                       "roofline": roof("f64"), "f32_speedup": rate[head] / rate["f64"]},
        "accuracy": acc_rep,
    }


# ---- config 3: 512 star systems, Euler / VV / RK4 ---------------------------
BATCH_EPS = 1e-3


def batched_systems():
    rng = np.random.default_rng(3)
    nsys = 512
    ns = np.exp(rng.uniform(np.log(64), np.log(8192), nsys)).astype(np.int64)
    systems = [nb.star_system(int(n), 10_000 + s) for s, n in enumerate(ns)]
    integ = [s % 3 for s in range(nsys)]
    return ns, systems, integ


def batch_force_accuracy(gpu, systems, ns, forms):
    """Max / median relative acceleration error against the oracle on 12
    systems spanning the size range (the smallest, the largest and ten
    quantiles), each evaluated alone in every form."""
    import oracle
    order = np.argsort(ns)
    pick = sorted(set(order[np.linspace(0, len(ns) - 1, 12).astype(int)].tolist()))
    errs = {k: [] for k in forms}
    for s in pick:
        b = systems[s]
        ref = oracle.newton_parallel(b.masses, b.pos_x, b.pos_y, b.pos_z, 1.0, BATCH_EPS,
                                     host_cpus()[0])
        for k in forms:
            db = nb.DeviceBodies(gpu, b)
            got = nb.GpuNewtonParallelGravity(1.0, BATCH_EPS ** 2, PRECS[k]) \
                .calculate_accelerations(gpu, db)
            db.close()
            errs[k].append(rel_err_vec(got, ref))
    out = {}
    for k in forms:
        e = np.concatenate(errs[k])
        out[f"{k}_max_rel"] = float(e.max())
        out[f"{k}_median_rel"] = float(np.median(e))
    out["force_sample"] = (f"{len(pick)} systems (N {int(ns[pick].min())}..{int(ns[pick].max())}), "
                           "every body, against the oracle's f64 direct sum")
    return out


def run_batched(args, gpu):
    ns, systems, integ = batched_systems()
    nsys = len(ns)
    evals = np.array([1, 1, 4])[integ]
    inter_step = float(np.sum(ns * (ns - 1) * evals))
    steps = args.batched_steps
    inter = inter_step * steps
    head = args.f32_batched
    forms = [head, other_f32(head), "f64"]
    rate, ms, acc = {}, {}, {}
    roofs = {}
    for name in forms:
        bt = nb.Batch(gpu, systems, integ, 1e-3, 1.0, BATCH_EPS ** 2, PRECS[name])
        bt.step(args.warmup)
        gpu.synchronize()
        e0 = bt.energy()
        gpu.record(0)
        bt.step(steps)
        gpu.record(1)
        ms[name] = gpu.elapsed_ms(0, 1)
        rate[name] = inter / (ms[name] / 1e3)
        e1 = bt.energy()
        d = np.abs(e1[:, 2] - e0[:, 2]) / np.abs(e0[:, 2])
        acc[f"{name}_energy_drift_max"] = float(d.max())
        acc[f"{name}_energy_drift_median"] = float(np.median(d))
        # launch-level timing: a separate pass of the same steps without the
        # graph, every force and combine launch bracketed by events
        kt_steps = 20
        gpu.kernel_timing(True)
        bt.step(kt_steps)
        kt = gpu.kernel_times()
        gpu.kernel_timing(False)
        roofs[name] = roofline("f64" if name == "f64" else "f32", inter_step * kt_steps, kt,
                               rate[name])
        roofs[name]["basis"] += f" (a separate {kt_steps}-step pass without the graph)"
        log(f"batched {name}: {ms[name]:.1f} ms for {steps} steps -> {rate[name]:.4e} int/s, "
            f"energy drift max {d.max():.2e}; force launches {roofs[name]['achieved']:.1f} TFLOP/s")
        bt.close()
    acc.update(batch_force_accuracy(gpu, systems, ns, forms))
    acc["energy_drift_over"] = f"the {steps} timed steps, every system"
    alt = forms[1]
    return {
        "metric": "pair-interactions/s (batched, f32-compensated forces)",
        "value": rate[head], "unit": "pair-interactions/s",
        "steps": steps, "warmup": args.warmup, "ms_per_step": ms[head] / steps,
        "dtype": DTYPE[head],
        "config": {"workload": "512 star systems, N log-uniform [64, 8192], "
                               "euler/vv/rk4 1:1:1, dt 1e-3, softening 1e-3, "
                               "graph-captured steps", "systems": nsys,
                   "bodies": int(ns.sum()), "interactions_per_step": inter_step,
                   "f32_form": head},
        "roofline": roofs[head],
        "f32_alt": {"form": alt, "value": rate[alt], "roofline": roofs[alt]},
        "fp64_plain": {"value": rate["f64"], "ms_per_step": ms["f64"] / steps,
                       "roofline": roofs["f64"], "f32_speedup": rate[head] / rate["f64"]},
        "accuracy": acc,
    }


# ---- config 4: served requests ------------------------------------------------
SERVE_EPS = 1e-2


def serve_requests(nreq, mix):
    rng = np.random.default_rng(11)
    kinds = mix.split(",")
    make = {"plummer": nb.plummer, "star": nb.star_system, "disc": nb.disc_galaxy}
    reqs = []
    for k in range(nreq):
        n = int(np.exp(rng.uniform(np.log(32), np.log(16384))))
        kind = kinds[int(rng.random() * len(kinds))]
        b = make[kind](n, 50_000 + k)
        reqs.append(dict(b=b, kind=kind, integ=nb.IntegratorMethod(int(rng.integers(0, 3))),
                         dt=float(rng.choice([1e-3, 5e-4])),
                         steps=int(rng.integers(10, 1001))))
    return reqs


def serve_stream(reqs, rate, prec, keep=()):
    """Submits the request list at Poisson times of the given rate and waits
    for every DONE.  Returns the per-rate figures and the final states of the
    requests listed in `keep` (by index)."""
    srv = nb.Server(capacity_bodies=4 << 20, max_systems=4096, snapshot_every=10,
                    precision=prec)
    gaps = np.random.default_rng(int(rate * 1000) + 17).exponential(1.0 / rate, len(reqs))
    lat, done = [], [0]
    finals, ids = {}, {}
    errors = []
    lock = threading.Lock()

    def consume():
        while done[0] < len(reqs) and not errors:
            r = srv.poll(timeout_ms=2000, capacity=16384)
            if r is None:
                continue
            ev, data = r
            if ev["kind"] in (nb.EVENT_DONE, nb.EVENT_FAILED, nb.EVENT_CANCELLED):
                if ev["kind"] != nb.EVENT_DONE:
                    errors.append(ev["message"])
                    return
                lat.append(ev["t_event"] - ev["t_submit"])
                with lock:
                    k = ids.get(ev["id"])
                if k is not None:
                    finals[k] = data
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
        rid = srv.submit(q["b"], q["integ"], q["dt"], q["steps"], eps2=SERVE_EPS ** 2,
                         want_snapshots=True)
        if k in keep:
            with lock:
                ids[rid] = k
    th.join(timeout=600)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    if errors or th.is_alive():
        raise RuntimeError(f"served stream failed: {errors or 'timed out'}")
    lat_ms = np.array(lat) * 1e3
    res = {"rate_req_s": rate, "requests": len(reqs), "wall_s": wall,
           "solves_s": len(reqs) / wall, "p50_ms": float(np.percentile(lat_ms, 50)),
           "p99_ms": float(np.percentile(lat_ms, 99)),
           "interactions_s": st["interactions"] / wall,
           "shared_steps": st["shared_steps"]}
    return res, finals


def serve_accuracy(gpu, reqs, finals, forms):
    """Force error of sampled requests' initial states in each form against
    the oracle, and the relative energy drift of their served results."""
    import oracle
    acc = {}
    errs = {k: [] for k in forms}
    for k in sorted(finals):
        b = reqs[k]["b"]
        ref = oracle.newton_parallel(b.masses, b.pos_x, b.pos_y, b.pos_z, 1.0, SERVE_EPS,
                                     host_cpus()[0])
        for f in forms:
            db = nb.DeviceBodies(gpu, b)
            got = nb.GpuNewtonParallelGravity(1.0, SERVE_EPS ** 2, PRECS[f]) \
                .calculate_accelerations(gpu, db)
            db.close()
            errs[f].append(rel_err_vec(got, ref))
    for f in forms:
        e = np.concatenate(errs[f])
        acc[f"{f}_max_rel"] = float(e.max())
        acc[f"{f}_median_rel"] = float(np.median(e))
    drifts = []
    for k in sorted(finals):
        b = reqs[k]["b"]
        x, y, z, vx, vy, vz = finals[k]
        e = []
        for bb in (b, nb.Bodies(b.masses, x, y, z, vx, vy, vz)):
            db = nb.DeviceBodies(gpu, bb)
            e.append(energy(gpu, db, 1.0, SERVE_EPS))
            db.close()
        drifts.append(abs(e[1][2] - e[0][2]) / abs(e[0][2]))
    acc["energy_drift_max"] = float(np.max(drifts))
    acc["energy_drift_median"] = float(np.median(drifts))
    acc["sample"] = (f"{len(finals)} requests (N {min(len(reqs[k]['b']) for k in finals)}.."
                     f"{max(len(reqs[k]['b']) for k in finals)}), every body; energy drift "
                     "of the served result after the request's own steps")
    return acc


def run_serve(args, gpu):
    reqs = serve_requests(args.requests, args.mix)
    rates = [float(r) for r in args.rates.split(",")]
    head = args.f32_serve
    keep = set(range(0, len(reqs), max(1, len(reqs) // 12)))
    results, finals = [], {}
    for rate in rates:
        res, fin = serve_stream(reqs, rate, PRECS[head], keep if not finals else ())
        finals = finals or fin
        log(f"serve {head} rate {rate}/s: {res['solves_s']:.1f} solves/s, "
            f"p50 {res['p50_ms']:.0f} ms, p99 {res['p99_ms']:.0f} ms, "
            f"{res['interactions_s']:.3e} int/s")
        results.append(res)
    sat = max(rates)
    res64, _ = serve_stream(reqs, sat, PRECS["f64"])
    log(f"serve f64 rate {sat}/s: {res64['solves_s']:.1f} solves/s, "
        f"{res64['interactions_s']:.3e} int/s")
    top = max(results, key=lambda r: r["rate_req_s"])
    acc = serve_accuracy(gpu, reqs, finals, [head, "f64"])
    return {
        "metric": "served solves/s at saturation (p50/p99 latency per rate in `rates`)",
        "value": top["solves_s"], "unit": "solves/s", "dtype": DTYPE[head],
        "config": {"workload": "served requests, Poisson arrivals, N log-uniform "
                               "[32, 16384], steps uniform [10, 1000], random "
                               "integrator, dt in {1e-3, 5e-4}, softening 1e-2, "
                               "snapshots every 10 steps",
                   "requests_per_rate": len(reqs), "mix": args.mix, "f32_form": head},
        "rates": results,
        "fp64_plain": {"rate_req_s": sat, "solves_s": res64["solves_s"],
                       "p50_ms": res64["p50_ms"], "p99_ms": res64["p99_ms"],
                       "interactions_s": res64["interactions_s"],
                       "f32_speedup": top["interactions_s"] / res64["interactions_s"]},
        "roofline": {"bound": "valu", "achieved": top["interactions_s"] * FLOPS_PER_INTERACTION / 1e12,
                     "peak": PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
                     "frac": top["interactions_s"] * FLOPS_PER_INTERACTION / 1e12 / PEAK_TFLOPS["f32"],
                     "traffic": None,
                     "basis": "interactions evaluated over the wall time of the saturated "
                              "stream (submit of the first request to DONE of the last)"},
        "accuracy": acc,
    }


# ---- config 5: N = 1,048,576 colliding spheres, RK4 ---------------------------
def run_large(args, gpu):
    n, g, eps, dt = 1 << 20, 1.0, 0.002, 2.5e-4
    b = nb.colliding(n, 5, separation=10.0, rel_speed=0.5)
    head = args.f32 or "f32r"
    forms = [head, other_f32(head), "f64"]
    steps = {head: args.large_steps, forms[1]: max(1, args.large_steps // 4),
             "f64": args.large_steps}
    acc_rep = {}
    if not args.no_cpu_baseline:
        import oracle
        idx = np.sort(np.random.default_rng(9).choice(n, 4096, replace=False))
        t0 = time.perf_counter()
        ref = oracle.newton_targets(b.masses, b.pos_x, b.pos_y, b.pos_z, g, eps, idx,
                                    host_cpus()[0])
        log(f"large: sampled oracle, 4096 targets in {time.perf_counter() - t0:.1f} s")
        for name in forms:
            db = nb.DeviceBodies(gpu, b)
            got = nb.GpuNewtonParallelGravity(g, eps * eps, PRECS[name]) \
                .calculate_accelerations(gpu, db)
            db.close()
            mx, med = rel_err(tuple(a[idx] for a in got), ref)
            acc_rep[f"{name}_max_rel_sampled"] = mx
            acc_rep[f"{name}_median_rel_sampled"] = med
        acc_rep["sample"] = "4096 seeded targets against all sources (oracle, f64)"
    inter_eval = n * (n - 1)
    rate, ms, roofs = {}, {}, {}
    for name in forms:
        grav = nb.GpuNewtonParallelGravity(g, eps * eps, PRECS[name])
        integ = nb.GpuRungeKuttaIntegrator(dt)
        db = nb.DeviceBodies(gpu, b)
        e0 = energy(gpu, db, g, eps)
        gpu.synchronize()
        gpu.kernel_timing(True)
        gpu.record(0)
        integ.step(gpu, db, grav, steps[name])
        gpu.record(1)
        ms[name] = gpu.elapsed_ms(0, 1)
        kt = gpu.kernel_times()
        gpu.kernel_timing(False)
        e1 = energy(gpu, db, g, eps)
        inter = inter_eval * 4 * steps[name]
        rate[name] = inter / (ms[name] / 1e3)
        roofs[name] = roofline("f64" if name == "f64" else "f32",
                               inter_eval * kt["force_launches"], kt, rate[name])
        acc_rep[f"{name}_energy_drift"] = abs(e1[2] - e0[2]) / abs(e0[2])
        acc_rep[f"{name}_energy_drift_steps"] = steps[name]
        log(f"large {name}: {ms[name]:.0f} ms for {steps[name]} RK4 steps -> "
            f"{rate[name]:.4e} int/s, energy drift {acc_rep[f'{name}_energy_drift']:.3e}")
        db.close()
    alt = forms[1]
    return {
        "metric": "pair-interactions/s (f32-compensated forces, f64 state)",
        "value": rate[head], "unit": "pair-interactions/s",
        "steps": steps[head], "warmup": 0, "ms_per_step": ms[head] / steps[head],
        "dtype": DTYPE[head],
        "config": {"workload": "two colliding Plummer spheres N=1048576, softening "
                               "0.002, dt 2.5e-4, runge-kutta", "n": n,
                   "f32_form": head},
        "roofline": roofs[head],
        "f32_alt": {"form": alt, "value": rate[alt], "steps": steps[alt],
                    "roofline": roofs[alt]},
        "fp64_plain": {"value": rate["f64"], "ms_per_step": ms["f64"] / steps["f64"],
                       "roofline": roofs["f64"], "f32_speedup": rate[head] / rate["f64"]},
        "accuracy": acc_rep,
    }


# ---- Barnes-Hut (not in the default run) ------------------------------------
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
    acc = {}
    mx, med = rel_err(acc_fast, acc_direct)
    acc.update(fast_vs_direct_max_rel=mx, fast_vs_direct_median_rel=med,
               strict_vs_direct_median_rel=rel_err(acc_strict, acc_direct)[1],
               energy_drift_rk4=abs(e1[2] - e0[2]) / abs(e0[2]))
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        idx = np.sort(np.random.default_rng(9).choice(n, 4096, replace=False))
        threads = host_cpus()[0]
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


SUB_COMMON = {"n_gpus": 1, "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "parallelism": "none (1 GPU)"}


def with_common(rec, data):
    out = dict(SUB_COMMON)
    out["data"] = data
    out.update(rec)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps of the headline (single) workload")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload",
                    choices=["all", "single", "batched", "serve", "large", "bh", "copy"],
                    default="all",
                    help="all (default): single as the headline with batched, "
                         "serving and large as sub-records of the same line")
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--large-steps", type=int, default=20)
    ap.add_argument("--rates", default="20,80,640",
                    help="serve: request rates (1/s); the last one saturates the GPU")
    ap.add_argument("--requests", type=int, default=200, help="serve: requests per rate")
    ap.add_argument("--mix", default="plummer,star",
                    help="serve: request scenarios, from plummer, star, disc")
    ap.add_argument("--f32", choices=sorted(F32_FORMS), default=None,
                    help="f32 form of the single and large headlines (default f32r; "
                         "the other form is timed beside it)")
    ap.add_argument("--f32-batched", choices=sorted(F32_FORMS), default="f32r")
    ap.add_argument("--f32-serve", choices=sorted(F32_FORMS), default="f32r")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--clock-warm", type=float, default=0.5,
                    help="seconds of untimed force evaluations before each "
                         "single-system form and before any other workload run "
                         "on its own (GPU clock ramp)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(independent systems shard by system with no communication; a single "
              "system would shard by target range with one all-gather of source "
              "positions per force evaluation -- not built here)", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = {"single": 200, "all": 200, "batched": 50, "serve": 0, "large": 2,
                      "bh": 5, "copy": 1}[args.workload]
    gpu = nb.GpuContext(0)
    name, cus = gpu.device_info()
    log(f"device {name}, {cus} CUs")
    t_start = time.perf_counter()
    if args.workload not in ("copy", "single", "all"):
        warm_clocks(gpu, args.clock_warm)
    if args.workload == "all":
        rec = run_single(args, gpu)
        rec["batched"] = with_common(run_batched(args, gpu),
                                     "synthetic (seeded star systems)")
        rec["serving"] = with_common(run_serve(args, gpu),
                                     f"synthetic Poisson request stream "
                                     f"({args.mix.replace(',', ' / ')})")
        rec["large"] = with_common(run_large(args, gpu),
                                   "synthetic (two seeded Plummer spheres)")
    elif args.workload == "single":
        rec = run_single(args, gpu)
    else:
        if args.workload == "batched":
            args.batched_steps = args.steps
        if args.workload == "large":
            args.large_steps = args.steps
        fn = {"batched": run_batched, "serve": run_serve, "large": run_large,
              "bh": run_bh, "copy": run_copy}[args.workload]
        rec = fn(args, gpu)
        for k, v in SUB_COMMON.items():
            if k != "parallelism":
                rec.setdefault(k, v)
        rec.setdefault("steps", args.steps)
        rec.setdefault("warmup", args.warmup)
        rec.setdefault("cpu_baseline", None)
        rec.setdefault("data", "synthetic")
        rec.setdefault("higher_is_better", True)
    rec["device"] = {"arch": name, "compute_units": cus}
    rec["bench_wall_s"] = time.perf_counter() - t_start
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
