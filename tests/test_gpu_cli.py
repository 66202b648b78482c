"""Trajectory recording and the command-line front end on the GPU.

* nbody_step_record == nbody_step one step at a time + download, bit for
  bit, in every precision and integrator, across several staging fills;
* F64_STRICT recording == the oracle's per-step trajectory, bit for bit;
* n-body-sim without --gpu (F64_STRICT) on the reference's example inputs
  and parameters (tests/golden/examples): the .csv and .nbody files equal,
  byte for byte, the files built from the oracle's trajectory by the
  test-side writer (tests/snapfmt.py) -- the files the reference binary
  writes for these inputs, since the oracle reproduces its f64 path exactly;
* n-body-sim --gpu (F64 arithmetic) within 1e-10 of the oracle trajectory.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import nbody_amd as nb
import oracle
from snapfmt import csv_bytes, nbody_bytes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "n-body-sim_amd", "bin", "n-body-sim")
EXAMPLES = os.path.join(ROOT, "tests", "golden", "examples")
PARAMS = json.load(open(os.path.join(EXAMPLES, "params.json")))
INTEG = ["euler", "velocity-verlet", "runge-kutta"]
PREC = [nb.Precision.F64_STRICT, nb.Precision.F64, nb.Precision.F32C,
        nb.Precision.F32R]


def times_of(dt, k):
    t, out = 0.0, [0.0]
    for _ in range(k):
        t += dt            # src/simulation/gpu_simulation.rs:48
        out.append(t)
    return np.array(out)


def oracle_traj(b, integ, g, eps, dt, k):
    xs, ys, zs = oracle.integrate_record(integ, b.masses, b.pos_x, b.pos_y, b.pos_z,
                                         b.vel_x, b.vel_y, b.vel_z, g, eps, dt, k,
                                         nthreads=1 if len(b) < 512 else None)
    return (np.vstack([b.pos_x[None], xs]), np.vstack([b.pos_y[None], ys]),
            np.vstack([b.pos_z[None], zs]))


@pytest.mark.parametrize("prec", PREC)
@pytest.mark.parametrize("integ", INTEG)
def test_record_equals_single_steps(gpu, prec, integ):
    b = nb.plummer(300, seed=11)
    params = nb.SimulationParameters(1.0, 1e-3, 0.01)
    im = nb.IntegratorMethod.from_str(integ)
    print(f"record vs steps: {prec.name} {integ} n=300", flush=True)
    rec = nb.GpuSimulation(b, params, integrator=im, precision=prec)
    nrec = 37  # crosses step 32, where F32R reorders its targets
    times, xs, ys, zs = rec.record(nrec, max_stage_bytes=4 * 3 * 300 * 8)  # fills of 4
    assert np.array_equal(np.concatenate([[0.0], times]), times_of(1e-3, nrec))
    # one call of all steps: the same launches without the snapshots
    whole = nb.GpuSimulation(b, params, integrator=im, precision=prec)
    whole.step(nrec)
    st = whole.snapshot()
    assert np.array_equal(xs[-1], st.pos_x) and np.array_equal(zs[-1], st.pos_z)
    one = nb.GpuSimulation(b, params, integrator=im, precision=prec)
    for s in range(nrec):
        one.step(1)
        st = one.snapshot()
        assert np.array_equal(xs[s], st.pos_x) and np.array_equal(ys[s], st.pos_y)
        assert np.array_equal(zs[s], st.pos_z)
    # the state after recording continues like the stepped one
    rec.step(2)
    one.step(2)
    assert np.array_equal(rec.snapshot().vel_x, one.snapshot().vel_x)


@pytest.mark.parametrize("integ", INTEG)
def test_record_strict_equals_oracle(gpu, integ):
    b = nb.plummer(100, seed=5)
    print(f"record strict vs oracle: {integ}", flush=True)
    sim = nb.GpuSimulation(b, nb.SimulationParameters(1.0, 2e-3, 0.02),
                           integrator=nb.IntegratorMethod.from_str(integ),
                           precision=nb.Precision.F64_STRICT)
    _, xs, ys, zs = sim.record(20)
    ox, oy, oz = oracle_traj(b, integ, 1.0, 0.02, 2e-3, 20)
    assert np.array_equal(xs, ox[1:]) and np.array_equal(ys, oy[1:])
    assert np.array_equal(zs, oz[1:])


def run_cli(*args, cwd=None):
    print("cli:", " ".join(map(str, args)), flush=True)
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True,
                          cwd=cwd, timeout=300)


def example_args(name, integ=None, steps=None):
    p = PARAMS[name]
    integ = integ or p["integrator"]
    steps = p["num_steps"] if steps is None else steps
    return (["-g", p["g_constant"], "-t", p["time_step"], "-n", steps,
             "--softening-factor", p["softening_factor"], "--gravity", p["gravity"],
             "--integrator", integ], integ, steps)


@pytest.mark.parametrize("name", sorted(PARAMS))
def test_cli_examples_bytes_equal_oracle(gpu, tmp_path, name):
    """The GUI's run of each example (gui/runners.py:94-114 flags)."""
    args, integ, steps = example_args(name)
    ic = os.path.join(EXAMPLES, name + ".csv")
    b = nb.load_bodies(ic)
    p = PARAMS[name]
    want = oracle_traj(b, integ, p["g_constant"], p["softening_factor"],
                       p["time_step"], steps)
    times = times_of(p["time_step"], steps)
    for ext, make in ((".nbody", nbody_bytes), (".csv", csv_bytes)):
        out = tmp_path / ("out" + ext)
        r = run_cli("-i", ic, "-o", out, "--theta", "0.5", "--progress", *args)
        assert r.returncode == 0, r.stderr
        got = out.read_bytes()
        assert got == make(times, *want), f"{name} {ext} differs"
    one = max(steps // 100, 1)
    assert r.stdout.split() == [str((s + 1) * 100 // steps)
                                for s in range(steps) if s % one == 0]


@pytest.mark.parametrize("name", ["figure-eight", "sun-earth-moon"])
@pytest.mark.parametrize("integ", INTEG)
def test_cli_all_integrators_bytes(gpu, tmp_path, name, integ):
    args, integ, steps = example_args(name, integ, steps=2000)
    ic = os.path.join(EXAMPLES, name + ".csv")
    b = nb.load_bodies(ic)
    p = PARAMS[name]
    want = oracle_traj(b, integ, p["g_constant"], p["softening_factor"],
                       p["time_step"], steps)
    out = tmp_path / "o.csv"
    r = run_cli("-i", ic, "-o", out, *args)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == csv_bytes(times_of(p["time_step"], steps), *want)


def write_ic(path, b):
    with open(path, "w") as f:
        f.write("mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z\n")
        for row in zip(*b.arrays()):
            f.write(",".join(repr(float(v)) for v in row) + "\n")


def test_cli_plummer_chunked_nbody(gpu, tmp_path):
    """n = 5000: the run crosses several 16 MB snapshot chunks."""
    b = nb.plummer(5000, seed=3)
    ic = tmp_path / "ic.csv"
    write_ic(ic, b)
    assert np.array_equal(nb.load_bodies(ic).pos_x, b.pos_x)
    steps, dt, eps = 200, 1e-3, 0.01     # 133 steps per 16 MB chunk
    out = tmp_path / "o.nbody"
    r = run_cli("-i", ic, "-o", out, "-n", steps, "-t", dt, "--softening-factor", eps,
                "--integrator", "velocity-verlet", "--gravity", "newton-parallel")
    assert r.returncode == 0, r.stderr
    want = oracle_traj(b, "velocity-verlet", 1.0, eps, dt, steps)
    assert out.read_bytes() == nbody_bytes(times_of(dt, steps), *want)
    t, ids, pos = nb.read_snapshots(out)
    assert pos.shape == (steps + 1, 5000, 3)


@pytest.mark.parametrize("integ", INTEG)
def test_cli_gpu_flag_f64(gpu, tmp_path, integ):
    b = nb.plummer(700, seed=9)
    ic = tmp_path / "ic.csv"
    write_ic(ic, b)
    steps, dt, eps = 100, 1e-3, 0.01
    out = tmp_path / "o.nbody"
    r = run_cli("-i", ic, "-o", out, "-n", steps, "-t", dt, "--softening-factor", eps,
                "--integrator", integ, "--gravity", "newton-parallel", "--gpu")
    assert r.returncode == 0, r.stderr
    t, _, pos = nb.read_snapshots(out)
    want = np.stack(oracle_traj(b, integ, 1.0, eps, dt, steps), -1)
    assert np.array_equal(t, times_of(dt, steps))
    scale = np.abs(want).max()
    assert np.abs(pos - want).max() <= 1e-10 * scale
