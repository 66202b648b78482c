"""CPU-only checks of the C ABI library and the host layer (no GPU compute).

* every function include/nbody.h declares is exported by libnbody_amd.so;
* without a GPU, context/server creation fails with NBODY_E_DEVICE (no CPU
  fallback exists);
* argument validation of the host-only entry points;
* the scenario generators (host code) against their defining properties;
* the reference's method-selection strings (src/args.rs:125-136, :176-187).
"""
import ctypes
import re

import numpy as np
import pytest

import nbody_amd as nb


def header_functions():
    text = open(nb.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbody_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = nb.load_library()
    names = header_functions()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in nb.load_library().nbody_version()


def _no_gpu():
    try:
        import torch
        return not torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return True


@pytest.mark.skipif(not _no_gpu(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    with pytest.raises(nb.NbodyError) as e:
        nb.GpuContext(0)
    assert e.value.code == nb.E_DEVICE
    with pytest.raises(nb.NbodyError) as e:
        nb.Server(capacity_bodies=1024)
    assert e.value.code == nb.E_DEVICE


def test_null_and_invalid_arguments():
    lib = nb.load_library()
    h = ctypes.c_void_p()
    assert lib.nbody_ctx_create(0, None) == nb.E_INVALID
    assert lib.nbody_ctx_synchronize(None) == nb.E_INVALID
    assert lib.nbody_step(None, None, 0, 1, 1, 1e-3, 1.0, 0.0, 1) == nb.E_INVALID
    assert lib.nbody_bodies_create(None, 4, *[None] * 7, ctypes.byref(h)) == nb.E_INVALID
    assert lib.nbody_ctx_destroy(None) == nb.OK          # null is a no-op
    assert lib.nbody_bodies_len(None) == 0
    arrs = [np.empty(4) for _ in range(7)]
    ptrs = [nb._p(a) for a in arrs]
    assert lib.nbody_scenario_plummer(0, 1, 1.0, 1.0, 1.0, *ptrs) == nb.E_INVALID
    assert lib.nbody_scenario_plummer(4, 1, -1.0, 1.0, 1.0, *ptrs) == nb.E_INVALID
    assert lib.nbody_scenario_colliding(1, 1, 10.0, 0.5, 1.0, *ptrs) == nb.E_INVALID
    assert lib.nbody_scenario_star_system(4, 1, 1.0, 1e-4, 1e-7, 0.5, 10.0, 5.0, 1.0,
                                          *ptrs) == nb.E_INVALID   # m_hi < m_lo


def test_scenarios_reject_non_finite_parameters():
    """Every generator parameter must be finite: an infinite scale, mass,
    radius, angle or G would otherwise give non-finite bodies under
    NBODY_OK."""
    lib = nb.load_library()
    arrs = [np.empty(4) for _ in range(7)]
    ptrs = [nb._p(a) for a in arrs]
    inf, nan = float("inf"), float("nan")
    plummer = [1.0, 1.0, 1.0]                        # a, total mass, g
    for k in range(3):
        for bad in (inf, nan):
            args = list(plummer)
            args[k] = bad
            assert lib.nbody_scenario_plummer(4, 1, *args, *ptrs) == nb.E_INVALID
    star = [1.0, 1e-7, 1e-4, 0.5, 10.0, 5.0, 1.0]    # central, m_lo, m_hi, r_lo, r_hi, incl, g
    for k in range(7):
        args = list(star)
        args[k] = inf
        assert lib.nbody_scenario_star_system(4, 1, *args, *ptrs) == nb.E_INVALID, k
    assert lib.nbody_scenario_colliding(4, 1, 10.0, 0.5, inf, *ptrs) == nb.E_INVALID
    disc = [1.0, 1e-3, 0.5, 0.0, 1.0]                # radius, mass, power, softening, g
    for k in range(5):
        args = list(disc)
        args[k] = inf
        assert lib.nbody_scenario_disc(4, 1, *args, *ptrs) == nb.E_INVALID, k
    # the valid parameter sets themselves succeed, with finite bodies
    assert lib.nbody_scenario_plummer(4, 1, *plummer, *ptrs) == nb.OK
    assert lib.nbody_scenario_star_system(4, 1, *star, *ptrs) == nb.OK
    assert lib.nbody_scenario_disc(4, 1, *disc, *ptrs) == nb.OK
    assert all(np.isfinite(a).all() for a in arrs)


def test_plummer_properties():
    b = nb.plummer(20000, 7)
    m = b.masses
    assert np.all(m == 1.0 / 20000)
    pos = np.stack([b.pos_x, b.pos_y, b.pos_z], 1)
    vel = np.stack([b.vel_x, b.vel_y, b.vel_z], 1)
    assert np.abs((m[:, None] * pos).sum(0)).max() < 1e-15   # centre-of-mass frame
    assert np.abs((m[:, None] * vel).sum(0)).max() < 1e-15
    r = np.linalg.norm(pos, axis=1)
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1) = 1.3048
    assert abs(np.median(r) - 1.3048) < 0.03
    assert r.max() <= 20.0 + 1e-9
    # virial equilibrium: K = 3*pi/64 for G = M = a = 1 (untruncated)
    K = 0.5 * (m * (vel ** 2).sum(1)).sum()
    assert abs(K - 3 * np.pi / 64) < 0.01
    # speeds below the local escape speed
    ve = np.sqrt(2.0) * (1 + r * r) ** -0.25
    assert np.all(np.linalg.norm(vel, axis=1) <= ve * (1 + 1e-12) + 1e-12)


def test_plummer_deterministic_and_seeded():
    a, b, c = nb.plummer(100, 1), nb.plummer(100, 1), nb.plummer(100, 2)
    assert all(np.array_equal(p, q) for p, q in zip(a.arrays(), b.arrays()))
    assert not np.array_equal(a.pos_x, c.pos_x)


def test_star_system_properties():
    b = nb.star_system(5000, 3, m_lo=1e-7, m_hi=1e-4, r_lo=0.5, r_hi=10.0,
                       max_incl_deg=5.0)
    assert b.masses[0] == 1.0 and b.pos_x[0] == 0.0 and b.vel_x[0] == 0.0
    m = b.masses[1:]
    assert m.min() >= 1e-7 and m.max() <= 1e-4
    # log-uniform: median of log10 mass at the midpoint -5.5
    assert abs(np.median(np.log10(m)) + 5.5) < 0.05
    pos = np.stack([b.pos_x, b.pos_y, b.pos_z], 1)[1:]
    vel = np.stack([b.vel_x, b.vel_y, b.vel_z], 1)[1:]
    r = np.linalg.norm(pos, axis=1)
    assert r.min() >= 0.5 and r.max() <= 10.0
    # circular: speed sqrt(G M / r), velocity perpendicular to the radius
    assert np.allclose(np.linalg.norm(vel, axis=1), np.sqrt(1.0 / r), rtol=1e-12)
    assert np.abs((pos * vel).sum(1)).max() < 1e-12
    # orbit normal within 5 degrees of +z
    h = np.cross(pos, vel)
    incl = np.degrees(np.arccos(h[:, 2] / np.linalg.norm(h, axis=1)))
    assert incl.max() < 5.0 + 1e-9


def test_disc_galaxy_properties():
    R, mb, p, soft = 20.0, 1e-3, 1.5, 0.5
    b = nb.disc_galaxy(20000, 4, radius=R, body_mass=mb, density_power=p, softening=soft)
    assert np.all(b.masses == mb)
    r = np.hypot(b.pos_x, b.pos_y)
    assert np.all(np.diff(r) >= 0) and r.max() <= R          # written by radius
    # radii R * U^p: median R * 0.5^p
    assert abs(np.median(r) / (R * 0.5 ** p) - 1) < 0.03
    assert abs(b.pos_z.std() / (0.01 * R) - 1) < 0.03         # thin layer
    # softened circular speed of the enclosed mass, 10 % dispersion
    menc = (np.arange(20000) + 1) * mb
    v = np.sqrt(menc * r / (r * r + soft * soft))
    th = np.arctan2(b.pos_y, b.pos_x)
    vt = -b.vel_x * np.sin(th) + b.vel_y * np.cos(th)
    vr = b.vel_x * np.cos(th) + b.vel_y * np.sin(th)
    assert abs(np.mean((vt - v) / v)) < 0.01
    assert abs(np.std((vt - v) / v) - 0.1) < 0.005
    assert abs(np.std(vr / v) - 0.1) < 0.005
    assert abs(np.std(b.vel_z / v) - 0.05) < 0.003
    a, c = nb.disc_galaxy(100, 1), nb.disc_galaxy(100, 1)
    assert all(np.array_equal(p1, q1) for p1, q1 in zip(a.arrays(), c.arrays()))


def test_colliding_properties():
    b = nb.colliding(2000, 5, separation=10.0, rel_speed=0.5)
    assert np.allclose(b.masses, 1.0 / 2000)
    left, right = slice(0, 1000), slice(1000, 2000)
    assert abs(b.pos_x[left].mean() + 5.0) < 1e-12
    assert abs(b.pos_x[right].mean() - 5.0) < 1e-12
    assert abs(b.vel_x[left].mean() - 0.25) < 1e-12
    assert abs(b.vel_x[right].mean() + 0.25) < 1e-12


def test_method_strings_match_reference():
    G, I = nb.GravityMethod, nb.IntegratorMethod
    assert G.from_str("newton") is G.NEWTON
    assert G.from_str("Newton-Parallel") is G.NEWTON_PARALLEL   # to_lowercase()
    assert G.from_str("barnes-hut") is G.BARNES_HUT
    assert I.from_str("euler") is I.EULER
    assert I.from_str("velocity-verlet") is I.VELOCITY_VERLET
    assert I.from_str("RUNGE-KUTTA") is I.RUNGE_KUTTA
    with pytest.raises(ValueError, match="Invalid gravity method: foo"):
        G.from_str("foo")
    with pytest.raises(ValueError, match="Invalid integrator method: rk5"):
        I.from_str("rk5")


def test_gpu_create_selection():
    p = nb.SimulationParameters(g_constant=2.0, time_step=0.01, softening_factor=0.1)
    grav = nb.GravityMethod.NEWTON_PARALLEL.gpu_create(p)
    assert grav.g_constant == 2.0 and grav.eps2 == 0.1 * 0.1   # src/args.rs:115
    with pytest.raises(nb.NbodyError) as e:
        nb.GravityMethod.NEWTON.gpu_create(p)
    assert e.value.code == nb.E_UNSUPPORTED
    p.theta = 0.7
    bh = nb.GravityMethod.BARNES_HUT.gpu_create(p, nb.Precision.F64_STRICT)
    assert isinstance(bh, nb.GpuBarnesHutGravity) and bh.theta == 0.7
    assert bh.method == nb.GravityMethod.BARNES_HUT and bh.eps2 == 0.1 * 0.1
    for im, cls in ((nb.IntegratorMethod.EULER, nb.GpuEulerIntegrator),
                    (nb.IntegratorMethod.VELOCITY_VERLET, nb.GpuVelocityVerletIntegrator),
                    (nb.IntegratorMethod.RUNGE_KUTTA, nb.GpuRungeKuttaIntegrator)):
        integ = im.gpu_create(0.01)
        assert isinstance(integ, cls) and integ.dt == 0.01


def test_bodies_length_check():
    with pytest.raises(ValueError):
        nb.Bodies(np.ones(3), np.ones(3), np.ones(2), np.ones(3), np.ones(3),
                  np.ones(3), np.ones(3))


def test_tuning_table_file_and_abi(tmp_path):
    """The split table loads from NBODY_AMD_TUNING (format `split b k`), is
    clamped to 1..32 parts, and set/get round-trip through the C ABI."""
    import subprocess
    import sys
    f = tmp_path / "t.txt"
    f.write_text("# test table\nsplit 10 3\nsplit 12 99\nsplit 5 1\n")
    code = ("import sys; sys.path.insert(0, %r); import nbody_amd as nb; "
            "t = nb.tuning_split(); print(t[10], t[12], t[5]); "
            "t[7] = 2; nb.set_tuning_split(t); print(nb.tuning_split()[7])"
            % nb._PKG_DIR)
    env = dict(__import__("os").environ, NBODY_AMD_TUNING=str(f))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["3", "32", "1", "2"]


def test_committed_tuning_table_parses():
    import os
    path = os.path.join(nb._PKG_DIR, "tuning", "gfx950.txt")
    rows = [ln.split() for ln in open(path) if ln.startswith("split ")]
    assert len(rows) == 32
    assert all(1 <= int(k) <= 32 for _, _, k in rows)


def test_tuning_variant_lines_and_abi(tmp_path):
    """`variant <mode> <bucket> <v>` lines load per precision (1 f64, 2 f32c,
    3 f32r); out-of-range entries are ignored; strict f64 has one kernel; the
    set/get calls round-trip and reject out-of-range variants."""
    import subprocess
    import sys
    f = tmp_path / "t.txt"
    f.write_text("split 10 3\nvariant 2 10 7\nvariant 3 11 4\nvariant 1 12 9\n"
                 "variant 0 10 5\nvariant 2 40 1\n")
    code = ("import sys; sys.path.insert(0, %r); import nbody_amd as nb; P = nb.Precision; "
            "print(nb.tuning_variant(P.F32C)[10], nb.tuning_variant(P.F32R)[11], "
            "nb.tuning_variant(P.F64)[12], nb.tuning_variant(P.F64_STRICT)[10]); "
            "t = nb.tuning_variant(P.F64); t[8] = 6; nb.set_tuning_variant(P.F64, t); "
            "print(nb.tuning_variant(P.F64)[8]); "
            "lib = nb.load_library(); import ctypes; a = (ctypes.c_int * 32)(*([9] * 32)); "
            "print(lib.nbody_tuning_set_variant(int(P.F64), a, 32), "
            "lib.nbody_tuning_set_variant(int(P.F64_STRICT), (ctypes.c_int * 32)(), 32))"
            % nb._PKG_DIR)
    env = dict(__import__("os").environ, NBODY_AMD_TUNING=str(f))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["7", "4", "0", "0", "6", str(nb.E_INVALID),
                                  str(nb.E_INVALID)]


def test_committed_tuning_variants_parse():
    import os
    path = os.path.join(nb._PKG_DIR, "tuning", "gfx950.txt")
    rows = [ln.split() for ln in open(path) if ln.startswith("variant ")]
    assert rows and all(1 <= int(m) <= 3 and 0 <= int(b) < 32 and 0 <= int(v) < 9
                        for _, m, b, v in rows)
