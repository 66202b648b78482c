"""CPU oracle for the direct-sum gravity + integrator path.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py -- as the checker, never as the thing measured
or shipped.  The product (the HIP library) never routes through it.

ctypes bindings to ``oracle/liboracle.so`` (built from ``oracle/oracle.c`` by
``oracle/Makefile``), a binary64 restatement of the reference's CPU path:

* ``newton``            -- src/gravity/cpu/newton.rs:69-141
* ``newton_parallel``   -- src/gravity/cpu/newton_parallel.rs:31-53
* ``newton_targets``    -- the same per-target sum for a sampled target list
* ``integrate``         -- src/integrators/cpu/{euler.rs:57-82,
                           velocity_verlet.rs:63-104, runge_kutta.rs:149-281}
* ``energy``            -- softened K and U (no reference counterpart)

The reference is Rust and cannot be built in this environment (no cargo or
rustc), so there is no ``oracle/_ref`` build; the restatement is pinned by
closed forms, brute force, invariants and an independent numpy evaluation in
tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

EULER, VELOCITY_VERLET, RUNGE_KUTTA = 0, 1, 2
INTEGRATORS = {"euler": EULER, "velocity-verlet": VELOCITY_VERLET,
               "runge-kutta": RUNGE_KUTTA}

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    i64, d, i = ctypes.c_int64, ctypes.c_double, ctypes.c_int
    lib.oracle_newton.argtypes = [i64, d, d, _D, _D, _D, _D, _D, _D, _D]
    lib.oracle_newton.restype = None
    lib.oracle_newton_parallel.argtypes = [i64, d, d, _D, _D, _D, _D, _D, _D,
                                           _D, i]
    lib.oracle_newton_parallel.restype = None
    lib.oracle_newton_targets.argtypes = [i64, d, d, _D, _D, _D, _D, _I64,
                                          i64, _D, _D, _D, i]
    lib.oracle_newton_targets.restype = None
    lib.oracle_integrate.argtypes = [i, i64, d, d, d, i64, _D, _D, _D, _D, _D,
                                     _D, _D, i]
    lib.oracle_integrate.restype = ctypes.c_int
    lib.oracle_integrate_gravity.argtypes = [i, i, d, i64, d, d, d, i64, _D, _D,
                                             _D, _D, _D, _D, _D, i]
    lib.oracle_integrate_gravity.restype = ctypes.c_int
    lib.oracle_integrate_record.argtypes = [i, i, d, i64, d, d, d, i64, _D, _D,
                                            _D, _D, _D, _D, _D, _D, _D, _D, i]
    lib.oracle_integrate_record.restype = ctypes.c_int
    lib.oracle_barnes_hut.argtypes = [i64, d, d, d, _D, _D, _D, _D, _D, _D, _D, i]
    lib.oracle_barnes_hut.restype = ctypes.c_int
    lib.oracle_barnes_hut_targets.argtypes = [i64, d, d, d, _D, _D, _D, _D, _I64,
                                              i64, _D, _D, _D, i]
    lib.oracle_barnes_hut_targets.restype = ctypes.c_int
    lib.oracle_bh_morton.argtypes = [ctypes.c_uint64] * 3
    lib.oracle_bh_morton.restype = ctypes.c_uint64
    _I32 = ctypes.POINTER(ctypes.c_int32)
    _U64 = ctypes.POINTER(ctypes.c_uint64)
    lib.oracle_bh_tree.argtypes = [i64, _D, _D, _D, _D, _D, _D, _D, _D, _D,
                                   _I64, _I64, _I32, _I64, _I64, _U64]
    lib.oracle_bh_tree.restype = i64
    lib.oracle_energy.argtypes = [i64, d, d, _D, _D, _D, _D, _D, _D, _D, _D]
    lib.oracle_energy.restype = None
    _lib = lib
    return lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def newton(m, x, y, z, g: float, softening: float):
    """Serial f64 accelerations (newton.rs:69-87). Returns (ax, ay, az)."""
    lib = _load()
    m, x, y, z = map(_f64, (m, x, y, z))
    n = m.shape[0]
    ax, ay, az = (np.empty(n) for _ in range(3))
    lib.oracle_newton(n, g, softening, _p(m), _p(x), _p(y), _p(z), _p(ax),
                      _p(ay), _p(az))
    return ax, ay, az


def newton_parallel(m, x, y, z, g: float, softening: float,
                    nthreads: int | None = None):
    """Threaded f64 accelerations (newton_parallel.rs:31-53)."""
    lib = _load()
    m, x, y, z = map(_f64, (m, x, y, z))
    n = m.shape[0]
    ax, ay, az = (np.empty(n) for _ in range(3))
    lib.oracle_newton_parallel(n, g, softening, _p(m), _p(x), _p(y), _p(z),
                               _p(ax), _p(ay), _p(az),
                               nthreads or default_threads())
    return ax, ay, az


def newton_targets(m, x, y, z, g: float, softening: float, idx,
                   nthreads: int | None = None):
    """f64 accelerations of the targets ``idx`` only (same arithmetic)."""
    lib = _load()
    m, x, y, z = map(_f64, (m, x, y, z))
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    k = idx.shape[0]
    ax, ay, az = (np.empty(k) for _ in range(3))
    lib.oracle_newton_targets(m.shape[0], g, softening, _p(m), _p(x), _p(y),
                              _p(z), idx.ctypes.data_as(_I64), k, _p(ax),
                              _p(ay), _p(az), nthreads or default_threads())
    return ax, ay, az


GRAVITY = {"newton": 0, "newton-parallel": 1, "barnes-hut": 2}


def _gravity_id(gravity) -> int:
    return GRAVITY[gravity] if isinstance(gravity, str) else int(gravity)


def integrate(integrator, m, x, y, z, vx, vy, vz, g: float, softening: float,
              dt: float, nsteps: int, nthreads: int | None = None,
              gravity: int | str = 1, theta: float = 0.5):
    """Advance copies of the state by ``nsteps`` reference steps.

    ``integrator`` is 0/1/2 or "euler"/"velocity-verlet"/"runge-kutta";
    ``gravity`` 0/1/2 or "newton"/"newton-parallel"/"barnes-hut" (with
    ``theta``).  Returns (x, y, z, vx, vy, vz) as new arrays.
    """
    lib = _load()
    if isinstance(integrator, str):
        integrator = INTEGRATORS[integrator]
    m = _f64(m)
    st = [np.array(a, dtype=np.float64, copy=True) for a in (x, y, z, vx, vy, vz)]
    rc = lib.oracle_integrate_gravity(int(integrator), _gravity_id(gravity), theta,
                                      m.shape[0], g, softening, dt,
                                      int(nsteps), _p(m), *[_p(a) for a in st],
                                      nthreads or default_threads())
    if rc != 0:
        raise ValueError("oracle_integrate failed")
    return tuple(st)


def integrate_record(integrator, m, x, y, z, vx, vy, vz, g: float,
                     softening: float, dt: float, nsteps: int,
                     nthreads: int | None = None, gravity: int | str = 1,
                     theta: float = 0.5):
    """Positions after every one of ``nsteps`` reference steps.

    Returns (xs, ys, zs), each of shape (nsteps, n): row s holds the
    positions after step s + 1, as the reference's run loop snapshots them.
    """
    lib = _load()
    if isinstance(integrator, str):
        integrator = INTEGRATORS[integrator]
    m = _f64(m)
    n = m.shape[0]
    st = [np.array(a, dtype=np.float64, copy=True) for a in (x, y, z, vx, vy, vz)]
    out = [np.empty((int(nsteps), n)) for _ in range(3)]
    rc = lib.oracle_integrate_record(int(integrator), _gravity_id(gravity), theta,
                                     n, g, softening, dt,
                                     int(nsteps), _p(m), *[_p(a) for a in st],
                                     *[_p(a) for a in out],
                                     nthreads or default_threads())
    if rc != 0:
        raise ValueError("oracle_integrate_record failed")
    return tuple(out)


def barnes_hut(m, x, y, z, g: float, softening: float, theta: float,
               nthreads: int | None = None):
    """BarnesHutGravity::calculate_accelerations (one tree build + traversal)."""
    lib = _load()
    m, x, y, z = (_f64(a) for a in (m, x, y, z))
    n = m.shape[0]
    ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = lib.oracle_barnes_hut(n, g, softening, theta, _p(m), _p(x), _p(y), _p(z),
                               _p(ax), _p(ay), _p(az), nthreads or default_threads())
    if rc != 0:
        raise ValueError("oracle_barnes_hut failed")
    return ax, ay, az


def barnes_hut_targets(m, x, y, z, g: float, softening: float, theta: float, idx,
                       nthreads: int | None = None):
    """barnes_hut() for the targets idx only (the full tree is built)."""
    lib = _load()
    m, x, y, z = (_f64(a) for a in (m, x, y, z))
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    k = idx.shape[0]
    ax, ay, az = np.zeros(k), np.zeros(k), np.zeros(k)
    rc = lib.oracle_barnes_hut_targets(m.shape[0], g, softening, theta, _p(m), _p(x),
                                       _p(y), _p(z), idx.ctypes.data_as(_I64), k,
                                       _p(ax), _p(ay), _p(az),
                                       nthreads or default_threads())
    if rc != 0:
        raise ValueError("oracle_barnes_hut_targets failed")
    return ax, ay, az


def bh_tree(m, x, y, z):
    """The octree of one build: dict of node arrays (preorder) and the
    Morton order of the bodies (perm[k] = original index, codes[k])."""
    lib = _load()
    m, x, y, z = (_f64(a) for a in (m, x, y, z))
    n = m.shape[0]
    cap = n * 22 + 1
    t = {k: np.zeros(cap) for k in ("mass", "com_x", "com_y", "com_z", "width")}
    t["start"] = np.zeros(cap, np.int64)
    t["count"] = np.zeros(cap, np.int64)
    t["nchild"] = np.zeros(cap, np.int32)
    t["child"] = np.zeros((cap, 8), np.int64)
    perm = np.zeros(n, np.int64)
    codes = np.zeros(n, np.uint64)
    c = lib.oracle_bh_tree(n, _p(m), _p(x), _p(y), _p(z),
                           *[_p(t[k]) for k in ("mass", "com_x", "com_y", "com_z",
                                                "width")],
                           t["start"].ctypes.data_as(_I64),
                           t["count"].ctypes.data_as(_I64),
                           t["nchild"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                           t["child"].ctypes.data_as(_I64),
                           perm.ctypes.data_as(_I64),
                           codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if c < 0:
        raise ValueError("oracle_bh_tree failed")
    out = {k: v[:c] for k, v in t.items()}
    out["perm"], out["codes"], out["node_count"] = perm, codes, int(c)
    return out


def morton(x: int, y: int, z: int) -> int:
    return int(_load().oracle_bh_morton(x, y, z))


def energy(m, x, y, z, vx, vy, vz, g: float, softening: float):
    """Softened (K, U) in f64."""
    lib = _load()
    arrs = list(map(_f64, (m, x, y, z, vx, vy, vz)))
    out = np.zeros(2)
    lib.oracle_energy(arrs[0].shape[0], g, softening, *[_p(a) for a in arrs],
                      _p(out))
    return float(out[0]), float(out[1])
