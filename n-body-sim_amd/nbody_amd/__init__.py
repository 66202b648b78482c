"""Python host layer over ``libnbody_amd.so`` (C ABI in ``include/nbody.h``).

Mirrors the reference's GPU selection and trait surface so callers read the
same way as the Rust code they replace:

* ``GravityMethod`` / ``IntegratorMethod`` with ``from_str`` and
  ``gpu_create`` -- src/args.rs:84-187
* ``CudaManager``-like context ``GpuContext`` -- src/gpu/cuda_manager.rs:12-62
* ``DeviceBodies`` -- src/gpu/device_bodies.rs:7-55
* ``GpuNewtonParallelGravity.calculate_accelerations`` -- src/gravity.rs:30-36,
  src/gravity/gpu/newton_parallel.rs:20-46
* ``GpuEulerIntegrator`` / ``GpuVelocityVerletIntegrator`` /
  ``GpuRungeKuttaIntegrator.step`` -- src/integrators/gpu/gpu_integrator.rs:8-15
* ``GpuSimulation.step`` / ``snapshot`` -- src/simulation/gpu_simulation.rs:45-65

There is no CPU fallback: every compute call goes through the HIP library and
raises ``NbodyError`` when it is missing or no GPU is present.
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# NBODY_AMD_LIB selects another build of the same library (e.g. an A/B build
# of a kernel variant); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("NBODY_AMD_LIB") or os.path.join(_PKG_DIR, "libnbody_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG_DIR), "include", "nbody.h")

OK = 0
E_INVALID, E_UNSUPPORTED, E_DEVICE, E_NOMEM = -1, -2, -3, -4
E_NOSPACE, E_NOTFOUND, E_STATE, E_TIMEOUT = -5, -6, -7, -8

EVENT_SNAPSHOT, EVENT_DONE, EVENT_FAILED, EVENT_CANCELLED = 1, 2, 3, 4


class NbodyError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


_D = ctypes.POINTER(ctypes.c_double)
_lib = None


def _bind(lib):
    i, i64, u64, d = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
    vp = ctypes.c_void_p
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "nbody_last_error": ([], ctypes.c_char_p),
        "nbody_version": ([], ctypes.c_char_p),
        "nbody_ctx_create": ([i, pp], i),
        "nbody_ctx_destroy": ([vp], i),
        "nbody_ctx_synchronize": ([vp], i),
        "nbody_ctx_device_info": ([vp, ctypes.c_char_p, i, ctypes.POINTER(i)], i),
        "nbody_ctx_record": ([vp, i], i),
        "nbody_ctx_elapsed": ([vp, i, i, ctypes.POINTER(ctypes.c_float)], i),
        "nbody_ctx_kernel_timing": ([vp, i], i),
        "nbody_ctx_kernel_times": ([vp, _D], i),
        "nbody_tuning_set_split": ([ctypes.POINTER(i), i], i),
        "nbody_tuning_get_split": ([ctypes.POINTER(i), i], i),
        "nbody_tuning_set_variant": ([i, ctypes.POINTER(i), i], i),
        "nbody_tuning_get_variant": ([i, ctypes.POINTER(i), i], i),
        "nbody_bodies_create": ([vp, i64] + [_D] * 7 + [pp], i),
        "nbody_bodies_download": ([vp, vp] + [_D] * 7, i),
        "nbody_bodies_destroy": ([vp], i),
        "nbody_bodies_len": ([vp], i64),
        "nbody_bodies_set_theta": ([vp, d], i),
        "nbody_bodies_tree_nodes": ([vp, ctypes.POINTER(i64)], i),
        "nbody_calculate_accelerations": ([vp, vp, i, d, d, i, _D, _D, _D], i),
        "nbody_calculate_accelerations_noalloc": ([vp, i64, vp, vp, vp, vp, d, d, i,
                                                   vp, vp, vp], i),
        "nbody_step": ([vp, vp, i, i, i, d, d, d, i64], i),
        "nbody_integrate_stage": ([vp, vp, i, d], i),
        "nbody_step_record": ([vp, vp, i, i, i, d, d, d, i64, _D, _D, _D, i64], i),
        "nbody_energy_momentum": ([vp, vp, d, d, _D], i),
        "nbody_scenario_plummer": ([i64, u64, d, d, d] + [_D] * 7, i),
        "nbody_scenario_star_system": ([i64, u64, d, d, d, d, d, d, d] + [_D] * 7, i),
        "nbody_scenario_colliding": ([i64, u64, d, d, d] + [_D] * 7, i),
        "nbody_scenario_disc": ([i64, u64, d, d, d, d, d] + [_D] * 7, i),
        "nbody_batch_create": ([vp, i, ctypes.POINTER(i64), ctypes.POINTER(i), _D, _D,
                                _D, i] + [_D] * 7 + [pp], i),
        "nbody_batch_step": ([vp, vp, i64], i),
        "nbody_batch_download": ([vp, vp] + [_D] * 6, i),
        "nbody_batch_energy": ([vp, vp, _D], i),
        "nbody_batch_destroy": ([vp], i),
        "nbody_server_create": ([ctypes.POINTER(ServerConfig), pp], i),
        "nbody_server_submit": ([vp, ctypes.POINTER(Request), ctypes.POINTER(u64)], i),
        "nbody_server_poll": ([vp, i, ctypes.POINTER(Event)] + [_D] * 6 + [i64], i),
        "nbody_server_cancel": ([vp, u64], i),
        "nbody_server_stats": ([vp, _D], i),
        "nbody_server_destroy": ([vp], i),
        "nbody_writer_open": ([ctypes.c_char_p, pp], i),
        "nbody_writer_append": ([vp, d, i64, _D, _D, _D], i),
        "nbody_writer_close": ([vp], i),
        "nbody_format_f64": ([d, ctypes.c_char_p], i),
        "nbody_load_initial_conditions": ([ctypes.c_char_p, ctypes.POINTER(i64)]
                                          + [_D] * 7, i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("precision", ctypes.c_int),
                ("capacity_bodies", ctypes.c_int64), ("max_systems", ctypes.c_int),
                ("snapshot_every", ctypes.c_int), ("max_snapshot_bytes", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("integrator", ctypes.c_int),
                ("gravity_method", ctypes.c_int), ("g", ctypes.c_double),
                ("dt", ctypes.c_double), ("eps2", ctypes.c_double),
                ("steps", ctypes.c_int64), ("want_snapshots", ctypes.c_int)] + [
        (f, _D) for f in ("masses", "pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z")]


class Event(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("request_id", ctypes.c_uint64),
                ("step", ctypes.c_int64), ("n", ctypes.c_int64),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_event", ctypes.c_double), ("message", ctypes.c_char * 128)]


def load_library():
    """Load the built HIP library (raises NbodyError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NbodyError(E_DEVICE, f"{LIB_PATH} is not built; run "
                             "__graft_entry__.build() (make -C n-body-sim_amd/csrc)")
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def _check(rc: int):
    if rc != OK:
        msg = load_library().nbody_last_error().decode(errors="replace")
        raise NbodyError(rc, msg)


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- reference selection enums (src/args.rs:84-187) -----------------------
class Precision(enum.IntEnum):
    F64_STRICT = 0   # reference operation order: bit-identical to the f64 CPU path
    F64 = 1          # f64, rsq + refinement, FMA
    F32C = 2         # f64 differences rounded to f32, f32 terms, compensated sums
    F32R = 3         # f32 relative to each wave's middle target, Hilbert order, near chunks corrected (fastest for large systems)


class GravityMethod(enum.IntEnum):
    NEWTON = 0
    NEWTON_PARALLEL = 1
    BARNES_HUT = 2

    @classmethod
    def from_str(cls, s: str) -> "GravityMethod":
        # src/args.rs:125-136
        table = {"newton": cls.NEWTON, "newton-parallel": cls.NEWTON_PARALLEL,
                 "barnes-hut": cls.BARNES_HUT}
        try:
            return table[s.lower()]
        except KeyError:
            raise ValueError(f"Invalid gravity method: {s}") from None

    def gpu_create(self, params: "SimulationParameters",
                   precision: Precision = Precision.F64):
        # src/args.rs:111-123 (newton-parallel only there; barnes-hut added
        # here); eps2 = eps^2
        eps2 = params.softening_factor * params.softening_factor
        if self is GravityMethod.NEWTON_PARALLEL:
            return GpuNewtonParallelGravity(params.g_constant, eps2, precision)
        if self is GravityMethod.BARNES_HUT:
            return GpuBarnesHutGravity(params.g_constant, eps2, params.theta, precision)
        raise NbodyError(E_UNSUPPORTED, "GPU-accelerated version of "
                         f"{self.name} gravity method is not implemented")


class IntegratorMethod(enum.IntEnum):
    EULER = 0
    VELOCITY_VERLET = 1
    RUNGE_KUTTA = 2

    @classmethod
    def from_str(cls, s: str) -> "IntegratorMethod":
        # src/args.rs:176-187
        table = {"euler": cls.EULER, "velocity-verlet": cls.VELOCITY_VERLET,
                 "runge-kutta": cls.RUNGE_KUTTA}
        try:
            return table[s.lower()]
        except KeyError:
            raise ValueError(f"Invalid integrator method: {s}") from None

    def gpu_create(self, time_step: float) -> "GpuIntegrator":
        # src/args.rs:165-174 (Euler only there; all three here)
        return {IntegratorMethod.EULER: GpuEulerIntegrator,
                IntegratorMethod.VELOCITY_VERLET: GpuVelocityVerletIntegrator,
                IntegratorMethod.RUNGE_KUTTA: GpuRungeKuttaIntegrator}[self](time_step)


@dataclass
class SimulationParameters:
    """src/args.rs SimulationParameters (the fields the hot path reads)."""
    g_constant: float = 1.0
    time_step: float = 0.01
    softening_factor: float = 0.005
    theta: float = 0.5


@dataclass
class Bodies:
    """SoA f64 bodies, src/body.rs:52-61."""
    masses: np.ndarray
    pos_x: np.ndarray
    pos_y: np.ndarray
    pos_z: np.ndarray
    vel_x: np.ndarray
    vel_y: np.ndarray
    vel_z: np.ndarray

    def __post_init__(self):
        for f in self.fields():
            setattr(self, f, _f64(getattr(self, f)))
        n = len(self.masses)
        if any(len(getattr(self, f)) != n for f in self.fields()):
            raise ValueError("all body arrays must have the same length")  # body.rs:74-82

    @staticmethod
    def fields():
        return ("masses", "pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z")

    def __len__(self):
        return len(self.masses)

    def arrays(self):
        return [getattr(self, f) for f in self.fields()]

    def copy(self) -> "Bodies":
        return Bodies(*[a.copy() for a in self.arrays()])


class GpuContext:
    """The CudaManager counterpart: device, stream, init check."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.nbody_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        _check(load_library().nbody_ctx_synchronize(self._h))

    def device_info(self):
        buf = ctypes.create_string_buffer(128)
        cu = ctypes.c_int()
        _check(load_library().nbody_ctx_device_info(self._h, buf, 128, ctypes.byref(cu)))
        return buf.value.decode(), cu.value

    def record(self, slot: int):
        """Record timing event `slot` (0..7) on the context's stream."""
        _check(load_library().nbody_ctx_record(self._h, slot))

    def elapsed_ms(self, slot_from: int, slot_to: int) -> float:
        ms = ctypes.c_float()
        _check(load_library().nbody_ctx_elapsed(self._h, slot_from, slot_to,
                                                ctypes.byref(ms)))
        return ms.value

    def kernel_timing(self, enable: bool):
        """Bracket every force and combine launch of this context's bodies
        and batches with HIP events (zeroes the sums)."""
        _check(load_library().nbody_ctx_kernel_timing(self._h, int(bool(enable))))

    def kernel_times(self) -> dict:
        """Summed launch times since kernel_timing(True): force and combine
        milliseconds and launch counts (waits for the recorded launches)."""
        out = np.zeros(4)
        _check(load_library().nbody_ctx_kernel_times(self._h, _p(out)))
        return {"force_ms": float(out[0]), "force_launches": int(out[1]),
                "combine_ms": float(out[2]), "combine_launches": int(out[3])}

    def close(self):
        if getattr(self, "_h", None):
            load_library().nbody_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBodies:
    """src/gpu/device_bodies.rs:7-55: device-resident SoA bodies."""

    def __init__(self, ctx: GpuContext, bodies: Bodies):
        lib = load_library()
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib.nbody_bodies_create(ctx.handle, len(bodies),
                                       *[_p(a) for a in bodies.arrays()],
                                       ctypes.byref(h)))
        self._h = h
        self._n = len(bodies)

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return self._n

    def download(self) -> Bodies:
        arrs = [np.empty(self._n) for _ in range(7)]
        _check(load_library().nbody_bodies_download(self.ctx.handle, self._h,
                                                    *[_p(a) for a in arrs]))
        return Bodies(*arrs)

    def close(self):
        if getattr(self, "_h", None):
            load_library().nbody_bodies_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuNewtonParallelGravity:
    """src/gravity/gpu/newton_parallel.rs: direct sum on the GPU."""

    method = GravityMethod.NEWTON_PARALLEL

    def __init__(self, g_constant: float, eps2: float,
                 precision: Precision = Precision.F64):
        self.g_constant = float(g_constant)
        self.eps2 = float(eps2)
        self.precision = Precision(precision)

    def configure(self, bodies: "DeviceBodies"):
        """Per-method settings of the bodies (none for the direct sum)."""

    def calculate_accelerations(self, gpu: GpuContext, bodies: DeviceBodies,
                                download: bool = True):
        """Accelerations into the device arrays; returned as (ax, ay, az)."""
        n = len(bodies)
        out = [np.empty(n) for _ in range(3)] if download else [None] * 3
        self.configure(bodies)
        _check(load_library().nbody_calculate_accelerations(
            gpu.handle, bodies.handle, int(self.method),
            self.g_constant, self.eps2, int(self.precision), *[_p(a) for a in out]))
        return tuple(out) if download else None


class GpuBarnesHutGravity(GpuNewtonParallelGravity):
    """src/gravity/cpu/barnes_hut.rs on the GPU: the reference's octree,
    rebuilt for every evaluation (bit-identical in Precision.F64_STRICT)."""

    method = GravityMethod.BARNES_HUT

    def __init__(self, g_constant: float, eps2: float, theta: float = 0.5,
                 precision: Precision = Precision.F64):
        super().__init__(g_constant, eps2, precision)
        self.theta = float(theta)

    def configure(self, bodies: "DeviceBodies"):
        _check(load_library().nbody_bodies_set_theta(bodies.handle, self.theta))

    @staticmethod
    def tree_nodes(bodies: "DeviceBodies") -> int:
        """Octree nodes of the last build."""
        c = ctypes.c_int64()
        _check(load_library().nbody_bodies_tree_nodes(bodies.handle, ctypes.byref(c)))
        return c.value


class GpuIntegrator:
    integrator: IntegratorMethod

    def __init__(self, dt: float):
        self.dt = float(dt)

    def step(self, gpu: GpuContext, bodies: DeviceBodies,
             gravity: GpuNewtonParallelGravity, nsteps: int = 1):
        """GpuIntegrator::step, nsteps times in one call."""
        gravity.configure(bodies)
        _check(load_library().nbody_step(
            gpu.handle, bodies.handle, int(self.integrator),
            int(gravity.method), int(gravity.precision), self.dt,
            gravity.g_constant, gravity.eps2, int(nsteps)))


class GpuEulerIntegrator(GpuIntegrator):
    integrator = IntegratorMethod.EULER


# NBODY_STAGE_* (include/nbody.h)
STAGE_EULER, STAGE_VV_DRIFT, STAGE_VV_KICK = 0, 1, 2
STAGE_RK4 = (3, 4, 5, 6)


class GpuStagedIntegrator:
    """A GpuIntegrator that steps through the gravity trait alone, as the
    reference's GpuEulerIntegrator does (src/integrators/gpu/euler.rs:17-43):
    gravity.calculate_accelerations(gpu, bodies) leaves a(x) in the device
    arrays, then nbody_integrate_stage applies the stage.  Velocity Verlet
    evaluates twice per step and RK4 four times, in the reference's order;
    the results are nbody_step's bits (velocity Verlet in F32R: to rounding
    at the steps where the bodies are re-sorted)."""

    SEQUENCE = {
        IntegratorMethod.EULER: (STAGE_EULER,),
        IntegratorMethod.VELOCITY_VERLET: (STAGE_VV_DRIFT, STAGE_VV_KICK),
        IntegratorMethod.RUNGE_KUTTA: STAGE_RK4,
    }

    def __init__(self, integrator: IntegratorMethod, dt: float):
        self.integrator = IntegratorMethod(integrator)
        self.dt = float(dt)

    def step(self, gpu: GpuContext, bodies: DeviceBodies, gravity, nsteps: int = 1):
        lib = load_library()
        for _ in range(int(nsteps)):
            for stage in self.SEQUENCE[self.integrator]:
                gravity.calculate_accelerations(gpu, bodies, download=False)
                _check(lib.nbody_integrate_stage(gpu.handle, bodies.handle, stage,
                                                 self.dt))


class GpuVelocityVerletIntegrator(GpuIntegrator):
    integrator = IntegratorMethod.VELOCITY_VERLET


class GpuRungeKuttaIntegrator(GpuIntegrator):
    integrator = IntegratorMethod.RUNGE_KUTTA


class GpuSimulation:
    """src/simulation/gpu_simulation.rs: owns context, bodies and methods."""

    def __init__(self, bodies: Bodies, params: SimulationParameters,
                 gravity: GravityMethod = GravityMethod.NEWTON_PARALLEL,
                 integrator: IntegratorMethod = IntegratorMethod.EULER,
                 precision: Precision = Precision.F64, device: int = 0):
        self.gpu = GpuContext(device)
        self.gravity = gravity.gpu_create(params, precision)
        self.integrator = integrator.gpu_create(params.time_step)
        self.bodies = DeviceBodies(self.gpu, bodies)
        self.time = 0.0
        self.dt = params.time_step

    def step(self, nsteps: int = 1):
        self.integrator.step(self.gpu, self.bodies, self.gravity, nsteps)
        for _ in range(nsteps):
            self.time += self.dt  # accumulated as at cpu_simulation.rs:37

    def snapshot(self) -> Bodies:
        return self.bodies.download()

    def record(self, nsteps: int, max_stage_bytes: int = 0):
        """Steps nsteps times keeping the positions after every step.

        The SimulationRunner::run loop (src/simulation_runner.rs:28-52) of
        step() + snapshot(), with the snapshots gathered on the device and
        copied out in bulk.  Returns (times, x, y, z): times has nsteps
        entries accumulated as ``time += dt``; x/y/z are (nsteps, n).
        """
        n = len(self.bodies)
        out = [np.empty((nsteps, n)) for _ in range(3)]
        self.gravity.configure(self.bodies)
        _check(load_library().nbody_step_record(
            self.gpu.handle, self.bodies.handle, int(self.integrator.integrator),
            int(self.gravity.method), int(self.gravity.precision),
            self.integrator.dt, self.gravity.g_constant, self.gravity.eps2,
            int(nsteps), *[_p(a) for a in out], int(max_stage_bytes)))
        times = np.empty(nsteps)
        for k in range(nsteps):
            self.time += self.dt
            times[k] = self.time
        return (times, *out)

    def energy_momentum(self):
        out = np.zeros(10)
        _check(load_library().nbody_energy_momentum(
            self.gpu.handle, self.bodies.handle, self.gravity.g_constant,
            self.gravity.eps2, _p(out)))
        return out


# ---- files ------------------------------------------------------------------
def load_bodies(path) -> Bodies:
    """input::load_bodies (src/input.rs:20-33): CSV initial conditions."""
    lib = load_library()
    bpath = os.fsencode(path)
    n = ctypes.c_int64(0)
    _check(lib.nbody_load_initial_conditions(bpath, ctypes.byref(n), *([None] * 7)))
    arrs = [np.empty(n.value) for _ in range(7)]
    _check(lib.nbody_load_initial_conditions(bpath, ctypes.byref(n),
                                             *[_p(a) for a in arrs]))
    return Bodies(*arrs)


def format_f64(v: float) -> str:
    """A float as the snapshot CSV writes it (shortest round trip, ryu layout)."""
    buf = ctypes.create_string_buffer(32)
    _check(load_library().nbody_format_f64(float(v), buf))
    return buf.value.decode()


class SnapshotWriter:
    """SimulationDataWriter (src/output.rs:73-139): .csv or .nbody by extension."""

    def __init__(self, path):
        h = ctypes.c_void_p()
        _check(load_library().nbody_writer_open(os.fsencode(path), ctypes.byref(h)))
        self._h = h

    def append(self, time: float, x, y, z):
        x, y, z = _f64(x), _f64(y), _f64(z)
        if not (len(x) == len(y) == len(z)):
            raise ValueError("x, y, z must have the same length")
        _check(load_library().nbody_writer_append(self._h, float(time), len(x),
                                                  _p(x), _p(y), _p(z)))

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            _check(load_library().nbody_writer_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NBODY_MAGIC = b"NBODY001"
NBODY_RECORD = np.dtype([("time", "<f8"), ("id", "<u8"), ("x", "<f8"),
                         ("y", "<f8"), ("z", "<f8")])


def read_snapshots(path):
    """Reads a .nbody or .csv snapshot file into (times, ids, positions).

    positions has shape (T, N, 3) in f64 -- the reshape of
    gui/storage.py:140-168 / :106-137, without its f32 narrowing.
    """
    path = os.fspath(path)
    if path.lower().endswith(".nbody"):
        with open(path, "rb") as f:
            if f.read(8) != NBODY_MAGIC:
                raise ValueError("Invalid binary format: missing magic header")
            arr = np.frombuffer(f.read(), dtype=NBODY_RECORD)
        t, ids = arr["time"], arr["id"]
        pos = np.stack((arr["x"], arr["y"], arr["z"]), axis=-1)
    elif path.lower().endswith(".csv"):
        with open(path, "rb") as f:
            head = f.readline()
        if not head:
            return np.empty(0), np.empty(0, np.uint64), np.empty((0, 0, 3))
        if head.strip() != b"time,id,x,y,z":
            raise ValueError("unexpected CSV header")
        raw = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2,
                         dtype=np.float64)
        t, ids, pos = raw[:, 0], raw[:, 1].astype(np.uint64), raw[:, 2:5]
    else:
        raise ValueError("Unsupported output file format")
    ut = np.unique(t)
    uid = np.unique(ids)
    if len(t) != len(ut) * len(uid):
        raise ValueError("records do not form a (time, id) grid")
    return ut, uid, pos.reshape(len(ut), len(uid), 3)


# ---- tuning -----------------------------------------------------------------
TUNING_BUCKETS = 32


def tuning_split() -> list:
    """Source-range parts per tile by bucket floor(log2 n) (process-wide)."""
    arr = (ctypes.c_int * TUNING_BUCKETS)()
    _check(load_library().nbody_tuning_get_split(arr, TUNING_BUCKETS))
    return list(arr)


def tuning_variant(precision: "Precision") -> list:
    """Force-kernel variant per size bucket for one precision (0..8)."""
    arr = (ctypes.c_int * 32)()
    _check(load_library().nbody_tuning_get_variant(int(precision), arr, 32))
    return list(arr)


def set_tuning_variant(precision: "Precision", variants) -> None:
    """Replaces one precision's variant table (systems created afterwards)."""
    v = list(variants) + [0] * (32 - len(variants))
    arr = (ctypes.c_int * 32)(*v[:32])
    _check(load_library().nbody_tuning_set_variant(int(precision), arr, 32))


def set_tuning_split(parts) -> None:
    """Replace the table; applies to systems uploaded afterwards."""
    arr = (ctypes.c_int * TUNING_BUCKETS)(*parts)
    _check(load_library().nbody_tuning_set_split(arr, TUNING_BUCKETS))


# ---- scenarios ------------------------------------------------------------
def _scenario(fn, n, *args) -> Bodies:
    arrs = [np.empty(n) for _ in range(7)]
    _check(fn(n, *args, *[_p(a) for a in arrs]))
    return Bodies(*arrs)


def plummer(n: int, seed: int, a: float = 1.0, total_mass: float = 1.0,
            g: float = 1.0) -> Bodies:
    return _scenario(load_library().nbody_scenario_plummer, n, seed, a, total_mass, g)


def disc_galaxy(n: int, seed: int, radius: float = 20.0, body_mass: float = 1e-3,
                density_power: float = 1.5, softening: float = 0.5,
                g: float = 1.0) -> Bodies:
    """gui/generators.py:101-153 generate_disc_galaxy (own seeded RNG)."""
    return _scenario(load_library().nbody_scenario_disc, n, seed, radius, body_mass,
                     density_power, softening, g)


def star_system(n: int, seed: int, central_mass: float = 1.0, m_lo: float = 1e-7,
                m_hi: float = 1e-4, r_lo: float = 0.5, r_hi: float = 10.0,
                max_incl_deg: float = 5.0, g: float = 1.0) -> Bodies:
    return _scenario(load_library().nbody_scenario_star_system, n, seed,
                     central_mass, m_lo, m_hi, r_lo, r_hi, max_incl_deg, g)


def colliding(n: int, seed: int, separation: float = 10.0, rel_speed: float = 0.5,
              g: float = 1.0) -> Bodies:
    return _scenario(load_library().nbody_scenario_colliding, n, seed,
                     separation, rel_speed, g)


# ---- batched form -----------------------------------------------------------
class Batch:
    """Many independent systems in one arena, stepped by graph-captured launches."""

    def __init__(self, ctx: GpuContext, systems, integrators, dts, gs, eps2s,
                 precision: Precision = Precision.F32C):
        lib = load_library()
        self.ctx = ctx
        self.n = np.array([len(b) for b in systems], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.n)[:-1]]).astype(np.int64)
        packed = [np.concatenate([getattr(b, f) for b in systems]) for f in Bodies.fields()]
        integ = np.ascontiguousarray(integrators, dtype=np.int32)
        dts, gs, eps2s = (_f64(np.broadcast_to(v, len(systems))) for v in (dts, gs, eps2s))
        h = ctypes.c_void_p()
        _check(lib.nbody_batch_create(
            ctx.handle, len(systems), self.n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            integ.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(dts), _p(gs), _p(eps2s),
            int(precision), *[_p(a) for a in packed], ctypes.byref(h)))
        self._h = h
        self.total = int(self.n.sum())

    def step(self, nsteps: int):
        _check(load_library().nbody_batch_step(self.ctx.handle, self._h, int(nsteps)))

    def download(self):
        arrs = [np.empty(self.total) for _ in range(6)]
        _check(load_library().nbody_batch_download(self.ctx.handle, self._h,
                                                   *[_p(a) for a in arrs]))
        return arrs

    def energy(self) -> np.ndarray:
        out = np.zeros((len(self.n), 10))
        _check(load_library().nbody_batch_energy(self.ctx.handle, self._h, _p(out)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            load_library().nbody_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- served form ------------------------------------------------------------
class Server:
    """Continuous-batching request server (thread-safe)."""

    def __init__(self, capacity_bodies: int, max_systems: int = 1024,
                 snapshot_every: int = 10, precision: Precision = Precision.F32C,
                 device: int = 0, max_snapshot_bytes: int = 1 << 30):
        lib = load_library()
        cfg = ServerConfig(device, int(precision), capacity_bodies, max_systems,
                           snapshot_every, max_snapshot_bytes)
        h = ctypes.c_void_p()
        _check(lib.nbody_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def submit(self, bodies: Bodies, integrator: IntegratorMethod, dt: float,
               steps: int, g: float = 1.0, eps2: float = 0.0,
               want_snapshots: bool = False,
               gravity: GravityMethod = GravityMethod.NEWTON_PARALLEL) -> int:
        arrs = bodies.arrays()
        req = Request(len(bodies), int(integrator), int(gravity), g, dt, eps2, steps,
                      int(want_snapshots), *[_p(a) for a in arrs])
        rid = ctypes.c_uint64()
        _check(load_library().nbody_server_submit(self._h, ctypes.byref(req),
                                                  ctypes.byref(rid)))
        return rid.value

    def poll(self, timeout_ms: int = -1, capacity: int = 0):
        """Returns (event dict, arrays) or None on timeout."""
        lib = load_library()
        ev = Event()
        while True:
            arrs = [np.empty(max(capacity, 1)) for _ in range(6)]
            rc = lib.nbody_server_poll(self._h, timeout_ms, ctypes.byref(ev),
                                       *[_p(a) for a in arrs], capacity)
            if rc == E_TIMEOUT:
                return None
            if rc == E_NOSPACE:
                capacity = int(ev.n)
                continue
            _check(rc)
            break
        n = int(ev.n)
        data = None
        if ev.kind == EVENT_SNAPSHOT:
            data = [a[:n].copy() for a in arrs[:3]]
        elif ev.kind == EVENT_DONE:
            data = [a[:n].copy() for a in arrs]
        return ({"kind": ev.kind, "id": ev.request_id, "step": ev.step, "n": n,
                 "t_submit": ev.t_submit, "t_admit": ev.t_admit, "t_event": ev.t_event,
                 "message": ev.message.decode(errors="replace")}, data)

    def cancel(self, request_id: int):
        _check(load_library().nbody_server_cancel(self._h, request_id))

    def stats(self) -> dict:
        out = np.zeros(8)
        _check(load_library().nbody_server_stats(self._h, _p(out)))
        keys = ("submitted", "admitted", "completed", "resident", "queued",
                "shared_steps", "interactions", "failed")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            load_library().nbody_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
