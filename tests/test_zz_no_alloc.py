"""The allocation-free force evaluation (nbody_calculate_accelerations_noalloc)
against the allocating entry point it mirrors (nbody_calculate_accelerations),
through the C ABI, on caller-owned device arrays.

Three calls in a row share the library's static scratch array: the largest
size it serves, then smaller systems (the second call must not see the
first's leftovers).  Each result must equal the allocating path's bit for bit
(the same kernels, split and order); strict f64 must also equal the oracle's.
"""
import ctypes

import numpy as np
import pytest

import nbody_amd as nb
import oracle

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int]
    return _hip


class DeviceArray:
    """n doubles of device memory owned by the caller (the test)."""

    def __init__(self, n):
        self.n = n
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), 8 * n) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert len(a) == self.n
        assert hip().hipMemcpy(self.p, a.ctypes.data, 8 * self.n, 1) == 0

    def download(self):
        out = np.empty(self.n)
        assert hip().hipMemcpy(out.ctypes.data, self.p, 8 * self.n, 2) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def noalloc(gpu, b, g, eps2, prec):
    n = len(b)
    ins = [DeviceArray(n) for _ in range(4)]
    outs = [DeviceArray(n) for _ in range(3)]
    for d, a in zip(ins, (b.masses, b.pos_x, b.pos_y, b.pos_z)):
        d.upload(a)
    for d in outs:   # poison the outputs: every element must be written
        d.upload(np.full(n, np.nan))
    lib = nb.load_library()
    rc = lib.nbody_calculate_accelerations_noalloc(
        gpu.handle, n, *[d.p for d in ins], g, eps2, int(prec), *[d.p for d in outs])
    assert rc == 0, nb.load_library().nbody_last_error()
    gpu.synchronize()
    got = tuple(d.download() for d in outs)
    for d in ins + outs:
        d.free()
    return got


def allocating(gpu, b, g, eps2, prec):
    db = nb.DeviceBodies(gpu, b)
    out = nb.GpuNewtonParallelGravity(g, eps2, prec).calculate_accelerations(gpu, db)
    db.close()
    return out


@pytest.mark.parametrize("prec", [nb.Precision.F64_STRICT, nb.Precision.F64,
                                  nb.Precision.F32C])
def test_noalloc_equals_allocating_path(gpu, prec):
    g, eps = 1.0, 0.005
    sizes = (65536, 5000, 700) if prec != nb.Precision.F64_STRICT else (20000, 5000, 700)
    for q, n in enumerate(sizes):
        b = nb.plummer(n, 90 + q)
        print(f"noalloc {prec.name} call {q}: n={n}", flush=True)
        got = noalloc(gpu, b, g, eps * eps, prec)
        want = allocating(gpu, b, g, eps * eps, prec)
        for c in range(3):
            assert np.array_equal(got[c], want[c]), (prec.name, n, c)
        if prec == nb.Precision.F64_STRICT and n <= 5000:
            ref = oracle.newton(b.masses, b.pos_x, b.pos_y, b.pos_z, g, eps)
            for c in range(3):
                assert np.array_equal(got[c], ref[c]), (n, c)


def test_noalloc_rejects_what_it_cannot_serve(gpu):
    lib = nb.load_library()
    null = ctypes.c_void_p(0)
    one = DeviceArray(8)
    p = one.p
    # beyond the scratch array's maximum, and F32R: refused before any launch
    assert lib.nbody_calculate_accelerations_noalloc(
        gpu.handle, 65537, p, p, p, p, 1.0, 0.0, int(nb.Precision.F64), p, p, p) == nb.E_INVALID
    assert b"65536" in lib.nbody_last_error()
    assert lib.nbody_calculate_accelerations_noalloc(
        gpu.handle, 8, p, p, p, p, 1.0, 0.0, int(nb.Precision.F32R), p, p, p) == nb.E_UNSUPPORTED
    assert lib.nbody_calculate_accelerations_noalloc(
        gpu.handle, 8, null, p, p, p, 1.0, 0.0, int(nb.Precision.F64), p, p, p) == nb.E_INVALID
    assert lib.nbody_calculate_accelerations_noalloc(
        gpu.handle, 0, p, p, p, p, 1.0, 0.0, int(nb.Precision.F64), p, p, p) == nb.E_INVALID
    one.free()
