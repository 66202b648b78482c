"""Test-side construction of the snapshot file bytes, independent of the library.

The digits of a float come from Python's ``repr`` (David Gay's shortest
round-trip algorithm); the layout is the Rust ryu crate's (the csv crate the
reference writes with, src/output.rs:106-122, formats every f64 with
``ryu::Buffer::format``).  The library builds the same text from
``std::to_chars`` digits, so agreement checks two digit generators and two
layout implementations against each other.
"""
import math

import numpy as np

NBODY_RECORD = np.dtype([("time", "<f8"), ("id", "<u8"), ("x", "<f8"),
                         ("y", "<f8"), ("z", "<f8")])


def ryu_text(v: float) -> str:
    if math.isnan(v):
        return "NaN"
    sign = "-" if math.copysign(1.0, v) < 0 else ""
    a = abs(v)
    if math.isinf(a):
        return sign + "inf"
    if a == 0.0:
        return sign + "0.0"
    mant, _, exp = repr(a).partition("e")
    ip, _, fp = mant.partition(".")
    digits = (ip + fp).lstrip("0")
    k = (int(exp) if exp else 0) - len(fp)    # value = int(digits) * 10**k
    core = digits.rstrip("0")
    k += len(digits) - len(core)
    L = len(core)
    kk = L + k                                # 10**(kk-1) <= value < 10**kk
    if 0 <= k and kk <= 16:
        body = core + "0" * (kk - L) + ".0"
    elif 0 < kk <= 16:
        body = core[:kk] + "." + core[kk:]
    elif -5 < kk <= 0:
        body = "0." + "0" * (-kk) + core
    elif L == 1:
        body = core + "e" + str(kk - 1)
    else:
        body = core[0] + "." + core[1:] + "e" + str(kk - 1)
    return sign + body


def csv_bytes(times, xs, ys, zs) -> bytes:
    """times: (T,), xs/ys/zs: (T, N) -> the .csv file the reference writes."""
    xs, ys, zs = (np.asarray(a, dtype=np.float64) for a in (xs, ys, zs))
    if xs.size == 0:
        return b""
    lines = ["time,id,x,y,z"]
    for s, t in enumerate(times):
        ts = ryu_text(float(t))
        for i in range(xs.shape[1]):
            lines.append(f"{ts},{i},{ryu_text(float(xs[s, i]))},"
                         f"{ryu_text(float(ys[s, i]))},{ryu_text(float(zs[s, i]))}")
    return ("\n".join(lines) + "\n").encode()


def nbody_bytes(times, xs, ys, zs) -> bytes:
    """The .nbody file: magic then 40-byte little-endian records."""
    xs, ys, zs = (np.asarray(a, dtype=np.float64) for a in (xs, ys, zs))
    T, N = xs.shape if xs.ndim == 2 else (len(times), 0)
    rec = np.zeros(T * N, dtype=NBODY_RECORD)
    rec["time"] = np.repeat(np.asarray(times, dtype=np.float64), N)
    rec["id"] = np.tile(np.arange(N, dtype=np.uint64), T)
    rec["x"], rec["y"], rec["z"] = xs.ravel(), ys.ravel(), zs.ravel()
    return b"NBODY001" + rec.tobytes()
