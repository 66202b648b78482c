"""Pins the CPU oracle (oracle/oracle.c) without the reference.

The reference ships no golden vectors for this path, and its Rust sources
cannot be built here, so each oracle function is pinned by checks chosen so a
plausible mistake (dropped term, wrong sign or index, transposed operand,
wrong power of r, wrong stage order) fails at least one of them:

* pair term          -- two-body closed form a = G m d / (d^2 + eps^2)^(3/2)
                        against a 50-digit decimal evaluation (sign, G*m, the
                        3/2 power, the direction, eps^2 vs eps); a 3-body
                        brute force with exact rational arithmetic.
* direct sum         -- an independent numpy evaluation (different order and
                        formula arrangement) to 1e-13 relative; Newton's third
                        law; translation invariance; serial == threaded bits;
                        the sampled-target form == the full form.
* j != i skip        -- softening 0 (sun-earth-moon) stays finite.
* Euler / VV / RK4   -- one step of a two-body system against hand-derived
                        values from the closed-form acceleration (stage order,
                        semi-implicit update, 0.5*a*dt^2, the 1-2-2-1 weights);
                        observed convergence orders 1 / 2 / 4 against the
                        analytic circular orbit; VV time reversibility.
* energy             -- two-body closed form and a numpy evaluation.
"""
import math
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest

import oracle

getcontext().prec = 50


def two_body_acc_decimal(m, d, eps):
    d, m, eps = Decimal(d), Decimal(m), Decimal(eps)
    return m * d / (d * d + eps * eps) ** Decimal("1.5")


@pytest.mark.parametrize("d,eps,m", [(2.0, 0.5, 3.0), (0.75, 0.0, 1.0),
                                     (4.0, 0.05, 10.0), (1e-3, 1e-2, 2.5e-6)])
def test_pair_term_closed_form(d, eps, m):
    # body 0 at the origin, body 1 (mass m) at +d on the y axis
    ms = np.array([1.0, m])
    x = np.zeros(2)
    y = np.array([0.0, d])
    ax, ay, az = oracle.newton(ms, x, y, x, 1.0, eps)
    want = float(two_body_acc_decimal(m, d, eps))
    assert ax[0] == 0.0 and az[0] == 0.0
    assert ay[0] == pytest.approx(want, rel=4e-16)      # pulled toward +y
    assert ay[1] == pytest.approx(-float(two_body_acc_decimal(1.0, d, eps)), rel=4e-16)


def test_gravitational_constant_scales_linearly():
    ms, x, y, z = np.array([1.0, 2.0]), np.array([0.0, 1.5]), np.zeros(2), np.zeros(2)
    a1 = oracle.newton(ms, x, y, z, 1.0, 0.1)[0]
    a3 = oracle.newton(ms, x, y, z, 4.0, 0.1)[0]
    assert np.array_equal(a3, 4.0 * a1)   # G*m_j*inv3 with G a power of two


def test_three_body_brute_force_rational():
    # eps chosen so every r^2 + eps^2 is a perfect square: exact rationals.
    pos = [(0, 0, 0), (3, 0, 0), (0, 4, 0)]
    ms = [1.0, 2.0, 3.0]
    eps = 0.0
    arr = np.array(pos, dtype=np.float64)
    ax, ay, az = oracle.newton(np.array(ms), arr[:, 0], arr[:, 1], arr[:, 2], 1.0, eps)
    for i in range(3):
        want = [Fraction(0)] * 3
        for j in range(3):
            if j == i:
                continue
            d = [Fraction(pos[j][k] - pos[i][k]) for k in range(3)]
            r = math.isqrt(int(sum(c * c for c in d)))
            for k in range(3):
                want[k] += Fraction(ms[j]) * d[k] / r ** 3
        got = (ax[i], ay[i], az[i])
        for k in range(3):
            assert got[k] == pytest.approx(float(want[k]), rel=1e-15, abs=0)


def numpy_direct(m, x, y, z, g, eps):
    p = np.stack([x, y, z], 1)
    d = p[None, :, :] - p[:, None, :]          # d[i, j] = p_j - p_i
    r2 = np.einsum("ijk,ijk->ij", d, d) + eps * eps
    np.fill_diagonal(r2, 1.0)
    w = g * m[None, :] / (r2 * np.sqrt(r2))
    np.fill_diagonal(w, 0.0)
    return tuple(np.einsum("ij,ijk->ik", w, d).T)


def random_cloud(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0.5, 2.0, n), *rng.normal(size=(3, n)))


@pytest.mark.parametrize("n,eps", [(257, 0.01), (100, 0.3)])
def test_direct_sum_matches_numpy(n, eps):
    m, x, y, z = random_cloud(n, n)
    got = np.stack(oracle.newton(m, x, y, z, 1.0, eps), 1)
    want = np.stack(numpy_direct(m, x, y, z, 1.0, eps), 1)
    scale = np.linalg.norm(want, axis=1)
    assert (np.linalg.norm(got - want, axis=1) / scale).max() < 1e-12


def test_third_law_and_translation():
    m, x, y, z = random_cloud(300, 5)
    a = np.stack(oracle.newton(m, x, y, z, 1.0, 0.05), 1)
    net = (m[:, None] * a).sum(0)
    assert np.abs(net).max() <= 1e-13 * (m[:, None] * np.abs(a)).sum()
    b = np.stack(oracle.newton(m, x + 1.0, y - 2.0, z + 0.5, 1.0, 0.05), 1)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_serial_parallel_targets_bitwise():
    m, x, y, z = random_cloud(513, 9)
    s = oracle.newton(m, x, y, z, 1.0, 0.01)
    for th in (1, 3, 8):
        p = oracle.newton_parallel(m, x, y, z, 1.0, 0.01, th)
        for q in range(3):
            assert np.array_equal(s[q], p[q])
    idx = np.array([512, 0, 17, 255], dtype=np.int64)
    t = oracle.newton_targets(m, x, y, z, 1.0, 0.01, idx, 2)
    for q in range(3):
        assert np.array_equal(t[q], s[q][idx])


def test_zero_softening_skips_self():
    m = np.array([1.0, 3.003e-06, 3.694e-08])   # sun-earth-moon
    x = np.array([0.0, 1.0, 1.00257])
    a = oracle.newton(m, x, np.zeros(3), np.zeros(3), 1.0, 0.0)
    assert all(np.isfinite(c).all() for c in a)


# ---- integrators ------------------------------------------------------------
def two_body(d=2.0, m=(10.0, 10.0), v=1.118):
    ms = np.array(m)
    return ms, np.array([-d / 2, d / 2]), np.zeros(2), np.zeros(2), \
        np.zeros(2), np.array([-v, v]), np.zeros(2)


def acc_x(x0, x1, m, eps):
    """x-accelerations of a two-body system on the x axis (closed form)."""
    d = x1 - x0
    k = 1.0 / (d * d + eps * eps) ** 1.5
    return m[1] * d * k, -m[0] * d * k


def test_euler_one_step_semi_implicit():
    ms, x, y, z, vx, vy, vz = two_body()
    dt, eps = 0.025, 0.05
    out = oracle.integrate("euler", ms, x, y, z, vx, vy, vz, 1.0, eps, dt, 1)
    a0, a1 = acc_x(x[0], x[1], ms, eps)
    v = [vx[0] + a0 * dt, vx[1] + a1 * dt]
    assert out[3][0] == pytest.approx(v[0], rel=1e-15)
    assert out[0][0] == pytest.approx(x[0] + v[0] * dt, rel=1e-15)   # new v used
    assert out[1][1] == pytest.approx(vy[1] * dt, rel=1e-15)


def test_vv_one_step():
    ms, x, y, z, vx, vy, vz = two_body(v=0.0)   # head-on: stays on the x axis
    vx = np.array([0.3, -0.2])
    dt, eps = 0.025, 0.05
    out = oracle.integrate("velocity-verlet", ms, x, y, z, vx, vy, vz, 1.0, eps, dt, 1)
    a0 = acc_x(x[0], x[1], ms, eps)
    xn = [x[i] + vx[i] * dt + 0.5 * a0[i] * dt * dt for i in range(2)]
    a1 = acc_x(xn[0], xn[1], ms, eps)
    assert out[0][0] == pytest.approx(xn[0], rel=1e-15)
    assert out[3][0] == pytest.approx(vx[0] + 0.5 * (a0[0] + a1[0]) * dt, abs=1e-15)


def test_rk4_one_step():
    ms, x, y, z, vx, vy, vz = two_body(v=0.0)
    vx = np.array([0.1, -0.1])
    dt, eps = 0.05, 0.05
    out = oracle.integrate("runge-kutta", ms, x, y, z, vx, vy, vz, 1.0, eps, dt, 1)
    h = dt / 2

    def f(px):
        return np.array(acc_x(px[0], px[1], ms, eps))
    k1r, k1v = vx, f(x)
    k2r, k2v = vx + h * k1v, f(x + h * k1r)
    k3r, k3v = vx + h * k2v, f(x + h * k2r)
    k4r, k4v = vx + dt * k3v, f(x + dt * k3r)
    xn = x + dt / 6 * (k1r + 2 * k2r + 2 * k3r + k4r)
    vn = vx + dt / 6 * (k1v + 2 * k2v + 2 * k3v + k4v)
    assert np.allclose(out[0], xn, rtol=1e-15, atol=0)
    assert np.allclose(out[3], vn, rtol=1e-14, atol=1e-16)


def circular_orbit_error(integrator, dt, t_end=2.0):
    # two equal masses m=0.5 at separation 1 on a circular orbit, eps = 0:
    # angular speed w = sqrt(G M / d^3) = 1, radius 0.5 each
    ms = np.array([0.5, 0.5])
    x, y = np.array([-0.5, 0.5]), np.zeros(2)
    vx, vy = np.zeros(2), np.array([-0.5, 0.5])
    steps = int(round(t_end / dt))
    out = oracle.integrate(integrator, ms, x, y, np.zeros(2), vx, vy, np.zeros(2),
                           1.0, 0.0, dt, steps)
    t = steps * dt
    exact = (0.5 * math.cos(t), 0.5 * math.sin(t))
    return math.hypot(out[0][1] - exact[0], out[1][1] - exact[1])


@pytest.mark.parametrize("integrator,order", [("euler", 1), ("velocity-verlet", 2),
                                              ("runge-kutta", 4)])
def test_convergence_order(integrator, order):
    e1 = circular_orbit_error(integrator, 0.02)
    e2 = circular_orbit_error(integrator, 0.01)
    observed = math.log2(e1 / e2)
    assert abs(observed - order) < 0.3, observed


def test_vv_time_reversible():
    m, x, y, z = random_cloud(16, 3)
    rng = np.random.default_rng(4)
    v = rng.normal(scale=0.1, size=(3, 16))
    fwd = oracle.integrate("velocity-verlet", m, x, y, z, *v, 1.0, 0.1, 1e-3, 50)
    back = oracle.integrate("velocity-verlet", m, *fwd[:3], *[-a for a in fwd[3:]],
                            1.0, 0.1, 1e-3, 50)
    for a, b in zip(back[:3], (x, y, z)):
        assert np.abs(a - b).max() < 1e-11


def test_binary_example_stays_circular():
    # data/examples/binary-system: m=10 at x=+-2 (separation 4), |v|=1.118,
    # v^2 = G m r / d^2 = 10*2/16 = 1.25: a circular orbit of period
    # 2*pi*2/1.118 ~ 11.24; integrate half a period with its own dt, eps.
    ms, x, y, z, vx, vy, vz = two_body(d=4.0)
    out = oracle.integrate("velocity-verlet", ms, x, y, z, vx, vy, vz, 1.0, 0.05,
                           0.025, 225)
    sep = np.hypot(out[0][1] - out[0][0], out[1][1] - out[1][0])
    assert abs(sep - 4.0) < 0.02
    # after ~half a period the bodies have swapped sides
    assert out[0][0] > 1.9 and out[0][1] < -1.9


def test_energy_two_body_and_numpy():
    ms, x, y, z, vx, vy, vz = two_body()
    K, U = oracle.energy(ms, x, y, z, vx, vy, vz, 1.0, 0.05)
    assert K == pytest.approx(0.5 * 10 * 1.118 ** 2 * 2, rel=1e-15)
    assert U == pytest.approx(-100.0 / math.sqrt(4.0 + 0.0025), rel=1e-15)
    m, x, y, z = random_cloud(200, 8)
    v = np.random.default_rng(1).normal(size=(3, 200))
    K, U = oracle.energy(m, x, y, z, *v, 2.0, 0.1)
    p = np.stack([x, y, z], 1)
    r = np.sqrt(((p[:, None] - p[None]) ** 2).sum(-1) + 0.01)
    iu = np.triu_indices(200, 1)
    assert U == pytest.approx(-2.0 * (m[:, None] * m[None] / r)[iu].sum(), rel=1e-13)
    assert K == pytest.approx(0.5 * (m * (v ** 2).sum(0)).sum(), rel=1e-14)


def test_invalid_integrator_rejected():
    m, x, y, z = random_cloud(4, 1)
    with pytest.raises(ValueError):
        oracle.integrate(7, m, x, y, z, x, y, z, 1.0, 0.0, 1e-3, 1)
