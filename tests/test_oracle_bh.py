"""Pins of the Barnes-Hut oracle (oracle/oracle_bh.c).

* the reference's own octree tests (src/gravity/cpu/barnes_hut/octree.rs:
  581-699): node counts, leaves, centres of mass of 1-, 2- and 3-body trees;
* Morton codes: the worked examples of docs/barnes-hut-octree.md and the
  loop-based encoder the reference keeps beside the magic-bits one
  (octree.rs:543-554), here as an independent Python loop;
* the opening criterion width^2/d^2 < theta^2 on configurations whose tree
  and answer follow by hand: a far pair taken as one mass at its centre of
  mass exactly when width/d < theta;
* theta = 0 opens every node: the direct sum to rounding;
* structure on a Plummer sphere: leaves partition the Morton order, node
  mass and centre of mass equal those of the node's bodies, widths halve per
  level, children lie inside their parent in octant order, codes are the
  quantised positions interleaved;
* theta > 0: error against the direct sum grows with theta and stays within
  the usual monopole bounds.
"""
import math

import numpy as np
import pytest

import oracle


def tree(m, x, y, z):
    return oracle.bh_tree(np.array(m, float), np.array(x, float),
                          np.array(y, float), np.array(z, float))


def leaves(t):
    return [i for i in range(t["node_count"]) if t["nchild"][i] == 0]


# ---- the reference's tests ---------------------------------------------------
def test_1_body_tree():
    t = tree([2.0], [1.0], [2.0], [3.0])
    assert t["node_count"] == 1
    assert t["mass"][0] == 2.0
    assert (t["com_x"][0], t["com_y"][0], t["com_z"][0]) == (1.0, 2.0, 3.0)
    assert t["count"][0] == 1 and t["nchild"][0] == 0


def test_2_body_tree():
    t = tree([1.0, 3.0], [0.0, 1.0], [1.0, 0.0], [0.0, 0.0])
    lv = leaves(t)
    assert len(lv) == 2
    found = set()
    for i in lv:
        assert t["count"][i] == 1
        p = (t["com_x"][i], t["com_y"][i], t["com_z"][i])
        if np.allclose(p, (0, 1, 0), atol=1e-12):
            found.add(0)
        elif np.allclose(p, (1, 0, 0), atol=1e-12):
            found.add(1)
    assert found == {0, 1}
    assert t["count"][0] == 2 and t["mass"][0] == 4.0
    assert abs(t["com_x"][0] - 0.75) < 1e-12 and abs(t["com_y"][0] - 0.25) < 1e-12


def test_3_body_tree():
    t = tree([1.0, 2.0, 3.0], [0.0, 1.0, 2.0], [0.0] * 3, [0.0] * 3)
    lv = leaves(t)
    assert len(lv) == 3
    assert sorted(round(t["com_x"][i], 12) for i in lv) == [0.0, 1.0, 2.0]
    assert all(t["count"][i] == 1 for i in lv)
    assert t["count"][0] == 3 and t["mass"][0] == 6.0
    assert abs(t["com_x"][0] - 8.0 / 6.0) < 1e-12


# ---- Morton codes --------------------------------------------------------------
def morton_loop(x, y, z):
    code = 0
    for i in range(21):
        code |= ((z >> i) & 1) << (3 * i + 2)
        code |= ((y >> i) & 1) << (3 * i + 1)
        code |= ((x >> i) & 1) << (3 * i)
    return code


def test_morton_doc_examples():
    assert oracle.morton(5, 3, 6) == 371                  # docs: 0b101110011
    assert oracle.morton(4, 6, 1) == 0b011010100


def test_morton_matches_loop():
    rng = np.random.default_rng(3)
    for x, y, z in rng.integers(0, 1 << 21, (2000, 3)):
        assert oracle.morton(int(x), int(y), int(z)) == morton_loop(int(x), int(y), int(z))
    top = (1 << 21) - 1
    assert oracle.morton(top, top, top) == (1 << 63) - 1
    assert oracle.morton(1 << 21, 0, 0) == 0               # 21 bits only


# ---- opening criterion by hand ---------------------------------------------------
def pair(g, eps2, m, d):
    r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
    return g * m * np.array(d) / (r2 + eps2) ** 1.5


def test_far_pair_taken_as_one_mass():
    # target at the origin; bodies at (10, 0, 0) and (10, 4, 0).  Root width
    # 10.01 (x-extent 10 times 1.001).  Both fall in the level-20 octant of
    # width 5.005 (x >= 5.005, y < 5.005) and split one level down (y against
    # 2.5025).  That node's centre of mass (10, 2.4, 0) is at distance
    # 10.28, so width/d = 0.487.
    m = [1.0, 2.0, 3.0]
    x, y, z = [0.0, 10.0, 10.0], [0.0, 0.0, 4.0], [0.0, 0.0, 0.0]
    g, eps = 1.0, 0.01
    com = np.array([10.0, (3.0 * 4.0) / 5.0, 0.0])
    width = 10.0 * 1.001 / 2
    assert 0.48 < width / np.linalg.norm(com) < 0.49
    t = tree(m, x, y, z)
    assert t["node_count"] == 5            # root, target leaf, pair node, 2 leaves
    ax = oracle.barnes_hut(m, x, y, z, g, eps, theta=0.6)
    got = np.array([ax[0][0], ax[1][0], ax[2][0]])
    want = pair(g, eps * eps, 5.0, com)
    assert np.allclose(got, want, rtol=1e-13, atol=0)
    # theta below width/d: the node is opened, both bodies act singly
    ax = oracle.barnes_hut(m, x, y, z, g, eps, theta=0.4)
    got = np.array([ax[0][0], ax[1][0], ax[2][0]])
    want = pair(g, eps * eps, 2.0, [10.0, 0.0, 0.0]) + pair(g, eps * eps, 3.0,
                                                            [10.0, 4.0, 0.0])
    assert np.allclose(got, want, rtol=1e-13, atol=0)
    assert not np.allclose(want, pair(g, eps * eps, 5.0, com), rtol=1e-3)
    # body 1 sees the target (a leaf) and body 2 (a leaf) singly either way
    for theta in (0.4, 0.6):
        ax = oracle.barnes_hut(m, x, y, z, g, eps, theta=theta)
        want = pair(g, eps * eps, 1.0, [-10.0, 0.0, 0.0]) + pair(g, eps * eps, 3.0,
                                                                 [0.0, 4.0, 0.0])
        assert np.allclose([ax[0][1], ax[1][1], ax[2][1]], want, rtol=1e-13)


def test_coincident_bodies_share_a_level0_leaf():
    t = tree([1.0, 1.0, 1.0], [0.0, 0.0, 5.0], [0.0, 0.0, 1.0], [0.0, 0.0, 2.0])
    lv = leaves(t)
    assert sorted(t["count"][i] for i in lv) == [1, 2]
    # the doubled leaf hangs at the bottom: depth 21 below the root
    deep = [i for i in lv if t["count"][i] == 2][0]
    assert t["width"][deep] == t["width"][0] * 2.0 ** -21
    # its centre of mass is the shared position, so d^2 == 0 for both bodies:
    # they feel only the third body
    ax = oracle.barnes_hut([1.0, 1.0, 1.0], [0.0, 0.0, 5.0], [0.0, 0.0, 1.0],
                           [0.0, 0.0, 2.0], 1.0, 0.1, theta=0.5)
    want = pair(1.0, 0.01, 1.0, [5.0, 1.0, 2.0])
    assert np.allclose([ax[0][0], ax[1][0], ax[2][0]], want, rtol=1e-13)


# ---- theta = 0 and structure -----------------------------------------------------
def plummer(n, seed):
    rng = np.random.default_rng(seed)
    r = 1.0 / np.sqrt(rng.uniform(0.05, 1.0, n) ** (-2.0 / 3.0) - 1.0)
    v = rng.standard_normal((n, 3))
    p = v / np.linalg.norm(v, axis=1)[:, None] * r[:, None]
    return np.full(n, 1.0 / n), p[:, 0], p[:, 1], p[:, 2]


def test_theta_zero_is_direct_sum():
    m, x, y, z = plummer(700, 1)
    bh = np.stack(oracle.barnes_hut(m, x, y, z, 1.0, 0.01, theta=0.0), 1)
    ds = np.stack(oracle.newton(m, x, y, z, 1.0, 0.01), 1)
    err = np.linalg.norm(bh - ds, axis=1) / np.linalg.norm(ds, axis=1)
    assert err.max() < 1e-12


def test_structure_plummer():
    m, x, y, z = plummer(2000, 2)
    m = m * np.linspace(0.5, 1.5, 2000)
    t = tree(m, x, y, z)
    c, n = t["node_count"], 2000
    assert c <= 22 * n + 1
    perm, codes = t["perm"], t["codes"]
    assert sorted(perm) == list(range(n))
    assert np.all(codes[1:] >= codes[:-1])
    # codes are the interleaved quantised positions (octree.rs:470-498)
    mn = [x.min(), y.min(), z.min()]
    w = max(x.max() - x.min(), y.max() - y.min(), z.max() - z.min()) * 1.001
    q = [((a - b) / w * 2097152.0).astype(np.uint64) for a, b in zip((x, y, z), mn)]
    for k in range(0, n, 37):
        i = perm[k]
        assert int(codes[k]) == morton_loop(int(q[0][i]), int(q[1][i]), int(q[2][i]))
    covered = np.zeros(n, int)
    depth = np.zeros(c, int)
    for node in range(c):
        s, cnt = t["start"][node], t["count"][node]
        bm, bx = m[perm[s:s + cnt]], x[perm[s:s + cnt]]
        assert math.isclose(t["mass"][node], bm.sum(), rel_tol=1e-13)
        assert math.isclose(t["com_x"][node], (bm * bx).sum() / bm.sum(),
                            rel_tol=1e-12, abs_tol=1e-14)
        assert t["width"][node] == w * 2.0 ** -depth[node]
        kids = t["child"][node, :t["nchild"][node]]
        if len(kids) == 0:
            covered[s:s + cnt] += 1
            assert cnt == 1 or depth[node] == 21
            continue
        assert cnt > 1
        assert kids[0] == node + 1                       # preorder
        pos = s
        octs = []
        for kid in kids:
            depth[kid] = depth[node] + 1
            assert t["start"][kid] == pos
            pos += t["count"][kid]
            octs.append((int(codes[t["start"][kid]]) >> (3 * (20 - depth[node]))) & 7)
        assert pos == s + cnt and octs == sorted(set(octs))
    assert np.all(covered == 1)


def test_theta_error_bounds():
    """Monopole error grows with theta: median below 0.03 theta^2 and worst
    below 0.8 theta^2 against the direct sum, increasing with theta."""
    m, x, y, z = plummer(3000, 4)
    ds = np.stack(oracle.newton_parallel(m, x, y, z, 1.0, 0.01), 1)
    meds = []
    for theta in (0.1, 0.3, 0.5, 0.8):
        bh = np.stack(oracle.barnes_hut(m, x, y, z, 1.0, 0.01, theta=theta), 1)
        err = np.linalg.norm(bh - ds, axis=1) / np.linalg.norm(ds, axis=1)
        assert np.median(err) < 0.03 * theta ** 2 and err.max() < 0.8 * theta ** 2
        meds.append(np.median(err))
    assert meds == sorted(meds) and meds[0] > 1e-6


def test_sampled_targets_equal_full_evaluation():
    m, x, y, z = plummer(1500, 8)
    full = oracle.barnes_hut(m, x, y, z, 1.0, 0.01, 0.5)
    idx = np.array([0, 7, 1499, 700, 3, 3])
    part = oracle.barnes_hut_targets(m, x, y, z, 1.0, 0.01, 0.5, idx)
    for p, q in zip(full, part):
        assert np.array_equal(p[idx], q)


def test_threads_do_not_change_bits():
    m, x, y, z = plummer(900, 5)
    a1 = oracle.barnes_hut(m, x, y, z, 1.0, 0.01, 0.5, nthreads=1)
    a8 = oracle.barnes_hut(m, x, y, z, 1.0, 0.01, 0.5, nthreads=8)
    for p, q in zip(a1, a8):
        assert np.array_equal(p, q)


def test_integrate_with_barnes_hut_theta0_tracks_direct_sum():
    m, x, y, z = plummer(300, 6)
    v = [np.zeros(300)] * 3
    a = oracle.integrate("velocity-verlet", m, x, y, z, *v, 1.0, 0.01, 1e-3, 5,
                         gravity="barnes-hut", theta=0.0)
    b = oracle.integrate("velocity-verlet", m, x, y, z, *v, 1.0, 0.01, 1e-3, 5)
    for p, q in zip(a, b):
        assert np.allclose(p, q, rtol=0, atol=1e-13)
