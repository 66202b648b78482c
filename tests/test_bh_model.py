"""CPU model of the GPU octree construction (n-body-sim_amd/csrc/bh.hip).

The kernels build the reference's preorder octree without recursion: the
nodes starting at sorted position k form a chain of levels [bot(k), top(k)]
(k_bh_chain), preorder ids come from a scan of chain lengths, each node's
run end comes from a binary search (k_bh_nodes), and each node finds its
parent by a binary search for the start of its parent's run (k_bh_link).
This replays those formulas in Python and compares every node with the
oracle's recursive build (oracle_bh.c, after octree.rs:257-352), and checks
that every index the kernels write stays inside its array -- on random
spheres, clustered and coincident bodies, and the smallest trees.
"""
import numpy as np
import pytest

import oracle

DEPTH = 21


def hibit(x):
    return int(x).bit_length() - 1


def model(codes):
    codes = [int(c) for c in codes]
    n = len(codes)
    top, bot, cnt = [0] * n, [0] * n, [0] * n
    for k in range(n):                                   # k_bh_chain
        c = codes[k]
        if k == 0:
            t = DEPTH
        else:
            x = codes[k - 1] ^ c
            t = -1 if x == 0 else min(DEPTH - 1, hibit(x) // 3)
        b = 0
        if t >= 0:
            if k == n - 1:
                b = t
            else:
                y = c ^ codes[k + 1]
                b = 0 if y == 0 else min(t, hibit(y) // 3)
        top[k], bot[k] = t, b
        cnt[k] = 0 if t < 0 else t - b + 1
    base = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(int)
    total = int(base[-1] + cnt[-1])
    start = [-1] * total
    count = [-1] * total
    level = [-1] * total
    leaf = [False] * total
    child = [[-1] * 8 for _ in range(total)]

    def run_end(k, b):
        p = codes[k] >> (3 * b)
        lo, hi = k + 1, n
        while lo < hi:
            mid = (lo + hi) // 2
            if (codes[mid] >> (3 * b)) == p:
                lo = mid + 1
            else:
                hi = mid
        return lo

    def run_start(k, b):
        p = codes[k] >> (3 * b)
        lo, hi = 0, k
        while lo < hi:
            mid = (lo + hi) // 2
            if (codes[mid] >> (3 * b)) < p:
                lo = mid + 1
            else:
                hi = mid
        return lo

    for k in range(n):                                   # k_bh_nodes
        t, lo = top[k], bot[k]
        if t < 0:
            continue
        for b in range(t, lo - 1, -1):
            i = base[k] + (t - b)
            assert 0 <= i < total and start[i] == -1     # each id written once
            start[i] = k
            single = b == lo and (k == n - 1 or codes[k + 1] != codes[k])
            count[i] = 1 if single else run_end(k, b) - k
            level[i] = b
            leaf[i] = b == lo
    for k in range(n):                                   # k_bh_link
        t, lo = top[k], bot[k]
        if t < 0:
            continue
        for b in range(min(t, DEPTH - 1), lo - 1, -1):
            i = base[k] + (t - b)
            if b < t:
                parent = i - 1
            else:
                s = run_start(k, b + 1)
                assert top[s] >= b + 1 and bot[s] <= b      # parent node exists
                parent = base[s] + (top[s] - (b + 1))
            assert 0 <= parent < total and not leaf[parent]
            slot = (codes[k] >> (3 * b)) & 7
            assert child[parent][slot] == -1
            child[parent][slot] = i
    return dict(total=total, start=start, count=count, level=level, leaf=leaf,
                child=child)


def compare(m, x, y, z):
    t = oracle.bh_tree(m, x, y, z)
    md = model(t["codes"])
    assert md["total"] == t["node_count"]
    assert md["total"] <= 22 * len(m)
    for i in range(md["total"]):
        assert md["start"][i] == t["start"][i] and md["count"][i] == t["count"][i]
        kids = [c for c in md["child"][i] if c >= 0]
        assert kids == list(t["child"][i, :t["nchild"][i]])
        assert md["leaf"][i] == (t["nchild"][i] == 0)
        # width = root * 2^(level - 21), the reference's repeated halving
        assert t["width"][i] == t["width"][0] * 2.0 ** (md["level"][i] - DEPTH)


def sphere(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, 3)) * rng.uniform(0.01, 3.0, (n, 1))
    return np.full(n, 1.0 / n), p[:, 0], p[:, 1], p[:, 2]


@pytest.mark.parametrize("n", [1, 2, 3, 8, 64, 65, 300, 2000])
def test_model_matches_oracle_tree(n):
    compare(*sphere(n, n))


def test_model_clustered_and_coincident():
    m, x, y, z = sphere(500, 7)
    x[:50] = x[0] + 1e-9 * np.arange(50)       # a tight cluster: deep chains
    y[:50] = y[0]
    z[:50] = z[0]
    x[100:110], y[100:110], z[100:110] = x[99], y[99], z[99]   # coincident
    compare(m, x, y, z)


def test_model_line_and_plane():
    n = 257
    compare(np.ones(n), np.linspace(0, 1, n), np.zeros(n), np.zeros(n))
    g = np.linspace(-1, 1, 16)
    X, Y = np.meshgrid(g, g)
    compare(np.ones(256), X.ravel(), Y.ravel(), np.zeros(256))


def test_walk_stack_bound():
    """The walk's shared stack (kStack = 8*(21+2) entries) holds every
    entry a depth-first walk can leave live: at most 7 pending siblings per
    level on the path plus the 8 children just pushed."""
    assert 1 + 7 * DEPTH + 8 <= 8 * (DEPTH + 2)
