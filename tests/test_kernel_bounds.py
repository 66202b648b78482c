"""CPU model of the force kernel's memory-access pattern (no GPU needed).

Replays, for every (n, tile, chunk size), the source regions the kernel walks
(force.h: make_bounds, part_range and the sweep loops, including the prefetch
of one chunk ahead) and checks that every source index stays inside the system's
arena span plus the arena's zeroed slack, that the regions cover the
system's sources exactly once in index order, and that the masked regions
are the only ones that can contain j == i or j >= n.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TYPES = open(os.path.join(ROOT, "n-body-sim_amd", "csrc", "nbody_types.h")).read()
K_PAD = int(re.search(r"kPad = (\d+);", TYPES).group(1))
K_WAVE = int(re.search(r"kWave = (\d+);", TYPES).group(1))
ALIGN = 64  # engine.cpp kAlign: systems start on 64-body boundaries


def bounds(off, n, t0, u):
    e_up = off + (n + u - 1) // u * u
    nfull = off + n // u * u
    b1 = min(t0 + K_WAVE, e_up)
    c1 = max(b1, nfull)
    return [(off, t0, False), (t0, b1, True), (b1, c1, False), (c1, e_up, True)]


def clip_part(regions, off, n, u, nsplit, part):
    """force.h part_range: restrict regions to one part's chunk run."""
    if nsplit <= 1:
        return regions
    nchunks = (n + u - 1) // u
    lo = off + (nchunks * part // nsplit) * u
    hi = off + (nchunks * (part + 1) // nsplit) * u
    out = []
    for a, e, m in regions:
        a, e = max(a, lo), min(e, hi)
        out.append((a, max(a, e), m))
    return out


def accessed(j0, j1, u, prefetch):
    if j0 >= j1:
        return []
    idx = list(range(j0, j1))
    if prefetch:  # sweep_f64 / sweep_f32 read the chunk after each chunk
        idx += list(range(j1, j1 + u))
    return idx


@pytest.mark.parametrize("u,prefetch", [(4, False), (4, True), (8, True)])
def test_source_indices_in_bounds_and_ordered(u, prefetch):
    for n in list(range(1, 140)) + [255, 256, 257, 1000, 1023, 1024, 1025, 4099]:
        span = (n + ALIGN - 1) // ALIGN * ALIGN
        off = 3 * ALIGN                     # system placed last in the arena
        cap = off + span
        tiles = (n + K_WAVE - 1) // K_WAVE
        for tile in range(tiles):
            t0 = off + tile * K_WAVE
            regions = bounds(off, n, t0, u)
            seen = []
            for j0, j1, masked in regions:
                assert (j0 - off) % u == 0 and (j1 - off) % u == 0
                for j in accessed(j0, j1, u, prefetch):
                    assert 0 <= j < cap + K_PAD, (n, tile, j)
                for j in range(j0, j1):
                    seen.append(j)
                    needs_mask = j >= off + n or t0 <= j < t0 + K_WAVE
                    if needs_mask:
                        assert masked, (n, tile, j)
            # every source visited once, in index order (reference j order)
            assert seen == list(range(off, off + (n + u - 1) // u * u))
            # the lanes' own loads: il = min(i, end - 1)
            for lane in range(K_WAVE):
                il = min(t0 + lane, off + n - 1)
                assert off <= il < off + n


def test_slack_covers_largest_prefetch():
    assert K_PAD >= 8          # one f32 chunk (U = 8) past a rounded-up end
    assert ALIGN % 8 == 0 and K_WAVE % 8 == 0


@pytest.mark.parametrize("u", [4, 8, 64])
def test_parts_cover_sources_once_in_order(u):
    for n in (513, 1000, 1024, 4096, 8191, 65536 // 16):
        off = 2 * ALIGN
        span = (n + ALIGN - 1) // ALIGN * ALIGN
        tiles = (n + K_WAVE - 1) // K_WAVE
        for nsplit in (1, 2, 3, 4, 8):
            for tile in (0, tiles // 2, tiles - 1):
                t0 = off + tile * K_WAVE
                seen = []
                for part in range(nsplit):
                    for j0, j1, _ in clip_part(bounds(off, n, t0, u), off, n, u,
                                               nsplit, part):
                        assert (j0 - off) % u == 0 and (j1 - off) % u == 0
                        for j in accessed(j0, j1, u, True):
                            assert j < off + span + K_PAD
                        seen.extend(range(j0, j1))
                assert seen == list(range(off, off + (n + u - 1) // u * u))


# ---- MODE_F32R target order (kernels.hip hilbert48 / k_sort_keys) -----------
def spread16(v):
    v &= 0xffff
    v = (v | (v << 16)) & 0x0000ff0000ff
    v = (v | (v << 8)) & 0x00f00f00f00f
    v = (v | (v << 4)) & 0x0c30c30c30c3
    v = (v | (v << 2)) & 0x249249249249
    return v


def hilbert48(a, b, c):
    """Line-for-line model of the device function (Skilling 2004)."""
    X = [a, b, c]
    Q = 1 << 15
    while Q > 1:
        P = Q - 1
        for i in range(3):
            if X[i] & Q:
                X[0] ^= P
            else:
                t = (X[0] ^ X[i]) & P
                X[0] ^= t
                X[i] ^= t
        Q >>= 1
    X[1] ^= X[0]
    X[2] ^= X[1]
    t, Q = 0, 1 << 15
    while Q > 1:
        if X[2] & Q:
            t ^= Q - 1
        Q >>= 1
    X = [v ^ t for v in X]
    return (spread16(X[0]) << 2) | (spread16(X[1]) << 1) | spread16(X[2])


def test_hilbert_order_is_a_curve_of_adjacent_cells():
    """On a 16^3 grid of 16-bit coordinates the index is a bijection onto
    0..4095 (in its top 12 bits) and consecutive indices are face-adjacent
    cells: runs of bodies along it stay spatially compact."""
    cells = {}
    step = 1 << 12
    for a in range(0, 1 << 16, step):
        for b in range(0, 1 << 16, step):
            for c in range(0, 1 << 16, step):
                cells[hilbert48(a, b, c) >> 36] = (a // step, b // step, c // step)
    keys = sorted(cells)
    assert keys == list(range(4096))
    for k0, k1 in zip(keys, keys[1:]):
        p, q = cells[k0], cells[k1]
        assert sum(abs(u - v) for u, v in zip(p, q)) == 1, (k0, p, q)


def test_sort_keys_leave_room_for_the_rank():
    top = (1 << 16) - 1
    assert hilbert48(top, top, top) < (1 << 48)
    assert hilbert48(0, 0, 0) == 0


# ---- work lists expanded on the device (kernels.hip k_expand) ---------------
def expand(spans, total):
    """Model of k_expand: item i -> the last span with first <= i (binary
    search), then (tile, part) = divmod(i - first, parts)."""
    out = []
    for i in range(total):
        lo, hi = 0, len(spans) - 1
        while lo < hi:
            mid = (lo + hi + 1) >> 1
            if spans[mid][0] <= i:
                lo = mid
            else:
                hi = mid - 1
        first, sys, tiles, parts = spans[lo]
        k = i - first
        assert 0 <= k < tiles * parts          # inside its span
        out.append((sys, k // parts, k % parts))
    return out


def test_expand_matches_per_system_item_order():
    """Spans laid out back to back expand to the host's per-system loop
    (tiles outer, parts inner), including one-tile and one-part systems."""
    systems = [(3, 5, 4), (0, 1, 1), (7, 2, 32), (1, 9, 1), (4, 1, 8)]
    spans, total, ref = [], 0, []
    for sys, tiles, parts in systems:
        spans.append((total, sys, tiles, parts))
        total += tiles * parts
        ref += [(sys, t, q) for t in range(tiles) for q in range(parts)]
    assert expand(spans, total) == ref
