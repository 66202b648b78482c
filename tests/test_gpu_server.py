"""GPU tests of the served form (continuous batching).

Each result must be the same alone and in the mix: every system is stepped,
re-centred, snapshotted and retired at the same own-step counts whatever
else shares its launches, so alone and mixed results are compared bit for bit
(stronger than the 1e-5 acceleration tolerance of the f32 form).
"""
import time

import numpy as np
import pytest

import nbody_amd as nb
import oracle

pytestmark = pytest.mark.gpu


def drain(srv, ids, timeout_s=120.0):
    done, snaps = {}, {i: [] for i in ids}
    t_end = time.time() + timeout_s
    while len(done) < len(ids):
        assert time.time() < t_end, "server did not finish in time"
        r = srv.poll(timeout_ms=1000, capacity=1 << 16)
        if r is None:
            continue
        ev, data = r
        if ev["kind"] == nb.EVENT_SNAPSHOT:
            assert ev["id"] not in done, "snapshot after the request's final event"
            snaps[ev["id"]].append((ev["step"], data))
        else:
            done[ev["id"]] = (ev, data)
    return done, snaps


def requests(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(32), np.log(900))))
        b = nb.plummer(n, 7000 + k) if k % 2 == 0 else nb.star_system(n, 7000 + k)
        out.append(dict(bodies=b, integrator=nb.IntegratorMethod(k % 3),
                        dt=(1e-3, 5e-4)[k % 2], steps=int(rng.integers(10, 60)),
                        eps2=1e-4))
    return out


def run_server(reqs, precision, snapshots=True, capacity=1 << 16):
    srv = nb.Server(capacity_bodies=capacity, max_systems=256, snapshot_every=10,
                    precision=precision)
    ids = [srv.submit(r["bodies"], r["integrator"], r["dt"], r["steps"],
                      eps2=r["eps2"], want_snapshots=snapshots) for r in reqs]
    done, snaps = drain(srv, ids)
    stats = srv.stats()
    srv.close()
    return ids, done, snaps, stats


@pytest.mark.parametrize("precision", [nb.Precision.F64_STRICT, nb.Precision.F32C,
                                       nb.Precision.F32R])
def test_alone_equals_mixed(gpu, precision):
    reqs = requests(20, 1)
    ids, done, snaps, stats = run_server(reqs, precision)
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    for k in (0, 7, 13, 19):
        aid, adone, asnaps, _ = run_server([reqs[k]], precision)
        ev_m, data_m = done[ids[k]]
        ev_a, data_a = adone[aid[0]]
        assert ev_m["kind"] == ev_a["kind"] == nb.EVENT_DONE
        assert ev_m["step"] == ev_a["step"] == reqs[k]["steps"]
        for q in range(6):
            assert np.array_equal(data_m[q], data_a[q]), (k, q)
        assert [s for s, _ in snaps[ids[k]]] == [s for s, _ in asnaps[aid[0]]]
        for (_, dm), (_, da) in zip(snaps[ids[k]], asnaps[aid[0]]):
            for q in range(3):
                assert np.array_equal(dm[q], da[q])


def test_strict_server_matches_oracle(gpu):
    reqs = requests(6, 2)
    ids, done, snaps, _ = run_server(reqs, nb.Precision.F64_STRICT)
    for rid, r in zip(ids, reqs):
        b = r["bodies"]
        ref = oracle.integrate(int(r["integrator"]), *b.arrays(), 1.0, 0.01, r["dt"],
                               r["steps"])
        ev, data = done[rid]
        for q in range(6):
            assert np.array_equal(data[q], ref[q])
        # snapshots every 10 own steps, in order, strictly before the end
        steps = [s for s, _ in snaps[rid]]
        assert steps == list(range(10, r["steps"], 10))[:len(steps)]
        assert steps == sorted(steps)


@pytest.mark.parametrize("precision", [nb.Precision.F32C, nb.Precision.F32R])
def test_f32_server_matches_oracle(gpu, precision):
    """Served results of the f32-compensated forms against the oracle's f64
    trajectories of the same requests (every integrator, N 32..900, 10 to 60
    steps): final positions within 1e-9 of the system's extent, final
    velocities within 1e-7 of its largest speed (measured: up to 1.0e-11
    and 7.2e-9)."""
    reqs = requests(6, 2)
    ids, done, _, _ = run_server(reqs, precision, snapshots=False)
    for rid, r in zip(ids, reqs):
        b = r["bodies"]
        ref = oracle.integrate(int(r["integrator"]), *b.arrays(), 1.0, 0.01, r["dt"],
                               r["steps"])
        ev, data = done[rid]
        assert ev["kind"] == nb.EVENT_DONE and ev["step"] == r["steps"]
        extent = max(np.abs(ref[q]).max() for q in range(3))
        speed = max(np.abs(ref[q]).max() for q in range(3, 6))
        dpos = max(np.abs(data[q] - ref[q]).max() for q in range(3)) / extent
        dvel = max(np.abs(data[q] - ref[q]).max() for q in range(3, 6)) / speed
        print(f"{precision.name} n={len(b)} {r['integrator'].name} {r['steps']} steps: "
              f"position {dpos:.2e} velocity {dvel:.2e}", flush=True)
        assert dpos <= 1e-9 and dvel <= 1e-7, (len(b), dpos, dvel)


def test_admission_by_memory_and_backlog(gpu):
    # arena of 1024 bodies: requests of 600 bodies cannot be resident together
    reqs = [dict(bodies=nb.plummer(600, 30 + k), integrator=nb.IntegratorMethod.EULER,
                 dt=1e-3, steps=20, eps2=1e-4) for k in range(3)]
    ids, done, _, stats = run_server(reqs, nb.Precision.F32C, snapshots=False,
                                     capacity=1024)
    assert stats["completed"] == 3
    evs = [done[i][0] for i in ids]
    # served one after another: each admitted no earlier than the previous ended
    for a, b in zip(evs, evs[1:]):
        assert b["t_admit"] >= a["t_event"] - 1e-3


def test_reject_oversized_and_bad_requests(gpu):
    srv = nb.Server(capacity_bodies=256, max_systems=4)
    with pytest.raises(nb.NbodyError) as e:
        srv.submit(nb.plummer(1000, 1), nb.IntegratorMethod.EULER, 1e-3, 10)
    assert e.value.code == nb.E_NOSPACE
    b = nb.plummer(16, 1)
    with pytest.raises(nb.NbodyError) as e:
        srv.submit(b, nb.IntegratorMethod.EULER, 0.0, 10)
    assert e.value.code == nb.E_INVALID
    bad = b.copy()
    bad.pos_x[3] = np.nan
    with pytest.raises(nb.NbodyError) as e:
        srv.submit(bad, nb.IntegratorMethod.EULER, 1e-3, 10)
    assert e.value.code == nb.E_INVALID
    with pytest.raises(nb.NbodyError) as e:
        srv.submit(b, nb.IntegratorMethod.EULER, 1e-3, 10,
                   gravity=nb.GravityMethod.BARNES_HUT)
    assert e.value.code == nb.E_UNSUPPORTED
    srv.close()


def test_cancel(gpu):
    srv = nb.Server(capacity_bodies=4096, max_systems=8, snapshot_every=10)
    big = srv.submit(nb.plummer(2000, 3), nb.IntegratorMethod.RUNGE_KUTTA, 1e-3, 100000)
    queued = srv.submit(nb.plummer(3000, 4), nb.IntegratorMethod.EULER, 1e-3, 10)
    srv.cancel(queued)     # cannot fit beside `big`: still queued
    time.sleep(0.2)
    srv.cancel(big)        # resident: retired at the next epoch boundary
    kinds = {}
    t_end = time.time() + 60
    while len(kinds) < 2 and time.time() < t_end:
        r = srv.poll(timeout_ms=500)
        if r and r[0]["kind"] in (nb.EVENT_CANCELLED, nb.EVENT_DONE):
            kinds[r[0]["id"]] = r[0]["kind"]
    assert kinds == {big: nb.EVENT_CANCELLED, queued: nb.EVENT_CANCELLED}
    with pytest.raises(nb.NbodyError) as e:
        srv.cancel(12345)
    assert e.value.code == nb.E_NOTFOUND
    srv.close()


def served_workload(count, seed):
    """Requests of the served workload's shape: Plummer or star systems with
    N log-uniform in [32, 16384], 10 to 1000 steps, any integrator, dt 1e-3
    or 5e-4."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(32), np.log(16384))))
        b = nb.plummer(n, 9000 + k) if k % 2 == 0 else nb.star_system(n, 9000 + k)
        out.append(dict(bodies=b, integrator=nb.IntegratorMethod(int(rng.integers(0, 3))),
                        dt=float(rng.choice([1e-3, 5e-4])),
                        steps=int(rng.integers(10, 1001)), eps2=1e-4))
    return out


@pytest.mark.parametrize("precision", [nb.Precision.F32C, nb.Precision.F32R])
def test_served_workload_full_shape(gpu, precision):
    """60 requests of the served workload submitted at once into a 131072-body
    arena (admission waits for memory, epochs are cut at completions): every
    snapshot arrives in order before its DONE, and the largest, the longest
    and a middle request end bit-identical to the same request served alone."""
    reqs = served_workload(60, 5)
    ids, done, snaps, stats = run_server(reqs, precision, capacity=1 << 17)
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    for rid, r in zip(ids, reqs):
        ev, _ = done[rid]
        assert ev["kind"] == nb.EVENT_DONE and ev["step"] == r["steps"]
        assert [s for s, _ in snaps[rid]] == list(range(10, r["steps"], 10))
    sizes = [r["bodies"].masses.shape[0] for r in reqs]
    picks = {int(np.argmax(sizes)), int(np.argmax([r["steps"] for r in reqs])), 30}
    for k in sorted(picks):
        aid, adone, asnaps, _ = run_server([reqs[k]], precision, capacity=1 << 17)
        data_m, data_a = done[ids[k]][1], adone[aid[0]][1]
        for q in range(6):
            assert np.array_equal(data_m[q], data_a[q]), (k, q)
        for (_, dm), (_, da) in zip(snaps[ids[k]], asnaps[aid[0]]):
            for q in range(3):
                assert np.array_equal(dm[q], da[q]), k


def test_served_workload_strict_matches_oracle(gpu):
    """The served workload's request shape in the strict form: 24 requests at
    once; the largest and the smallest of those the oracle replays in
    seconds (at most 3e10 pair terms) equal its trajectory bit for bit."""
    reqs = served_workload(24, 6)
    ids, done, _, stats = run_server(reqs, nb.Precision.F64_STRICT, snapshots=False,
                                     capacity=1 << 17)
    assert stats["completed"] == len(reqs)
    sizes = [r["bodies"].masses.shape[0] for r in reqs]
    cheap = [k for k, r in enumerate(reqs) if sizes[k] ** 2 * r["steps"] * 4 <= 3e10]
    for k in {max(cheap, key=lambda q: sizes[q]), min(cheap, key=lambda q: sizes[q])}:
        r = reqs[k]
        ref = oracle.integrate(int(r["integrator"]), *r["bodies"].arrays(), 1.0, 0.01,
                               r["dt"], r["steps"], min(16, oracle.default_threads()))
        data = done[ids[k]][1]
        for q in range(6):
            assert np.array_equal(data[q], ref[q]), (k, q)


def test_admission_replays_placement(gpu):
    """Admission must place a pass's requests as the upload will.  A 640-body
    arena takes 384 + 64 + 192 bodies; the 384 and the 192 finish first while
    the 64 runs on, leaving free ranges of 384 and 192.  Of the waiting
    requests, 256 takes the front of the 384 range; 320 then fits in no range
    (though 320 bodies are free in total and it fit before the 256 was
    placed), so it must wait for the 256 to finish, not fail."""
    def req(n, steps, seed):
        return dict(bodies=nb.plummer(n, seed), integrator=nb.IntegratorMethod.EULER,
                    dt=1e-3, steps=steps, eps2=1e-4)
    reqs = [req(384, 10, 1), req(64, 200, 2), req(192, 10, 3), req(256, 10, 4),
            req(320, 10, 5)]
    ids, done, _, stats = run_server(reqs, nb.Precision.F32C, snapshots=False,
                                     capacity=640)
    for rid, r in zip(ids, reqs):
        ev, _ = done[rid]
        assert ev["kind"] == nb.EVENT_DONE, (r["bodies"].masses.shape[0], ev["message"])
    assert stats["completed"] == len(reqs) and stats["failed"] == 0


def test_served_workload_cancel_under_backpressure(gpu):
    """40 requests of the served workload's shape with a 4 MB cap on queued
    snapshot bytes (the worker pauses until the client drains) and every
    fifth request cancelled, half at once and half later.  Each request ends
    with one final event and no snapshot after it; every finished request has
    all its snapshots in order; uncancelled requests equal their runs alone."""
    reqs = served_workload(40, 7)
    srv = nb.Server(capacity_bodies=1 << 17, max_systems=256, snapshot_every=10,
                    precision=nb.Precision.F32C, max_snapshot_bytes=4 << 20)
    ids = [srv.submit(r["bodies"], r["integrator"], r["dt"], r["steps"],
                      eps2=r["eps2"], want_snapshots=True) for r in reqs]
    cancel_now = ids[0::10]
    cancel_later = ids[5::10]

    def cancel(rid):   # a request already retired is no longer known
        try:
            srv.cancel(rid)
        except nb.NbodyError as e:
            assert e.code == nb.E_NOTFOUND

    for rid in cancel_now:
        cancel(rid)
    final, snaps = {}, {rid: [] for rid in ids}
    polled = 0
    t_end = time.time() + 240
    while len(final) < len(ids):
        assert time.time() < t_end, "server did not finish in time"
        if polled < 40:
            time.sleep(0.01)   # a slow reader first: the snapshot cap fills
        r = srv.poll(timeout_ms=1000, capacity=1 << 16)
        if r is None:
            continue
        polled += 1
        if polled == 60:
            for rid in cancel_later:
                if rid not in final:
                    cancel(rid)
        ev, data = r
        rid = ev["id"]
        assert rid not in final, "event after the request's final event"
        if ev["kind"] == nb.EVENT_SNAPSHOT:
            snaps[rid].append((ev["step"], data))
        else:
            final[rid] = (ev, data)
    stats = srv.stats()
    srv.close()
    assert stats["failed"] == 0
    cancelled = set(cancel_now) | set(cancel_later)
    for rid, r in zip(ids, reqs):
        ev, _ = final[rid]
        steps = [s for s, _ in snaps[rid]]
        assert steps == list(range(10, 10 * len(steps) + 1, 10))[:len(steps)]
        if ev["kind"] == nb.EVENT_DONE:
            assert ev["step"] == r["steps"]
            assert steps == list(range(10, r["steps"], 10))
        else:
            assert ev["kind"] == nb.EVENT_CANCELLED and rid in cancelled
    for k in (1, 13, 27):      # never cancelled
        assert final[ids[k]][0]["kind"] == nb.EVENT_DONE
        aid, adone, asnaps, _ = run_server([reqs[k]], nb.Precision.F32C,
                                           capacity=1 << 17)
        for q in range(6):
            assert np.array_equal(final[ids[k]][1][q], adone[aid[0]][1][q]), (k, q)
        for (_, dm), (_, da) in zip(snaps[ids[k]], asnaps[aid[0]]):
            for q in range(3):
                assert np.array_equal(dm[q], da[q]), k
