"""Continuous-batching server: requests admitted while others run give the
same results as each run alone; admission by memory budget; cancellation."""
import math
import time

import numpy as np
import pytest

import coolbubble_amd as cb

pytestmark = pytest.mark.gpu


def req(D, N, L, c, mode, nt, snaps, **kw):
    p = dict(D=D, nGrids=N, L=L, c=c, readxns=mode, dt=0.2 * L / N)
    p.update(kw)
    return p, nt, snaps


REQS = [
    req(2, 64, 8.0, 4.0, 1, 130, 3),
    req(3, 32, 4.0, 2.0, -2, 97, 2),
    req(3, 24, 3.0, 2.0, -3, 61, 2, Trad=0.3, expansion=1, Mpl=10.0),
    req(1, 256, 16.0, 4.0, 1, 88, 2),
    req(2, 45, 5.625, 4.0, 1, 50, 1, nBubbles=3),
    req(2, 128, 16.0, 6.0, -3, 170, 4, Trad=0.2),
    req(3, 32, 4.0, 4.0, 1, 45, 5, nBubbles=4, rand_seed=9),
    # a lattice on which a lone run and a batch take three steps per launch
    # (FP32): the server gives it two
    req(3, 192, 24.0, 4.0, -2, 47, 2),
]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_results_identical_alone_and_in_the_mix(precision):
    sv = cb.Server(precision=precision, max_active=3)
    ids = {}
    for i, (p, nt, snaps) in enumerate(REQS):
        ids[sv.submit(nt, snaps, tag=i, **p)] = i
        time.sleep(0.002 * (i % 3))
    sv.drain()
    results = sv.poll()
    assert len(results) == len(REQS)
    for r in results:
        assert r.status == 0, r.status
        p, nt, snaps = REQS[r.tag]
        alone = cb.Bubble(precision=precision, **p)
        v = cb.server_variant(alone)
        alone.select_kernel(v)            # -1: the general kernel, as the server runs it
        if p["nGrids"] == 192:
            assert alone.variant_info(v)["fuse"] == 2
        rows = [alone.energy_row()]
        spectra = [alone.spectrum()]
        steps_done = 0
        dt = p["dt"]
        T = nt * dt          # SnapshotSchedule, src/main.cpp:201-223
        sched = sorted(set(min(nt, max(1, int(math.floor(T * k / snaps / dt + 0.5))))
                           for k in range(1, snaps + 1)))
        events = sorted(set(list(range(40, nt + 1, 40)) + sched + [nt]))
        for ev in events:
            alone.step(ev - steps_done)
            steps_done = ev
            row = alone.energy_row()
            if ev % 40 == 0:
                rows.append(row)
            if ev in sched:
                spectra.append(alone.spectrum())
        np.testing.assert_array_equal(sv.rows(r.id), np.array(rows))
        np.testing.assert_array_equal(np.array(r.final_row[:]), row)
        assert r.n_spectra == len(spectra)
        for w, (k, _, d) in enumerate(spectra):
            ks, ds = sv.spectrum(r.id, w)
            # bins are summed in a fixed order (no atomics): bit-identical too
            np.testing.assert_array_equal(ks, k)
            np.testing.assert_array_equal(ds, d)
        sv.release(r.id)
        alone.close()
    sv.close()


def test_admission_rejects_what_can_never_fit_and_queues_the_rest():
    sv = cb.Server(precision="fp64", mem_budget=64 << 20, max_active=2)
    big = sv.submit(10, 1, tag=1, D=3, nGrids=256, L=32.0, dt=0.025, readxns=-2)
    small = [sv.submit(30, 1, tag=2 + i, D=2, nGrids=64, L=8.0, dt=0.025, readxns=-2) for i in range(4)]
    sv.drain()
    res = {r.id: r for r in sv.poll()}
    assert res[big].status == 7                       # CB_ADMISSION_REJECTED
    for i in small:
        assert res[i].status == 0
        assert res[i].t_admit >= res[i].t_submit and res[i].t_done > res[i].t_admit
    sv.close()


def test_cancel_queued_request():
    sv = cb.Server(precision="fp32", max_active=1)
    first = sv.submit(400, 1, D=3, nGrids=64, L=8.0, dt=0.025, readxns=-2)
    second = sv.submit(400, 1, D=3, nGrids=64, L=8.0, dt=0.025, readxns=-2)
    sv.cancel(second)
    sv.drain()
    res = {r.id: r for r in sv.poll()}
    assert res[first].status == 0
    assert res[second].status == 8                    # CB_NOT_READY: cancelled
    sv.close()


def test_cancel_active_request_while_others_finish_unchanged():
    """A request cancelled while it runs is retired at its next tick with
    CB_NOT_READY; the requests beside it finish with the results they have
    alone.  Cancelling an id that is neither queued nor active is an error."""
    sv = cb.Server(precision="fp64", max_active=4)
    long_id = sv.submit(100000000, 1, tag=99, D=3, nGrids=64, L=8.0, dt=0.025, readxns=-2)
    others = {sv.submit(nt, snaps, tag=i, **p): i for i, (p, nt, snaps) in enumerate(REQS[:3])}
    for _ in range(2000):                           # wait until the long one is active
        if sv.stats()["active"] >= 1 and sv.stats()["queued"] <= 3:
            break
        time.sleep(0.005)
    time.sleep(0.05)
    sv.cancel(long_id)
    sv.drain()
    res = {r.id: r for r in sv.poll()}
    assert res[long_id].status == 8                 # CB_NOT_READY: cancelled
    assert res[long_id].site_updates == 0
    for rid, i in others.items():
        assert res[rid].status == 0
        p, nt, snaps = REQS[i]
        alone = cb.Bubble(precision="fp64", **p)
        v = cb.server_variant(alone)
        alone.select_kernel(v)            # -1: the general kernel, as the server runs it
        done = 0
        dt = p["dt"]
        T = nt * dt
        sched = sorted(set(min(nt, max(1, int(math.floor(T * k / snaps / dt + 0.5)))) for k in range(1, snaps + 1)))
        for ev in sorted(set(list(range(40, nt + 1, 40)) + sched + [nt])):
            alone.step(ev - done)
            done = ev
        np.testing.assert_array_equal(np.array(res[rid].final_row[:]), alone.energy_row())
        alone.close()
    with pytest.raises(cb.CBError) as e:
        sv.cancel(long_id)                          # finished: neither queued nor active
    assert e.value.status == 1
    with pytest.raises(cb.CBError):
        sv.cancel(123456789)
    sv.close()
