"""The batched-ensemble and serving workloads at their configured sizes.

* Ensemble: the 80 members of the measured ensemble (64 thermal d=3 N=128
  L=16 realisations, c in {1,2,4,8} x 16 seeds, Trad=0.1, and 16 d=2 N=1024
  L=64 five-bubble runs, c ~ U[1,8]), FP32, stepped through graph-captured
  40-step segments: members equal their lone runs bit for bit.  FP64 batches
  of a 128^3 thermal member and a 1024^2 five-bubble member follow the
  oracle (the reference's Bubble ctor and oneStep, src/bubble.cpp:156-316,
  849-876) to 1e-12 in every energy row and in the fields.
* Serving: requests of the serving workload's largest shapes (d=1 N=2^16,
  d=2 N=2048, d=3 N=384) and odd N, run in the mix, equal their lone runs
  bit for bit (rows and spectra); in FP64 their energy rows and spectra
  follow the oracle to 1e-12.
"""
import math

import numpy as np
import pytest

import coolbubble_amd as cb
import oracle
from test_gpu_parity import TOL, field_err, row_errors

pytestmark = pytest.mark.gpu


def ensemble_members(precision):
    rng = np.random.default_rng(2024)
    out = []
    for c in (1.0, 2.0, 4.0, 8.0):
        for s in range(16):
            out.append(dict(D=3, nGrids=128, L=16.0, dt=0.025, c=c, readxns=-3, Trad=0.1,
                            thermal_base=12345 + s, precision=precision))
    for c in rng.uniform(1.0, 8.0, 16):
        out.append(dict(D=2, nGrids=1024, L=64.0, dt=0.0125, c=float(c), readxns=1, nBubbles=5,
                        precision=precision))
    return out


def test_c4_ensemble_fp32_members_equal_lone_runs():
    ms = ensemble_members("fp32")
    print("C4 ensemble: %d members" % len(ms), flush=True)
    b = cb.Batch(ms)
    picks = [0, 21, 63, 64, 79]
    alone = []
    for i in picks:
        s = cb.Bubble(**ms[i])
        s.select_kernel(b.members[i].kernel_variants()[1])
        alone.append(s)
    for chunk in (40, 40, 40, 13):
        b.step(chunk)
        for s in alone:
            s.step(chunk)
        for i, s in zip(picks, alone):
            np.testing.assert_array_equal(b.members[i].energy_row(), s.energy_row())
    for m in b.members:
        assert np.all(np.isfinite(m.energy_row()))
    for i, s in zip(picks, alone):
        pa, qa = s.fields()
        pb, qb = b.members[i].fields()
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(qa, qb)
        s.close()
    b.close()


def test_c4_members_fp64_follow_the_oracle():
    ms = ensemble_members("fp64")
    pick = [ms[19], ms[70]]          # thermal c=2 seed 12348; a 1024^2 five-bubble run
    print("C4 FP64 members vs oracle:", pick, flush=True)
    b = cb.Batch(pick)
    orc = []
    for m in pick:
        kw = {k: v for k, v in m.items() if k != "precision"}
        orc.append(oracle.Sim(**kw))
    for o, v in zip(orc, b.members):
        assert row_errors(o.energy_row()[None], v.energy_row()[None]).max() <= TOL
    for chunk in (40, 7, 40):
        b.step(chunk)
        for o, v in zip(orc, b.members):
            o.step(chunk)
            err = row_errors(o.energy_row()[None], v.energy_row()[None]).max()
            assert err <= TOL, (chunk, err)
    for o, v in zip(orc, b.members):
        phi, pi = v.fields()
        assert field_err(o.phi(), phi) <= TOL
        assert field_err(o.pi(), pi) <= TOL
        o.close()
    b.close()


def _req(D, N, dx, c, mode, nt, snaps, **kw):
    p = dict(D=D, nGrids=N, L=N * dx, c=c, readxns=mode, dt=0.2 * dx)
    p.update(kw)
    return p, nt, snaps


C5_REQS = [
    _req(1, 65536, 0.0625, 4.0, 1, 120, 2),
    _req(2, 2048, 0.0625, 3.0, -1, 60, 2, DeltaR=0.5),
    _req(3, 384, 0.125, 2.0, -3, 40, 1, Trad=0.1),
    _req(3, 65, 0.125, 5.0, 1, 80, 2, nBubbles=3),
    _req(2, 257, 0.0625, 7.0, -2, 100, 3),
]


def _schedule(nt, dt, snaps):
    T = nt * dt   # SnapshotSchedule, src/main.cpp:201-223
    return sorted(set(min(nt, max(1, int(math.floor(T * k / snaps / dt + 0.5)))) for k in range(1, snaps + 1)))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c5_sized_requests_alone_in_the_mix_and_vs_oracle(precision):
    sv = cb.Server(precision=precision, max_active=3)
    ids = {}
    for i, (p, nt, snaps) in enumerate(C5_REQS):
        ids[sv.submit(nt, snaps, tag=i, **p)] = i
    sv.drain()
    results = sv.poll()
    assert len(results) == len(C5_REQS)
    for r in sorted(results, key=lambda r: r.tag):
        assert r.status == 0, r.status
        p, nt, snaps = C5_REQS[r.tag]
        print("C5 request", precision, p, "nt", nt, flush=True)
        alone = cb.Bubble(precision=precision, **p)
        v = cb.server_variant(alone)
        alone.select_kernel(v)            # -1: the general kernel, as the server runs it
        o = oracle.Sim(**p) if precision == "fp64" else None
        sched = _schedule(nt, p["dt"], snaps)
        rows, spectra = [alone.energy_row()], [alone.spectrum()]
        done = 0
        for ev in sorted(set(list(range(40, nt + 1, 40)) + sched + [nt])):
            alone.step(ev - done)
            if o is not None:
                o.step(ev - done)
            done = ev
            row = alone.energy_row()
            if ev % 40 == 0:
                rows.append(row)
            if ev in sched:
                spectra.append(alone.spectrum())
                if o is not None and p["nGrids"] != 384:
                    ko, _, do = o.spectrum()
                    kg, dg = sv.spectrum(r.id, len(spectra) - 1)
                    np.testing.assert_allclose(kg, ko, rtol=1e-13)
                    floor = 1e-13 * np.max(do)
                    assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + floor)
            if o is not None:
                err = row_errors(o.energy_row()[None], row[None]).max()
                assert err <= TOL, (p, ev, err)
        np.testing.assert_array_equal(sv.rows(r.id), np.array(rows))
        np.testing.assert_array_equal(np.array(r.final_row[:]), row)
        assert r.n_spectra == len(spectra)
        for w, (k, _, d) in enumerate(spectra):
            ks, ds = sv.spectrum(r.id, w)
            np.testing.assert_array_equal(ks, k)
            np.testing.assert_array_equal(ds, d)
        if o is not None:
            phi, _ = alone.fields()
            assert field_err(o.phi(), phi) <= TOL
            o.close()
        sv.release(r.id)
        alone.close()
    sv.close()
