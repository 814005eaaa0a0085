"""Parity at the full sizes of the measured configurations.

* d=2 N=512 L=32 c=4 mode 1 (5 bubbles) dt=0.0125, nt=1000, 20 snapshots:
  the whole run, FP64 GPU vs the oracle (energy rows every 40 steps, spectra
  at every snapshot, fields at the end) to 1e-12; FP32 energies within 1e-5
  of the FP64 oracle.
* d=3 N=512 L=64 c=4 mode 1 (32 bubbles, radius x1.3) dt=0.025: the seeded
  phi on one full plane and at random sites vs the oracle's seeding, the
  ctor sequence and 4 steps over the full lattice (the oracle started from
  the GPU's seeded phi) to 1e-12; then 400 steps: static-box energy
  conservation and FP32 vs FP64 energies within 1e-5.
* d=3 N=384 L=48 c=2 mode -3 (Trad=0.1) expanding (Mpl=10) dt=0.025: the
  oracle's whole ctor (thermal phi and pi) and 3 steps vs FP64 GPU to 1e-12;
  FP32 energies within 1e-5.

The row layout and tolerances follow the reference's outEP columns
(src/bubble.cpp:977-989); SnapshotSchedule is src/main.cpp:201-223.
"""
import math

import numpy as np
import pytest

import coolbubble_amd as cb
import oracle
from test_gpu_parity import TOL, field_err, row_errors

pytestmark = pytest.mark.gpu


def snapshot_steps(nt, dt, n):
    """SnapshotSchedule (SnapshotSpacing = 0): step = min(nt, max(1,
    lround(T k / n / dt))), deduplicated."""
    T = nt * dt
    out = []
    for k in range(1, n + 1):
        st = min(nt, max(1, int(math.floor(T * k / n / dt + 0.5))))
        if not out or st > out[-1]:
            out.append(st)
    return out


def assert_fp32_energies(r32, r64, tol=1e-5):
    for col in (2, 4, 5):                       # Etot, Egrad, Epot
        rel = abs(r32[col] - r64[col]) / abs(r64[col])
        assert rel <= tol, (col, rel)
    assert abs(r32[3] - r64[3]) <= tol * abs(r64[2])   # Ekin, relative to Etot


def test_2d_n512_whole_run_with_snapshots():
    kw = dict(D=2, nGrids=512, L=32.0, c=4.0, readxns=1, nBubbles=5, dt=0.0125)
    nt = 1000
    snaps = snapshot_steps(nt, kw["dt"], 20)
    assert snaps == list(range(50, 1001, 50))
    events = sorted(set(range(40, nt + 1, 40)) | set(snaps))
    print("2d N=512: events", len(events), flush=True)
    o = oracle.Sim(**kw)
    g = cb.Bubble(precision="fp64", **kw)
    g32 = cb.Bubble(precision="fp32", **kw)
    np.testing.assert_array_equal(o.xns(), g.xns())
    assert row_errors(o.energy_row()[None], g.energy_row()[None]).max() <= TOL
    done = 0
    for ev in events:
        o.step(ev - done)
        g.step(ev - done)
        g32.step(ev - done)
        done = ev
        ro, rg = o.energy_row(), g.energy_row()
        err = row_errors(ro[None], rg[None]).max()
        assert err <= TOL, (ev, err)
        assert_fp32_energies(g32.energy_row(), ro)
        if ev in snaps:
            ko, _, do = o.spectrum()
            kg, _, dg = g.spectrum()
            np.testing.assert_allclose(kg, ko, rtol=1e-13)
            floor = 1e-13 * np.max(do)
            assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + floor), ev
            _, _, d32 = g32.spectrum()
            big = do > 1e-6 * np.max(do)
            np.testing.assert_allclose(d32[big], do[big], rtol=2e-3)
    phi_g, pi_g = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    assert field_err(o.pi(), pi_g) <= TOL


def test_3d_n512_seeding_ctor_and_steps():
    kw = dict(D=3, nGrids=512, L=64.0, c=4.0, readxns=1, nBubbles=32, dt=0.025)
    N = kw["nGrids"]
    print("3d N=512 mode 1: create", flush=True)
    g = cb.Bubble(precision="fp64", **kw)
    xs = g.xns()
    np.testing.assert_array_equal(xs, oracle.generate_xns(3, 64.0, 4.0, 32))
    row0 = g.energy_row()
    phi0, pim = g.fields()
    # seeding: the middle plane i = N/2 and 4096 random sites
    rng = np.random.default_rng(512)
    idx = np.concatenate([(N // 2) * N * N + np.arange(N * N, dtype=np.int64),
                          rng.integers(0, N ** 3, 4096, dtype=np.int64)])
    ref = oracle.seed_sites(xs, idx, **kw)
    flat = phi0.reshape(-1)
    assert np.max(np.abs(flat[idx] - ref)) <= TOL * np.max(np.abs(ref))
    assert np.max(ref) > 0.5          # the plane crosses bubbles
    # ctor sequence on the full lattice from the GPU's seeded phi (pi0 = 0)
    print("3d N=512: oracle ctor", flush=True)
    o = oracle.Sim(blank=True, **kw)
    o.phi()[...] = phi0
    o.ctor()
    assert row_errors(o.energy_row()[None], row0[None]).max() <= TOL
    assert field_err(o.pi(), pim) <= TOL
    del phi0, pim, flat
    print("3d N=512: 4 steps", flush=True)
    o.step(4)
    g.step(4)
    assert row_errors(o.energy_row()[None], g.energy_row()[None]).max() <= TOL
    phi_g, pi_g = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    assert field_err(o.pi(), pi_g) <= TOL
    o.close()
    del phi_g, pi_g
    # 400 steps (energies every 40): conservation and FP32 vs FP64
    print("3d N=512: 400 steps fp64 / fp32", flush=True)
    g.close()
    g = cb.Bubble(precision="fp64", **kw)
    g32 = cb.Bubble(precision="fp32", **kw)
    e0 = g.energy_row()[2]
    drift = 0.0
    for _ in range(10):
        g.step(40)
        g32.step(40)
        r64, r32 = g.energy_row(), g32.energy_row()
        drift = max(drift, abs(r64[2] - e0) / e0)
        assert_fp32_energies(r32, r64)
    print("relative Etot drift over 400 steps: %.3e" % drift, flush=True)
    assert drift <= 1e-2


def test_3d_n384_thermal_expanding_ctor_and_steps():
    kw = dict(D=3, nGrids=384, L=48.0, c=2.0, readxns=-3, Trad=0.1, expansion=1, Mpl=10.0, dt=0.025)
    print("3d N=384 thermal expanding: oracle create", flush=True)
    o = oracle.Sim(**kw)
    g = cb.Bubble(precision="fp64", **kw)
    g32 = cb.Bubble(precision="fp32", **kw)
    assert row_errors(o.energy_row()[None], g.energy_row()[None]).max() <= TOL
    phi_g, pi_g = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    assert field_err(o.pi(), pi_g) <= TOL
    del phi_g, pi_g
    for n in range(3):
        print("step", n + 1, flush=True)
        o.step(1)
        g.step(1)
        g32.step(1)
        ro = o.energy_row()
        err = row_errors(ro[None], g.energy_row()[None]).max()
        assert err <= TOL, (n, err)
        assert_fp32_energies(g32.energy_row(), ro)
    phi_g, pi_g = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    assert field_err(o.pi(), pi_g) <= TOL
