"""Whole runs of the two large measured configurations against full-length
oracle fixtures (tests/golden/make_fullsize.py writes them; a missing
fixture skips its case).

* c2: d=3 N=512 L=64 c=4 mode 1 (32 bubbles) dt=0.025, static, nt=2000.
* c3: d=3 N=384 L=48 c=2 mode -3 (Trad=0.1), expanding (Mpl=10), dt=0.025,
  nt=3000.

The FP64 GPU run issues the reference driver's sequence (src/main.cpp:274-288:
oneStep every step, outEP every 40 steps, SaveSnapshot at the 20 steps of
SnapshotSchedule, src/main.cpp:201-223) and is compared, at every row and
snapshot and after the last step, with the oracle's run of the same
parameters: energy rows (outEP, src/bubble.cpp:977-989) and spectrum bins
(FourierTransform + ComputePowerSpectrum, :642-709: mean |k| and power) to
1e-12 relative (the power with an absolute floor of 1e-13 of the largest
bin, the FFT's rounding level), the field checksums (sum, sum of squares, max |x|) to 1e-12, and phi
and pi at 4096 fixed sites to 1e-12 of max |phi| / max |pi|.  The FP32 run of
the same configuration reports its errors against the same fixture: energies
within 1e-5 (Etot, Egrad, Epot relative; Ekin relative to Etot, the bar the
FP32 path is held to), spectrum bins within 1e-3 of the largest bin (and
the worst relative error of the bins above 1e-6 of it, reported).
"""
import json
import os

import numpy as np
import pytest

import coolbubble_amd as cb
from test_gpu_parity import TOL, row_errors

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = {"c2": "c2_d3_n512_nt2000.npz", "c3": "c3_d3_n384_nt3000.npz"}


def load(name):
    path = os.path.join(HERE, "golden", FIXTURES[name])
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated (python tests/golden/make_fullsize.py %s)" % (path, name))
    z = np.load(path)
    meta = json.loads(str(z["meta"]))
    return meta, z


def checksums(x):
    x = x.reshape(-1)
    return np.array([np.sum(x), np.sum(x * x), np.max(np.abs(x))])


def assert_checksums(got, want, n, what):
    """sum within 1e-12 of sqrt(n sum x^2) (a bound on sum |x|), sum of
    squares and max |x| within 1e-12 relative"""
    assert abs(got[0] - want[0]) <= TOL * np.sqrt(n * want[1]), (what, got[0], want[0])
    assert abs(got[1] - want[1]) <= TOL * want[1], (what, got[1], want[1])
    assert abs(got[2] - want[2]) <= TOL * want[2], (what, got[2], want[2])


def assert_fp32_energies(r32, r64, tol=1e-5):
    for col in (2, 4, 5):                       # Etot, Egrad, Epot
        rel = abs(r32[col] - r64[col]) / abs(r64[col])
        assert rel <= tol, (col, rel)
    assert abs(r32[3] - r64[3]) <= tol * abs(r64[2])   # Ekin, relative to Etot


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_whole_run_matches_full_length_fixture(name):
    meta, z = load(name)
    kw, nt = meta["params"], meta["nt"]
    row_steps = [int(s) for s in z["row_steps"]]
    spec_steps = [int(s) for s in z["spec_steps"]]
    rows, spec_delta, spec_k, sums = z["rows"], z["spec_delta"], z["spec_k"], z["checksums"]
    print("%s: %s, nt=%d, %d rows, %d spectra" % (name, kw, nt, len(row_steps), len(spec_steps)), flush=True)
    g = cb.Bubble(precision="fp64", **kw)
    g32 = cb.Bubble(precision="fp32", **kw)
    n = g.nsites
    worst = {"row": 0.0, "bin": 0.0, "fp32_row": 0.0, "fp32_bin": 0.0, "fp32_spec": 0.0}

    def check(step):
        if step in row_steps:
            ro = rows[row_steps.index(step)]
            err = row_errors(ro[None], g.energy_row()[None]).max()
            worst["row"] = max(worst["row"], err)
            assert err <= TOL, (name, step, err)
            r32 = g32.energy_row()
            assert_fp32_energies(r32, ro)
            worst["fp32_row"] = max(worst["fp32_row"], abs(r32[2] - ro[2]) / abs(ro[2]))
        if step in spec_steps:
            i = spec_steps.index(step)
            kg, _, dg = g.spectrum()
            # each bin's mean |k| sums up to ~1e6 modes at 512^3 (a different
            # order on each side): 1e-12, the bar of every spectrum column
            np.testing.assert_allclose(kg, spec_k[i], rtol=TOL, atol=0)
            do = spec_delta[i]
            floor = 1e-13 * np.max(do)
            err = np.max(np.abs(dg - do) / (TOL * np.abs(do) + floor))
            worst["bin"] = max(worst["bin"], float(np.max(np.abs(dg - do) / np.maximum(np.abs(do), floor / TOL))))
            assert err <= 1.0, (name, step, err)
            _, _, d32 = g32.spectrum()
            big = do > 1e-6 * np.max(do)
            rel32 = np.abs(d32[big] - do[big]) / do[big]
            worst["fp32_bin"] = max(worst["fp32_bin"], float(np.max(rel32)))
            norm32 = float(np.max(np.abs(d32 - do)) / np.max(do))
            worst["fp32_spec"] = max(worst["fp32_spec"], norm32)
            assert norm32 <= 1e-3, (name, step, norm32)
            phi, pi = g.fields()
            assert_checksums(checksums(phi), sums[i][:3], n, (name, step, "phi"))
            assert_checksums(checksums(pi), sums[i][3:], n, (name, step, "pi"))

    check(0)
    done = 0
    for ev in sorted(set(row_steps) | set(spec_steps)):
        if ev == 0:
            continue
        g.step(ev - done)
        g32.step(ev - done)
        done = ev
        check(ev)
    assert done == nt
    phi, pi = g.fields()
    sites = z["sites"]
    for got, want, full in ((phi.reshape(-1)[sites], z["phi_sites"], sums[-1][2]),
                            (pi.reshape(-1)[sites], z["pi_sites"], sums[-1][5])):
        assert np.max(np.abs(got - want)) <= TOL * full
    print("%s: worst FP64 row error %.2e, bin error %.2e; FP32 Etot error %.2e, spectrum error %.2e of the "
          "largest bin (%.2e relative in bins above 1e-6 of it)"
          % (name, worst["row"], worst["bin"], worst["fp32_row"], worst["fp32_spec"], worst["fp32_bin"]), flush=True)
    g.close()
    g32.close()
