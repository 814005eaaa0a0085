"""cb_power_spectrum_noalloc: the power spectrum with its work arrays in the
library's static device scratch (no hipMalloc).  Called through the C ABI on
two calls in a row -- the second on a smaller lattice -- and compared with
cb_power_spectrum on the same state, which allocates its own plan: the
same k means, k / a and bins (bit for bit: same transforms, same summation
order), on the power-of-two fused path (2D, 3D), the generic path (odd N)
and the four-step long 1D rows.  A lattice whose work arrays exceed the
scratch is refused with CB_UNSUPPORTED."""
import ctypes

import numpy as np
import pytest

import coolbubble_amd as cb

pytestmark = pytest.mark.gpu


def _noalloc(sim):
    p = sim.p
    cap = int(np.sqrt(p.D) * p.nGrids / 2) + 2
    k, kp, d = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    nb = ctypes.c_int()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    st = cb.lib().cb_power_spectrum_noalloc(sim.handle, dp(k), dp(kp), dp(d), cap, ctypes.byref(nb))
    return st, k[:nb.value], kp[:nb.value], d[:nb.value]


CASES = [   # (first, second): the second lattice is smaller
    (dict(D=3, nGrids=128, L=16.0, c=2.0, readxns=-3, Trad=0.2, dt=0.025),
     dict(D=3, nGrids=64, L=8.0, c=4.0, readxns=-2, dt=0.025)),
    (dict(D=2, nGrids=2048, L=128.0, c=4.0, readxns=1, nBubbles=5, dt=0.0125),
     dict(D=2, nGrids=45, L=5.625, c=4.0, readxns=1, nBubbles=3, dt=0.0125)),
    (dict(D=1, nGrids=65536, L=4096.0, c=4.0, dt=0.0125),
     dict(D=1, nGrids=1000, L=62.5, c=3.0, dt=0.0125)),
    (dict(D=3, nGrids=96, L=12.0, c=2.0, readxns=-3, Trad=0.2, dt=0.025, expansion=1, Mpl=10.0),
     dict(D=3, nGrids=33, L=4.125, c=2.0, readxns=-3, Trad=0.2, dt=0.025, expansion=1, Mpl=10.0)),
]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_spectrum_equals_cb_power_spectrum(case, precision):
    for kw in CASES[case]:
        print("noalloc spectrum", precision, kw, flush=True)
        s = cb.Bubble(precision=precision, **kw)
        s.step(5)
        st, k, kp, d = _noalloc(s)
        assert st == 0, cb.lib().cb_last_error()
        k0, kp0, d0 = s.spectrum()        # cb_power_spectrum (allocates its plan)
        assert len(k) == len(k0) > 0
        np.testing.assert_array_equal(k, k0)
        np.testing.assert_array_equal(kp, kp0)
        np.testing.assert_array_equal(d, d0)
        assert np.all(np.isfinite(d)) and d.max() > 0
        s.close()


def test_noalloc_spectrum_refuses_lattices_beyond_the_scratch():
    s = cb.Bubble(precision="fp32", D=3, nGrids=256, L=32.0, c=4.0, readxns=-2, dt=0.025)
    st, _, _, _ = _noalloc(s)
    assert st == 9                                     # CB_UNSUPPORTED
    assert b"scratch" in cb.lib().cb_last_error()
    st, k, _, d = _noalloc(cb.Bubble(precision="fp32", D=2, nGrids=64, L=8.0, c=4.0, readxns=-2, dt=0.025))
    assert st == 0 and len(k) > 0                      # still usable afterwards
