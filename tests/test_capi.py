"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, validates parameters and reports errors through status
codes (the reference exits instead, src/main.cpp:105-138,171-183)."""
import ctypes
import os
import re

import pytest

import coolbubble_amd as cb

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "coolbubble.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const char \*|cb_status |long long )\s*(cb_\w+)\s*\(",
                                 text, re.M)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 15
    L = cb.lib()
    for s in syms:
        assert hasattr(L, s), s


def test_default_params_are_reference_defaults():
    p = cb.default_params()
    # src/def.h:30-58
    assert (p.D, p.nGrids, p.L, p.c, p.readxns, p.dt) == (2, 512, 32.0, 4.0, 1, 0.004)
    assert (p.nBubbles, p.Trad, p.expansion, p.alpha, p.Mpl) == (5, 0.2, 0, 0, 200.0)
    assert (p.rand_seed, p.thermal_base) == (42, 12345)


@pytest.mark.parametrize("D,n,L,dt,ok", [
    (2, 512, 32.0, 0.004, True),
    (2, 512, 32.0, 0.0625 / 2 ** 0.5 * 0.999, True),
    (2, 512, 32.0, 0.0625 / 2 ** 0.5 * 1.001, False),
    (3, 64, 8.0, 0.125 / 3 ** 0.5, False),   # dt == dx/sqrt(D) is not < (src/main.cpp:175)
    (1, 100, 10.0, 0.099, True),
])
def test_cfl_check(D, n, L, dt, ok):
    p = cb.default_params(D=D, nGrids=n, L=L, dt=dt)
    st = cb.lib().cb_check_stability(ctypes.byref(p))
    assert st == (0 if ok else 2)


@pytest.mark.parametrize("field,value", [("D", 4), ("D", 0), ("nGrids", 1), ("L", -1.0),
                                         ("readxns", 2), ("readxns", -4), ("precision", 7),
                                         ("c", 0.0), ("alpha", 2)])
def test_invalid_params_rejected_without_device(field, value):
    p = cb.default_params()
    setattr(p, field, value)
    h = ctypes.c_void_p()
    st = cb.lib().cb_create(ctypes.byref(p), None, 0, ctypes.byref(h))
    assert st == 1
    assert not h.value
    assert cb.lib().cb_last_error()


def test_cfl_violation_rejected_by_create():
    p = cb.default_params(D=2, nGrids=64, L=4.0, dt=0.1)
    h = ctypes.c_void_p()
    assert cb.lib().cb_create(ctypes.byref(p), None, 0, ctypes.byref(h)) == 2


def test_status_strings():
    for code in range(10):
        assert cb.lib().cb_status_string(code)


@pytest.mark.parametrize("kw,name", [
    (dict(D=2, nGrids=512, L=32.0, dt=0.0125, c=4.0, readxns=1, nBubbles=5, nt=1000),
     "Sim2_D_nt_1000_dt_0.01250_nGrids_512_L_32.0_Mode_1_nB_5_c_4.00_FFT_1_Mpl_200"),
    (dict(D=3, nGrids=384, L=48.0, dt=0.025, c=2.0, readxns=-3, nt=3000, Mpl=10.0, expansion=1),
     "Sim3_D_nt_3000_dt_0.02500_nGrids_384_L_48.0_Mode_-3_nB_5_c_2.00_FFT_1_Mpl_10"),
    (dict(D=1, nGrids=65536, L=4096.0, dt=0.0125, c=7.25, nt=500, fft=0, full1d_every=50),
     "Sim1_D_nt_500_dt_0.01250_nGrids_65536_L_4096.0_Mode_1_nB_5_c_7.25_FFT_0_Full1DFieldEvery_50_Mpl_200"),
    (dict(), "Sim2_D_nt_1000_dt_0.00400_nGrids_512_L_32.0_Mode_1_nB_5_c_4.00_FFT_1_Mpl_200"),
])
def test_run_folder_name_is_the_references(kw, name):
    """Run-folder names as src/main.cpp:243-247 builds them (to_precistr cuts
    std::to_string's six decimals; Mpl is printed as an int)."""
    assert cb.run_folder_name(**kw) == name


def test_run_rejects_bad_options_and_cfl_without_device(tmp_path):
    with pytest.raises(cb.CBError) as e:
        cb.run(str(tmp_path), nt=0, D=2, nGrids=64, L=8.0, dt=0.025)
    assert e.value.status == 1
    with pytest.raises(cb.CBError) as e:
        cb.run(str(tmp_path), nt=10, D=2, nGrids=64, L=4.0, dt=0.1)
    assert e.value.status == 2
    assert not os.listdir(tmp_path)       # nothing created before the stability check passes


@pytest.mark.parametrize("opt", [dict(field_snapshots=2), dict(memorysave=-1), dict(deflate=10),
                                 dict(wall_tracking=3), dict(wall_track_every=0), dict(snapshot_spacing=2)])
def test_run_rejects_bad_output_options(tmp_path, opt):
    """Output options are checked before anything is created (no device needed)."""
    with pytest.raises(cb.CBError) as e:
        cb.run(str(tmp_path), nt=10, D=2, nGrids=64, L=8.0, dt=0.025, **opt)
    assert e.value.status == 1
    assert not os.listdir(tmp_path)


def test_run_defaults_are_the_references():
    o = cb.RunOpts()
    assert cb.lib().cb_default_run_opts(ctypes.byref(o)) == 0
    assert (o.nt, o.snapshots, o.snapshot_spacing, o.energy_every, o.fft, o.full1d_every) == (1000, 20, 0, 40, 1, 100)
    assert (o.field_snapshots, o.memorysave, o.deflate, o.wall_tracking, o.wall_track_every) == (1, 0, 9, 0, 400)


def test_thermal_spectrum_option_is_validated():
    p = cb.default_params(thermal_spectrum=2)
    assert cb.lib().cb_check_stability(ctypes.byref(p)) == 1        # CB_INVALID_ARG
    for v in (0, 1):
        p = cb.default_params(thermal_spectrum=v)
        assert cb.lib().cb_check_stability(ctypes.byref(p)) == 0
