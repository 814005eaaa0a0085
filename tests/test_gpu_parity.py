"""FP64 GPU path vs the oracle on identical seeded inputs.

Bar: fields, energy rows and spectrum bins agree to 1e-12 relative (norm-wise
for fields: max|diff| <= 1e-12 max|oracle|), the tolerance of the FP64 path.
Each case runs the reference sequence: Bubble ctor (src/bubble.cpp:156-316),
then oneStep n times (:849-876), reading the outEP row (:977-989) and the
fields (OutField, :887-917).
"""
import numpy as np
import pytest

import coolbubble_amd as cb
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def params(**kw):
    base = dict(D=2, nGrids=64, L=8.0, c=4.0, readxns=1, nBubbles=5)
    base.update(kw)
    if "dt" not in base:
        base["dt"] = 0.2 * base["L"] / base["nGrids"]
    return base


def run_pair(kw, steps, variant=None, precision="fp64"):
    o = oracle.Sim(**kw)
    g = cb.Bubble(precision=precision, **kw)
    if variant is not None:
        g.select_kernel(variant)
    rows_o, rows_g = [o.energy_row()], [g.energy_row()]
    done = 0
    for n in steps:
        o.step(n - done)
        g.step(n - done)
        done = n
        rows_o.append(o.energy_row())
        rows_g.append(g.energy_row())
    return o, g, np.array(rows_o), np.array(rows_g)


def field_err(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(a)), 1e-300)


def row_errors(ro, rg):
    """Per-column relative error; <phi'> (col 12) is scaled by Etot and the
    Hubble constraint C (col 15) by H^2, both being differences of larger
    terms."""
    scale = np.abs(ro).copy()
    scale[:, 12] = np.maximum(scale[:, 12], np.abs(ro[:, 2]))
    scale[:, 15] = np.maximum(scale[:, 15], ro[:, 14] ** 2)
    scale = np.maximum(scale, 1e-300)
    return np.abs(ro - rg) / scale


CASES = {
    "2d_mode1": params(),
    "2d_mode-1_deltaR": params(readxns=-1, nGrids=48, L=6.0, DeltaR=0.3),
    "2d_test_wave": params(test_wave=1, nGrids=64, L=8.0),
    "2d_mode-2": params(readxns=-2, nGrids=64, L=8.0, c=2.0),
    "3d_mode-2": params(D=3, readxns=-2, nGrids=32, L=4.0),
    "3d_mode1": params(D=3, readxns=1, nGrids=24, L=4.0, nBubbles=4),
    "1d": params(D=1, nGrids=256, L=16.0),
    "2d_odd_n": params(nGrids=45, L=5.625),
}


def has_tiled(kw):
    g = cb.Bubble(**kw)
    n = g.kernel_variants()[0]
    g.close()
    return n > 0


# each case on the kernel cb_create selects, and on a tiled variant (small
# lattices start on the general kernel)
@pytest.mark.parametrize("variant", [None, 0])
@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_matches_oracle(name, variant):
    kw = CASES[name]
    print("case", name, kw, variant, flush=True)
    if variant is not None and not has_tiled(kw):
        pytest.skip("no tiled variant for this lattice")
    o, g, ro, rg = run_pair(kw, [1, 40, 80], variant)
    err = row_errors(ro, rg)
    assert err.max() <= TOL, (name, err.max(axis=0))
    phi_g, pi_g = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    assert field_err(o.pi(), pi_g) <= TOL
    if kw["readxns"] in (0, 1) and kw["D"] > 1:
        np.testing.assert_array_equal(o.xns(), g.xns())


@pytest.mark.parametrize("name", ["2d_mode1", "3d_mode-2", "3d_mode1"])
def test_every_kernel_variant_gives_the_same_fields(name):
    """Tiled variants and the general kernel do the same per-site arithmetic:
    phi and pi must be bit-identical whatever the tile shape."""
    kw = CASES[name]
    ref = cb.Bubble(**kw)
    ref.select_kernel(-1)
    ref.step(12)
    phi0, pi0 = ref.fields()
    e0 = ref.energy_row()
    n, _ = ref.kernel_variants()
    assert n >= 1
    for v in range(n):
        g = cb.Bubble(**kw)
        g.select_kernel(v)
        g.step(12)
        phi, pi = g.fields()
        np.testing.assert_array_equal(phi, phi0)
        np.testing.assert_array_equal(pi, pi0)
        assert np.max(row_errors(e0[None], g.energy_row()[None])) <= 1e-13
        g.close()


@pytest.mark.parametrize("name", ["2d_mode1", "3d_mode-2"])
def test_every_fp32_kernel_variant_gives_the_same_fields(name):
    """The FP32 path (fused multiply-adds, neighbours summed in one fixed
    order) is the same arithmetic in the general kernel and in every tiled
    one-, two- and three-step variant: bit-identical fields."""
    kw = CASES[name]
    ref = cb.Bubble(precision="fp32", **kw)
    ref.select_kernel(-1)
    ref.step(12)
    phi0, pi0 = ref.fields()
    e0 = ref.energy_row()
    n, _ = ref.kernel_variants()
    fuses, prefetch = set(), {2: set(), 3: set()}
    for v in range(n):
        g = cb.Bubble(precision="fp32", **kw)
        g.select_kernel(v)
        info = g.variant_info(v)
        fuses.add(info["fuse"])
        if info["fuse"] >= 2:
            prefetch[info["fuse"]].add(info["pf"])
        g.step(12)
        phi, pi = g.fields()
        np.testing.assert_array_equal(phi, phi0)
        np.testing.assert_array_equal(pi, pi0)
        assert np.max(row_errors(e0[None], g.energy_row()[None])) <= 1e-13
        g.close()
    assert fuses == {1, 2, 3}
    # multi-step tiles with and without the plane prefetch (3D three-step
    # tiles also with loads two planes ahead)
    assert prefetch == {2: {1, 2}, 3: {1, 2, 3} if kw["D"] == 3 else {1, 2}}


THERMAL = {
    "2d_static": params(readxns=-3, nGrids=64, L=8.0, c=2.0, Trad=0.2),
    "3d_static": params(D=3, readxns=-3, nGrids=32, L=4.0, c=2.0, Trad=0.2),
    "3d_expanding": params(D=3, readxns=-3, nGrids=32, L=4.0, c=2.0, Trad=0.2, expansion=1, Mpl=10.0),
    "3d_expanding_n24": params(D=3, readxns=-3, nGrids=24, L=3.0, c=2.0, Trad=0.3, expansion=1, Mpl=10.0),
    # conformal time, ALPHAVAL = 1 (src/def.h:57, src/bubble.cpp:741-755)
    "3d_expanding_conformal": params(D=3, readxns=-3, nGrids=32, L=4.0, c=2.0, Trad=0.2, expansion=1, Mpl=10.0,
                                     alpha=1),
    "3d_bubble_expanding_conformal": params(D=3, readxns=-2, nGrids=32, L=4.0, expansion=1, Mpl=10.0, alpha=1),
}


@pytest.mark.parametrize("variant", [None, 0])
@pytest.mark.parametrize("name", sorted(THERMAL))
def test_thermal_and_expanding_match_oracle(name, variant):
    kw = THERMAL[name]
    print("case", name, kw, variant, flush=True)
    o, g, ro, rg = run_pair(kw, [1, 2, 30], variant)
    err = row_errors(ro, rg)
    assert err.max() <= TOL, (name, err.max(axis=0))
    phi_g, pi_g = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    assert field_err(o.pi(), pi_g) <= TOL


@pytest.mark.parametrize("variant", [None, 0])
def test_expanding_bubble_matches_oracle(variant):
    kw = params(D=3, readxns=-2, nGrids=32, L=4.0, expansion=1, Mpl=10.0)
    o, g, ro, rg = run_pair(kw, [1, 7, 40], variant)
    assert row_errors(ro, rg).max() <= TOL
    phi_g, pi_g = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    # pending lagged drift materialised; stepping on continues to agree
    o.step(5)
    g.step(5)
    assert row_errors(o.energy_row()[None], g.energy_row()[None]).max() <= TOL


@pytest.mark.parametrize("name", ["2d_mode1", "3d_mode-2", "1d", "2d_odd_n"])
def test_spectrum_matches_oracle(name):
    kw = CASES[name]
    o, g, _, _ = run_pair(kw, [20])
    ko, kpo, do = o.spectrum()
    kg, kpg, dg = g.spectrum()
    assert len(ko) == len(kg)
    np.testing.assert_allclose(kg, ko, rtol=1e-13, atol=0)
    np.testing.assert_allclose(kpg, kpo, rtol=1e-13, atol=0)
    # bins: relative 1e-12, with an absolute floor at the FFT's rounding level
    floor = 1e-13 * np.max(do)
    assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + floor), np.max(np.abs(dg - do) / (np.abs(do) + floor))
    _, _, dg2 = g.spectrum()
    np.testing.assert_array_equal(dg2, dg)


def test_fp32_energies_within_1e5_of_fp64():
    for kw in (CASES["2d_mode1"], CASES["3d_mode-2"]):
        o, g, ro, rg = run_pair(kw, [40, 120], precision="fp32")
        for col in (2, 4, 5):       # Etot, Egrad, Epot
            rel = np.abs(rg[:, col] - ro[:, col]) / np.abs(ro[:, col])
            assert rel.max() <= 1e-5, (col, rel)
        assert abs(rg[-1, 3] - ro[-1, 3]) <= 1e-5 * ro[-1, 2]


def test_autotune_preserves_state_and_selects_a_variant():
    kw = CASES["3d_mode-2"]
    g = cb.Bubble(**kw)
    g.step(3)
    phi0, pi0 = g.fields()
    st0 = g.state()
    best, ms = g.autotune(steps=2)
    n, sel = g.kernel_variants()
    assert sel == best and -1 <= best < n
    assert np.all(ms > 0)
    phi1, pi1 = g.fields()
    np.testing.assert_array_equal(phi0, phi1)
    np.testing.assert_array_equal(pi0, pi1)
    st1 = g.state()
    assert (st0.t, st0.step, st0.a) == (st1.t, st1.step, st1.a)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["2d_mode1", "3d_mode-2"])
def test_nan_injected_through_set_state_is_reported(name, precision):
    """A non-finite value in the field reaches the energy sums; the row after
    the next measured step reports CB_NAN_DETECTED (the reference has no
    check and would print nan, src/bubble.cpp:977-989)."""
    kw = CASES[name]
    g = cb.Bubble(precision=precision, **kw)
    g.step(3)
    phi, pi = g.fields()
    st = g.state()
    phi.reshape(-1)[phi.size // 3] = np.nan
    g.set_state(st, phi, pi)
    g.step(2)
    with pytest.raises(cb.CBError) as e:
        g.energy_row()
    assert e.value.status == 5                       # CB_NAN_DETECTED
    # a clean state clears the flag
    phi.reshape(-1)[phi.size // 3] = 0.0
    g.set_state(st, phi, pi)
    g.step(2)
    assert np.all(np.isfinite(g.energy_row()))


def test_set_state_roundtrip_and_stepping_from_it():
    kw = CASES["2d_mode1"]
    o = oracle.Sim(**kw)
    o.step(10)
    g = cb.Bubble(**kw)
    s = o.scalars()
    st = cb.State(a=s[0], ap=s[1], H=s[2], C=s[3], t=s[4], EkinHalfPrev=s[5], step=10)
    g.set_state(st, o.phi(), o.pi())
    o.step(25)
    g.step(25)
    phi_g, pi_g = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    assert row_errors(o.energy_row()[None], g.energy_row()[None]).max() <= TOL


@pytest.mark.parametrize("D,N", [(2, 8), (2, 16), (3, 16), (3, 64), (3, 128), (2, 1024), (2, 2048)])
def test_power_of_two_sizes_thermal_fields_and_spectra(D, N):
    """Power-of-two lattices take the compile-time radix-8/4/2 transforms
    (thermal c2r, r2c, column passes, the last one fused with the binning):
    FP64 fields and spectrum vs the oracle, FP32 spectrum within 2e-3."""
    dx = 0.0625 if D == 2 else 0.125
    kw = dict(D=D, nGrids=N, L=N * dx, dt=0.2 * dx, c=2.0, readxns=-3, Trad=0.2)
    print("case", kw, flush=True)
    o, g, ro, rg = run_pair(kw, [3])
    assert row_errors(ro, rg).max() <= TOL
    phi_g, _ = g.fields()
    assert field_err(o.phi(), phi_g) <= TOL
    ko, _, do = o.spectrum()
    kg, _, dg = g.spectrum()
    assert len(kg) == len(ko)
    np.testing.assert_allclose(kg, ko, rtol=1e-13, atol=0)
    floor = 1e-13 * np.max(do)
    assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + floor), np.max(np.abs(dg - do) / (np.abs(do) + floor))
    g32 = cb.Bubble(precision="fp32", **kw)
    g32.step(3)
    k32, _, d32 = g32.spectrum()
    np.testing.assert_allclose(k32, ko, rtol=1e-13)
    big = do > 1e-6 * np.max(do)
    np.testing.assert_allclose(d32[big], do[big], rtol=2e-3)
    # a second spectrum of the same state gives the same bins bit for bit
    # (fixed summation order, no atomics)
    _, _, dg2 = g.spectrum()
    np.testing.assert_array_equal(dg2, dg)
    _, _, d32b = g32.spectrum()
    np.testing.assert_array_equal(d32b, d32)


@pytest.mark.parametrize("N", [8192, 65536])
def test_long_1d_spectrum_matches_oracle(N):
    """1D rows longer than one in-LDS transform go through the four-step
    path (the serving workload's d=1 requests have N = 65536)."""
    kw = params(D=1, nGrids=N, L=N * 0.0625, dt=0.0125)
    o, g, _, _ = run_pair(kw, [40])
    ko, _, do = o.spectrum()
    kg, _, dg = g.spectrum()
    assert len(kg) == len(ko) == N // 2 + 1
    np.testing.assert_allclose(kg, ko, rtol=1e-13)
    floor = 1e-13 * np.max(do)
    assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + floor)
    g32 = cb.Bubble(precision="fp32", **kw)
    g32.step(40)
    k32, _, d32 = g32.spectrum()
    big = do > 1e-6 * np.max(do)
    np.testing.assert_allclose(d32[big], do[big], rtol=1e-3)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("N", [100, 3000, 65536])
def test_1d_multi_step_launches_bit_identical(N, precision):
    """A 1D call runs its steps up to 32 per launch (each block's sites and
    their light cone in LDS), the last one on the one-step kernel with the
    energies.  Fields and energy row equal those of one-step calls bit for
    bit: a lattice shorter than a block's window (100 sites, 100 + 2 x 32
    loaded) and one that is not a multiple of a block's 1024 sites included.
    75 steps = launches of 32, 32 and 10 steps, then one."""
    kw = params(D=1, nGrids=N, L=N * 0.0625, dt=0.0125)
    a = cb.Bubble(precision=precision, **kw)
    b = cb.Bubble(precision=precision, **kw)
    a.step(75)
    for _ in range(75):
        b.step(1)
    (pa, qa), (pb, qb) = a.fields(), b.fields()
    assert np.any(pa != pa.flat[0])
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(qa, qb)
    np.testing.assert_array_equal(a.energy_row(), b.energy_row())
    assert a.state().t == b.state().t
    a.close()
    b.close()


@pytest.mark.parametrize("fuse", [2, 3])
@pytest.mark.parametrize("name", ["2d_mode1", "3d_mode-2", "3d_mode1"])
def test_multi_step_variants_bit_identical_with_odd_counts(name, fuse):
    """Variants that fuse two or three leapfrog steps per pass (static box)
    give the single-step fields bit for bit, also when a call's step count is
    not a multiple of theirs (the rest runs on a one-step companion)."""
    kw = CASES[name]
    ref = cb.Bubble(**kw)
    ref.select_kernel(-1)
    n, _ = ref.kernel_variants()
    fused = [v for v in range(n) if ref.variant_info(v)["fuse"] == fuse]
    assert fused, "no %d-step variant for %s" % (fuse, name)
    for chunk in (7, 40, 1, 2):
        ref.step(chunk)
    phi0, pi0 = ref.fields()
    e0 = ref.energy_row()
    for v in fused[:6] + fused[-2:]:
        g = cb.Bubble(**kw)
        g.select_kernel(v)
        for chunk in (7, 40, 1, 2):
            g.step(chunk)
        phi, pi = g.fields()
        np.testing.assert_array_equal(phi, phi0)
        np.testing.assert_array_equal(pi, pi0)
        assert np.max(row_errors(e0[None], g.energy_row()[None])) <= 1e-13
        g.close()


@pytest.mark.parametrize("D,N,precision", [(3, 256, "fp32"), (3, 384, "fp32"), (3, 512, "fp32"), (3, 256, "fp64"),
                                         (3, 512, "fp64"), (2, 512, "fp32"), (2, 1024, "fp64")])
def test_wide_row_variants_bit_identical(D, N, precision):
    """The prefetching multi-step tiles at lattice sizes where a 64-lane wave
    holds whole rows or parts of one row only (N = 256 and up), the shapes
    the measured workloads use, and the direct-load tiles (phi_n loaded
    straight into LDS: prefetch kind 4; 3D N = 512, FP32 three and two
    steps of eight sites per thread, FP64 two steps of four): every such
    variant gives the general kernel's fields bit for bit over a call of 7
    steps (multi-step launches, then the leftover step on the one-step
    companion), and at N = 512 also over calls of 8 and then 6 more (the
    three-step tiles end 3q + 1, 3q + 2 and 3q steps on one to three
    launches of the direct-load two-step form, the last one measuring the
    energies: rows within 1e-13 of the general kernel's).  At N = 512 only
    the direct-load tiles and the 1024-thread four-row three-step tiles are
    run (the others are the C2 tests')."""
    kw = params(D=D, readxns=-2, nGrids=N, L=N / 8.0)
    ref = cb.Bubble(precision=precision, **kw)
    ref.select_kernel(-1)
    ref.step(7)
    phi0, pi0 = ref.fields()
    n, _ = ref.kernel_variants()
    if N == 512 and D == 3:
        ref.step(8)
        phi1, pi1 = ref.fields()
        e1 = ref.energy_row()
        ref.step(6)
        phi2, pi2 = ref.fields()
    ref.close()
    g = cb.Bubble(precision=precision, **kw)
    infos = [g.variant_info(v) for v in range(n)]
    g.close()
    multi = [v for v in range(n) if infos[v]["fuse"] >= 2 and (infos[v]["pf"] >= 2 or infos[v]["vw"] == 8)]
    if N == 512 and D == 3:
        # the direct-load tiles, and the 1024-thread four-row three-step
        # tiles, whose calls end on the direct-load tile's two-step form
        multi = [v for v in multi if infos[v]["pf"] == 4 or
                 (infos[v]["fuse"] == 3 and infos[v]["nthr"] == 1024 and infos[v]["TR"] == 4)]
        assert any(infos[v]["pf"] == 4 for v in multi)
    assert multi

    for v in multi:
        g = cb.Bubble(precision=precision, **kw)
        g.select_kernel(v)
        g.step(7)
        phi, pi = g.fields()
        np.testing.assert_array_equal(phi, phi0, err_msg=str(infos[v]))
        np.testing.assert_array_equal(pi, pi0, err_msg=str(infos[v]))
        if N == 512 and D == 3:
            g.step(8)
            phi, pi = g.fields()
            np.testing.assert_array_equal(phi, phi1, err_msg=str(infos[v]))
            np.testing.assert_array_equal(pi, pi1, err_msg=str(infos[v]))
            assert np.max(row_errors(e1[None], g.energy_row()[None])) <= 1e-13, infos[v]
            g.step(6)
            phi, pi = g.fields()
            np.testing.assert_array_equal(phi, phi2, err_msg=str(infos[v]))
            np.testing.assert_array_equal(pi, pi2, err_msg=str(infos[v]))
        print("variant", v, infos[v], "bit-identical", flush=True)
        g.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("N,expanding", [(96, False), (192, True), (384, True)])
def test_three_wave_one_step_tiles_bit_identical(N, expanding, precision):
    """One-step tiles of 192 and 384 threads (3D lattices whose rows hold a
    multiple of three 64-lane waves of vectors, N = 3 x 2^m): fields bit for
    bit the general kernel's over a call of 7 steps, energy rows within
    1e-13; static and expanding (lagged drift, Friedmann sums only on the
    steps before the last).  N = 384 runs the two-vector 384-thread tiles
    only (the others are covered at N = 96 and 192)."""
    if expanding:
        kw = params(D=3, readxns=-3, nGrids=N, L=N / 8.0, c=2.0, Trad=0.2, expansion=1, Mpl=10.0)
    else:
        kw = params(D=3, readxns=-2, nGrids=N, L=N / 8.0)
    ref = cb.Bubble(precision=precision, **kw)
    ref.select_kernel(-1)
    ref.step(7)
    phi0, pi0 = ref.fields()
    e0 = ref.energy_row()
    n, _ = ref.kernel_variants()
    infos = [ref.variant_info(v) for v in range(n)]
    ref.close()
    three = [v for v in range(n) if infos[v]["fuse"] == 1 and infos[v]["nthr"] in (192, 384)]
    if N == 384:
        three = [v for v in three if infos[v]["nthr"] == 384 and infos[v]["items"] == 2]
    assert three
    for v in three:
        g = cb.Bubble(precision=precision, **kw)
        g.select_kernel(v)
        g.step(7)
        phi, pi = g.fields()
        np.testing.assert_array_equal(phi, phi0, err_msg=str(infos[v]))
        np.testing.assert_array_equal(pi, pi0, err_msg=str(infos[v]))
        assert np.max(row_errors(e0[None], g.energy_row()[None])) <= 1e-13, infos[v]
        g.close()


def test_default_variant_follows_the_lattice():
    """cb_create starts small lattices on the general kernel, large static
    boxes on a three-step tile, expanding ones on a one-step tile."""
    small = cb.Bubble(**CASES["2d_mode1"])
    assert small.kernel_variants()[1] == -1
    big = cb.Bubble(precision="fp32", **params(D=3, readxns=-2, nGrids=256, L=32.0))
    n, sel = big.kernel_variants()
    assert 0 <= sel < n and big.variant_info(sel)["fuse"] == 3
    exp = cb.Bubble(precision="fp32", **params(D=3, readxns=-2, nGrids=128, L=16.0, expansion=1, Mpl=10.0))
    n, sel = exp.kernel_variants()
    assert 0 <= sel < n and exp.variant_info(sel)["fuse"] == 1
    wide = cb.Bubble(precision="fp32", **params(D=2, readxns=-2, nGrids=2048, L=128.0))
    n, sel = wide.kernel_variants()
    assert 0 <= sel < n and wide.variant_info(sel)["fuse"] == 3


def test_default_variant_at_512_cubed():
    """3D N = 512 static boxes start on the tiles the autotuner picks there:
    FP32 the 1024-thread three-step tile in 256-plane chunks (one block per
    CU), FP64 the direct-load two-step tile (prefetch kind 4) in 256-plane
    chunks; a 20-step call on the default equals the general kernel's."""
    kw = params(D=3, readxns=-2, nGrids=512, L=64.0)
    want = {"fp32": dict(fuse=3, nthr=1024, TR=4, NI=256, pf=2, nt=1),
            "fp64": dict(fuse=2, nthr=768, TR=4, NI=256, pf=4, nt=1)}
    for precision, w in want.items():
        ref = cb.Bubble(precision=precision, **kw)
        ref.select_kernel(-1)
        ref.step(20)
        phi0, pi0 = ref.fields()
        e0 = ref.energy_row()
        ref.close()
        g = cb.Bubble(precision=precision, **kw)
        n, sel = g.kernel_variants()
        assert 0 <= sel < n
        info = g.variant_info(sel)
        assert {k: info[k] for k in w} == w, (precision, info)
        g.step(20)
        phi, pi = g.fields()
        np.testing.assert_array_equal(phi, phi0)
        np.testing.assert_array_equal(pi, pi0)
        assert np.max(row_errors(e0[None], g.energy_row()[None])) <= 1e-13
        g.close()
        print(precision, "default", info, "bit-identical", flush=True)


def test_spectrum_with_more_bins_than_64k_of_lds_matches_oracle():
    """2D N = 4096 has 2896 |k| bins: the binning histogram (3 doubles per
    bin) needs more than 64 KiB of LDS per block, which gfx950 allows."""
    kw = params(D=2, readxns=-2, nGrids=4096, L=512.0)
    o, g, _, _ = run_pair(kw, [2])
    ko, kpo, do = o.spectrum()
    kg, kpg, dg = g.spectrum()
    assert len(ko) == len(kg) > 2730
    np.testing.assert_allclose(kg, ko, rtol=1e-13, atol=0)
    floor = 1e-13 * np.max(do)
    assert np.all(np.abs(dg - do) <= TOL * np.abs(do) + floor), np.max(np.abs(dg - do) / (np.abs(do) + floor))


def test_slab_probe_times_one_ranks_share():
    """The slab probe runs the multi-step tile on the first planes of the
    lattice only (one rank's share of a run sharded along the slowest axis)
    and the halo-exchange copies; it refuses shapes without room for the
    halo planes or off the chunk length."""
    s = cb.Bubble(precision="fp32", **params(D=3, readxns=-2, nGrids=64, L=8.0))
    n, _ = s.kernel_variants()
    v = next(v for v in range(n) if s.variant_info(v)["fuse"] == 3 and s.variant_info(v)["NI"] == 16)
    s.select_kernel(v)
    ml, mx = s.slab_probe(32, 5)
    assert ml > 0 and mx > 0
    for bad in (8, 24, 64):
        with pytest.raises(cb.CBError):
            s.slab_probe(bad, 1)
    s.close()
    e = cb.Bubble(precision="fp32", **params(D=3, readxns=-2, nGrids=64, L=8.0, expansion=1, Mpl=10.0))
    with pytest.raises(cb.CBError):
        e.slab_probe(32, 1)
    e.close()


@pytest.mark.parametrize("D,N,L", [(3, 32, 8.0), (2, 256, 16.0)])
def test_rayleigh_jeans_thermal_option_spectrum(D, N, L):
    """Build option thermal_spectrum = 1 (no reference counterpart): the
    seeded modes of phi have <|phi_k|^2> = T / omega^2 with T = Trad and the
    lattice frequency omega^2 = sum_i (2/dx)^2 sin^2(pi s_i / N) + 2c.  The
    modes are recovered from phi at t = 0 (phi = L^(-D/2) c2r(modes)); each
    |phi_k|^2 / P(k) is exponentially distributed with mean 1, so its mean
    over the generic modes (the planes j = 0 and N/2 are symmetrised) and
    over the lower and upper halves of the momenta must be 1 within 5 sigma.
    The continuum momentum instead of the lattice one, or T = Trad mT, misses
    the upper half by far more."""
    c, Trad = 2.0, 0.5
    g = cb.Bubble(precision="fp64", D=D, nGrids=N, L=L, dt=0.2 * L / N, c=c, readxns=-3, Trad=Trad,
                  thermal_spectrum=1)
    phi, _ = g.fields()
    X = np.fft.rfftn(phi) * L ** (D / 2) / N ** D
    s = np.fft.fftfreq(N, 1.0 / N)
    dx = L / N
    kl = (2 / dx) * np.sin(np.pi * s / N)
    grids = np.meshgrid(*([kl] * (D - 1) + [kl[:N // 2 + 1]]), indexing="ij")
    w2 = sum(q * q for q in grids) + 2 * c
    P = Trad / w2
    ratio = (np.abs(X) ** 2 / P)[..., 1:N // 2]
    w2g = w2[..., 1:N // 2]
    assert abs(ratio.mean() - 1.0) <= 5.0 / np.sqrt(ratio.size), ratio.mean()
    cut = np.median(w2g)
    for band in (ratio[w2g <= cut], ratio[w2g > cut]):
        assert abs(band.mean() - 1.0) <= 5.0 / np.sqrt(band.size), band.mean()
    # against the continuum momentum (k >= k_lat) the upper half is 15 % or more too strong
    kc = 2 * np.pi * s / L
    gc = np.meshgrid(*([kc] * (D - 1) + [kc[:N // 2 + 1]]), indexing="ij")
    Pc = (Trad / (sum(q * q for q in gc) + 2 * c))[..., 1:N // 2]
    up = w2g > cut
    assert (np.abs(X[..., 1:N // 2]) ** 2 / Pc)[up].mean() > 1.15
