"""Pins of the oracle (oracle/cb_oracle.c) before it is trusted as the checker.

The reference ships no tests, fixtures or golden data.  Its src/bubble.cpp
needs FFTW3, which is absent here, so only src/gridfuncs.cpp can be built
(oracle/_ref).  Each part of the oracle is therefore pinned by one of:
  * the compiled reference (lattice indexing, periodic images);
  * a known-answer test or the library routine the reference itself calls
    (C++ standard mt19937_64 KAT, libstdc++ normal_distribution, glibc rand,
    numpy's FFT for the DFT conventions);
  * closed forms and invariants (potential, plane-wave Laplacian eigenvalue,
    linear dispersion of the leapfrog, de Sitter scale factor, Parseval,
    lattice-point counts of the spectrum shells, energy conservation).
"""
import ctypes
import math
import os
import shutil
import subprocess
import tempfile
from fractions import Fraction

import numpy as np
import pytest

import oracle

HAVE_GXX = shutil.which("g++") is not None


# --------------------------------------------------------------------------
# random numbers
# --------------------------------------------------------------------------
def test_mt19937_64_known_answer():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed
    # mt19937_64 (seed 5489) is 9981545732273789042
    assert int(oracle.mt64(5489, 10000)[-1]) == 9981545732273789042


_CXX_NORMALS = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char **argv) {
    unsigned long long seed = std::strtoull(argv[1], nullptr, 10);
    long n = std::atol(argv[2]);
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> d(0.0, 1.0);
    for (long i = 0; i < n; ++i) std::printf("%a\n", d(gen));
    return 0;
}
"""


@pytest.fixture(scope="module")
def cxx_normals():
    if not HAVE_GXX:
        pytest.skip("g++ absent")
    d = tempfile.mkdtemp()
    src, exe = os.path.join(d, "n.cpp"), os.path.join(d, "n")
    open(src, "w").write(_CXX_NORMALS)
    subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", exe], check=True)
    yield exe
    shutil.rmtree(d)


@pytest.mark.parametrize("seed", [12345, 12345 + 7, 2 ** 63 + 11, 0])
def test_normals_bitwise_equal_libstdcxx(cxx_normals, seed):
    n = 3001
    out = subprocess.run([cxx_normals, str(seed), str(n)], capture_output=True, text=True,
                         check=True).stdout.split()
    ref = np.array([float.fromhex(x) for x in out])
    np.testing.assert_array_equal(oracle.normals(seed, n), ref)


def _libc_rand_xns(D, L, c, nb, seed):
    """Centres from glibc's own srand/rand, formatted as the reference writes
    them, with the periodic images written by the reference's check_PBC."""
    ref = oracle.ref_gridfuncs()
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    RAND_MAX = 2147483647
    phiC, Rn, lw = oracle.bubble_fit(D, c)
    text = ""
    for _ in range(nb):
        xn = [L * float(libc.rand()) / RAND_MAX for _ in range(3)]
        text += "%.5e %.5e %.5e\n" % tuple(xn)
        buf = ctypes.create_string_buffer(4096)
        arr = (ctypes.c_double * 3)(*xn)
        ref.ref_check_pbc(D, arr, L, Rn + 5 * lw, buf, 4096)
        text += buf.value.decode()
    vals = [float(v) for v in text.split()]
    return np.array(vals).reshape(-1, 3)


@pytest.mark.parametrize("D,L,c,nb", [(2, 32.0, 4.0, 5), (3, 64.0, 4.0, 32), (2, 8.0, 4.0, 5),
                                      (3, 4.0, 2.0, 6)])
def test_nucleation_points_match_glibc_rand_and_reference_check_pbc(D, L, c, nb):
    if oracle.ref_gridfuncs() is None:
        pytest.skip("reference gridfuncs not built (no reference tree)")
    np.testing.assert_array_equal(oracle.generate_xns(D, L, c, nb, 42), _libc_rand_xns(D, L, c, nb, 42))


# --------------------------------------------------------------------------
# lattice indexing (src/gridfuncs.cpp)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("N", [2, 3, 4, 5, 7])
def test_indexing_matches_reference_and_brute_force(D, N):
    L = oracle.lib()
    ref = oracle.ref_gridfuncs()
    total = N ** D
    for ig in range(total):
        idx = np.unravel_index(ig, (N,) * D)
        for s in range(D if D > 1 else 2):
            m = L.cbo_iGridM(D, N, ig, s)
            if D == 1:
                want = (idx[0] + (1 if s == 1 else -1)) % N
                assert m == want
            else:
                p = L.cbo_iGridP(D, N, ig, s)
                ip = list(idx); ip[s] = (ip[s] + 1) % N
                im = list(idx); im[s] = (im[s] - 1) % N
                assert p == np.ravel_multi_index(ip, (N,) * D)
                assert m == np.ravel_multi_index(im, (N,) * D)
                if ref is not None:
                    assert p == ref.ref_iGridP(D, N, ig, s)
            if ref is not None:
                assert m == ref.ref_iGridM(D, N, ig, s)


# --------------------------------------------------------------------------
# potential (src/bubble.cpp:69-77)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("c", [1.0, 2.0, 4.0, 7.3])
def test_potential_closed_forms(c):
    L = oracle.lib()
    assert L.cbo_V(0.0, c, 1.0) == 1.0
    assert L.cbo_V(1.0, c, 1.0) == 0.0
    assert L.cbo_Vpr(0.0, c) == 0.0
    assert L.cbo_Vpr(1.0, c) == 0.0
    assert abs(L.cbo_Vpr(c / (2 * c + 6), c)) < 1e-15        # barrier top
    phi_zero = (2 + c - math.sqrt(4 + c)) / (3 + c)             # V(phiZero) = V(0)
    assert abs(L.cbo_V(phi_zero, c, 1.0) - 1.0) < 1e-14
    rng = np.random.default_rng(1)
    for phi in rng.uniform(-0.6, 1.6, 200):
        f, cc = Fraction(phi), Fraction(c)
        v = cc * f ** 2 - (2 * cc + 4) * f ** 3 + (cc + 3) * f ** 4 + 1
        vp = 2 * cc * f - 3 * (2 * cc + 4) * f ** 2 + 4 * (cc + 3) * f ** 3
        scale_v = float(abs(cc * f ** 2) + abs((2 * cc + 4) * f ** 3) + abs((cc + 3) * f ** 4) + 1)
        scale_p = float(abs(2 * cc * f) + abs(3 * (2 * cc + 4) * f ** 2) + abs(4 * (cc + 3) * f ** 3))
        assert abs(L.cbo_V(phi, c, 1.0) - float(v)) <= 4e-16 * scale_v
        assert abs(L.cbo_Vpr(phi, c) - float(vp)) <= 4e-16 * max(scale_p, 1e-300)


# --------------------------------------------------------------------------
# stepper
# --------------------------------------------------------------------------
def _vprime(phi, c):
    return 2 * c * phi - 3 * (2 * c + 4) * phi ** 2 + 4 * (c + 3) * phi ** 3


@pytest.mark.parametrize("D,N,m", [(1, 32, (3,)), (2, 16, (2, 5)), (3, 8, (1, 2, 3)), (3, 9, (4, 0, 2))])
def test_laplacian_plane_wave_eigenvalue(D, N, m):
    L = 2.0
    s = oracle.Sim(blank=True, D=D, nGrids=N, L=L, dt=0.001, c=3.0)
    idx = np.indices((N,) * D)
    theta = sum(2 * np.pi * m[a] * idx[a] / N for a in range(D))
    phi = 0.2 + 1e-3 * np.cos(theta)
    s.phi()[...] = phi
    dx = L / N
    lam = -(4 / dx ** 2) * sum(math.sin(math.pi * m[a] / N) ** 2 for a in range(D))
    lap = s.forces() + _vprime(phi, 3.0)
    np.testing.assert_allclose(lap, lam * (phi - 0.2), atol=1e-9 * abs(lam) * 1e-3 + 1e-12)


def test_uniform_field_follows_scalar_leapfrog():
    """A uniform field has an exactly zero Laplacian, so every site follows
    the scalar leapfrog phi'' = -V'(phi) started by the ctor's half back-step
    (src/bubble.cpp:280-289) -- replayed here with Python floats and glibc pow."""
    c, dt, phi0 = 4.0, 0.01, 0.31
    s = oracle.Sim(readxns=0, xns=np.zeros((0, 3)), D=2, nGrids=6, L=1.0, dt=dt, c=c)
    s.phi()[...] = phi0
    # re-run the back-step from the uniform field by hand on the blank copy
    b = oracle.Sim(blank=True, D=2, nGrids=6, L=1.0, dt=dt, c=c)
    b.phi()[...] = phi0
    vp = lambda f: 2 * c * f - 3 * (2 * c + 4) * (f * f) + 4 * (c + 3) * math.pow(f, 3)
    F = 1.0 * 0.0 / ((1.0 / 6) * (1.0 / 6)) - 1.0 * vp(phi0)
    np.testing.assert_array_equal(b.forces(), F)
    pi = 0.0 - 0.5 * dt * F
    b.pi()[...] = pi
    f = phi0
    for _ in range(50):
        F = -vp(f) if False else (1.0 * 0.0 / ((1.0 / 6) * (1.0 / 6)) - 1.0 * vp(f))
        pi = pi + dt * F
        f = f + dt * pi * 1.0
    b.step(50)
    np.testing.assert_array_equal(b.phi(), f)
    np.testing.assert_array_equal(b.pi(), pi)


@pytest.mark.parametrize("D,N,m", [(2, 32, (1, 3)), (3, 16, (2, 1, 1))])
def test_linear_waves_follow_discrete_dispersion(D, N, m):
    """Small waves about the true vacuum phi = 1 (mass^2 = V''(1) = 2(c+6))
    obey the leapfrog dispersion cos(theta) = 1 - dt^2 Omega^2 / 2 with
    Omega^2 = m^2 + (4/dx^2) sum sin^2(k dx / 2); from the ctor's half back-step
    the amplitude after n steps is exactly A cos(n theta)."""
    c, L, A = 4.0, 4.0, 1e-7
    dx = L / N
    dt = 0.2 * dx
    s = oracle.Sim(blank=True, D=D, nGrids=N, L=L, dt=dt, c=c)
    idx = np.indices((N,) * D)
    shape = np.cos(sum(2 * np.pi * m[a] * idx[a] / N for a in range(D)))
    s.phi()[...] = 1.0 + A * shape
    s.pi()[...] = -0.5 * dt * s.forces()
    omega2 = 2 * (c + 6) + (4 / dx ** 2) * sum(math.sin(math.pi * m[a] / N) ** 2 for a in range(D))
    theta = math.acos(1 - 0.5 * dt ** 2 * omega2)
    for n in (10, 37, 80):
        s.step(n - (0 if n == 10 else {37: 10, 80: 37}[n]))
        amp = np.sum((s.phi() - 1.0) * shape) / np.sum(shape * shape)
        assert abs(amp - A * math.cos(n * theta)) <= 2e-5 * A, (n, amp / A, math.cos(n * theta))


def test_de_sitter_scale_factor_closed_form():
    """False vacuum at rest (no bubbles: phi = 0, V = 1, K = G = 0), expanding:
    a'' = a / (3 Mpl^2), and the leapfrog of a (src/bubble.cpp:858-865) from
    the ctor's a' = H0 - dt a''/2 (:273-276) is a_n = cosh(n mu) + B sinh(n mu)
    with cosh(mu) = 1 + dt^2 / (6 Mpl^2)."""
    Mpl, dt = 10.0, 0.05
    s = oracle.Sim(readxns=0, xns=np.zeros((0, 3)), D=3, nGrids=4, L=1.0, dt=dt, c=2.0,
                   expansion=1, Mpl=Mpl)
    kappa = 1.0 / (3 * Mpl ** 2)
    H0 = math.sqrt(kappa)
    # cosh(mu) - 1 = 2 sinh^2(mu/2) = dt^2 kappa / 2; a_1 - cosh(mu) = dt H0
    mu = 2 * math.asinh(0.5 * dt * math.sqrt(kappa))
    B = dt * H0 / math.sinh(mu)
    for n in range(1, 201):
        s.step(1)
        a = s.energy_row()[1]
        assert abs(a - (math.cosh(n * mu) + B * math.sinh(n * mu))) <= 1e-13 * a
    # continuum limit: a(t) ~ exp(H0 t)
    assert abs(a - math.exp(H0 * 200 * dt)) / a < 1e-4
    row = s.energy_row()
    assert row[3] == 0.0 and row[4] == 0.0 and abs(row[5] - 1.0) < 1e-14


def test_conformal_time_de_sitter_converges_to_the_closed_form():
    """Conformal time (ALPHAVAL = 1, src/bubble.cpp:751-755): false vacuum at
    rest gives a'' = 2 a^3 V / (3 Mpl^2), whose solution from a(0) = 1,
    a'(0) = H0 is a(tau) = 1 / (1 - H0 tau).  The leapfrog of a approaches it
    at second order in dt (a wrong power of a would miss it)."""
    Mpl, tau = 10.0, 10.0
    H0 = math.sqrt(1.0 / (3 * Mpl ** 2))
    exact = 1.0 / (1.0 - H0 * tau)
    errs = []
    for dt in (0.05, 0.025):
        s = oracle.Sim(readxns=0, xns=np.zeros((0, 3)), D=3, nGrids=4, L=1.0, dt=dt, c=2.0,
                       expansion=1, Mpl=Mpl, alpha=1)
        s.step(int(round(tau / dt)))
        errs.append(abs(s.energy_row()[1] - exact) / exact)
    assert errs[0] < 1e-3
    assert 3.5 < errs[0] / errs[1] < 4.5, errs


def test_gradient_energy_of_a_plane_wave():
    """Forward-difference gradient energy (src/bubble.cpp:790-808) of
    A cos(2 pi m i / N) along one axis: A^2 sin^2(pi m / N) / dx^2."""
    N, L, A, m = 32, 4.0, 0.01, 3
    dx = L / N
    s = oracle.Sim(blank=True, D=2, nGrids=N, L=L, dt=0.2 * dx, c=4.0)
    i = np.arange(N)[:, None] * np.ones((1, N))
    s.phi()[...] = A * np.cos(2 * np.pi * m * i / N)
    s.step(1)   # energies are measured at the incoming phi
    eg = s.energy_row()[4]
    assert abs(eg - A ** 2 * math.sin(math.pi * m / N) ** 2 / dx ** 2) <= 1e-13 * eg


def test_energy_conservation_static_box():
    s = oracle.Sim(D=2, nGrids=256, L=16.0, dt=0.0125, c=4.0, readxns=1, nBubbles=5)
    e0 = s.energy_row()[2]
    worst = 0.0
    for _ in range(20):
        s.step(40)
        worst = max(worst, abs(s.energy_row()[2] - e0) / e0)
    assert worst < 2e-3


def test_empty_box_energies():
    s = oracle.Sim(readxns=0, xns=np.zeros((0, 3)), D=3, nGrids=6, L=2.0, dt=0.05, c=4.0)
    row = s.energy_row()
    assert abs(row[2] - 1.0) < 1e-14 and row[3] == 0.0 and row[4] == 0.0
    s.step(3)
    assert np.all(s.phi() == 0.0) and np.all(s.pi() == 0.0)


# --------------------------------------------------------------------------
# Fourier transforms and spectrum (src/bubble.cpp:642-709)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 15, 16, 49, 64, 97, 100])
def test_dft_matches_direct_sum(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    np.testing.assert_allclose(oracle.dft(x), W @ x, atol=1e-13 * np.abs(x).sum())
    np.testing.assert_allclose(oracle.dft(x, +1), W.conj() @ x, atol=1e-13 * np.abs(x).sum())


@pytest.mark.parametrize("shape", [(16,), (15,), (16, 16), (15, 15), (8, 8, 8), (9, 9, 9), (6, 6, 6)])
def test_r2c_matches_numpy_rfftn(shape):
    rng = np.random.default_rng(len(shape))
    f = rng.standard_normal(shape)
    np.testing.assert_allclose(oracle.r2c(f), np.fft.rfftn(f), atol=1e-13 * np.abs(f).sum())


def _shell_counts(D, N, L):
    """Weights and k sums of each |k| shell by enumerating the half-complex
    modes with the reference's weights (1 on j = 0 and j = N/2, else 2)."""
    nb = int(math.sqrt(D) * N / 2) + 1
    ws, ks = np.zeros(nb), np.zeros(nb)
    signed = lambda i: i if i <= N // 2 else i - N
    rows = list(np.ndindex(*((N,) * (D - 1)))) if D > 1 else [()]
    for r in rows:
        n2 = sum(signed(i) ** 2 for i in r)
        for j in range(N // 2 + 1):
            n = math.sqrt(n2 + j * j)
            b = int(n)
            if b >= nb:
                continue
            w = 1.0 if (j == 0 or j == N // 2) else 2.0
            ws[b] += w
            ks[b] += w * n * 2 * math.pi / L
    return ws, ks


@pytest.mark.parametrize("D,N,m", [(2, 16, (3, 2)), (3, 8, (1, 2, 1)), (2, 15, (4, 3)), (1, 64, (5,)),
                                   (3, 9, (2, 0, 3))])
def test_spectrum_of_a_single_mode(D, N, m):
    L = 2.0
    s = oracle.Sim(blank=True, D=D, nGrids=N, L=L, dt=0.01, c=4.0)
    idx = np.indices((N,) * D)
    s.phi()[...] = np.cos(sum(2 * np.pi * m[a] * idx[a] / N for a in range(D)))
    k, kp, delta = s.spectrum()
    ws, ks = _shell_counts(D, N, L)
    nonempty = ws > 0
    np.testing.assert_allclose(k, ks[nonempty] / ws[nonempty], rtol=1e-13)
    np.testing.assert_allclose(kp, k, rtol=0)
    dx = L / N
    nm = math.sqrt(sum(x * x for x in m))
    pref = {1: 1 / math.pi, 2: 1 / (2 * math.pi), 3: 1 / (2 * math.pi ** 2)}[D]
    Pnorm = dx ** (2 * D) / L ** D
    w = 1.0 if (m[-1] == 0 or 2 * m[-1] == N) else 2.0
    amp2 = (N ** D / 2.0) ** 2 * (4.0 if m[-1] == 0 and all(x == 0 for x in m) else 1.0)
    if m[-1] == 0:
        amp2 = (N ** D / 2.0) ** 2   # the j = 0 plane stores both k and -k
        w_total = 2.0
    else:
        w_total = w
    want = w_total * pref * (nm * 2 * math.pi / L) ** D * Pnorm * amp2
    b = int(nm)
    full = np.zeros(len(ws))
    full[nonempty] = delta
    assert abs(full[b] * ws[b] - want) <= 1e-12 * want
    others = np.delete(full * ws, b)
    assert np.all(np.abs(others) <= 1e-20 * want + 1e-25)


# --------------------------------------------------------------------------
# thermal initial conditions (src/bubble.cpp:459-523)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("D,N", [(2, 32), (3, 16), (2, 15), (3, 9), (2, 33), (3, 7)])
def test_thermal_modes_hermitian_and_c2r_matches_numpy(D, N):
    L, T, m = 4.0, 0.6, 2.0
    X = oracle.thermal_modes(D, N, L, T, m, velocity=False)
    rows = X.reshape(-1, N // 2 + 1)
    for j in sorted({0, N // 2}):
        for r in range(rows.shape[0]):
            idx = np.unravel_index(r, (N,) * (D - 1))
            pr = np.ravel_multi_index(tuple((-np.array(idx)) % N), (N,) * (D - 1))
            if pr == r:
                assert rows[r, j].imag == 0.0
            else:
                assert rows[r, j] == np.conj(rows[pr, j])
    field = oracle.thermal_field(D, N, L, T, m, velocity=False)
    # numpy's irfftn takes the half array as the stored half of a Hermitian
    # array, as FFTW's c2r does; for odd N the reference's "fix" of plane
    # j = N/2 (a regular plane there) only changes the data, not the rule
    want = np.fft.irfftn(X, s=(N,) * D, axes=tuple(range(D))) * N ** D / math.sqrt(L ** D)
    np.testing.assert_allclose(field, want, atol=1e-13 * np.abs(want).max())
    # Parseval: sum_x field^2 = N^D L^-D sum over the Hermitian-extended array
    w = np.full(N // 2 + 1, 2.0)
    w[0] = 1.0
    if N % 2 == 0:
        w[N // 2] = 1.0
    want_sq = N ** D / L ** D * np.sum(w * np.abs(rows) ** 2)
    assert abs(np.sum(field ** 2) - want_sq) <= 1e-12 * want_sq


def test_thermal_mode_variance_is_the_bose_einstein_spectrum():
    D, N, L, T, m = 3, 16, 4.0, 0.8, 2.0
    for velocity in (False, True):
        X = oracle.thermal_modes(D, N, L, T, m, velocity=velocity).reshape(-1, N // 2 + 1)
        ratio, count = 0.0, 0
        signed = lambda i: i if i <= N // 2 else i - N
        for r in range(X.shape[0]):
            i0, i1 = divmod(r, N)
            for j in range(1, N // 2):   # generic modes: <|X|^2> = P(k)
                k2 = (2 * math.pi / L) ** 2 * (signed(i0) ** 2 + signed(i1) ** 2 + j * j)
                om = math.sqrt(k2 + m * m)
                nk = 1.0 / (math.exp(om / T) - 1.0)
                P = nk * om if velocity else nk / om
                ratio += abs(X[r, j]) ** 2 / P
                count += 1
        assert abs(ratio / count - 1.0) < 0.05, ratio / count


def test_box_sums_are_exact_to_a_few_ulp_whatever_the_thread_split():
    """The oracle's box sums are compensated: the ctor's <phi> (sum of
    phi dxLD over the lattice) equals the exactly rounded sum (math.fsum) of
    the same per-site products to 2 ulp, with 1 or 7 threads."""
    kw = dict(D=3, nGrids=48, L=6.0, c=2.0, readxns=-3, Trad=0.3, dt=0.025)
    dxLD = (1.0 / 48) ** 3
    vals = []
    for nth in (1, 7):
        o = oracle.Sim(nthreads=nth, **kw)
        exact = math.fsum((o.phi().ravel() * dxLD).tolist())
        got = o.energy_row()[10]
        assert abs(got - exact) <= 2 * np.spacing(abs(exact)), (nth, got, exact)
        vals.append(got)
    assert abs(vals[0] - vals[1]) <= 2 * np.spacing(abs(vals[0]))


@pytest.mark.parametrize("kw", [
    dict(D=2, nGrids=48, L=6.0, c=4.0, readxns=1, nBubbles=5, dt=0.025),
    dict(D=3, nGrids=24, L=4.0, c=4.0, readxns=1, nBubbles=4, dt=0.025),
    dict(D=3, nGrids=20, L=2.5, c=2.0, readxns=1, nBubbles=3, dt=0.025, expansion=1, Mpl=10.0),
])
def test_seed_sites_and_ctor_pieces_reproduce_the_whole_ctor(kw):
    """The pieces the full-size parity tests use -- seeding evaluated at
    chosen sites, and the ctor sequence run on fields copied in -- give the
    whole ctor's phi, pi, energy row and scalars bit for bit."""
    o = oracle.Sim(**kw)
    phi0 = o.phi().copy()
    sites = oracle.seed_sites(o.xns(), np.arange(o.n), **kw)
    np.testing.assert_array_equal(sites.reshape(o.shape()), phi0)
    b = oracle.Sim(blank=True, **kw)
    b.phi()[...] = phi0                 # mode 1 starts at rest: raw pi = 0
    b.ctor()
    np.testing.assert_array_equal(b.pi(), o.pi())
    np.testing.assert_array_equal(b.energy_row(), o.energy_row())
    np.testing.assert_array_equal(b.scalars(), o.scalars())
    o.step(3)
    b.step(3)
    np.testing.assert_array_equal(b.phi(), o.phi())


@pytest.mark.parametrize("D,N", [(1, 64), (2, 24), (3, 10)])
def test_ctor_energy_maps_sum_to_the_initial_energies(D, N):
    """The ctor's local energy densities (src/bubble.cpp:239-265) times the
    cell volume sum to the initial energy row, with pi before the back-step."""
    kw = dict(D=D, nGrids=N, L=4.0, c=3.0, dt=0.01)
    s = oracle.Sim(blank=True, **kw)
    rng = np.random.default_rng(7 + D)
    phi = 0.4 + 0.3 * rng.standard_normal((N,) * D)
    pi0 = 0.2 * rng.standard_normal((N,) * D)
    s.phi()[...] = phi
    s.pi()[...] = pi0
    s.ctor()
    row = s.energy_row()
    ekin, egrad, epot = s.energy_maps(phi, pi0, ctor=True)
    dxLD = (1.0 / N) ** D
    np.testing.assert_allclose([ekin.sum() * dxLD, egrad.sum() * dxLD, epot.sum() * dxLD], row[3:6], rtol=1e-13)
    # a plane wave has the closed-form gradient energy of forward differences
    m = (2,) + (1,) * (D - 1)
    idx = np.indices((N,) * D)
    wave = np.cos(sum(2 * np.pi * m[a] * idx[a] / N for a in range(D)))
    _, g, _ = s.energy_maps(wave, np.zeros_like(wave), ctor=True)
    dx = 4.0 / N
    fwd = [np.roll(wave, -1, axis=a) - wave for a in range(D)]
    np.testing.assert_allclose(g, 0.5 * sum(f * f for f in fwd) / dx ** 2, rtol=1e-12, atol=1e-14)
    # the maps of the ctor and of measureEnergies at a = a_kin = 1 agree
    for x, y in zip(s.energy_maps(phi, pi0, ctor=True), s.energy_maps(phi, pi0, a=1.0, a_kin=1.0)):
        np.testing.assert_allclose(x, y, rtol=2e-16, atol=0)


@pytest.mark.parametrize("D,N", [(1, 64), (2, 24), (3, 10)])
def test_step_energy_maps_sum_to_the_measured_energies(D, N):
    """The maps of measureEnergies (src/bubble.cpp:790-833) see phi before the
    step's drift, pi after its kick and the scale factors before the step:
    their sums are the step's gradient and potential energies and its raw
    half-step kinetic energy (EkinHalfPrev after the step)."""
    kw = dict(D=D, nGrids=N, L=4.0, c=3.0, dt=0.01, readxns=-3, Trad=0.3, expansion=1, Mpl=2.0)
    s = oracle.Sim(**kw)
    s.step(5)
    phi_prev = s.phi().copy()
    a, ap = s.scalars()[:2]
    s.step(1)
    row, sc = s.energy_row(), s.scalars()
    assert (a != 1.0) == (D == 3)        # the background expands in 3D only (src/bubble.hpp:77)
    ekin, egrad, epot = s.energy_maps(phi_prev, s.pi(), a=a, a_kin=a + 0.5 * 0.01 * ap)
    dxLD = (1.0 / N) ** D
    np.testing.assert_allclose([egrad.sum() * dxLD, epot.sum() * dxLD], row[4:6], rtol=1e-12)
    np.testing.assert_allclose(ekin.sum() * dxLD, sc[5], rtol=1e-12)
    # a slip of the phi or pi time level is seen
    assert abs(s.energy_maps(s.phi(), s.pi(), a=a, a_kin=a + 0.005 * ap)[1].sum() * dxLD - row[4]) > 1e-9 * row[4]


@pytest.mark.parametrize("D,readxns,N", [(1, 1, 512), (2, -2, 128), (2, -1, 128), (3, -2, 64)])
@pytest.mark.parametrize("R", [3.1, 7.6, 12.05])
def test_wall_finder_matches_a_brute_force_scan(D, readxns, N, R):
    """WallFinder (src/bubble.cpp:576-621) on a monotone tanh wall at radius R
    along the tracked line: Rmin = the outermost radius with phi > phiZero,
    Rmax = the first with phi <= phiMax, found here over all radii at once."""
    L, c = 32.0, 4.0
    w = oracle.WallTracker(D, N, L, c, readxns, dt=0.05)
    dx = L / N
    shape = (N,) * D
    phi = np.zeros(shape)
    i = np.arange(N)
    r = dx * np.sqrt((i + 0.5) ** 2 + 0.25 * (D - 1))
    prof = 0.5 * (1 - np.tanh((r - R) / 0.7))
    idx = (i if D == 1 else (w.centre[0] + i) % N,) + tuple(np.full(N, w.centre[d]) for d in range(1, D))
    phi[idx] = prof
    w.find(phi)
    ok = (r < w.max_radius) & (np.arange(N) >= int(w.Rn / (2 * dx)))
    out = np.nonzero(ok & (prof <= w.phi_max))[0]
    inside = np.nonzero(ok & (prof > w.phi_zero))[0]
    if len(out):
        assert w.rmax == r[out[0]] and w.rmin == r[inside[inside < out[0]][-1]]
        assert w.width == w.rmax - w.rmin > 0
    else:   # no point at the barrier top before the scan's end: Rmax < Rmin, the width is kept
        assert w.rmax == 0 and w.rmin == r[inside[-1]] and w.width == 0 and w.gamma(w.width) == 0
    # the next scan starts just inside the wall and finds it again
    before = (w.rmin, w.rmax)
    w.find(phi)
    assert (w.rmin, w.rmax) == before or w.rmax >= 0.98 * w.max_radius


def test_wall_tracker_schedule_and_line_format():
    w = oracle.WallTracker(1, 256, 16.0, 4.0, 1, dt=0.0125, every=100)
    assert w.it_start == int(math.pi / math.sqrt(20.0) / 0.0125) == 56
    phi = 0.5 * (1 - np.tanh((np.arange(256) * 16.0 / 256 - 3.0) / 0.5))
    assert w.track(55, 0.6875, phi) == ""
    first = w.track(56, 0.7, phi)
    assert w.dini == w.width > 0
    assert first == "0.700 1.0000 %.3f %.3f %.3f\n " % (w.width, w.rmin, w.rmax)
    assert w.track(150, 1.875, phi) == ""                          # not a multiple of every
    assert w.track(200, 2.5, phi).startswith("2.500 1.0000 ")
    assert w.track(700, 8.75, phi) == ""                           # t >= L / 2
    assert oracle.WallTracker(2, 64, 16.0, 4.0, 1, dt=0.05).track(14, 0.7, np.zeros((64, 64))) == ""


# --------------------------------------------------------------------------
# critical-bubble fits and seeded profiles (src/bubble.cpp:81-108,324-393)
# --------------------------------------------------------------------------
# The fits as src/bubble.cpp:81-99 states them, evaluated once at 40 digits
# with Python's decimal module (independent of glibc's exp/pow):
#   phiC = 0.997 (1 - exp(-0.688 c)), R = 0.823 c^0.203 (x 1.3 in 3D),
#   lw = 0.496 c^-0.531.
FIT_KAT = {   # c: (phiC, R in 1D/2D, R in 3D, lw)
    1.0: (0.49592751564964788, 0.82299999999999995, 1.0699000000000001, 0.496),
    2.0: (0.74517087806114957, 0.94734665136274865, 1.2315506467715733, 0.34326912958076705),
    4.0: (0.93339126714494247, 1.0904807750282057, 1.4176250075366674, 0.23756793411922872),
    7.5: (0.9912755254134028, 1.2389029061226764, 1.6105737779594793, 0.17014694104868963),
}


@pytest.mark.parametrize("c", sorted(FIT_KAT))
def test_bubble_fit_known_answers(c):
    phiC, R2, R3, lw = FIT_KAT[c]
    for D, R in ((1, R2), (2, R2), (3, R3)):
        got = oracle.bubble_fit(D, c)
        np.testing.assert_allclose(got, [phiC, R, lw], rtol=4e-16, atol=0)
    # nucleated 1.3 times larger in 3D (src/bubble.cpp:87-94)
    assert abs(oracle.bubble_fit(3, c)[1] / oracle.bubble_fit(2, c)[1] - 1.3) <= 4e-16


def _o3_bounce_centre(c):
    """Centre value of the O(3)-symmetric critical bubble of V (the thermal
    bounce: phi'' + (2/r) phi' = V'(phi), phi'(0) = 0, phi -> 0 far away),
    found by overshoot/undershoot shooting on phi(0)."""
    from scipy.integrate import solve_ivp

    def vp(f):
        return 2 * c * f - 3 * (2 * c + 4) * f * f + 4 * (c + 3) * f ** 3

    def overshoots(f0):
        r0 = 1e-6
        y0 = [f0 + vp(f0) * r0 * r0 / 6, vp(f0) * r0 / 3]
        crossed = lambda r, y: y[0]
        crossed.terminal = True
        turned = lambda r, y: y[1]
        turned.terminal, turned.direction = True, 1
        sol = solve_ivp(lambda r, y: [y[1], vp(y[0]) - 2 / r * y[1]], [r0, 200], y0,
                        events=[crossed, turned], rtol=1e-9, atol=1e-12)
        return len(sol.t_events[0]) > 0

    lo, hi = c / (2 * c + 6), 1.0 - 1e-15     # barrier top .. true vacuum
    while hi - lo > 1e-9:
        mid = 0.5 * (lo + hi)
        lo, hi = (lo, mid) if overshoots(mid) else (mid, hi)
    return 0.5 * (lo + hi)


@pytest.mark.parametrize("c", [1.0, 2.0, 4.0, 8.0])
def test_bubble_centre_fit_is_the_thermal_critical_bubble(c):
    """phizero (src/bubble.cpp:81-84) fits the centre value of the critical
    bubble: the O(3) bounce of this potential reproduces it within 0.5 %
    over c = 1..8 (a changed constant, sign or exponent misses by far more)."""
    got = oracle.bubble_fit(3, c)[0]
    bounce = _o3_bounce_centre(c)
    assert abs(got - bounce) <= 5e-3 * bounce, (got, bounce)


def _profile(r, c, D):
    phiC, R2, R3, lw = FIT_KAT[c]
    R = R3 if D == 3 else R2
    return (phiC / 2.0) * (np.tanh((R - r) / lw) + 1.0)


def _coords(D, N, L):
    """Cell-centred coordinates x = L (i + 1/2) / N (src/gridfuncs.cpp:77-85)."""
    x = (np.arange(N) + 0.5) * L / N
    return np.meshgrid(*([x] * D), indexing="ij")


@pytest.mark.parametrize("D,N,L,c,mode,DR", [(2, 40, 8.0, 4.0, -2, 0.0), (3, 24, 6.0, 2.0, -2, 0.0),
                                            (2, 48, 6.0, 4.0, -1, 0.3), (3, 20, 5.0, 7.5, -1, 0.25),
                                            (1, 96, 6.0, 4.0, 1, 0.0), (1, 64, 1.5, 1.0, 1, 0.0)])
def test_seeded_profiles_at_known_distances(D, N, L, c, mode, DR):
    """phi after seeding = the tanh profile of the fits at each site's
    distance from the centres, added in order and capped at phiC
    (NucleateOneBubble :356-369, NucleateTwoDiagonalBubbles :373-393,
    Initialize1D :324-335: the 1D bubble sits on the box edge, profiles
    around x = 0 and x = L).  In the 1D box with L = 1.5 the two profiles
    overlap and the cap phiC is reached."""
    s = oracle.Sim(D=D, nGrids=N, L=L, c=c, readxns=mode, DeltaR=DR, dt=0.2 * L / N)
    phiC = FIT_KAT[c][0]
    X = _coords(D, N, L)
    if D == 1:
        centres = [(0.0,), (L,)]
    elif mode == -2:
        centres = [(L / 2,) * D]
    else:
        centres = [(L / 4 + DR, L / 4 + DR, L / 2)[:D], (3 * L / 4 - DR, 3 * L / 4 - DR, L / 2)[:D]]
    want = np.zeros((N,) * D)
    for ctr in centres:
        r = np.sqrt(sum((X[a] - ctr[a]) ** 2 for a in range(D)))
        want = np.minimum(want + _profile(r, c, D), phiC)
    np.testing.assert_allclose(s.phi(), want, rtol=0, atol=4e-15)   # ~ulp of 1 + tanh, times phiC
    assert s.phi().max() <= phiC
    if L == 1.5:
        assert np.sum(s.phi() == phiC) > 0
    assert s.phi().max() > 0.9 * phiC     # the lattice crosses the bubble interior


@pytest.mark.parametrize("D,c", [(2, 4.0), (3, 2.0), (3, 7.5)])
def test_profile_is_half_the_centre_value_at_the_radius(D, c):
    """A centre placed exactly R from a site gives phi = phiC / 2 there."""
    phiC, R2, R3, _ = FIT_KAT[c]
    R = R3 if D == 3 else R2
    N, L = 16, 4.0
    site = (3, 5, 7)[:D]
    x = [(i + 0.5) * L / N for i in site]
    ctr = np.array([[x[0] + R] + x[1:] + [0.0] * (3 - D)])
    idx = np.ravel_multi_index(site, (N,) * D)
    got = oracle.seed_sites(ctr, [idx], D=D, nGrids=N, L=L, c=c)[0]
    # (x + R) - x is R to an ulp of x + R; the profile's slope there is phiC / (2 lw)
    tol = 2 * np.spacing(x[0] + R) * phiC / (2 * FIT_KAT[c][3]) + 2 * np.spacing(phiC)
    assert abs(got - phiC / 2) <= tol
    # one wall width further out the profile is (phiC / 2)(1 - tanh 1)
    ctr[0, 0] = x[0] + R + FIT_KAT[c][3]
    got = oracle.seed_sites(ctr, [idx], D=D, nGrids=N, L=L, c=c)[0]
    assert abs(got - phiC / 2 * (1 - math.tanh(1.0))) <= 2 * tol


# --------------------------------------------------------------------------
# ctor back-step and initial scale factor (src/bubble.cpp:239-291)
# --------------------------------------------------------------------------
def _numpy_ctor(phi, pi0, D, N, L, c, dt, expanding, Mpl, alpha):
    """The ctor's sequence evaluated independently: energies with forward
    differences / dx, H0 from the Friedmann equation, a' = a H0 - dt/2 a'',
    pi_{-1/2} = pi_0 - dt/2 F(phi_0), and the kinetic energy of pi_{-1/2}."""
    dx = L / N
    dxLD = (1.0 / N) ** D
    ax = tuple(range(D))
    lap = sum(np.roll(phi, -1, a) - 2 * phi + np.roll(phi, 1, a) for a in ax)
    grad = 0.5 * sum(((np.roll(phi, -1, a) - phi) / dx) ** 2 for a in ax)
    V = c * phi ** 2 - (2 * c + 4) * phi ** 3 + (c + 3) * phi ** 4 + 1.0
    K, G, P = (0.5 * pi0 ** 2).sum() * dxLD, grad.sum() * dxLD, V.sum() * dxLD
    E = K + G + P
    H0 = math.sqrt(E / (3 * Mpl ** 2)) if expanding else 0.0
    if not expanding:
        add = 0.0
    elif alpha == 0:
        add = -(2 * K - P) / (3 * Mpl ** 2)
    else:
        add = (2 * P - K + G) / (3 * Mpl ** 2)
    ap = H0 - 0.5 * dt * add
    F = lap / dx ** 2 - _vprime(phi, c)
    pim = pi0 - 0.5 * dt * F
    a_mh = 1.0 - 0.5 * dt * ap
    ekh = (0.5 * a_mh ** (-2 * D) * pim ** 2).sum() * dxLD
    return dict(row=(E, K, G, P), H0=H0, ap=ap, pim=pim, ekh=ekh, F=F)


@pytest.mark.parametrize("D,N,expanding,alpha", [(2, 20, 0, 0), (3, 10, 0, 0), (3, 10, 1, 0), (3, 12, 1, 1)])
def test_ctor_back_step_and_scale_factor_on_a_random_field(D, N, expanding, alpha):
    L, c, dt, Mpl = 5.0, 3.0, 0.05, 0.8
    kw = dict(D=D, nGrids=N, L=L, c=c, dt=dt, expansion=expanding, Mpl=Mpl, alpha=alpha)
    rng = np.random.default_rng(100 * D + N + alpha)
    phi = 0.5 + 0.3 * rng.standard_normal((N,) * D)
    pi0 = 0.4 * rng.standard_normal((N,) * D)
    s = oracle.Sim(blank=True, **kw)
    s.phi()[...] = phi
    s.pi()[...] = pi0
    s.ctor()
    want = _numpy_ctor(phi, pi0, D, N, L, c, dt, expanding, Mpl, alpha)
    row, sc = s.energy_row(), s.scalars()
    np.testing.assert_allclose(row[2:6], want["row"], rtol=1e-13)
    assert abs(sc[1] - want["ap"]) <= 1e-13 * max(abs(want["ap"]), 1e-300) + (0 if expanding else 0.0)
    assert row[14] == sc[2] and abs(sc[2] - want["H0"]) <= 1e-14 * max(want["H0"], 1e-300)
    assert sc[0] == 1.0 and sc[4] == 0.0
    scale = np.abs(pi0).max() + 0.5 * dt * np.abs(want["F"]).max()
    assert np.max(np.abs(s.pi() - want["pim"])) <= 1e-14 * scale
    assert abs(sc[5] - want["ekh"]) <= 1e-13 * want["ekh"]
    # the back-step moves pi by half a step of the force; EkinHalfPrev is the
    # kinetic energy of pi_{-1/2} with a_{-1/2}, not of pi_0
    assert abs(sc[5] - row[3]) > 1e-6 * row[3]
    if expanding:
        assert want["ap"] != want["H0"]


@pytest.mark.parametrize("alpha", [0, 1])
def test_whole_thermal_ctor_back_step(alpha):
    """The whole ctor of a thermal start (mode -3: pi_0 drawn as well) leaves
    pi_{-1/2}, the scalars and EkinHalfPrev of the independent evaluation of
    its own phi_0 and pi_0 (thermal fields, pinned against numpy above)."""
    D, N, L, c, dt, Mpl, Trad = 3, 16, 8.0, 2.0, 0.04, 1.5, 0.5
    s = oracle.Sim(D=D, nGrids=N, L=L, c=c, dt=dt, readxns=-3, Trad=Trad, expansion=1, Mpl=Mpl, alpha=alpha)
    T, m = Trad * math.sqrt(2 * (c + 6)), math.sqrt(2 * c)
    phi = oracle.thermal_field(D, N, L, T, m, velocity=False)
    pi0 = oracle.thermal_field(D, N, L, T, m, velocity=True)
    np.testing.assert_array_equal(s.phi(), phi)
    want = _numpy_ctor(phi, pi0, D, N, L, c, dt, True, Mpl, alpha)
    row, sc = s.energy_row(), s.scalars()
    np.testing.assert_allclose(row[2:6], want["row"], rtol=1e-13)
    assert abs(sc[1] - want["ap"]) <= 1e-13 * abs(want["ap"])
    scale = np.abs(pi0).max() + 0.5 * dt * np.abs(want["F"]).max()
    assert np.max(np.abs(s.pi() - want["pim"])) <= 1e-14 * scale
    assert abs(sc[5] - want["ekh"]) <= 1e-13 * want["ekh"]


# --------------------------------------------------------------------------
# Friedmann update (src/bubble.cpp:747-757,858-875)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("alpha,Mpl,T", [(0, 1.0, 2.0), (1, 2.0, 1.0)])
def test_hubble_constraint_converges_at_second_order(alpha, Mpl, T):
    """With a'' from the second Friedmann equation, the lattice system
    conserves the first one, H^2 = a^(2 alpha) Etot / (3 Mpl^2), exactly in
    continuous time (the leapfrog forces are the gradients of the lattice
    energies).  So the constraint C written in column 15 (:873-875) is the
    integrator's O(dt^2) error: C / H^2 at a fixed time falls 4x when dt
    halves.  The start has K ~ V, so a wrong kinetic term in a'' (cosmic
    time -a (2K - V), conformal a^3 (2V - K + G)) would leave an O(1)
    constraint violation that does not converge."""
    vals, a_end = [], []
    for dt in (0.05, 0.025, 0.0125):
        s = oracle.Sim(D=3, nGrids=16, L=8.0, c=2.0, readxns=-3, Trad=0.5, dt=dt, expansion=1, Mpl=Mpl,
                       alpha=alpha)
        r0 = s.energy_row()
        assert r0[3] > 0.5 * r0[5] and r0[15] == 0.0           # kinetic ~ potential at the start
        s.step(int(round(T / dt)))
        r = s.energy_row()
        vals.append(r[15] / r[14] ** 2)
        a_end.append(r[1])
    assert a_end[-1] > 1.5                                      # the background has expanded
    assert abs(vals[-1]) < 1e-3
    for coarse, fine in zip(vals, vals[1:]):
        assert 3.6 < coarse / fine < 4.6, vals


@pytest.mark.parametrize("kw", [dict(D=2, nGrids=40, L=5.0, c=4.0, readxns=1, nBubbles=4, dt=0.025),
                                dict(D=3, nGrids=16, L=4.0, c=2.0, readxns=-3, Trad=0.4, dt=0.025)])
def test_energy_cadence_gives_the_every_step_rows(kw):
    """A static box measured only on every 40th step (and the kinetic sum on
    the step before) prints the every-step oracle's rows bit for bit, on
    the same fields (the full-length fixtures are generated this way)."""
    a, b = oracle.Sim(**kw), oracle.Sim(**kw)
    b.set_energy_cadence(40)
    for _ in range(3):
        a.step(40)
        b.step(40)
        np.testing.assert_array_equal(a.energy_row(), b.energy_row())
    a.step(17)
    b.step(17)
    np.testing.assert_array_equal(a.phi(), b.phi())
    np.testing.assert_array_equal(a.pi(), b.pi())
    # resumed mid-way at step 137 with the step count
    c = oracle.Sim(blank=True, **kw)
    c.phi()[...] = b.phi()
    c.pi()[...] = b.pi()
    c.set_scalars(b.scalars())
    c.set_energy_cadence(40, 137)
    a.step(23)
    c.step(23)
    np.testing.assert_array_equal(a.energy_row()[2:], c.energy_row()[2:])
