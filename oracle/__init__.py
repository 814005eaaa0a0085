"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker.  The product package
(coolbubble_amd) never imports this module.

The oracle restates the reference's algorithm (Bubble ctor, oneStep,
measureEnergies, ComputePowerSpectrum; src/bubble.cpp:156-876) in plain C.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Params(ctypes.Structure):
    """Mirror of cbo_params (oracle/cb_oracle.h)."""
    _fields_ = [
        ("D", ctypes.c_int), ("nGrids", ctypes.c_int), ("L", ctypes.c_double),
        ("c", ctypes.c_double), ("readxns", ctypes.c_int), ("dt", ctypes.c_double),
        ("nBubbles", ctypes.c_int), ("Trad", ctypes.c_double), ("DeltaR", ctypes.c_double),
        ("test_wave", ctypes.c_int), ("expansion", ctypes.c_int), ("alpha", ctypes.c_int),
        ("Mpl", ctypes.c_double), ("rand_seed", ctypes.c_uint), ("thermal_base", ctypes.c_uint64),
    ]


def default_params(**kw):
    """Reference defaults (src/def.h:29-58), overridable by keyword."""
    p = dict(D=2, nGrids=512, L=32.0, c=4.0, readxns=1, dt=0.004, nBubbles=5, Trad=0.2,
             DeltaR=0.0, test_wave=0, expansion=0, alpha=0, Mpl=200.0, rand_seed=42,
             thermal_base=12345)
    p.update(kw)
    return Params(**p)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.cbo_create.restype = ctypes.c_void_p
        L.cbo_create.argtypes = [ctypes.POINTER(Params), dp, ctypes.c_int, ctypes.c_int]
        L.cbo_destroy.argtypes = [ctypes.c_void_p]
        L.cbo_create_blank.restype = ctypes.c_void_p
        L.cbo_create_blank.argtypes = [ctypes.POINTER(Params), ctypes.c_int]
        L.cbo_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cbo_energy_row.argtypes = [ctypes.c_void_p, dp]
        L.cbo_spectrum.argtypes = [ctypes.c_void_p, dp, dp, dp, ctypes.c_int]
        L.cbo_spectrum.restype = ctypes.c_int
        L.cbo_nsites.argtypes = [ctypes.c_void_p]
        L.cbo_nsites.restype = ctypes.c_longlong
        L.cbo_phi.argtypes = [ctypes.c_void_p]
        L.cbo_phi.restype = dp
        L.cbo_pi.argtypes = [ctypes.c_void_p]
        L.cbo_pi.restype = dp
        L.cbo_xns.argtypes = [ctypes.c_void_p, dp, ctypes.c_int]
        L.cbo_xns.restype = ctypes.c_int
        L.cbo_get_scalars.argtypes = [ctypes.c_void_p, dp]
        L.cbo_set_scalars.argtypes = [ctypes.c_void_p, dp]
        L.cbo_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cbo_set_energy_cadence.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong]
        L.cbo_ctor.argtypes = [ctypes.c_void_p]
        L.cbo_seed_sites.argtypes = [ctypes.POINTER(Params), dp, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_longlong), ctypes.c_int, dp]
        L.cbo_forces.argtypes = [ctypes.c_void_p, dp]
        L.cbo_energy_maps.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      dp, dp, dp]
        L.cbo_V.restype = ctypes.c_double
        L.cbo_V.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.cbo_Vpr.restype = ctypes.c_double
        L.cbo_Vpr.argtypes = [ctypes.c_double, ctypes.c_double]
        for f in (L.cbo_iGridM, L.cbo_iGridP):
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int] * 4
        L.cbo_bubble_fit.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.cbo_generate_xns.restype = ctypes.c_int
        L.cbo_generate_xns.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_uint, dp, ctypes.c_int]
        L.cbo_mt64.argtypes = [ctypes.c_uint64, ctypes.c_long, ctypes.POINTER(ctypes.c_uint64)]
        L.cbo_normals.argtypes = [ctypes.c_uint64, ctypes.c_long, dp]
        L.cbo_dft_ld.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
        L.cbo_r2c.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.cbo_thermal_modes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_uint64, dp, dp]
        L.cbo_thermal_field.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_uint64, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Sim:
    """One oracle run: the reference's Bubble object (src/bubble.hpp:34-135)."""

    def __init__(self, params=None, xns=None, nthreads=0, blank=False, **kw):
        self.p = params if params is not None else default_params(**kw)
        L = lib()
        if blank:
            self._h = L.cbo_create_blank(ctypes.byref(self.p), nthreads)
        elif xns is not None:
            xa = np.ascontiguousarray(np.asarray(xns, dtype=np.float64).reshape(-1, 3))
            self._h = L.cbo_create(ctypes.byref(self.p), _dp(xa), xa.shape[0], nthreads)
        else:
            self._h = L.cbo_create(ctypes.byref(self.p), None, 0, nthreads)
        if not self._h:
            raise RuntimeError("cbo_create failed")
        self.n = L.cbo_nsites(self._h)

    def close(self):
        if self._h:
            lib().cbo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        lib().cbo_step(self._h, n)

    def shape(self):
        return (self.p.nGrids,) * self.p.D

    def phi(self):
        return np.ctypeslib.as_array(lib().cbo_phi(self._h), shape=(self.n,)).reshape(self.shape())

    def pi(self):
        return np.ctypeslib.as_array(lib().cbo_pi(self._h), shape=(self.n,)).reshape(self.shape())

    def energy_row(self):
        row = np.zeros(16)
        lib().cbo_energy_row(self._h, _dp(row))
        return row

    def scalars(self):
        s = np.zeros(6)
        lib().cbo_get_scalars(self._h, _dp(s))
        return s

    def set_scalars(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        lib().cbo_set_scalars(self._h, _dp(s))

    def set_energy_cadence(self, every, nstep=0):
        """Static box: energies only on every `every`-th step counted from
        nstep (cbo_set_energy_cadence); rows there are the every-step rows."""
        lib().cbo_set_energy_cadence(self._h, int(every), int(nstep))

    def ctor(self):
        """Run the ctor's sequence (src/bubble.cpp:239-291) on the current
        phi and raw t = 0 pi (e.g. fields copied in from elsewhere)."""
        lib().cbo_ctor(self._h)

    def forces(self):
        out = np.zeros(self.n)
        lib().cbo_forces(self._h, _dp(out))
        return out.reshape(self.shape())

    def energy_maps(self, phi, pi, a=1.0, a_kin=1.0, ctor=False):
        """(Ekin, gradphi, potEnergy) per site of the snapshot maps for phi,
        pi: the ctor's formulas (ctor=True) or measureEnergies' with a, a_kin."""
        ph = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        pp = np.ascontiguousarray(pi, dtype=np.float64).reshape(-1)
        out = [np.zeros(self.n) for _ in range(3)]
        lib().cbo_energy_maps(self._h, _dp(ph), _dp(pp), a, a_kin, int(ctor), *(_dp(o) for o in out))
        return tuple(o.reshape(self.shape()) for o in out)

    def xns(self):
        n = lib().cbo_xns(self._h, None, 0)
        out = np.zeros((max(n, 1), 3))
        lib().cbo_xns(self._h, _dp(out), n)
        return out[:n]

    def spectrum(self):
        cap = int(np.sqrt(self.p.D) * self.p.nGrids / 2) + 2
        k, kp, d = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        nb = lib().cbo_spectrum(self._h, _dp(k), _dp(kp), _dp(d), cap)
        return k[:nb].copy(), kp[:nb].copy(), d[:nb].copy()


class WallTracker:
    """The reference's wall tracking on the oracle's fields, in plain Python:
    CanTrackWall / TrackWall / WallTrackStart / UpdateWall / WallFinder /
    GammaFactor (src/bubble.cpp:533-621), the centre (:310-313), the
    thresholds phiMax, phiZero and mass mT (:167-169) and outGamma's line
    format (:1009-1014)."""

    def __init__(self, D, nGrids, L, c, readxns, dt, DeltaR=0.0, every=400):
        self.D, self.N, self.L, self.every = D, nGrids, L, every
        self.on = D == 1 or readxns in (-1, -2)
        mid = nGrids // 2 if D == 3 else 0
        if readxns == -2 and D >= 2:
            self.centre = (nGrids // 2, nGrids // 2, mid)
        elif readxns == -1 and D >= 2:
            self.centre = (nGrids // 4, nGrids // 4, mid)
        else:
            self.centre = (0, 0, 0)
        self.phi_max = c / (6 + 2 * c)
        self.phi_zero = (2 + c - math.sqrt(4 + c)) / (3 + c)
        self.dx = L / nGrids
        _, self.Rn, self.lw = bubble_fit(D, c)
        self.max_radius = L / 4.0 - DeltaR if (readxns == -1 and D >= 2) else L / 2.0
        self.it_start = int(math.pi / math.sqrt(2 * (c + 6)) / dt)
        self.dini = self.width = self.rmin = self.rmax = 0.0

    def radius(self, i):
        return self.dx * math.sqrt((i + 0.5) * (i + 0.5) + 0.25 * (self.D - 1))

    def along(self, phi, i):
        if self.D == 1:
            return float(phi[i])
        i0 = ((self.centre[0] + i) % self.N + self.N) % self.N
        return float(phi[(i0,) + tuple(self.centre[1:self.D])])

    def find(self, phi):
        if self.rmax >= 0.98 * self.max_radius or self.rmin >= 0.98 * self.max_radius:
            return
        istart = max(int((self.rmin - self.lw) / self.dx), int(self.Rn / (2 * self.dx)))
        self.rmin = self.rmax = 0.0
        i = istart
        while (self.rmax == 0 or self.rmin == 0) and self.radius(i) < self.max_radius:
            v = self.along(phi, i)
            if v > self.phi_zero:
                self.rmin = self.radius(i)
            if v <= self.phi_max and self.rmax == 0:
                self.rmax = self.radius(i)
            i += 1
        if self.rmax < self.rmin:
            return
        if self.rmax * self.rmin == 0 and self.rmin + self.rmax != 0:
            self.rmax = self.rmin
        self.width = self.rmax - self.rmin

    def gamma(self, d):
        return 0.0 if d == 0 else self.dini / d

    def line(self, t):
        return "%.3f %.4f %.3f %.3f %.3f\n " % (t, self.gamma(self.width), self.width, self.rmin, self.rmax)

    def track(self, it, t, phi):
        """TrackWall after step it: the gamma.dat text it appends."""
        if not self.on:
            return ""
        if it == self.it_start:
            self.find(phi)
            self.dini = self.width
            return self.line(t)
        if it > self.it_start and it % self.every == 0 and t < self.L / 2:
            self.find(phi)
            return self.line(t)
        return ""

    def update(self, it, t, phi):
        """UpdateWall before snapshot output."""
        if self.on and it > self.it_start and t < self.L / 2:
            self.find(phi)


def seed_sites(xns, idx, **kw):
    """phi after seeding the centres xns (n x 3, in order) at the flat
    lattice indices idx, without building the lattice (src/bubble.cpp:420-453)."""
    p = default_params(**kw)
    xa = np.ascontiguousarray(np.asarray(xns, dtype=np.float64).reshape(-1, 3))
    ia = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
    out = np.zeros(ia.shape[0])
    lib().cbo_seed_sites(ctypes.byref(p), _dp(xa), xa.shape[0],
                         ia.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), ia.shape[0], _dp(out))
    return out


def mt64(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().cbo_mt64(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out


def normals(seed, n):
    out = np.zeros(n)
    lib().cbo_normals(seed, n, _dp(out))
    return out


def dft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    re, im = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    ore, oim = np.zeros(len(x)), np.zeros(len(x))
    lib().cbo_dft_ld(len(x), sign, _dp(re), _dp(im), _dp(ore), _dp(oim))
    return ore + 1j * oim


def r2c(field):
    field = np.ascontiguousarray(field, dtype=np.float64)
    D, N = field.ndim, field.shape[0]
    rows = field.size // N
    re, im = np.zeros(rows * (N // 2 + 1)), np.zeros(rows * (N // 2 + 1))
    lib().cbo_r2c(D, N, _dp(field), _dp(re), _dp(im))
    shape = field.shape[:-1] + (N // 2 + 1,)
    return (re + 1j * im).reshape(shape)


def thermal_modes(D, N, L, T, m, velocity, base=12345):
    rows = N * N if D == 3 else N
    re, im = np.zeros(rows * (N // 2 + 1)), np.zeros(rows * (N // 2 + 1))
    lib().cbo_thermal_modes(D, N, L, T, m, int(velocity), base, _dp(re), _dp(im))
    return (re + 1j * im).reshape(((N,) * (D - 1)) + (N // 2 + 1,))


def thermal_field(D, N, L, T, m, velocity, base=12345):
    out = np.zeros(N ** D)
    lib().cbo_thermal_field(D, N, L, T, m, int(velocity), base, _dp(out))
    return out.reshape((N,) * D)


def bubble_fit(D, c):
    out = np.zeros(3)
    lib().cbo_bubble_fit(D, c, _dp(out))
    return out


def generate_xns(D, L, c, nBubbles, seed=42):
    n = lib().cbo_generate_xns(D, L, c, nBubbles, seed, None, 0)
    out = np.zeros((max(n, 1), 3))
    lib().cbo_generate_xns(D, L, c, nBubbles, seed, _dp(out), n)
    return out[:n]


def ref_gridfuncs():
    """The reference's own src/gridfuncs.cpp, compiled into oracle/_ref (or None)."""
    path = os.path.join(_HERE, "_ref", "libref_gridfuncs.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    for f in (L.ref_iGridM, L.ref_iGridP):
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int] * 4
    L.ref_xfunc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                            ctypes.POINTER(ctypes.c_double)]
    L.ref_ijk.restype = ctypes.c_int
    L.ref_ijk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.ref_check_pbc.restype = ctypes.c_int
    L.ref_check_pbc.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                ctypes.c_double, ctypes.c_char_p, ctypes.c_int]
    return L
