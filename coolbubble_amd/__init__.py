"""coolbubble_amd -- MI355X lattice time-stepper (ctypes binding of libcoolbubble.so).

Python mirror of the reference's `Bubble` object (src/bubble.hpp:34-135) over
the C ABI declared in include/coolbubble.h.  Every call goes to the HIP
library; there is no CPU fallback: if libcoolbubble.so is missing or cannot
be loaded, importing the library raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcoolbubble.so")

CB_FP64 = 0
CB_FP32 = 1

STATUS = {0: "CB_OK", 1: "CB_INVALID_ARG", 2: "CB_CFL_VIOLATION", 3: "CB_OUT_OF_MEMORY",
          4: "CB_HIP_ERROR", 5: "CB_NAN_DETECTED", 6: "CB_NO_DEVICE", 7: "CB_ADMISSION_REJECTED",
          8: "CB_NOT_READY", 9: "CB_UNSUPPORTED"}

# Column names of an EnergyDensities.dat row (src/bubble.cpp:962-964).
ENERGY_COLUMNS = ("t", "a", "Etot", "Ekin", "Egrad", "Epot", "EtotSpread", "EkinSpread",
                  "EgradSpread", "EpotSpread", "avphi", "avphiSpread", "avphiP",
                  "avphiPSpread", "H", "C")


class CBError(RuntimeError):
    def __init__(self, status, msg):
        self.status = status
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))


class Params(ctypes.Structure):
    """cb_params (include/coolbubble.h)."""
    _fields_ = [
        ("D", ctypes.c_int), ("nGrids", ctypes.c_int), ("L", ctypes.c_double),
        ("c", ctypes.c_double), ("readxns", ctypes.c_int), ("dt", ctypes.c_double),
        ("nBubbles", ctypes.c_int), ("Trad", ctypes.c_double), ("DeltaR", ctypes.c_double),
        ("test_wave", ctypes.c_int), ("expansion", ctypes.c_int), ("alpha", ctypes.c_int),
        ("Mpl", ctypes.c_double), ("rand_seed", ctypes.c_uint),
        ("thermal_base", ctypes.c_ulonglong), ("precision", ctypes.c_int), ("device", ctypes.c_int),
        ("thermal_spectrum", ctypes.c_int),
    ]


class ServerOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("precision", ctypes.c_int),
                ("mem_budget", ctypes.c_longlong), ("max_active", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("params", Params), ("nt", ctypes.c_int), ("snapshots", ctypes.c_int),
                ("tag", ctypes.c_ulonglong)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_ulonglong), ("tag", ctypes.c_ulonglong), ("status", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("site_updates", ctypes.c_longlong), ("n_rows", ctypes.c_int), ("n_spectra", ctypes.c_int),
                ("final_row", ctypes.c_double * 16)]


class Energies(ctypes.Structure):
    _fields_ = [("col", ctypes.c_double * 16)]


class State(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("ap", ctypes.c_double), ("H", ctypes.c_double),
                ("C", ctypes.c_double), ("t", ctypes.c_double), ("EkinHalfPrev", ctypes.c_double),
                ("step", ctypes.c_longlong)]


_lib = None


def lib():
    """The loaded libcoolbubble.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcoolbubble.so not built: run python -m coolbubble_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.cb_default_params.argtypes = [ctypes.POINTER(Params)]
        L.cb_check_stability.argtypes = [ctypes.POINTER(Params)]
        L.cb_create.argtypes = [ctypes.POINTER(Params), dp, ctypes.c_int, ctypes.POINTER(vp)]
        L.cb_step.argtypes = [vp, ctypes.c_int]
        L.cb_energies_row.argtypes = [vp, ctypes.POINTER(Energies)]
        L.cb_power_spectrum.argtypes = [vp, dp, dp, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_power_spectrum_noalloc.argtypes = [vp, dp, dp, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_get_fields.argtypes = [vp, dp, dp]
        L.cb_get_state.argtypes = [vp, ctypes.POINTER(State)]
        L.cb_set_state.argtypes = [vp, ctypes.POINTER(State), dp, dp]
        L.cb_get_xns.argtypes = [vp, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_stream.argtypes = [vp, ctypes.POINTER(vp)]
        L.cb_kernel_variants.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.cb_select_kernel.argtypes = [vp, ctypes.c_int]
        L.cb_kernel_variant_info.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_kernel_variant_info.restype = ctypes.c_int
        L.cb_autotune.argtypes = [vp, ctypes.c_int, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_bandwidth_probe.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, dp]
        L.cb_bandwidth_probe.restype = ctypes.c_int
        L.cb_slab_probe.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.cb_slab_probe.restype = ctypes.c_int
        L.cb_batch_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(vp)]
        L.cb_batch_step.argtypes = [vp, ctypes.c_int]
        L.cb_batch_member.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.cb_batch_size.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)]
        L.cb_batch_stream.argtypes = [vp, ctypes.POINTER(vp)]
        L.cb_batch_destroy.argtypes = [vp]
        for name in ("cb_batch_create", "cb_batch_step", "cb_batch_member", "cb_batch_size",
                     "cb_batch_stream", "cb_batch_destroy"):
            getattr(L, name).restype = ctypes.c_int
        L.cb_batch_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp]
        L.cb_batch_autotune.restype = ctypes.c_int
        L.cb_batch_set_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.cb_batch_set_shape.restype = ctypes.c_int
        ull = ctypes.c_ulonglong
        L.cb_server_create.argtypes = [ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)]
        L.cb_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ull)]
        L.cb_server_cancel.argtypes = [vp, ull]
        L.cb_server_poll.argtypes = [vp, ctypes.POINTER(Result), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_server_rows.argtypes = [vp, ull, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_server_spectrum.argtypes = [vp, ull, ctypes.c_int, dp, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.cb_server_release.argtypes = [vp, ull]
        L.cb_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)]
        L.cb_server_drain.argtypes = [vp]
        L.cb_server_destroy.argtypes = [vp]
        L.cb_server_variant.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        for name in ("cb_server_create", "cb_server_submit", "cb_server_cancel", "cb_server_poll",
                     "cb_server_rows", "cb_server_spectrum", "cb_server_release", "cb_server_stats",
                     "cb_server_drain", "cb_server_destroy", "cb_server_variant"):
            getattr(L, name).restype = ctypes.c_int
        L.cb_default_run_opts.argtypes = [vp]
        L.cb_run_folder_name.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.cb_run.argtypes = [vp]
        L.cb_hdf5_available.argtypes = []
        L.cb_write_snapshot_h5.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.POINTER(ctypes.POINTER(ctypes.c_float)), ctypes.c_int]
        for name in ("cb_default_run_opts", "cb_run_folder_name", "cb_run", "cb_hdf5_available",
                     "cb_write_snapshot_h5"):
            getattr(L, name).restype = ctypes.c_int
        L.cb_nsites.argtypes = [vp]
        L.cb_nsites.restype = ctypes.c_longlong
        L.cb_destroy.argtypes = [vp]
        L.cb_status_string.argtypes = [ctypes.c_int]
        L.cb_status_string.restype = ctypes.c_char_p
        L.cb_last_error.restype = ctypes.c_char_p
        for name in ("cb_default_params", "cb_check_stability", "cb_create", "cb_step",
                     "cb_energies_row", "cb_power_spectrum", "cb_power_spectrum_noalloc", "cb_get_fields",
                     "cb_get_state",
                     "cb_set_state", "cb_get_xns", "cb_stream", "cb_destroy",
                     "cb_kernel_variants", "cb_select_kernel", "cb_autotune"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise CBError(status, lib().cb_last_error().decode(errors="replace"))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_params(**kw):
    """Reference defaults (src/def.h), overridable by keyword ('precision': 'fp32'/'fp64')."""
    p = Params()
    check(lib().cb_default_params(ctypes.byref(p)))
    for key, val in kw.items():
        if key == "precision" and isinstance(val, str):
            val = CB_FP32 if val.lower() == "fp32" else CB_FP64
        setattr(p, key, val)
    return p


class RunOpts(ctypes.Structure):
    """cb_run_opts (include/coolbubble.h)."""
    _fields_ = [("params", Params), ("nt", ctypes.c_int), ("snapshots", ctypes.c_int),
                ("snapshot_spacing", ctypes.c_int), ("energy_every", ctypes.c_int), ("fft", ctypes.c_int),
                ("full1d_every", ctypes.c_int), ("field_snapshots", ctypes.c_int), ("memorysave", ctypes.c_int),
                ("deflate", ctypes.c_int), ("wall_tracking", ctypes.c_int), ("wall_track_every", ctypes.c_int),
                ("out_dir", ctypes.c_char_p),
                ("xns", ctypes.POINTER(ctypes.c_double)), ("n_xns", ctypes.c_int),
                ("folder", ctypes.c_char * 1024), ("elapsed_s", ctypes.c_double)]


def _run_opts(out_dir, params, **opts):
    o = RunOpts()
    check(lib().cb_default_run_opts(ctypes.byref(o)))
    o.params = params
    for key, val in opts.items():
        if val is not None:
            setattr(o, key, int(val))
    o.out_dir = os.fsencode(out_dir)
    return o


def run_folder_name(nt=None, snapshots=None, fft=None, full1d_every=None, **kw):
    """The reference's run-folder name for these options (src/main.cpp:243-247)."""
    o = _run_opts(".", default_params(**kw), nt=nt, snapshots=snapshots, fft=fft, full1d_every=full1d_every)
    buf = ctypes.create_string_buffer(1024)
    check(lib().cb_run_folder_name(ctypes.byref(o), buf, 1024))
    return buf.value.decode()


def run(out_dir, nt=None, snapshots=None, snapshot_spacing=None, energy_every=None, fft=None,
        full1d_every=None, field_snapshots=None, memorysave=None, deflate=None, wall_tracking=None,
        wall_track_every=None, xns=None, **kw):
    """The reference's whole run (src/main.cpp:229-310): creates the run folder
    under out_dir, writes parameters.dat, EnergyDensities.dat, __xns.dat,
    _phi_k_t=NNNN.dat, _phi_t=NNNN.h5 (field_snapshots), gamma.dat
    (wall_tracking) and (1D) FullField.dat.  Returns (folder, elapsed s)."""
    o = _run_opts(out_dir, default_params(**kw), nt=nt, snapshots=snapshots, snapshot_spacing=snapshot_spacing,
                  energy_every=energy_every, fft=fft, full1d_every=full1d_every,
                  field_snapshots=field_snapshots, memorysave=memorysave, deflate=deflate,
                  wall_tracking=wall_tracking, wall_track_every=wall_track_every)
    keep = None
    if xns is not None:
        keep = np.ascontiguousarray(np.asarray(xns, dtype=np.float64).reshape(-1, 3))
        o.xns = _dp(keep)
        o.n_xns = keep.shape[0]
    check(lib().cb_run(ctypes.byref(o)))
    return o.folder.decode(), o.elapsed_s


def hdf5_available():
    """True when the HDF5 C library loads (the field snapshots need it)."""
    return lib().cb_hdf5_available() == 0


def snapshot_points(D, nGrids):
    """(reduc, n): the thinning step and saved points per axis of a snapshot
    (src/bubble.cpp:891-894)."""
    reduc = 1
    if D == 2 and nGrids > 2048:
        reduc = nGrids // 2048
    if D == 3 and nGrids > 512:
        reduc = nGrids // 512
    return reduc, (nGrids + reduc - 1) // reduc


def write_snapshot_h5(path, fields, D, nGrids, L, c, t, a, deflate=9):
    """Bubble::OutField's file (src/bubble.cpp:887-954) from host arrays:
    fields = [phi, pi] or [phi, pi, Ekin, gradphi, potEnergy], each n^D values
    of the saved points in lattice order."""
    if len(fields) not in (2, 5):
        raise ValueError("fields must hold 2 or 5 arrays")
    _, n = snapshot_points(D, nGrids)
    keep = [np.ascontiguousarray(np.asarray(f, dtype=np.float32).reshape(-1)) for f in fields]
    for f in keep:
        if f.size != n ** D:
            raise ValueError("each field must hold %d values" % n ** D)
    fp = ctypes.POINTER(ctypes.c_float)
    arr = (fp * len(keep))(*[f.ctypes.data_as(fp) for f in keep])
    check(lib().cb_write_snapshot_h5(os.fsencode(path), D, nGrids, L, c, t, a, len(keep), arr, deflate))


def bandwidth_probe(nbytes=1 << 30, reps=20, device=0):
    """GB/s (read + write) of a plain device-to-device copy kernel."""
    g = ctypes.c_double()
    check(lib().cb_bandwidth_probe(device, int(nbytes), int(reps), ctypes.byref(g)))
    return g.value


class Bubble:
    """One simulation on the GPU: Bubble::Bubble + oneStep + outEP + spectrum."""

    def __init__(self, params=None, xns=None, _handle=None, **kw):
        self.p = params if params is not None else default_params(**kw)
        self._owned = _handle is None
        if _handle is not None:
            self._h = _handle
            self.nsites = lib().cb_nsites(_handle)
            return
        h = ctypes.c_void_p()
        if xns is not None:
            xa = np.ascontiguousarray(np.asarray(xns, dtype=np.float64).reshape(-1, 3))
            check(lib().cb_create(ctypes.byref(self.p), _dp(xa), xa.shape[0], ctypes.byref(h)))
        else:
            check(lib().cb_create(ctypes.byref(self.p), None, 0, ctypes.byref(h)))
        self._h = h
        self.nsites = lib().cb_nsites(h)

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                lib().cb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def shape(self):
        return (self.p.nGrids,) * self.p.D

    def step(self, n=1):
        check(lib().cb_step(self._h, int(n)))

    def energy_row(self):
        e = Energies()
        check(lib().cb_energies_row(self._h, ctypes.byref(e)))
        return np.array(e.col[:])

    def fields(self):
        phi = np.empty(self.nsites)
        pi = np.empty(self.nsites)
        check(lib().cb_get_fields(self._h, _dp(phi), _dp(pi)))
        return phi.reshape(self.shape()), pi.reshape(self.shape())

    def state(self):
        st = State()
        check(lib().cb_get_state(self._h, ctypes.byref(st)))
        return st

    def set_state(self, st, phi=None, pi=None):
        ph = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64).ravel()
        pp = None if pi is None else np.ascontiguousarray(pi, dtype=np.float64).ravel()
        check(lib().cb_set_state(self._h, ctypes.byref(st), None if ph is None else _dp(ph),
                                 None if pp is None else _dp(pp)))

    def xns(self):
        n = ctypes.c_int()
        check(lib().cb_get_xns(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 3))
        check(lib().cb_get_xns(self._h, _dp(out), n.value, ctypes.byref(n)))
        return out[:n.value]

    def spectrum(self):
        cap = int(np.sqrt(self.p.D) * self.p.nGrids / 2) + 2
        k, kp, d = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        nb = ctypes.c_int()
        check(lib().cb_power_spectrum(self._h, _dp(k), _dp(kp), _dp(d), cap, ctypes.byref(nb)))
        n = nb.value
        return k[:n].copy(), kp[:n].copy(), d[:n].copy()

    def slab_probe(self, planes, launches=20):
        """(ms per launch, ms per exchange) of this simulation's multi-step
        variant on its first `planes` planes, as one rank's slab of a run
        sharded along the slowest axis; the fields are scratch afterwards."""
        ml, mx = ctypes.c_double(), ctypes.c_double()
        check(lib().cb_slab_probe(self._h, int(planes), int(launches), ctypes.byref(ml), ctypes.byref(mx)))
        return ml.value, mx.value

    def kernel_variants(self):
        """(number of tiled variants, selected variant; -1 = general kernel)."""
        n, sel = ctypes.c_int(), ctypes.c_int()
        check(lib().cb_kernel_variants(self._h, ctypes.byref(n), ctypes.byref(sel)))
        return n.value, sel.value

    def variant_info(self, variant):
        """dict(vw, items, nthr, TR, NI, pf, nt) of a tiled variant."""
        info = (ctypes.c_int * 8)()
        check(lib().cb_kernel_variant_info(self._h, int(variant), info))
        return dict(zip(("vw", "items", "nthr", "TR", "NI", "pf", "nt", "fuse"), info[:]))

    def select_kernel(self, variant):
        check(lib().cb_select_kernel(self._h, int(variant)))

    def autotune(self, steps=5):
        """Time every variant (state preserved), select the fastest; returns
        (best variant, ms per step of each variant with the general kernel last)."""
        n, _ = self.kernel_variants()
        ms = np.zeros(n + 1)
        best = ctypes.c_int()
        check(lib().cb_autotune(self._h, int(steps), _dp(ms), n + 1, ctypes.byref(best)))
        return best.value, ms

    def stream(self):
        s = ctypes.c_void_p()
        check(lib().cb_stream(self._h, ctypes.byref(s)))
        return s


class Batch:
    """An ensemble stepped together (cb_batch_*): one persistent launch per
    step over all members, 40-step segments replayed as a HIP graph."""

    def __init__(self, members):
        plist = [m if isinstance(m, Params) else default_params(**m) for m in members]
        arr = (Params * len(plist))(*plist)
        h = ctypes.c_void_p()
        check(lib().cb_batch_create(arr, len(plist), ctypes.byref(h)))
        self._h = h
        self.params = plist
        self.members = []
        for i, pp in enumerate(plist):
            mh = ctypes.c_void_p()
            check(lib().cb_batch_member(h, i, ctypes.byref(mh)))
            self.members.append(Bubble(params=pp, _handle=mh))
        n, sites = ctypes.c_int(), ctypes.c_longlong()
        check(lib().cb_batch_size(h, ctypes.byref(n), ctypes.byref(sites)))
        self.sites = sites.value

    def step(self, n):
        check(lib().cb_batch_step(self._h, int(n)))

    def autotune(self, steps=10):
        """Pick the batched launch shape; returns ((items, pf, nt, NI, two_step), ms per step)."""
        shape = (ctypes.c_int * 5)()
        ms = ctypes.c_double()
        check(lib().cb_batch_autotune(self._h, int(steps), shape, ctypes.byref(ms)))
        return tuple(shape[:]), ms.value

    def set_shape(self, shape):
        """Set the batched launch shape (items, pf, nt, NI, two_step)."""
        arr = (ctypes.c_int * 5)(*[int(x) for x in shape])
        check(lib().cb_batch_set_shape(self._h, arr))

    def stream(self):
        s = ctypes.c_void_p()
        check(lib().cb_batch_stream(self._h, ctypes.byref(s)))
        return s

    def close(self):
        if getattr(self, "_h", None):
            for m in self.members:
                m._h = None
            lib().cb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server (cb_server_*)."""

    def __init__(self, precision="fp32", mem_budget=0, max_active=0, device=0):
        o = ServerOpts(device=device, precision=CB_FP32 if precision == "fp32" else CB_FP64,
                       mem_budget=int(mem_budget), max_active=int(max_active))
        h = ctypes.c_void_p()
        check(lib().cb_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h

    def submit(self, nt, snapshots=5, tag=0, params=None, **kw):
        r = Request(params=params if params is not None else default_params(**kw), nt=int(nt),
                    snapshots=int(snapshots), tag=int(tag))
        rid = ctypes.c_ulonglong()
        check(lib().cb_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        return rid.value

    def cancel(self, rid):
        check(lib().cb_server_cancel(self._h, int(rid)))

    def poll(self, cap=256):
        out = (Result * cap)()
        n = ctypes.c_int()
        check(lib().cb_server_poll(self._h, out, cap, ctypes.byref(n)))
        return [out[i] for i in range(n.value)]

    def rows(self, rid):
        n = ctypes.c_int()
        check(lib().cb_server_rows(self._h, int(rid), None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 16))
        check(lib().cb_server_rows(self._h, int(rid), _dp(out), n.value, ctypes.byref(n)))
        return out[:n.value]

    def spectrum(self, rid, which):
        n = ctypes.c_int()
        check(lib().cb_server_spectrum(self._h, int(rid), int(which), None, None, 0, ctypes.byref(n)))
        k, d = np.zeros(max(n.value, 1)), np.zeros(max(n.value, 1))
        check(lib().cb_server_spectrum(self._h, int(rid), int(which), _dp(k), _dp(d), n.value,
                                       ctypes.byref(n)))
        return k[:n.value], d[:n.value]

    def release(self, rid):
        check(lib().cb_server_release(self._h, int(rid)))

    def stats(self):
        q, a, d, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        check(lib().cb_server_stats(self._h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(d),
                                    ctypes.byref(b)))
        return dict(queued=q.value, active=a.value, done=d.value, bytes=b.value)

    def drain(self):
        check(lib().cb_server_drain(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().cb_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def server_variant(sim):
    """The kernel variant the server gives a request with sim's parameters
    (cb_server_variant), or -1: selecting it on a lone simulation reproduces
    the request's results bit for bit."""
    v = ctypes.c_int()
    check(lib().cb_server_variant(sim.handle, ctypes.byref(v)))
    return v.value
