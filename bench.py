#!/usr/bin/env python3
"""Headline benchmark: site-updates/s of the lattice time-stepper on one MI355X.

Workload: d=3, N=512^3, L=64, c=4, mode 1 with 32 seeded bubbles (radius 1.3x
the fitted critical radius), dt = 0.2 dx, static background -- the north
star's FP32 vs FP64 configuration.  A "step" is one leapfrog step
(Bubble::oneStep, src/bubble.cpp:849-876) of the whole lattice; the steps are
issued as the reference's driver issues them, in runs of 40 with the
box-averaged energies measured at the end of each run (OutEnergyEvery = 40,
src/def.h:65, src/main.cpp:280).

Timing: HIP events recorded on the library's own stream around exactly K
steps, inputs resident in HBM.  value = FP32 site-updates/s.  The FP64 GPU
path runs the same workload in the same process; the CPU baseline times the
oracle (the C/OpenMP FP64 restatement of the reference) on the host cores
over a bounded sample of the same lattice.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HBM_PEAK = 8.0e12  # MI355X datasheet, bytes/s
CPU_SHARE = 16     # host cores that go with one GPU

CONFIG = dict(D=3, nGrids=512, L=64.0, c=4.0, readxns=1, nBubbles=32, expansion=0)
CONFIG["dt"] = 0.2 * CONFIG["L"] / CONFIG["nGrids"]
ENERGY_EVERY = 40
# the expanding-background workload: d=3 N=384 L=48 c=2, thermal start (mode -3,
# Trad=0.1, seeded), Mpl=10, FP32 field with FP64 scale factor and sums
CONFIG_EXP = dict(D=3, nGrids=384, L=48.0, c=2.0, readxns=-3, Trad=0.1, expansion=1, Mpl=10.0, dt=0.025)


class HipEvents:
    """hipEvent timing on an explicit stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []

    def event(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0


def run_precision(cb, hip, precision, steps, warmup, log, variant=None):
    t0 = time.time()
    sim = cb.Bubble(precision=precision, **CONFIG)
    log("%s: seeded %d^3 in %.1f s" % (precision, CONFIG["nGrids"], time.time() - t0))
    if variant is None:
        # trials in calls of the length then timed (the output cadence, or
        # fewer steps when fewer are timed): a variant's tail launches count
        best, ms = sim.autotune(steps=min(ENERGY_EVERY, steps))
        nvar, sel = sim.kernel_variants()
        log("%s: autotune picked variant %d of %d (%.3f ms/step; general kernel %.3f ms/step)"
            % (precision, sel, nvar, ms[sel] if sel >= 0 else ms[-1], ms[-1]))
        log("%s: selected %s" % (precision, sim.variant_info(sel) if sel >= 0 else "general kernel"))
        order = np.argsort(ms)[:5]
        log("%s: fastest variants: %s" % (precision, "; ".join(
            "%d %.4f ms %s" % (v, ms[v], sim.variant_info(v) if v < nvar else "general") for v in order)))
    else:
        sim.select_kernel(variant)
    # warm-up calls no longer than the timed ones, so that every kernel the
    # timed calls launch (their tails included) has been launched once
    # before the timed region (a kernel's first launch loads its code)
    done = 0
    while done < warmup:
        n = min(ENERGY_EVERY, steps, warmup - done)
        sim.step(n)
        done += n
    stream = sim.stream()
    a, b = hip.event(), hip.event()
    hip.sync()
    hip.record(a, stream)
    done = 0
    while done < steps:
        n = min(ENERGY_EVERY, steps - done)
        sim.step(n)
        done += n
    hip.record(b, stream)
    ms_total = hip.elapsed_ms(a, b)
    row = sim.energy_row()
    log("%s: %d steps in %.3f ms (%.4f ms/step)" % (precision, steps, ms_total, ms_total / steps))
    return sim, ms_total, row


def run_expanding(cb, hip, steps, log):
    """FP32 steps of the expanding-background thermal workload: every step
    needs the previous step's global energy sum (the Friedmann update), so
    each step is one pass of the lagged-drift kernel with its energies."""
    t0 = time.time()
    sim = cb.Bubble(precision="fp32", **CONFIG_EXP)
    sites = CONFIG_EXP["nGrids"] ** 3
    log("expanding: thermal start %d^3 in %.1f s" % (CONFIG_EXP["nGrids"], time.time() - t0))
    sim.autotune(steps=ENERGY_EVERY)
    sel = sim.kernel_variants()[1]
    sim.step(ENERGY_EVERY)
    stream = sim.stream()
    a, b = hip.event(), hip.event()
    hip.sync()
    hip.record(a, stream)
    done = 0
    while done < steps:
        n = min(ENERGY_EVERY, steps - done)
        sim.step(n)
        done += n
    hip.record(b, stream)
    ms = hip.elapsed_ms(a, b)
    su = sites * steps / (ms * 1e-3)
    log("expanding: %d steps in %.3f ms (%.4f ms/step), a = %.6f" % (steps, ms, ms / steps, sim.state().a))
    out = {"config": "d=3 N=384^3 L=48 c=2 mode -3 thermal Trad=0.1 expansion Mpl=10 dt=0.025, FP32 field, "
                     "FP64 scale factor and sums, energies every step (lagged drift)",
           "steps": steps, "ms_per_step": ms / steps, "site_updates_per_s": su,
           "achieved_GBps": 16.0 * su / 1e9, "frac": 16.0 * su / HBM_PEAK,
           "variant": sim.variant_info(sel) if sel >= 0 else "general kernel", "final_a": sim.state().a}
    sim.close()
    return out


def launch_time_live(sim, hip, fuse, m=13, reps=3):
    """Average duration of one plain launch of the selected variant (fuse
    steps), live with HIP events on the simulation's stream: calls of
    fuse * m and fuse * 2 steps end on the same tail launches (the one that
    measures the energies, and for a three-step tile with a two-step form
    the launches of that form), so they differ by m - 2 plain launches:
    (T(fuse * m) - T(fuse * 2)) / (m - 2) is one plain launch."""
    stream = sim.stream()

    def timed(n):
        best = None
        for _ in range(reps):
            a, b = hip.event(), hip.event()
            hip.sync()
            hip.record(a, stream)
            sim.step(n)
            hip.record(b, stream)
            ms = hip.elapsed_ms(a, b)
            best = ms if best is None else min(best, ms)
        return best

    return (timed(fuse * m) - timed(fuse * 2)) / (m - 2) * 1e3


def slab_share(sim, info, log, planes=64):
    """One GPU's slab of the same lattice sharded 8 ways along the slowest
    axis (512 x 512 x 64 planes): the selected multi-step tile and the
    1024-thread three-step tile, each with the chunk lengths that divide the
    slab, restricted to the slab's planes (a rank re-tunes for its slab: the
    direct-load tile's 256 row tiles of a whole box are only 64 blocks of a
    64-plane slab), and the device copies standing in for the halo
    exchange."""
    if not info or info.get("fuse", 1) < 2:
        return None
    n, _ = sim.kernel_variants()
    infos = [sim.variant_info(v) for v in range(n)]
    out = {"planes": planes, "shape": "%d x %d x %d" % (CONFIG["nGrids"], CONFIG["nGrids"], planes), "tiles": []}
    bases = [info]
    tile1024 = {"vw": 4, "items": 1, "nthr": 1024, "TR": 4, "NI": 64, "pf": 2, "nt": 1, "fuse": 3}
    if any(i == tile1024 for i in infos) and dict(info, NI=64) != tile1024:
        bases.append(tile1024)
    for base in bases:
        for ni in (64, 32, 16):
            want = dict(base, NI=ni)
            v = next((v for v in range(n) if infos[v] == want), None)
            if v is None or planes % ni:
                continue
            sim.select_kernel(v)
            ml, mx = sim.slab_probe(planes, 20)
            su = CONFIG["nGrids"] ** 2 * planes * base["fuse"] / (ml * 1e-3)
            log("slab %s, %d threads, NI %d: %.1f us per launch (%d steps), %.3e site-updates/s; exchange copies %.1f us"
                % (out["shape"], base["nthr"], ni, ml * 1e3, base["fuse"], su, mx * 1e3))
            out["tiles"].append({"variant": want, "NI": ni, "us_per_launch": ml * 1e3, "site_updates_per_s": su,
                                 "exchange_copies_us": mx * 1e3})
    if out["tiles"]:
        out["best_site_updates_per_s"] = max(t["site_updates_per_s"] for t in out["tiles"])
    out["exchange_bytes_per_launch"] = 2 * 2 * info["fuse"] * CONFIG["nGrids"] ** 2 * 4
    return out


def spectrum_time(sim, hip, reps=3):
    stream = sim.stream()
    sim.spectrum()   # plan + warm
    a, b = hip.event(), hip.event()
    hip.record(a, stream)
    for _ in range(reps):
        sim.spectrum()
    hip.record(b, stream)
    return hip.elapsed_ms(a, b) / reps


def ensemble_members():
    """64 thermal d=3 realisations (N=128, L=16, c in {1,2,4,8} x 16 seeds,
    Trad = 0.1) and 16 d=2 five-bubble runs (N=1024, L=64, c ~ U[1,8] from a
    seeded generator), static, FP32, dt = 0.2 dx."""
    rng = np.random.default_rng(2024)
    out = []
    for c in (1.0, 2.0, 4.0, 8.0):
        for s in range(16):
            out.append(dict(D=3, nGrids=128, L=16.0, dt=0.025, c=c, readxns=-3, Trad=0.1,
                            thermal_base=12345 + s, precision="fp32"))
    for c in rng.uniform(1.0, 8.0, 16):
        out.append(dict(D=2, nGrids=1024, L=64.0, dt=0.0125, c=float(c), readxns=1, nBubbles=5,
                        precision="fp32"))
    return out


def run_ensemble(cb, hip, nt, log):
    t0 = time.time()
    batch = cb.Batch(ensemble_members())
    log("ensemble: %d members, %d sites, created in %.1f s" % (len(batch.members), batch.sites,
                                                                time.time() - t0))
    shape, tms = batch.autotune(steps=10)
    log("ensemble: autotuned launch shape (items, prefetch, nt, NI, two-step) = %s, %.3f ms/step" % (shape, tms))
    batch.step(40)   # graph capture + warm
    stream = batch.stream()
    a, b = hip.event(), hip.event()
    hip.sync()
    hip.record(a, stream)
    batch.step(nt)
    hip.record(b, stream)
    ms = hip.elapsed_ms(a, b)
    rows = [m.energy_row() for m in batch.members]
    assert all(np.all(np.isfinite(r)) for r in rows)
    batch.close()
    log("ensemble: %d steps of all members in %.1f ms" % (nt, ms))
    return {"members": 80, "steps": nt, "wall_ms": ms, "runs_per_hour": 80 * 3600.0 / (ms * 1e-3),
            "site_updates_per_s": batch.sites * nt / (ms * 1e-3),
            "config": "64 x d=3 N=128 L=16 thermal (c in {1,2,4,8} x 16 seeds, Trad=0.1) + "
                      "16 x d=2 N=1024 L=64 five bubbles (c ~ U[1,8]); FP32, static, "
                      "energies every 40 steps, 40-step HIP graph segments"}


def serving_workload(n, rate, seed=7):
    """Poisson arrivals at `rate` per second; d in {1,2,3}; N = 2^16 (d=1),
    N in [256, 2048] (d=2, multiples of 32), N in [64, 384] (d=3, multiples
    of 16); c ~ U[1,8]; mode in {-2,-1,1,-3}; nt ~ U[500,5000]; 5 snapshots;
    dx = 0.0625 (d=1,2) or 0.125 (d=3), L = N dx, dt = 0.2 dx."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        d = int(rng.integers(1, 4))
        if d == 1:
            N, dx = 65536, 0.0625
        elif d == 2:
            N, dx = 32 * int(rng.integers(8, 65)), 0.0625
        else:
            N, dx = 16 * int(rng.integers(4, 25)), 0.125
        p = dict(D=d, nGrids=N, L=N * dx, dt=0.2 * dx, c=float(rng.uniform(1.0, 8.0)),
                 readxns=int(rng.choice([-2, -1, 1, -3])))
        out.append((t, p, int(rng.integers(500, 5001))))
    return out


def run_serving(cb, n, rate, log):
    work = serving_workload(n, rate)
    sv = cb.Server(precision="fp32")
    t0 = time.time()
    ids = {}
    for i, (ta, p, nt) in enumerate(work):
        delay = ta - (time.time() - t0)
        if delay > 0:
            time.sleep(delay)
        ids[sv.submit(nt, 5, tag=i, **p)] = i
    sv.drain()
    res = sv.poll(cap=n + 8)
    sv.close()
    ok = [r for r in res if r.status == 0]
    lat = np.array([r.t_done - r.t_submit for r in ok])
    span = max(r.t_done for r in ok) - min(r.t_submit for r in ok)
    su = sum(r.site_updates for r in ok)
    # the work the arrivals bring per second of the arrival window: at or
    # below the served rate the server keeps up and the latencies are
    # service times, not a growing backlog
    offered = sum(p["nGrids"] ** p["D"] * nt for _, p, nt in work) / work[-1][0]
    log("serving: %d requests (%d ok) at %.1f/s: p50 %.3f s, p99 %.3f s, %.3e site-updates/s over %.1f s "
        "(offered %.3e site-updates/s)" % (n, len(ok), rate, np.percentile(lat, 50), np.percentile(lat, 99),
                                          su / span, span, offered))
    return {"requests": n, "completed": len(ok), "arrival_rate_per_s": rate,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "site_updates_per_s": su / span, "makespan_s": span, "arrival_window_s": work[-1][0],
            "offered_site_updates_per_s": offered, "precision": "fp32",
            "config": "Poisson arrivals; d in {1,2,3} (N=2^16 | [256,2048] | [64,384]); c ~ U[1,8]; "
                      "mode in {-2,-1,1,-3}; nt ~ U[500,5000]; 5 snapshots with spectra"}


def run_driver(cb, log):
    """Whole runs of the 3D N=512 32-bubble workload through cb_run: the
    reference's time loop with energy rows every 40 steps, 20 spectra and the
    text files (no HDF5), then the HDF5 field snapshots alone (two files of
    5 x 512^3 floats at deflate 9 and 1; one more run with five files at
    deflate 1 shows the writer keeping up with the stepping)."""
    import shutil
    import tempfile
    out = tempfile.mkdtemp(prefix="cb_run_")
    kw = dict(CONFIG)
    legs = [("text_fp32", dict(field_snapshots=0, snapshots=20, precision="fp32")),
            ("text_fp64", dict(field_snapshots=0, snapshots=20, precision="fp64")),
            ("h5_deflate9_fp32", dict(field_snapshots=1, snapshots=1, deflate=9, precision="fp32")),
            ("h5_deflate1_fp32", dict(field_snapshots=1, snapshots=1, deflate=1, precision="fp32")),
            ("h5x5_deflate1_fp32", dict(field_snapshots=1, snapshots=4, deflate=1, precision="fp32"))]
    res = {}
    try:
        for name, o in legs:
            d = os.path.join(out, name)
            folder, el = cb.run(d, nt=2000, **o, **kw)
            h5 = [os.path.getsize(os.path.join(folder, f)) for f in os.listdir(folder) if f.endswith(".h5")]
            res[name] = {"elapsed_s": el, "h5_files": len(h5), "h5_bytes": sum(h5)}
            log("run %s: %.2f s, %d .h5 files, %.2f GB" % (name, el, len(h5), sum(h5) / 1e9))
            shutil.rmtree(d, ignore_errors=True)
    finally:
        shutil.rmtree(out, ignore_errors=True)
    res["config"] = "3D N=512 L=64 c=4 mode 1 with 32 bubbles, dt=0.2 dx, nt=2000"
    return res


def cpu_model():
    """The host CPU as /proc/cpuinfo names it (model name, family, model,
    stepping) and the logical CPUs the machine shows."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" in line:
                    k, v = (x.strip() for x in line.split(":", 1))
                    if k in ("model name", "cpu family", "model", "stepping", "vendor_id") and k not in info:
                        info[k] = v
    except OSError:
        pass
    return "%s (vendor %s, family %s, model %s, stepping %s; %d logical CPUs on the host)" % (
        info.get("model name", "unknown"), info.get("vendor_id", "?"), info.get("cpu family", "?"),
        info.get("model", "?"), info.get("stepping", "?"), os.cpu_count() or 0)


def cpu_baseline(sim64, nsteps, log):
    """Oracle (FP64 C/OpenMP restatement, 4 passes per step with energies
    every step, like the reference) stepping the same lattice from the GPU's
    seeded state, on the host cores."""
    import oracle
    # the host share of one GPU on the MI355X boxes is 16 cores (the machine
    # shows more, shared with the other GPUs' jobs)
    threads = min(CPU_SHARE, os.cpu_count() or 1)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    phi, pi = sim64.fields()
    st = sim64.state()
    p = oracle.default_params(**CONFIG)
    o = oracle.Sim(params=p, blank=True, nthreads=threads)
    o.phi()[...] = phi
    o.pi()[...] = pi
    o.set_scalars([st.a, st.ap, st.H, st.C, st.t, st.EkinHalfPrev])
    del phi, pi
    o.step(1)   # first touch
    t0 = time.time()
    o.step(nsteps)
    dt = time.time() - t0
    o.close()
    log("cpu baseline: %d steps in %.2f s on %d threads" % (nsteps, dt, threads))
    return dt, threads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["both", "fp32", "fp64"], default="both",
                    help="profiling aid: run one precision only")
    ap.add_argument("--variant", type=int, default=None,
                    help="profiling aid: fixed kernel variant instead of autotuning")
    ap.add_argument("--no-spectrum", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-expanding", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--only", choices=["ensemble", "serving", "run", "expanding"], default=None,
                    help="run one leg alone and print its record (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the lattice time-stepper runs on a single MI355X "
              "(multi-GPU slab decomposition is a design note, not built); use --gpus 1",
              file=sys.stderr)
        return 2

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    import coolbubble_amd as cb
    hip = HipEvents()
    N = CONFIG["nGrids"]
    sites = N ** 3
    K, W = args.steps, args.warmup

    if args.only == "ensemble":
        print(json.dumps({"ensemble": run_ensemble(cb, hip, args.ensemble_steps, log)}), flush=True)
        return 0
    if args.only == "serving":
        print(json.dumps({"serving": run_serving(cb, args.serve_requests, args.serve_rate, log)}), flush=True)
        return 0
    if args.only == "run":
        print(json.dumps({"run": run_driver(cb, log)}), flush=True)
        return 0
    if args.only == "expanding":
        print(json.dumps({"expanding": run_expanding(cb, hip, K, log)}), flush=True)
        return 0
    copy_gbps = cb.bandwidth_probe(1 << 31, 20)
    log("device-to-device copy: %.1f GB/s (read + write)" % copy_gbps)
    if args.precision != "both":
        sim, ms, row = run_precision(cb, hip, args.precision, K, W, log, args.variant)
        if not args.no_spectrum:
            log("%s spectrum: %.3f ms" % (args.precision, spectrum_time(sim, hip)))
        sim.close()
        w = 4 if args.precision == "fp32" else 8
        print(json.dumps({"precision": args.precision, "ms_per_step": ms / K,
                          "site_updates_per_s": sites * K / (ms * 1e-3),
                          "achieved_GBps": 4 * w * sites / (ms * 1e-3 / K) / 1e9}), flush=True)
        return 0
    sim32, ms32, row32 = run_precision(cb, hip, "fp32", K, W, log, args.variant)
    # FP32 state after the same W + K steps, for its errors against FP64
    phi32, _ = sim32.fields()
    _, _, d32 = sim32.spectrum()
    sel32 = sim32.kernel_variants()[1]
    info32 = sim32.variant_info(sel32) if sel32 >= 0 else None
    fuse32 = info32.get("fuse", 1) if info32 else 1
    launch_us32 = launch_time_live(sim32, hip, fuse32)
    log("fp32: one plain launch (%d steps) %.1f us, live" % (fuse32, launch_us32))
    spec32 = spectrum_time(sim32, hip)
    log("fp32 spectrum: %.3f ms" % spec32)
    slab = slab_share(sim32, info32, log)   # last: only the slab advances
    sim32.close()
    sim64, ms64, row64 = run_precision(cb, hip, "fp64", K, W, log, args.variant)
    phi64, _ = sim64.fields()
    _, _, d64 = sim64.spectrum()
    field_err32 = float(np.max(np.abs(phi32 - phi64)) / np.max(np.abs(phi64)))
    big = d64 > 1e-6 * np.max(d64)
    spec_err32 = float(np.max(np.abs(d32[big] - d64[big]) / d64[big])) if len(d32) == len(d64) else None
    del phi32, phi64
    spec64 = spectrum_time(sim64, hip)
    log("fp64 spectrum: %.3f ms" % spec64)

    su32 = sites * K / (ms32 * 1e-3)
    traffic32, steps_per_launch32 = None, info32.get("fuse", 1) if info32 else 1
    prof32 = {}
    try:
        recs = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                           "hbm_traffic.json")))["stepper"]
        for rec in recs:
            if info32 and rec["precision"] == "fp32" and rec["variant_config"] == info32:
                traffic32 = rec["traffic_bytes_per_launch"]
                prof32 = {"kernel": rec["kernel"], "kernel_avg_launch_us_profiled": rec["avg_launch_us"],
                          "delivered_GBps_profiled": traffic32 / (rec["avg_launch_us"] * 1e-6) / 1e9,
                          "algorithmic_GBps_profiled": 16.0 * sites * rec["steps_per_launch"]
                          / (rec["avg_launch_us"] * 1e-6) / 1e9}
    except (OSError, KeyError, ValueError, TypeError):
        pass
    su64 = sites * K / (ms64 * 1e-3)
    bytes32 = 16.0 * sites          # compulsory bytes of one pass: read phi, pi; write pi, phi (FP32)
    bytes64 = 32.0 * sites
    # the dominant kernel moves the lattice's compulsory bytes once per
    # launch, whatever the steps it takes in that pass
    ach32 = bytes32 / (launch_us32 * 1e-6)
    one_pass_su32 = HBM_PEAK / 16.0      # site-updates/s of one HBM pass per step at peak (FP32)
    one_pass_su64 = HBM_PEAK / 32.0
    # energies after the same W+K steps from the same seeded lattice
    err = {name: abs(row32[i] - row64[i]) / abs(row64[i])
           for i, name in ((2, "Etot"), (3, "Ekin"), (4, "Egrad"), (5, "Epot"))}
    err["field_max_abs_over_max"] = field_err32
    err["spectrum_bins_max_rel"] = spec_err32
    fft_bytes = lambda w: (2 * 3 - 1) * w * sites   # (2d-1) w N^d per spectrum
    cpu = None
    if not args.no_cpu:
        cdt, threads = cpu_baseline(sim64, args.cpu_steps, log)
        cpu = {"value": sites * args.cpu_steps / cdt, "unit": "site-updates/s", "cores": threads,
               "kind": "port", "cpu_model": cpu_model(),
               "sample": "%d oracle steps (FP64, energies every step) of the same 512^3 lattice "
                         "from the GPU-seeded state" % args.cpu_steps}
    sim64.close()
    expd = None if args.no_expanding else run_expanding(cb, hip, K, log)
    ens = None if args.no_ensemble else run_ensemble(cb, hip, args.ensemble_steps, log)
    srv = None if args.no_serving else run_serving(cb, args.serve_requests, args.serve_rate, log)
    out = {
        "metric": "site-updates/s (FP32 stepper, d=3 N=512 static, mode 1 32 bubbles)",
        "value": su32,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms32 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded bubbles, reference mode 1 generator)",
        "config": {"workload": "lattice time-stepper d=3 N=512^3 L=64 c=4 mode 1 nBubbles=32 "
                               "dt=0.025 static, energies every 40 steps",
                   "sites": sites, "energy_every": ENERGY_EVERY},
        "roofline": dict({"bound": "hbm", "achieved": ach32 / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": ach32 / HBM_PEAK, "traffic": traffic32,
                          "steps_per_launch": steps_per_launch32,
                          "kernel_avg_launch_us_live": launch_us32,
                          "compulsory_bytes_per_launch": bytes32,
                          "one_pass_roofline_site_updates_per_s": one_pass_su32,
                          "site_updates_over_one_pass_roofline": su32 / one_pass_su32,
                          "note": "achieved = compulsory HBM bytes of one launch of the dominant kernel (16 B per "
                                  "site: read phi, pi; write pi, phi; once per launch, whatever the steps it "
                                  "takes) / its average plain launch, timed live with HIP events on the "
                                  "simulation's stream (kernel_avg_launch_us_live); frac = achieved / 8 TB/s. "
                                  "The kernel takes steps_per_launch leapfrog steps per pass over HBM, so the "
                                  "site-update rate can exceed the one-pass roofline (8 TB/s / 16 B): "
                                  "site_updates_over_one_pass_roofline. traffic = measured HBM bytes per launch "
                                  "(profiles/hbm_traffic.json, rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy "
                                  "kernel); delivered_GBps_profiled = traffic / the profiled average launch"},
                         **prof32),
        "cpu_baseline": cpu,
        "fp64": {"value": su64, "ms_per_step": ms64 / K,
                 "site_updates_over_one_pass_roofline": su64 / one_pass_su64},
        "fp32_over_fp64": su32 / su64,
        "expanding": expd,
        "slab_share_8_gpus": slab,
        "ensemble": ens,
        "serving": srv,
        "measured_copy_GBps": copy_gbps,
        "frac_of_measured_copy": ach32 / (copy_gbps * 1e9),
        "fp32_energy_rel_err_vs_fp64": err,
        "spectrum": {"fp32_ms": spec32, "fp64_ms": spec64,
                     "fp32_GBps_algorithmic": fft_bytes(4) / (spec32 * 1e-3) / 1e9,
                     "fp64_GBps_algorithmic": fft_bytes(8) / (spec64 * 1e-3) / 1e9},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
