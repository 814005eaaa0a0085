"""Write the full-length oracle fixtures of the two large measured workloads.

    python tests/golden/make_fullsize.py c2 [threads] [checkpoint dir]   # ~2 h on 8 cores
    python tests/golden/make_fullsize.py c3 [threads] [checkpoint dir]   # ~1.5 h on 8 cores

With a checkpoint directory the run saves its state there every 200 steps
(phi, pi, the scalars and the outputs so far) and resumes from it when
started again; the oracle's step is a function of that state alone, so a
resumed run writes the same numbers.

Each runs the oracle (oracle/cb_oracle.c, the CPU restatement of the
reference's Bubble ctor and oneStep, src/bubble.cpp:156-316,849-876) for the
whole run and stores, as a small .npz next to this script:

* rows: the outEP row (src/bubble.cpp:977-989) at t = 0 and after every 40th
  step (the cadence of src/main.cpp:280);
* spectra: FourierTransform + ComputePowerSpectrum (:642-709) at t = 0 and at
  every step of SnapshotSchedule with 20 snapshots (src/main.cpp:201-223);
* checksums at every snapshot: sum, sum of squares and max |x| of phi and pi;
* site values: phi and pi at 4096 fixed-stride sites after the last step.

c2: d=3, N=512, L=64, c=4, mode 1 (32 bubbles), dt=0.025, static, nt=2000.
c3: d=3, N=384, L=48, c=2, mode -3 (Trad=0.1), expanding (Mpl=10), dt=0.025,
    nt=3000.
The fixtures hold no reference code: inputs are the parameters above, the
outputs are numbers.
"""
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

CONFIGS = {
    "c2": (dict(D=3, nGrids=512, L=64.0, c=4.0, readxns=1, nBubbles=32, dt=0.025), 2000),
    "c3": (dict(D=3, nGrids=384, L=48.0, c=2.0, readxns=-3, Trad=0.1, expansion=1, Mpl=10.0, dt=0.025), 3000),
}
N_SITES = 4096


def snapshot_steps(nt, dt, n):
    """SnapshotSchedule, SnapshotSpacing = 0 (src/main.cpp:201-223)."""
    T = nt * dt
    out = []
    for k in range(1, n + 1):
        st = min(nt, max(1, int(math.floor(T * k / n / dt + 0.5))))
        if not out or st > out[-1]:
            out.append(st)
    return out


def fixed_sites(n_total):
    """N_SITES site indices at a fixed stride, offset so that every plane
    and row position is visited."""
    stride = n_total // N_SITES
    return (np.arange(N_SITES, dtype=np.int64) * stride + (np.arange(N_SITES) * 7919) % stride) % n_total


def checksums(x):
    x = x.reshape(-1)
    return [float(np.sum(x)), float(np.sum(x * x)), float(np.max(np.abs(x)))]


def path_of(name):
    kw, nt = CONFIGS[name]
    return os.path.join(HERE, "%s_d%d_n%d_nt%d.npz" % (name, kw["D"], kw["nGrids"], nt))


def _save(ckpt, o, done, rows, row_steps, spec_steps, spec_k, spec_d, sums):
    tmp = os.path.join(ckpt, "state.tmp.npz")
    np.savez(tmp, done=done, phi=o.phi(), pi=o.pi(), scalars=o.scalars(), rows=np.array(rows),
             row_steps=np.array(row_steps), spec_steps=np.array(spec_steps),
             spec_k=np.concatenate(spec_k),
             spec_d=np.concatenate(spec_d), spec_len=np.array([len(k) for k in spec_k]),
             sums=np.array(sums))
    os.replace(tmp, os.path.join(ckpt, "state.npz"))


def main(name, nthreads=0, ckpt=None):
    kw, nt = CONFIGS[name]
    snaps = snapshot_steps(nt, kw["dt"], 20)
    events = sorted(set(range(40, nt + 1, 40)) | set(snaps))
    t0 = time.time()
    state = os.path.join(ckpt, "state.npz") if ckpt else None
    if state and os.path.exists(state):
        z = np.load(state)
        o = oracle.Sim(nthreads=nthreads, blank=True, **kw)
        o.phi()[...] = z["phi"]
        o.pi()[...] = z["pi"]
        o.set_scalars(z["scalars"])
        done = int(z["done"])
        rows, row_steps = list(z["rows"]), list(z["row_steps"])
        spec_steps, sums = list(z["spec_steps"]), [list(x) for x in z["sums"]]
        cut = np.cumsum(z["spec_len"])[:-1]
        spec_k, spec_d = np.split(z["spec_k"], cut), np.split(z["spec_d"], cut)
        print("%s: resumed at step %d" % (name, done), flush=True)
    else:
        o = oracle.Sim(nthreads=nthreads, **kw)
        print("%s: created in %.0f s" % (name, time.time() - t0), flush=True)
        rows, row_steps = [o.energy_row()], [0]
        spec_steps, spec_k, spec_d, sums = [0], [], [], []
        k, _, d = o.spectrum()
        spec_k.append(k)
        spec_d.append(d)
        sums.append(checksums(o.phi()) + checksums(o.pi()))
        done = 0
    if not kw.get("expansion"):
        # a static box feeds nothing back from its energies: measure them on
        # the rows' steps only (and the kinetic sum the step before), which
        # prints the every-step rows bit for bit (tests/test_oracle.py)
        o.set_energy_cadence(40, done)
    for ev in events:
        if ev <= done:
            continue
        o.step(ev - done)
        done = ev
        if ev % 40 == 0:
            rows.append(o.energy_row())
            row_steps.append(ev)
        if ev in snaps:
            k, _, d = o.spectrum()
            spec_steps.append(ev)
            spec_k.append(k)
            spec_d.append(d)
            sums.append(checksums(o.phi()) + checksums(o.pi()))
            print("%s: step %d, %.0f s" % (name, ev, time.time() - t0), flush=True)
        if ckpt and ev % 200 == 0 and ev < nt:
            _save(ckpt, o, done, rows, row_steps, spec_steps, spec_k, spec_d, sums)
    sites = fixed_sites(o.n)
    meta = dict(params=kw, nt=nt, snapshots=20, generator="tests/golden/make_fullsize.py")
    np.savez_compressed(path_of(name), meta=np.array(json.dumps(meta)), row_steps=np.array(row_steps),
                        rows=np.array(rows), spec_steps=np.array(spec_steps), spec_k=np.array(spec_k),
                        spec_delta=np.array(spec_d), checksums=np.array(sums), sites=sites,
                        phi_sites=o.phi().reshape(-1)[sites].copy(), pi_sites=o.pi().reshape(-1)[sites].copy())
    print("%s: wrote %s in %.0f s" % (name, path_of(name), time.time() - t0), flush=True)


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 0, sys.argv[3] if len(sys.argv) > 3 else None)
