/*
 * coolbubble.h -- C ABI of the MI355X lattice time-stepper (libcoolbubble.so).
 *
 * Drop-in for CoolBubble's `Bubble` object and lattice helpers
 * (src/bubble.hpp:34-135, src/gridfuncs.h:27-41).  The reference has no FFI:
 * its boundary is C++ linkage plus process globals (src/gridfuncs.h:27-32,
 * defined at src/main.cpp:45-50) read by member initialisers, and
 * compile-time constants (src/def.h).  Here the same state lives behind an
 * opaque handle, every parameter is a run-time field of cb_params, and no
 * call exits, throws or prints: each returns a cb_status.
 *
 * Threading: one handle owns one HIP stream; calls on one handle must be
 * serialised by the caller; distinct handles may be driven from different
 * threads.  All device memory is owned by the library and freed by
 * cb_destroy.  Host buffers are owned by the caller.
 */
#ifndef COOLBUBBLE_H
#define COOLBUBBLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cb_status {
    CB_OK = 0,
    CB_INVALID_ARG = 1,     /* reference: exit(1) in parseArguments, src/main.cpp:105-138 */
    CB_CFL_VIOLATION = 2,   /* reference: CheckStability returns false, src/main.cpp:171-183 */
    CB_OUT_OF_MEMORY = 3,
    CB_HIP_ERROR = 4,
    CB_NAN_DETECTED = 5,
    CB_NO_DEVICE = 6,
    CB_ADMISSION_REJECTED = 7,
    CB_NOT_READY = 8,
    CB_UNSUPPORTED = 9
} cb_status;

typedef enum cb_precision {
    CB_FP64 = 0,            /* phi, pi stored and stepped in double (reference arithmetic) */
    CB_FP32 = 1             /* phi, pi stored and stepped in float; reductions, scale factor
                               and spectrum bins in double */
} cb_precision;

/* Run parameters.  Field comments give the reference global / #define. */
typedef struct cb_params {
    int D;                       /* D, src/gridfuncs.h:27 (1, 2 or 3) */
    int nGrids;                  /* nGrids, src/gridfuncs.h:28 (>= 2) */
    double L;                    /* L, src/gridfuncs.h:29 */
    double c;                    /* c, src/gridfuncs.h:32 */
    int readxns;                 /* readxns (IC mode 1, 0, -1, -2, -3), src/gridfuncs.h:31 */
    double dt;                   /* TIME_STEP, src/def.h:34 */
    int nBubbles;                /* nBubbles, src/def.h:46 */
    double Trad;                 /* Trad, src/def.h:48 */
    double DeltaR;               /* DeltaR, src/def.h:50 */
    int test_wave;               /* TEST, src/def.h:51 */
    int expansion;               /* expansion, src/def.h:56 (3D only, bubble.hpp:77) */
    int alpha;                   /* ALPHAVAL, src/def.h:57 (0 cosmic, 1 conformal time) */
    double Mpl;                  /* Mpl, src/def.h:58 */
    unsigned int rand_seed;      /* srand seed of mode 1 (42: src/bubble.cpp:401) */
    unsigned long long thermal_base; /* thermal seed base (12345: src/bubble.cpp:137) */
    int precision;               /* cb_precision (no counterpart: the reference is FP64) */
    int device;                  /* HIP device ordinal */
    int thermal_spectrum;        /* mode -3 occupancy: 0 the reference's Bose-Einstein spectrum
                                    (GenerateThermalFluc, src/bubble.cpp:116-130,219-228: <|phi_k|^2> =
                                    n_k / omega, <|pi_k|^2> = n_k omega, continuum k, T = Trad mT);
                                    1 Rayleigh-Jeans (no counterpart: <|phi_k|^2> = T / omega^2,
                                    <|pi_k|^2> = T, omega^2 = k_lat^2 + 2c with the lattice momentum
                                    k_lat^2 = sum (2/dx)^2 sin^2(k dx / 2), T = Trad) */
} cb_params;

/* The 16 columns of one EnergyDensities.dat row (outEP, src/bubble.cpp:977-989):
 * t, a, Etot, Ekin, Egrad, Epot, sqrt(EtotVar), sqrt(EkinVar), sqrt(EgradVar),
 * sqrt(EpotVar), <phi>, sqrt(<phi>Var), <phi'>, sqrt(<phi'>Var), H, C. */
typedef struct cb_energies {
    double col[16];
} cb_energies;

/* Scalar state beside the fields (src/bubble.hpp:75-84). */
typedef struct cb_state {
    double a, ap, H, C, t, EkinHalfPrev;
    long long step;
} cb_state;

typedef struct cb_sim cb_sim;

/* Reference defaults from src/def.h:29-58 (D=2, nGrids=512, L=32, c=4, mode 1,
 * TIME_STEP 0.004, 5 bubbles, Trad 0.2, static, Mpl 200), FP64, device 0. */
cb_status cb_default_params(cb_params *p);

/* CheckStability, src/main.cpp:171-183: CB_OK iff dt < dx/sqrt(D). */
cb_status cb_check_stability(const cb_params *p);

/* Bubble::Bubble, src/bubble.cpp:156-316: allocate the lattice on the
 * device, seed the initial condition (mode 1 draws its centres like
 * GenerateNucleationPoints, :397-417; mode 0 takes the caller's centres,
 * which replace <run>/__xns.dat, :420-453), measure the initial energies and
 * step pi back by half a step.  xns: n_xns rows of (x, y, z); ignored unless
 * readxns == 0. */
cb_status cb_create(const cb_params *p, const double *xns, int n_xns, cb_sim **out);

/* Bubble::oneStep, src/bubble.cpp:849-876, nsteps times.  Energies are
 * measured on the last step, so cb_energies afterwards gives the row the
 * reference would print after that step (src/main.cpp:280). */
cb_status cb_step(cb_sim *s, int nsteps);

/* outEP, src/bubble.cpp:977-989: the current energy row. */
cb_status cb_energies_row(cb_sim *s, cb_energies *out);

/* FourierTransform + ComputePowerSpectrum, src/bubble.cpp:642-709: for each
 * non-empty |k| shell (width 2 pi / L) the mean comoving k, k / a and
 * Delta(k).  cap = capacity of each output array; *nbins = bins written. */
cb_status cb_power_spectrum(cb_sim *s, double *kmean, double *kphys, double *delta,
                            int cap, int *nbins);

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): cb_power_spectrum with no allocation.  cb_power_spectrum
 * allocates its work arrays (the half-complex array, as large as the
 * lattice, twiddles, bin partials) on its first call on a handle; this
 * variant carves them from a static device array of
 * CB_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends, and
 * calls no hipMalloc, hipFree or other allocating routine.  Lattices whose
 * work arrays exceed it (e.g. 3D above about 192^3 in FP64, 2D above 2048^2)
 * return CB_UNSUPPORTED.  Results agree with cb_power_spectrum's.  Calls run
 * one at a time in a process (a lock held until the call's stream drains);
 * calls on one GPU that overlapped would race on the shared scratch. */
#define CB_NOALLOC_SCRATCH_BYTES (64u << 20)
cb_status cb_power_spectrum_noalloc(cb_sim *s, double *kmean, double *kphys, double *delta,
                                    int cap, int *nbins);

/* The fields a snapshot writes (OutField, src/bubble.cpp:887-917): phi at
 * the current whole step and pi half a step behind, as double (FP32 runs are
 * widened).  Either pointer may be NULL.  Each holds nGrids^D values in
 * lattice order (src/gridfuncs.cpp:25-26). */
cb_status cb_get_fields(cb_sim *s, double *phi, double *pi);

/* Scalar state, and state import (phi, pi given as double; FP32 runs round
 * them).  cb_set_state replaces the fields and scalars; EkinHalfPrev must be
 * the kinetic energy of pi (the previous half step), as the ctor and
 * oneStep leave it.  No reference counterpart (the reference cannot
 * restart, its snapshots being float32). */
cb_status cb_get_state(cb_sim *s, cb_state *st);
cb_status cb_set_state(cb_sim *s, const cb_state *st, const double *phi, const double *pi);

/* Centres used by mode 1 / 0 (the contents of __xns.dat, src/bubble.cpp:408-416):
 * *n = rows available; min(cap, *n) rows of (x, y, z) written to xyz. */
cb_status cb_get_xns(cb_sim *s, double *xyz, int cap, int *n);

/* Step-kernel choice (no reference counterpart).  variant -1 = the general
 * kernel (any lattice); variant k >= 0 = the k-th LDS-tiled marching
 * configuration valid for this (D, nGrids, precision).  cb_kernel_variants
 * reports how many tiled configurations exist and which one is selected.
 * cb_create selects the configuration nearest to what cb_autotune picks on
 * MI355X for lattices of this kind (static boxes: three steps per launch,
 * FP64 at 3D nGrids >= 512 two; expanding: one), or the general kernel for 3D lattices below 2^19 sites
 * and 2D lattices below 512^2. */
cb_status cb_kernel_variants(cb_sim *s, int *count, int *selected);
cb_status cb_select_kernel(cb_sim *s, int variant);
/* Shape of tiled variant `variant`: info = {vector width, vectors per thread,
 * threads per block, rows per tile, planes per block, prefetch (1 none, 2
 * the next plane's loads one plane ahead, 3 two planes ahead, 4 phi loaded
 * straight into LDS buffers two planes ahead), non-temporal stores, steps
 * per launch (1, or 2 / 3 leapfrog steps per pass over the lattice, static
 * box only)}. */
cb_status cb_kernel_variant_info(cb_sim *s, int variant, int info[8]);

/* Autotune: time every variant over one call of `steps` steps (after an
 * untimed call of the same length, so that its kernels are loaded), then the
 * 8 fastest again over four calls of `steps` steps in a row (each call
 * measures the energies on its last step, as a run printing every `steps`
 * steps does; the lattice state is saved and restored around the trials),
 * and select the fastest.  ms_per_step (optional, capacity cap) receives
 * each variant's time per step (the longer trial where there was one),
 * general kernel last. */
cb_status cb_autotune(cb_sim *s, int steps, double *ms_per_step, int cap, int *best);

/* The HIP stream (hipStream_t) every kernel of this handle runs on. */
cb_status cb_stream(cb_sim *s, void **stream);

/* ---- Ensembles (no reference counterpart: the reference runs one
 * simulation per process) ----
 * A batch steps n independent simulations together: one persistent launch
 * per step covers every tile of every member (members may differ in D,
 * nGrids, L, c, dt, mode, seeds; they share the precision and device), the
 * energies of all members are finalised by one launch, and runs of 40 steps
 * are captured once as a HIP graph and replayed.  A member steps
 * bit-identically alone (cb_create with the same params, same kernel
 * variant) or in the batch.  cb_batch_member gives a handle for reading a
 * member (energies, fields, spectrum, state); it must not be destroyed on
 * its own. */
typedef struct cb_batch cb_batch;
cb_status cb_batch_create(const cb_params *members, int n, cb_batch **out);
cb_status cb_batch_step(cb_batch *b, int nsteps);
cb_status cb_batch_member(cb_batch *b, int i, cb_sim **member);
cb_status cb_batch_size(cb_batch *b, int *n, long long *sites);
cb_status cb_batch_stream(cb_batch *b, void **stream);
/* Autotune the batched launch: time `steps` steps of the whole batch for a
 * set of launch shapes (vectors per thread, prefetch depth, non-temporal
 * stores, preferred chunk length, several steps per launch for static
 * members: three where a three-step tile fits the batched launch, else
 * two; each member then takes its tallest tile of that shape),
 * state saved and restored around the trials; keep the fastest.  shape
 * (optional, 5 ints) receives it, best_ms its time per step. */
cb_status cb_batch_autotune(cb_batch *b, int steps, int shape[5], double *best_ms);
/* Set the batched launch shape directly (the five values cb_batch_autotune
 * reports: vectors per thread 1/2/4, prefetch 1/2, non-temporal 0/1,
 * chunk planes >= 1, several steps per launch for static members 0/1). */
cb_status cb_batch_set_shape(cb_batch *b, const int shape[5]);
cb_status cb_batch_destroy(cb_batch *b);

/* ---- Serving (no reference counterpart) ----
 * A request is one run: parameters, nt steps and a number of snapshots
 * (spaced as SnapshotSchedule, src/main.cpp:201-223, each with a power
 * spectrum).  The server admits queued requests in arrival order while the
 * device-memory budget and the slot count allow, keeps their lattices
 * resident, and advances every active request together: one batched launch
 * per step over all of them, in ticks that end at the nearest event of any
 * request (an energy row every 40 steps, a snapshot, completion).  Requests
 * arriving while others run are admitted at the next tick.  Each request's
 * results are bit-identical to running it alone with the same kernel
 * variant.  Requests' device memory comes from the device's default memory
 * pool, allocated and freed in order on the server's stream; the server
 * sets that pool to keep what it frees (for the next requests) and trims it
 * when it is destroyed.  Times are seconds on the server's steady clock. */
typedef struct cb_server cb_server;
typedef struct cb_server_opts {
    int device;
    int precision;               /* all requests run in this precision */
    long long mem_budget;        /* bytes of device memory for lattices (0: 90% of free) */
    int max_active;              /* slot count (0: unlimited) */
} cb_server_opts;
typedef struct cb_request {
    cb_params params;            /* precision and device are taken from the server */
    int nt;                      /* steps */
    int snapshots;               /* snapshots after the initial state, each with a spectrum */
    unsigned long long tag;      /* caller's tag, returned in the result */
} cb_request;
typedef struct cb_result {
    unsigned long long id, tag;
    cb_status status;            /* CB_OK, CB_ADMISSION_REJECTED (can never fit), CB_NOT_READY (cancelled) */
    double t_submit, t_admit, t_done;
    long long site_updates;
    int n_rows;                  /* energy rows recorded (t = 0 and every 40 steps) */
    int n_spectra;               /* spectra recorded (initial + snapshots) */
    double final_row[16];        /* energy row after the last step */
} cb_result;
cb_status cb_server_create(const cb_server_opts *o, cb_server **out);
cb_status cb_server_submit(cb_server *s, const cb_request *r, unsigned long long *id);
/* A queued request is dropped at once, an active one at its next tick; its
 * result has status CB_NOT_READY.  CB_INVALID_ARG for an id that is neither
 * queued nor active (unknown, or finished). */
cb_status cb_server_cancel(cb_server *s, unsigned long long id);
/* completed requests since the last poll: up to cap results, *n written */
cb_status cb_server_poll(cb_server *s, cb_result *out, int cap, int *n);
/* stored outputs of a completed (polled) request */
cb_status cb_server_rows(cb_server *s, unsigned long long id, double *rows, int cap_rows, int *n);
cb_status cb_server_spectrum(cb_server *s, unsigned long long id, int which, double *kmean, double *delta,
                             int cap, int *nbins);
cb_status cb_server_release(cb_server *s, unsigned long long id);   /* drop stored outputs */
/* queued, active, completed-but-unpolled requests; device bytes in use */
cb_status cb_server_stats(cb_server *s, int *queued, int *active, int *done, long long *bytes);
/* block until no request is queued or active */
cb_status cb_server_drain(cb_server *s);
/* The kernel variant the server gives a request with the parameters of s
 * (for a static box of 2^23 sites or more the variant a lone run starts on:
 * such requests step on their own kernels; else a two-step variant where one
 * fits the batched launch, else a one-step one; -1 = the general kernel).
 * Selecting it on a lone simulation reproduces the server's results for that
 * request bit for bit. */
cb_status cb_server_variant(const cb_sim *s, int *variant);
cb_status cb_server_destroy(cb_server *s);

/* ---- Run driver (the time loop of src/main.cpp:229-310) ----
 * cb_run runs a whole simulation as the reference's main does and writes
 * its text outputs into <out_dir>/<run folder> (the folder name of
 * src/main.cpp:243-247): parameters.dat (OutParameters, src/bubble.cpp:
 * 1018-1032, plus the elapsed-time line), EnergyDensities.dat (header and
 * outEP rows at t = 0 and every energy_every steps, :958-989), __xns.dat
 * (mode 1; mode 0 reads it from the run folder unless xns is given,
 * :397-453), _phi_k_t=NNNN.dat at the initial state and every snapshot
 * (:642-709), FullField.dat in 1D (:993-1006), and (field_snapshots = 1)
 * the HDF5 field snapshots _phi_t=NNNN.h5 at the same times (OutField,
 * :883-954; see cb_write_snapshot_h5), and (wall_tracking = 1) gamma.dat
 * (TrackWall / WallFinder / outGamma, :535-621, :1009-1014).  The snapshot
 * files are written by a host thread while the time steps go on.  Plots and
 * videos are not produced.  Nothing is printed. */
typedef struct cb_run_opts {
    cb_params params;
    int nt;                      /* time steps (nt_default 1000, src/def.h:32) */
    int snapshots;               /* snapshots after the initial one (nSnapshots 20) */
    int snapshot_spacing;        /* SnapshotSpacing 0 | 1 (src/def.h:67) */
    int energy_every;            /* OutEnergyEvery 40 */
    int fft;                     /* FFT 1: a spectrum file per snapshot, 0: none */
    int full1d_every;            /* Full1DFieldEvery 100 (1D) */
    int field_snapshots;         /* 1: _phi_t=NNNN.h5 with every spectrum (the reference always writes them) */
    int memorysave;              /* memorysave (src/def.h:64) 0: the .h5 files hold phi, pi, Ekin, gradphi,
                                    potEnergy; 1: phi, pi */
    int deflate;                 /* gzip level of the .h5 datasets, 0..9 (the reference: 9) */
    int wall_tracking;           /* WallTracking (src/def.h:72) 1: gamma.dat (1D, modes -1 and -2) */
    int wall_track_every;        /* WallTrackEvery 400 */
    const char *out_dir;         /* folder in which the run folder is created (Path_default "./") */
    const double *xns;           /* mode 0: n_xns centres (x, y, z); NULL: read <run folder>/__xns.dat */
    int n_xns;
    char folder[1024];           /* out: the run folder */
    double elapsed_s;            /* out: wall time of cb_run */
} cb_run_opts;
/* src/def.h defaults (and cb_default_params for params). */
cb_status cb_default_run_opts(cb_run_opts *o);
/* The run folder name (no directory part), e.g.
 * "Sim2_D_nt_1000_dt_0.01250_nGrids_512_L_32.0_Mode_1_nB_5_c_4.00_FFT_1_Mpl_200". */
cb_status cb_run_folder_name(const cb_run_opts *o, char *buf, int cap);
cb_status cb_run(cb_run_opts *o);

/* ---- HDF5 field snapshots (Bubble::OutField, src/bubble.cpp:883-954) ----
 * The HDF5 C library (1.10.3 or later) is loaded at run time: the path in
 * $CB_HDF5_LIBRARY, else libhdf5.so on the loader's path, else the usual
 * install locations.  CB_OK when it loads, else CB_UNSUPPORTED (reason in
 * cb_last_error). */
cb_status cb_hdf5_available(void);
/* Writes one snapshot file as OutField does: nfields = 2 (phi, pi) or 5 (phi,
 * pi, Ekin, gradphi, potEnergy) float datasets of shape (n), (n, n) or
 * (n, n, n), fields[f] holding n^D values in lattice order, where n is the
 * number of saved points per axis: nGrids, or for 2D above 2048 and 3D above
 * 512 points ceil(nGrids / reduc) with reduc = nGrids / 2048 (2D) or
 * nGrids / 512 (3D).  Chunked by whole rows (1D the whole field, 2D
 * min(n, max(1, 65536 / n)) rows, 3D one plane), shuffle + deflate (level
 * 0..9, the reference: 9); double attributes D, L, nGrids, reduc, c, t, a. */
cb_status cb_write_snapshot_h5(const char *path, int D, int nGrids, double L, double c, double t, double a,
                               int nfields, const float *const *fields, int deflate);

/* Measurement helper (no reference counterpart): achieved bandwidth of a
 * plain 16-byte-vector device-to-device copy of `bytes` bytes, repeated
 * `reps` times, counting read + write bytes, in GB/s. */
cb_status cb_bandwidth_probe(int device, long long bytes, int reps, double *gbps);

/* Measurement helper (no reference counterpart): one GPU's share of a run
 * sharded into slabs along the slowest axis, at its real shape.  Times
 * `launches` launches of s's multi-step tiled variant (static 3D box)
 * restricted to the first `planes` planes of the lattice, which read their
 * halo planes (steps-per-launch planes per face) where a rank would hold
 * the planes its neighbours sent; then the device-to-device copies that
 * stand in for that exchange (those planes of phi and pi, both faces).
 * planes: a multiple of the variant's chunk length, leaving room for the
 * halo planes.  Only the slab advances, so s's fields are scratch
 * afterwards.  ms_launch / ms_exchange: average per launch. */
cb_status cb_slab_probe(cb_sim *s, int planes, int launches, double *ms_launch, double *ms_exchange);

/* Number of lattice sites N = nGrids^D (src/bubble.hpp:43). */
long long cb_nsites(const cb_sim *s);

cb_status cb_destroy(cb_sim *s);

/* Text of a status code, and the message of the last failed call on this
 * host thread. */
const char *cb_status_string(cb_status st);
const char *cb_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* COOLBUBBLE_H */
