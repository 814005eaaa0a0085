// cb_tiled.hpp -- LDS-tiled marching step kernel: configurations and launcher.
#pragma once

#include "cb_kernels.hpp"

namespace cb {

// One launch shape of the tiled kernel.  A block owns TR rows (3D; 1 in 2D)
// of every plane of an i-chunk of NI planes and marches through the chunk:
// the planes i-1, i, i+1 of its own columns sit in registers, the plane i
// tile plus one halo row on each side sits in LDS (double-buffered) for the j
// and k taps.
struct TiledCfg {
    int vw;        // elements per vector load (16 B: 4 float / 2 double); 1 = scalar
    int items;     // vectors per thread per plane (template: 1, 2, 4)
    int nthr;      // threads per block (template: 128, 256)
    int TR;        // rows per tile (3D)
    int NI;        // planes per block
    int pf;        // 2: next-but-one plane (and next halo rows) prefetched into registers; 1: none
    int nt;        // 1: non-temporal stores of phi and pi
    int fuse;      // steps per launch: 1, or 2 / 3 (temporal blocking, static box only)
};

// Geometry of one launch, derived from (N, cfg).
struct TiledGeom {
    int TR, NI, ntr, nic, NV, nt;
};

// Is cfg usable for this lattice?  (vector width divides N, tile fits LDS,
// every vector of a tile plane has an owner thread)
bool tiled_valid(const SimConst &k, int fp32, const TiledCfg &c);

// Blocks a launch of cfg uses (energy partials are written per block).
int tiled_blocks(const SimConst &k, const TiledCfg &c);

// Candidate configurations for (D, N, precision), best-guess first.
int tiled_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap);

hipError_t launch_step_tiled_cfg(const SimConst &k, int fp32, int lagged, int energy, const TiledCfg &c,
                                 const StepPtrs &p, const DevScalars *sc, double *partials,
                                 hipStream_t st);

// Launch with the configuration registered under index `which` for this sim
// (see tiled_register); *nblocks receives the block count.
hipError_t launch_step_tiled(const SimConst &k, int fp32, int lagged, int energy, int which,
                             const StepPtrs &p, const DevScalars *sc, double *partials, hipStream_t st,
                             int *nblocks);

// One ensemble member as the batched step kernel sees it: constants, tile
// geometry, both buffers of each field and the parity of the current ones at
// the start of a launch sequence.
struct BatchMember {
    SimConst k;
    TiledGeom g;
    int lagged;
    int tile_begin;     // first work item of this member in the batch
    int ntiles;
    int par_phi, par_pi;
    int nsteps;         // steps this member takes in the current launch sequence
    int energy_it;      // step index of this member's energy step; -1: every energy launch
    // multi-step members (launch_batch_multi): fused = steps per launch (2
    // or 3; 0 for one-step members); g / ntiles are the multi-step tile's,
    // g1 / ntiles1 the one-step companion's (the leftover steps)
    int fused;
    int ntiles1;
    TiledGeom g1;
    void *phi[2];
    void *pi[2];
    DevScalars *sc;
    double *partials;
};

// Batched persistent step: every tile of every member, one launch.  All
// members share the precision; their tiled configs must have 256 threads and
// at most `items` vectors per thread; `pf` is the launch's prefetch depth.
// A member's results depend only on its (TR, NI, threads), not on items, pf
// or the non-temporal flag, so it steps bit-identically alone with a variant
// of the same (TR, NI, threads).
// step_it = index of this launch in the sequence: a member takes part while
// step_it < its nsteps (its buffers flip with step_it & 1).  With `energy`
// set, a member measures energies if it expands, if its energy_it is -1, or
// if its energy_it equals step_it; with `split` also set, the other members
// take the plain step in a launch of their own first (an energy launch runs
// only the members that measure energies, so the plain ones keep the lighter
// kernel).
hipError_t launch_batch_step(int fp32, int energy, int split, int items, int pf, const BatchMember *table, int nmem,
                             int total_tiles, int grid, size_t shm, int step_it, hipStream_t st);
// Finalise the energies of the members that measured them in launch step_it.
hipError_t launch_batch_finalize(const BatchMember *table, int nmem, int step_it, hipStream_t st);
// Batched multi-step launch number L of a sequence, for static members with
// a two- or three-step variant (BatchMember::fused = steps per launch) of 256
// threads and at most `items` vectors per thread.  A member taking nsteps
// steps runs nsteps / F multi-step launches (phi and pi flip once each), then
// nsteps % F launches of its one-step companion tile (as a lone run takes
// the leftover steps), and is skipped after that.  Energies as
// launch_batch_step, at the member's energy step.  A member spans
// max(ntiles, ntiles1) work items.  fuses: 2 or 3 when every member has
// that F (a leaner kernel), 0 for a mix.
hipError_t launch_batch_multi(int fp32, int energy, int split, int items, int fuses, const BatchMember *table,
                              int nmem, int total_tiles, int grid, size_t shm, int L, hipStream_t st);
hipError_t launch_batch_finalize_multi(const BatchMember *table, int nmem, int L, hipStream_t st);
// launches a member with F steps per launch takes for n steps, and the index
// of the last step that its launch L covers
inline int multi_launches(int F, int n) { return n / F + n % F; }
inline int multi_last_step(int F, int n, int L) { return L < n / F ? F * L + F - 1 : F * (n / F) + (L - n / F); }
size_t tiled_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c);
TiledGeom tiled_geom(const SimConst &k, const TiledCfg &c);

// Two steps per launch (cb_step2.hip): static box only; pi must ping-pong.
bool step2_valid(const SimConst &k, int fp32, const TiledCfg &c);
int step2_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap);
size_t step2_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c);
hipError_t launch_step2(const SimConst &k, int fp32, int energy, const TiledCfg &c, const StepPtrs &p,
                        const DevScalars *sc, double *partials, hipStream_t st, int *nblocks, int nchunks = 0);

// Three steps per launch (cb_step3.hip): static box only; pi must ping-pong.
// nchunks > 0: only the tiles of the first nchunks chunks of NI planes (one
// GPU's slab of a run sharded along the slowest axis); the same for two.
bool step3_valid(const SimConst &k, int fp32, const TiledCfg &c);
int step3_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap);
size_t step3_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c);
hipError_t launch_step3(const SimConst &k, int fp32, int energy, const TiledCfg &c, const StepPtrs &p,
                        const DevScalars *sc, double *partials, hipStream_t st, int *nblocks, int nchunks = 0);

// Registry of chosen configurations (the autotuner writes, the step reads).
int tiled_register(const TiledCfg &c);
TiledCfg tiled_get(int which);

} // namespace cb
