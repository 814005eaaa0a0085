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
    // two-step members (launch_batch_step2): g / ntiles are the two-step
    // tile's, g1 / ntiles1 the one-step companion's (odd last step)
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
// Batched two-step launch number pit of a sequence (static members with a
// two-step variant of 256 threads and at most `items` vectors per thread):
// it covers steps 2 pit and 2 pit + 1.  A member with both steps left runs
// its two-step tile (phi and pi flip once per launch); one with only step
// 2 pit left runs its one-step companion tile (the odd last step, as a lone
// run takes it); the others are skipped.  Energies as launch_batch_step, at
// the member's energy step.  A member spans max(ntiles, ntiles1) work items.
hipError_t launch_batch_step2(int fp32, int energy, int split, int items, const BatchMember *table, int nmem,
                              int total_tiles, int grid, size_t shm, int pit, hipStream_t st);
hipError_t launch_batch_finalize2(const BatchMember *table, int nmem, int pit, hipStream_t st);
size_t tiled_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c);
TiledGeom tiled_geom(const SimConst &k, const TiledCfg &c);

// Two steps per launch (cb_step2.hip): static box only; pi must ping-pong.
bool step2_valid(const SimConst &k, int fp32, const TiledCfg &c);
int step2_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap);
size_t step2_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c);
hipError_t launch_step2(const SimConst &k, int fp32, int energy, const TiledCfg &c, const StepPtrs &p,
                        const DevScalars *sc, double *partials, hipStream_t st, int *nblocks);

// Three steps per launch (cb_step3.hip): static box only; pi must ping-pong.
bool step3_valid(const SimConst &k, int fp32, const TiledCfg &c);
int step3_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap);
size_t step3_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c);
hipError_t launch_step3(const SimConst &k, int fp32, int energy, const TiledCfg &c, const StepPtrs &p,
                        const DevScalars *sc, double *partials, hipStream_t st, int *nblocks);

// Registry of chosen configurations (the autotuner writes, the step reads).
int tiled_register(const TiledCfg &c);
TiledCfg tiled_get(int which);

} // namespace cb
