// cb_kernels.hpp -- host-side launchers of the HIP kernels (C++ linkage,
// internal to libcoolbubble.so).
#pragma once

#include "cb_common.hpp"

namespace cb {

// Field pointers of one step.  Static box: phi ping-pongs, pi is updated in
// place (pi_in == pi_out).  Expanding box (lagged drift): both ping-pong.
struct StepPtrs {
    const void *phi_in;
    void *phi_out;
    const void *pi_in;
    void *pi_out;
};

// Launch configuration of the general step kernel.
struct SimpleCfg {
    int block;      // threads per block
    int grid;       // blocks (grid-stride over sites)
};

// One fused leapfrog step (force + kick + drift, energy partials when
// `energy`), general kernel: one site per thread per iteration.
hipError_t launch_step_simple(const SimConst &k, int fp32, int lagged, int energy,
                              const StepPtrs &p, const DevScalars *sc, double *partials,
                              SimpleCfg cfg, hipStream_t st);

// Up to kMaxSteps1d non-lagged 1D steps in one launch, no energies; blocks
// of kThreads1d threads own kSeg1d sites each.  phi and pi are read from the
// *_in buffers and written to distinct *_out buffers.
constexpr int kSeg1d = 1024, kMaxSteps1d = 32, kThreads1d = 256;
hipError_t launch_step1d_multi(const SimConst &k, int fp32, const StepPtrs &p, const DevScalars *sc, int steps,
                               hipStream_t st);

// Constructor pieces (src/bubble.cpp:239-291).
hipError_t launch_ctor_energy(const SimConst &k, int fp32, const void *phi, const void *pi,
                              double *partials, SimpleCfg cfg, hipStream_t st);
hipError_t launch_backstep(const SimConst &k, int fp32, const void *phi, void *pi, double pk2,
                           double *partials, SimpleCfg cfg, hipStream_t st);

// Sum nblocks x NQ partials into tot[NQ] (deterministic, compensated).
hipError_t launch_sum_partials(const double *partials, int nblocks, double *tot, hipStream_t st);

// Sum partials of one step, then the scalar chain of oneStep
// (src/bubble.cpp:839-875): time-centred kinetic energy, a', a, a_half, H, C,
// and the factors of the next step.  use_old_kin: EkinHalfPrev is taken from
// the Q_KIN_OLD partials (static box) instead of the chained value.
hipError_t launch_finalize(const SimConst &k, const double *partials, int nblocks,
                           DevScalars *sc, int use_old_kin, hipStream_t st);

// phi += (dt * pi) * s_pending; s_pending = 0 (materialise a lagged drift).
hipError_t launch_apply_drift(const SimConst &k, int fp32, void *phi, const void *pi,
                              DevScalars *sc, SimpleCfg cfg, hipStream_t st);

// Bubble seeding: phi = min(phi + profile(c), phiC) per centre in order
// (src/bubble.cpp:324-453).  cen/sub: ncen x 3 doubles each; the distance
// along an axis is (x - cen) - sub.
hipError_t launch_seed_bubbles(const SimConst &k, int fp32, void *phi, const double *cen,
                               const double *sub, int ncen, double L, double phiC, double Rn,
                               double lw, SimpleCfg cfg, hipStream_t st);

// TEST=1 sine wave, src/bubble.cpp:338-353.
hipError_t launch_seed_wave(const SimConst &k, int fp32, void *phi, double L, double kwave,
                            double amplitude, double phi0, SimpleCfg cfg, hipStream_t st);

// Widen / narrow between the storage type and double.
hipError_t launch_convert(int to_double, const void *src, void *dst, long long n, hipStream_t st);

} // namespace cb
