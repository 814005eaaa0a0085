// cb_sim.hpp -- the simulation handle behind the C ABI (internal).
#pragma once

#include <string>
#include <vector>

#include "../../include/coolbubble.h"
#include "cb_kernels.hpp"
#include "cb_tiled.hpp"

namespace cb {

void set_error(const std::string &msg);

struct FftPlan;   // cb_fft.hip

} // namespace cb

struct cb_sim {
    cb_params p{};
    cb::SimConst k{};
    bool fp32 = false;
    bool lagged = false;        // expanding background: lagged-drift step
    size_t esz = 8;             // bytes per stored value
    long long nsites = 0;
    int nkmax = 0;              // src/bubble.hpp:49
    double dxD = 0, phiC = 0, Rn = 0, lw = 0, mT = 0, phi0 = 1.0;
    hipStream_t st = nullptr;
    bool own_stream = false;    // false when the stream belongs to a batch / server
    // non-null: device memory comes from the device's default memory pool,
    // allocated and freed in order on this stream (a server's requests)
    hipStream_t pool_st = nullptr;
    bool batch_owned = false;   // member of a cb_batch (destroyed with it)
    void *phi[2] = {nullptr, nullptr};
    void *pi[2] = {nullptr, nullptr};
    int iphi = 0, ipi = 0;
    double *partials = nullptr;
    double *dtot = nullptr;
    cb::DevScalars *dsc = nullptr;
    cb::SimpleCfg cfg{256, 1};
    int tiled = -1;             // registry index of the tiled config in use, -1 = general kernel
    int variant = -1;           // index into `variants`, -1 = general kernel
    std::vector<cb::TiledCfg> variants;   // tiled configurations valid for this lattice
    int max_blocks = 0;         // capacity of `partials`, in blocks
    int single_variant = -1;    // one-step variant used for odd steps when `variant` fuses two
    int tiled_energy = -1;      // registry index of the energy launches of a multi-step variant
    int tiled_pair = -1;        // registry index of the two-step tile a 1024-thread three-step variant
                                // ends a call with (-1: none)
    bool pair_energy = false;   // that two-step tile also measures the energies (one vector per thread)
    bool pair_dma = false;      // that two-step tile is the direct-load one (cheaper than a one-step launch)
    double t = 0.0;
    long long step = 0;
    std::vector<double> xns;    // centres (the __xns.dat rows)
    std::string xns_text;       // mode 1: the __xns.dat text as the reference writes it
    cb::FftPlan *fft = nullptr;
    bool keep_pi0 = false;      // keep pi at t = 0 (before the back-step) in pi0
    void *pi0 = nullptr;
    // ctor-time energy row (spreads are zero there, src/bubble.cpp:44)
    bool row_valid = false;
};

namespace cb {
// spectrum of the current phi (cb_fft.hip); bins filled with the reference's
// non-empty-shell convention
// nsteps leapfrog steps.  multi1d: a 1D simulation takes all but the last
// step in launches of several steps, which flip its buffers once per launch;
// a caller that replays captured steps and expects one flip of phi per step
// and pi in place (the batch's general members) passes false.
cb_status step_n(cb_sim *s, int nsteps, bool multi1d = true);
// cb_create, optionally keeping pi at t = 0 (the initial snapshot's kinetic map)
cb_status create_sim(const cb_params *p, const double *xns, int n_xns, bool keep_pi0, cb_sim **out);
// apply a lagged drift still pending on phi (expanding runs)
cb_status materialise_drift(cb_sim *s);
// device memory of s: from the memory pool in order on s->pool_st when that
// is set (after one trim of the pool's free memory if the first try fails),
// else hipMalloc / hipFree
hipError_t sim_alloc(const cb_sim *s, void **p, size_t bytes);
void sim_free(const cb_sim *s, void *p);
// the device scalars of s (synchronises its stream)
cb_status sim_scalars(cb_sim *s, DevScalars &h);
// the 16 outEP columns of s at its current step from its device scalars h
// (read back by the caller); false when h flags a non-finite energy
bool row_from_scalars(const cb_sim *s, const DevScalars &h, double col[16]);
cb_status construct_sim(cb_sim *s, const double *xns, int n_xns);
void release_sim(cb_sim *s);
bool params_ok(const cb_params *p, std::string &why);
// default batched launch shape: at most 2 vectors per thread, prefetch 1,
// non-temporal stores, chunks near 32 planes; batch_variant is a member's
// tile for it (the server's choice)
constexpr int kBatchItems = 2;
int batch_variant(const cb_sim *s);
int pick_variant(const cb_sim *s, int items_max, int pf, int nt, int ni_pref);
// a static member's two- or three-step tile (at most max_fuse steps) for
// the batched multi-step launch, chunks nearest ni3 planes in 3D, ni2 in
// 2D, or -1
int pick_variant_multi(const cb_sim *s, int max_fuse, int ni3 = 16, int ni2 = 8);
// the variant the server gives a request: a lone run's for the large
// static lattices that step on their own (server_solo), else a two-step
// tile from pick_variant_multi when there is one, else batch_variant
bool server_solo(const cb_sim *s);
int server_variant(const cb_sim *s);
int companion_variant(const cb_sim *s, const TiledCfg &f);
// table entry of a member with its selected tiled variant (one- or two-step)
void fill_member(BatchMember &m, const cb_sim *s, int tile_begin, int fp32, size_t *shm);
// work items a member spans in a batched launch
int member_work(const BatchMember &m);
// vectors per thread a batched kernel needs for a member: its variant's and,
// for two steps per launch, its one-step companion's (a kernel with more
// steps every tile of it the same way, but holds more registers)
int member_items(const cb_sim *s);
// host side of n steps run by batched launches (one-step / two-step members)
void advance_host(cb_sim *s, int n);
void advance_host_multi(cb_sim *s, int n);
cb_status select_variant(cb_sim *s, int variant);
// the variant cb_create selects (-1 = the general kernel)
int default_variant(const cb_sim *s);
cb_status spectrum(cb_sim *s, double *kmean, double *kphys, double *delta, int cap, int *nbins);
// the same spectrum with the work arrays in the static no-allocation scratch
cb_status spectrum_noalloc(cb_sim *s, double *kmean, double *kphys, double *delta, int cap, int *nbins);
// thermal initial condition of one field (cb_fft.hip)
cb_status thermal_field(cb_sim *s, void *field, double T, double m, int velocity);
void fft_destroy(FftPlan *p);
} // namespace cb
