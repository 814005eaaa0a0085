// cb_slab.cpp -- measurement helper: one GPU's share of a run sharded into
// slabs along the slowest axis i (the layout is i-major), at its real shape.
//
// A rank holding planes [0, S) of an N^3 box advances them with the same
// multi-step tile as a whole box, reading F halo planes on each face that
// its neighbours sent (F = steps per launch).  Here the slab is the first S
// planes of a simulation's own lattice: its halo planes are the lattice's
// planes [N - F, N) and [S, S + F), which the kernel reads through its
// periodic plane index exactly where a rank would read received planes.
// The exchange itself (F planes of phi and of pi per face per launch) is
// stood in for by device-to-device copies into those halo planes, timed
// separately: they show the bytes, not the link.
#include <string>

#include "../../include/coolbubble.h"
#include "cb_sim.hpp"
#include "cb_tiled.hpp"

using cb::set_error;

extern "C" cb_status cb_slab_probe(cb_sim *s, int planes, int launches, double *ms_launch, double *ms_exchange)
{
    if (!s || !ms_launch || !ms_exchange || launches < 1) { set_error("invalid argument"); return CB_INVALID_ARG; }
    if (s->variant < 0 || s->lagged || s->p.D != 3) {
        set_error("slab probe: needs a static 3D lattice on a multi-step tiled variant");
        return CB_UNSUPPORTED;
    }
    const cb::TiledCfg c = s->variants[(size_t)s->variant];
    const int N = s->p.nGrids, F = c.fuse;
    if (F < 2) { set_error("slab probe: needs a multi-step variant"); return CB_UNSUPPORTED; }
    if (planes < c.NI || planes < F || planes % c.NI != 0 || planes + 2 * F > N) {
        set_error("slab probe: planes must be a multiple of the variant's chunk (" + std::to_string(c.NI) +
                  ") and leave room for the halo planes");
        return CB_INVALID_ARG;
    }
    const int nchunks = planes / c.NI;
    const size_t plane_bytes = (size_t)N * N * s->esz;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventCreate(&e2);
    float ms_l = 0.f, ms_x = 0.f;
    auto one_launch = [&]() -> hipError_t {
        cb::StepPtrs p;
        p.phi_in = s->phi[s->iphi];
        p.phi_out = s->phi[s->iphi ^ 1];
        p.pi_in = s->pi[s->ipi];
        p.pi_out = s->pi[s->ipi ^ 1];
        int nb = 0;
        const hipError_t r =
            F >= 3 ? cb::launch_step3(s->k, s->fp32, 0, c, p, s->dsc, s->partials, s->st, &nb, nchunks)
                   : cb::launch_step2(s->k, s->fp32, 0, c, p, s->dsc, s->partials, s->st, &nb, nchunks);
        s->iphi ^= 1;
        s->ipi ^= 1;
        return r;
    };
    auto exchange = [&]() -> hipError_t {
        // F planes per face into the halo slots of the fields the next launch reads
        hipError_t r = hipSuccess;
        for (void *buf : {s->phi[s->iphi], s->pi[s->ipi]}) {
            char *b = static_cast<char *>(buf);
            if (r == hipSuccess)   // below the slab: planes N-F .. N-1
                r = hipMemcpyAsync(b + (size_t)(N - F) * plane_bytes, b + (size_t)(planes - F) * plane_bytes,
                                   (size_t)F * plane_bytes, hipMemcpyDeviceToDevice, s->st);
            if (r == hipSuccess)   // above it: planes S .. S+F-1
                r = hipMemcpyAsync(b + (size_t)planes * plane_bytes, b, (size_t)F * plane_bytes,
                                   hipMemcpyDeviceToDevice, s->st);
        }
        return r;
    };
    if (e == hipSuccess) e = one_launch();   // warm
    if (e == hipSuccess) e = exchange();
    if (e == hipSuccess) e = hipEventRecord(e0, s->st);
    for (int i = 0; i < launches && e == hipSuccess; ++i) e = one_launch();
    if (e == hipSuccess) e = hipEventRecord(e1, s->st);
    for (int i = 0; i < launches && e == hipSuccess; ++i) e = exchange();
    if (e == hipSuccess) e = hipEventRecord(e2, s->st);
    if (e == hipSuccess) e = hipEventSynchronize(e2);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms_l, e0, e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms_x, e1, e2);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e2) hipEventDestroy(e2);
    if (e != hipSuccess) { set_error(std::string("slab probe: ") + hipGetErrorString(e)); return CB_HIP_ERROR; }
    *ms_launch = ms_l / launches;
    *ms_exchange = ms_x / launches;
    return CB_OK;
}
