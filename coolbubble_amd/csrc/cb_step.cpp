// cb_step.cpp -- the time loop of one simulation (Bubble::oneStep,
// src/bubble.cpp:849-876, called nt times by src/main.cpp:274-276).
//
// Static box: one fused kernel per step; the energy reduction (and the
// scalar chain) runs only on the last step of a cb_step call, because in a
// static box a == 1 exactly and the energies feed nothing back
// (src/bubble.cpp:747-757 returns 0).  The kinetic energy of the previous
// half step, which the time-centred Ekin needs, is recomputed from the
// incoming pi in the same pass.
// Expanding box: every step runs the lagged-drift kernel and the
// finalisation (energies -> a'' -> a -> drift factor), all on the device.
#include "cb_sim.hpp"
#include "cb_tiled.hpp"

using cb::set_error;

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                \
            return CB_HIP_ERROR;                                                         \
        }                                                                                \
    } while (0)

namespace cb {

static hipError_t launch_one(cb_sim *s, bool energy)
{
    StepPtrs p;
    p.phi_in = s->phi[s->iphi];
    p.phi_out = s->phi[s->iphi ^ 1];
    p.pi_in = s->pi[s->ipi];
    p.pi_out = s->lagged ? s->pi[s->ipi ^ 1] : s->pi[s->ipi];
    int nblocks = s->cfg.grid;
    hipError_t e;
    if (s->tiled >= 0) {
        e = launch_step_tiled(s->k, s->fp32, s->lagged, energy, s->tiled, p, s->dsc, s->partials, s->st,
                              &nblocks);
    } else {
        e = launch_step_simple(s->k, s->fp32, s->lagged, energy, p, s->dsc, s->partials, s->cfg, s->st);
    }
    if (e != hipSuccess) return e;
    if (energy) e = launch_finalize(s->k, s->partials, nblocks, s->dsc, s->lagged ? 0 : 1, s->st);
    return e;
}

cb_status step_n(cb_sim *s, int nsteps)
{
    for (int it = 0; it < nsteps; ++it) {
        const bool energy = s->lagged || it == nsteps - 1;
        HIP_TRY(launch_one(s, energy));
        s->iphi ^= 1;
        if (s->lagged) s->ipi ^= 1;
        s->t += s->p.dt;
        s->step += 1;
    }
    return CB_OK;
}

} // namespace cb
