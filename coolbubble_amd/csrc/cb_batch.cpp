// cb_batch.cpp -- ensembles of independent simulations stepped together.
//
// Members whose lattice suits the tiled kernel are stepped by one persistent
// launch per step (k_batch_step) over the concatenated tiles of all of them;
// the others (1D, N not a multiple of the vector width) by their own general
// kernel launches on the same stream.  Energies of every tiled member are
// finalised by one launch.  A run of kSegment steps (energies on its last
// step, or every step when a member has an expanding background) is captured
// once into a HIP graph and replayed; kSegment is even, so every member's
// ping-pong buffers are back in place after a replay and the captured
// kernel arguments stay valid.
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "cb_sim.hpp"

using cb::set_error;

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                \
            return CB_HIP_ERROR;                                                         \
        }                                                                                \
    } while (0)

namespace {
constexpr int kSegment = 40;   // OutEnergyEvery, src/def.h:65
}

struct cb_batch {
    std::vector<cb_sim *> mem;
    std::vector<int> tiled;        // member indices on the batched kernel, table order
    std::vector<int> general;      // member indices on their own general kernel
    hipStream_t st = nullptr;
    int fp32 = 0;
    bool any_lagged = false;
    cb::BatchMember *h_tab = nullptr;   // pinned host copy
    cb::BatchMember *d_tab = nullptr;
    int total_tiles = 0, grid = 0;
    size_t shm = 0;
    long long sites = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void batch_free(cb_batch *b)
{
    if (!b) return;
    if (b->st) hipStreamSynchronize(b->st);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (cb_sim *s : b->mem) {
        if (!s) continue;
        s->batch_owned = false;
        cb::release_sim(s);
    }
    if (b->d_tab) hipFree(b->d_tab);
    if (b->h_tab) hipHostFree(b->h_tab);
    if (b->st) hipStreamDestroy(b->st);
    delete b;
}

// Refresh the parities of the table from the members' current buffers.
cb_status upload_table(cb_batch *b)
{
    for (size_t t = 0; t < b->tiled.size(); ++t) {
        const cb_sim *s = b->mem[(size_t)b->tiled[t]];
        b->h_tab[t].par_phi = s->iphi;
        b->h_tab[t].par_pi = s->ipi;
    }
    if (!b->tiled.empty())
        HIP_TRY(hipMemcpyAsync(b->d_tab, b->h_tab, sizeof(cb::BatchMember) * b->tiled.size(), hipMemcpyHostToDevice,
                               b->st));
    return CB_OK;
}

// Enqueue n steps of every member (energies on the last, or on all when a
// member expands).  Host-side member state (buffer parities, t, step) is
// advanced as if the launches ran.
cb_status enqueue_steps(cb_batch *b, int n)
{
    for (int it = 0; it < n; ++it) {
        const bool energy = b->any_lagged || it == n - 1;
        if (!b->tiled.empty()) {
            HIP_TRY(cb::launch_batch_step(b->fp32, energy, b->d_tab, (int)b->tiled.size(), b->total_tiles, b->grid,
                                          b->shm, it & 1, b->st));
            if (energy) HIP_TRY(cb::launch_batch_finalize(b->d_tab, (int)b->tiled.size(), b->st));
        }
    }
    for (int i : b->general) {
        cb_status r = cb::step_n(b->mem[(size_t)i], n);
        if (r != CB_OK) return r;
    }
    for (int i : b->tiled) {
        cb_sim *s = b->mem[(size_t)i];
        if (n & 1) {
            s->iphi ^= 1;
            if (s->lagged) s->ipi ^= 1;
        }
        for (int it = 0; it < n; ++it) s->t += s->p.dt;
        s->step += n;
    }
    return CB_OK;
}

cb_status capture_segment(cb_batch *b)
{
    // host state is advanced by enqueue_steps during capture; put it back
    struct Saved { int iphi, ipi; double t; long long step; };
    std::vector<Saved> saved;
    for (cb_sim *s : b->mem) saved.push_back({s->iphi, s->ipi, s->t, s->step});
    HIP_TRY(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    cb_status r = enqueue_steps(b, kSegment);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->st, &g);
    for (size_t i = 0; i < b->mem.size(); ++i) {
        cb_sim *s = b->mem[i];
        s->iphi = saved[i].iphi;
        s->ipi = saved[i].ipi;
        s->t = saved[i].t;
        s->step = saved[i].step;
    }
    if (r != CB_OK) {
        if (g) hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) { set_error(std::string("graph capture: ") + hipGetErrorString(e)); return CB_HIP_ERROR; }
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return CB_OK;
}

} // namespace

extern "C" {

cb_status cb_batch_create(const cb_params *members, int n, cb_batch **out)
{
    if (!out || !members || n < 1) { set_error("invalid argument"); return CB_INVALID_ARG; }
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        std::string why;
        if (!cb::params_ok(&members[i], why)) { set_error("member " + std::to_string(i) + ": " + why); return CB_INVALID_ARG; }
        cb_status r = cb_check_stability(&members[i]);
        if (r != CB_OK) return r;
        if (members[i].precision != members[0].precision || members[i].device != members[0].device) {
            set_error("batch members must share precision and device");
            return CB_INVALID_ARG;
        }
    }
    cb_batch *b = new (std::nothrow) cb_batch();
    if (!b) { set_error("host allocation failed"); return CB_OUT_OF_MEMORY; }
    b->fp32 = members[0].precision == CB_FP32;
    if (hipSetDevice(members[0].device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        set_error("no HIP device");
        return CB_NO_DEVICE;
    }
    for (int i = 0; i < n; ++i) {
        cb_sim *s = new (std::nothrow) cb_sim();
        if (!s) { batch_free(b); set_error("host allocation failed"); return CB_OUT_OF_MEMORY; }
        s->p = members[i];
        s->st = b->st;
        s->own_stream = false;
        b->mem.push_back(s);
        cb_status r = cb::construct_sim(s, nullptr, 0);
        if (r != CB_OK) { batch_free(b); return r; }
        s->batch_owned = true;
        b->sites += s->nsites;
        if (s->lagged) b->any_lagged = true;
        // the batched kernel's shape: 256 threads, <= 4 vectors per thread, prefetch 2
        int pick = -1;
        for (size_t v = 0; v < s->variants.size(); ++v) {
            const cb::TiledCfg &c = s->variants[v];
            if (c.nthr == 256 && c.items <= 4 && c.pf == 2 && c.nt == 0) { pick = (int)v; break; }
        }
        cb::select_variant(s, pick);
        if (pick >= 0) b->tiled.push_back(i);
        else b->general.push_back(i);
    }
    const size_t nt = b->tiled.size();
    if (nt) {
        if (hipHostMalloc((void **)&b->h_tab, sizeof(cb::BatchMember) * nt, hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void **)&b->d_tab, sizeof(cb::BatchMember) * nt) != hipSuccess) {
            batch_free(b);
            set_error("batch table allocation failed");
            return CB_OUT_OF_MEMORY;
        }
        int tiles = 0;
        for (size_t t = 0; t < nt; ++t) {
            cb_sim *s = b->mem[(size_t)b->tiled[t]];
            const cb::TiledCfg c = s->variants[(size_t)s->variant];
            cb::BatchMember &m = b->h_tab[t];
            std::memset(&m, 0, sizeof m);
            m.k = s->k;
            m.g = cb::tiled_geom(s->k, c);
            m.lagged = s->lagged ? 1 : 0;
            m.tile_begin = tiles;
            m.ntiles = cb::tiled_blocks(s->k, c);
            m.phi[0] = s->phi[0];
            m.phi[1] = s->phi[1];
            m.pi[0] = s->pi[0];
            m.pi[1] = s->lagged ? s->pi[1] : s->pi[0];
            m.sc = s->dsc;
            m.partials = s->partials;
            tiles += m.ntiles;
            const size_t shm = cb::tiled_lds_bytes(s->k, b->fp32, c);
            if (shm > b->shm) b->shm = shm;
        }
        b->total_tiles = tiles;
        int dev = 0, cus = 256;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const int cap = cus * 8;
        b->grid = tiles < cap ? tiles : cap;
        cb_status r = upload_table(b);
        if (r != CB_OK) { batch_free(b); return r; }
    }
    if (hipStreamSynchronize(b->st) != hipSuccess) { batch_free(b); set_error("batch setup failed"); return CB_HIP_ERROR; }
    *out = b;
    return CB_OK;
}

cb_status cb_batch_step(cb_batch *b, int nsteps)
{
    if (!b || nsteps < 0) { set_error("invalid argument"); return CB_INVALID_ARG; }
    cb_status r = upload_table(b);
    if (r != CB_OK) return r;
    int left = nsteps;
    if (left >= kSegment && !b->exec) {
        r = capture_segment(b);
        if (r != CB_OK) return r;
    }
    while (left >= kSegment) {
        HIP_TRY(hipGraphLaunch(b->exec, b->st));
        for (cb_sim *s : b->mem) {
            for (int it = 0; it < kSegment; ++it) s->t += s->p.dt;
            s->step += kSegment;
        }
        left -= kSegment;
    }
    if (left > 0) return enqueue_steps(b, left);
    return CB_OK;
}

cb_status cb_batch_member(cb_batch *b, int i, cb_sim **member)
{
    if (!b || !member || i < 0 || i >= (int)b->mem.size()) { set_error("invalid argument"); return CB_INVALID_ARG; }
    *member = b->mem[(size_t)i];
    return CB_OK;
}

cb_status cb_batch_size(cb_batch *b, int *n, long long *sites)
{
    if (!b) { set_error("invalid handle"); return CB_INVALID_ARG; }
    if (n) *n = (int)b->mem.size();
    if (sites) *sites = b->sites;
    return CB_OK;
}

cb_status cb_batch_stream(cb_batch *b, void **stream)
{
    if (!b || !stream) { set_error("invalid argument"); return CB_INVALID_ARG; }
    *stream = (void *)b->st;
    return CB_OK;
}

cb_status cb_batch_destroy(cb_batch *b)
{
    batch_free(b);
    return CB_OK;
}

} // extern "C"
