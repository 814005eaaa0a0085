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
#include <algorithm>
#include <cmath>
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

namespace cb {

// The tiled variant a member uses inside a batched launch of shape
// (items_max, pf, nt): 256 threads, at most items_max vectors per thread;
// among those the tallest tile (most rows per block), then the chunk length
// closest to ni_pref.  -1 if the lattice has no tiled shape.
int pick_variant(const cb_sim *s, int items_max, int pf, int nt, int ni_pref)
{
    int best = -1;
    double best_score = -1e300;
    for (size_t v = 0; v < s->variants.size(); ++v) {
        const TiledCfg &c = s->variants[v];
        if (c.fuse != 1 || c.nthr != 256 || c.items > items_max || c.pf != pf || c.nt != nt) continue;
        const double score = 1000.0 * c.TR - std::fabs(std::log2((double)c.NI / ni_pref));
        if (score > best_score) { best_score = score; best = (int)v; }
    }
    return best;
}

int batch_variant(const cb_sim *s) { return pick_variant(s, kBatchItems, 1, 1, 32); }

// Multi-step tile of a static member for the batched multi-step launch:
// 256 threads, at most kBatchItems vectors per thread, prefetch 1, its
// one-step companion fitting the batched launches too.  Three steps per
// launch where such a tile of 2 rows or more exists (at most max_fuse),
// else two; the tallest tile of at most 4 rows, chunks nearest ni3 planes
// (ni2 in 2D; the server's requests 16 and 8, an ensemble's its autotuned
// length), non-temporal stores preferred.  An ensemble of 64 x 3D N = 128
// and 16 x 2D N = 1024 members steps 1000 times in 249 ms with three steps
// per launch, 335 ms with two.  -1 if there is none.
int pick_variant_multi(const cb_sim *s, int max_fuse, int ni3, int ni2)
{
    if (s->lagged || s->p.D < 2) return -1;
    const double ni_pref = s->p.D == 3 ? (double)ni3 : (double)ni2;
    for (int fuse : {3, 2}) {
        if (fuse > max_fuse) continue;
        int best = -1;
        double best_score = -1e300;
        for (size_t v = 0; v < s->variants.size(); ++v) {
            const TiledCfg &c = s->variants[v];
            if (c.fuse != fuse || c.nthr != 256 || c.pf != 1 || c.items > kBatchItems || c.TR > 4) continue;
            if (fuse == 3 && s->p.D == 3 && c.TR < 2) continue;   // overlap of 3 rows per owned row
            const int comp = companion_variant(s, c);
            if (comp < 0) continue;
            const TiledCfg &c1 = s->variants[(size_t)comp];
            if (c1.nthr != 256 || c1.items > kBatchItems || c1.pf != 1) continue;
            const double score = 1000.0 * c.TR - std::fabs(std::log2(c.NI / ni_pref)) + (c.nt ? 0.5 : 0.0);
            if (score > best_score) { best_score = score; best = (int)v; }
        }
        if (best >= 0) return best;
    }
    return -1;
}

// Static lattices of 2^23 sites or more (3D N >= 208), and those whose rows
// no batched tile holds, step on their own kernels, with the variant a lone
// run starts on (the three-step tiles of
// 512-1024 threads with the plane prefetch): the batched launch's
// 256-thread tiles run them 1.3x slower or more (8 x 3D N = 256: 0.234 ms
// per step batched, 0.175 ms as lone runs).  The rest take two steps per
// batched launch: with three, a serving mix of 200 Poisson-arriving
// requests of every size ran at p50 2.6 s against 1.8 s; a request's short
// runs to its next event leave steps over for the one-step companion.
bool server_solo(const cb_sim *s)
{
    if (s->lagged || s->p.D < 2) return false;
    if (s->nsites >= (1LL << 23)) return true;
    // rows too wide for a batched tile (e.g. FP64 2D N = 2048): a lone
    // run's tiles rather than the general kernel
    return pick_variant_multi(s, 2) < 0 && batch_variant(s) < 0 && default_variant(s) >= 0;
}

int server_variant(const cb_sim *s)
{
    if (server_solo(s)) return default_variant(s);
    const int vm = pick_variant_multi(s, 2);
    return vm >= 0 ? vm : batch_variant(s);
}

// Table entry of a member stepping with its selected tiled variant.
void fill_member(BatchMember &m, const cb_sim *s, int tile_begin, int fp32, size_t *shm)
{
    const TiledCfg c = s->variants[(size_t)s->variant];
    std::memset(&m, 0, sizeof m);
    m.k = s->k;
    m.g = tiled_geom(s->k, c);
    m.lagged = s->lagged ? 1 : 0;
    m.tile_begin = tile_begin;
    m.ntiles = tiled_blocks(s->k, c);
    m.phi[0] = s->phi[0];
    m.phi[1] = s->phi[1];
    m.pi[0] = s->pi[0];
    m.pi[1] = s->pi[1] ? s->pi[1] : s->pi[0];
    m.sc = s->dsc;
    m.partials = s->partials;
    m.par_phi = s->iphi;
    m.par_pi = s->ipi;
    m.nsteps = 1 << 30;
    m.energy_it = -1;
    size_t b = 0;
    if (c.fuse >= 2) {
        const TiledCfg c1 = s->variants[(size_t)s->single_variant];
        m.fused = c.fuse;
        m.g1 = tiled_geom(s->k, c1);
        m.ntiles1 = tiled_blocks(s->k, c1);
        const size_t bm = c.fuse == 3 ? step3_lds_bytes(s->k, fp32, c) : step2_lds_bytes(s->k, fp32, c);
        b = std::max(bm, tiled_lds_bytes(s->k, fp32, c1));
    } else {
        b = tiled_lds_bytes(s->k, fp32, c);
    }
    if (shm && b > *shm) *shm = b;
}

int member_work(const BatchMember &m) { return m.fused ? std::max(m.ntiles, m.ntiles1) : m.ntiles; }

int member_items(const cb_sim *s)
{
    if (s->variant < 0) return 1;
    const TiledCfg &c = s->variants[(size_t)s->variant];
    int it = c.items;
    if (c.fuse >= 2 && s->single_variant >= 0) it = std::max(it, s->variants[(size_t)s->single_variant].items);
    return it;
}

// Host side of n steps run by batched launches: buffer parities, t, step.
void advance_host(cb_sim *s, int n)
{
    if (n & 1) {
        s->iphi ^= 1;
        if (s->lagged) s->ipi ^= 1;
    }
    for (int it = 0; it < n; ++it) s->t += s->p.dt;
    s->step += n;
}

// n steps of a multi-step member (F steps per launch): n / F multi-step
// launches (phi and pi flip once each), then n % F one-step launches (phi
// flips).
void advance_host_multi(cb_sim *s, int n)
{
    const int F = s->variants[(size_t)s->variant].fuse;
    const int q = n / F, r = n % F;
    if ((q + r) & 1) s->iphi ^= 1;
    if (q & 1) s->ipi ^= 1;
    for (int it = 0; it < n; ++it) s->t += s->p.dt;
    s->step += n;
}

} // namespace cb

struct cb_batch {
    std::vector<cb_sim *> mem;
    std::vector<int> tiled;        // member indices on the batched one-step kernel, table order
    std::vector<int> paired;       // member indices on the batched multi-step kernel, after them
    std::vector<int> general;      // member indices on their own general kernel
    hipStream_t st = nullptr;
    int fp32 = 0;
    bool any_lagged = false;
    std::vector<cb::BatchMember> h_tab; // host copy (members' entries, parities filled in at upload)
    cb::BatchMember *h_ring = nullptr;  // pinned staging slots of uploads in flight
    hipEvent_t ring_ev[4] = {};         // each slot's copy done
    int ring_next = 0;
    cb::BatchMember *d_tab = nullptr;
    int total_tiles = 0, grid = 0;
    size_t shm = 0;
    int total_tiles2 = 0, grid2 = 0;    // the multi-step part of the table
    size_t shm2 = 0;
    int fused = 0;                      // static members take two or three steps per launch
    long long sites = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // buffer parities of the general-kernel members when the graph was
    // captured (their pointers are baked into the graph)
    std::vector<int> cap_par;
    int items = cb::kBatchItems, pf = 1, nt = 1, ni = 32;   // batched launch shape (cb_batch_autotune)
    int items1 = 1, items2 = 1;         // vectors per thread the members need (one- / two-step kernels)
    int fuses2 = 0;                     // F of every multi-step member, 0 for a mix
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
    if (b->h_ring) hipHostFree(b->h_ring);
    for (hipEvent_t &e : b->ring_ev)
        if (e) hipEventDestroy(e);
    if (b->st) hipStreamDestroy(b->st);
    delete b;
}

// Refresh the parities of the table from the members' current buffers.
// nsteps: length of the launch sequence that follows (the multi-step
// members take its leftover steps on the companion tile, and measure
// energies at its last step).
// The copy is asynchronous: it goes through one of four pinned staging slots,
// which is written again only once its previous copy has run.
cb_status upload_table(cb_batch *b, int nsteps)
{
    const size_t n1 = b->tiled.size(), n = n1 + b->paired.size();
    if (!n) return CB_OK;
    for (size_t t = 0; t < n; ++t) {
        const cb_sim *s = b->mem[(size_t)(t < n1 ? b->tiled[t] : b->paired[t - n1])];
        b->h_tab[t].par_phi = s->iphi;
        b->h_tab[t].par_pi = s->ipi;
        if (t >= n1) {
            b->h_tab[t].nsteps = nsteps;
            b->h_tab[t].energy_it = nsteps - 1;
        }
    }
    const int k = b->ring_next;
    b->ring_next = (k + 1) % 4;
    HIP_TRY(hipEventSynchronize(b->ring_ev[k]));
    cb::BatchMember *slot = b->h_ring + (size_t)k * n;
    std::memcpy(slot, b->h_tab.data(), sizeof(cb::BatchMember) * n);
    HIP_TRY(hipMemcpyAsync(b->d_tab, slot, sizeof(cb::BatchMember) * n, hipMemcpyHostToDevice, b->st));
    HIP_TRY(hipEventRecord(b->ring_ev[k], b->st));
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
            HIP_TRY(cb::launch_batch_step(b->fp32, energy, 0, b->items1, b->pf, b->d_tab, (int)b->tiled.size(),
                                          b->total_tiles, b->grid, b->shm, it, b->st));
            if (energy) HIP_TRY(cb::launch_batch_finalize(b->d_tab, (int)b->tiled.size(), it, b->st));
        }
    }
    // multi-step members: F steps per launch, then the leftover steps on
    // the one-step companion; energies in each member's last launch (the
    // others step plainly beside it)
    if (!b->paired.empty()) {
        const cb::BatchMember *d2 = b->d_tab + b->tiled.size();
        int nl = 0;
        std::vector<int> lastl;
        for (int i : b->paired) {
            const int F = b->mem[(size_t)i]->variants[(size_t)b->mem[(size_t)i]->variant].fuse;
            lastl.push_back(cb::multi_launches(F, n) - 1);
            nl = std::max(nl, lastl.back() + 1);
        }
        for (int L = 0; L < nl; ++L) {
            bool energy = false, plain = false;
            for (int ll : lastl) {
                if (ll == L) energy = true;
                else if (ll > L) plain = true;
            }
            HIP_TRY(cb::launch_batch_multi(b->fp32, energy, energy && plain, b->items2, b->fuses2, d2, (int)b->paired.size(),
                                           b->total_tiles2, b->grid2, b->shm2, L, b->st));
            if (energy) HIP_TRY(cb::launch_batch_finalize_multi(d2, (int)b->paired.size(), L, b->st));
        }
    }
    for (int i : b->general) {
        cb_status r = cb::step_n(b->mem[(size_t)i], n, false);   // one buffer flip per step, as replays assume
        if (r != CB_OK) return r;
    }
    for (int i : b->tiled) cb::advance_host(b->mem[(size_t)i], n);
    for (int i : b->paired) cb::advance_host_multi(b->mem[(size_t)i], n);
    return CB_OK;
}

void drop_graph(cb_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Pick every member's variant for the batch's launch shape and lay out the
// work table (members without a tiled shape step on their general kernel).
cb_status build_table(cb_batch *b)
{
    if (b->st) HIP_TRY(hipStreamSynchronize(b->st));
    drop_graph(b);
    b->tiled.clear();
    b->paired.clear();
    b->general.clear();
    for (size_t i = 0; i < b->mem.size(); ++i) {
        cb_sim *s = b->mem[i];
        const int pick2 = b->fused ? cb::pick_variant_multi(s, 3, b->ni, b->ni) : -1;
        if (pick2 >= 0 && cb::select_variant(s, pick2) == CB_OK) {
            b->paired.push_back((int)i);
            continue;
        }
        const int pick = cb::pick_variant(s, b->items, b->pf, b->nt, b->ni);
        cb::select_variant(s, pick);
        (pick >= 0 ? b->tiled : b->general).push_back((int)i);
    }
    b->items1 = b->items2 = 1;
    for (int i : b->tiled) b->items1 = std::max(b->items1, cb::member_items(b->mem[(size_t)i]));
    for (int i : b->paired) b->items2 = std::max(b->items2, cb::member_items(b->mem[(size_t)i]));
    bool two = false, three = false;
    for (int i : b->paired) (b->mem[(size_t)i]->variants[(size_t)b->mem[(size_t)i]->variant].fuse == 3 ? three : two) = true;
    b->fuses2 = two && three ? 0 : (three ? 3 : 2);
    const size_t n1 = b->tiled.size(), n2 = b->paired.size();
    if (b->d_tab) { hipFree(b->d_tab); b->d_tab = nullptr; }
    if (b->h_ring) { hipHostFree(b->h_ring); b->h_ring = nullptr; }
    b->h_tab.clear();
    b->total_tiles = b->total_tiles2 = 0;
    b->shm = b->shm2 = 0;
    b->grid = b->grid2 = 0;
    if (!n1 && !n2) return CB_OK;
    b->h_tab.assign(n1 + n2, cb::BatchMember{});
    for (hipEvent_t &e : b->ring_ev)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            set_error("batch event creation failed");
            return CB_HIP_ERROR;
        }
    if (hipHostMalloc((void **)&b->h_ring, sizeof(cb::BatchMember) * (n1 + n2) * 4, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&b->d_tab, sizeof(cb::BatchMember) * (n1 + n2)) != hipSuccess) {
        set_error("batch table allocation failed");
        return CB_OUT_OF_MEMORY;
    }
    int tiles = 0;
    for (size_t t = 0; t < n1; ++t) {
        cb::fill_member(b->h_tab[t], b->mem[(size_t)b->tiled[t]], tiles, b->fp32, &b->shm);
        tiles += cb::member_work(b->h_tab[t]);
    }
    b->total_tiles = tiles;
    tiles = 0;
    for (size_t t = 0; t < n2; ++t) {
        cb::fill_member(b->h_tab[n1 + t], b->mem[(size_t)b->paired[t]], tiles, b->fp32, &b->shm2);
        tiles += cb::member_work(b->h_tab[n1 + t]);
    }
    b->total_tiles2 = tiles;
    int dev = 0, cus = 256;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int cap = cus * 8;
    b->grid = b->total_tiles < cap ? b->total_tiles : cap;
    b->grid2 = b->total_tiles2 < cap ? b->total_tiles2 : cap;
    return upload_table(b, kSegment);
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
    b->cap_par.clear();
    for (int i : b->general) b->cap_par.push_back(b->mem[(size_t)i]->iphi * 2 + b->mem[(size_t)i]->ipi);
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
    }
    cb_status r = build_table(b);
    if (r != CB_OK) { batch_free(b); return r; }
    if (hipStreamSynchronize(b->st) != hipSuccess) { batch_free(b); set_error("batch setup failed"); return CB_HIP_ERROR; }
    *out = b;
    return CB_OK;
}

cb_status cb_batch_step(cb_batch *b, int nsteps)
{
    if (!b || nsteps < 0) { set_error("invalid argument"); return CB_INVALID_ARG; }
    cb_status r = upload_table(b, kSegment);
    if (r != CB_OK) return r;
    int left = nsteps;
    if (b->exec) {   // general members moved to other buffers since the capture: capture again
        bool stale = false;
        for (size_t g = 0; g < b->general.size(); ++g) {
            const cb_sim *s = b->mem[(size_t)b->general[g]];
            if (s->iphi * 2 + s->ipi != b->cap_par[g]) stale = true;
        }
        if (stale) {
            HIP_TRY(hipStreamSynchronize(b->st));
            drop_graph(b);
        }
    }
    if (left >= kSegment && !b->exec) {
        r = capture_segment(b);
        if (r != CB_OK) return r;
    }
    while (left >= kSegment) {
        HIP_TRY(hipGraphLaunch(b->exec, b->st));
        // host side of the segment, as enqueue_steps advanced it during the
        // capture (a general member flips its buffers once per step: an even
        // segment leaves them)
        for (int i : b->tiled) cb::advance_host(b->mem[(size_t)i], kSegment);
        for (int i : b->paired) cb::advance_host_multi(b->mem[(size_t)i], kSegment);
        for (int i : b->general) {
            cb_sim *s = b->mem[(size_t)i];
            for (int it = 0; it < kSegment; ++it) s->t += s->p.dt;
            s->step += kSegment;
        }
        left -= kSegment;
        if (left >= kSegment) {   // the parities for the next replay
            r = upload_table(b, kSegment);
            if (r != CB_OK) return r;
        }
    }
    if (left > 0) {
        r = upload_table(b, left);
        if (r != CB_OK) return r;
        return enqueue_steps(b, left);
    }
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

cb_status cb_batch_autotune(cb_batch *b, int steps, int shape[5], double *best_ms)
{
    if (!b || steps < 1) { set_error("invalid argument"); return CB_INVALID_ARG; }
    HIP_TRY(hipStreamSynchronize(b->st));
    // save every member's current fields, scalars and host state
    struct Saved { void *phi, *pi; cb::DevScalars sc; int iphi, ipi; double t; long long step; };
    std::vector<Saved> sv(b->mem.size());
    cb_status r = CB_OK;
    for (size_t i = 0; i < b->mem.size() && r == CB_OK; ++i) {
        cb_sim *s = b->mem[i];
        const size_t bytes = (size_t)s->nsites * s->esz;
        Saved &q = sv[i];
        q = Saved{nullptr, nullptr, {}, s->iphi, s->ipi, s->t, s->step};
        if (hipMalloc(&q.phi, bytes) != hipSuccess || hipMalloc(&q.pi, bytes) != hipSuccess) { r = CB_OUT_OF_MEMORY; break; }
        hipMemcpyAsync(q.phi, s->phi[s->iphi], bytes, hipMemcpyDeviceToDevice, b->st);
        hipMemcpyAsync(q.pi, s->pi[s->ipi], bytes, hipMemcpyDeviceToDevice, b->st);
        hipMemcpyAsync(&q.sc, s->dsc, sizeof q.sc, hipMemcpyDeviceToHost, b->st);
    }
    auto restore = [&]() -> hipError_t {
        for (size_t i = 0; i < b->mem.size(); ++i) {
            cb_sim *s = b->mem[i];
            Saved &q = sv[i];
            const size_t bytes = (size_t)s->nsites * s->esz;
            s->iphi = q.iphi;
            s->ipi = q.ipi;
            s->t = q.t;
            s->step = q.step;
            hipMemcpyAsync(s->phi[s->iphi], q.phi, bytes, hipMemcpyDeviceToDevice, b->st);
            hipMemcpyAsync(s->pi[s->ipi], q.pi, bytes, hipMemcpyDeviceToDevice, b->st);
            hipMemcpyAsync(s->dsc, &q.sc, sizeof q.sc, hipMemcpyHostToDevice, b->st);
        }
        return hipStreamSynchronize(b->st);
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (r == CB_OK && (hipStreamSynchronize(b->st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                       hipEventCreate(&e1) != hipSuccess))
        r = CB_HIP_ERROR;
    // (vectors per thread, prefetch, non-temporal stores, chunk planes, two steps per launch)
    const int shapes[][5] = {{4, 1, 1, 32, 0}, {4, 2, 1, 32, 0}, {4, 1, 0, 32, 0}, {2, 1, 1, 32, 0},
                             {1, 1, 1, 32, 0}, {4, 1, 1, 16, 0}, {4, 1, 1, 64, 0}, {1, 2, 1, 32, 0},
                             {2, 2, 1, 16, 0}, {4, 2, 0, 16, 0}, {1, 1, 1, 32, 1}, {2, 1, 1, 32, 1},
                             {1, 1, 1, 16, 1}, {1, 1, 1, 64, 1}, {2, 1, 1, 64, 1}};
    double best = 1e300;
    int best_i = 0;
    for (int c = 0; c < (int)(sizeof shapes / sizeof shapes[0]) && r == CB_OK; ++c) {
        b->items = shapes[c][0];
        b->pf = shapes[c][1];
        b->nt = shapes[c][2];
        b->ni = shapes[c][3];
        b->fused = shapes[c][4];
        r = build_table(b);
        if (r != CB_OK) break;
        double per = 1e300;
        for (int rep = 0; rep < 2 && r == CB_OK; ++rep) {
            if (restore() != hipSuccess) { r = CB_HIP_ERROR; break; }
            r = upload_table(b, steps);
            if (r != CB_OK) break;
            hipEventRecord(e0, b->st);
            r = enqueue_steps(b, steps);
            hipEventRecord(e1, b->st);
            float ms = 0.f;
            if (r == CB_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                r = CB_HIP_ERROR;
            if (ms / steps < per) per = ms / steps;
        }
        if (per < best) { best = per; best_i = c; }
    }
    if (restore() != hipSuccess && r == CB_OK) r = CB_HIP_ERROR;
    for (Saved &q : sv) {
        if (q.phi) hipFree(q.phi);
        if (q.pi) hipFree(q.pi);
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    b->items = shapes[best_i][0];
    b->pf = shapes[best_i][1];
    b->nt = shapes[best_i][2];
    b->ni = shapes[best_i][3];
    b->fused = shapes[best_i][4];
    cb_status r2 = build_table(b);
    if (r == CB_OK) r = r2;
    if (r != CB_OK) return r;
    if (shape) for (int i = 0; i < 5; ++i) shape[i] = shapes[best_i][i];
    if (best_ms) *best_ms = best;
    return CB_OK;
}

cb_status cb_batch_set_shape(cb_batch *b, const int shape[5])
{
    if (!b || !shape) { set_error("invalid argument"); return CB_INVALID_ARG; }
    const int items = shape[0], pf = shape[1], nt = shape[2], ni = shape[3], fused = shape[4];
    if ((items != 1 && items != 2 && items != 4) || (pf != 1 && pf != 2) || (nt != 0 && nt != 1) || ni < 1 ||
        (fused != 0 && fused != 1)) {
        set_error("batch shape out of range");
        return CB_INVALID_ARG;
    }
    b->items = items;
    b->pf = pf;
    b->nt = nt;
    b->ni = ni;
    b->fused = fused;
    return build_table(b);
}

cb_status cb_batch_destroy(cb_batch *b)
{
    batch_free(b);
    return CB_OK;
}

} // extern "C"
