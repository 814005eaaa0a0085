// cb_server.cpp -- continuous-batching request server.
//
// One scheduler thread owns the device work.  Each tick it
//   1. admits queued requests in arrival order while the slot count and the
//      device-memory budget allow (a request larger than the whole budget is
//      rejected), building each one's lattice (Bubble ctor) on the server
//      stream and recording its initial energy row and spectrum;
//   2. gives every active request its own step count: the steps to its next
//      event -- an energy row (every 40 steps, src/main.cpp:280), a snapshot
//      (SnapshotSchedule, src/main.cpp:201-223) or its last step -- at most
//      40;
//   3. runs one batched launch per step over all of their tiles
//      (k_batch_step); a request whose count is used up skips its tiles;
//      energies are computed on the steps where some request stops (every
//      step if one expands) and finalised for the requests still taking
//      part; requests without a tiled shape run their general kernel;
//   4. records rows / spectra of the requests at an event, retires finished
//      ones and frees their memory.
// A request's arithmetic does not depend on its neighbours or on where tick
// boundaries fall (a static box never feeds energies back; an expanding one
// computes them every step), so its results are bit-identical to a lone run
// with the same kernel variant.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "cb_sim.hpp"

using cb::set_error;

namespace {

struct Req {
    unsigned long long id = 0;
    cb_request r{};
    double t_submit = 0, t_admit = 0, t_done = 0;
    cb_sim *sim = nullptr;
    long long footprint = 0;
    std::vector<int> snaps;         // snapshot steps
    size_t next_snap = 0;
    std::vector<double> rows;       // 16 per row
    std::vector<std::vector<double>> spec_k, spec_d;
    double final_row[16] = {0};
    bool cancelled = false;
    cb_status status = CB_OK;
};

// SnapshotSchedule, src/main.cpp:201-223 (SnapshotSpacing = 0)
std::vector<int> snapshot_steps(int nt, double dt, int n)
{
    const double T = nt * dt;
    std::vector<int> steps;
    for (int k = 1; k <= n; ++k) {
        const double time = T * k / n;
        int st = (int)std::lround(time / dt);
        st = st < 1 ? 1 : st;
        st = st > nt ? nt : st;
        if (steps.empty() || st > steps.back()) steps.push_back(st);
    }
    return steps;
}

// Device bytes a request holds while active (fields, partials, spectrum
// work arrays) and transiently while its thermal IC is built.
long long footprint_of(const cb_params &p)
{
    const long long N = p.nGrids;
    long long sites = 1;
    for (int d = 0; d < p.D; ++d) sites *= N;
    const long long esz = p.precision == CB_FP32 ? 4 : 8;
    const bool lagged = p.expansion && p.D == 3;
    // phi and pi double-buffered (expanding box, a multi-step variant, or
    // the 1D launches of several steps)
    (void)lagged;
    const long long fields = sites * esz * 4;
    const long long rows = sites / N, nc = N / 2 + 1;
    const long long ncp = (nc + 7) / 8 * 8;   // padded rows of the fused power-of-two spectrum
    const long long fft = rows * ncp * 2 * esz + N * 2 * esz + 256 * 3 * 2048 * 8;
    const long long partials = (sites / 64 + 8192) * 12 * 8;
    long long transient = 0;
    if (p.readxns == -3 && p.D >= 2) transient = rows * 312 * 8 + rows * nc * 16;
    return fields + partials + (fft > transient ? fft : transient) + (1 << 20);
}

} // namespace

struct cb_server {
    cb_server_opts o{};
    hipStream_t st = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv, idle_cv;
    std::deque<Req *> queue;
    std::vector<Req *> active;               // scheduler thread only
    std::vector<Req *> done;                 // completed, not yet polled
    std::map<unsigned long long, Req *> kept;   // polled, outputs stored until released
    std::map<unsigned long long, bool> cancel_req;   // cancels of admitted requests, taken at their next tick
    std::set<unsigned long long> live;               // admitted and not finished
    long long used = 0, budget = 0;
    bool stop = false;
    int busy = 0;                            // requests admitted and not finished
    unsigned long long next_id = 1;
    std::chrono::steady_clock::time_point t0;
    cb::BatchMember *h_tab = nullptr, *d_tab = nullptr;
    size_t tab_cap = 0;
    cb::DevScalars *h_sc = nullptr;          // pinned: each active request's scalars after a tick
    size_t sc_cap = 0;
    int cus = 256;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

bool read_row(cb_sim *s, double out[16])
{
    cb_energies e;
    cb_status r = cb_energies_row(s, &e);
    std::memcpy(out, e.col, sizeof e.col);
    return r == CB_OK;
}

void record_spectrum(Req *q)
{
    const int cap = (int)(std::sqrt((double)q->r.params.D) * q->r.params.nGrids / 2) + 2;
    std::vector<double> k((size_t)cap), kp((size_t)cap), d((size_t)cap);
    int nb = 0;
    if (cb_power_spectrum(q->sim, k.data(), kp.data(), d.data(), cap, &nb) != CB_OK) nb = 0;
    k.resize((size_t)nb);
    d.resize((size_t)nb);
    q->spec_k.push_back(std::move(k));
    q->spec_d.push_back(std::move(d));
}

void finish(cb_server *sv, Req *q, cb_status st)
{
    q->status = st;
    q->t_done = sv->now();
    if (q->sim) {
        cb::release_sim(q->sim);
        q->sim = nullptr;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->used -= q->footprint;
    sv->busy -= 1;
    sv->live.erase(q->id);
    sv->cancel_req.erase(q->id);
    sv->done.push_back(q);
    sv->idle_cv.notify_all();
}

void admit(cb_server *sv)
{
    for (;;) {
        Req *q = nullptr;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            if (sv->queue.empty()) return;
            Req *head = sv->queue.front();
            if (head->footprint > sv->budget) {          // can never fit
                sv->queue.pop_front();
                sv->busy += 1;
                sv->used += head->footprint;
                q = head;
                q->status = CB_ADMISSION_REJECTED;
            } else if ((sv->o.max_active > 0 && (int)sv->active.size() >= sv->o.max_active) ||
                       sv->used + head->footprint > sv->budget) {
                return;                                  // wait for slots / memory
            } else {
                sv->queue.pop_front();
                sv->used += head->footprint;
                sv->busy += 1;
                sv->live.insert(head->id);
                q = head;
            }
        }
        if (q->status == CB_ADMISSION_REJECTED) {
            finish(sv, q, CB_ADMISSION_REJECTED);
            continue;
        }
        q->t_admit = sv->now();
        cb_sim *s = new (std::nothrow) cb_sim();
        cb_status r = s ? CB_OK : CB_OUT_OF_MEMORY;
        if (s) {
            s->p = q->r.params;
            s->p.precision = sv->o.precision;
            s->p.device = sv->o.device;
            s->st = sv->st;
            s->own_stream = false;
            s->pool_st = sv->st;    // freed requests' memory serves the next ones
            r = cb::construct_sim(s, nullptr, 0);
            if (r != CB_OK) { cb::release_sim(s); s = nullptr; }
        }
        if (r != CB_OK) { finish(sv, q, r); continue; }
        q->sim = s;
        if (cb::select_variant(s, cb::server_variant(s)) != CB_OK) cb::select_variant(s, cb::batch_variant(s));
        double row[16];
        read_row(s, row);
        q->rows.insert(q->rows.end(), row, row + 16);
        record_spectrum(q);
        q->snaps = snapshot_steps(q->r.nt, s->p.dt, q->r.snapshots);
        sv->active.push_back(q);
    }
}

// One tick over the active requests: each advances to its own next event
// (at most 40 steps); the batched launches skip the tiles of requests that
// already stopped.
cb_status tick(cb_server *sv)
{
    std::vector<Req *> &act = sv->active;
    std::vector<int> m(act.size());
    int L = 1;
    bool any_lagged = false;
    for (size_t i = 0; i < act.size(); ++i) {
        Req *q = act[i];
        const long long step = q->sim->step;
        long long ev = q->r.nt;
        const long long next40 = (step / 40 + 1) * 40;
        if (next40 < ev) ev = next40;
        if (q->next_snap < q->snaps.size() && q->snaps[q->next_snap] < ev) ev = q->snaps[q->next_snap];
        long long n = ev - step;
        if (n < 1) n = 1;
        if (n > 40) n = 40;
        m[i] = (int)n;
        if (m[i] > L) L = m[i];
        if (q->sim->lagged) any_lagged = true;
    }
    // batched tables: one-step tiled requests first, then multi-step ones;
    // the general-kernel requests and the large ones step on their own
    std::vector<size_t> single, pair, general;
    for (size_t i = 0; i < act.size(); ++i) {
        const cb_sim *s = act[i]->sim;
        if (s->variant < 0 || cb::server_solo(s)) general.push_back(i);
        else (s->variants[(size_t)s->variant].fuse >= 2 ? pair : single).push_back(i);
    }
    const size_t ntab = single.size() + pair.size();
    if (ntab > sv->tab_cap) {
        if (sv->h_tab) hipHostFree(sv->h_tab);
        if (sv->d_tab) hipFree(sv->d_tab);
        sv->h_tab = nullptr;
        sv->d_tab = nullptr;
        sv->tab_cap = ntab * 2;
        if (hipHostMalloc((void **)&sv->h_tab, sizeof(cb::BatchMember) * sv->tab_cap, hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void **)&sv->d_tab, sizeof(cb::BatchMember) * sv->tab_cap) != hipSuccess) {
            sv->tab_cap = 0;
            set_error("server table allocation failed");
            return CB_OUT_OF_MEMORY;
        }
    }
    if (act.size() > sv->sc_cap) {
        if (sv->h_sc) hipHostFree(sv->h_sc);
        sv->h_sc = nullptr;
        sv->sc_cap = act.size() * 2;
        if (hipHostMalloc((void **)&sv->h_sc, sizeof(cb::DevScalars) * sv->sc_cap, hipHostMallocDefault) != hipSuccess) {
            sv->sc_cap = 0;
            set_error("server scalar staging allocation failed");
            return CB_OUT_OF_MEMORY;
        }
    }
    const int fp32 = sv->o.precision == CB_FP32;
    // one-step table: L launches, energies where a member ends (every step if one expands)
    int tiles1 = 0, L1 = 0;
    size_t shm1 = 0;
    std::vector<char> energy1((size_t)L, any_lagged ? 1 : 0);
    for (size_t t = 0; t < single.size(); ++t) {
        cb::fill_member(sv->h_tab[t], act[single[t]]->sim, tiles1, fp32, &shm1);
        const int n = m[single[t]];
        sv->h_tab[t].nsteps = n;
        sv->h_tab[t].energy_it = n - 1;   // energies at its own event only
        energy1[(size_t)n - 1] = 1;
        tiles1 += cb::member_work(sv->h_tab[t]);
        if (n > L1) L1 = n;
    }
    // multi-step table: a member's launch covers F steps (2 or 3), then its
    // leftover steps run one per launch on its companion tile; energies in
    // its last launch
    cb::BatchMember *tab2 = sv->h_tab + single.size();
    int tiles2 = 0, L2 = 0;     // L2: launches of the longest member
    size_t shm2 = 0;
    std::vector<int> last2(pair.size());
    for (size_t t = 0; t < pair.size(); ++t) {
        cb::fill_member(tab2[t], act[pair[t]]->sim, tiles2, fp32, &shm2);
        const int n = m[pair[t]];
        tab2[t].nsteps = n;
        tab2[t].energy_it = n - 1;
        last2[t] = cb::multi_launches(tab2[t].fused, n) - 1;
        tiles2 += cb::member_work(tab2[t]);
        if (last2[t] + 1 > L2) L2 = last2[t] + 1;
    }
    std::vector<char> energy2((size_t)L2, 0);
    for (size_t t = 0; t < pair.size(); ++t) energy2[(size_t)last2[t]] = 1;
    hipError_t e = hipSuccess;
    if (ntab) e = hipMemcpyAsync(sv->d_tab, sv->h_tab, sizeof(cb::BatchMember) * ntab, hipMemcpyHostToDevice, sv->st);
    // members taking part in each launch without measuring energies (they
    // get a plain launch of their own beside the energy launch)
    std::vector<char> plain1((size_t)L, 0), plain2((size_t)L2, 0);
    for (size_t t = 0; t < single.size(); ++t) {
        if (sv->h_tab[t].lagged) continue;
        for (int it = 0; it < sv->h_tab[t].nsteps - 1; ++it) plain1[(size_t)it] = 1;
    }
    for (size_t t = 0; t < pair.size(); ++t)
        for (int l = 0; l < last2[t]; ++l) plain2[(size_t)l] = 1;
    const int cap = sv->cus * 8;
    // the smallest kernels that step every member (fewer registers)
    int items1 = 1, items2 = 1;
    for (size_t i : single) items1 = std::max(items1, cb::member_items(act[i]->sim));
    for (size_t i : pair) items2 = std::max(items2, cb::member_items(act[i]->sim));
    bool two = false, three = false;    // F of the multi-step requests: one kernel for each, or both
    for (size_t t = 0; t < pair.size(); ++t) (tab2[t].fused == 3 ? three : two) = true;
    const int fuses2 = two && three ? 0 : (three ? 3 : 2);
    for (int it = 0; it < L1 && e == hipSuccess; ++it) {
        const bool energy = energy1[(size_t)it] != 0;
        e = cb::launch_batch_step(fp32, energy, plain1[(size_t)it], items1, 1, sv->d_tab, (int)single.size(),
                                  tiles1, std::min(tiles1, cap), shm1, it, sv->st);
        if (e == hipSuccess && energy) e = cb::launch_batch_finalize(sv->d_tab, (int)single.size(), it, sv->st);
    }
    const cb::BatchMember *d_tab2 = sv->d_tab + single.size();
    for (int l = 0; l < L2 && e == hipSuccess; ++l) {
        const bool energy = energy2[(size_t)l] != 0;
        e = cb::launch_batch_multi(fp32, energy, plain2[(size_t)l], items2, fuses2, d_tab2, (int)pair.size(), tiles2,
                                   std::min(tiles2, cap), shm2, l, sv->st);
        if (e == hipSuccess && energy) e = cb::launch_batch_finalize_multi(d_tab2, (int)pair.size(), l, sv->st);
    }
    if (e != hipSuccess) { set_error(std::string("server tick: ") + hipGetErrorString(e)); return CB_HIP_ERROR; }
    for (size_t i : general) {
        cb_status r = cb::step_n(act[i]->sim, m[i]);
        if (r != CB_OK) return r;
    }
    for (size_t i : single) cb::advance_host(act[i]->sim, m[i]);
    for (size_t i : pair) cb::advance_host_multi(act[i]->sim, m[i]);
    // the scalars of the requests with a row at this step, copied after the
    // tick's launches (one stream) and waited for with them
    for (size_t i = 0; i < act.size(); ++i) {
        const long long step = act[i]->sim->step;
        if (step % 40 != 0 && step != act[i]->r.nt) continue;
        if (hipMemcpyAsync(&sv->h_sc[i], act[i]->sim->dsc, sizeof(cb::DevScalars), hipMemcpyDeviceToHost, sv->st) !=
            hipSuccess) {
            set_error("server tick: scalar copy failed");
            return CB_HIP_ERROR;
        }
    }
    if (hipStreamSynchronize(sv->st) != hipSuccess) { set_error("server tick: stream failed"); return CB_HIP_ERROR; }
    // events
    std::vector<Req *> keep;
    for (size_t i = 0; i < act.size(); ++i) {
        Req *q = act[i];
        const long long step = q->sim->step;
        double row[16];
        bool have = false;
        if (step % 40 == 0 || step == q->r.nt) { cb::row_from_scalars(q->sim, sv->h_sc[i], row); have = true; }
        if (step % 40 == 0) q->rows.insert(q->rows.end(), row, row + 16);
        if (q->next_snap < q->snaps.size() && q->snaps[q->next_snap] == step) {
            record_spectrum(q);
            q->next_snap += 1;
        }
        bool cancel = false;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            auto it = sv->cancel_req.find(q->id);
            if (it != sv->cancel_req.end()) { cancel = true; sv->cancel_req.erase(it); }
        }
        if (step >= q->r.nt || cancel) {
            if (have) std::memcpy(q->final_row, row, sizeof row);
            finish(sv, q, cancel ? CB_NOT_READY : CB_OK);
        } else {
            keep.push_back(q);
        }
    }
    act.swap(keep);
    return CB_OK;
}

void run(cb_server *sv)
{
    hipSetDevice(sv->o.device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv.wait(lk, [&] { return sv->stop || !sv->queue.empty() || !sv->active.empty(); });
            if (sv->stop) return;
        }
        admit(sv);
        if (sv->active.empty()) continue;
        if (tick(sv) != CB_OK) {
            for (Req *q : sv->active) finish(sv, q, CB_HIP_ERROR);
            sv->active.clear();
        }
    }
}

} // namespace

extern "C" {

cb_status cb_server_create(const cb_server_opts *o, cb_server **out)
{
    if (!o || !out) { set_error("invalid argument"); return CB_INVALID_ARG; }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error("no HIP device visible"); return CB_NO_DEVICE; }
    if (o->device < 0 || o->device >= ndev || (o->precision != CB_FP32 && o->precision != CB_FP64)) {
        set_error("invalid device or precision");
        return CB_INVALID_ARG;
    }
    cb_server *sv = new (std::nothrow) cb_server();
    if (!sv) { set_error("host allocation failed"); return CB_OUT_OF_MEMORY; }
    sv->o = *o;
    hipSetDevice(o->device);
    if (hipStreamCreateWithFlags(&sv->st, hipStreamNonBlocking) != hipSuccess) {
        delete sv;
        set_error("stream creation failed");
        return CB_HIP_ERROR;
    }
    size_t freeb = 0, totalb = 0;
    hipMemGetInfo(&freeb, &totalb);
    {   // requests allocate from the device's memory pool in stream order; the
        // pool keeps what they free (a request's construction then costs no
        // hipMalloc, its end no device-wide hipFree)
        hipMemPool_t pool = nullptr;
        unsigned long long keep = ~0ULL;
        if (hipDeviceGetDefaultMemPool(&pool, o->device) == hipSuccess)
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    sv->budget = o->mem_budget > 0 ? o->mem_budget : (long long)(0.9 * (double)freeb);
    hipDeviceGetAttribute(&sv->cus, hipDeviceAttributeMultiprocessorCount, o->device);
    sv->t0 = std::chrono::steady_clock::now();
    sv->th = std::thread(run, sv);
    *out = sv;
    return CB_OK;
}

cb_status cb_server_submit(cb_server *sv, const cb_request *r, unsigned long long *id)
{
    if (!sv || !r) { set_error("invalid argument"); return CB_INVALID_ARG; }
    cb_params p = r->params;
    p.precision = sv->o.precision;
    p.device = sv->o.device;
    std::string why;
    if (!cb::params_ok(&p, why)) { set_error(why); return CB_INVALID_ARG; }
    cb_status st = cb_check_stability(&p);
    if (st != CB_OK) return st;
    if (r->nt < 1 || r->snapshots < 1) { set_error("nt and snapshots must be >= 1"); return CB_INVALID_ARG; }
    Req *q = new (std::nothrow) Req();
    if (!q) { set_error("host allocation failed"); return CB_OUT_OF_MEMORY; }
    q->r = *r;
    q->r.params = p;
    q->footprint = footprint_of(p);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        q->id = sv->next_id++;
        q->t_submit = sv->now();
        sv->queue.push_back(q);
    }
    sv->cv.notify_all();
    if (id) *id = q->id;
    return CB_OK;
}

cb_status cb_server_cancel(cb_server *sv, unsigned long long id)
{
    if (!sv) { set_error("invalid handle"); return CB_INVALID_ARG; }
    std::lock_guard<std::mutex> lk(sv->mu);
    for (auto it = sv->queue.begin(); it != sv->queue.end(); ++it) {
        if ((*it)->id == id) {
            Req *q = *it;
            sv->queue.erase(it);
            q->status = CB_NOT_READY;
            q->t_done = sv->now();
            sv->done.push_back(q);
            return CB_OK;
        }
    }
    if (!sv->live.count(id)) {
        set_error("no queued or active request with this id");
        return CB_INVALID_ARG;
    }
    sv->cancel_req[id] = true;   // active: retired at its next tick
    return CB_OK;
}

cb_status cb_server_poll(cb_server *sv, cb_result *out, int cap, int *n)
{
    if (!sv || !n || (cap > 0 && !out)) { set_error("invalid argument"); return CB_INVALID_ARG; }
    std::lock_guard<std::mutex> lk(sv->mu);
    int k = 0;
    while (k < cap && !sv->done.empty()) {
        Req *q = sv->done.front();
        sv->done.erase(sv->done.begin());
        cb_result &o = out[k++];
        std::memset(&o, 0, sizeof o);
        o.id = q->id;
        o.tag = q->r.tag;
        o.status = q->status;
        o.t_submit = q->t_submit;
        o.t_admit = q->t_admit;
        o.t_done = q->t_done;
        long long sites = 1;
        for (int d = 0; d < q->r.params.D; ++d) sites *= q->r.params.nGrids;
        o.site_updates = q->status == CB_OK ? sites * (long long)q->r.nt : 0;
        o.n_rows = (int)(q->rows.size() / 16);
        o.n_spectra = (int)q->spec_k.size();
        std::memcpy(o.final_row, q->final_row, sizeof o.final_row);
        sv->kept[q->id] = q;
    }
    *n = k;
    return CB_OK;
}

cb_status cb_server_rows(cb_server *sv, unsigned long long id, double *rows, int cap_rows, int *n)
{
    if (!sv || !n) { set_error("invalid argument"); return CB_INVALID_ARG; }
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->kept.find(id);
    if (it == sv->kept.end()) { set_error("unknown or unpolled request"); return CB_INVALID_ARG; }
    const Req *q = it->second;
    *n = (int)(q->rows.size() / 16);
    if (rows) std::memcpy(rows, q->rows.data(), sizeof(double) * 16 * (size_t)(*n < cap_rows ? *n : cap_rows));
    return CB_OK;
}

cb_status cb_server_spectrum(cb_server *sv, unsigned long long id, int which, double *kmean, double *delta, int cap,
                             int *nbins)
{
    if (!sv || !nbins) { set_error("invalid argument"); return CB_INVALID_ARG; }
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->kept.find(id);
    if (it == sv->kept.end() || which < 0 || which >= (int)it->second->spec_k.size()) {
        set_error("unknown request or spectrum");
        return CB_INVALID_ARG;
    }
    const Req *q = it->second;
    const std::vector<double> &k = q->spec_k[(size_t)which], &d = q->spec_d[(size_t)which];
    *nbins = (int)k.size();
    const size_t m = (size_t)(*nbins < cap ? *nbins : cap);
    if (kmean) std::memcpy(kmean, k.data(), sizeof(double) * m);
    if (delta) std::memcpy(delta, d.data(), sizeof(double) * m);
    return CB_OK;
}

cb_status cb_server_release(cb_server *sv, unsigned long long id)
{
    if (!sv) { set_error("invalid handle"); return CB_INVALID_ARG; }
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->kept.find(id);
    if (it == sv->kept.end()) { set_error("unknown request"); return CB_INVALID_ARG; }
    delete it->second;
    sv->kept.erase(it);
    return CB_OK;
}

cb_status cb_server_stats(cb_server *sv, int *queued, int *active, int *done, long long *bytes)
{
    if (!sv) { set_error("invalid handle"); return CB_INVALID_ARG; }
    std::lock_guard<std::mutex> lk(sv->mu);
    if (queued) *queued = (int)sv->queue.size();
    if (active) *active = sv->busy;
    if (done) *done = (int)sv->done.size();
    if (bytes) *bytes = sv->used;
    return CB_OK;
}

cb_status cb_server_drain(cb_server *sv)
{
    if (!sv) { set_error("invalid handle"); return CB_INVALID_ARG; }
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->idle_cv.wait(lk, [&] { return sv->queue.empty() && sv->busy == 0; });
    return CB_OK;
}

cb_status cb_server_variant(const cb_sim *s, int *variant)
{
    if (!s || !variant) { set_error("invalid argument"); return CB_INVALID_ARG; }
    *variant = cb::server_variant(s);
    return CB_OK;
}

cb_status cb_server_destroy(cb_server *sv)
{
    if (!sv) return CB_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv.notify_all();
    if (sv->th.joinable()) sv->th.join();
    for (Req *q : sv->active) {
        if (q->sim) cb::release_sim(q->sim);
        delete q;
    }
    for (Req *q : sv->queue) delete q;
    for (Req *q : sv->done) delete q;
    for (auto &kv : sv->kept) delete kv.second;
    if (sv->d_tab) hipFree(sv->d_tab);
    if (sv->h_tab) hipHostFree(sv->h_tab);
    if (sv->h_sc) hipHostFree(sv->h_sc);
    if (sv->st) {
        hipStreamSynchronize(sv->st);   // the requests' frees have run
        hipStreamDestroy(sv->st);
    }
    {   // what the requests freed goes back to the device
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, sv->o.device) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
    }
    delete sv;
    return CB_OK;
}

} // extern "C"
