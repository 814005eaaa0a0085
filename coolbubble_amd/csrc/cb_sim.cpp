// cb_sim.cpp -- host side of the drop-in boundary: the Bubble constructor
// sequence, the step loop and the state accessors behind include/coolbubble.h.
//
// Host scalars that the reference computes once (lattice constants, bubble
// fits, the initial Hubble rate) are computed here with the same glibc calls
// as the reference, so they are the reference's values bit for bit.
#include "cb_sim.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

namespace cb {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

} // namespace cb

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

// Nucleation points of mode 1 (GenerateNucleationPoints + check_PBC,
// src/bubble.cpp:397-417, src/gridfuncs.cpp:93-119): glibc's rand() stream
// after srand(seed) (drawn from a private random_r state, so the caller's
// rand() state is untouched), written as text with "%.5e" / "%.6e" and read
// back with strtod exactly as the reference's fscanf("%lf") does.  The text
// (the __xns.dat file) goes to *text_out when given.
std::vector<double> generate_xns(int D, double L, double Rcut, int nBubbles, unsigned seed, std::string *text_out)
{
    struct random_data rd;
    char state[128];
    std::memset(&rd, 0, sizeof rd);
    std::memset(state, 0, sizeof state);
    initstate_r(seed, state, sizeof state, &rd);
    std::string text;
    char line[160];
    for (int b = 0; b < nBubbles; ++b) {
        int32_t r[3];
        for (int q = 0; q < 3; ++q) random_r(&rd, &r[q]);
        double xn[3];
        for (int q = 0; q < 3; ++q) xn[q] = L * (double)r[q] / (double)RAND_MAX;
        std::snprintf(line, sizeof line, "%.5e %.5e %.5e\n", xn[0], xn[1], xn[2]);
        text += line;
        int allowed[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) {
            if (xn[d] + Rcut > L) allowed[d] = -1;
            else if (xn[d] - Rcut < 0) allowed[d] = +1;
        }
        for (int mask = 1; mask < (1 << D); ++mask) {
            double im[3] = {xn[0], xn[1], xn[2]};
            bool valid = true;
            for (int d = 0; d < D; ++d)
                if (mask & (1 << d)) {
                    if (allowed[d] == 0) { valid = false; break; }
                    im[d] += allowed[d] * L;
                }
            if (valid) {
                std::snprintf(line, sizeof line, "%.6e %.6e %.6e\n", im[0], im[1], im[2]);
                text += line;
            }
        }
    }
    std::vector<double> out;
    const char *p = text.c_str();
    for (;;) {
        double v[3];
        char *end = nullptr;
        bool ok = true;
        for (int q = 0; q < 3 && ok; ++q) {
            v[q] = std::strtod(p, &end);
            if (end == p) ok = false;
            p = end;
        }
        if (!ok) break;
        out.insert(out.end(), v, v + 3);
    }
    if (text_out) *text_out = text;
    return out;
}

bool valid_params(const cb_params *p, std::string &why)
{
    if (!p) { why = "params is NULL"; return false; }
    if (p->D < 1 || p->D > 3) { why = "D must be 1, 2 or 3"; return false; }
    if (p->nGrids < 2 || !(p->L > 0) || !(p->dt > 0)) { why = "nGrids >= 2, L > 0 and dt > 0 required"; return false; }
    if (p->readxns < -3 || p->readxns > 1) { why = "readxns must be between -3 and 1"; return false; }
    if (!(p->c > 0)) { why = "c must be positive"; return false; }
    if (p->precision != CB_FP64 && p->precision != CB_FP32) { why = "precision must be CB_FP64 or CB_FP32"; return false; }
    if (p->alpha != 0 && p->alpha != 1) { why = "alpha must be 0 or 1"; return false; }
    if (p->readxns == 1 && p->nBubbles < 0) { why = "nBubbles must be >= 0"; return false; }
    double n = std::pow((double)p->nGrids, p->D);
    if (n > 4.0e10) { why = "lattice too large"; return false; }
    if (p->expansion && p->D == 3 && !(p->Mpl > 0)) { why = "Mpl must be positive"; return false; }
    if (p->thermal_spectrum != 0 && p->thermal_spectrum != 1) { why = "thermal_spectrum must be 0 or 1"; return false; }
    return true;
}

cb::SimpleCfg simple_cfg(long long nsites)
{
    long long blocks = (nsites + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    return cb::SimpleCfg{256, (int)blocks};
}

void release(cb_sim *s)
{
    if (!s) return;
    if (s->st) hipStreamSynchronize(s->st);
    if (s->fft) cb::fft_destroy(s->fft);
    for (int i = 0; i < 2; ++i) {
        cb::sim_free(s, s->phi[i]);
        cb::sim_free(s, s->pi[i]);
    }
    cb::sim_free(s, s->pi0);
    cb::sim_free(s, s->partials);
    cb::sim_free(s, s->dtot);
    cb::sim_free(s, s->dsc);
    if (s->st && s->own_stream) hipStreamDestroy(s->st);
    delete s;
}

// Host copies of the per-step factors for a given (a, ap) (src/bubble.cpp:741-742,765).
void host_factors(const cb_sim *s, cb::DevScalars &h)
{
    const int D = s->p.D, al = s->p.alpha;
    h.fac_lap = std::pow(h.a, D - 2 + al);
    h.fac_pot = std::pow(h.a, D + al);
    h.adx = h.a * s->k.dx;
    h.inv_adx = 1.0 / h.adx;
    double a_kin = h.a + 0.5 * s->p.dt * h.ap;
    h.pk2 = std::pow(a_kin, -2 * D);
    h.pk1 = std::pow(a_kin, -D);
}

cb_status sum_to_host(cb_sim *s, double tot[cb::NQ])
{
    HIP_TRY(cb::launch_sum_partials(s->partials, s->cfg.grid, s->dtot, s->st));
    HIP_TRY(hipMemcpyAsync(tot, s->dtot, sizeof(double) * cb::NQ, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return CB_OK;
}

// Bubble::Bubble, src/bubble.cpp:156-316.
cb_status construct(cb_sim *s, const double *xns, int n_xns)
{
    const cb_params &p = s->p;
    const int D = p.D, g = p.nGrids;
    const double L = p.L, c = p.c;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return CB_NO_DEVICE;
    }
    if (p.device < 0 || p.device >= ndev) { set_error("device ordinal out of range"); return CB_INVALID_ARG; }
    HIP_TRY(hipSetDevice(p.device));
    if (!s->st) {
        HIP_TRY(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        s->own_stream = true;
    }

    // lattice constants (src/bubble.hpp:43-49)
    s->nsites = (long long)std::pow((double)g, D);
    s->fp32 = p.precision == CB_FP32;
    s->esz = s->fp32 ? 4 : 8;
    s->lagged = p.expansion != 0 && D == 3;
    cb::SimConst &k = s->k;
    k.nsites = s->nsites;
    k.N = g;
    k.D = D;
    k.alpha = p.alpha;
    k.expanding = s->lagged ? 1 : 0;
    k.dt = p.dt;
    k.dx = (double)L / (double)g;
    k.dx2 = k.dx * k.dx;
    {
        int ex = 0;
        k.inv_dx2 = std::frexp(k.dx2, &ex) == 0.5 ? 1.0 / k.dx2 : 0.0;
    }
    k.dxLD = std::pow(k.dx / L, D);
    k.c = c;
    k.DeltaV = 1.;
    k.c2 = 2 * c;
    k.c3 = 3 * (2 * c + 4);
    k.c4 = 4 * (c + 3);
    k.c5 = 2 * c + 4;
    k.c6 = c + 3;
    k.Mpl = p.Mpl;
    s->dxD = std::pow(k.dx, D);
    s->nkmax = (int)(std::sqrt(double(D)) * g / 2);
    // potential and critical bubble (src/bubble.cpp:81-99,160-171)
    s->phiC = 0.997 * (1 - std::exp(-c * 0.688));
    s->Rn = (D == 3) ? 0.823 * std::pow(c, 0.203) * 1.3 : 0.823 * std::pow(c, 0.203);
    s->lw = 0.496 * std::pow(c, -0.531);
    s->mT = std::sqrt(2 * (c + 6));
    s->cfg = simple_cfg(s->nsites);

    {   // step-kernel variants and the energy-partials capacity they need
        static const int kMaxVariants = 512;
        std::vector<cb::TiledCfg> cand((size_t)kMaxVariants);
        int nc = cb::tiled_candidates(k, s->fp32, cand.data(), kMaxVariants);
        s->variants.assign(cand.begin(), cand.begin() + nc);
        s->max_blocks = s->cfg.grid;
        for (const auto &c : s->variants) {
            int nb = cb::tiled_blocks(k, c);
            if (nb > s->max_blocks) s->max_blocks = nb;
        }
    }
    const size_t bytes = (size_t)s->nsites * s->esz;
    for (int i = 0; i < 2; ++i) {
        HIP_TRY(cb::sim_alloc(s, &s->phi[i], bytes));
        if (i == 0 || s->lagged || s->p.D == 1) HIP_TRY(cb::sim_alloc(s, &s->pi[i], bytes));
    }
    HIP_TRY(cb::sim_alloc(s, (void **)&s->partials, sizeof(double) * cb::NQ * (size_t)s->max_blocks));
    HIP_TRY(cb::sim_alloc(s, (void **)&s->dtot, sizeof(double) * cb::NQ));
    HIP_TRY(cb::sim_alloc(s, (void **)&s->dsc, sizeof(cb::DevScalars)));
    HIP_TRY(hipMemsetAsync(s->phi[0], 0, bytes, s->st));
    HIP_TRY(hipMemsetAsync(s->pi[0], 0, bytes, s->st));

    // initial condition (src/bubble.cpp:191-234)
    std::vector<double> cen, sub;
    auto add = [&](double x0, double x1, double x2, double s0, double s1, double s2) {
        cen.insert(cen.end(), {x0, x1, x2});
        sub.insert(sub.end(), {s0, s1, s2});
    };
    bool bubbles = true;
    if (D == 1) {                                  // Initialize1D: profiles at x and x - L
        add(0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
        add(L, 0.0, 0.0, 0.0, 0.0, 0.0);
    } else if (p.test_wave == 1) {
        bubbles = false;
        double kw = 2.0 * M_PI / (L / 40.0);
        HIP_TRY(cb::launch_seed_wave(k, s->fp32, s->phi[0], L, kw, 0.01 * s->phi0, s->phi0, s->cfg, s->st));
    } else if (p.readxns == -2) {                  // NucleateOneBubble
        add(L / 2., L / 2., L / 2., 0.0, 0.0, 0.0);
    } else if (p.readxns == -1) {                  // NucleateTwoDiagonalBubbles
        add(L / 4, L / 4., L / 2., p.DeltaR, p.DeltaR, 0.0);
        add(3. * L / 4., 3. * L / 4, L / 2., -p.DeltaR, -p.DeltaR, 0.0);
    } else if (p.readxns == 1 || p.readxns == 0) {
        if (p.readxns == 1) {
            s->xns = generate_xns(D, L, s->Rn + 5 * s->lw, p.nBubbles, p.rand_seed, &s->xns_text);
        } else {
            if (n_xns < 0 || (n_xns > 0 && !xns)) { set_error("mode 0 needs the bubble centres"); return CB_INVALID_ARG; }
            s->xns.assign(xns, xns + 3 * (size_t)n_xns);
        }
        for (size_t b = 0; b + 2 < s->xns.size(); b += 3) add(s->xns[b], s->xns[b + 1], s->xns[b + 2], 0.0, 0.0, 0.0);
    } else {                                       // readxns == -3: thermal fluctuations
        bubbles = false;
        // the reference's temperature is Trad in units of the true-vacuum
        // mass (src/bubble.cpp:223); the Rayleigh-Jeans option takes T = Trad
        double Temp = p.thermal_spectrum == 1 ? p.Trad : p.Trad * s->mT;
        double mF = std::sqrt(2 * c);
        cb_status r = cb::thermal_field(s, s->phi[0], Temp, mF, 0);
        if (r != CB_OK) return r;
        r = cb::thermal_field(s, s->pi[0], Temp, mF, 1);
        if (r != CB_OK) return r;
    }
    if (bubbles && !cen.empty()) {
        const int ncen = (int)(cen.size() / 3);
        double *dcen = nullptr;
        HIP_TRY(cb::sim_alloc(s, (void **)&dcen, sizeof(double) * 6 * (size_t)ncen));
        hipError_t e = hipMemcpyAsync(dcen, cen.data(), sizeof(double) * 3 * ncen, hipMemcpyHostToDevice, s->st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dcen + 3 * ncen, sub.data(), sizeof(double) * 3 * ncen, hipMemcpyHostToDevice, s->st);
        if (e == hipSuccess)
            e = cb::launch_seed_bubbles(k, s->fp32, s->phi[0], dcen, dcen + 3 * ncen, ncen, L, s->phiC, s->Rn, s->lw,
                                        s->cfg, s->st);
        if (e == hipSuccess) e = hipStreamSynchronize(s->st);
        cb::sim_free(s, dcen);
        if (e != hipSuccess) { set_error(std::string("seeding: ") + hipGetErrorString(e)); return CB_HIP_ERROR; }
    }

    // pi at t = 0, before the back-step below: the kinetic-energy map of the
    // initial snapshot reads it (src/bubble.cpp:253-261)
    if (s->keep_pi0) {
        HIP_TRY(cb::sim_alloc(s, &s->pi0, s->esz * (size_t)s->nsites));
        HIP_TRY(hipMemcpyAsync(s->pi0, s->pi[0], s->esz * (size_t)s->nsites, hipMemcpyDeviceToDevice, s->st));
    }

    // initial energies (src/bubble.cpp:239-269)
    double tot[cb::NQ];
    HIP_TRY(cb::launch_ctor_energy(k, s->fp32, s->phi[0], s->pi[0], s->partials, s->cfg, s->st));
    cb_status r = sum_to_host(s, tot);
    if (r != CB_OK) return r;
    cb::DevScalars h{};
    h.avphi = tot[0];
    h.avphiP = tot[1];
    h.Epottot = tot[2];
    h.Egradtot = tot[3];
    h.Ekintot = tot[4];
    h.Etot = h.Ekintot + h.Epottot + h.Egradtot;
    // expansion (src/bubble.cpp:273-276)
    h.a = 1;
    const bool expanding = s->lagged;
    const double Mpl = p.Mpl;
    double H0 = expanding ? std::sqrt(std::pow(h.a, 2 * p.alpha) * h.Etot / (3.0 * Mpl * Mpl)) : 0.0;
    h.ap = h.a * H0;
    double addot = 0.0;
    if (expanding) {
        if (p.alpha == 0) addot = -h.a * (2.0 * h.Ekintot - h.Epottot) / (3.0 * Mpl * Mpl);
        else addot = h.a * h.a * h.a * (2 * h.Epottot - h.Ekintot + h.Egradtot) / (3.0 * Mpl * Mpl);
    }
    h.ap -= 0.5 * p.dt * addot;
    // pi back-step (src/bubble.cpp:280-289)
    double a_mhalf = h.a - 0.5 * p.dt * h.ap;
    HIP_TRY(cb::launch_backstep(k, s->fp32, s->phi[0], s->pi[0], std::pow(a_mhalf, -2 * D), s->partials, s->cfg, s->st));
    r = sum_to_host(s, tot);
    if (r != CB_OK) return r;
    h.EkinHalfPrev = tot[0];
    h.H = H0;
    h.C = 0.;
    h.s_pending = 0.0;
    host_factors(s, h);
    HIP_TRY(hipMemcpyAsync(s->dsc, &h, sizeof h, hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    s->t = 0.;
    s->step = 0;
    s->row_valid = true;
    return cb::select_variant(s, cb::default_variant(s));
}

cb_status read_scalars(cb_sim *s, cb::DevScalars &h)
{
    HIP_TRY(hipMemcpyAsync(&h, s->dsc, sizeof h, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return CB_OK;
}

cb_status materialise(cb_sim *s)
{
    if (!s->lagged) return CB_OK;
    HIP_TRY(cb::launch_apply_drift(s->k, s->fp32, s->phi[s->iphi], s->pi[s->ipi], s->dsc, s->cfg, s->st));
    return CB_OK;
}

} // namespace

namespace cb {
cb_status create_sim(const cb_params *p, const double *xns, int n_xns, bool keep_pi0, cb_sim **out)
{
    if (!out) { set_error("out is NULL"); return CB_INVALID_ARG; }
    *out = nullptr;
    std::string why;
    if (!valid_params(p, why)) { set_error(why); return CB_INVALID_ARG; }
    cb_status r = cb_check_stability(p);
    if (r != CB_OK) return r;
    cb_sim *s = new (std::nothrow) cb_sim();
    if (!s) { set_error("host allocation failed"); return CB_OUT_OF_MEMORY; }
    s->p = *p;
    s->keep_pi0 = keep_pi0;
    r = construct(s, xns, n_xns);
    if (r != CB_OK) {
        if (r == CB_HIP_ERROR && hipGetLastError() == hipErrorOutOfMemory) r = CB_OUT_OF_MEMORY;
        release(s);
        return r;
    }
    *out = s;
    return CB_OK;
}
cb_status materialise_drift(cb_sim *s) { return materialise(s); }
cb_status sim_scalars(cb_sim *s, DevScalars &h) { return read_scalars(s, h); }
hipError_t sim_alloc(const cb_sim *s, void **p, size_t bytes)
{
    if (!s->pool_st) return hipMalloc(p, bytes);
    hipError_t e = hipMallocAsync(p, bytes, s->pool_st);
    if (e != hipErrorOutOfMemory) return e;
    (void)hipGetLastError();
    // free memory held by the pool back to the device, then once more
    hipMemPool_t pool = nullptr;
    if (hipStreamSynchronize(s->pool_st) != hipSuccess ||
        hipDeviceGetDefaultMemPool(&pool, s->p.device) != hipSuccess || hipMemPoolTrimTo(pool, 0) != hipSuccess)
        return hipErrorOutOfMemory;
    return hipMallocAsync(p, bytes, s->pool_st);
}
void sim_free(const cb_sim *s, void *p)
{
    if (!p) return;
    if (s->pool_st) hipFreeAsync(p, s->pool_st);
    else hipFree(p);
}
bool row_from_scalars(const cb_sim *s, const DevScalars &h, double c[16])
{
    c[0] = s->t;
    c[1] = h.a;
    c[2] = h.Etot;
    c[3] = h.Ekintot;
    c[4] = h.Egradtot;
    c[5] = h.Epottot;
    c[6] = std::sqrt(h.EtotVar);
    c[7] = std::sqrt(h.EkintotVar);
    c[8] = std::sqrt(h.EgradtotVar);
    c[9] = std::sqrt(h.EpottotVar);
    c[10] = h.avphi;
    c[11] = std::sqrt(h.avphiVar);
    c[12] = h.avphiP;
    c[13] = std::sqrt(h.avphiPVar);
    c[14] = h.H;
    c[15] = h.C;
    return !h.nan_flag;
}
cb_status construct_sim(cb_sim *s, const double *xns, int n_xns) { return construct(s, xns, n_xns); }
void release_sim(cb_sim *s) { release(s); }
bool params_ok(const cb_params *p, std::string &why) { return valid_params(p, why); }
} // namespace cb

extern "C" {

cb_status cb_default_params(cb_params *p)
{
    if (!p) return CB_INVALID_ARG;
    std::memset(p, 0, sizeof *p);
    p->D = 2;
    p->nGrids = 512;
    p->L = 32;
    p->c = 4.;
    p->readxns = 1;
    p->dt = 0.004;
    p->nBubbles = 5;
    p->Trad = 0.2;
    p->DeltaR = 0.;
    p->test_wave = 0;
    p->expansion = 0;
    p->alpha = 0;
    p->Mpl = 200.0;
    p->rand_seed = 42;
    p->thermal_base = 12345ULL;
    p->precision = CB_FP64;
    p->device = 0;
    return CB_OK;
}

cb_status cb_check_stability(const cb_params *p)
{
    std::string why;
    if (!valid_params(p, why)) { set_error(why); return CB_INVALID_ARG; }
    double dx = double(p->L) / p->nGrids;
    double dtmax = dx / std::sqrt(p->D);
    if (p->dt < dtmax) return CB_OK;
    set_error("stability condition dt < dx/sqrt(D) violated");
    return CB_CFL_VIOLATION;
}

cb_status cb_create(const cb_params *p, const double *xns, int n_xns, cb_sim **out)
{
    return cb::create_sim(p, xns, n_xns, false, out);
}

cb_status cb_step(cb_sim *s, int nsteps)
{
    if (!s || nsteps < 0) { set_error("invalid handle or step count"); return CB_INVALID_ARG; }
    if (nsteps == 0) return CB_OK;
    return cb::step_n(s, nsteps);
}

cb_status cb_energies_row(cb_sim *s, cb_energies *out)
{
    if (!s || !out) { set_error("invalid argument"); return CB_INVALID_ARG; }
    cb::DevScalars h;
    cb_status r = read_scalars(s, h);
    if (r != CB_OK) return r;
    if (!cb::row_from_scalars(s, h, out->col)) { set_error("non-finite energy detected"); return CB_NAN_DETECTED; }
    return CB_OK;
}

cb_status cb_get_fields(cb_sim *s, double *phi, double *pi)
{
    if (!s) { set_error("invalid handle"); return CB_INVALID_ARG; }
    cb_status r = materialise(s);
    if (r != CB_OK) return r;
    void *src[2] = {s->phi[s->iphi], s->pi[s->ipi]};
    double *dst[2] = {phi, pi};
    for (int f = 0; f < 2; ++f) {
        if (!dst[f]) continue;
        if (!s->fp32) {
            HIP_TRY(hipMemcpyAsync(dst[f], src[f], sizeof(double) * s->nsites, hipMemcpyDeviceToHost, s->st));
        } else {
            std::vector<float> tmp((size_t)s->nsites);
            HIP_TRY(hipMemcpyAsync(tmp.data(), src[f], sizeof(float) * s->nsites, hipMemcpyDeviceToHost, s->st));
            HIP_TRY(hipStreamSynchronize(s->st));
            for (long long i = 0; i < s->nsites; ++i) dst[f][i] = (double)tmp[(size_t)i];
        }
    }
    HIP_TRY(hipStreamSynchronize(s->st));
    return CB_OK;
}

cb_status cb_get_state(cb_sim *s, cb_state *st)
{
    if (!s || !st) { set_error("invalid argument"); return CB_INVALID_ARG; }
    cb::DevScalars h;
    cb_status r = read_scalars(s, h);
    if (r != CB_OK) return r;
    st->a = h.a;
    st->ap = h.ap;
    st->H = h.H;
    st->C = h.C;
    st->t = s->t;
    st->EkinHalfPrev = h.EkinHalfPrev;
    st->step = s->step;
    return CB_OK;
}

cb_status cb_set_state(cb_sim *s, const cb_state *st, const double *phi, const double *pi)
{
    if (!s || !st) { set_error("invalid argument"); return CB_INVALID_ARG; }
    HIP_TRY(hipStreamSynchronize(s->st));
    const double *srcs[2] = {phi, pi};
    void *dsts[2] = {s->phi[s->iphi], s->pi[s->ipi]};
    for (int f = 0; f < 2; ++f) {
        if (!srcs[f]) continue;
        if (!s->fp32) {
            HIP_TRY(hipMemcpy(dsts[f], srcs[f], sizeof(double) * s->nsites, hipMemcpyHostToDevice));
        } else {
            std::vector<float> tmp((size_t)s->nsites);
            for (long long i = 0; i < s->nsites; ++i) tmp[(size_t)i] = (float)srcs[f][i];
            HIP_TRY(hipMemcpy(dsts[f], tmp.data(), sizeof(float) * s->nsites, hipMemcpyHostToDevice));
        }
    }
    cb::DevScalars h;
    cb_status r = read_scalars(s, h);
    if (r != CB_OK) return r;
    h.a = st->a;
    h.ap = st->ap;
    h.H = st->H;
    h.C = st->C;
    h.EkinHalfPrev = st->EkinHalfPrev;
    h.s_pending = 0.0;
    h.nan_flag = 0;
    host_factors(s, h);
    HIP_TRY(hipMemcpyAsync(s->dsc, &h, sizeof h, hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    s->t = st->t;
    s->step = st->step;
    return CB_OK;
}

cb_status cb_get_xns(cb_sim *s, double *xyz, int cap, int *n)
{
    if (!s || !n) { set_error("invalid argument"); return CB_INVALID_ARG; }
    int rows = (int)(s->xns.size() / 3);
    *n = rows;
    if (xyz)
        for (int i = 0; i < rows && i < cap; ++i) std::memcpy(xyz + 3 * i, s->xns.data() + 3 * i, 3 * sizeof(double));
    return CB_OK;
}

cb_status cb_power_spectrum(cb_sim *s, double *kmean, double *kphys, double *delta, int cap, int *nbins)
{
    if (!s || !nbins) { set_error("invalid argument"); return CB_INVALID_ARG; }
    cb_status r = materialise(s);
    if (r != CB_OK) return r;
    return cb::spectrum(s, kmean, kphys, delta, cap, nbins);
}

cb_status cb_power_spectrum_noalloc(cb_sim *s, double *kmean, double *kphys, double *delta, int cap, int *nbins)
{
    if (!s || !nbins) { set_error("invalid argument"); return CB_INVALID_ARG; }
    cb_status r = materialise(s);
    if (r != CB_OK) return r;
    return cb::spectrum_noalloc(s, kmean, kphys, delta, cap, nbins);
}

cb_status cb_kernel_variants(cb_sim *s, int *count, int *selected)
{
    if (!s) { set_error("invalid handle"); return CB_INVALID_ARG; }
    if (count) *count = (int)s->variants.size();
    if (selected) *selected = s->variant;
    return CB_OK;
}

cb_status cb_kernel_variant_info(cb_sim *s, int variant, int info[8])
{
    if (!s || !info || variant < 0 || variant >= (int)s->variants.size()) {
        set_error("invalid handle or variant");
        return CB_INVALID_ARG;
    }
    const cb::TiledCfg &c = s->variants[(size_t)variant];
    const int v[8] = {c.vw, c.items, c.nthr, c.TR, c.NI, c.pf, c.nt, c.fuse};
    for (int i = 0; i < 8; ++i) info[i] = v[i];
    return CB_OK;
}

cb_status cb_select_kernel(cb_sim *s, int variant)
{
    if (!s) { set_error("invalid handle"); return CB_INVALID_ARG; }
    return cb::select_variant(s, variant);
}

cb_status cb_stream(cb_sim *s, void **stream)
{
    if (!s || !stream) { set_error("invalid argument"); return CB_INVALID_ARG; }
    *stream = (void *)s->st;
    return CB_OK;
}

long long cb_nsites(const cb_sim *s) { return s ? s->nsites : 0; }

cb_status cb_destroy(cb_sim *s)
{
    if (s && s->batch_owned) {
        set_error("this simulation belongs to a batch: destroy the batch");
        return CB_INVALID_ARG;
    }
    release(s);
    return CB_OK;
}

const char *cb_status_string(cb_status st)
{
    switch (st) {
    case CB_OK: return "ok";
    case CB_INVALID_ARG: return "invalid argument";
    case CB_CFL_VIOLATION: return "stability condition dt < dx/sqrt(D) violated";
    case CB_OUT_OF_MEMORY: return "out of device memory";
    case CB_HIP_ERROR: return "HIP runtime error";
    case CB_NAN_DETECTED: return "non-finite value detected";
    case CB_NO_DEVICE: return "no HIP device";
    case CB_ADMISSION_REJECTED: return "admission rejected";
    case CB_NOT_READY: return "not ready";
    case CB_UNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

const char *cb_last_error(void) { return cb::g_err.c_str(); }

} // extern "C"
