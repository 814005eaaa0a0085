/*
 * cb_oracle.c -- CPU FP64 restatement of CoolBubble's lattice time-stepper.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * Each function follows the reference file:line given in its comment, in the
 * reference's evaluation order.  OpenMP loops use a static split with
 * per-thread partial sums combined in thread order, so a fixed thread count
 * gives a deterministic result.  The box sums (energies, averages, spreads)
 * are Neumaier-compensated: the reference combines its OpenMP reduction
 * partials in an unspecified order (src/bubble.cpp:776), so over 1e8 sites
 * its own sums scatter by ~sqrt(n) ulp relative (~1e-12 at 512^3); the
 * oracle evaluates each sum of the same per-site terms to within a few ulp
 * of the exact sum, the value every summation order approximates.
 */
#define _GNU_SOURCE
#include "cb_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct cbo_sim {
    cbo_params p;
    int nthreads;
    long long N;       /* src/bubble.hpp:43 */
    int n_cplx;        /* src/bubble.hpp:44 */
    double dx, dxD, dxLD;
    int nkmax;         /* src/bubble.hpp:49 */
    double DeltaV, phi0, phiC, Rn, lw, mT;
    double *phi, *pi, *dpi;
    double a, ap, H, C, t;
    double avphi, avphiP, Etot, Ekintot, Epottot, Egradtot, EkinHalfPrev;
    double EtotVar, EkintotVar, EpottotVar, EgradtotVar, avphiPVar, avphiVar;
    double *xns;       /* centres read back from the __xns.dat text */
    int n_xns;
    /* static box only: measure the energies on steps n with n % energy_every
     * == 0 and the kinetic half-step sum on the steps before them (0: every
     * step, as the reference does); nstep counts the steps taken */
    int energy_every;
    long long nstep;
};

/* ------------------------------------------------------------------ */
/* Lattice indexing, src/gridfuncs.cpp:31-51                           */
/* ------------------------------------------------------------------ */
int cbo_iGridM(int D, int nGrids, int iGrid, int iShift)
{
    int n = nGrids;
    if (D == 1) {                                   /* :33-37 */
        if (iShift == 1) return (iGrid + n + 1) % n;
        if (iShift == 0) return (iGrid + n - 1) % n;
    }
    if (D == 2)                                     /* :38 */
        return n * ((iGrid / n + n - (iShift == 0 ? 1 : 0)) % n)
               + (iGrid + n - (iShift == 1 ? 1 : 0)) % n;
    if (D == 3)                                     /* :39 */
        return n * (n * ((iGrid / (n * n) + n - (iShift == 0 ? 1 : 0)) % n)
                    + ((iGrid / n + n - (iShift == 1 ? 1 : 0)) % n))
               + (iGrid + n - (iShift == 2 ? 1 : 0)) % n;
    return 0;
}

int cbo_iGridP(int D, int nGrids, int iGrid, int iShift)
{
    int n = nGrids;
    if (D == 2)                                     /* :48 */
        return n * ((iGrid / n + (iShift == 0 ? 1 : 0)) % n)
               + (iGrid + (iShift == 1 ? 1 : 0)) % n;
    if (D == 3)                                     /* :49 */
        return n * (n * ((iGrid / (n * n) + (iShift == 0 ? 1 : 0)) % n)
                    + (iGrid / n + (iShift == 1 ? 1 : 0)) % n)
               + (iGrid + (iShift == 2 ? 1 : 0)) % n;
    return 0;
}

/* xFunc, src/gridfuncs.cpp:77-85: cell-centred coordinates */
static void xfunc(int D, int n, double L, long long iGrid, double x[3])
{
    long long i0 = 0, i1 = 0, i2 = 0;
    if (D == 1) { i0 = iGrid % n; }
    if (D == 2) { i0 = (iGrid / n) % n; i1 = iGrid % n; }
    if (D == 3) { i0 = (iGrid / ((long long)n * n)) % n; i1 = (iGrid / n) % n; i2 = iGrid % n; }
    x[0] = x[1] = x[2] = 0.;
    x[0] = L * ((double)i0 + 0.5) / (double)n;
    if (D >= 2) x[1] = L * ((double)i1 + 0.5) / (double)n;
    if (D == 3) x[2] = L * ((double)i2 + 0.5) / (double)n;
}

/* ------------------------------------------------------------------ */
/* Potential and critical-bubble fits, src/bubble.cpp:69-108           */
/* ------------------------------------------------------------------ */
double cbo_V(double phi, double cc, double DeltaV)          /* :69-72 */
{
    return cc * (phi * phi) - (2 * cc + 4) * pow(phi, 3) + (cc + 3) * pow(phi, 4) + DeltaV;
}

double cbo_Vpr(double phi, double cc)                       /* :74-77 */
{
    return 2 * cc * phi - 3 * (2 * cc + 4) * (phi * phi) + 4 * (cc + 3) * pow(phi, 3);
}

static double phizero(double c) { return 0.997 * (1 - exp(-c * 0.688)); }   /* :81-84 */
static double cradius(int D, double c)                                      /* :87-94 */
{
    if (D == 3) return 0.823 * pow(c, 0.203) * 1.3;
    return 0.823 * pow(c, 0.203);
}
static double clw(double c) { return 0.496 * pow(c, -0.531); }              /* :96-99 */

void cbo_bubble_fit(int D, double c, double out[3])
{
    out[0] = phizero(c);
    out[1] = cradius(D, c);
    out[2] = clw(c);
}

/* ThickBubble, src/bubble.cpp:102-108 (fits recomputed per call there; the
 * values are identical, so they are hoisted here) */
static double thick_bubble(double x2, double phiC, double Rn, double lw)
{
    return (phiC / 2.) * (tanh((Rn - sqrt(x2)) / lw) + 1.);
}

static double dmin(double a, double b) { return (b < a) ? b : a; }  /* std::min */

/* ------------------------------------------------------------------ */
/* Thermal spectra and seed, src/bubble.cpp:116-142                    */
/* ------------------------------------------------------------------ */
static double thermalspectrum(double T, double k2, double m)
{
    double omega_k = sqrt(k2 + m * m);
    double n_k = 1.0 / (exp(omega_k / T) - 1.0);
    return n_k / omega_k;
}
static double velocitythermalspectrum(double T, double k2, double m)
{
    double omega_k = sqrt(k2 + m * m);
    double n_k = 1.0 / (exp(omega_k / T) - 1.0);
    return n_k * omega_k;
}
static uint64_t thermal_seed(uint64_t base, const char *spec)        /* :134-142 */
{
    uint64_t off = 0;
    for (const char *ch = spec; *ch; ++ch) off = off * 131 + (unsigned char)*ch;
    return base + off;
}

/* ------------------------------------------------------------------ */
/* mt19937_64 (Matsumoto & Nishimura 64-bit Mersenne Twister with the  */
/* parameters of std::mt19937_64) and libstdc++'s                      */
/* generate_canonical<double,53> + Marsaglia polar normal_distribution */
/* (/usr/include/c++/11/bits/random.tcc:1802-1835, 3348-3378), which   */
/* the reference draws from at src/bubble.cpp:476-485.                 */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t mt[312]; int mti; } mt64;

static void mt64_seed(mt64 *g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->mti = 312;
}

static uint64_t mt64_next(mt64 *g)
{
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    const uint64_t A = 0xB5026F5AA96619E9ULL;
    if (g->mti >= 312) {
        int i;
        for (i = 0; i < 312 - 156; ++i) {
            uint64_t x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
            g->mt[i] = g->mt[i + 156] ^ (x >> 1) ^ ((x & 1ULL) ? A : 0ULL);
        }
        for (; i < 311; ++i) {
            uint64_t x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
            g->mt[i] = g->mt[i + 156 - 312] ^ (x >> 1) ^ ((x & 1ULL) ? A : 0ULL);
        }
        uint64_t x = (g->mt[311] & UM) | (g->mt[0] & LM);
        g->mt[311] = g->mt[155] ^ (x >> 1) ^ ((x & 1ULL) ? A : 0ULL);
        g->mti = 0;
    }
    uint64_t x = g->mt[g->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

static double canonical(mt64 *g)
{
    double r = (double)mt64_next(g) / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

typedef struct { double saved; int have; } polar;

static double polar_next(polar *st, mt64 *g)
{
    if (st->have) { st->have = 0; return st->saved; }
    double x, y, r2;
    do {
        x = 2.0 * canonical(g) - 1.0;
        y = 2.0 * canonical(g) - 1.0;
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    double mult = sqrt(-2 * log(r2) / r2);
    st->saved = x * mult;
    st->have = 1;
    return y * mult;
}

void cbo_mt64(uint64_t seed, long n, uint64_t *out)
{
    mt64 g; mt64_seed(&g, seed);
    for (long i = 0; i < n; ++i) out[i] = mt64_next(&g);
}

void cbo_normals(uint64_t seed, long n, double *out)
{
    mt64 g; mt64_seed(&g, seed);
    polar st = {0., 0};
    for (long i = 0; i < n; ++i) out[i] = polar_next(&st, &g) * 1.0 + 0.0;
}

/* ------------------------------------------------------------------ */
/* Exact DFT in long double (stand-in for the mathematics FFTW          */
/* evaluates; FFTW's own rounding is not reproduced).                  */
/* sign = -1: forward (FFTW_FORWARD), +1: backward, unnormalised.      */
/* ------------------------------------------------------------------ */
typedef long double ld;

/* W_n^e = exp(sign * 2 pi i e / n) from a table of the root length */
typedef struct { int n; const ld *c, *s; } twtab;

static void tw_at(const twtab *t, long long e, long long n, int sign, ld *c, ld *s)
{
    long long step = t->n / n;                 /* n divides the root length */
    long long k = ((e % n) + n) % n * step;
    *c = t->c[k];
    *s = sign * t->s[k];
}

/* recursive mixed-radix decimation in time on strided input; y = output,
 * t = scratch of the same length */
static void fft_rec(const twtab *tw, int n, int sign, const ld *xr, const ld *xi, long long stride,
                    ld *yr, ld *yi, ld *tr, ld *ti)
{
    if (n == 1) { yr[0] = xr[0]; yi[0] = xi[0]; return; }
    int p = 2;
    while (p * p <= n && n % p) ++p;
    if (n % p) p = n;                          /* n is prime */
    int m = n / p;
    if (m == 1) {                              /* direct DFT of prime length */
        for (int k = 0; k < n; ++k) {
            ld sr = 0, si = 0;
            for (int j = 0; j < n; ++j) {
                ld c, s; tw_at(tw, (long long)j * k, n, sign, &c, &s);
                ld ar = xr[j * stride], ai = xi[j * stride];
                sr += ar * c - ai * s;
                si += ar * s + ai * c;
            }
            yr[k] = sr; yi[k] = si;
        }
        return;
    }
    for (int q = 0; q < p; ++q)
        fft_rec(tw, m, sign, xr + q * stride, xi + q * stride, stride * p,
                tr + (long long)q * m, ti + (long long)q * m, yr, yi);
    for (int k = 0; k < m; ++k)
        for (int s2 = 0; s2 < p; ++s2) {
            long long kk = k + (long long)m * s2;
            ld sr = 0, si = 0;
            for (int q = 0; q < p; ++q) {
                ld c, s; tw_at(tw, (long long)q * kk, n, sign, &c, &s);
                ld ar = tr[(long long)q * m + k], ai = ti[(long long)q * m + k];
                sr += ar * c - ai * s;
                si += ar * s + ai * c;
            }
            yr[kk] = sr; yi[kk] = si;
        }
}

/* table of cos / sin (2 pi e / n), e < n, in long double: tab[0..n) cos,
 * tab[n..2n) sin */
static void make_twtab(int n, ld *tab, twtab *tw)
{
    const ld twopi = 2.0L * 3.141592653589793238462643383279502884L;
    for (int e = 0; e < n; ++e) {
        ld ang = twopi * (ld)e / (ld)n;
        tab[e] = cosl(ang); tab[n + e] = sinl(ang);
    }
    tw->n = n; tw->c = tab; tw->s = tab + n;
}

/* one line of n elements at stride st, transformed in place with the
 * table tw of length n; work holds 6n long doubles (copy, output, scratch) */
static void fft_line_ld(int n, int sign, ld *re, ld *im, long long st, ld *work, const twtab *tw)
{
    ld *xr = work, *xi = work + n, *yr = work + 2 * n, *yi = work + 3 * n;
    ld *tr = work + 4 * n, *ti = work + 5 * n;
    for (int e = 0; e < n; ++e) { xr[e] = re[e * st]; xi[e] = im[e * st]; }
    fft_rec(tw, n, sign, xr, xi, 1, yr, yi, tr, ti);
    for (int e = 0; e < n; ++e) { re[e * st] = yr[e]; im[e * st] = yi[e]; }
}

void cbo_dft_ld(int n, int sign, const double *in_re, const double *in_im,
                double *out_re, double *out_im)
{
    ld *re = malloc(sizeof(ld) * n), *im = malloc(sizeof(ld) * n);
    ld *work = malloc(sizeof(ld) * 6 * n), *tab = malloc(sizeof(ld) * 2 * n);
    twtab tw;
    make_twtab(n, tab, &tw);
    for (int e = 0; e < n; ++e) { re[e] = in_re[e]; im[e] = in_im ? in_im[e] : 0; }
    fft_line_ld(n, sign, re, im, 1, work, &tw);
    for (int e = 0; e < n; ++e) { out_re[e] = (double)re[e]; out_im[e] = (double)im[e]; }
    free(re); free(im); free(work); free(tab);
}

/* all-axes transform of a D-dim complex array of side N (row-major, last
 * axis fastest); axes >= first_axis and < last_axis */
static void fft_axes_ld(int D, int N, ld *re, ld *im, int first_axis, int last_axis,
                        int sign, int nthreads)
{
    long long total = 1;
    for (int d = 0; d < D; ++d) total *= N;
    ld *tab = malloc(sizeof(ld) * 2 * N);
    twtab tw;
    make_twtab(N, tab, &tw);
    for (int ax = first_axis; ax < last_axis; ++ax) {
        long long st = 1;
        for (int d = ax + 1; d < D; ++d) st *= N;
        long long lines = total / N;
#pragma omp parallel num_threads(nthreads)
        {
            ld *work = malloc(sizeof(ld) * 6 * N);
#pragma omp for schedule(static)
            for (long long l = 0; l < lines; ++l) {
                long long outer = l / st, inner = l % st;
                long long base = outer * st * N + inner;
                fft_line_ld(N, sign, re + base, im + base, st, work, &tw);
            }
            free(work);
        }
    }
    free(tab);
}

/* r2c of a real D-dim field (FFTW_FORWARD, unnormalised), half-complex
 * output rows x (N/2+1); src/bubble.cpp:648-651 */
void cbo_r2c(int D, int N, const double *field, double *out_re, double *out_im)
{
    long long total = 1;
    for (int d = 0; d < D; ++d) total *= N;
    ld *re = malloc(sizeof(ld) * total), *im = malloc(sizeof(ld) * total);
    for (long long i = 0; i < total; ++i) { re[i] = field[i]; im[i] = 0; }
    fft_axes_ld(D, N, re, im, 0, D, -1, omp_get_max_threads());
    int nc = N / 2 + 1;
    long long rows = total / N;
    for (long long r = 0; r < rows; ++r)
        for (int j = 0; j < nc; ++j) {
            out_re[r * nc + j] = (double)re[r * N + j];
            out_im[r * nc + j] = (double)im[r * N + j];
        }
    free(re); free(im);
}

/* c2r of a half-complex array (FFTW_BACKWARD, unnormalised) into a real
 * field; the half array is taken as the stored half of a Hermitian array,
 * which is what FFTW's c2r assumes (src/bubble.cpp:465-467,515). */
static void c2r_ld(int D, int N, const double *hre, const double *him, double *field)
{
    long long total = 1;
    for (int d = 0; d < D; ++d) total *= N;
    int nc = N / 2 + 1;
    long long rows = total / N;
    /* leading axes on the half array: treat as (D-1)-dim batch of nc columns */
    ld *re = malloc(sizeof(ld) * rows * nc), *im = malloc(sizeof(ld) * rows * nc);
    ld *tab = malloc(sizeof(ld) * 2 * N);
    twtab tw;
    make_twtab(N, tab, &tw);
    for (long long i = 0; i < rows * nc; ++i) { re[i] = hre[i]; im[i] = him[i]; }
    if (D >= 2) {
        for (int ax = 0; ax < D - 1; ++ax) {
            long long st = nc;
            for (int d = ax + 1; d < D - 1; ++d) st *= N;
            long long lines = rows * nc / N;
#pragma omp parallel
            {
                ld *work = malloc(sizeof(ld) * 6 * N);
#pragma omp for schedule(static)
                for (long long l = 0; l < lines; ++l) {
                    long long outer = l / st, inner = l % st;
                    long long base = outer * st * N + inner;
                    fft_line_ld(N, +1, re + base, im + base, st, work, &tw);
                }
                free(work);
            }
        }
    }
    /* last axis: Hermitian-extended complex line, real part of the inverse */
#pragma omp parallel
    {
        ld *lr = malloc(sizeof(ld) * N), *li = malloc(sizeof(ld) * N);
        ld *work = malloc(sizeof(ld) * 6 * N);
#pragma omp for schedule(static)
        for (long long r = 0; r < rows; ++r) {
            ld *xr = re + r * nc, *xi = im + r * nc;
            lr[0] = xr[0]; li[0] = 0;
            for (int j = 1; j < N; ++j) {
                if (2 * j < N)      { lr[j] = xr[j];     li[j] = xi[j]; }
                else if (2 * j == N) { lr[j] = xr[j];     li[j] = 0; }
                else                 { lr[j] = xr[N - j]; li[j] = -xi[N - j]; }
            }
            fft_line_ld(N, +1, lr, li, 1, work, &tw);
            for (int k = 0; k < N; ++k) field[r * N + k] = (double)lr[k];
        }
        free(lr); free(li); free(work);
    }
    free(tab); free(re); free(im);
}

/* ------------------------------------------------------------------ */
/* Thermal initial conditions, GenerateThermalFluc, src/bubble.cpp:459-523 */
/* ------------------------------------------------------------------ */
void cbo_thermal_modes(int D, int N, double L, double T, double m, int velocity,
                       uint64_t base, double *re, double *im)
{
    int nRows = (D == 3) ? N * N : N;                    /* :462 */
    int nc = N / 2 + 1;
    uint64_t seed = thermal_seed(base, velocity ? "velocitythermalspec" : "thermalspec");
#define KPHYS(i) (2.0 * M_PI * (double)(((i) <= N / 2) ? (i) : (i) - N) / L)   /* :469 */
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nRows; ++r) {                    /* :473-488 */
        mt64 g; mt64_seed(&g, seed + (uint64_t)r);
        polar st = {0., 0};
        double k2rows;
        if (D == 2) { double k = KPHYS(r); k2rows = k * k; }
        else { double k1 = KPHYS(r / N), k2 = KPHYS(r % N); k2rows = k1 * k1 + k2 * k2; }
        for (int j = 0; j < nc; ++j) {
            double kj = 2.0 * M_PI * j / L;
            double P = velocity ? velocitythermalspectrum(T, k2rows + kj * kj, m)
                                : thermalspectrum(T, k2rows + kj * kj, m);
            double sigma = sqrt(P / 2.);
            re[(long long)r * nc + j] = sigma * (polar_next(&st, &g) * 1.0 + 0.0);
            im[(long long)r * nc + j] = sigma * (polar_next(&st, &g) * 1.0 + 0.0);
        }
    }
#undef KPHYS
    /* Hermitian fix on planes j = 0 and j = N/2, :493-511 */
    int planes[2] = {0, N / 2};
    for (int q = 0; q < 2; ++q) {
        int j = planes[q];
        for (int r = 0; r < nRows; ++r) {
            int partner = (D == 2) ? (N - r) % N
                                   : ((N - r / N) % N) * N + (N - r % N) % N;
            long long id = (long long)r * nc + j;
            if (partner == r) { re[id] *= sqrt(2.); im[id] = 0.; }
            else if (partner < r) {
                re[id] = re[(long long)partner * nc + j];
                im[id] = -im[(long long)partner * nc + j];
            }
        }
    }
}

void cbo_thermal_field(int D, int N, double L, double T, double m, int velocity,
                       uint64_t base, double *field)
{
    int nRows = (D == 3) ? N * N : N;
    int nc = N / 2 + 1;
    double *re = malloc(sizeof(double) * (size_t)nRows * nc);
    double *im = malloc(sizeof(double) * (size_t)nRows * nc);
    cbo_thermal_modes(D, N, L, T, m, velocity, base, re, im);
    c2r_ld(D, N, re, im, field);
    long long total = (long long)nRows * N;
    double norm = 1.0 / sqrt(pow(L, D));                 /* :516 */
    for (long long i = 0; i < total; ++i) field[i] = field[i] * norm + 0.0;
    free(re); free(im);
}

/* ------------------------------------------------------------------ */
/* Nucleation points, GenerateNucleationPoints + check_PBC              */
/* (src/bubble.cpp:397-417, src/gridfuncs.cpp:93-119): text written     */
/* "%.5e" / "%.6e" and read back with "%lf" (src/bubble.cpp:431).       */
/* ------------------------------------------------------------------ */
static int append_line(char **buf, size_t *len, size_t *cap, const char *fmt,
                       double a, double b, double c)
{
    char line[128];
    int k = snprintf(line, sizeof line, fmt, a, b, c);
    if (*len + k + 1 > *cap) { *cap = (*cap + k + 1) * 2; *buf = realloc(*buf, *cap); }
    memcpy(*buf + *len, line, k + 1);
    *len += k;
    return k;
}

static char *xns_text(int D, double L, double c, int nBubbles, unsigned seed)
{
    struct random_data rd; char state[128];
    memset(&rd, 0, sizeof rd);
    initstate_r(seed, state, sizeof state, &rd);     /* same stream as srand(seed); rand() */
    double fit[3]; cbo_bubble_fit(D, c, fit);
    double Rcut = fit[1] + 5 * fit[2];               /* :414 */
    size_t len = 0, cap = 256; char *buf = malloc(cap); buf[0] = 0;
    for (int b = 0; b < nBubbles; ++b) {
        int32_t r0, r1, r2;
        random_r(&rd, &r0); random_r(&rd, &r1); random_r(&rd, &r2);
        double xn[3] = {L * (double)r0 / (double)RAND_MAX, L * (double)r1 / (double)RAND_MAX,
                        L * (double)r2 / (double)RAND_MAX};
        append_line(&buf, &len, &cap, "%.5e %.5e %.5e\n", xn[0], xn[1], xn[2]);
        int allowed[3] = {0, 0, 0};                  /* check_PBC :96-101 */
        for (int d = 0; d < D; ++d) {
            if (xn[d] + Rcut > L) allowed[d] = -1;
            else if (xn[d] - Rcut < 0) allowed[d] = +1;
        }
        for (int mask = 1; mask < (1 << D); ++mask) {  /* :104-118 */
            double im[3] = {xn[0], xn[1], xn[2]};
            int valid = 1;
            for (int d = 0; d < D; ++d)
                if (mask & (1 << d)) {
                    if (allowed[d] == 0) { valid = 0; break; }
                    im[d] += allowed[d] * L;
                }
            if (valid) append_line(&buf, &len, &cap, "%.6e %.6e %.6e\n", im[0], im[1], im[2]);
        }
    }
    return buf;
}

static int parse_xns(const char *text, double *xyz, int cap)
{
    int n = 0;
    const char *p = text;
    for (;;) {
        double v[3]; int k;
        if (sscanf(p, "%lf %lf %lf%n", &v[0], &v[1], &v[2], &k) != 3) break;
        if (n < cap) { xyz[3 * n] = v[0]; xyz[3 * n + 1] = v[1]; xyz[3 * n + 2] = v[2]; }
        ++n; p += k;
    }
    return n;
}

int cbo_generate_xns(int D, double L, double c, int nBubbles, unsigned seed, double *xyz, int cap)
{
    char *t = xns_text(D, L, c, nBubbles, seed);
    int n = parse_xns(t, xyz, cap);
    free(t);
    return n;
}

/* ------------------------------------------------------------------ */
/* Stepper                                                              */
/* ------------------------------------------------------------------ */
static int expanding(const cbo_sim *s) { return s->p.expansion != 0 && s->p.D == 3; }  /* bubble.hpp:77 */

static void forces(cbo_sim *s, double *dpi)                  /* src/bubble.cpp:716-743 */
{
    const int D = s->p.D, n = s->p.nGrids;
    const double *phi = s->phi;
    const double a = s->a, dx = s->dx, c = s->p.c;
    const int al = s->p.alpha;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (long long i = 0; i < s->N; ++i) {
        int ig = (int)i;
        double dd = 0.0;
        if (D == 1) {
            dd = phi[cbo_iGridM(D, n, ig, 1)] - 2.0 * phi[ig] + phi[cbo_iGridM(D, n, ig, 0)];
        } else {
            for (int sft = 0; sft < D; ++sft)
                dd += phi[cbo_iGridP(D, n, ig, sft)] - 2.0 * phi[ig] + phi[cbo_iGridM(D, n, ig, sft)];
        }
        dpi[i] = pow(a, D - 2 + al) * dd / (dx * dx) - pow(a, D + al) * cbo_Vpr(phi[ig], c);
    }
}

void cbo_forces(cbo_sim *s, double *dpi_out) { forces(s, dpi_out); }

static double addot(const cbo_sim *s)                        /* src/bubble.cpp:747-757 */
{
    if (!expanding(s)) return 0.0;
    const double a = s->a, Mpl = s->p.Mpl;
    if (s->p.alpha == 0) return -a * (2.0 * s->Ekintot - s->Epottot) / (3.0 * Mpl * Mpl);
    return a * a * a * (2 * s->Epottot - s->Ekintot + s->Egradtot) / (3.0 * Mpl * Mpl);
}

/* Neumaier-compensated running sum */
typedef struct { double s, c; } ksum;
static inline void ks_add(ksum *k, double x)
{
    double t = k->s + x;
    if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x;
    else k->c += (x - t) + k->s;
    k->s = t;
}
static inline void ks_merge(ksum *k, ksum o) { ks_add(k, o.s); ks_add(k, o.c); }
static inline double ks_val(ksum k) { return k.s + k.c; }

enum { E_KIN, E_POT, E_GRAD, E_PHI, E_PHIP, V_KIN, V_POT, V_GRAD, V_TOT, V_PHI, V_PHIP, E_NQ };

static void measure_energies(cbo_sim *s)                     /* src/bubble.cpp:761-844 */
{
    const int D = s->p.D, n = s->p.nGrids, nth = s->nthreads;
    const double *phi = s->phi, *pi = s->pi;
    const double a = s->a, dx = s->dx, dxLD = s->dxLD, c = s->p.c, DeltaV = s->DeltaV;
    const double a_kin = a + 0.5 * s->p.dt * s->ap;          /* :765 */
    ksum *part = calloc((size_t)nth * E_NQ, sizeof(ksum));
#pragma omp parallel num_threads(nth)
    {
        ksum acc[E_NQ] = {{0, 0}};
        int tid = omp_get_thread_num();
#pragma omp for schedule(static)
        for (long long i = 0; i < s->N; ++i) {
            int ig = (int)i;
            double gx = 0., gy = 0., gz = 0.;
            if (D == 1) gx = (phi[cbo_iGridM(D, n, ig, 1)] - phi[ig]) / (a * dx);
            if (D == 2 || D == 3) {
                gx = (phi[cbo_iGridP(D, n, ig, 1)] - phi[ig]) / (a * dx);
                gy = (phi[cbo_iGridP(D, n, ig, 0)] - phi[ig]) / (a * dx);
            }
            if (D == 3) gz = (phi[cbo_iGridP(D, n, ig, 2)] - phi[ig]) / (a * dx);
            double gx2 = gx * gx, gy2 = gy * gy, gz2 = gz * gz;
            double Eg = 0.5 * (gx2 + gy2 + gz2);
            double Ep = cbo_V(phi[ig], c, DeltaV);
            double Ek = 0.5 * pow(a_kin, -2 * D) * pi[ig] * pi[ig];
            ks_add(&acc[E_KIN], Ek * dxLD);
            ks_add(&acc[E_POT], Ep * dxLD);
            ks_add(&acc[E_GRAD], Eg * dxLD);
            ks_add(&acc[E_PHI], phi[ig] * dxLD);
            ks_add(&acc[E_PHIP], pi[ig] * pow(a_kin, -D) * dxLD);
            ks_add(&acc[V_KIN], Ek * dxLD * Ek * dxLD);
            ks_add(&acc[V_POT], Ep * dxLD * Ep * dxLD);
            ks_add(&acc[V_GRAD], Eg * dxLD * Eg * dxLD);
            ks_add(&acc[V_TOT], (Ek + Eg + Ep) * dxLD * (Ek + Eg + Ep) * dxLD);
            ks_add(&acc[V_PHI], phi[ig] * dxLD * phi[ig] * dxLD);
            ks_add(&acc[V_PHIP], pi[ig] * pow(a_kin, -D) * dxLD * pi[ig] * pow(a_kin, -D) * dxLD);
        }
        memcpy(part + (size_t)tid * E_NQ, acc, sizeof acc);
    }
    double tot[E_NQ];
    for (int q = 0; q < E_NQ; ++q) {
        ksum k = {0, 0};
        for (int t = 0; t < nth; ++t) ks_merge(&k, part[(size_t)t * E_NQ + q]);
        tot[q] = ks_val(k);
    }
    free(part);
    s->Epottot = tot[E_POT]; s->Egradtot = tot[E_GRAD];
    s->avphi = tot[E_PHI]; s->avphiP = tot[E_PHIP];
    s->EkintotVar = tot[V_KIN]; s->EpottotVar = tot[V_POT]; s->EgradtotVar = tot[V_GRAD];
    s->EtotVar = tot[V_TOT]; s->avphiVar = tot[V_PHI]; s->avphiPVar = tot[V_PHIP];
    double EkinHalf = tot[E_KIN];                            /* :839-843 */
    s->Ekintot = 0.5 * (EkinHalf + s->EkinHalfPrev);
    s->EkinHalfPrev = EkinHalf;
    s->Etot = s->Ekintot + s->Epottot + s->Egradtot;
}

/* The kinetic half-step sum of measure_energies alone (the same per-site
 * terms, thread split and compensated order as its E_KIN sum, so the same
 * double), chained into EkinHalfPrev as measure_energies chains it. */
static void measure_kinetic(cbo_sim *s)
{
    const int D = s->p.D, nth = s->nthreads;
    const double *pi = s->pi;
    const double dxLD = s->dxLD;
    const double a_kin = s->a + 0.5 * s->p.dt * s->ap;
    ksum *part = calloc((size_t)nth, sizeof(ksum));
#pragma omp parallel num_threads(nth)
    {
        ksum acc = {0, 0};
        int tid = omp_get_thread_num();
#pragma omp for schedule(static)
        for (long long i = 0; i < s->N; ++i) {
            double Ek = 0.5 * pow(a_kin, -2 * D) * pi[i] * pi[i];
            ks_add(&acc, Ek * dxLD);
        }
        part[tid] = acc;
    }
    ksum k = {0, 0};
    for (int t = 0; t < nth; ++t) ks_merge(&k, part[t]);
    free(part);
    s->EkinHalfPrev = ks_val(k);
}

static void one_step(cbo_sim *s, double *dpi)                /* src/bubble.cpp:849-876 */
{
    const double dt = s->p.dt;
    const int D = s->p.D, al = s->p.alpha, nth = s->nthreads;
    forces(s, dpi);
    double *pi = s->pi, *phi = s->phi;
#pragma omp parallel for schedule(static) num_threads(nth)
    for (long long i = 0; i < s->N; ++i) pi[i] += dt * dpi[i];
    s->nstep += 1;
    if (s->energy_every > 0 && !expanding(s)) {
        /* a static box: nothing feeds back from the energies (a'' = 0) */
        if (s->nstep % s->energy_every == 0) measure_energies(s);
        else if ((s->nstep + 1) % s->energy_every == 0) measure_kinetic(s);
    } else {
        measure_energies(s);
    }
    double ap_prev = s->ap;
    s->ap += dt * addot(s);
    double ap_atn = (ap_prev + s->ap) / 2.0;
    double a_old = s->a;
    s->a += dt * s->ap;
    double a_half = 0.5 * (s->a + a_old);
    double f = pow(a_half, -D + al);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (long long i = 0; i < s->N; ++i) phi[i] += dt * pi[i] * f;
    s->t += dt;
    s->H = ap_atn / a_old;
    s->C = expanding(s) ? s->H * s->H - pow(a_old, 2 * al) * s->Etot / (3.0 * s->p.Mpl * s->p.Mpl) : 0.0;
}

void cbo_step(cbo_sim *s, int nsteps)
{
    for (int it = 0; it < nsteps; ++it) one_step(s, s->dpi);
}

/* ------------------------------------------------------------------ */
/* Constructor, src/bubble.cpp:156-316                                  */
/* ------------------------------------------------------------------ */
static void seed_centres(cbo_sim *s, const double *xyz, int n)   /* :420-453 */
{
    const int D = s->p.D, g = s->p.nGrids;
    const double L = s->p.L;
    for (int b = 0; b < n; ++b) {
        const double *xn = xyz + 3 * b;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
        for (long long i = 0; i < s->N; ++i) {
            double x[3]; xfunc(D, g, L, i, x);
            double d0 = x[0] - xn[0], d1 = x[1] - xn[1], d2 = x[2] - xn[2];
            if (D == 2)
                s->phi[i] = dmin(s->phi[i] + thick_bubble(d0 * d0 + d1 * d1, s->phiC, s->Rn, s->lw), s->phiC);
            else if (D == 3)
                s->phi[i] = dmin(s->phi[i] + thick_bubble(d0 * d0 + d1 * d1 + d2 * d2, s->phiC, s->Rn, s->lw), s->phiC);
        }
    }
}

static void ctor_sequence(cbo_sim *s);

cbo_sim *cbo_create(const cbo_params *p, const double *xns, int n_xns, int nthreads)
{
    cbo_sim *s = calloc(1, sizeof *s);
    s->p = *p;
    s->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
    const int D = p->D, g = p->nGrids;
    const double L = p->L, c = p->c;
    s->N = (long long)pow(g, D);                           /* bubble.hpp:43-49 */
    s->n_cplx = g / 2 + 1;
    s->dx = (double)L / (double)g;
    s->dxD = pow(s->dx, D);
    s->dxLD = pow(s->dx / L, D);
    s->nkmax = (int)(sqrt((double)D) * g / 2);
    s->DeltaV = 1.;                                        /* :160-171 */
    s->phi0 = 1.;
    s->phiC = phizero(c);
    s->Rn = cradius(D, c);
    s->lw = clw(c);
    s->mT = sqrt(2 * (c + 6));
    s->phi = calloc(s->N, sizeof(double));
    s->pi = calloc(s->N, sizeof(double));
    s->dpi = calloc(s->N, sizeof(double));
    s->a = 1.; s->ap = 0.; s->H = 0.; s->C = 0.; s->t = 0.;

    if (D == 1) {                                          /* Initialize1D :324-335 */
        for (long long i = 0; i < g; ++i) {
            double x[3]; xfunc(D, g, L, i, x);
            double xxn = x[0];
            s->phi[i] = dmin(s->phi[i] + thick_bubble(xxn * xxn, s->phiC, s->Rn, s->lw), s->phiC);
            xxn = x[0] - L;
            s->phi[i] = dmin(s->phi[i] + thick_bubble(xxn * xxn, s->phiC, s->Rn, s->lw), s->phiC);
        }
    } else if (p->test_wave == 1) {                        /* InitializeWave :338-353 */
        double k = 2.0 * M_PI / (L / 40.0);
        double amplitude = 0.01 * s->phi0;
        for (long long i = 0; i < s->N; ++i) {
            double x[3]; xfunc(D, g, L, i, x);
            double w = 0.0;
            for (int d = 0; d < D; ++d) w += amplitude * sin(k * x[d]);
            s->phi[i] = s->phi0 + w / (double)D;
        }
    } else if (p->readxns == -2) {                         /* NucleateOneBubble :356-369 */
        for (long long i = 0; i < s->N; ++i) {
            double x[3]; xfunc(D, g, L, i, x);
            double e0 = x[0] - L / 2., e1 = x[1] - L / 2., e2 = x[2] - L / 2.;
            double rn = (D == 2) ? (e0 * e0 + e1 * e1) : (e0 * e0 + e1 * e1 + e2 * e2);
            s->phi[i] = dmin(s->phi[i] + thick_bubble(rn, s->phiC, s->Rn, s->lw), s->phiC);
        }
    } else if (p->readxns == -1) {                         /* NucleateTwoDiagonalBubbles :373-393 */
        const double DR = p->DeltaR;
        for (long long i = 0; i < s->N; ++i) {
            double x[3]; xfunc(D, g, L, i, x);
            double a0 = x[0] - L / 4 - DR, a1 = x[1] - L / 4. - DR, a2 = x[2] - L / 2.;
            double b0 = x[0] - 3. * L / 4. + DR, b1 = x[1] - 3. * L / 4 + DR, b2 = x[2] - L / 2.;
            double r1 = (D == 2) ? (a0 * a0 + a1 * a1) : (a0 * a0 + a1 * a1 + a2 * a2);
            double r2 = (D == 2) ? (b0 * b0 + b1 * b1) : (b0 * b0 + b1 * b1 + b2 * b2);
            s->phi[i] = dmin(s->phi[i] + thick_bubble(r1, s->phiC, s->Rn, s->lw), s->phiC);
            s->phi[i] = dmin(s->phi[i] + thick_bubble(r2, s->phiC, s->Rn, s->lw), s->phiC);
        }
    } else if (p->readxns == 1 || p->readxns == 0) {       /* :205-213 */
        if (p->readxns == 1) {
            char *t = xns_text(D, L, c, p->nBubbles, p->rand_seed);
            int n = parse_xns(t, NULL, 0);
            s->xns = malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
            s->n_xns = parse_xns(t, s->xns, n);
            free(t);
        } else {
            s->xns = malloc(sizeof(double) * 3 * (n_xns > 0 ? n_xns : 1));
            memcpy(s->xns, xns, sizeof(double) * 3 * n_xns);
            s->n_xns = n_xns;
        }
        seed_centres(s, s->xns, s->n_xns);
    } else if (p->readxns == -3) {                         /* :214-229 */
        double Temp = p->Trad * s->mT;
        double mF = sqrt(2 * c);
        cbo_thermal_field(D, g, L, Temp, mF, 0, p->thermal_base, s->phi);
        cbo_thermal_field(D, g, L, Temp, mF, 1, p->thermal_base, s->pi);
    }
    ctor_sequence(s);
    return s;
}

/* The ctor's sequence after the fields are set (src/bubble.cpp:239-291):
 * initial energies, H0, a', the pi back-step and EkinHalfPrev. */
static void ctor_sequence(cbo_sim *s)
{
    const cbo_params *p = &s->p;
    const int D = p->D, g = p->nGrids;
    const double c = p->c;
    s->a = 1.; s->ap = 0.; s->H = 0.; s->C = 0.; s->t = 0.;
    /* initial energies, :239-269 */
    {
        const int nth = s->nthreads;
        ksum *part = calloc((size_t)nth * 5, sizeof(ksum));
        const double dx = s->dx, dxLD = s->dxLD;
#pragma omp parallel num_threads(nth)
        {
            ksum acc[5] = {{0, 0}};
            int tid = omp_get_thread_num();
#pragma omp for schedule(static)
            for (long long i = 0; i < s->N; ++i) {
                int ig = (int)i;
                double tmp = 0;
                if (D == 1) {
                    double q = (s->phi[cbo_iGridM(D, g, ig, 1)] - s->phi[ig]) / dx;
                    tmp += q * q;
                } else {
                    for (int sft = 0; sft < D; ++sft) {
                        double q = (s->phi[cbo_iGridP(D, g, ig, sft)] - s->phi[ig]) / dx;
                        tmp += q * q;
                    }
                }
                double egrad = 0.5 * tmp;
                double epot = cbo_V(s->phi[ig], c, s->DeltaV);
                double ekin = 0.5 * s->pi[ig] * s->pi[ig];
                ks_add(&acc[0], s->phi[ig] * dxLD);
                ks_add(&acc[1], s->pi[ig] * dxLD);
                ks_add(&acc[2], epot * dxLD);
                ks_add(&acc[3], egrad * dxLD);
                ks_add(&acc[4], ekin * dxLD);
            }
            memcpy(part + (size_t)tid * 5, acc, sizeof acc);
        }
        double tot[5];
        for (int q = 0; q < 5; ++q) {
            ksum k = {0, 0};
            for (int t = 0; t < nth; ++t) ks_merge(&k, part[(size_t)t * 5 + q]);
            tot[q] = ks_val(k);
        }
        free(part);
        s->avphi = tot[0]; s->avphiP = tot[1]; s->Epottot = tot[2];
        s->Egradtot = tot[3]; s->Ekintot = tot[4];
        s->Etot = s->Ekintot + s->Epottot + s->Egradtot;
    }
    /* expansion and pi back-step, :273-291 */
    s->a = 1;
    double H0 = expanding(s) ? sqrt(pow(s->a, 2 * p->alpha) * s->Etot / (3.0 * p->Mpl * p->Mpl)) : 0.0;
    s->ap = s->a * H0;
    s->ap -= 0.5 * p->dt * addot(s);
    forces(s, s->dpi);
    double a_mhalf = s->a - 0.5 * p->dt * s->ap;
    {
        const int nth = s->nthreads;
        ksum *part = calloc((size_t)nth, sizeof(ksum));
        const double dt = p->dt, dxLD = s->dxLD;
#pragma omp parallel num_threads(nth)
        {
            ksum acc = {0, 0};
            int tid = omp_get_thread_num();
#pragma omp for schedule(static)
            for (long long i = 0; i < s->N; ++i) {
                s->pi[i] -= 0.5 * dt * s->dpi[i];
                ks_add(&acc, 0.5 * pow(a_mhalf, -2 * D) * s->pi[i] * s->pi[i] * dxLD);
            }
            part[tid] = acc;
        }
        ksum k = {0, 0};
        for (int t = 0; t < nth; ++t) ks_merge(&k, part[t]);
        double EkinHalf = ks_val(k);
        free(part);
        s->EkinHalfPrev = EkinHalf;
    }
    s->H = H0;
}

void cbo_ctor(cbo_sim *s) { ctor_sequence(s); }

/* phi after seeding the centres xyz[0..n) in order (src/bubble.cpp:420-453)
 * at the flat lattice indices idx[0..m), without building the lattice. */
void cbo_seed_sites(const cbo_params *p, const double *xyz, int n, const long long *idx, int m, double *out)
{
    const int D = p->D, g = p->nGrids;
    const double L = p->L;
    const double phiC = phizero(p->c), Rn = cradius(D, p->c), lw = clw(p->c);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < m; ++q) {
        double x[3]; xfunc(D, g, L, idx[q], x);
        double phi = 0.0;
        for (int b = 0; b < n; ++b) {
            const double *xn = xyz + 3 * b;
            double d0 = x[0] - xn[0], d1 = x[1] - xn[1], d2 = x[2] - xn[2];
            if (D == 2) phi = dmin(phi + thick_bubble(d0 * d0 + d1 * d1, phiC, Rn, lw), phiC);
            else if (D == 3) phi = dmin(phi + thick_bubble(d0 * d0 + d1 * d1 + d2 * d2, phiC, Rn, lw), phiC);
        }
        out[q] = phi;
    }
}

cbo_sim *cbo_create_blank(const cbo_params *p, int nthreads)
{
    cbo_sim *s = calloc(1, sizeof *s);
    s->p = *p;
    s->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
    const int D = p->D, g = p->nGrids;
    s->N = (long long)pow(g, D);
    s->n_cplx = g / 2 + 1;
    s->dx = (double)p->L / (double)g;
    s->dxD = pow(s->dx, D);
    s->dxLD = pow(s->dx / p->L, D);
    s->nkmax = (int)(sqrt((double)D) * g / 2);
    s->DeltaV = 1.; s->phi0 = 1.;
    s->phiC = phizero(p->c); s->Rn = cradius(D, p->c); s->lw = clw(p->c);
    s->mT = sqrt(2 * (p->c + 6));
    s->phi = calloc(s->N, sizeof(double));
    s->pi = calloc(s->N, sizeof(double));
    s->dpi = calloc(s->N, sizeof(double));
    s->a = 1.;
    return s;
}

void cbo_destroy(cbo_sim *s)
{
    if (!s) return;
    free(s->phi); free(s->pi); free(s->dpi); free(s->xns);
    free(s);
}

/* outEP columns, src/bubble.cpp:977-989 (spreads are printed as sqrt) */
void cbo_energy_row(const cbo_sim *s, double row[16])
{
    row[0] = s->t; row[1] = s->a; row[2] = s->Etot; row[3] = s->Ekintot;
    row[4] = s->Egradtot; row[5] = s->Epottot;
    row[6] = sqrt(s->EtotVar); row[7] = sqrt(s->EkintotVar);
    row[8] = sqrt(s->EgradtotVar); row[9] = sqrt(s->EpottotVar);
    row[10] = s->avphi; row[11] = sqrt(s->avphiVar);
    row[12] = s->avphiP; row[13] = sqrt(s->avphiPVar);
    row[14] = s->H; row[15] = s->C;
}

/* FourierTransform + ComputePowerSpectrum, src/bubble.cpp:642-709 */
int cbo_spectrum(cbo_sim *s, double *kmean, double *kphys, double *delta, int cap)
{
    const int D = s->p.D, g = s->p.nGrids, nc = s->n_cplx;
    const double L = s->p.L;
    long long nRows = s->N / g;
    double *re = malloc(sizeof(double) * nRows * nc), *im = malloc(sizeof(double) * nRows * nc);
    cbo_r2c(D, g, s->phi, re, im);
    const int nBins = s->nkmax + 1;
    const double dk = 2.0 * M_PI / L;
    const double prefactor = (D == 1) ? 1.0 / M_PI : (D == 2) ? 1.0 / (2.0 * M_PI) : 1.0 / (2.0 * M_PI * M_PI);
    const double Pnorm = s->dxD * s->dxD / pow(L, D);
    double *pw = calloc(nBins, sizeof(double)), *ws = calloc(nBins, sizeof(double)),
           *ks = calloc(nBins, sizeof(double));
#define SIGNED(i) (((i) <= g / 2) ? (i) : (i) - g)
    for (long long r = 0; r < nRows; ++r) {
        double n2rows = 0.;
        if (D == 2) { double q = SIGNED((int)r); n2rows = q * q; }
        if (D == 3) { double q1 = SIGNED((int)(r / g)), q2 = SIGNED((int)(r % g)); n2rows = q1 * q1 + q2 * q2; }
        for (int j = 0; j < nc; ++j) {
            double n = sqrt(n2rows + (double)j * j);
            int bin = (int)n;
            if (bin >= nBins) continue;
            double w = (j == 0 || j == g / 2) ? 1.0 : 2.0;
            double f0 = re[r * nc + j], f1 = im[r * nc + j];
            double kc = n * dk;
            pw[bin] += w * prefactor * pow(kc, D) * Pnorm * (f0 * f0 + f1 * f1);
            ws[bin] += w;
            ks[bin] += w * kc;
        }
    }
#undef SIGNED
    int nout = 0;
    for (int b = 0; b < nBins; ++b)
        if (ws[b] > 0) {
            double km = ks[b] / ws[b];
            if (nout < cap) { kmean[nout] = km; kphys[nout] = km / s->a; delta[nout] = pw[b] / ws[b]; }
            ++nout;
        }
    free(pw); free(ws); free(ks); free(re); free(im);
    return nout;
}

long long cbo_nsites(const cbo_sim *s) { return s->N; }
double *cbo_phi(cbo_sim *s) { return s->phi; }
double *cbo_pi(cbo_sim *s) { return s->pi; }
void cbo_set_threads(cbo_sim *s, int n) { s->nthreads = n > 0 ? n : omp_get_max_threads(); }
void cbo_set_energy_cadence(cbo_sim *s, int every, long long nstep) { s->energy_every = every; s->nstep = nstep; }

int cbo_xns(const cbo_sim *s, double *xyz, int cap)
{
    for (int i = 0; i < s->n_xns && i < cap; ++i) memcpy(xyz + 3 * i, s->xns + 3 * i, 3 * sizeof(double));
    return s->n_xns;
}

void cbo_get_scalars(const cbo_sim *s, double out[6])
{
    out[0] = s->a; out[1] = s->ap; out[2] = s->H; out[3] = s->C; out[4] = s->t; out[5] = s->EkinHalfPrev;
}

void cbo_set_scalars(cbo_sim *s, const double in[6])
{
    s->a = in[0]; s->ap = in[1]; s->H = in[2]; s->C = in[3]; s->t = in[4]; s->EkinHalfPrev = in[5];
}

void cbo_energy_maps(const cbo_sim *s, const double *phi, const double *pi, double a, double a_kin, int ctor,
                     double *ekin, double *egrad, double *epot)
{
    const int D = s->p.D, n = s->p.nGrids, nth = s->nthreads;
    const double dx = s->dx, c = s->p.c, DeltaV = s->DeltaV;
#pragma omp parallel for schedule(static) num_threads(nth)
    for (long long i = 0; i < s->N; ++i) {
        int ig = (int)i;
        if (ctor) {                                              /* src/bubble.cpp:239-265 */
            double tmp = 0;
            if (D == 1) {
                tmp += pow((phi[cbo_iGridM(D, n, ig, 1)] - phi[ig]) / dx, 2);
            } else {
                for (int sft = 0; sft < D; ++sft) tmp += pow((phi[cbo_iGridP(D, n, ig, sft)] - phi[ig]) / dx, 2);
            }
            egrad[i] = 0.5 * tmp;
            epot[i] = cbo_V(phi[ig], c, DeltaV);
            ekin[i] = 0.5 * pi[ig] * pi[ig];
        } else {                                                 /* :790-833 */
            double gx = 0., gy = 0., gz = 0.;
            if (D == 1) gx = (phi[cbo_iGridM(D, n, ig, 1)] - phi[ig]) / (a * dx);
            if (D == 2 || D == 3) {
                gx = (phi[cbo_iGridP(D, n, ig, 1)] - phi[ig]) / (a * dx);
                gy = (phi[cbo_iGridP(D, n, ig, 0)] - phi[ig]) / (a * dx);
            }
            if (D == 3) gz = (phi[cbo_iGridP(D, n, ig, 2)] - phi[ig]) / (a * dx);
            egrad[i] = 0.5 * (gx * gx + gy * gy + gz * gz);
            epot[i] = cbo_V(phi[ig], c, DeltaV);
            ekin[i] = 0.5 * pow(a_kin, -2 * D) * pi[ig] * pi[ig];
        }
    }
}
