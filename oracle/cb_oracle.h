/*
 * cb_oracle.h -- CPU FP64 restatement of CoolBubble's lattice time-stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * nowhere else.  The product (coolbubble_amd, libcoolbubble.so) never links,
 * loads or calls it.
 *
 * Every function restates the reference algorithm in plain C, in the
 * reference's operation order (left-to-right C++ evaluation, no FMA), citing
 * the reference file:line it follows.  It is built with gcc -O2
 * -ffp-contract=off and calls glibc pow/tanh/exp/log/sqrt where the
 * reference does, so its per-site arithmetic is the reference's.
 *
 * The FFTs (reference: FFTW3 r2c / c2r, src/bubble.cpp:465-467,515,648-651)
 * are an exact mathematical DFT evaluated in long double (mixed radix,
 * recursive), rounded to double; FFTW itself is absent from the image.
 */
#ifndef CB_ORACLE_H
#define CB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters: the reference's process globals (src/gridfuncs.h:27-32)
 * and compile-time constants (src/def.h:29-76) as one struct. */
typedef struct cbo_params {
    int D;                 /* src/def.h:30   */
    int nGrids;            /* src/def.h:31   */
    double L;              /* src/def.h:32   */
    double c;              /* src/def.h:39   */
    int readxns;           /* src/def.h:40   initial-condition mode */
    double dt;             /* src/def.h:34   TIME_STEP */
    int nBubbles;          /* src/def.h:46   */
    double Trad;           /* src/def.h:48   */
    double DeltaR;         /* src/def.h:50   */
    int test_wave;         /* src/def.h:51   TEST */
    int expansion;         /* src/def.h:56   */
    int alpha;             /* src/def.h:57   ALPHAVAL */
    double Mpl;            /* src/def.h:58   */
    unsigned int rand_seed;      /* srand(42), src/bubble.cpp:401 */
    uint64_t thermal_base;       /* 12345, src/bubble.cpp:137 */
} cbo_params;

typedef struct cbo_sim cbo_sim;

/* ---- whole-run API (Bubble ctor / oneStep / outEP / spectrum) ---- */
cbo_sim *cbo_create(const cbo_params *p, const double *xns, int n_xns, int nthreads);
/* A lattice with the ctor's derived constants and zero fields, a = 1, no
 * initial condition (for timing the stepper from a state set by the caller). */
cbo_sim *cbo_create_blank(const cbo_params *p, int nthreads);
/* Run the ctor's sequence after seeding (initial energies, H0, a', pi
 * back-step, EkinHalfPrev; src/bubble.cpp:239-291) on the current phi and
 * the raw t = 0 pi of s. */
void cbo_ctor(cbo_sim *s);
/* phi after seeding the n centres xyz (modes 0 / 1) at the flat indices
 * idx[0..m), without building the lattice. */
void cbo_seed_sites(const cbo_params *p, const double *xyz, int n, const long long *idx, int m, double *out);
void cbo_destroy(cbo_sim *s);
void cbo_step(cbo_sim *s, int nsteps);
void cbo_energy_row(const cbo_sim *s, double row[16]);
int cbo_spectrum(cbo_sim *s, double *kmean, double *kphys, double *delta, int cap);
long long cbo_nsites(const cbo_sim *s);
double *cbo_phi(cbo_sim *s);
double *cbo_pi(cbo_sim *s);
int cbo_xns(const cbo_sim *s, double *xyz, int cap);
/* scalars: a, ap, H, C, t, EkinHalfPrev */
void cbo_get_scalars(const cbo_sim *s, double out[6]);
void cbo_set_scalars(cbo_sim *s, const double in[6]);
void cbo_set_threads(cbo_sim *s, int nthreads);
/* Static box: measure the energies only on steps n (counting from nstep)
 * with n % every == 0, and on the step before each the kinetic half-step
 * sum they need (every = 0: every step).  The rows read after those steps
 * are the every-step rows bit for bit; the fields do not depend on it.
 * Expanding runs measure every step whatever the setting. */
void cbo_set_energy_cadence(cbo_sim *s, int every, long long nstep);

/* ---- pieces, exported for the pinning tests ---- */
double cbo_V(double phi, double c, double DeltaV);
double cbo_Vpr(double phi, double c);
int cbo_iGridM(int D, int nGrids, int iGrid, int iShift);
int cbo_iGridP(int D, int nGrids, int iGrid, int iShift);
void cbo_bubble_fit(int D, double c, double out[3]); /* phiC, Rn, lw */
int cbo_generate_xns(int D, double L, double c, int nBubbles, unsigned seed,
                     double *xyz, int cap);
void cbo_mt64(uint64_t seed, long n, uint64_t *out);
void cbo_normals(uint64_t seed, long n, double *out);
void cbo_dft_ld(int n, int sign, const double *in_re, const double *in_im,
                double *out_re, double *out_im);
void cbo_r2c(int D, int N, const double *field, double *out_re, double *out_im);
void cbo_thermal_modes(int D, int N, double L, double T, double m, int velocity,
                       uint64_t base, double *re, double *im);
void cbo_thermal_field(int D, int N, double L, double T, double m, int velocity,
                       uint64_t base, double *field);
void cbo_forces(cbo_sim *s, double *dpi_out);
/* Local energy densities per site for the HDF5 snapshots: ctor = 1 the ctor's
 * (src/bubble.cpp:239-265, a = 1, pi at t = 0), ctor = 0 measureEnergies'
 * (:790-833) with the scale factors a and a_kin.  phi, pi: N^D values. */
void cbo_energy_maps(const cbo_sim *s, const double *phi, const double *pi, double a, double a_kin, int ctor,
                     double *ekin, double *egrad, double *epot);

#ifdef __cplusplus
}
#endif
#endif
