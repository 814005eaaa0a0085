// cb_common.hpp -- shared definitions of the MI355X lattice time-stepper.
//
// Device-side arithmetic follows the reference's evaluation order
// (src/bubble.cpp:69-77, 716-876).  All HIP sources are compiled with
// -ffp-contract=off, so a*b+c stays two roundings as in the reference's
// g++ -O2 build (no FMA on x86-64 without -march).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cb {

// ---------------------------------------------------------------------------
// Energy accumulators of one step (measureEnergies, src/bubble.cpp:776-836),
// plus the kinetic energy of the incoming pi (the previous half step).
// ---------------------------------------------------------------------------
enum EnergyQ {
    Q_KIN = 0,   // sum Ek dxLD       at pi_{n+1/2}
    Q_POT,       // sum V dxLD
    Q_GRAD,      // sum Eg dxLD
    Q_PHI,       // sum phi dxLD
    Q_PHIP,      // sum pi a_kin^-D dxLD
    Q_VKIN,      // spreads: sum (e dxLD) e dxLD, src/bubble.cpp:815-820
    Q_VPOT,
    Q_VGRAD,
    Q_VTOT,
    Q_VPHI,
    Q_VPHIP,
    Q_KIN_OLD,   // sum Ek dxLD at pi_{n-1/2} (static box: EkinHalfPrev)
    NQ
};

// Per-simulation scalars kept on the device, so the expanding-background
// chain (energies -> a'' -> a_{n+1/2} -> drift) never waits on the host.
struct DevScalars {
    double a, ap, H, C, EkinHalfPrev;
    double s_pending;       // drift factor a_half^(-D+alpha) not yet applied to phi
    double fac_lap;         // a^(D-2+alpha) for the next force evaluation
    double fac_pot;         // a^(D+alpha)
    double adx;             // a * dx
    double inv_adx;         // 1 / (a dx), for the energies' gradients
    double pk2, pk1;        // a_kin^(-2D), a_kin^(-D)
    double Etot, Ekintot, Epottot, Egradtot, avphi, avphiP;
    double EtotVar, EkintotVar, EpottotVar, EgradtotVar, avphiVar, avphiPVar;
    int nan_flag;
    int pad_;
};

// Constants of one simulation, passed by value to every kernel.
struct SimConst {
    long long nsites;       // N^D
    int N, D;
    int alpha, expanding;
    double dt, dx, dx2, dxLD;
    double inv_dx2;         // 1 / dx^2 when dx^2 is a power of two (the product is then the quotient), else 0
    double c, DeltaV;
    double c2, c3, c4;      // V' coefficients 2c, 3(2c+4), 4(c+3)
    double c5, c6;          // V coefficients 2c+4, c+3
    double Mpl;
    // 1: this launch measures only what the Friedmann update needs (the
    // kinetic and potential sums, the gradient sum in conformal time); an
    // expanding run's steps before the last of a call, whose rows nobody
    // reads, are measured so
    int energy_lite;
};

// ---------------------------------------------------------------------------
// Double-double helpers for correctly rounded small integer powers.  glibc's
// pow(x,3), pow(x,4) and pow(a, n) (src/bubble.cpp:70,76,741-742,813,870)
// are within 0.52 ulp of the exact value; these return the correctly rounded
// value (exact product carried in two doubles, one final rounding), so the
// two agree except where glibc misrounds.
// ---------------------------------------------------------------------------
struct ddbl { double hi, lo; };

__host__ __device__ inline ddbl two_prod(double a, double b)
{
    double p = a * b;
    double e = __builtin_fma(a, b, -p);
    return {p, e};
}

__host__ __device__ inline ddbl quick_two_sum(double a, double b)
{
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}

__host__ __device__ inline ddbl dd_mul_d(ddbl x, double b)
{
    ddbl p = two_prod(x.hi, b);
    double lo = p.lo + x.lo * b;
    return quick_two_sum(p.hi, lo);
}

__host__ __device__ inline ddbl dd_mul(ddbl x, ddbl y)
{
    ddbl p = two_prod(x.hi, y.hi);
    double lo = p.lo + (x.hi * y.lo + x.lo * y.hi);
    return quick_two_sum(p.hi, lo);
}

__host__ __device__ inline double cube_rn(double x)
{
    return dd_mul_d(two_prod(x, x), x).hi;
}

__host__ __device__ inline double quart_rn(double x)
{
    ddbl x2 = two_prod(x, x);
    return dd_mul(x2, x2).hi;
}

// x^n for |n| <= 16, correctly rounded (to within 2^-100 relative before the
// final rounding).
__host__ __device__ inline double ipow_rn(double x, int n)
{
    if (n == 0) return 1.0;
    int m = n < 0 ? -n : n;
    ddbl r = {x, 0.0};
    for (int i = 1; i < m; ++i) r = dd_mul_d(r, x);
    if (n > 0) return r.hi;
    // 1 / (hi + lo): one Newton correction in double-double
    double q1 = 1.0 / r.hi;
    double t = __builtin_fma(-q1, r.hi, 1.0);
    t = t - q1 * r.lo;
    double q2 = t / r.hi;
    return q1 + q2;
}

// V'(phi), Vepspr, src/bubble.cpp:74-77, in the reference's order.
__device__ inline double vprime(double phi, const SimConst &k)
{
    return (k.c2 * phi - k.c3 * (phi * phi)) + k.c4 * cube_rn(phi);
}

// V(phi), Veps, src/bubble.cpp:69-72.
__device__ inline double vpot(double phi, const SimConst &k)
{
    return ((k.c * (phi * phi) - k.c5 * cube_rn(phi)) + k.c6 * quart_rn(phi)) + k.DeltaV;
}

// (a^(D-2+alpha) dd) / dx^2 as the reference divides (src/bubble.cpp:741);
// a multiplication by the exact reciprocal when dx^2 is a power of two
// (e.g. L = 64, N = 512), which gives the same double.
__device__ inline double lap_term(double fl_dd, const SimConst &k)
{
    return k.inv_dx2 != 0.0 ? fl_dd * k.inv_dx2 : fl_dd / k.dx2;
}

// FP32 path: the force of one site from its neighbours, with fused
// multiply-adds (13 operations per site-update instead of 24; the FP64 path
// keeps the reference's separately rounded operations).  a = the pair along
// axis 0, b = along the fastest axis, c = along axis 1 of a 3D lattice; every
// kernel sums them in this order, so all variants agree bit for bit.
// FP32 leapfrog constants of one step.  The kick is written so that each
// site costs 2D - 1 adds and five fused multiply-adds:
//   g = s - x (A + x (B + x C)),  s = the sum of the 2D neighbours,
//   pi' = pi + P g,               phi' = phi + dt pi',
// with P = dt lapc, A = 2D + fp c2 / lapc, B = -fp c3 / lapc, C = fp c4 / lapc
// (lapc = a^(D-2+alpha) / dx^2, fp = a^(D+alpha)): P g is dt times the
// force lapc (s - 2D x) - fp V'(x), V'(x) = x (c2 + x (-c3 + x c4)).
struct StepF32 {
    float P, A, B, C;
    float dt;
};

__device__ inline StepF32 step_f32(const SimConst &k, const DevScalars *sc)
{
    const double lapc = sc->fac_lap / k.dx2, fp = sc->fac_pot;
    return StepF32{(float)(k.dt * lapc), (float)(2.0 * k.D + fp * k.c2 / lapc), (float)(-fp * k.c3 / lapc),
                   (float)(fp * k.c4 / lapc), (float)k.dt};
}

// pi after the kick of one site (FP32): x the field, aP/aM, bP/bM, cP/cM its
// neighbours along the lattice axes (added in this order in every kernel,
// so that all FP32 kernels agree bit for bit), pi the incoming momentum.
template <int D>
__device__ inline float kick_f32(float x, float aP, float aM, float bP, float bM, float cP, float cM, float pi,
                                 const StepF32 &f)
{
    float s = aP + aM;
    if constexpr (D >= 2) s = s + (bP + bM);
    if constexpr (D == 3) s = s + (cP + cM);
    const float t = __builtin_fmaf(__builtin_fmaf(f.C, x, f.B), x, f.A);
    const float g = __builtin_fmaf(-x, t, s);
    return __builtin_fmaf(f.P, g, pi);
}

} // namespace cb
