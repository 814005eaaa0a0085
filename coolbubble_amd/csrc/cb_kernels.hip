// cb_kernels.hip -- general stepper, constructor, reduction and seeding
// kernels for gfx950 (CDNA4, 64-wide wavefronts).
//
// The general step kernel runs one lattice site per thread per grid-stride
// iteration and reads its 2D+1 taps straight from global memory (the L1/L2
// caches absorb the reuse).  It serves every (D, N, precision); the
// LDS-tiled marching kernels in cb_step_tiled.hip take the hot shapes.
#include "cb_kernels.hpp"
#include "cb_device.hpp"

namespace cb {

// ---------------------------------------------------------------------------
// General fused step: force (src/bubble.cpp:716-743), kick (:854-856),
// energies (:761-844, when ENERGY) and drift (:868-870).
//  static box  : reads phi_n, pi_{n-1/2}; writes pi_{n+1/2} (in place) and
//                phi_{n+1} = phi_n + dt pi_{n+1/2} (a_half = 1).
//  LAGGED      : stored phi lags by one drift: each tap is recomputed as
//                phi_n = phi_{n-1} + (dt pi_{n-1/2}) s_{n-1} (the reference's
//                drift of the previous step); writes phi_n and pi_{n+1/2};
//                the drift of this step follows once a_half is known.
// ---------------------------------------------------------------------------
template <typename T, int D, bool LAGGED, int ENERGY>
__global__ void __launch_bounds__(256)
k_step_simple(SimConst k, const T *__restrict__ phi_in, T *__restrict__ phi_out, const T *pi_in,
              T *pi_out, const DevScalars *__restrict__ sc, double *__restrict__ partials)
{
    using C = T;
    const C dt = (C)k.dt;
    const C s_prev = LAGGED ? (C)sc->s_pending : (C)0;
    const C fac_lap = (C)sc->fac_lap, fac_pot = (C)sc->fac_pot;
    const StepF32 sf = step_f32(k, sc);
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;

    auto tap = [&](long long idx) -> C {
        C v = (C)phi_in[idx];
        if constexpr (LAGGED) v = v + dt * (C)pi_in[idx] * s_prev;
        return v;
    };

    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < k.nsites; s += stride) {
        long long nb[6];
        neighbours<D>(s, k.N, nb);
        const C ph = tap(s);
        C P[3], M[3];
#pragma unroll
        for (int a = 0; a < D; ++a) { P[a] = tap(nb[2 * a]); M[a] = tap(nb[2 * a + 1]); }
        const C po = (C)pi_in[s];
        C pn;
        if constexpr (sizeof(T) == 8) {
            C dd;
            if constexpr (D == 1) {
                dd = P[0] - (C)2.0 * ph + M[0];
            } else {
                dd = (C)0.0;
#pragma unroll
                for (int a = 0; a < D; ++a) dd += P[a] - (C)2.0 * ph + M[a];
            }
            const C F = lap_term(fac_lap * dd, k) - fac_pot * vprime(ph, k);
            pn = po + dt * F;
            pi_out[s] = (T)pn;
            if constexpr (LAGGED) phi_out[s] = (T)ph;
            else phi_out[s] = (T)(ph + dt * pn);
        } else {
            // axis 0, the fastest axis, then axis 1 of 3D
            pn = kick_f32<D>(ph, P[0], M[0], P[D - 1], M[D - 1], P[D == 3 ? 1 : 0], M[D == 3 ? 1 : 0], po, sf);
            pi_out[s] = pn;
            if constexpr (LAGGED) phi_out[s] = ph;
            else phi_out[s] = __builtin_fmaf(sf.dt, pn, ph);
        }
        if constexpr (ENERGY) {
            double fwd[3] = {(double)P[0], D >= 2 ? (double)P[1] : 0.0, D == 3 ? (double)P[2] : 0.0};
            energy_terms<D, sizeof(T) == 8, ENERGY == 2>((double)ph, fwd, (double)pn, (double)po, k, sc, acc);
        }
    }
    if constexpr (ENERGY) energy_partials(acc, k, sc, partials + (long long)blockIdx.x * NQ);
}

// ---------------------------------------------------------------------------
// Constructor: initial energies (src/bubble.cpp:239-269, a different
// gradient formula from measureEnergies: /dx, axes in order 0..D-1).
// Slots: 0 <phi>, 1 <pi>, 2 Epot, 3 Egrad, 4 Ekin.
// ---------------------------------------------------------------------------
template <typename T, int D>
__global__ void __launch_bounds__(256)
k_ctor_energy(SimConst k, const T *__restrict__ phi, const T *__restrict__ pi, double *partials)
{
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < k.nsites; s += stride) {
        long long nb[6];
        neighbours<D>(s, k.N, nb);
        const double ph = (double)phi[s], p = (double)pi[s];
        double tmp = 0;
        if constexpr (D == 1) {
            double q = ((double)phi[nb[0]] - ph) / k.dx;
            tmp += q * q;
        } else {
#pragma unroll
            for (int a = 0; a < D; ++a) {
                double q = ((double)phi[nb[2 * a]] - ph) / k.dx;
                tmp += q * q;
            }
        }
        double egrad = 0.5 * tmp;
        double epot = vpot(ph, k);
        double ekin = 0.5 * p * p;
        acc[0] += ph * k.dxLD;
        acc[1] += p * k.dxLD;
        acc[2] += epot * k.dxLD;
        acc[3] += egrad * k.dxLD;
        acc[4] += ekin * k.dxLD;
    }
    block_partials<NQ>(acc, partials + (long long)blockIdx.x * NQ);
}

// pi_{-1/2} = pi_0 - (dt/2) F(phi_0) with a = 1, and its kinetic energy
// (src/bubble.cpp:280-289).  Slot 0 holds the sum.
template <typename T, int D>
__global__ void __launch_bounds__(256)
k_backstep(SimConst k, const T *__restrict__ phi, T *pi, double pk2, double *partials)
{
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < k.nsites; s += stride) {
        long long nb[6];
        neighbours<D>(s, k.N, nb);
        const double ph = (double)phi[s];
        double dd;
        if constexpr (D == 1) {
            dd = (double)phi[nb[0]] - 2.0 * ph + (double)phi[nb[1]];
        } else {
            dd = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) dd += (double)phi[nb[2 * a]] - 2.0 * ph + (double)phi[nb[2 * a + 1]];
        }
        double F = 1.0 * dd / k.dx2 - 1.0 * vprime(ph, k);
        double p = (double)pi[s] - 0.5 * k.dt * F;
        pi[s] = (T)p;
        p = (double)(T)p;
        acc[0] += 0.5 * pk2 * p * p * k.dxLD;
    }
    block_partials<NQ>(acc, partials + (long long)blockIdx.x * NQ);
}

__global__ void __launch_bounds__(256) k_sum_partials(const double *partials, int nblocks, double *tot)
{
    sum_partials_block(partials, nblocks, tot);
}

// Energies of the step and the scalar chain of oneStep (finalize_member in
// cb_device.hpp), one simulation.
__global__ void __launch_bounds__(256)
k_finalize(SimConst k, const double *partials, int nblocks, DevScalars *sc, int use_old_kin)
{
    finalize_member(k, partials, nblocks, sc, use_old_kin);
}

template <typename T>
__global__ void __launch_bounds__(256) k_apply_drift(long long n, double dt, T *phi, const T *pi, DevScalars *sc)
{
    const T s = (T)sc->s_pending;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        phi[i] = phi[i] + (T)dt * pi[i] * s;
}

__global__ void k_clear_pending(DevScalars *sc) { sc->s_pending = 0.0; }

// ---------------------------------------------------------------------------
// Seeding, src/bubble.cpp:324-453 (profile ThickBubble :102-108, coordinates
// xFunc src/gridfuncs.cpp:77-85).
// ---------------------------------------------------------------------------
template <typename T, int D>
__global__ void __launch_bounds__(256)
k_seed_bubbles(SimConst k, T *phi, const double *__restrict__ cen, const double *__restrict__ sub,
               int ncen, double L, double phiC, double Rn, double lw)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    const int N = k.N;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < k.nsites; s += stride) {
        long long idx[3] = {0, 0, 0};
        if constexpr (D == 1) idx[0] = s;
        else if constexpr (D == 2) { idx[0] = s / N; idx[1] = s - idx[0] * N; }
        else { long long NN = (long long)N * N; idx[0] = s / NN; long long r = s - idx[0] * NN; idx[1] = r / N; idx[2] = r - idx[1] * N; }
        double x[3] = {0., 0., 0.};
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = L * ((double)idx[a] + 0.5) / (double)N;
        double v = (double)phi[s];
        for (int c = 0; c < ncen; ++c) {
            double r2 = 0.0;
            double d0 = x[0] - cen[3 * c] - sub[3 * c];
            if constexpr (D == 1) r2 = d0 * d0;
            else {
                double d1 = x[1] - cen[3 * c + 1] - sub[3 * c + 1];
                if constexpr (D == 2) r2 = d0 * d0 + d1 * d1;
                else {
                    double d2 = x[2] - cen[3 * c + 2] - sub[3 * c + 2];
                    r2 = d0 * d0 + d1 * d1 + d2 * d2;
                }
            }
            double prof = (phiC / 2.) * (tanh((Rn - sqrt(r2)) / lw) + 1.);
            double w = v + prof;
            v = (phiC < w) ? phiC : w;
        }
        phi[s] = (T)v;
    }
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
k_seed_wave(SimConst k, T *phi, double L, double kw, double amp, double phi0)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    const int N = k.N;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < k.nsites; s += stride) {
        long long idx[3] = {0, 0, 0};
        if constexpr (D == 2) { idx[0] = s / N; idx[1] = s - idx[0] * N; }
        else { long long NN = (long long)N * N; idx[0] = s / NN; long long r = s - idx[0] * NN; idx[1] = r / N; idx[2] = r - idx[1] * N; }
        double w = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double x = L * ((double)idx[a] + 0.5) / (double)N;
            w += amp * sin(kw * x);
        }
        phi[s] = (T)(phi0 + w / (double)D);
    }
}

template <typename S, typename Dd>
__global__ void __launch_bounds__(256) k_convert(const S *src, Dd *dst, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = (Dd)src[i];
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
#define CB_DISPATCH_D(D_, ...)                                  \
    switch (D_) {                                               \
    case 1: { constexpr int DD = 1; __VA_ARGS__; break; }       \
    case 2: { constexpr int DD = 2; __VA_ARGS__; break; }       \
    default: { constexpr int DD = 3; __VA_ARGS__; break; }      \
    }

template <typename T>
static void step_simple_t(const SimConst &k, int lagged, int energy, const StepPtrs &p,
                          const DevScalars *sc, double *partials, SimpleCfg cfg, hipStream_t st)
{
    const T *pin = (const T *)p.phi_in;
    T *pout = (T *)p.phi_out;
    const T *qin = (const T *)p.pi_in;
    T *qout = (T *)p.pi_out;
    dim3 g(cfg.grid), b(cfg.block);
    CB_DISPATCH_D(k.D,
        if (lagged) {
            if (energy && k.energy_lite) k_step_simple<T, DD, true, 2><<<g, b, 0, st>>>(k, pin, pout, qin, qout, sc, partials);
            else if (energy) k_step_simple<T, DD, true, true><<<g, b, 0, st>>>(k, pin, pout, qin, qout, sc, partials);
            else k_step_simple<T, DD, true, false><<<g, b, 0, st>>>(k, pin, pout, qin, qout, sc, partials);
        } else {
            if (energy) k_step_simple<T, DD, false, true><<<g, b, 0, st>>>(k, pin, pout, qin, qout, sc, partials);
            else k_step_simple<T, DD, false, false><<<g, b, 0, st>>>(k, pin, pout, qin, qout, sc, partials);
        })
}

hipError_t launch_step_simple(const SimConst &k, int fp32, int lagged, int energy, const StepPtrs &p,
                              const DevScalars *sc, double *partials, SimpleCfg cfg, hipStream_t st)
{
    if (fp32) step_simple_t<float>(k, lagged, energy, p, sc, partials, cfg, st);
    else step_simple_t<double>(k, lagged, energy, p, sc, partials, cfg, st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Several 1D steps per launch (temporal blocking; force, kick and drift of
// src/bubble.cpp:716-743, :854-856, :868-870 per step).  A 1D lattice of 2^16
// sites is 0.5 MiB a field: one launch per step
// leaves the GPU idle between launches.  Block b owns sites
// [b kSeg1d, b kSeg1d + seg) and loads them with `steps` sites either side
// (periodic) into LDS; step t updates window positions [t, W - t), whose
// neighbours are still exact after t - 1 steps; the owned sites are written
// after the last step.  The site update is the expression of k_step_simple's
// non-lagged 1D case in the same translation unit, so the fields equal
// `steps` one-step launches bit for bit.  No energies: a call's last step
// runs on k_step_simple with them.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads1d)
k_step1d_multi(SimConst k, const T *__restrict__ phi_in, T *__restrict__ phi_out, const T *__restrict__ pi_in,
               T *__restrict__ pi_out, const DevScalars *__restrict__ sc, int steps)
{
    using C = T;
    constexpr int WMAX = kSeg1d + 2 * kMaxSteps1d;
    constexpr int R = (WMAX + kThreads1d - 1) / kThreads1d;
    __shared__ T sphi[WMAX], spi[WMAX];
    const C dt = (C)k.dt;
    const C fac_lap = (C)sc->fac_lap, fac_pot = (C)sc->fac_pot;
    const StepF32 sf = step_f32(k, sc);
    const long long N = k.N;
    const long long base = (long long)blockIdx.x * kSeg1d;
    const int seg = (int)(N - base < kSeg1d ? N - base : kSeg1d);
    const int W = seg + 2 * steps;
    for (int j = threadIdx.x; j < W; j += kThreads1d) {
        long long g = (base - steps + j) % N;
        if (g < 0) g += N;
        sphi[j] = phi_in[g];
        spi[j] = pi_in[g];
    }
    __syncthreads();
    for (int t = 1; t <= steps; ++t) {
        T nphi[R], npi[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = t + threadIdx.x + r * kThreads1d;
            if (j < W - t) {
                const C ph = sphi[j], P0 = sphi[j + 1], M0 = sphi[j - 1], po = spi[j];
                C pn;
                if constexpr (sizeof(T) == 8) {
                    const C dd = P0 - (C)2.0 * ph + M0;
                    const C F = lap_term(fac_lap * dd, k) - fac_pot * vprime(ph, k);
                    pn = po + dt * F;
                    npi[r] = (T)pn;
                    nphi[r] = (T)(ph + dt * pn);
                } else {
                    pn = kick_f32<1>(ph, P0, M0, P0, M0, P0, M0, po, sf);
                    npi[r] = pn;
                    nphi[r] = __builtin_fmaf(sf.dt, pn, ph);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = t + threadIdx.x + r * kThreads1d;
            if (j < W - t) {
                sphi[j] = nphi[r];
                spi[j] = npi[r];
            }
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < seg; j += kThreads1d) {
        phi_out[base + j] = sphi[steps + j];
        pi_out[base + j] = spi[steps + j];
    }
}

hipError_t launch_step1d_multi(const SimConst &k, int fp32, const StepPtrs &p, const DevScalars *sc, int steps,
                               hipStream_t st)
{
    if (k.D != 1 || steps < 1 || steps > kMaxSteps1d || p.pi_in == p.pi_out || p.phi_in == p.phi_out)
        return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((k.nsites + kSeg1d - 1) / kSeg1d);
    if (fp32)
        k_step1d_multi<float><<<grid, kThreads1d, 0, st>>>(k, (const float *)p.phi_in, (float *)p.phi_out,
                                                          (const float *)p.pi_in, (float *)p.pi_out, sc, steps);
    else
        k_step1d_multi<double><<<grid, kThreads1d, 0, st>>>(k, (const double *)p.phi_in, (double *)p.phi_out,
                                                           (const double *)p.pi_in, (double *)p.pi_out, sc, steps);
    return hipGetLastError();
}

hipError_t launch_ctor_energy(const SimConst &k, int fp32, const void *phi, const void *pi,
                              double *partials, SimpleCfg cfg, hipStream_t st)
{
    dim3 g(cfg.grid), b(cfg.block);
    if (fp32) { CB_DISPATCH_D(k.D, k_ctor_energy<float, DD><<<g, b, 0, st>>>(k, (const float *)phi, (const float *)pi, partials)) }
    else { CB_DISPATCH_D(k.D, k_ctor_energy<double, DD><<<g, b, 0, st>>>(k, (const double *)phi, (const double *)pi, partials)) }
    return hipGetLastError();
}

hipError_t launch_backstep(const SimConst &k, int fp32, const void *phi, void *pi, double pk2,
                           double *partials, SimpleCfg cfg, hipStream_t st)
{
    dim3 g(cfg.grid), b(cfg.block);
    if (fp32) { CB_DISPATCH_D(k.D, k_backstep<float, DD><<<g, b, 0, st>>>(k, (const float *)phi, (float *)pi, pk2, partials)) }
    else { CB_DISPATCH_D(k.D, k_backstep<double, DD><<<g, b, 0, st>>>(k, (const double *)phi, (double *)pi, pk2, partials)) }
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double *partials, int nblocks, double *tot, hipStream_t st)
{
    k_sum_partials<<<1, 256, 0, st>>>(partials, nblocks, tot);
    return hipGetLastError();
}

hipError_t launch_finalize(const SimConst &k, const double *partials, int nblocks, DevScalars *sc,
                           int use_old_kin, hipStream_t st)
{
    k_finalize<<<1, 256, 0, st>>>(k, partials, nblocks, sc, use_old_kin);
    return hipGetLastError();
}

hipError_t launch_apply_drift(const SimConst &k, int fp32, void *phi, const void *pi, DevScalars *sc,
                              SimpleCfg cfg, hipStream_t st)
{
    dim3 g(cfg.grid), b(cfg.block);
    if (fp32) k_apply_drift<float><<<g, b, 0, st>>>(k.nsites, k.dt, (float *)phi, (const float *)pi, sc);
    else k_apply_drift<double><<<g, b, 0, st>>>(k.nsites, k.dt, (double *)phi, (const double *)pi, sc);
    k_clear_pending<<<1, 1, 0, st>>>(sc);
    return hipGetLastError();
}

hipError_t launch_seed_bubbles(const SimConst &k, int fp32, void *phi, const double *cen,
                               const double *sub, int ncen, double L, double phiC, double Rn, double lw,
                               SimpleCfg cfg, hipStream_t st)
{
    dim3 g(cfg.grid), b(cfg.block);
    if (fp32) { CB_DISPATCH_D(k.D, k_seed_bubbles<float, DD><<<g, b, 0, st>>>(k, (float *)phi, cen, sub, ncen, L, phiC, Rn, lw)) }
    else { CB_DISPATCH_D(k.D, k_seed_bubbles<double, DD><<<g, b, 0, st>>>(k, (double *)phi, cen, sub, ncen, L, phiC, Rn, lw)) }
    return hipGetLastError();
}

hipError_t launch_seed_wave(const SimConst &k, int fp32, void *phi, double L, double kwave,
                            double amplitude, double phi0, SimpleCfg cfg, hipStream_t st)
{
    dim3 g(cfg.grid), b(cfg.block);
    if (k.D == 2) {
        if (fp32) k_seed_wave<float, 2><<<g, b, 0, st>>>(k, (float *)phi, L, kwave, amplitude, phi0);
        else k_seed_wave<double, 2><<<g, b, 0, st>>>(k, (double *)phi, L, kwave, amplitude, phi0);
    } else {
        if (fp32) k_seed_wave<float, 3><<<g, b, 0, st>>>(k, (float *)phi, L, kwave, amplitude, phi0);
        else k_seed_wave<double, 3><<<g, b, 0, st>>>(k, (double *)phi, L, kwave, amplitude, phi0);
    }
    return hipGetLastError();
}

hipError_t launch_convert(int to_double, const void *src, void *dst, long long n, hipStream_t st)
{
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    if (to_double) k_convert<float, double><<<(unsigned)blocks, 256, 0, st>>>((const float *)src, (double *)dst, n);
    else k_convert<double, float><<<(unsigned)blocks, 256, 0, st>>>((const double *)src, (float *)dst, n);
    return hipGetLastError();
}

} // namespace cb
