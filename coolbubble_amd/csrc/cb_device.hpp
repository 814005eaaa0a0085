// cb_device.hpp -- device helpers shared by the kernel files: block
// reductions, compensated sums, lattice neighbours and the per-site energy
// terms of measureEnergies.
#pragma once

#include "cb_common.hpp"

namespace cb {

// ---------------------------------------------------------------------------
// Reductions: 64-lane wavefront shuffles, then the waves of the block in
// order through LDS.  Deterministic for a fixed launch shape.
// ---------------------------------------------------------------------------
template <int NV>
__device__ inline void block_partials(double (&acc)[NV], double *out)
{
    __shared__ double sm[NV][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double v = acc[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sm[q][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += sm[threadIdx.x][i];
        out[threadIdx.x] = s;
    }
}

// Neumaier-compensated running sum.
struct Neu {
    double s, c;
    __device__ void add(double x)
    {
        double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ void merge(Neu o)
    {
        add(o.s);
        c += o.c;
    }
    __device__ double value() const { return s + c; }
};

// Sum nblocks x NQ partials (row-major) into tot[0, NV) (the first NV
// quantities of each row); one block of 256.
template <int NV = NQ>
__device__ inline void sum_partials_block(const double *partials, int nblocks, double *tot)
{
    __shared__ double ss[NV][256];
    __shared__ double sc[NV][256];
    const int t = threadIdx.x;
    Neu n[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) n[q] = {0.0, 0.0};
    // Each thread adds rows t, t + 256, t + 512, ... in that order; the rows
    // are loaded U at a time so that their latencies overlap (one block
    // alone on the GPU has nothing else to hide them behind): 16 when a row
    // holds few quantities (the Friedmann sums of an expanding step), 4
    // for all of them.
    constexpr int U = NV <= 4 ? 16 : 4;
    const int stride = blockDim.x;
    int b = t;
    for (; b + (U - 1) * stride < nblocks; b += U * stride) {
        double v[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < NV; ++q) v[u][q] = partials[(long long)(b + u * stride) * NQ + q];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < NV; ++q) n[q].add(v[u][q]);
    }
    for (; b < nblocks; b += stride)
#pragma unroll
        for (int q = 0; q < NV; ++q) n[q].add(partials[(long long)b * NQ + q]);
#pragma unroll
    for (int q = 0; q < NV; ++q) { ss[q][t] = n[q].s; sc[q][t] = n[q].c; }
    __syncthreads();
    for (int half = blockDim.x >> 1; half > 0; half >>= 1) {
        if (t < half) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                Neu a = {ss[q][t], sc[q][t]};
                a.merge(Neu{ss[q][t + half], sc[q][t + half]});
                ss[q][t] = a.s;
                sc[q][t] = a.c;
            }
        }
        __syncthreads();
    }
    if (t < NV) tot[t] = ss[t][0] + sc[t][0];
}

// ---------------------------------------------------------------------------
// Lattice neighbours, src/gridfuncs.cpp:31-51: nb = {P0, M0, P1, M1, P2, M2},
// axis 0 slowest.  In 1D nb[0] is the next point and nb[1] the previous one.
// ---------------------------------------------------------------------------
template <int D>
__device__ inline void neighbours(long long s, int N, long long (&nb)[6])
{
    if constexpr (D == 1) {
        int i = (int)s;
        nb[0] = (i + 1 == N) ? 0 : i + 1;
        nb[1] = (i == 0) ? N - 1 : i - 1;
    } else if constexpr (D == 2) {
        long long i = s / N;
        int j = (int)(s - i * N);
        long long ip = (i + 1 == N) ? 0 : i + 1, im = (i == 0) ? N - 1 : i - 1;
        int jp = (j + 1 == N) ? 0 : j + 1, jm = (j == 0) ? N - 1 : j - 1;
        nb[0] = ip * N + j;
        nb[1] = im * N + j;
        nb[2] = i * N + jp;
        nb[3] = i * N + jm;
    } else {
        long long NN = (long long)N * N;
        long long i = s / NN;
        long long r = s - i * NN;
        int j = (int)(r / N);
        int kk = (int)(r - (long long)j * N);
        long long ip = (i + 1 == N) ? 0 : i + 1, im = (i == 0) ? N - 1 : i - 1;
        int jp = (j + 1 == N) ? 0 : j + 1, jm = (j == 0) ? N - 1 : j - 1;
        int kp = (kk + 1 == N) ? 0 : kk + 1, km = (kk == 0) ? N - 1 : kk - 1;
        long long base = (long long)j * N + kk;
        nb[0] = ip * NN + base;
        nb[1] = im * NN + base;
        nb[2] = i * NN + (long long)jp * N + kk;
        nb[3] = i * NN + (long long)jm * N + kk;
        nb[4] = i * NN + (long long)j * N + kp;
        nb[5] = i * NN + (long long)j * N + km;
    }
}

// Energy terms of one site (measureEnergies, src/bubble.cpp:790-827), in
// double.  gP = {forward neighbour along axis 1, 0, 2} per the reference's
// gx, gy, gz naming (1D: gx only).  EXACT: the FP64 path.
// LITE: only the sums the Friedmann update reads (src/bubble.cpp:747-757):
// kinetic, potential and gradient (read in conformal time only, but summed
// whatever alpha is: a branch on alpha here raised the one-step FP32 tile
// from 78 to 151 VGPRs, three waves per SIMD instead of six).
template <int D, bool EXACT, bool LITE = false>
__device__ inline void energy_terms(double ph, const double (&fwd)[3], double pn, double po,
                                    const SimConst &k, const DevScalars *sc, double (&acc)[NQ])
{
    // The gradient divides by a dx through its reciprocal (within 1 ulp of the
    // reference's quotient; its terms are positive, so the sums keep that).
    // The potential is a difference of O(c) terms that cancel near phi = 0
    // and phi = 1, so the FP64 path keeps the reference's correctly rounded
    // powers; the FP32 path, whose fields carry 1e-7 errors, uses products.
    const double iadx = sc->inv_adx, pk2 = sc->pk2, pk1 = sc->pk1;
    if constexpr (LITE) {
        if constexpr (EXACT) {
            acc[Q_POT] += vpot(ph, k);
        } else {
            const double ph2 = ph * ph;
            acc[Q_POT] += ((k.c * ph2 - k.c5 * (ph2 * ph)) + k.c6 * (ph2 * ph2)) + k.DeltaV;
        }
        acc[Q_KIN] += 0.5 * pk2 * pn * pn;
        {
            double g0 = 0., g1 = 0., g2 = 0.;
            if constexpr (D == 1) {
                g0 = (fwd[0] - ph) * iadx;
            } else {
                g0 = (fwd[1] - ph) * iadx;
                g1 = (fwd[0] - ph) * iadx;
                if constexpr (D == 3) g2 = (fwd[2] - ph) * iadx;
            }
            acc[Q_GRAD] += 0.5 * ((g0 * g0 + g1 * g1) + g2 * g2);
        }
        return;
    }
    double gx = 0., gy = 0., gz = 0.;
    if constexpr (D == 1) {
        gx = (fwd[0] - ph) * iadx;
    } else {
        gx = (fwd[1] - ph) * iadx;
        gy = (fwd[0] - ph) * iadx;
        if constexpr (D == 3) gz = (fwd[2] - ph) * iadx;
    }
    double gx2 = gx * gx, gy2 = gy * gy, gz2 = gz * gz;
    double Eg = 0.5 * (gx2 + gy2 + gz2);
    double Ep;
    if constexpr (EXACT) {
        Ep = vpot(ph, k);
    } else {
        const double ph2 = ph * ph;
        Ep = ((k.c * ph2 - k.c5 * (ph2 * ph)) + k.c6 * (ph2 * ph2)) + k.DeltaV;
    }
    double Ek = 0.5 * pk2 * pn * pn;
    double Et = Ek + Eg + Ep;
    double pp = pn * pk1;
    // The reference weights every term by the site volume dxLD (dxLD^2 for
    // the squares) and the old kinetic term by 0.5 pk2 as it adds it; these
    // factors are the same for every site of a step, so the sums here are
    // unweighted and energy_partials applies them once per thread.
    acc[Q_KIN] += Ek;
    acc[Q_POT] += Ep;
    acc[Q_GRAD] += Eg;
    acc[Q_PHI] += ph;
    acc[Q_PHIP] += pp;
    acc[Q_VKIN] += Ek * Ek;
    acc[Q_VPOT] += Ep * Ep;
    acc[Q_VGRAD] += Eg * Eg;
    acc[Q_VTOT] += Et * Et;
    acc[Q_VPHI] += ph * ph;
    acc[Q_VPHIP] += pp * pp;
    acc[Q_KIN_OLD] += po * po;
}

// Weight a thread's energy_terms sums (src/bubble.cpp:790-827 weights each
// site's terms) and write the block's partials.
__device__ inline void energy_partials(double (&acc)[NQ], const SimConst &k, const DevScalars *sc, double *out)
{
    const double v = k.dxLD, v2 = v * v;
    acc[Q_KIN] *= v;
    acc[Q_POT] *= v;
    acc[Q_GRAD] *= v;
    acc[Q_PHI] *= v;
    acc[Q_PHIP] *= v;
    acc[Q_VKIN] *= v2;
    acc[Q_VPOT] *= v2;
    acc[Q_VGRAD] *= v2;
    acc[Q_VTOT] *= v2;
    acc[Q_VPHI] *= v2;
    acc[Q_VPHIP] *= v2;
    acc[Q_KIN_OLD] *= 0.5 * sc->pk2 * v;
    block_partials<NQ>(acc, out);
}

// Energies of the step and the scalar chain of oneStep, src/bubble.cpp:839-875.
__device__ inline void finalize_member(const SimConst &k, const double *partials, int nblocks, DevScalars *sc,
                                       int use_old_kin)
{
    __shared__ double tot[NQ];
    if (k.energy_lite) sum_partials_block<3>(partials, nblocks, tot);   // Q_KIN, Q_POT, Q_GRAD
    else sum_partials_block<NQ>(partials, nblocks, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    DevScalars s = *sc;
    const double EkinHalf = tot[Q_KIN];
    if (use_old_kin) s.EkinHalfPrev = tot[Q_KIN_OLD];
    s.Ekintot = 0.5 * (EkinHalf + s.EkinHalfPrev);
    s.EkinHalfPrev = EkinHalf;
    s.Epottot = tot[Q_POT];
    s.Egradtot = tot[Q_GRAD];
    if (!k.energy_lite) {   // a lite step's row is never read: its other sums are left as they were
        s.avphi = tot[Q_PHI];
        s.avphiP = tot[Q_PHIP];
        s.EkintotVar = tot[Q_VKIN];
        s.EpottotVar = tot[Q_VPOT];
        s.EgradtotVar = tot[Q_VGRAD];
        s.EtotVar = tot[Q_VTOT];
        s.avphiVar = tot[Q_VPHI];
        s.avphiPVar = tot[Q_VPHIP];
    }
    s.Etot = s.Ekintot + s.Epottot + s.Egradtot;
    // addot, src/bubble.cpp:747-757
    double addot = 0.0;
    if (k.expanding) {
        if (k.alpha == 0) addot = -s.a * (2.0 * s.Ekintot - s.Epottot) / (3.0 * k.Mpl * k.Mpl);
        else addot = s.a * s.a * s.a * (2 * s.Epottot - s.Ekintot + s.Egradtot) / (3.0 * k.Mpl * k.Mpl);
    }
    const double ap_prev = s.ap;
    s.ap += k.dt * addot;
    const double ap_atn = (ap_prev + s.ap) / 2.0;
    const double a_old = s.a;
    s.a += k.dt * s.ap;
    const double a_half = 0.5 * (s.a + a_old);
    s.s_pending = ipow_rn(a_half, -k.D + k.alpha);
    s.H = ap_atn / a_old;
    s.C = k.expanding ? s.H * s.H - ipow_rn(a_old, 2 * k.alpha) * s.Etot / (3.0 * k.Mpl * k.Mpl) : 0.0;
    // factors of the next step's force and energies (src/bubble.cpp:741-742,765)
    s.fac_lap = ipow_rn(s.a, k.D - 2 + k.alpha);
    s.fac_pot = ipow_rn(s.a, k.D + k.alpha);
    s.adx = s.a * k.dx;
    s.inv_adx = 1.0 / s.adx;
    const double a_kin = s.a + 0.5 * k.dt * s.ap;
    s.pk2 = ipow_rn(a_kin, -2 * k.D);
    s.pk1 = ipow_rn(a_kin, -k.D);
    if (!isfinite(s.Etot)) s.nan_flag = 1;
    *sc = s;
}

} // namespace cb
