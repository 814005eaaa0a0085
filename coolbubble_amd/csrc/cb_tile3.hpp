// cb_tile3.hpp -- one tile of the three-steps-per-pass leapfrog for a static
// box, 2D and 3D (internal; included by cb_step3.hip).
//
// The deeper form of the two-step tile (cb_tile2.hpp): in a static box step
// n+1 at a site needs only step n at its neighbours, so a block marching the
// slow axis i can chain three steps plane by plane:
//   step A (n   -> n+1) on plane p   for TR+4 rows (two extra each side),
//   step B (n+1 -> n+2) on plane p-1 for TR+2 rows,
//   step C (n+2 -> n+3) on plane p-2 for the TR owned rows,
// each from the previous level's planes p-1, p, p+1 (relative to its own
// plane) in registers and its plane in LDS for the j and k taps:
//   L0 = phi_n     plane p,   rows r0-3 .. r0+TR+2 (TR+6 rows)
//   L1 = phi_n+1   plane p-1, rows r0-2 .. r0+TR+1 (TR+4 rows, step A's)
//   L2 = phi_n+2   plane p-2, rows r0-1 .. r0+TR   (TR+2 rows, step B's)
// The pass reads phi_n, pi_{n-1/2} and writes phi_{n+3}, pi_{n+5/2}: one
// pass of 16 B (FP32) per site for three site-updates, plus the overlap
// (rows and planes +-3 per tile and chunk) that L2 mostly serves.  Each
// site's arithmetic is that of three single steps (src/bubble.cpp:716-743,
// :854-856, :868-870), so results are bit-identical to the one-step kernels.
// Energies, when asked for, are those of step C with the kinetic energy of
// its incoming pi_{n+3/2} (the time-centred Ekin).  pi ping-pongs.  2D: one
// row, the three levels on rows p, p-1, p-2 of the marching axis.
// PF = 2 issues the HBM loads of the next plane (phi_n at p+2, pi at p+1,
// the outer L0 rows of p+1) one plane ahead, so their latency overlaps a
// plane of arithmetic instead of stalling the block at every plane (12 more
// VGPRs per vector item).
#pragma once

#include <type_traits>

#include "cb_device.hpp"
#include "cb_tile_common.hpp"
#include "cb_tiled.hpp"

namespace cb {

// One leapfrog step of VW adjacent sites: c the field, mM / mP its i-1 / i+1
// values, lrow its row in LDS (j taps at lrow -+ N in 3D, k taps at the
// vector's edges), piv the incoming pi.  Writes the kicked pi and drifted phi.
template <typename T, int D, int VW, bool ENERGY>
__device__ __forceinline__ void leap_vec(const SimConst &k, const Vec<T, VW> &c, const Vec<T, VW> &mM, const Vec<T, VW> &mP,
                                const T *lrow, int col, const Vec<T, VW> &piv, T fac_lap, T fac_pot,
                                const StepF32 &sf, T dt, Vec<T, VW> &pout, Vec<T, VW> &phout,
                                const DevScalars *sc, double (&acc)[NQ])
{
    using V = Vec<T, VW>;
    using C = T;
    const int N = k.N;
    V rm, rp;
    if constexpr (D == 3) {
        rm = *reinterpret_cast<const V *>(lrow - N + col);
        rp = *reinterpret_cast<const V *>(lrow + N + col);
    }
    const C fl = lrow[col == 0 ? N - 1 : col - 1];
    const C fr = lrow[(col + VW == N) ? 0 : col + VW];
#pragma unroll
    for (int e = 0; e < VW; ++e) {
        const C x = c.v[e];
        const C fM = (e == 0) ? fl : c.v[e - 1];
        const C fP = (e == VW - 1) ? fr : c.v[e + 1];
        C pnew;
        if constexpr (sizeof(T) == 8) {
            C dd = (C)0.0;
            dd += mP.v[e] - (C)2.0 * x + mM.v[e];
            if constexpr (D == 3) dd += rp.v[e] - (C)2.0 * x + rm.v[e];
            dd += fP - (C)2.0 * x + fM;
            const C F = lap_term(fac_lap * dd, k) - fac_pot * vprime(x, k);
            pnew = piv.v[e] + dt * F;
            phout.v[e] = x + dt * pnew;
        } else {
            C rP = 0, rM = 0;
            if constexpr (D == 3) { rP = rp.v[e]; rM = rm.v[e]; }
            pnew = kick_f32<D>(x, mP.v[e], mM.v[e], fP, fM, rP, rM, piv.v[e], sf);
            phout.v[e] = __builtin_fmaf(sf.dt, pnew, x);
        }
        pout.v[e] = pnew;
        if constexpr (ENERGY) {
            double fwd[3];
            fwd[0] = (double)mP.v[e];
            if constexpr (D == 3) { fwd[1] = (double)rp.v[e]; fwd[2] = (double)fP; }
            else { fwd[1] = (double)fP; fwd[2] = 0.0; }
            energy_terms<D, sizeof(T) == 8>((double)x, fwd, (double)pnew, (double)piv.v[e], k, sc, acc);
        }
    }
}

template <typename T, int D, int ITEMS, int NTHR, bool ENERGY, int PF = 1>
__device__ __forceinline__ void tile_step3(const SimConst &k, const TiledGeom &g, int lt, const T *__restrict__ phi_in,
                                  T *__restrict__ phi_out, const T *__restrict__ pi_in, T *__restrict__ pi_out,
                                  const DevScalars *__restrict__ sc, double *__restrict__ partials, T *lds)
{
    constexpr int VW = sizeof(T) == 4 ? 4 : 2;
    constexpr int H = (D == 3) ? 1 : 0;           // extra rows per level on each side
    using V = Vec<T, VW>;

    const int N = k.N;
    const int NV = g.NV;
    const long long NN = (D == 3) ? (long long)N * N : (long long)N;
    const int TRv = (D == 3) ? g.TR : 1;
    const int rt = (D == 3) ? lt % g.ntr : 0;
    const int ic = (D == 3) ? lt / g.ntr : lt;
    const int r0 = rt * TRv;
    const int i0 = ic * g.NI;
    const int i1 = min(i0 + g.NI, N);
    const int rowsA = TRv + 4 * H;
    const int nvecA = rowsA * NV;
    T *L0 = lds;                                  // phi_n   plane p,   rows r0-3H .. r0+TR-1+3H
    T *L1 = L0 + (size_t)(TRv + 6 * H) * N;       // phi_n+1 plane p-1, rows r0-2H .. r0+TR-1+2H
    T *L2 = L1 + (size_t)(TRv + 4 * H) * N;       // phi_n+2 plane p-2, rows r0-H  .. r0+TR-1+H
    const bool nt = g.nt != 0;

    auto wrapN = [&](int i) { return i < 0 ? i + N : (i >= N ? i - N : i); };

    const T dt = (T)k.dt;
    const T fac_lap = (T)sc->fac_lap, fac_pot = (T)sc->fac_pot;
    const StepF32 sf = step_f32(k, sc);

    // item u: row rowA of the step-A rows (tile rows r0-2H+rowA), vector col
    int rowA[ITEMS], col[ITEMS];
    bool ok[ITEMS], inB[ITEMS], inC[ITEMS];
    long long off[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
        const int q = threadIdx.x + u * NTHR;
        ok[u] = q < nvecA;
        const int qq = ok[u] ? q : 0;
        rowA[u] = qq / NV;
        col[u] = (qq - rowA[u] * NV) * VW;
        const int grow = (D == 3) ? wrapN(r0 - 2 * H + rowA[u]) : 0;
        off[u] = (D == 3) ? (long long)grow * N + col[u] : (long long)col[u];
        inB[u] = ok[u] && rowA[u] >= H && rowA[u] < 3 * H + TRv;
        inC[u] = ok[u] && rowA[u] >= 2 * H && rowA[u] < 2 * H + TRv;
    }
    // outer rows of L0 (3D): rows r0-3 and r0+TR+2 at L0 rows 0 and TR+5
    bool hok[ITEMS];
    long long hoff[ITEMS];
    int hlds[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
        const int h = threadIdx.x + u * NTHR;
        hok[u] = (D == 3) && h < 2 * NV;
        const int hh = hok[u] ? h : 0;
        const int side = hh >= NV;
        const int c0 = (hh - side * NV) * VW;
        const int grow = side ? wrapN(r0 + TRv + 2) : wrapN(r0 - 3);
        hoff[u] = (long long)grow * N + c0;
        hlds[u] = (side ? TRv + 5 : 0) * N + c0;
    }

    // Each level's three planes sit in a ring indexed by the plane counter
    // t = p - (i0 - 2) modulo 3, so moving to the next plane moves no
    // registers for them: the loop body is instantiated for t % 3 = 0, 1, 2
    // and runs three planes per trip (a plane loop with barriers is not
    // unrolled by the compiler, and the copies p = c, c = n were a third of
    // its VALU instructions; 5% faster at 3D N = 512 FP32, while unrolling
    // six times to drop the pi copies as well ran slower).
    //   f0[t % 3] phi_n at plane p   (f0[(t+2) % 3] at p-1, f0[(t+1) % 3] at p+1)
    //   f1[t % 3] phi_n+1 at plane p-1 (likewise at p-2 and p)
    //   f2[t % 3] phi_n+2 at plane p-2 (likewise at p-3 and p-1)
    // PF == 2: phi_n at p+1, pi_n-1/2 and the outer L0 rows at p arrive in
    // nx / pr / hr (native vectors: arrays of Vec carried around the loop
    // went to scratch), loaded during plane p-1.
    typedef T nvec __attribute__((ext_vector_type(VW)));
    V f0[3][ITEMS], f1[3][ITEMS], f2[3][ITEMS];
    V q1c[ITEMS], q1n[ITEMS];               // pi_n+1/2 at planes p-1, p
    V q2c[ITEMS], q2n[ITEMS];               // pi_n+3/2 at planes p-2, p-1
    nvec nx[ITEMS], pr[ITEMS], hr[ITEMS];
    nvec nx2[ITEMS], pr2[ITEMS], hr2[ITEMS];   // PF == 3: the loads of the plane after

    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    double nacc[NQ];                         // steps A and B measure nothing

    {
        const long long bm3 = (long long)wrapN(i0 - 3) * NN, bm2 = (long long)wrapN(i0 - 2) * NN;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u]) {
                f0[2][u] = *reinterpret_cast<const V *>(phi_in + bm3 + off[u]);
                f0[0][u] = *reinterpret_cast<const V *>(phi_in + bm2 + off[u]);
            }
    }
    if constexpr (PF >= 2) {
        const long long b0 = (long long)wrapN(i0 - 2) * NN, b1 = (long long)wrapN(i0 - 1) * NN;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) {
            nx[u] = pr[u] = hr[u] = (nvec)(T)0;
            if (ok[u]) {
                nx[u] = *reinterpret_cast<const nvec *>(phi_in + b1 + off[u]);
                pr[u] = *reinterpret_cast<const nvec *>(pi_in + b0 + off[u]);
            }
            if (D == 3 && hok[u]) hr[u] = *reinterpret_cast<const nvec *>(phi_in + b0 + hoff[u]);
        }
    }
    if constexpr (PF == 3) {
        const long long b1 = (long long)wrapN(i0 - 1) * NN, b2 = (long long)wrapN(i0) * NN;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) {
            nx2[u] = pr2[u] = hr2[u] = (nvec)(T)0;
            if (ok[u]) {
                nx2[u] = *reinterpret_cast<const nvec *>(phi_in + b2 + off[u]);
                pr2[u] = *reinterpret_cast<const nvec *>(pi_in + b1 + off[u]);
            }
            if (D == 3 && hok[u]) hr2[u] = *reinterpret_cast<const nvec *>(phi_in + b1 + hoff[u]);
        }
    }

    // one plane p, with S3 = t % 3
    auto plane = [&](auto s3, int p) __attribute__((always_inline)) {
        constexpr int S3 = decltype(s3)::value;
        constexpr int PV = (S3 + 2) % 3, NX = (S3 + 1) % 3;
        const long long base = (long long)wrapN(p) * NN;
        const long long baseN = (long long)wrapN(p + 1) * NN;
        V pi0[ITEMS], hv[ITEMS];
        if constexpr (PF == 3) {
            // plane p+1's values were loaded two planes ago, plane p+2's one
            // plane ago; plane p+3's loads go out now (p + 3 <= i1 + 4 <= N + 4:
            // wrapN keeps every plane in the box)
            const long long bq2 = (long long)wrapN(p + 3) * NN, bp2 = (long long)wrapN(p + 2) * NN;
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    f0[NX][u].v[e] = nx[u][e];
                    pi0[u].v[e] = pr[u][e];
                    hv[u].v[e] = hr[u][e];
                }
                nx[u] = nx2[u];
                pr[u] = pr2[u];
                hr[u] = hr2[u];
                if (ok[u]) {
                    nx2[u] = *reinterpret_cast<const nvec *>(phi_in + bq2 + off[u]);
                    pr2[u] = *reinterpret_cast<const nvec *>(pi_in + bp2 + off[u]);
                }
                if (D == 3 && hok[u]) hr2[u] = *reinterpret_cast<const nvec *>(phi_in + bp2 + hoff[u]);
            }
        } else if constexpr (PF == 2) {
            // plane p+1's loads (p + 2 <= i1 + 3 <= N + 3: wrapN keeps every plane in the box)
            const long long bq1 = (long long)wrapN(p + 2) * NN;
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    f0[NX][u].v[e] = nx[u][e];
                    pi0[u].v[e] = pr[u][e];
                    hv[u].v[e] = hr[u][e];
                }
                if (ok[u]) {
                    nx[u] = *reinterpret_cast<const nvec *>(phi_in + bq1 + off[u]);
                    pr[u] = *reinterpret_cast<const nvec *>(pi_in + baseN + off[u]);
                }
                if (D == 3 && hok[u]) hr[u] = *reinterpret_cast<const nvec *>(phi_in + baseN + hoff[u]);
            }
        } else {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (ok[u]) {
                    f0[NX][u] = *reinterpret_cast<const V *>(phi_in + baseN + off[u]);
                    pi0[u] = *reinterpret_cast<const V *>(pi_in + base + off[u]);
                }
        }
        const bool doB = p >= i0;            // step B on plane p-1 in [i0-1, i1]
        const bool doC = p >= i0 + 2;        // step C on plane p-2 in [i0, i1-1]
        // stage phi_n(p), phi_n+1(p-1), phi_n+2(p-2)
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u]) *reinterpret_cast<V *>(L0 + (rowA[u] + H) * N + col[u]) = f0[S3][u];
        if constexpr (D == 3) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (hok[u]) {
                    if constexpr (PF >= 2) *reinterpret_cast<V *>(L0 + hlds[u]) = hv[u];
                    else *reinterpret_cast<V *>(L0 + hlds[u]) = *reinterpret_cast<const V *>(phi_in + base + hoff[u]);
                }
        }
        if (doB) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (ok[u]) *reinterpret_cast<V *>(L1 + rowA[u] * N + col[u]) = f1[S3][u];
        }
        if (doC) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (inB[u]) *reinterpret_cast<V *>(L2 + (rowA[u] - H) * N + col[u]) = f2[S3][u];
        }
        lds_barrier();
        // step A on plane p, all TR + 4H rows
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u])
                leap_vec<T, D, VW, false>(k, f0[S3][u], f0[PV][u], f0[NX][u], L0 + (rowA[u] + H) * N, col[u], pi0[u],
                                          fac_lap, fac_pot, sf, dt, q1n[u], f1[NX][u], sc, nacc);
        // step B on plane p-1, TR + 2H rows
        if (doB) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (inB[u])
                    leap_vec<T, D, VW, false>(k, f1[S3][u], f1[PV][u], f1[NX][u], L1 + rowA[u] * N, col[u], q1c[u],
                                              fac_lap, fac_pot, sf, dt, q2n[u], f2[NX][u], sc, nacc);
        }
        // step C on plane p-2, owned rows: to HBM
        if (doC) {
            const long long baseq = (long long)wrapN(p - 2) * NN;
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
                if (!inC[u]) continue;
                V pn, phn;
                leap_vec<T, D, VW, ENERGY>(k, f2[S3][u], f2[PV][u], f2[NX][u], L2 + (rowA[u] - H) * N, col[u],
                                           q2c[u], fac_lap, fac_pot, sf, dt, pn, phn, sc, acc);
                store_vec<T, VW>(pi_out + baseq + off[u], pn, nt);
                store_vec<T, VW>(phi_out + baseq + off[u], phn, nt);
            }
        }
        lds_barrier();
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) {
            q1c[u] = q1n[u];
            q2c[u] = q2n[u];
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    // (p is uniform over the block: every wave leaves at the same plane)
    for (int p = i0 - 2; p <= i1 + 1; p += 3) {
        plane(I0{}, p);
        if (p + 1 > i1 + 1) break;
        plane(I1{}, p + 1);
        if (p + 2 > i1 + 1) break;
        plane(I2{}, p + 2);
    }
    if constexpr (ENERGY) energy_partials(acc, k, sc, partials + (long long)lt * NQ);
}

} // namespace cb
