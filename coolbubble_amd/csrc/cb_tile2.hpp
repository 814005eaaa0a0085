// cb_tile2.hpp -- one tile of the two-steps-per-pass (temporal blocking)
// leapfrog for a static box, 2D and 3D (internal; included by the .hip
// files that launch it).
//
// In a static box a == 1 and nothing global feeds back into the step
// (src/bubble.cpp:749), so step n+1 at a site only needs step n at its
// neighbours.  The tile is TR rows x all N columns of the (j, k) plane,
// marching the slow axis i over NI planes:
//   step A (n -> n+1) on plane p for TR+2 rows (one extra row each side),
//          from phi_n planes p-1, p, p+1 in registers and plane p with two
//          extra rows each side in LDS (L0);
//   step B (n+1 -> n+2) on plane p-1 for the TR owned rows, from the step-A
//          results of planes p-2, p-1, p in registers and plane p-1 in LDS
//          (L1).
// Only phi_n, pi_{n-1/2} are read and phi_{n+2}, pi_{n+3/2} written: half the
// HBM bytes of two single steps, plus the overlap (rows +-2, planes +-2 per
// chunk) mostly served by L2.  Each site's arithmetic is exactly that of two
// single steps (src/bubble.cpp:716-743, :854-856, :868-870), so results are
// bit-identical.  Energies, when asked for, are those of step B (the second
// step), with the kinetic energy of the incoming pi_{n+1/2} for the
// time-centred Ekin.  pi ping-pongs (the halo rows of pi_n are read while
// neighbours write pi_{n+2}).  The LDS holds (TR+4) + (TR+2) rows of N (3D)
// or 2 rows (2D); the caller's block has NTHR threads.  PF = 2 issues the
// HBM loads of the next plane one plane ahead (as the three-step tile).
#pragma once

#include "cb_device.hpp"
#include "cb_tile_common.hpp"
#include "cb_tiled.hpp"

namespace cb {

template <typename T, int D, int ITEMS, int NTHR, bool ENERGY, int PF = 1>
__device__ __forceinline__ void tile_step2(const SimConst &k, const TiledGeom &g, int lt, const T *__restrict__ phi_in,
                                  T *__restrict__ phi_out, const T *__restrict__ pi_in, T *__restrict__ pi_out,
                                  const DevScalars *__restrict__ sc, double *__restrict__ partials, T *lds)
{
    constexpr int VW = sizeof(T) == 4 ? 4 : 2;
    constexpr int H = (D == 3) ? 1 : 0;           // extra rows of step A on each side
    using V = Vec<T, VW>;
    using C = T;

    const int N = k.N;
    const int NV = g.NV;
    const long long NN = (D == 3) ? (long long)N * N : (long long)N;
    const int TRv = (D == 3) ? g.TR : 1;
    const int rt = (D == 3) ? lt % g.ntr : 0;
    const int ic = (D == 3) ? lt / g.ntr : lt;
    const int r0 = rt * TRv;
    const int i0 = ic * g.NI;
    const int i1 = min(i0 + g.NI, N);
    const int rowsA = TRv + 2 * H;
    const int nvecA = rowsA * NV;
    T *L0 = lds;                                  // phi_n plane p, rows r0-2H .. r0+TR-1+2H
    T *L1 = lds + (size_t)(TRv + 4 * H) * N;      // phi_{n+1} plane p-1, rows r0-H .. r0+TR-1+H
    const bool nt = g.nt != 0;

    auto wrapN = [&](int i) { return i < 0 ? i + N : (i >= N ? i - N : i); };

    const C dt = (C)k.dt;
    const C fac_lap = (C)sc->fac_lap, fac_pot = (C)sc->fac_pot;
    const StepF32 sf = step_f32(k, sc);

    int rowA[ITEMS], col[ITEMS];
    bool ok[ITEMS], inner[ITEMS];
    long long off[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
        const int q = threadIdx.x + u * NTHR;
        ok[u] = q < nvecA;
        const int qq = ok[u] ? q : 0;
        rowA[u] = qq / NV;
        col[u] = (qq - rowA[u] * NV) * VW;
        const int grow = (D == 3) ? wrapN(r0 - H + rowA[u]) : 0;
        off[u] = (D == 3) ? (long long)grow * N + col[u] : (long long)col[u];
        inner[u] = ok[u] && rowA[u] >= H && rowA[u] < H + TRv;
    }
    // outer rows of L0 (3D): rows r0-2 and r0+TR+1 at L0 rows 0 and TR+3
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
        const int grow = side ? wrapN(r0 + TRv + 1) : wrapN(r0 - 2);
        hoff[u] = (long long)grow * N + c0;
        hlds[u] = (side ? TRv + 3 : 0) * N + c0;
    }

    V p0[ITEMS], c0[ITEMS], n0[ITEMS];      // phi_n   at planes p-1, p, p+1
    V p1[ITEMS], c1[ITEMS], n1[ITEMS];      // phi_n+1 at planes p-2, p-1, p
    V q1c[ITEMS], q1n[ITEMS];               // pi_n+1/2 at planes p-1, p

    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;

    {
        const long long bm2 = (long long)wrapN(i0 - 2) * NN, bm1 = (long long)wrapN(i0 - 1) * NN;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u]) {
                p0[u] = *reinterpret_cast<const V *>(phi_in + bm2 + off[u]);
                c0[u] = *reinterpret_cast<const V *>(phi_in + bm1 + off[u]);
            }
    }

    // PF == 2: loads for the next plane, in flight during this one (native
    // vectors: arrays of Vec carried around the loop went to scratch)
    typedef T nvec __attribute__((ext_vector_type(VW)));
    nvec nx_n0[ITEMS], nx_pi[ITEMS], nx_h[ITEMS];
    if constexpr (PF == 2) {
        const long long b0 = (long long)wrapN(i0 - 1) * NN, b1 = (long long)wrapN(i0) * NN;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) {
            nx_n0[u] = nx_pi[u] = nx_h[u] = (nvec)(T)0;
            if (ok[u]) {
                nx_n0[u] = *reinterpret_cast<const nvec *>(phi_in + b1 + off[u]);
                nx_pi[u] = *reinterpret_cast<const nvec *>(pi_in + b0 + off[u]);
            }
            if (D == 3 && hok[u]) nx_h[u] = *reinterpret_cast<const nvec *>(phi_in + b0 + hoff[u]);
        }
    }

    for (int p = i0 - 1; p <= i1; ++p) {
        const long long base = (long long)wrapN(p) * NN;
        const long long baseN = (long long)wrapN(p + 1) * NN;
        V pi0[ITEMS], hv[ITEMS];
        if constexpr (PF == 2) {
            // (p + 2 <= i1 + 2 <= N + 2: wrapN keeps every plane in the box)
            const long long bq1 = (long long)wrapN(p + 2) * NN;
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    n0[u].v[e] = nx_n0[u][e];
                    pi0[u].v[e] = nx_pi[u][e];
                    hv[u].v[e] = nx_h[u][e];
                }
                if (ok[u]) {
                    nx_n0[u] = *reinterpret_cast<const nvec *>(phi_in + bq1 + off[u]);
                    nx_pi[u] = *reinterpret_cast<const nvec *>(pi_in + baseN + off[u]);
                }
                if (D == 3 && hok[u]) nx_h[u] = *reinterpret_cast<const nvec *>(phi_in + baseN + hoff[u]);
            }
        } else {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (ok[u]) {
                    n0[u] = *reinterpret_cast<const V *>(phi_in + baseN + off[u]);
                    pi0[u] = *reinterpret_cast<const V *>(pi_in + base + off[u]);
                }
        }
        // stage phi_n(p) and phi_n+1(p-1)
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u]) *reinterpret_cast<V *>(L0 + (rowA[u] + H) * N + col[u]) = c0[u];
        if constexpr (D == 3) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (hok[u]) {
                    if constexpr (PF == 2) *reinterpret_cast<V *>(L0 + hlds[u]) = hv[u];
                    else *reinterpret_cast<V *>(L0 + hlds[u]) = *reinterpret_cast<const V *>(phi_in + base + hoff[u]);
                }
        }
        const bool doB = p - 1 >= i0;
        if (doB) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (ok[u]) *reinterpret_cast<V *>(L1 + rowA[u] * N + col[u]) = c1[u];
        }
        lds_barrier();
        // step A on plane p, all TR + 2H rows
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) {
            if (!ok[u]) continue;
            const T *lrow = L0 + (rowA[u] + H) * N;
            V rm, rp;
            if constexpr (D == 3) {
                rm = *reinterpret_cast<const V *>(lrow - N + col[u]);
                rp = *reinterpret_cast<const V *>(lrow + N + col[u]);
            }
            const C fl = lrow[col[u] == 0 ? N - 1 : col[u] - 1];
            const C fr = lrow[(col[u] + VW == N) ? 0 : col[u] + VW];
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                const C c = c0[u].v[e];
                const C fM = (e == 0) ? fl : c0[u].v[e - 1];
                const C fP = (e == VW - 1) ? fr : c0[u].v[e + 1];
                if constexpr (sizeof(T) == 8) {
                    C dd = (C)0.0;
                    dd += n0[u].v[e] - (C)2.0 * c + p0[u].v[e];
                    if constexpr (D == 3) dd += rp.v[e] - (C)2.0 * c + rm.v[e];
                    dd += fP - (C)2.0 * c + fM;
                    const C F = lap_term(fac_lap * dd, k) - fac_pot * vprime(c, k);
                    const C pnew = pi0[u].v[e] + dt * F;
                    q1n[u].v[e] = pnew;
                    n1[u].v[e] = c + dt * pnew;
                } else {
                    C rP = 0, rM = 0;
                    if constexpr (D == 3) { rP = rp.v[e]; rM = rm.v[e]; }
                    const C pnew = kick_f32<D>(c, n0[u].v[e], p0[u].v[e], fP, fM, rP, rM, pi0[u].v[e], sf);
                    q1n[u].v[e] = pnew;
                    n1[u].v[e] = __builtin_fmaf(sf.dt, pnew, c);
                }
            }
        }
        // step B on plane p-1, owned rows
        if (doB) {
            const long long baseq = (long long)wrapN(p - 1) * NN;
#pragma unroll
            for (int u = 0; u < ITEMS; ++u) {
                if (!inner[u]) continue;
                const T *lrow = L1 + rowA[u] * N;
                V rm, rp;
                if constexpr (D == 3) {
                    rm = *reinterpret_cast<const V *>(lrow - N + col[u]);
                    rp = *reinterpret_cast<const V *>(lrow + N + col[u]);
                }
                const C fl = lrow[col[u] == 0 ? N - 1 : col[u] - 1];
                const C fr = lrow[(col[u] + VW == N) ? 0 : col[u] + VW];
                V pn, phn;
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    const C c = c1[u].v[e];
                    const C fM = (e == 0) ? fl : c1[u].v[e - 1];
                    const C fP = (e == VW - 1) ? fr : c1[u].v[e + 1];
                    const C mM = p1[u].v[e], mP = n1[u].v[e];
                    C pnew;
                    if constexpr (sizeof(T) == 8) {
                        C dd = (C)0.0;
                        dd += mP - (C)2.0 * c + mM;
                        if constexpr (D == 3) dd += rp.v[e] - (C)2.0 * c + rm.v[e];
                        dd += fP - (C)2.0 * c + fM;
                        const C F = lap_term(fac_lap * dd, k) - fac_pot * vprime(c, k);
                        pnew = q1c[u].v[e] + dt * F;
                        phn.v[e] = c + dt * pnew;
                    } else {
                        C rP = 0, rM = 0;
                        if constexpr (D == 3) { rP = rp.v[e]; rM = rm.v[e]; }
                        pnew = kick_f32<D>(c, mP, mM, fP, fM, rP, rM, q1c[u].v[e], sf);
                        phn.v[e] = __builtin_fmaf(sf.dt, pnew, c);
                    }
                    pn.v[e] = pnew;
                    if constexpr (ENERGY) {
                        double fwd[3];
                        fwd[0] = (double)mP;
                        if constexpr (D == 3) { fwd[1] = (double)rp.v[e]; fwd[2] = (double)fP; }
                        else { fwd[1] = (double)fP; fwd[2] = 0.0; }
                        energy_terms<D, sizeof(T) == 8>((double)c, fwd, (double)pnew, (double)q1c[u].v[e], k, sc, acc);
                    }
                }
                store_vec<T, VW>(pi_out + baseq + off[u], pn, nt);
                store_vec<T, VW>(phi_out + baseq + off[u], phn, nt);
            }
        }
        lds_barrier();
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) {
            p0[u] = c0[u];
            c0[u] = n0[u];
            p1[u] = c1[u];
            c1[u] = n1[u];
            q1c[u] = q1n[u];
        }
    }
    if constexpr (ENERGY) energy_partials(acc, k, sc, partials + (long long)lt * NQ);
}

} // namespace cb
