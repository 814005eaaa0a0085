// cb_step2.hip -- two leapfrog steps per HBM pass (temporal blocking) for a
// static box, 2D and 3D, on gfx950.
//
// In a static box a == 1 and nothing global feeds back into the step
// (src/bubble.cpp:749), so step n+1 at a site only needs step n at its
// neighbours.  A block owns TR rows x all N columns of the (j, k) plane and
// marches the slow axis i:
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
// neighbours write pi_{n+2}).
#include "cb_device.hpp"
#include "cb_tile_common.hpp"
#include "cb_tiled.hpp"

namespace cb {

template <typename T, int D, int ITEMS, int NTHR, bool ENERGY>
__global__ void __launch_bounds__(NTHR)
k_step2_tiled(SimConst k, TiledGeom g, const T *__restrict__ phi_in, T *__restrict__ phi_out,
              const T *__restrict__ pi_in, T *__restrict__ pi_out, const DevScalars *__restrict__ sc,
              double *__restrict__ partials)
{
    constexpr int VW = sizeof(T) == 4 ? 4 : 2;
    constexpr int H = (D == 3) ? 1 : 0;           // extra rows of step A on each side
    using V = Vec<T, VW>;
    using C = T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);

    const int N = k.N;
    const int NV = g.NV;
    const long long NN = (D == 3) ? (long long)N * N : (long long)N;
    const int G = gridDim.x, b = blockIdx.x;
    const int lt = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
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
    const C inv_dx2 = (C)(1.0 / k.dx2);
    const C fc2 = (C)k.c2, fc3 = (C)k.c3, fc4 = (C)k.c4;

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

    for (int p = i0 - 1; p <= i1; ++p) {
        const long long base = (long long)wrapN(p) * NN;
        const long long baseN = (long long)wrapN(p + 1) * NN;
        V pi0[ITEMS];
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u]) {
                n0[u] = *reinterpret_cast<const V *>(phi_in + baseN + off[u]);
                pi0[u] = *reinterpret_cast<const V *>(pi_in + base + off[u]);
            }
        // stage phi_n(p) and phi_n+1(p-1)
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u]) *reinterpret_cast<V *>(L0 + (rowA[u] + H) * N + col[u]) = c0[u];
        if constexpr (D == 3) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (hok[u]) *reinterpret_cast<V *>(L0 + hlds[u]) = *reinterpret_cast<const V *>(phi_in + base + hoff[u]);
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
                C dd = (C)0.0;
                dd += n0[u].v[e] - (C)2.0 * c + p0[u].v[e];
                if constexpr (D == 3) dd += rp.v[e] - (C)2.0 * c + rm.v[e];
                dd += fP - (C)2.0 * c + fM;
                C F;
                if constexpr (sizeof(T) == 8) F = fac_lap * dd / k.dx2 - fac_pot * vprime(c, k);
                else F = fac_lap * dd * inv_dx2 - fac_pot * vprime_f(c, fc2, fc3, fc4);
                const C pnew = pi0[u].v[e] + dt * F;
                q1n[u].v[e] = pnew;
                n1[u].v[e] = c + dt * pnew;
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
                    C dd = (C)0.0;
                    dd += mP - (C)2.0 * c + mM;
                    if constexpr (D == 3) dd += rp.v[e] - (C)2.0 * c + rm.v[e];
                    dd += fP - (C)2.0 * c + fM;
                    C F;
                    if constexpr (sizeof(T) == 8) F = fac_lap * dd / k.dx2 - fac_pot * vprime(c, k);
                    else F = fac_lap * dd * inv_dx2 - fac_pot * vprime_f(c, fc2, fc3, fc4);
                    const C pnew = q1c[u].v[e] + dt * F;
                    pn.v[e] = pnew;
                    phn.v[e] = c + dt * pnew;
                    if constexpr (ENERGY) {
                        double fwd[3];
                        fwd[0] = (double)mP;
                        if constexpr (D == 3) { fwd[1] = (double)rp.v[e]; fwd[2] = (double)fP; }
                        else { fwd[1] = (double)fP; fwd[2] = 0.0; }
                        energy_terms<D>((double)c, fwd, (double)pnew, (double)q1c[u].v[e], k, sc, acc);
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
    if constexpr (ENERGY) block_partials<NQ>(acc, partials + (long long)lt * NQ);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static TiledGeom geom2(const SimConst &k, const TiledCfg &c)
{
    TiledGeom g;
    g.TR = (k.D == 3) ? c.TR : 1;
    g.NI = c.NI;
    g.NV = k.N / c.vw;
    g.ntr = (k.D == 3) ? k.N / c.TR : 1;
    g.nic = (k.N + c.NI - 1) / c.NI;
    g.nt = c.nt;
    return g;
}

size_t step2_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c)
{
    const size_t esz = fp32 ? 4 : 8;
    const int rows = (k.D == 3) ? (c.TR + 4) + (c.TR + 2) : 2;
    return (size_t)rows * k.N * esz;
}

bool step2_valid(const SimConst &k, int fp32, const TiledCfg &c)
{
    if (k.D < 2 || k.expanding) return false;
    const int want_vw = fp32 ? 4 : 2;
    if (c.vw != want_vw || k.N % c.vw != 0 || k.N < 4) return false;
    if (c.items != 1 && c.items != 2 && c.items != 4) return false;
    if (c.nthr != 256 && c.nthr != 512 && c.nthr != 1024) return false;
    if (c.NI < 1) return false;
    if (k.D == 3 && (c.TR < 1 || k.N % c.TR != 0)) return false;
    const int NV = k.N / c.vw;
    const int rowsA = (k.D == 3) ? c.TR + 2 : 1;
    if ((long long)rowsA * NV > (long long)c.items * c.nthr) return false;
    if (k.D == 3 && 2 * NV > c.items * c.nthr) return false;
    if (step2_lds_bytes(k, fp32, c) > 65536) return false;
    return true;
}

template <typename T, int D, int ITEMS, int NTHR>
static void launch2_t(const SimConst &k, int energy, const TiledCfg &c, const StepPtrs &p, const DevScalars *sc,
                      double *partials, hipStream_t st, size_t shm)
{
    TiledGeom g = geom2(k, c);
    dim3 grid(g.ntr * g.nic), block(NTHR);
    if (energy)
        k_step2_tiled<T, D, ITEMS, NTHR, true><<<grid, block, shm, st>>>(
            k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
    else
        k_step2_tiled<T, D, ITEMS, NTHR, false><<<grid, block, shm, st>>>(
            k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
}

template <typename T, int D>
static void launch2_d(const SimConst &k, int energy, const TiledCfg &c, const StepPtrs &p, const DevScalars *sc,
                      double *partials, hipStream_t st, size_t shm)
{
#define CB_L2(I, NT) launch2_t<T, D, I, NT>(k, energy, c, p, sc, partials, st, shm)
    if (c.nthr == 256) {
        if (c.items == 1) CB_L2(1, 256);
        else if (c.items == 2) CB_L2(2, 256);
        else CB_L2(4, 256);
    } else if (c.nthr == 512) {
        if (c.items == 1) CB_L2(1, 512);
        else if (c.items == 2) CB_L2(2, 512);
        else CB_L2(4, 512);
    } else {
        if (c.items == 1) CB_L2(1, 1024);
        else if (c.items == 2) CB_L2(2, 1024);
        else CB_L2(4, 1024);
    }
#undef CB_L2
}

hipError_t launch_step2(const SimConst &k, int fp32, int energy, const TiledCfg &c, const StepPtrs &p,
                        const DevScalars *sc, double *partials, hipStream_t st, int *nblocks)
{
    if (!step2_valid(k, fp32, c)) return hipErrorInvalidValue;
    const size_t shm = step2_lds_bytes(k, fp32, c);
    TiledGeom g = geom2(k, c);
    if (nblocks) *nblocks = g.ntr * g.nic;
    if (fp32) {
        if (k.D == 3) launch2_d<float, 3>(k, energy, c, p, sc, partials, st, shm);
        else launch2_d<float, 2>(k, energy, c, p, sc, partials, st, shm);
    } else {
        if (k.D == 3) launch2_d<double, 3>(k, energy, c, p, sc, partials, st, shm);
        else launch2_d<double, 2>(k, energy, c, p, sc, partials, st, shm);
    }
    return hipGetLastError();
}

int step2_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap)
{
    int n = 0;
    const int vw = fp32 ? 4 : 2;
    const int NV = k.N / vw;
    auto push = [&](int nthr, int TR, int NI, int nt) {
        const int rowsA = (k.D == 3) ? TR + 2 : 1;
        int items = (rowsA * NV + nthr - 1) / nthr;
        if (items == 3) items = 4;
        if (k.D == 3) {
            int hitems = (2 * NV + nthr - 1) / nthr;
            if (hitems == 3) hitems = 4;
            if (hitems > items) items = hitems;
        }
        if (items < 1) items = 1;
        TiledCfg c{vw, items, nthr, TR, NI, 1, nt, 2};
        if (n < cap && step2_valid(k, fp32, c)) out[n++] = c;
    };
    const int trs3[] = {6, 4, 2, 8, 12, 1};
    const int nis[] = {32, 16, 64};
    if (k.D == 3) {
        for (int TR : trs3)
            for (int nthr : {256, 512, 1024})
                for (int NI : nis)
                    for (int nt : {1, 0}) push(nthr, TR, NI, nt);
    } else {
        for (int nthr : {256, 512})
            for (int NI : {16, 8, 32})
                for (int nt : {1, 0}) push(nthr, 1, NI, nt);
    }
    return n;
}

} // namespace cb
