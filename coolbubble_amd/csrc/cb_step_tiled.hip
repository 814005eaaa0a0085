// cb_step_tiled.hip -- LDS-tiled, plane-marching fused step for 2D and 3D
// lattices on gfx950.
//
// Layout (src/gridfuncs.cpp:25-26): site (i, j, k) at (i N + j) N + k, k
// fastest.  A block owns a tile of TR rows x all N columns of the (j, k)
// plane (2D: one row) and walks the slow axis i through a chunk of NI planes.
// Per plane it loads the next plane's tile once from HBM (16-byte vector
// loads, whole rows, so the k taps never leave the tile), keeps planes
// i-1, i, i+1 in registers for the i taps, and stages plane i plus one halo
// row above and below in LDS for the j taps and the k taps across vector
// boundaries.  Each site is read once per step apart from the halo rows
// (2/TR) and the chunk-end planes (2/NI), which adjacent blocks scheduled on
// the same XCD (block ids are remapped so an XCD owns a contiguous run of
// tiles) mostly find in that XCD's L2.
//
// The arithmetic per site is the reference's, in its order
// (src/bubble.cpp:716-743 force, :854-856 kick, :868-870 drift, :776-836
// energies), identical to the general kernel in cb_kernels.hip.
#include <mutex>
#include <vector>

#include "cb_device.hpp"
#include "cb_tile2.hpp"
#include "cb_tile3.hpp"
#include "cb_tile_common.hpp"
#include "cb_tiled.hpp"

namespace cb {

// One tile of one step: TR rows x all N columns of every plane i in
// [lt's i-chunk].  lds holds two plane tiles (double buffer).  Ends with an
// LDS barrier so a caller may run the next tile on the same LDS.
template <typename T, int D, int VW, int ITEMS, int NTHR, bool LAGGED, int ENERGY, int PF>
__device__ __forceinline__ void tile_step(const SimConst &k, const TiledGeom &g, int lt, const T *__restrict__ phi_in,
                                 T *__restrict__ phi_out, const T *pi_in, T *pi_out,
                                 const DevScalars *__restrict__ sc, double *__restrict__ partials, T *lds)
{
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
    const int nvec = TRv * NV;
    const int lds_plane = (D == 3) ? (TRv + 2) * N : N;
    const bool nt = g.nt != 0;

    const C dt = (C)k.dt;
    const C s_prev = LAGGED ? (C)sc->s_pending : (C)0;
    const C fac_lap = (C)sc->fac_lap, fac_pot = (C)sc->fac_pot;
    const StepF32 sf = step_f32(k, sc);

    int rowq[ITEMS], col[ITEMS];
    bool ok[ITEMS];
    long long off[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
        int q = threadIdx.x + u * NTHR;
        ok[u] = q < nvec;
        int qq = ok[u] ? q : 0;
        rowq[u] = qq / NV;
        col[u] = (qq - rowq[u] * NV) * VW;
        off[u] = (D == 3) ? (long long)(r0 + rowq[u]) * N + col[u] : (long long)col[u];
    }
    // halo rows (3D): vector h of 2 NV, row r0-1 (h < NV) or r0+TR
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
        const int grow = side ? ((r0 + TRv == N) ? 0 : r0 + TRv) : (r0 == 0 ? N - 1 : r0 - 1);
        hoff[u] = (long long)grow * N + c0;
        hlds[u] = (side ? TRv + 1 : 0) * N + c0;
    }

    auto wrap = [&](int i) { return i >= N ? i - N : (i < 0 ? i + N : i); };
    auto load_v = [&](long long idx, V &out, V &pout) {
        V f = *reinterpret_cast<const V *>(phi_in + idx);
        if constexpr (LAGGED) {
            V p = *reinterpret_cast<const V *>(pi_in + idx);
#pragma unroll
            for (int e = 0; e < VW; ++e) f.v[e] = f.v[e] + dt * p.v[e] * s_prev;
            pout = p;
        }
        out = f;
    };
    auto load_plane = [&](int i, V (&out)[ITEMS], V (&pout)[ITEMS]) {
        const long long base = (long long)i * NN;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u]) load_v(base + off[u], out[u], pout[u]);
    };
    // The halo rows are held in native vectors: held as arrays of Vec
    // written under the hok branch, the compiler kept them in scratch and
    // waited for every load in flight to store them there.
    typedef T hvec __attribute__((ext_vector_type(VW)));
    auto load_halo = [&](int i, hvec (&out)[ITEMS]) {
        const long long base = (long long)i * NN;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (hok[u]) {
                hvec f = *reinterpret_cast<const hvec *>(phi_in + base + hoff[u]);
                if constexpr (LAGGED) {
                    const hvec p = *reinterpret_cast<const hvec *>(pi_in + base + hoff[u]);
#pragma unroll
                    for (int e = 0; e < VW; ++e) f[e] = f[e] + dt * p[e] * s_prev;
                }
                out[u] = f;
            }
    };

    V prev[ITEMS], cur[ITEMS], nxt[ITEMS], nx2[ITEMS];
    V pcur[ITEMS], pnxt[ITEMS], pnx2[ITEMS];
    hvec hcur[ITEMS], hnxt[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) hcur[u] = hnxt[u] = (hvec)(T)0;

    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;

    load_plane(wrap(i0 - 1), prev, pnxt);
    load_plane(i0, cur, pcur);
    if constexpr (PF == 2) {
        load_plane(wrap(i0 + 1), nxt, pnxt);
        if constexpr (D == 3) load_halo(i0, hcur);
    }

    int buf = 0;
    for (int i = i0; i < i1; ++i) {
        const long long base = (long long)i * NN;
        if constexpr (PF == 2) {
            if (i + 1 < i1) {
                load_plane(wrap(i + 2), nx2, pnx2);
                if constexpr (D == 3) load_halo(i + 1, hnxt);
            }
        } else {
            load_plane(wrap(i + 1), nxt, pnxt);
            if constexpr (D == 3) load_halo(i, hcur);
        }
        if constexpr (!LAGGED) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (ok[u]) pcur[u] = *reinterpret_cast<const V *>(pi_in + base + off[u]);
        }
        T *L = lds + buf * lds_plane;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u)
            if (ok[u]) *reinterpret_cast<V *>(L + (D == 3 ? (rowq[u] + 1) * N : 0) + col[u]) = cur[u];
        if constexpr (D == 3) {
#pragma unroll
            for (int u = 0; u < ITEMS; ++u)
                if (hok[u]) *reinterpret_cast<hvec *>(L + hlds[u]) = hcur[u];
        }
        lds_barrier();
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) {
            if (!ok[u]) continue;
            const T *lrow = L + (D == 3 ? (rowq[u] + 1) * N : 0);
            V rm, rp;
            if constexpr (D == 3) {
                rm = *reinterpret_cast<const V *>(lrow - N + col[u]);
                rp = *reinterpret_cast<const V *>(lrow + N + col[u]);
            }
            const C fl = lrow[col[u] == 0 ? N - 1 : col[u] - 1];
            const C fr = lrow[(col[u] + VW == N) ? 0 : col[u] + VW];
            V po = pcur[u], pn, phn;
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                const C c = cur[u].v[e];
                const C fM = (e == 0) ? fl : cur[u].v[e - 1];
                const C fP = (e == VW - 1) ? fr : cur[u].v[e + 1];
                const C mM = prev[u].v[e], mP = nxt[u].v[e];
                C pnew;
                if constexpr (sizeof(T) == 8) {
                    C dd = (C)0.0;
                    dd += mP - (C)2.0 * c + mM;
                    if constexpr (D == 3) dd += rp.v[e] - (C)2.0 * c + rm.v[e];
                    dd += fP - (C)2.0 * c + fM;
                    const C F = lap_term(fac_lap * dd, k) - fac_pot * vprime(c, k);
                    pnew = po.v[e] + dt * F;
                    if constexpr (LAGGED) phn.v[e] = c;
                    else phn.v[e] = c + dt * pnew;
                } else {
                    C rP = 0, rM = 0;
                    if constexpr (D == 3) { rP = rp.v[e]; rM = rm.v[e]; }
                    pnew = kick_f32<D>(c, mP, mM, fP, fM, rP, rM, po.v[e], sf);
                    if constexpr (LAGGED) phn.v[e] = c;
                    else phn.v[e] = __builtin_fmaf(sf.dt, pnew, c);
                }
                pn.v[e] = pnew;
                if constexpr (ENERGY) {
                    double fwd[3];
                    fwd[0] = (double)mP;
                    if constexpr (D == 3) { fwd[1] = (double)rp.v[e]; fwd[2] = (double)fP; }
                    else { fwd[1] = (double)fP; fwd[2] = 0.0; }
                    energy_terms<D, sizeof(T) == 8, ENERGY == 2>((double)c, fwd, (double)pnew, (double)po.v[e], k, sc, acc);
                }
            }
            store_vec<T, VW>(pi_out + base + off[u], pn, nt);
            store_vec<T, VW>(phi_out + base + off[u], phn, nt);
        }
        buf ^= 1;
#pragma unroll
        for (int u = 0; u < ITEMS; ++u) {
            prev[u] = cur[u];
            cur[u] = nxt[u];
            if constexpr (LAGGED) pcur[u] = pnxt[u];
            if constexpr (PF == 2) {
                nxt[u] = nx2[u];
                if constexpr (LAGGED) pnxt[u] = pnx2[u];
                if constexpr (D == 3) hcur[u] = hnxt[u];
            }
        }
    }
    if constexpr (ENERGY) energy_partials(acc, k, sc, partials + (long long)lt * NQ);
    lds_barrier();
}

template <typename T, int D, int VW, int ITEMS, int NTHR, bool LAGGED, int ENERGY, int PF>
__global__ void __launch_bounds__(NTHR)
k_step_tiled(SimConst k, TiledGeom g, const T *__restrict__ phi_in, T *__restrict__ phi_out,
             const T *pi_in, T *pi_out, const DevScalars *__restrict__ sc, double *__restrict__ partials)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int G = gridDim.x, b = blockIdx.x;
    const int lt = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;   // XCD-contiguous tiles
    tile_step<T, D, VW, ITEMS, NTHR, LAGGED, ENERGY, PF>(k, g, lt, phi_in, phi_out, pi_in, pi_out, sc, partials,
                                                         reinterpret_cast<T *>(smem_raw));
}

// Batched persistent step: block b takes work items lt(b), lt(b) + G, ...
// (XCD-contiguous order), each one tile of one member (binary search on the
// members' first work items).  Per-member D and lagging are uniform per
// block.  The per-tile arithmetic and partials are those of k_step_tiled
// with the same config, so a member steps bit-identically alone or batched.
// DSEL = 2 or 3 runs only the members of that dimension: a kernel then holds
// one dimension's tile code, and the energy kernels, whose FP64 sums need
// many registers, keep their occupancy (0: every member).
template <typename T, bool ENERGY, int ITEMS, int PF, int DSEL>
__global__ void __launch_bounds__(256)
k_batch_step(const BatchMember *__restrict__ table, int nmem, int total, int step_it, int sel)
{
    constexpr int VW = sizeof(T) == 4 ? 4 : 2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const int G = gridDim.x, b = blockIdx.x;
    const int lb = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
    for (int w = lb; w < total; w += G) {
        int lo = 0, hi = nmem - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (table[mid].tile_begin <= w) lo = mid;
            else hi = mid - 1;
        }
        const BatchMember &m = table[lo];
        if (DSEL != 0 && m.k.D != DSEL) continue;
        if (step_it >= m.nsteps) continue;     // this member stopped at an event earlier in the sequence
        // members measure energies every step when expanding, else at their energy step
        const bool want = m.lagged || m.energy_it < 0 || m.energy_it == step_it;
        if (ENERGY ? !want : (sel == 1 && want)) continue;
        const int lt = w - m.tile_begin;
        const int pphi = (m.par_phi + step_it) & 1;
        const int ppi = m.lagged ? ((m.par_pi + step_it) & 1) : m.par_pi;
        const T *phi_in = (const T *)m.phi[pphi];
        T *phi_out = (T *)m.phi[pphi ^ 1];
        const T *pi_in = (const T *)m.pi[ppi];
        T *pi_out = m.lagged ? (T *)m.pi[ppi ^ 1] : (T *)m.pi[ppi];
        if (DSEL != 2 && m.k.D == 3) {
            if (m.lagged)
                tile_step<T, 3, VW, ITEMS, 256, true, ENERGY, PF>(m.k, m.g, lt, phi_in, phi_out, pi_in, pi_out, m.sc,
                                                             m.partials, lds);
            else
                tile_step<T, 3, VW, ITEMS, 256, false, ENERGY, PF>(m.k, m.g, lt, phi_in, phi_out, pi_in, pi_out, m.sc,
                                                              m.partials, lds);
        } else if (DSEL != 3) {
            tile_step<T, 2, VW, ITEMS, 256, false, ENERGY, PF>(m.k, m.g, lt, phi_in, phi_out, pi_in, pi_out, m.sc,
                                                          m.partials, lds);
        }
    }
}

__global__ void __launch_bounds__(256) k_batch_finalize(const BatchMember *__restrict__ table, int step_it)
{
    const BatchMember &m = table[blockIdx.x];
    if (step_it >= m.nsteps) return;
    if (!m.lagged && m.energy_it >= 0 && m.energy_it != step_it) return;
    finalize_member(m.k, m.partials, m.ntiles, m.sc, m.lagged ? 0 : 1);
}

// Blocks of a persistent batched launch of kernel `kern`: at most as many as
// the GPU holds at once for it (by its registers and LDS), a multiple of 8
// (one share per XCD).  A block beyond those would start only when a
// resident one ends and then walk its whole share of tiles while most CUs
// idle.  The tiles' arithmetic and partials do not depend on the grid.
template <typename K>
static int resident_grid(K kern, size_t shm, int grid)
{
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(kern), 256, shm) != hipSuccess ||
        nb < 1 || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return grid;
    int cap = nb * cus;
    cap -= cap % 8;
    return (cap >= 8 && cap < grid) ? cap : grid;
}

template <typename T, bool ENERGY, int DSEL>
static void batch_launch_ip(int items, int pf, const BatchMember *table, int nmem, int total, int grid, size_t shm,
                            int step_it, int sel, hipStream_t st)
{
#define CB_BATCH_L(I, P)                                                                                          \
    k_batch_step<T, ENERGY, I, P, DSEL><<<resident_grid(k_batch_step<T, ENERGY, I, P, DSEL>, shm, grid), 256, shm, st>>>( \
        table, nmem, total, step_it, sel)
    if (pf == 2) {
        if (items == 1) CB_BATCH_L(1, 2);
        else if (items == 2) CB_BATCH_L(2, 2);
        else CB_BATCH_L(4, 2);
    } else {
        if (items == 1) CB_BATCH_L(1, 1);
        else if (items == 2) CB_BATCH_L(2, 1);
        else CB_BATCH_L(4, 1);
    }
#undef CB_BATCH_L
}

hipError_t launch_batch_step(int fp32, int energy, int split, int items, int pf, const BatchMember *table, int nmem,
                             int total_tiles, int grid, size_t shm, int step_it, hipStream_t st)
{
    // sel 0: every member, plain; 1: members not measuring energies, plain;
    // 2 (energy kernel): members measuring energies
    // the energy kernels run per dimension (see k_batch_step)
    if (fp32) {
        if (!energy || split)
            batch_launch_ip<float, false, 0>(items, pf, table, nmem, total_tiles, grid, shm, step_it, energy ? 1 : 0, st);
        if (energy) {
            batch_launch_ip<float, true, 2>(items, pf, table, nmem, total_tiles, grid, shm, step_it, 2, st);
            batch_launch_ip<float, true, 3>(items, pf, table, nmem, total_tiles, grid, shm, step_it, 2, st);
        }
    } else {
        if (!energy || split)
            batch_launch_ip<double, false, 0>(items, pf, table, nmem, total_tiles, grid, shm, step_it, energy ? 1 : 0, st);
        if (energy) {
            batch_launch_ip<double, true, 2>(items, pf, table, nmem, total_tiles, grid, shm, step_it, 2, st);
            batch_launch_ip<double, true, 3>(items, pf, table, nmem, total_tiles, grid, shm, step_it, 2, st);
        }
    }
    return hipGetLastError();
}

hipError_t launch_batch_finalize(const BatchMember *table, int nmem, int step_it, hipStream_t st)
{
    if (nmem > 0) k_batch_finalize<<<nmem, 256, 0, st>>>(table, step_it);
    return hipGetLastError();
}

// Batched multi-step launch L of a sequence: every tile of every static
// member with a two- or three-step tile (BatchMember::fused = F, DSEL as in
// k_batch_step).  A member taking n steps runs n / F multi-step launches
// (phi and pi flip once each), then n % F launches of its one-step
// companion (phi flips, pi in place), then is skipped; each launch is the
// same tile of the same variant as a lone run, so a member steps
// bit-identically alone or batched.  Energies as k_batch_step, at the
// member's energy step (the last one its launch covers).  FSEL = 2 or 3
// holds only the tile of that many steps (every member in the table has
// F = FSEL), 0 both: the kernel that holds both takes more registers (90
// VGPRs against 74 for the two-step one alone, FP32, one vector per thread).
template <typename T, bool ENERGY, int ITEMS, int DSEL, int FSEL>
__global__ void __launch_bounds__(256)
k_batch_multi(const BatchMember *__restrict__ table, int nmem, int total, int L, int sel)
{
    constexpr int VW = sizeof(T) == 4 ? 4 : 2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const int G = gridDim.x, b = blockIdx.x;
    const int lb = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
    for (int w = lb; w < total; w += G) {
        int lo = 0, hi = nmem - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (table[mid].tile_begin <= w) lo = mid;
            else hi = mid - 1;
        }
        const BatchMember &m = table[lo];
        if (DSEL != 0 && m.k.D != DSEL) continue;
        const int F = m.fused, q = m.nsteps / F, r = m.nsteps - q * F;
        if (L >= q + r) continue;
        const bool multi = L < q;
        const int last = multi ? F * L + F - 1 : F * q + (L - q);
        const bool want = m.energy_it < 0 || m.energy_it == last;
        if (ENERGY ? !want : (sel == 1 && want)) continue;
        const int lt = w - m.tile_begin;
        const int pphi = (m.par_phi + L) & 1;
        const int ppi = (m.par_pi + (multi ? L : q)) & 1;
        const T *phi_in = (const T *)m.phi[pphi];
        T *phi_out = (T *)m.phi[pphi ^ 1];
        if (multi) {
            if (lt >= m.ntiles) continue;
            const T *pi_in = (const T *)m.pi[ppi];
            T *pi_out = (T *)m.pi[ppi ^ 1];
            if (FSEL != 2 && (FSEL == 3 || F == 3)) {
                if (DSEL != 2 && m.k.D == 3) tile_step3<T, 3, ITEMS, 256, ENERGY>(m.k, m.g, lt, phi_in, phi_out, pi_in, pi_out, m.sc, m.partials, lds);
                else if (DSEL != 3) tile_step3<T, 2, ITEMS, 256, ENERGY>(m.k, m.g, lt, phi_in, phi_out, pi_in, pi_out, m.sc, m.partials, lds);
            } else if (FSEL != 3) {
                if (DSEL != 2 && m.k.D == 3) tile_step2<T, 3, ITEMS, 256, ENERGY>(m.k, m.g, lt, phi_in, phi_out, pi_in, pi_out, m.sc, m.partials, lds);
                else if (DSEL != 3) tile_step2<T, 2, ITEMS, 256, ENERGY>(m.k, m.g, lt, phi_in, phi_out, pi_in, pi_out, m.sc, m.partials, lds);
            }
        } else {                   // a step left over: the one-step companion, pi in place
            if (lt >= m.ntiles1) continue;
            T *pi_io = (T *)m.pi[ppi];
            if (DSEL != 2 && m.k.D == 3) tile_step<T, 3, VW, ITEMS, 256, false, ENERGY, 1>(m.k, m.g1, lt, phi_in, phi_out, pi_io, pi_io, m.sc, m.partials, lds);
            else if (DSEL != 3) tile_step<T, 2, VW, ITEMS, 256, false, ENERGY, 1>(m.k, m.g1, lt, phi_in, phi_out, pi_io, pi_io, m.sc, m.partials, lds);
        }
    }
}

__global__ void __launch_bounds__(256) k_batch_finalize_multi(const BatchMember *__restrict__ table, int L)
{
    const BatchMember &m = table[blockIdx.x];
    const int F = m.fused, q = m.nsteps / F, r = m.nsteps - q * F;
    if (L >= q + r) return;
    const bool multi = L < q;
    const int last = multi ? F * L + F - 1 : F * q + (L - q);
    if (m.energy_it >= 0 && m.energy_it != last) return;
    finalize_member(m.k, m.partials, multi ? m.ntiles : m.ntiles1, m.sc, 1);
}

template <typename T, bool ENERGY, int DSEL, int FSEL>
static void batch_multi_launch_i(int items, const BatchMember *table, int nmem, int total, int grid, size_t shm, int L,
                                 int sel, hipStream_t st)
{
    if (items == 1)
        k_batch_multi<T, ENERGY, 1, DSEL, FSEL><<<resident_grid(k_batch_multi<T, ENERGY, 1, DSEL, FSEL>, shm, grid), 256,
                                                 shm, st>>>(table, nmem, total, L, sel);
    else
        k_batch_multi<T, ENERGY, 2, DSEL, FSEL><<<resident_grid(k_batch_multi<T, ENERGY, 2, DSEL, FSEL>, shm, grid), 256,
                                                 shm, st>>>(table, nmem, total, L, sel);
}

template <typename T, int FSEL>
static void batch_multi_launch_f(int energy, int split, int items, const BatchMember *table, int nmem, int total,
                                 int grid, size_t shm, int L, hipStream_t st)
{
    if (!energy || split) batch_multi_launch_i<T, false, 0, FSEL>(items, table, nmem, total, grid, shm, L, energy ? 1 : 0, st);
    if (energy) {
        batch_multi_launch_i<T, true, 2, FSEL>(items, table, nmem, total, grid, shm, L, 2, st);
        batch_multi_launch_i<T, true, 3, FSEL>(items, table, nmem, total, grid, shm, L, 2, st);
    }
}

template <typename T>
static void batch_multi_launch_t(int fuses, int energy, int split, int items, const BatchMember *table, int nmem,
                                 int total, int grid, size_t shm, int L, hipStream_t st)
{
    if (fuses == 2) batch_multi_launch_f<T, 2>(energy, split, items, table, nmem, total, grid, shm, L, st);
    else if (fuses == 3) batch_multi_launch_f<T, 3>(energy, split, items, table, nmem, total, grid, shm, L, st);
    else batch_multi_launch_f<T, 0>(energy, split, items, table, nmem, total, grid, shm, L, st);
}

hipError_t launch_batch_multi(int fp32, int energy, int split, int items, int fuses, const BatchMember *table,
                              int nmem, int total_tiles, int grid, size_t shm, int L, hipStream_t st)
{
    if (items > 2 || (fuses != 0 && fuses != 2 && fuses != 3)) return hipErrorInvalidValue;
    if (fp32) batch_multi_launch_t<float>(fuses, energy, split, items, table, nmem, total_tiles, grid, shm, L, st);
    else batch_multi_launch_t<double>(fuses, energy, split, items, table, nmem, total_tiles, grid, shm, L, st);
    return hipGetLastError();
}

hipError_t launch_batch_finalize_multi(const BatchMember *table, int nmem, int L, hipStream_t st)
{
    if (nmem > 0) k_batch_finalize_multi<<<nmem, 256, 0, st>>>(table, L);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static TiledGeom geom(const SimConst &k, const TiledCfg &c)
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

static size_t lds_bytes(const SimConst &k, int fp32, const TiledCfg &c);

size_t tiled_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c) { return lds_bytes(k, fp32, c); }

static TiledGeom geom(const SimConst &k, const TiledCfg &c);

TiledGeom tiled_geom(const SimConst &k, const TiledCfg &c) { return geom(k, c); }

static size_t lds_bytes(const SimConst &k, int fp32, const TiledCfg &c)
{
    const size_t esz = fp32 ? 4 : 8;
    const int rows = (k.D == 3) ? c.TR + 2 : 1;
    return 2 * (size_t)rows * k.N * esz;     // double-buffered plane tile
}

bool tiled_valid(const SimConst &k, int fp32, const TiledCfg &c)
{
    if (k.D < 2) return false;
    const int want_vw = fp32 ? 4 : 2;
    if (c.vw != want_vw) return false;
    if (k.N % c.vw != 0 || k.N < 4) return false;
    if (c.items != 1 && c.items != 2 && c.items != 4) return false;
    // 192 and 384 threads (3D, one or two vectors per thread): tiles whose
    // vector count is a multiple of 3 x 64 (N = 3 x 2^m) without idle lanes
    const bool three = k.D == 3 && (c.nthr == 192 || c.nthr == 384) && c.items <= 2;
    if (c.nthr != 128 && c.nthr != 256 && !three) return false;
    if (c.NI < 1) return false;
    if (k.D == 3 && (c.TR < 1 || k.N % c.TR != 0)) return false;
    const int rows = (k.D == 3) ? c.TR : 1;
    if ((long long)rows * (k.N / c.vw) > (long long)c.items * c.nthr) return false;
    if (lds_bytes(k, fp32, c) > 65536) return false;
    if (c.pf != 1 && c.pf != 2) return false;
    if (c.nt != 0 && c.nt != 1) return false;
    if (k.D == 3 && 2 * (k.N / c.vw) > c.items * c.nthr) return false;   // halo rows need owners
    return true;
}

int tiled_blocks(const SimConst &k, const TiledCfg &c)
{
    TiledGeom g = geom(k, c);
    return g.ntr * g.nic;
}

static int single_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap);

int tiled_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap)
{
    int n = single_candidates(k, fp32, out, cap);
    if (!k.expanding && n < cap) n += step2_candidates(k, fp32, out + n, cap - n);
    if (!k.expanding && n < cap) n += step3_candidates(k, fp32, out + n, cap - n);
    return n;
}

static int single_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap)
{
    int n = 0;
    const int vw = fp32 ? 4 : 2;
    auto push = [&](int items, int nthr, int TR, int NI) {
        for (int pf : {2, 1})
            for (int nt : {0, 1}) {
                TiledCfg c{vw, items, nthr, TR, NI, pf, nt, 1};
                if (n < cap && tiled_valid(k, fp32, c)) out[n++] = c;
            }
    };
    if (k.D == 3) {
        const int trs[] = {8, 4, 2, 1};
        const int nis[] = {32, 16, 64, 8};
        for (int TR : trs) {
            const int nv = TR * (k.N / vw);
            for (int nthr : {256, 128}) {
                int items = (nv + nthr - 1) / nthr;
                if (items == 3) items = 4;
                for (int NI : nis) push(items, nthr, TR, NI);
            }
            // whole multiples of 192 or 384 vectors: every lane has a vector;
            // chunks of 48 planes too where they tile N (N = 384: 8 chunks,
            // 6 two-row tiles of 192 threads per CU, all resident at once)
            for (int nthr : {192, 384})
                if (nv % nthr == 0 && nv / nthr <= 2) {
                    for (int NI : nis) push(nv / nthr, nthr, TR, NI);
                    if (k.N % 48 == 0) push(nv / nthr, nthr, TR, 48);
                }
        }
    } else {
        const int nis[] = {16, 8, 32, 4};
        for (int nthr : {128, 256}) {
            const int nv = k.N / vw;
            int items = (nv + nthr - 1) / nthr;
            if (items == 3) items = 4;
            for (int NI : nis) push(items, nthr, 1, NI);
        }
    }
    return n;
}

template <typename T, int D, int ITEMS, int NTHR, int PF>
static void launch_t(const SimConst &k, int lagged, int energy, const TiledCfg &c, const StepPtrs &p,
                     const DevScalars *sc, double *partials, hipStream_t st, size_t shm)
{
    constexpr int VW = sizeof(T) == 4 ? 4 : 2;
    TiledGeom g = geom(k, c);
    dim3 grid(g.ntr * g.nic), block(NTHR);
    const T *pin = (const T *)p.phi_in;
    T *pout = (T *)p.phi_out;
    const T *qin = (const T *)p.pi_in;
    T *qout = (T *)p.pi_out;
    if constexpr (D == 3) {
        if (lagged) {
            if (energy && k.energy_lite) k_step_tiled<T, D, VW, ITEMS, NTHR, true, 2, PF><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
            else if (energy) k_step_tiled<T, D, VW, ITEMS, NTHR, true, true, PF><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
            else k_step_tiled<T, D, VW, ITEMS, NTHR, true, false, PF><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
            return;
        }
    }
    if (energy) k_step_tiled<T, D, VW, ITEMS, NTHR, false, true, PF><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
    else k_step_tiled<T, D, VW, ITEMS, NTHR, false, false, PF><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
}

template <typename T, int D, int PF>
static void launch_p(const SimConst &k, int lagged, int energy, const TiledCfg &c, const StepPtrs &p,
                     const DevScalars *sc, double *partials, hipStream_t st, size_t shm)
{
    if constexpr (D == 3) {
        if (c.nthr == 192 || c.nthr == 384) {
            if (c.nthr == 192) {
                if (c.items == 1) launch_t<T, D, 1, 192, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
                else launch_t<T, D, 2, 192, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
            } else {
                if (c.items == 1) launch_t<T, D, 1, 384, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
                else launch_t<T, D, 2, 384, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
            }
            return;
        }
    }
    if (c.nthr == 256) {
        if (c.items == 1) launch_t<T, D, 1, 256, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
        else if (c.items == 2) launch_t<T, D, 2, 256, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
        else launch_t<T, D, 4, 256, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
    } else {
        if (c.items == 1) launch_t<T, D, 1, 128, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
        else if (c.items == 2) launch_t<T, D, 2, 128, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
        else launch_t<T, D, 4, 128, PF>(k, lagged, energy, c, p, sc, partials, st, shm);
    }
}

template <typename T, int D>
static void launch_d(const SimConst &k, int lagged, int energy, const TiledCfg &c, const StepPtrs &p,
                     const DevScalars *sc, double *partials, hipStream_t st, size_t shm)
{
    if (c.pf == 2) launch_p<T, D, 2>(k, lagged, energy, c, p, sc, partials, st, shm);
    else launch_p<T, D, 1>(k, lagged, energy, c, p, sc, partials, st, shm);
}

hipError_t launch_step_tiled_cfg(const SimConst &k, int fp32, int lagged, int energy, const TiledCfg &c,
                                 const StepPtrs &p, const DevScalars *sc, double *partials, hipStream_t st)
{
    if (c.fuse != 1) return hipErrorInvalidValue;   // multi-step shapes launch through launch_step2 / 3
    if (!tiled_valid(k, fp32, c)) return hipErrorInvalidValue;
    if (lagged && k.D != 3) return hipErrorInvalidValue;
    const size_t shm = lds_bytes(k, fp32, c);
    if (fp32) {
        if (k.D == 3) launch_d<float, 3>(k, lagged, energy, c, p, sc, partials, st, shm);
        else launch_d<float, 2>(k, lagged, energy, c, p, sc, partials, st, shm);
    } else {
        if (k.D == 3) launch_d<double, 3>(k, lagged, energy, c, p, sc, partials, st, shm);
        else launch_d<double, 2>(k, lagged, energy, c, p, sc, partials, st, shm);
    }
    return hipGetLastError();
}

static std::mutex g_reg_mu;
static std::vector<TiledCfg> g_reg;

int tiled_register(const TiledCfg &c)
{
    std::lock_guard<std::mutex> lk(g_reg_mu);
    for (size_t i = 0; i < g_reg.size(); ++i) {
        const TiledCfg &o = g_reg[i];
        if (o.vw == c.vw && o.items == c.items && o.nthr == c.nthr && o.TR == c.TR && o.NI == c.NI &&
            o.pf == c.pf && o.nt == c.nt && o.fuse == c.fuse)
            return (int)i;
    }
    g_reg.push_back(c);
    return (int)g_reg.size() - 1;
}

TiledCfg tiled_get(int which)
{
    std::lock_guard<std::mutex> lk(g_reg_mu);
    return g_reg.at((size_t)which);
}

hipError_t launch_step_tiled(const SimConst &k, int fp32, int lagged, int energy, int which,
                             const StepPtrs &p, const DevScalars *sc, double *partials, hipStream_t st,
                             int *nblocks)
{
    TiledCfg c = tiled_get(which);
    if (nblocks) *nblocks = tiled_blocks(k, c);
    return launch_step_tiled_cfg(k, fp32, lagged, energy, c, p, sc, partials, st);
}

} // namespace cb
