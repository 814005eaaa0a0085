// cb_tile3d.hpp -- the three-step tile of cb_tile3.hpp with phi_n staged in
// LDS by direct global-to-LDS loads (internal; included by cb_step3.hip).
//
// 3D FP32 only: eight sites per thread vector, 768 threads, tiles of TR rows
// (TR = 8 at N = 512: 12 rows of step A x 64 vectors), rows of a multiple of
// 256 sites.  The levels and steps are those of tile_step3 (step A on plane
// p for TR+4 rows, B on p-1 for TR+2, C on p-2 for the TR owned rows, each
// site's arithmetic that of three single steps, so fields are bit-identical
// to the one-step kernels).  What differs is where phi_n lives: instead of a
// ring of three planes in registers plus a register prefetch of the next
// plane and the outer rows, four LDS buffers hold the TR+6 rows of phi_n at
// planes p-1, p, p+1 (step A's i taps and its plane's j and k taps) and
// p+2, which arrives while plane p is computed.  The loads of plane p+2 go
// from HBM straight into LDS (global_load_lds_dwordx4: each lane's 16 bytes
// land at the wave's LDS base + 16 x lane, so a wave moves 1 KiB of a row),
// issued after the wait at the top of plane p into the buffer plane p-2 held
// (last read before plane p-1's closing barrier) and waited for at the top
// of plane p+1, before the barrier that publishes it.  Without the
// registers of the phi_n ring and of its prefetch the tile fits 768 threads
// at 3 waves per SIMD without spilling; pi_{n-1/2} of the next plane is
// still prefetched into registers.
#pragma once

#include <type_traits>

#include "cb_device.hpp"
#include "cb_tile3.hpp"
#include "cb_tile_common.hpp"
#include "cb_tiled.hpp"

namespace cb {

// One wave-wide direct load: lane l's 16 bytes at src (its own address)
// land at the wave-uniform LDS address dst + 16 l bytes.
__device__ __forceinline__ void lds_dma16(const float *src, float *dst)
{
    __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}

constexpr int kDmaThreads = 768;

template <bool ENERGY>
__device__ __forceinline__ void tile_step3_dma(const SimConst &k, const TiledGeom &g, int lt,
                                               const float *__restrict__ phi_in, float *__restrict__ phi_out,
                                               const float *__restrict__ pi_in, float *__restrict__ pi_out,
                                               const DevScalars *__restrict__ sc, double *__restrict__ partials,
                                               float *lds)
{
    constexpr int VW = 8;
    constexpr int D = 3;
    constexpr int W = kDmaThreads / 64;      // waves per block
    using T = float;
    using V = Vec<T, VW>;
    typedef T nvec __attribute__((ext_vector_type(VW)));

    const int N = k.N;
    const int NV = g.NV;
    const long long NN = (long long)N * N;
    const int TR = g.TR;
    const int rt = lt % g.ntr;
    const int ic = lt / g.ntr;
    const int r0 = rt * TR;
    const int i0 = ic * g.NI;
    const int i1 = min(i0 + g.NI, N);
    const int rows0 = TR + 6;                         // phi_n rows r0-3 .. r0+TR+2
    const size_t pl0 = (size_t)rows0 * N;
    float *L0b = lds;                                 // four phi_n plane buffers
    float *L1 = L0b + 4 * pl0;                        // phi_n+1 plane p-1, rows r0-2 .. r0+TR+1
    float *L2 = L1 + (size_t)(TR + 4) * N;            // phi_n+2 plane p-2, rows r0-1 .. r0+TR
    const bool nt = g.nt != 0;

    auto wrapN = [&](int i) { return i < 0 ? i + N : (i >= N ? i - N : i); };

    const T dt = (T)k.dt;
    const T fac_lap = (T)sc->fac_lap, fac_pot = (T)sc->fac_pot;
    const StepF32 sf = step_f32(k, sc);

    // this thread's vector: row rowA of the step-A rows (tile row r0-2+rowA)
    const int q = threadIdx.x;
    const bool ok = q < (TR + 4) * NV;
    const int qq = ok ? q : 0;
    const int rowA = qq / NV;
    const int col = (qq - rowA * NV) * VW;
    const long long off = (long long)wrapN(r0 - 2 + rowA) * N + col;
    const bool inB = ok && rowA >= 1 && rowA < 3 + TR;
    const bool inC = ok && rowA >= 2 && rowA < 2 + TR;
    const int l0row = (rowA + 1) * N + col;           // the vector's place in an L0 buffer

    // direct loads of one plane of phi_n into buffer b: the plane's rows0
    // rows in pieces of 256 floats, piece d by wave d mod W
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int segs = N / 256;
    const int npieces = rows0 * segs;
    auto dma_plane = [&](int p, int b) __attribute__((always_inline)) {
        const long long base = (long long)wrapN(p) * NN;
        float *dst = L0b + (size_t)b * pl0;
        for (int d = wv; d < npieces; d += W) {
            const int r = d / segs, s = d - r * segs;
            lds_dma16(phi_in + base + (long long)wrapN(r0 - 3 + r) * N + s * 256 + lane * 4, dst + r * N + s * 256);
        }
    };

    V f1[3], f2[3];                  // phi_n+1, phi_n+2 rings (as in tile_step3)
    V q1c, q1n, q2c, q2n;            // pi_n+1/2 at planes p-1, p; pi_n+3/2 at p-2, p-1
    nvec pr;                         // pi_n-1/2 of the next plane, loaded a plane ahead

    double acc[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) acc[u] = 0.0;
    double nacc[NQ];

    // buffers by the plane counter t = p - (i0 - 2): plane p in buffer (t + 1) & 3
    dma_plane(i0 - 3, 0);
    dma_plane(i0 - 2, 1);
    dma_plane(i0 - 1, 2);
    pr = (nvec)0.0f;
    if (ok) pr = *reinterpret_cast<const nvec *>(pi_in + (long long)wrapN(i0 - 2) * NN + off);

    auto plane = [&](auto s3, int p, int t) __attribute__((always_inline)) {
        constexpr int S3 = decltype(s3)::value;
        constexpr int PV = (S3 + 2) % 3, NX = (S3 + 1) % 3;
        const float *Lm = L0b + (size_t)(t & 3) * pl0;          // plane p-1
        const float *Lc = L0b + (size_t)((t + 1) & 3) * pl0;    // plane p
        const float *Lp = L0b + (size_t)((t + 2) & 3) * pl0;    // plane p+1
        // this wave's direct loads of plane p+1 and its pi prefetch have landed
        vm_wait_all();
        V pi0;
#pragma unroll
        for (int e = 0; e < VW; ++e) pi0.v[e] = pr[e];
        const bool doB = p >= i0;            // step B on plane p-1 in [i0-1, i1]
        const bool doC = p >= i0 + 2;        // step C on plane p-2 in [i0, i1-1]
        if (doB && ok) *reinterpret_cast<V *>(L1 + rowA * N + col) = f1[S3];
        if (doC && inB) *reinterpret_cast<V *>(L2 + (rowA - 1) * N + col) = f2[S3];
        // plane p+2 into the buffer plane p-2 held; pi of plane p+1
        dma_plane(p + 2, (t + 3) & 3);
        if (ok) pr = *reinterpret_cast<const nvec *>(pi_in + (long long)wrapN(p + 1) * NN + off);
        lds_barrier();
        // step A on plane p, all TR + 4 rows: phi_n at p-1, p, p+1 from LDS
        if (ok) {
            const V c = *reinterpret_cast<const V *>(Lc + l0row);
            const V mM = *reinterpret_cast<const V *>(Lm + l0row);
            const V mP = *reinterpret_cast<const V *>(Lp + l0row);
            leap_vec<T, D, VW, false>(k, c, mM, mP, Lc + (rowA + 1) * N, col, pi0, fac_lap, fac_pot, sf, dt, q1n,
                                      f1[NX], sc, nacc);
        }
        // step B on plane p-1, TR + 2 rows
        if (doB && inB)
            leap_vec<T, D, VW, false>(k, f1[S3], f1[PV], f1[NX], L1 + rowA * N, col, q1c, fac_lap, fac_pot, sf, dt,
                                      q2n, f2[NX], sc, nacc);
        // step C on plane p-2, owned rows: to HBM
        if (doC && inC) {
            const long long baseq = (long long)wrapN(p - 2) * NN;
            V pn, phn;
            leap_vec<T, D, VW, ENERGY>(k, f2[S3], f2[PV], f2[NX], L2 + (rowA - 1) * N, col, q2c, fac_lap, fac_pot,
                                       sf, dt, pn, phn, sc, acc);
            store_vec<T, VW>(pi_out + baseq + off, pn, nt);
            store_vec<T, VW>(phi_out + baseq + off, phn, nt);
        }
        lds_barrier();
        q1c = q1n;
        q2c = q2n;
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    // (p is uniform over the block: every wave leaves at the same plane)
    for (int p = i0 - 2, t = 0; p <= i1 + 1; p += 3, t += 3) {
        plane(I0{}, p, t);
        if (p + 1 > i1 + 1) break;
        plane(I1{}, p + 1, t + 1);
        if (p + 2 > i1 + 1) break;
        plane(I2{}, p + 2, t + 2);
    }
    // the loads of the last two planes' look-ahead (beyond the chunk) are
    // waited for before the block ends
    vm_wait_all();
    if constexpr (ENERGY) energy_partials(acc, k, sc, partials + (long long)lt * NQ);
}

} // namespace cb
