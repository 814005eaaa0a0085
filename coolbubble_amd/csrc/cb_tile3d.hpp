// cb_tile3d.hpp -- the multi-step tile of cb_tile3.hpp with phi_n staged in
// LDS by direct global-to-LDS loads, S = 2, 3 or 4 steps per pass (internal;
// included by cb_step3.hip).
//
// 3D only: 32-byte thread vectors (eight FP32 sites, four FP64), one vector
// per thread, rows of a multiple of 256 sites.  A block marching the slow
// axis i chains S steps plane by plane, as tile_step3 does for three: step 0 (n -> n+1) on plane p
// for TR + 2(S-1) rows, step l on plane p-l for TR + 2(S-1-l) rows, the
// last step (S-1) on plane p-S+1 for the TR owned rows, which go to HBM.
// Each site's arithmetic is that of S single steps, so fields are
// bit-identical to the one-step kernels.  Level l (phi_{n+l}, l >= 1) keeps
// its planes p-l-1, p-l, p-l+1 in a register ring and its plane p-l in LDS
// for the next step's j and k taps.  Level 0 (phi_n) lives in four LDS
// buffers instead: planes p-1, p, p+1 (step 0's i taps and its plane's j
// and k taps) and p+2, which arrives while plane p is computed.  The loads
// of plane p+2 go from HBM straight into LDS (global_load_lds_dwordx4: each
// lane's 16 bytes land at the wave's LDS base + 16 x lane, so a wave moves
// 1 KiB of a row), issued after the wait at the top of plane p into the
// buffer plane p-2 held (last read before plane p-1's closing barrier) and
// waited for at the top of plane p+1, before the barrier that publishes it.
// Without a register ring and prefetch for phi_n the tile fits eight sites
// per thread at three waves per SIMD without spilling; pi_{n-1/2} of the
// next plane is still prefetched into registers.
//   S = 2: TR = 8 at N = 512 (10 rows of step 0 x 64 vectors, 640 threads),
//          1.125x the owned rows recomputed, 116 KiB of LDS; the two-step
//          tail of the FP32 three-step tiles' calls;
//   S = 3: TR = 8 at N = 512 (12 rows of step 0 x 64 vectors, 768 threads),
//          1.25x the owned rows recomputed, 158 VGPRs;
//   S = 2, FP64: TR = 4 at N = 512 (6 rows of step 0 x 128 vectors, 768
//          threads), 1.25x the owned rows recomputed, 152 KiB of LDS;
//   S = 4: TR = 4 at N = 512 (10 rows x 64 vectors, 640 threads) would move
//          three quarters of the HBM bytes per step, but its three rings and
//          six pi vectors need more than the 168 VGPRs of three waves per
//          SIMD (200 bytes of scratch per lane: 0.575 ms/step at 3D N = 512
//          against 0.135), so S = 4 is not launched.
#pragma once

#include <type_traits>

#include "cb_device.hpp"
#include "cb_tile3.hpp"
#include "cb_tile_common.hpp"
#include "cb_tiled.hpp"

namespace cb {

// One wave-wide direct load: lane l's 16 bytes at src (its own address)
// land at the wave-uniform LDS address dst + 16 l bytes.
__device__ __forceinline__ void lds_dma16(const void *src, void *dst)
{
    __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}

// threads per block of the S-step direct-load tile: TR + 2S - 2 rows of the
// N / VW vectors of a row at N = 512 (FP32, 64 vectors a row: S = 2 and 3
// TR = 8, S = 4 TR = 4; FP64, 128 vectors a row: S = 2, TR = 4)
template <typename T, int S>
constexpr int kDmaThreads = sizeof(T) == 8 ? 768 : (S == 3 ? 768 : 640);

template <typename T, int S, bool ENERGY>
__device__ __forceinline__ void tile_stepS_dma(const SimConst &k, const TiledGeom &g, int lt,
                                               const T *__restrict__ phi_in, T *__restrict__ phi_out,
                                               const T *__restrict__ pi_in, T *__restrict__ pi_out,
                                               const DevScalars *__restrict__ sc, double *__restrict__ partials,
                                               float *lds)
{
    static_assert(S >= 2 && S <= 4, "two to four steps per pass");
    constexpr int VW = 32 / sizeof(T);      // 32-byte thread vectors
    constexpr int D = 3;
    constexpr int W = kDmaThreads<T, S> / 64;   // waves per block
    constexpr int NL = S - 1;                // register levels 1 .. S-1
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
    const int rows0 = TR + 2 * S;                     // phi_n rows r0-S .. r0+TR+S-1
    const size_t pl0 = (size_t)rows0 * N;
    T *L0b = reinterpret_cast<T *>(lds);              // four phi_n plane buffers
    // Ll[l-1]: phi_n+l at plane p-l, rows r0-S+l .. r0+TR+S-1-l (TR + 2S - 2l rows)
    T *Ll[NL];
    {
        T *q = L0b + 4 * pl0;
#pragma unroll
        for (int l = 1; l <= NL; ++l) {
            Ll[l - 1] = q;
            q += (size_t)(TR + 2 * S - 2 * l) * N;
        }
    }
    const bool nt = g.nt != 0;

    auto wrapN = [&](int i) { return i < 0 ? i + N : (i >= N ? i - N : i); };

    const T dt = (T)k.dt;
    const T fac_lap = (T)sc->fac_lap, fac_pot = (T)sc->fac_pot;
    const StepF32 sf = step_f32(k, sc);

    // this thread's vector: row rowA of step 0's rows (tile row r0-S+1+rowA)
    const int q = threadIdx.x;
    const int rowsA = TR + 2 * S - 2;
    const bool ok = q < rowsA * NV;
    const int qq = ok ? q : 0;
    const int rowA = qq / NV;
    const int col = (qq - rowA * NV) * VW;
    const long long off = (long long)wrapN(r0 - S + 1 + rowA) * N + col;
    // in[l]: the thread takes part in step l (rows l .. rowsA-1-l)
    bool in[S];
#pragma unroll
    for (int l = 0; l < S; ++l) in[l] = ok && rowA >= l && rowA < rowsA - l;
    const int l0row = (rowA + 1) * N + col;           // the vector's place in an L0 buffer

    // direct loads of one plane of phi_n into buffer b: the plane's rows0
    // rows in pieces of 1 KiB (PE elements, 16 bytes a lane), piece d by
    // wave d mod W
    constexpr int LE = 16 / sizeof(T), PE = 64 * LE;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int segs = N / PE;
    const int npieces = rows0 * segs;
    auto dma_plane = [&](int p, int b) __attribute__((always_inline)) {
        const long long base = (long long)wrapN(p) * NN;
        T *dst = L0b + (size_t)b * pl0;
        for (int d = wv; d < npieces; d += W) {
            const int r = d / segs, s = d - r * segs;
            lds_dma16(phi_in + base + (long long)wrapN(r0 - S + r) * N + s * PE + lane * LE, dst + r * N + s * PE);
        }
    };

    V f[NL][3];                      // level l+1 ring: planes p-l-2, p-l-1, p-l (by S3)
    V qc[NL], qn[NL];                // pi_{n+l+1/2} at planes p-l-1 (this trip), p-l (next)
    nvec pr;                         // pi_n-1/2 of the next plane, loaded a plane ahead

    double acc[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) acc[u] = 0.0;
    double nacc[NQ];

    // plane counter t = p - p_first: plane p in buffer (t + 1) & 3
    const int p_first = i0 - S + 1, p_last = i1 + S - 2;
    dma_plane(p_first - 1, 0);
    dma_plane(p_first, 1);
    dma_plane(p_first + 1, 2);
    pr = (nvec)(T)0;
    if (ok) pr = *reinterpret_cast<const nvec *>(pi_in + (long long)wrapN(p_first) * NN + off);

    auto plane = [&](auto s3, int p, int t) __attribute__((always_inline)) {
        constexpr int S3 = decltype(s3)::value;
        constexpr int PV = (S3 + 2) % 3, NX = (S3 + 1) % 3;
        const T *Lm = L0b + (size_t)(t & 3) * pl0;              // plane p-1
        const T *Lc = L0b + (size_t)((t + 1) & 3) * pl0;        // plane p
        const T *Lp = L0b + (size_t)((t + 2) & 3) * pl0;    // plane p+1
        // this wave's direct loads of plane p+1 and its pi prefetch have landed
        vm_wait_all();
        V pi0;
#pragma unroll
        for (int e = 0; e < VW; ++e) pi0.v[e] = pr[e];
        // step l on plane p-l (1 <= l <= S-1) runs from p >= i0 - S + 1 + 2l
        bool doS[S];
#pragma unroll
        for (int l = 0; l < S; ++l) doS[l] = p >= p_first + 2 * l;
        // stage each level's plane p-l for its step's j and k taps
#pragma unroll
        for (int l = 1; l < S; ++l)
            if (doS[l] && in[l - 1]) *reinterpret_cast<V *>(Ll[l - 1] + (rowA - l + 1) * N + col) = f[l - 1][S3];
        // plane p+2 into the buffer plane p-2 held; pi of plane p+1
        dma_plane(p + 2, (t + 3) & 3);
        if (ok) pr = *reinterpret_cast<const nvec *>(pi_in + (long long)wrapN(p + 1) * NN + off);
        lds_barrier();
        // step 0 on plane p: phi_n at p-1, p, p+1 from LDS
        if (ok) {
            const V c = *reinterpret_cast<const V *>(Lc + l0row);
            const V mM = *reinterpret_cast<const V *>(Lm + l0row);
            const V mP = *reinterpret_cast<const V *>(Lp + l0row);
            leap_vec<T, D, VW, false>(k, c, mM, mP, Lc + (rowA + 1) * N, col, pi0, fac_lap, fac_pot, sf, dt, qn[0],
                                      f[0][NX], sc, nacc);
        }
        // steps 1 .. S-2 on planes p-l
#pragma unroll
        for (int l = 1; l < S - 1; ++l)
            if (doS[l] && in[l])
                leap_vec<T, D, VW, false>(k, f[l - 1][S3], f[l - 1][PV], f[l - 1][NX], Ll[l - 1] + (rowA - l + 1) * N,
                                          col, qc[l - 1], fac_lap, fac_pot, sf, dt, qn[l], f[l][NX], sc, nacc);
        // the last step on plane p-S+1, owned rows: to HBM
        if (doS[S - 1] && in[S - 1]) {
            const long long baseq = (long long)wrapN(p - S + 1) * NN;
            V pn, phn;
            leap_vec<T, D, VW, ENERGY>(k, f[NL - 1][S3], f[NL - 1][PV], f[NL - 1][NX],
                                       Ll[NL - 1] + (rowA - S + 2) * N, col, qc[NL - 1], fac_lap, fac_pot, sf, dt,
                                       pn, phn, sc, acc);
            store_vec<T, VW>(pi_out + baseq + off, pn, nt);
            store_vec<T, VW>(phi_out + baseq + off, phn, nt);
        }
        lds_barrier();
#pragma unroll
        for (int l = 0; l < NL; ++l) qc[l] = qn[l];
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    // (p is uniform over the block: every wave leaves at the same plane)
    for (int p = p_first, t = 0; p <= p_last; p += 3, t += 3) {
        plane(I0{}, p, t);
        if (p + 1 > p_last) break;
        plane(I1{}, p + 1, t + 1);
        if (p + 2 > p_last) break;
        plane(I2{}, p + 2, t + 2);
    }
    // the direct loads of the look-ahead planes (beyond the chunk) land
    // before the block ends and its LDS goes to another block
    vm_wait_all();
    if constexpr (ENERGY) energy_partials(acc, k, sc, partials + (long long)lt * NQ);
}

} // namespace cb
