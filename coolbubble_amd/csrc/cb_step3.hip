// cb_step3.hip -- three leapfrog steps per HBM pass (temporal blocking) for a
// static box, 2D and 3D, on gfx950: one block per tile of cb_tile3.hpp, the
// launch's configurations, validity and candidates.
#include "cb_device.hpp"
#include "cb_tile3.hpp"
#include "cb_tile3d.hpp"
#include "cb_tile_common.hpp"
#include "cb_tiled.hpp"

namespace cb {

template <typename T, int D, int ITEMS, int NTHR, bool ENERGY, int PF>
__global__ void __launch_bounds__(NTHR)
k_step3_tiled(SimConst k, TiledGeom g, const T *__restrict__ phi_in, T *__restrict__ phi_out,
              const T *__restrict__ pi_in, T *__restrict__ pi_out, const DevScalars *__restrict__ sc,
              double *__restrict__ partials)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int G = gridDim.x, b = blockIdx.x;
    const int lt = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;   // XCD-contiguous tiles
    tile_step3<T, D, ITEMS, NTHR, ENERGY, PF>(k, g, lt, phi_in, phi_out, pi_in, pi_out, sc, partials,
                                          reinterpret_cast<T *>(smem_raw));
}

template <typename T, int S, bool ENERGY>
__global__ void __launch_bounds__((kDmaThreads<T, S>))
k_stepS_dma(SimConst k, TiledGeom g, const T *__restrict__ phi_in, T *__restrict__ phi_out,
            const T *__restrict__ pi_in, T *__restrict__ pi_out, const DevScalars *__restrict__ sc,
            double *__restrict__ partials)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int G = gridDim.x, b = blockIdx.x;
    const int lt = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;   // XCD-contiguous tiles
    tile_stepS_dma<T, S, ENERGY>(k, g, lt, phi_in, phi_out, pi_in, pi_out, sc, partials,
                              reinterpret_cast<float *>(smem_raw));
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
constexpr size_t kLdsWide = 160 * 1024;   // gfx950: the LDS one block may use

static TiledGeom geom3(const SimConst &k, const TiledCfg &c)
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

size_t step3_lds_bytes(const SimConst &k, int fp32, const TiledCfg &c)
{
    const size_t esz = fp32 ? 4 : 8;
    if (c.pf == 4) {
        // the direct-load form: four planes of phi_n (TR + 2S rows), then
        // level l's plane (TR + 2S - 2l rows), l = 1 .. S-1
        const int S = c.fuse;
        int rows = 4 * (c.TR + 2 * S);
        for (int l = 1; l < S; ++l) rows += c.TR + 2 * S - 2 * l;
        return (size_t)rows * k.N * esz;
    }
    const int rows = (k.D == 3) ? (c.TR + 6) + (c.TR + 4) + (c.TR + 2) : 3;
    return (size_t)rows * k.N * esz;
}

// The wide tile (cb_tile3d.hpp): 3D, 32-byte thread vectors (eight FP32
// sites, four FP64), one vector per thread, phi_n loaded straight into four
// LDS plane buffers (prefetch kind 4).  FP32, three steps: a tile of TR = 8
// rows fits one block (12 rows of step A x 64 vectors at N = 512, 768
// threads), so the rows recomputed for steps A and B are 1.25x the owned
// ones instead of 1.5x at TR = 4; two steps: 10 rows, 640 threads.  FP64,
// two steps: TR = 4 (6 rows x 128 vectors, 768 threads, 152 KiB of LDS).
// At most 160 KiB of LDS; rows of a multiple of 256 sites (whole waves of
// 16-byte pieces).
static bool step3_wide(const SimConst &k, int fp32, const TiledCfg &c)
{
    if (k.D != 3 || c.items != 1 || c.vw != (fp32 ? 8 : 4)) return false;
    if (fp32) return (c.fuse == 3 && c.nthr == kDmaThreads<float, 3>) || (c.fuse == 2 && c.nthr == kDmaThreads<float, 2>);
    return c.fuse == 2 && c.nthr == kDmaThreads<double, 2>;
}

bool step3_valid(const SimConst &k, int fp32, const TiledCfg &c)
{
    const bool wide = step3_wide(k, fp32, c);
    if (k.D < 2 || k.expanding || (c.fuse != 3 && !wide)) return false;
    const int want_vw = fp32 ? 4 : 2;
    if ((c.vw != want_vw && !wide) || k.N % c.vw != 0 || k.N < 8) return false;
    if (c.items != 1 && c.items != 2 && c.items != 4) return false;
    if (c.nthr != 256 && c.nthr != 512 && c.nthr != 1024 && !wide) return false;
    if (wide && !(c.pf == 4 && k.N % 256 == 0)) return false;
    if (wide && c.fuse == 2 && (c.TR + 2) * (k.N / c.vw) != c.nthr) return false;   // every thread a vector
    if (c.NI < 1 || (c.nt != 0 && c.nt != 1)) return false;
    if (!wide && c.pf != 1 && !((c.pf == 2 || c.pf == 3) && c.items == 1)) return false;   // prefetch: one vector per thread
    if (k.D == 3 && (c.TR < 1 || k.N % c.TR != 0)) return false;
    const int NV = k.N / c.vw;
    const int rowsA = (k.D == 3) ? c.TR + 2 * c.fuse - 2 : 1;
    if ((long long)rowsA * NV > (long long)c.items * c.nthr) return false;
    if (k.D == 3 && 2 * NV > c.items * c.nthr) return false;
    if (step3_lds_bytes(k, fp32, c) > (wide ? kLdsWide : 65536)) return false;
    return true;
}

template <typename T, int D, int ITEMS, int NTHR>
static void launch3_t(const SimConst &k, int energy, const TiledCfg &c, const StepPtrs &p, const DevScalars *sc,
                      double *partials, hipStream_t st, size_t shm, int nic)
{
    TiledGeom g = geom3(k, c);
    g.nic = nic;
    dim3 grid(g.ntr * g.nic), block(NTHR);
    // (the energy launches do not prefetch: their FP64 sums take the registers)
    if (energy)
        k_step3_tiled<T, D, ITEMS, NTHR, true, 1><<<grid, block, shm, st>>>(
            k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
    else if constexpr (ITEMS == 1) {
        if (c.pf == 3)
            k_step3_tiled<T, D, 1, NTHR, false, 3><<<grid, block, shm, st>>>(
                k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
        else if (c.pf == 2)
            k_step3_tiled<T, D, 1, NTHR, false, 2><<<grid, block, shm, st>>>(
                k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
        else
            k_step3_tiled<T, D, 1, NTHR, false, 1><<<grid, block, shm, st>>>(
                k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
    } else
        k_step3_tiled<T, D, ITEMS, NTHR, false, 1><<<grid, block, shm, st>>>(
            k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
}

template <typename T, int D>
static void launch3_d(const SimConst &k, int energy, const TiledCfg &c, const StepPtrs &p, const DevScalars *sc,
                      double *partials, hipStream_t st, size_t shm, int nic)
{
#define CB_L3(I, NT) launch3_t<T, D, I, NT>(k, energy, c, p, sc, partials, st, shm, nic)
    if (c.nthr == 256) {
        if (c.items == 1) CB_L3(1, 256);
        else if (c.items == 2) CB_L3(2, 256);
        else CB_L3(4, 256);
    } else if (c.nthr == 512) {
        if (c.items == 1) CB_L3(1, 512);
        else if (c.items == 2) CB_L3(2, 512);
        else CB_L3(4, 512);
    } else {
        if (c.items == 1) CB_L3(1, 1024);
        else if (c.items == 2) CB_L3(2, 1024);
        else CB_L3(4, 1024);
    }
#undef CB_L3
}

hipError_t launch_step3(const SimConst &k, int fp32, int energy, const TiledCfg &c, const StepPtrs &p,
                        const DevScalars *sc, double *partials, hipStream_t st, int *nblocks, int nchunks)
{
    if (!step3_valid(k, fp32, c)) return hipErrorInvalidValue;
    const size_t shm = step3_lds_bytes(k, fp32, c);
    TiledGeom g = geom3(k, c);
    if (nchunks > 0 && nchunks < g.nic) g.nic = nchunks;   // the first nchunks chunks of planes only
    if (nblocks) *nblocks = g.ntr * g.nic;
    if (step3_wide(k, fp32, c)) {
        // phi_n staged by direct loads into LDS (cb_tile3d.hpp); the
        // energies on the same tile
        dim3 grid(g.ntr * g.nic), block(c.nthr);
        if (!fp32) {
            const double *pin = (const double *)p.phi_in, *qin = (const double *)p.pi_in;
            double *pout = (double *)p.phi_out, *qout = (double *)p.pi_out;
            const void *kern = energy ? (const void *)k_stepS_dma<double, 2, true> : (const void *)k_stepS_dma<double, 2, false>;
            hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            if (e != hipSuccess) return e;
            if (energy) k_stepS_dma<double, 2, true><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
            else k_stepS_dma<double, 2, false><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
            return hipGetLastError();
        }
        const float *pin = (const float *)p.phi_in, *qin = (const float *)p.pi_in;
        float *pout = (float *)p.phi_out, *qout = (float *)p.pi_out;
        const void *kern = c.fuse == 2 ? (energy ? (const void *)k_stepS_dma<float, 2, true> : (const void *)k_stepS_dma<float, 2, false>)
                                       : (energy ? (const void *)k_stepS_dma<float, 3, true> : (const void *)k_stepS_dma<float, 3, false>);
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
        if (c.fuse == 2) {
            if (energy) k_stepS_dma<float, 2, true><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
            else k_stepS_dma<float, 2, false><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
        } else {
            if (energy) k_stepS_dma<float, 3, true><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
            else k_stepS_dma<float, 3, false><<<grid, block, shm, st>>>(k, g, pin, pout, qin, qout, sc, partials);
        }
        return hipGetLastError();
    }
    if (fp32) {
        if (k.D == 3) launch3_d<float, 3>(k, energy, c, p, sc, partials, st, shm, g.nic);
        else launch3_d<float, 2>(k, energy, c, p, sc, partials, st, shm, g.nic);
    } else {
        if (k.D == 3) launch3_d<double, 3>(k, energy, c, p, sc, partials, st, shm, g.nic);
        else launch3_d<double, 2>(k, energy, c, p, sc, partials, st, shm, g.nic);
    }
    return hipGetLastError();
}

int step3_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap)
{
    int n = 0;
    const int vw = fp32 ? 4 : 2;
    const int NV = k.N / vw;
    auto push = [&](int nthr, int TR, int NI, int nt, int pf) {
        const int rowsA = (k.D == 3) ? TR + 4 : 1;
        int items = (rowsA * NV + nthr - 1) / nthr;
        if (items == 3) items = 4;
        if (k.D == 3) {
            int hitems = (2 * NV + nthr - 1) / nthr;
            if (hitems == 3) hitems = 4;
            if (hitems > items) items = hitems;
        }
        if (items < 1) items = 1;
        // (more vectors per thread than this spill registers)
        if (items > (nthr == 256 ? 4 : nthr == 512 ? 2 : 1)) return;
        if (pf >= 2 && items != 1) return;
        TiledCfg c{vw, items, nthr, TR, NI, pf, nt, 3};
        if (n < cap && step3_valid(k, fp32, c)) out[n++] = c;
    };
    // without prefetch, then (one vector per thread) with it
    if (k.D == 3) {
        for (int TR : {4, 2, 6, 8, 1})
            for (int nthr : {512, 1024, 256})
                for (int NI : {32, 16})
                    for (int nt : {1, 0}) push(nthr, TR, NI, nt, 1);
        for (int TR : {4, 2, 6, 8, 1})
            for (int nthr : {1024, 512, 256})
                for (int NI : {32, 16, 64})
                    for (int nt : {1, 0}) push(nthr, TR, NI, nt, 2);
        // loads two planes ahead (1024 threads: one block per CU, so the
        // loads of one plane ahead land after the plane's arithmetic)
        for (int TR : {4, 2})
            for (int NI : {64, 32})
                for (int nt : {1, 0}) push(1024, TR, NI, nt, 3);
        // one chunk of 128 planes per tile (half the chunk-end planes of 64)
        for (int nt : {1, 0}) push(1024, 4, 128, nt, 2);
        // the wide tile (FP32): eight sites per thread, 768 threads, phi_n
        // loaded straight into LDS
        if (fp32)
            for (int NI : {128, 64})
                for (int nt : {1, 0}) {
                    TiledCfg c{8, 1, kDmaThreads<float, 3>, 8, NI, 4, nt, 3};
                    if (n < cap && step3_valid(k, fp32, c) && 12 * (k.N / 8) > 512) out[n++] = c;
                }
        // its two-step form (ten rows of 64 vectors, 640 threads), listed
        // last so that the other variants keep their indices
        if (fp32)
            for (int NI : {128, 64})
                for (int nt : {1, 0}) {
                    TiledCfg c{8, 1, kDmaThreads<float, 2>, 8, NI, 4, nt, 2};
                    if (n < cap && step3_valid(k, fp32, c)) out[n++] = c;
                }
        // 256-thread tiles without prefetch in 64-plane chunks (an
        // ensemble's members), also listed last
        for (int TR : {4, 2})
            for (int nt : {1, 0}) push(256, TR, 64, nt, 1);
        // the FP64 two-step direct-load tile (six rows of 128 vectors of
        // four sites, 768 threads; at NI = 256 one block per CU), last
        if (!fp32)
            for (int NI : {64, 128, 256})
                for (int nt : {1, 0}) {
                    TiledCfg c{4, 1, kDmaThreads<double, 2>, 4, NI, 4, nt, 2};
                    if (n < cap && step3_valid(k, fp32, c)) out[n++] = c;
                }
        // 1024-thread tiles in chunks of 256 planes (FP32 3D N = 512: 256
        // blocks, one round of one block per CU), last
        for (int pf : {2, 3})
            for (int nt : {1, 0}) push(1024, 4, 256, nt, pf);

    } else {
        for (int nthr : {256, 512})
            for (int NI : {16, 32, 8})
                for (int nt : {1, 0}) push(nthr, 1, NI, nt, 1);
        for (int nthr : {256, 512})
            for (int NI : {16, 32, 8})
                for (int nt : {1, 0}) push(nthr, 1, NI, nt, 2);
    }
    return n;
}

} // namespace cb
