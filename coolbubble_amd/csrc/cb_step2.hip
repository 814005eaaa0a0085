// cb_step2.hip -- two leapfrog steps per HBM pass (temporal blocking) for a
// static box, 2D and 3D, on gfx950: one block per tile of cb_tile2.hpp,
// the launch's configurations, validity and candidates.
#include "cb_device.hpp"
#include "cb_tile2.hpp"
#include "cb_tile_common.hpp"
#include "cb_tiled.hpp"

namespace cb {

template <typename T, int D, int ITEMS, int NTHR, bool ENERGY, int PF>
__global__ void __launch_bounds__(NTHR)
k_step2_tiled(SimConst k, TiledGeom g, const T *__restrict__ phi_in, T *__restrict__ phi_out,
              const T *__restrict__ pi_in, T *__restrict__ pi_out, const DevScalars *__restrict__ sc,
              double *__restrict__ partials)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int G = gridDim.x, b = blockIdx.x;
    const int lt = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;   // XCD-contiguous tiles
    tile_step2<T, D, ITEMS, NTHR, ENERGY, PF>(k, g, lt, phi_in, phi_out, pi_in, pi_out, sc, partials,
                                          reinterpret_cast<T *>(smem_raw));
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
    if (c.pf == 4) return step3_lds_bytes(k, fp32, c);   // the direct-load tile
    const size_t esz = fp32 ? 4 : 8;
    const int rows = (k.D == 3) ? (c.TR + 4) + (c.TR + 2) : 2;
    return (size_t)rows * k.N * esz;
}

bool step2_valid(const SimConst &k, int fp32, const TiledCfg &c)
{
    if (c.pf == 4) return c.fuse == 2 && step3_valid(k, fp32, c);   // the direct-load tile, S = 2
    if (k.D < 2 || k.expanding) return false;
    const int want_vw = fp32 ? 4 : 2;
    if (c.vw != want_vw || k.N % c.vw != 0 || k.N < 4) return false;
    if (c.items != 1 && c.items != 2 && c.items != 4) return false;
    if (c.nthr != 256 && c.nthr != 512 && c.nthr != 1024) return false;
    if (c.NI < 1) return false;
    if (c.pf != 1 && !(c.pf == 2 && c.items <= 2)) return false;   // prefetch: up to two vectors per thread
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
                      double *partials, hipStream_t st, size_t shm, int nic)
{
    TiledGeom g = geom2(k, c);
    g.nic = nic;
    dim3 grid(g.ntr * g.nic), block(NTHR);
    // (energy launches prefetch only with one vector per thread: with two
    // their FP64 sums leave no registers for it; with one, FP32 at 1024
    // threads, the prefetching energy tile takes 122 VGPRs, no spills)
    if (energy && ITEMS == 1 && c.pf == 2)
        k_step2_tiled<T, D, 1, NTHR, true, 2><<<grid, block, shm, st>>>(
            k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
    else if (energy)
        k_step2_tiled<T, D, ITEMS, NTHR, true, 1><<<grid, block, shm, st>>>(
            k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
    else if constexpr (ITEMS <= 2) {
        if (c.pf == 2)
            k_step2_tiled<T, D, ITEMS, NTHR, false, 2><<<grid, block, shm, st>>>(
                k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
        else
            k_step2_tiled<T, D, ITEMS, NTHR, false, 1><<<grid, block, shm, st>>>(
                k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
    } else
        k_step2_tiled<T, D, ITEMS, NTHR, false, 1><<<grid, block, shm, st>>>(
            k, g, (const T *)p.phi_in, (T *)p.phi_out, (const T *)p.pi_in, (T *)p.pi_out, sc, partials);
}

template <typename T, int D>
static void launch2_d(const SimConst &k, int energy, const TiledCfg &c, const StepPtrs &p, const DevScalars *sc,
                      double *partials, hipStream_t st, size_t shm, int nic)
{
#define CB_L2(I, NT) launch2_t<T, D, I, NT>(k, energy, c, p, sc, partials, st, shm, nic)
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
                        const DevScalars *sc, double *partials, hipStream_t st, int *nblocks, int nchunks)
{
    if (!step2_valid(k, fp32, c)) return hipErrorInvalidValue;
    if (c.pf == 4) return launch_step3(k, fp32, energy, c, p, sc, partials, st, nblocks, nchunks);
    const size_t shm = step2_lds_bytes(k, fp32, c);
    TiledGeom g = geom2(k, c);
    if (nchunks > 0 && nchunks < g.nic) g.nic = nchunks;   // the first nchunks chunks of planes only
    if (nblocks) *nblocks = g.ntr * g.nic;
    if (fp32) {
        if (k.D == 3) launch2_d<float, 3>(k, energy, c, p, sc, partials, st, shm, g.nic);
        else launch2_d<float, 2>(k, energy, c, p, sc, partials, st, shm, g.nic);
    } else {
        if (k.D == 3) launch2_d<double, 3>(k, energy, c, p, sc, partials, st, shm, g.nic);
        else launch2_d<double, 2>(k, energy, c, p, sc, partials, st, shm, g.nic);
    }
    return hipGetLastError();
}

int step2_candidates(const SimConst &k, int fp32, TiledCfg *out, int cap)
{
    int n = 0;
    const int vw = fp32 ? 4 : 2;
    const int NV = k.N / vw;
    auto push = [&](int nthr, int TR, int NI, int nt, int pf) {
        const int rowsA = (k.D == 3) ? TR + 2 : 1;
        int items = (rowsA * NV + nthr - 1) / nthr;
        if (items == 3) items = 4;
        if (k.D == 3) {
            int hitems = (2 * NV + nthr - 1) / nthr;
            if (hitems == 3) hitems = 4;
            if (hitems > items) items = hitems;
        }
        if (items < 1) items = 1;
        if (pf == 2 && items > 2) return;
        TiledCfg c{vw, items, nthr, TR, NI, pf, nt, 2};
        if (n < cap && step2_valid(k, fp32, c)) out[n++] = c;
    };
    const int trs3[] = {6, 4, 2, 8, 12, 1};
    const int nis[] = {32, 16, 64};
    // without prefetch, then (up to two vectors per thread) with it
    for (int pf : {1, 2}) {
        if (k.D == 3) {
            for (int TR : trs3)
                for (int nthr : {256, 512, 1024})
                    for (int NI : nis)
                        for (int nt : {1, 0}) push(nthr, TR, NI, nt, pf);
        } else {
            for (int nthr : {256, 512})
                for (int NI : {16, 8, 32})
                    for (int nt : {1, 0}) push(nthr, 1, NI, nt, pf);
        }
    }
    return n;
}

} // namespace cb
