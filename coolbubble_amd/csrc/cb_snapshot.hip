// cb_snapshot.hip -- device side of the run driver's field read-outs: the
// HDF5 snapshots (thinning, conversion to float, the local energy densities
// and HDF5's byte shuffle, written straight into the chunk images the host
// deflates) and wall tracking's line of the field.
//
// One thread per saved point, in lattice order, so that a wavefront's 64
// points are consecutive elements of one chunk row and each of the four
// byte planes it writes is one contiguous 64-byte run.
#include "cb_snapshot.hpp"

namespace cb {

const char *const kSnapNames[kSnapFieldsMax] = {"phi", "pi", "Ekin", "gradphi", "potEnergy"};

SnapGeom snap_geom(int D, int nGrids)
{
    SnapGeom g{};
    g.D = D;
    g.nGrids = nGrids;
    g.reduc = 1;
    if (D == 2 && nGrids > 2048) g.reduc = nGrids / 2048;
    if (D == 3 && nGrids > 512) g.reduc = nGrids / 512;
    g.n = (nGrids + g.reduc - 1) / g.reduc;
    const long long n = g.n;
    for (int d = 0; d < 3; ++d) g.ch[d] = 1;
    if (D == 1) g.ch[0] = n;
    if (D == 2) {
        long long rows = 65536 / n;
        g.ch[0] = rows < 1 ? 1 : (rows > n ? n : rows);
        g.ch[1] = n;
    }
    if (D == 3) {
        g.ch[0] = 1;
        g.ch[1] = n;
        g.ch[2] = n;
    }
    g.chunk_elems = 1;
    g.nchunks = 1;
    g.points = 1;
    for (int d = 0; d < 3; ++d) {
        g.nch[d] = d < D ? (n + g.ch[d] - 1) / g.ch[d] : 1;
        if (d < D) {
            g.chunk_elems *= g.ch[d];
            g.nchunks *= g.nch[d];
            g.points *= n;
        }
    }
    return g;
}

namespace {

// Potential V of src/bubble.cpp:69-72 in the reference's operation order.
__device__ inline double site_potential(double ph, const SimConst &k) { return vpot(ph, k); }

template <typename T, int D>
__global__ void __launch_bounds__(256) k_snapshot_pack(SimConst k, SnapGeom g, int nfields, const T *__restrict__ phi,
                                                       const T *__restrict__ pi, const T *__restrict__ phi_m,
                                                       const T *__restrict__ pi_m, const DevScalars *msc,
                                                       unsigned char *__restrict__ out)
{
    const long long N = k.N, n = g.n, r = g.reduc;
    const long long ce = g.chunk_elems;
    const size_t cbytes = g.chunk_bytes(), dbytes = g.dataset_bytes();
    double adx = 0., pk2 = 0.;
    if (nfields > 2) {
        adx = msc->adx;
        pk2 = msc->pk2;
    }
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < g.points;
         o += (long long)gridDim.x * blockDim.x) {
        // saved point -> its lattice site and its place in the chunk images
        long long c[3] = {0, 0, 0};
        long long rem = o;
        for (int d = D - 1; d >= 0; --d) {
            c[d] = rem % n;
            rem /= n;
        }
        long long site = 0, q = 0, e = 0;
        for (int d = 0; d < D; ++d) {
            site = site * N + c[d] * r;
            q = q * g.nch[d] + c[d] / g.ch[d];
            e = e * g.ch[d] + c[d] % g.ch[d];
        }
        float v[kSnapFieldsMax];
        v[0] = (float)phi[site];
        v[1] = (float)pi[site];
        if (nfields > 2) {
            // measureEnergies' local densities (src/bubble.cpp:790-833):
            // forward differences over (a dx), gx along axis 1, gy along
            // axis 0, gz along axis 2 (1D: the next point)
            const double ph = (double)phi_m[site];
            long long lc[3];
            for (int d = 0; d < D; ++d) lc[d] = c[d] * r;
            double g2[3] = {0., 0., 0.};
            for (int ax = 0; ax < D; ++ax) {
                long long nb = 0;
                for (int d = 0; d < D; ++d) {
                    long long x = lc[d];
                    if (d == ax) x = (x + 1 == N) ? 0 : x + 1;
                    nb = nb * N + x;
                }
                const double gr = ((double)phi_m[nb] - ph) / adx;
                g2[ax] = gr * gr;
            }
            double gx2, gy2, gz2;
            if (D == 1) {
                gx2 = g2[0];
                gy2 = 0.;
                gz2 = 0.;
            } else {
                gx2 = g2[1];
                gy2 = g2[0];
                gz2 = D == 3 ? g2[2] : 0.;
            }
            const double pm = (double)pi_m[site];
            v[2] = (float)(0.5 * pk2 * pm * pm);
            v[3] = (float)(0.5 * (gx2 + gy2 + gz2));
            v[4] = (float)site_potential(ph, k);
        }
        for (int f = 0; f < nfields; ++f) {
            unsigned char *img = out + (size_t)f * dbytes + (size_t)q * cbytes + (size_t)e;
            const unsigned int bits = __float_as_uint(v[f]);
            img[0] = (unsigned char)(bits & 0xffu);
            img[ce] = (unsigned char)((bits >> 8) & 0xffu);
            img[2 * ce] = (unsigned char)((bits >> 16) & 0xffu);
            img[3 * ce] = (unsigned char)(bits >> 24);
        }
    }
}

template <typename T>
hipError_t pack_t(const SimConst &k, const SnapGeom &g, int nfields, const void *phi, const void *pi,
                  const void *phi_m, const void *pi_m, const DevScalars *msc, unsigned char *out, hipStream_t st)
{
    const int block = 256;
    long long blocks = (g.points + block - 1) / block;
    if (blocks > 65536) blocks = 65536;
    const dim3 grid((unsigned)blocks);
    const T *a = (const T *)phi, *b = (const T *)pi, *am = (const T *)phi_m, *bm = (const T *)pi_m;
    switch (g.D) {
    case 1: k_snapshot_pack<T, 1><<<grid, block, 0, st>>>(k, g, nfields, a, b, am, bm, msc, out); break;
    case 2: k_snapshot_pack<T, 2><<<grid, block, 0, st>>>(k, g, nfields, a, b, am, bm, msc, out); break;
    default: k_snapshot_pack<T, 3><<<grid, block, 0, st>>>(k, g, nfields, a, b, am, bm, msc, out); break;
    }
    return hipGetLastError();
}

} // namespace

namespace {

template <typename T>
__global__ void k_gather_line(SimConst k, const T *__restrict__ phi, int c0, int c1, int c2, double *out)
{
    const long long N = k.N;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        long long site;
        if (k.D == 1) {
            site = i;
        } else {
            const long long i0 = (c0 + i) % N;
            site = k.D == 2 ? i0 * N + c1 : (i0 * N + c1) * N + c2;
        }
        out[i] = (double)phi[site];
    }
}

} // namespace

hipError_t launch_gather_line(const SimConst &k, int fp32, const void *phi, int c0, int c1, int c2, double *out,
                              hipStream_t st)
{
    const int block = 256;
    const unsigned grid = (unsigned)((k.N + block - 1) / block);
    if (fp32) k_gather_line<float><<<grid, block, 0, st>>>(k, (const float *)phi, c0, c1, c2, out);
    else k_gather_line<double><<<grid, block, 0, st>>>(k, (const double *)phi, c0, c1, c2, out);
    return hipGetLastError();
}

hipError_t launch_snapshot_pack(const SimConst &k, int fp32, const SnapGeom &g, int nfields, const void *phi,
                                const void *pi, const void *phi_m, const void *pi_m, const DevScalars *msc,
                                unsigned char *out, hipStream_t st)
{
    if (g.padded()) {
        hipError_t e = hipMemsetAsync(out, 0, (size_t)nfields * g.dataset_bytes(), st);
        if (e != hipSuccess) return e;
    }
    if (fp32) return pack_t<float>(k, g, nfields, phi, pi, phi_m, pi_m, msc, out, st);
    return pack_t<double>(k, g, nfields, phi, pi, phi_m, pi_m, msc, out, st);
}

} // namespace cb
