// cb_fft.hip -- power spectrum and thermal initial conditions on gfx950:
// a hand-written Stockham FFT in LDS (no FFT library), radial k-binning, and
// the seeded Gaussian random field of mode -3.
//
// Reference: FourierTransform + ComputePowerSpectrum (src/bubble.cpp:642-709)
// use an FFTW r2c of phi (FFTW_FORWARD, unnormalised) into the half-complex
// layout (rows = N^(D-1), N/2+1 columns), then bin |phi_k|^2 in shells of
// width 2 pi / L.  GenerateThermalFluc (src/bubble.cpp:459-523) fills that
// layout with mt19937_64 + normal_distribution draws, enforces Hermitian
// symmetry on the planes j = 0 and j = N/2 and runs an FFTW c2r.
//
// Transforms here:
//   rows   : length-N transform of each contiguous row, r2c or c2r
//   columns: length-N transform along a strided axis of the half-complex
//            array, CB adjacent columns per block so every row segment is a
//            coalesced load
// each as a radix-4/2/3 (+ any prime) Stockham sweep between two LDS
// buffers with twiddles from a table computed in long double on the host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "cb_device.hpp"
#include "cb_fft.hpp"
#include "cb_sim.hpp"

namespace cb {

// ---------------------------------------------------------------------------
// Device: batched Stockham FFT in LDS.  Sequence b, element e at
// buf[b * bs + e * es].  sign = -1 forward, +1 backward; tw[e] =
// exp(-2 pi i e / n).  Returns the buffer holding the result.
// ---------------------------------------------------------------------------
template <typename R>
__device__ inline cpx<R> twid(const cpx<R> *tw, int e, int n, int sign)
{
    e %= n;
    cpx<R> w = tw[e];
    if (sign > 0) w.y = -w.y;
    return w;
}

template <typename R>
__device__ cpx<R> *lds_fft(cpx<R> *A, cpx<R> *B, int batch, int bs, int es, const Stages &st,
                           const cpx<R> *__restrict__ tw, int sign)
{
    const int n = st.n;
    int Ns = 1;
    for (int s = 0; s < st.ns; ++s) {
        const int Rd = st.radix[s];
        const int m = n / Rd;
        const int total = batch * m;
        const int tstep = n / (Ns * Rd);      // twiddle index step for angle 1/(Ns*Rd)
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int b = idx / m;
            const int j = idx - b * m;
            const int k = j % Ns;
            const cpx<R> *in = A + b * bs;
            cpx<R> *out = B + b * bs;
            const int dst = (j / Ns) * Ns * Rd + k;
            if (Rd == 2 || Rd == 4 || Rd == 3) {
                cpx<R> v[4];
                for (int r = 0; r < Rd; ++r) {
                    cpx<R> a = in[(j + r * m) * es];
                    if (r > 0 && k > 0) a = cmul(a, twid(tw, r * k * tstep, n, sign));
                    v[r] = a;
                }
                if (Rd == 2) {
                    out[dst * es] = {v[0].x + v[1].x, v[0].y + v[1].y};
                    out[(dst + Ns) * es] = {v[0].x - v[1].x, v[0].y - v[1].y};
                } else if (Rd == 4) {
                    cpx<R> a0 = {v[0].x + v[2].x, v[0].y + v[2].y};
                    cpx<R> a1 = {v[0].x - v[2].x, v[0].y - v[2].y};
                    cpx<R> b0 = {v[1].x + v[3].x, v[1].y + v[3].y};
                    cpx<R> b1 = {v[1].x - v[3].x, v[1].y - v[3].y};
                    // (-sign i) * b1 : forward multiplies by -i
                    cpx<R> rb1 = (sign < 0) ? cpx<R>{b1.y, -b1.x} : cpx<R>{-b1.y, b1.x};
                    out[dst * es] = {a0.x + b0.x, a0.y + b0.y};
                    out[(dst + Ns) * es] = {a1.x + rb1.x, a1.y + rb1.y};
                    out[(dst + 2 * Ns) * es] = {a0.x - b0.x, a0.y - b0.y};
                    out[(dst + 3 * Ns) * es] = {a1.x - rb1.x, a1.y - rb1.y};
                } else {
                    const R h = (R)0.86602540378443864676;   // sqrt(3)/2
                    cpx<R> t1 = {v[1].x + v[2].x, v[1].y + v[2].y};
                    cpx<R> t2 = {v[0].x - (R)0.5 * t1.x, v[0].y - (R)0.5 * t1.y};
                    cpx<R> d = {v[1].x - v[2].x, v[1].y - v[2].y};
                    // (-sign i) * h * d
                    cpx<R> t3 = (sign < 0) ? cpx<R>{h * d.y, -h * d.x} : cpx<R>{-h * d.y, h * d.x};
                    out[dst * es] = {v[0].x + t1.x, v[0].y + t1.y};
                    out[(dst + Ns) * es] = {t2.x + t3.x, t2.y + t3.y};
                    out[(dst + 2 * Ns) * es] = {t2.x - t3.x, t2.y - t3.y};
                }
            } else {
                // any radix: direct length-Rd DFT of the twiddled inputs
                for (int q = 0; q < Rd; ++q) {
                    R sx = 0, sy = 0;
                    for (int r = 0; r < Rd; ++r) {
                        cpx<R> a = in[(j + r * m) * es];
                        if (r > 0 && k > 0) a = cmul(a, twid(tw, r * k * tstep, n, sign));
                        cpx<R> w = twid(tw, ((r * q) % Rd) * m, n, sign);
                        cpx<R> p = cmul(a, w);
                        sx += p.x;
                        sy += p.y;
                    }
                    out[(dst + q * Ns) * es] = {sx, sy};
                }
            }
        }
        __syncthreads();
        cpx<R> *t = A;
        A = B;
        B = t;
        Ns *= Rd;
    }
    return A;
}

// r2c of RB contiguous rows of length N: real field -> half-complex rows.
template <typename T, typename R>
__global__ void __launch_bounds__(256)
k_rows_r2c(const T *__restrict__ field, cpx<R> *__restrict__ H, long long rows, int RB, Stages st,
           const cpx<R> *__restrict__ tw)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N = st.n, nc = N / 2 + 1;
    cpx<R> *A = reinterpret_cast<cpx<R> *>(smem_raw);
    cpx<R> *B = A + (size_t)RB * N;
    const long long r0 = (long long)blockIdx.x * RB;
    const int nr = (int)((rows - r0) < RB ? (rows - r0) : RB);
    const T *src = field + r0 * N;
    // loads in batches of 8 per thread, all issued before any is waited for
    for (int i0 = threadIdx.x; i0 < RB * N; i0 += 8 * blockDim.x) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = i0 + u * (int)blockDim.x;
            v[u] = idx < nr * N ? src[idx] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = i0 + u * (int)blockDim.x;
            if (idx < RB * N) A[idx] = cpx<R>{(R)v[u], (R)0};
        }
    }
    __syncthreads();
    cpx<R> *res = lds_fft<R>(A, B, RB, N, 1, st, tw, -1);
    cpx<R> *dst = H + r0 * nc;
    for (int idx = threadIdx.x; idx < nr * nc; idx += blockDim.x) {
        const int rr = idx / nc, j = idx - rr * nc;
        dst[idx] = res[rr * N + j];
    }
}

// c2r of RB rows: Hermitian-extended row, backward transform, real part,
// scaled (src/bubble.cpp:515-519): field = x * norm + shift.
template <typename T, typename R>
__global__ void __launch_bounds__(256)
k_rows_c2r(const cpx<R> *__restrict__ H, T *__restrict__ field, long long rows, int RB, Stages st,
           const cpx<R> *__restrict__ tw, double norm, double shift)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N = st.n, nc = N / 2 + 1;
    cpx<R> *A = reinterpret_cast<cpx<R> *>(smem_raw);
    cpx<R> *B = A + (size_t)RB * N;
    const long long r0 = (long long)blockIdx.x * RB;
    const int nr = (int)((rows - r0) < RB ? (rows - r0) : RB);
    for (int idx = threadIdx.x; idx < RB * N; idx += blockDim.x) {
        const int rr = idx / N, e = idx - rr * N;
        cpx<R> v = {(R)0, (R)0};
        if (rr < nr) {
            const cpx<R> *row = H + (r0 + rr) * nc;
            if (e == 0) v = {row[0].x, (R)0};
            else if (2 * e < N) v = row[e];
            else if (2 * e == N) v = {row[e].x, (R)0};
            else v = {row[N - e].x, -row[N - e].y};
        }
        A[idx] = v;
    }
    __syncthreads();
    cpx<R> *res = lds_fft<R>(A, B, RB, N, 1, st, tw, +1);
    T *dst = field + r0 * N;
    for (int idx = threadIdx.x; idx < nr * N; idx += blockDim.x)
        dst[idx] = (T)((double)res[idx].x * norm + shift);
}

// Transform along a strided axis: array [outer][len][inner], CB adjacent
// columns of one outer index per block, in place.
template <typename R>
__global__ void __launch_bounds__(256)
k_cols(cpx<R> *H, long long inner, int CB, Stages st, const cpx<R> *__restrict__ tw, int sign)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int len = st.n;
    cpx<R> *A = reinterpret_cast<cpx<R> *>(smem_raw);
    cpx<R> *B = A + (size_t)CB * len;
    const long long tiles = (inner + CB - 1) / CB;
    const long long o = blockIdx.x / tiles;
    const long long c0 = (blockIdx.x - o * tiles) * CB;
    const int nc = (int)((inner - c0) < CB ? (inner - c0) : CB);
    cpx<R> *base = H + o * (long long)len * inner + c0;
    // loads in batches of 8 per thread, all issued before any is waited for
    for (int i0 = threadIdx.x; i0 < len * CB; i0 += 8 * blockDim.x) {
        cpx<R> v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = i0 + u * (int)blockDim.x;
            const int e = idx / CB, cc = idx - e * CB;
            v[u] = (idx < len * CB && cc < nc) ? base[(long long)e * inner + cc] : cpx<R>{(R)0, (R)0};
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int idx = i0 + u * (int)blockDim.x;
            const int e = idx / CB, cc = idx - e * CB;
            if (idx < len * CB) A[cc * len + e] = v[u];
        }
    }
    __syncthreads();
    cpx<R> *res = lds_fft<R>(A, B, CB, len, 1, st, tw, sign);
    for (int idx = threadIdx.x; idx < len * CB; idx += blockDim.x) {
        const int e = idx / CB, cc = idx - e * CB;
        if (cc < nc) base[(long long)e * inner + cc] = res[cc * len + e];
    }
}

// Radial binning of the half-complex spectrum, ComputePowerSpectrum
// (src/bubble.cpp:668-709), in a fixed order with no atomics.  A block
// takes a contiguous run of rows, RI rows at a time; each row is cut into K
// chunks of E consecutive modes, one chunk per thread (K = ceil(nc / E),
// RI K <= 256).  Along a row the bin never decreases with j, so the modes
// of one bin form one run (a segment).  A thread sums the runs in its
// chunk; a run that crosses into the next chunks is completed by the
// thread where it starts, from those chunks' first partial sums, in order.
// Each (row, bin) gets one total, written by one thread into the row's
// slot; the RI slots of each bin are added in a fixed order at the end, one
// row of (power, weight, weighted k) partials per block.
// GEOM: weights and weighted k only (they do not depend on the field; H is
// not read).
template <typename R, bool GEOM = false>
__global__ void __launch_bounds__(256)
k_bin(const cpx<R> *__restrict__ H, long long rows, int N, int D, int nBins, double dk, double pref,
      double Pnorm, int E, int K, int RI, double *__restrict__ binpart)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *hp = reinterpret_cast<double *>(smem_raw);   // [RI][nBins] power
    double *hw = hp + (size_t)RI * nBins;                  // weights
    double *hk = hw + (size_t)RI * nBins;                  // weighted k
    double *rfs = hk + (size_t)RI * nBins;                 // [3][256] first-run sums
    int *rfk = reinterpret_cast<int *>(rfs + 3 * 256);     // first-run key (-1: none at the chunk start)
    int *rlk = rfk + 256;                                  // last-run key (-1: none at the chunk end)
    int *rone = rlk + 256;                                 // one run over the whole chunk
    for (int i = threadIdx.x; i < 3 * RI * nBins; i += blockDim.x) hp[i] = 0.0;
    const int nc = N / 2 + 1;
    const long long per = (rows + gridDim.x - 1) / gridDim.x;
    const long long r0 = per * blockIdx.x;
    const long long r1 = (r0 + per < rows) ? r0 + per : rows;
    const int t = threadIdx.x, slot = t / K, chunk = t - slot * K;
    for (long long rb = r0; rb < r1; rb += RI) {
        __syncthreads();   // the previous rows' records are read
        const long long r = rb + slot;
        const bool rowok = slot < RI && r < r1;
        long long S = 0;   // n2rows, an integer
        if (rowok && D == 2) {
            const long long s = (r <= N / 2) ? r : r - N;
            S = s * s;
        } else if (rowok && D == 3) {
            const long long a = r / N, c = r % N;
            const long long s1 = (a <= N / 2) ? a : a - N, s2 = (c <= N / 2) ? c : c - N;
            S = s1 * s1 + s2 * s2;
        }
        double *slot_p = hp + (size_t)slot * nBins, *slot_w = hw + (size_t)slot * nBins, *slot_k = hk + (size_t)slot * nBins;
        // runs of the chunk: A (first) and Z (last) kept, the others written
        int nseg = 0, ak = -1, zk = -1;
        bool a0 = false, zE = false;
        double ap = 0, aw = 0, akk = 0, zp = 0, zw = 0, zkk = 0;
        const int j0 = chunk * E;
        for (int u = 0; u < E; ++u) {
            const long long j = j0 + u;
            int key = -1;
            double p = 0.0, w = 0.0, wk = 0.0;
            if (rowok && j < nc) {
                const double n = sqrt((double)S + (double)j * j);
                const int bin = (int)n;
                if (bin < nBins) {
                    key = bin;
                    w = (j == 0 || j == N / 2) ? 1.0 : 2.0;
                    const double kc = n * dk;
                    if (!GEOM) {
                        const cpx<R> f = H[r * nc + j];
                        const double fx = (double)f.x, fy = (double)f.y;
                        p = w * pref * ipow_rn(kc, D) * Pnorm * (fx * fx + fy * fy);
                    }
                    wk = w * kc;
                }
            }
            if (key >= 0 && key == zk && nseg > 0 && zE) {          // the current run goes on
                zp += p;
                zw += w;
                zkk += wk;
                continue;
            }
            // the current run (Z) closes: write it unless it is the first run from u = 0
            if (nseg > 0 && zE && !(nseg == 1 && a0)) {
                slot_p[zk] += zp;
                slot_w[zk] += zw;
                slot_k[zk] += zkk;
            }
            if (nseg == 1 && a0 && zE) {   // the first run from u = 0 ends inside the chunk
                ap = zp;
                aw = zw;
                akk = zkk;
            }
            zE = false;
            if (key >= 0) {                // a new run
                if (nseg == 0) {
                    ak = key;
                    a0 = (u == 0);
                }
                ++nseg;
                zk = key;
                zp = p;
                zw = w;
                zkk = wk;
                zE = true;                 // open; stays true while it runs to u = E - 1
            }
        }
        // zE: the last run is open at u = E - 1 (it ends at the chunk end)
        const bool one = nseg == 1 && a0 && zE;
        if (nseg == 1 && a0 && zE) {
            ap = zp;
            aw = zw;
            akk = zkk;
        }
        rfk[t] = (nseg > 0 && a0) ? ak : -1;
        rfs[t] = ap;
        rfs[256 + t] = aw;
        rfs[512 + t] = akk;
        rlk[t] = (nseg > 0 && zE) ? zk : -1;
        rone[t] = one ? 1 : 0;
        __syncthreads();
        auto walk = [&](int key, double &p, double &w, double &k) {
            for (int u2 = t + 1; u2 < 256 && u2 / K == slot && rfk[u2] == key; ++u2) {
                p += rfs[u2];
                w += rfs[256 + u2];
                k += rfs[512 + u2];
                if (!rone[u2]) break;
            }
        };
        if (nseg > 0 && a0) {
            const bool continued = chunk > 0 && rlk[t - 1] == ak;
            if (!continued) {
                double p = ap, w = aw, k = akk;
                if (one) walk(ak, p, w, k);
                slot_p[ak] += p;
                slot_w[ak] += w;
                slot_k[ak] += k;
            }
        }
        if (nseg > 0 && zE && !one) {
            double p = zp, w = zw, k = zkk;
            walk(zk, p, w, k);
            slot_p[zk] += p;
            slot_w[zk] += w;
            slot_k[zk] += k;
        }
    }
    __syncthreads();
    double *dst = binpart + (long long)blockIdx.x * 3 * nBins;
    for (int b = threadIdx.x; b < nBins; b += blockDim.x) {
        double p = 0.0, w = 0.0, k = 0.0;
        for (int q = 0; q < RI; ++q) {
            p += hp[(size_t)q * nBins + b];
            w += hw[(size_t)q * nBins + b];
            k += hk[(size_t)q * nBins + b];
        }
        dst[b] = p;
        dst[nBins + b] = w;
        dst[2 * nBins + b] = k;
    }
}

// Chunk shape of k_bin for rows of N / 2 + 1 modes: E modes per thread (at
// least 8), K chunks per row, RI rows per pass of 256 threads; LDS bytes.
struct BinShape {
    int E, K, RI;
    size_t shm;
};
static BinShape bin_shape(int N, int nBins)
{
    const int nc = N / 2 + 1;
    BinShape b;
    b.E = std::max(8, (nc + 255) / 256);
    b.K = (nc + b.E - 1) / b.E;
    b.RI = std::max(1, 256 / b.K);
    b.shm = sizeof(double) * (3 * (size_t)b.RI * nBins + 3 * 256) + sizeof(int) * 3 * 256;
    return b;
}

// Fixed-order sum of the per-block bin partials: 64 bins per block of
// 64 * BR_WAVES threads.  Wave w sums rows w, w + W, w + 2W, ... in that
// order (Neumaier), eight rows' loads issued before any is added (a lone
// small grid has nothing else to hide their latency behind); the W wave sums
// are then merged in wave order.  The order of every addition depends only
// on (nblk, W), so the sums are the same on every call.
constexpr int BR_WAVES = 16;
__global__ void __launch_bounds__(64 * BR_WAVES)
k_bin_reduce(const double *__restrict__ binpart, int nblk, int nq, double *__restrict__ out)
{
    constexpr int W = BR_WAVES;
    __shared__ double ss[W][64], sc[W][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q = blockIdx.x * 64 + lane;
    Neu s = {0.0, 0.0};
    if (q < nq) {
        int b = w;
        for (; b + 7 * W < nblk; b += 8 * W) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = binpart[(long long)(b + u * W) * nq + q];
#pragma unroll
            for (int u = 0; u < 8; ++u) s.add(v[u]);
        }
        for (; b < nblk; b += W) s.add(binpart[(long long)b * nq + q]);
    }
    ss[w][lane] = s.s;
    sc[w][lane] = s.c;
    __syncthreads();
    if (w == 0 && q < nq) {
        Neu t = {ss[0][lane], sc[0][lane]};
        for (int k = 1; k < W; ++k) t.merge(Neu{ss[k][lane], sc[k][lane]});
        out[q] = t.value();
    }
}

// ---------------------------------------------------------------------------
// Thermal modes (src/bubble.cpp:469-488): one thread per row r, an
// mt19937_64 seeded with seed + r (state words interleaved across rows, so
// the 312-word twists are coalesced), libstdc++'s generate_canonical and
// Marsaglia polar method, sigma = sqrt(P(T, k^2, m) / 2), re drawn before im.
// The state lives in global memory, where the compiler cannot tell the
// words apart, so every access would wait for the one before it: the twist
// loads 4 words of each operand before it stores 4 new ones (larger blocks
// cost more registers than the occupancy they leave is worth), and the
// outputs are read 8 words at a time into registers.
// ---------------------------------------------------------------------------
struct MtRow {
    unsigned long long *st;
    long long rows, r;
    int mti;                      // next word to read into the ring
    int rpos;                     // next word of the ring to temper
    unsigned long long ring[8];
    __device__ unsigned long long &w(int i) { return st[(long long)i * rows + r]; }
    __device__ void seed(unsigned long long s)
    {
        unsigned long long x = s;
        w(0) = x;
        for (int i = 1; i < 312; ++i) {
            x = 6364136223846793005ULL * (x ^ (x >> 62)) + (unsigned long long)i;
            w(i) = x;
        }
        mti = 312;
        rpos = 8;
    }
    static __device__ unsigned long long mix(unsigned long long hi, unsigned long long lo, unsigned long long far)
    {
        const unsigned long long UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL, A = 0xB5026F5AA96619E9ULL;
        const unsigned long long x = (hi & UM) | (lo & LM);
        return far ^ (x >> 1) ^ ((x & 1ULL) ? A : 0ULL);
    }
    // new words i0 .. i0+K-1 from old words i0 .. i0+K and words i0+off ..
    // (old ones above the block, new ones below it)
    template <int K>
    __device__ void twist_block(int i0, int off)
    {
        unsigned long long a[K + 1], b[K];
#pragma unroll
        for (int k = 0; k <= K; ++k) a[k] = w(i0 + k);
#pragma unroll
        for (int k = 0; k < K; ++k) b[k] = w(i0 + off + k);
#pragma unroll
        for (int k = 0; k < K; ++k) w(i0 + k) = mix(a[k], a[k + 1], b[k]);
    }
    __device__ void twist()
    {
#pragma unroll 1
        for (int i0 = 0; i0 < 156; i0 += 4) twist_block<4>(i0, 156);      // 39 blocks
#pragma unroll 1
        for (int i0 = 156; i0 < 300; i0 += 4) twist_block<4>(i0, -156);    // 36 blocks
#pragma unroll 1
        for (int i = 300; i < 311; ++i) w(i) = mix(w(i), w(i + 1), w(i - 156));
        w(311) = mix(w(311), w(0), w(155));
        mti = 0;
    }
    __device__ unsigned long long next()
    {
        if (rpos == 8) {
            if (mti >= 312) twist();
#pragma unroll
            for (int k = 0; k < 8; ++k) ring[k] = w(mti + k);   // 312 = 39 x 8
            mti += 8;
            rpos = 0;
        }
        unsigned long long y = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k == rpos) y = ring[k];
        rpos += 1;
        y ^= (y >> 29) & 0x5555555555555555ULL;
        y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
        y ^= (y << 37) & 0xFFF7EEE000000000ULL;
        y ^= (y >> 43);
        return y;
    }
    __device__ double canonical()
    {
        double r = (double)next() / 18446744073709551616.0;
        if (r >= 1.0) r = 0x1.fffffffffffffp-1;
        return r;
    }
};

__global__ void __launch_bounds__(256)
k_thermal_modes(cpx<double> *H, unsigned long long *state, long long rows, int N, int D, double L, double T,
                double m, int velocity, unsigned long long seed, int rj)
{
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const double PI = 3.14159265358979323846;
    const int nc = N / 2 + 1;
    MtRow g{state, rows, r, 0, 8, {}};
    g.seed(seed + (unsigned long long)r);
    auto kphys = [&](long long i) -> double {
        const long long s = (i <= N / 2) ? i : i - N;
        return 2.0 * PI * (double)s / L;
    };
    // Rayleigh-Jeans option: the lattice momentum of mode s along an axis,
    // (2 / dx) sin(pi s / N), the eigenvalue of the force's Laplacian
    const double tdx = 2.0 * (double)N / L;
    auto klat = [&](long long i) -> double { return tdx * sin(PI * (double)i / (double)N); };
    double k2rows, kl2rows;
    if (D == 2) {
        const double q = kphys(r), ql = klat(r);
        k2rows = q * q;
        kl2rows = ql * ql;
    } else {
        const double q1 = kphys(r / N), q2 = kphys(r % N), l1 = klat(r / N), l2 = klat(r % N);
        k2rows = q1 * q1 + q2 * q2;
        kl2rows = l1 * l1 + l2 * l2;
    }
    bool have = false;
    double saved = 0.0;
    auto normal = [&]() -> double {
        if (have) { have = false; return saved; }
        double x, y, r2;
        do {
            x = 2.0 * g.canonical() - 1.0;
            y = 2.0 * g.canonical() - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double mult = sqrt(-2 * log(r2) / r2);
        saved = x * mult;
        have = true;
        return y * mult;
    };
    for (int j = 0; j < nc; ++j) {
        double P;
        if (rj) {
            const double lj = klat(j);
            const double w2 = kl2rows + lj * lj + m * m;
            P = velocity ? T : T / w2;
        } else {
            const double kj = 2.0 * PI * j / L;
            const double k2 = k2rows + kj * kj;
            const double omega = sqrt(k2 + m * m);
            const double n_k = 1.0 / (exp(omega / T) - 1.0);
            P = velocity ? n_k * omega : n_k / omega;
        }
        const double sigma = sqrt(P / 2.);
        const double re = sigma * (normal() * 1.0 + 0.0);
        const double im = sigma * (normal() * 1.0 + 0.0);
        H[r * nc + j] = {re, im};
    }
}

// Hermitian symmetry on the planes j = 0 and j = N/2 (src/bubble.cpp:493-511).
// Each row only reads a partner with a smaller index, which it never writes,
// so the rows are independent.
__global__ void __launch_bounds__(256) k_hermitian(cpx<double> *H, long long rows, int N, int D)
{
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int nc = N / 2 + 1;
    const long long partner = (D == 2) ? (N - r) % N : ((N - r / N) % N) * N + (N - r % N) % N;
    const int planes[2] = {0, N / 2};
    for (int q = 0; q < 2; ++q) {
        const int j = planes[q];
        cpx<double> *mode = H + r * nc + j;
        if (partner == r) {
            mode->x *= 1.4142135623730951;   // sqrt(2.)
            mode->y = 0.;
        } else if (partner < r) {
            const cpx<double> p = H[partner * nc + j];
            mode->x = p.x;
            mode->y = -p.y;
        }
    }
}

// ---------------------------------------------------------------------------
// Long 1D transforms (N beyond one in-LDS transform): four-step.  The real
// row x[N2 n1 + n2] is viewed as A[n1][n2]; transform along n1 (strided
// columns), multiply by W_N^(k1 n2), transform along n2 (rows); then
// X[k1 + N1 k2] = C[k1][k2].
// ---------------------------------------------------------------------------
template <typename T, typename R>
__global__ void __launch_bounds__(256) k_real_to_cpx(const T *__restrict__ f, cpx<R> *__restrict__ H, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) H[i] = {(R)f[i], (R)0};
}

template <typename R>
__global__ void __launch_bounds__(256)
k_twiddle4(cpx<R> *H, int N1, int N2, const cpx<R> *__restrict__ twN, long long N)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)N1 * N2) return;
    const long long k1 = i / N2, n2 = i - k1 * N2;
    const long long e = (k1 * n2) % N;
    H[i] = cmul(H[i], twN[e]);
}

// 1D bins: shell j holds the single mode j (n = sqrt(j^2) = j), so each
// bin is written directly (ComputePowerSpectrum, src/bubble.cpp:680-701).
template <typename R>
__global__ void __launch_bounds__(256)
k_bin1d(const cpx<R> *__restrict__ C, int N1, int N2, int N, int nBins, double dk, double pref, double Pnorm,
        double *__restrict__ bins)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nBins) return;
    double pw = 0.0, ws = 0.0, ks = 0.0;
    if (j <= N / 2) {
        const int k1 = j % N1, k2 = j / N1;
        const cpx<R> f = C[(long long)k1 * N2 + k2];
        const double n = sqrt(0. + (double)j * j);
        const double w = (j == 0 || j == N / 2) ? 1.0 : 2.0;
        const double kc = n * dk;
        const double fx = (double)f.x, fy = (double)f.y;
        pw = w * pref * ipow_rn(kc, 1) * Pnorm * (fx * fx + fy * fy);
        ws = w;
        ks = w * kc;
    }
    bins[j] = pw;
    bins[nBins + j] = ws;
    bins[2 * nBins + j] = ks;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct FftPlan {
    int N = 0, D = 0;
    bool fp32 = false;
    Stages st{};
    void *tw = nullptr;        // cpx<R> table, n entries
    void *H = nullptr;         // half-complex work array
    long long rows = 0;
    int nBins = 0, binblocks = 0;
    double *binpart = nullptr, *bins = nullptr;
    // power-of-two D = 2 / 3: H transposed ([N/2+1][N][ld] or [N/2+1][ld],
    // ld = p2_pitch(N, fp32)),
    // binning fused into the last pass, the field-independent weights and
    // weighted k of each bin computed once into bins[nBins, 3 nBins) (the
    // pass itself writes bins[0, nBins))
    bool fused = false;
    int fused_blocks = 0;
    // long 1D rows (four-step): N = N1 N2
    int N1 = 0, N2 = 0;
    Stages st1{}, st2{};
    void *tw1 = nullptr, *tw2 = nullptr;
    // device buffers from hipMalloc (a simulation's own plan, freed by
    // fft_destroy), or carved from the no-allocation scratch
    bool owned = true;
    hipStream_t pool_st = nullptr;   // owned buffers from the memory pool, in order on this stream
};

// Device memory for a plan: hipMalloc when arena is null, the memory pool
// in order on `pool` when the arena has no base, else the next 256-byte-
// aligned piece of the arena.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    hipStream_t pool = nullptr;
};

static hipError_t dev_get(Arena *a, void **out, size_t bytes)
{
    if (!a) return hipMalloc(out, bytes);
    if (!a->base) return hipMallocAsync(out, bytes, a->pool);
    const size_t off = (a->used + 255) & ~(size_t)255;
    if (off > a->cap || bytes > a->cap - off) return hipErrorOutOfMemory;
    *out = a->base + off;
    a->used = off + bytes;
    return hipSuccess;
}

static bool factorise(int n, Stages &st)
{
    st.n = n;
    st.ns = 0;
    int m = n;
    while (m % 4 == 0) { st.radix[st.ns++] = 4; m /= 4; }
    while (m % 2 == 0) { st.radix[st.ns++] = 2; m /= 2; }
    while (m % 3 == 0) { st.radix[st.ns++] = 3; m /= 3; }
    for (int p = 5; m > 1 && p * p <= m; p += 2)
        while (m % p == 0) { st.radix[st.ns++] = p; m /= p; if (st.ns >= kMaxStages) return false; }
    if (m > 1) st.radix[st.ns++] = m;
    return st.ns <= kMaxStages;
}

// Twiddles exp(-2 pi i e / n), e < n, computed in long double on the host
// (x87 cosl / sinl: a mix of served requests, 1D N = 65536 among them, spent
// 5 ms per request's plan on them before the tables were kept).
template <typename R>
static void fill_twiddles(cpx<R> *h, int n)
{
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int e = 0; e < n; ++e) {
        long double a = twopi * (long double)e / (long double)n;
        h[e] = {(R)cosl(a), (R)-sinl(a)};
    }
}

// The table for n, computed once per process and kept (entries are never
// removed, so the pointer stays valid); null beyond 1024 distinct lengths.
template <typename R>
static const cpx<R> *kept_twiddles(int n)
{
    static std::mutex mu;
    static std::map<int, std::vector<cpx<R>>> kept;
    std::lock_guard<std::mutex> lk(mu);
    auto it = kept.find(n);
    if (it != kept.end()) return it->second.data();
    if (kept.size() >= 1024) return nullptr;
    std::vector<cpx<R>> &v = kept[n];
    v.resize((size_t)n);
    fill_twiddles(v.data(), n);
    return v.data();
}

// The table copied to device memory from dev_get.  stage: a host buffer of
// at least n entries to compute the table in (the no-allocation path: no
// host allocation either); null: the kept table.
template <typename R>
static hipError_t make_twiddles(int n, void **out, hipStream_t s, Arena *arena = nullptr, void *stage = nullptr)
{
    std::vector<cpx<R>> tmp;
    const cpx<R> *h = nullptr;
    if (stage) {
        fill_twiddles(reinterpret_cast<cpx<R> *>(stage), n);
        h = reinterpret_cast<const cpx<R> *>(stage);
    } else {
        h = kept_twiddles<R>(n);
        if (!h) {
            tmp.resize((size_t)n);
            fill_twiddles(tmp.data(), n);
            h = tmp.data();
        }
    }
    hipError_t err = dev_get(arena, out, sizeof(cpx<R>) * (size_t)n);
    if (err != hipSuccess) return err;
    err = hipMemcpyAsync(*out, h, sizeof(cpx<R>) * (size_t)n, hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    return err;
}

// A kernel launched with more than 64 KiB of dynamic LDS asks for it first.
template <typename K>
static hipError_t lds_opt_in(K kernel, size_t shm)
{
    if (shm <= 65536) return hipSuccess;
    if (shm > kLdsBlockMax) return hipErrorInvalidValue;
    return hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
}

// Rows / columns per block for an LDS budget.
static int batch_for(int n, size_t csz)
{
    size_t per = 2 * (size_t)n * csz;
    int b = (int)(kLdsBlockMax / per);
    return b < 1 ? 0 : (b > 16 ? 16 : b);
}

// All axis transforms of one D-dim half-complex array, forward from a real
// field (r2c) or backward into one (c2r).
template <typename R, typename T>
static hipError_t forward_r2c(const Stages &st, const cpx<R> *tw, const T *field, cpx<R> *H, int D,
                              hipStream_t s)
{
    const int N = st.n, nc = N / 2 + 1;
    const long long rows = (D == 3) ? (long long)N * N : (D == 2 ? N : 1);
    if (p2_supported(N)) {
        const int f32 = sizeof(R) == 4;
        hipError_t e = p2_rows_r2c(f32, field, H, rows, nc, N, tw, s);
        if (e == hipSuccess && D == 3) e = p2_cols(f32, H, N, nc, nc, N, tw, -1, s);        // axis 1
        if (e == hipSuccess && D >= 2) e = p2_cols(f32, H, 1, rows / N * nc, rows / N * nc, N, tw, -1, s); // axis 0
        return e;
    }
    const int RB = batch_for(N, sizeof(cpx<R>));
    if (RB < 1) return hipErrorInvalidValue;
    size_t shm = 2 * (size_t)RB * N * sizeof(cpx<R>);
    hipError_t e = lds_opt_in(k_rows_r2c<T, R>, shm);
    if (e != hipSuccess) return e;
    k_rows_r2c<T, R><<<(unsigned)((rows + RB - 1) / RB), 256, shm, s>>>(field, H, rows, RB, st, tw);
    const int CB = batch_for(N, sizeof(cpx<R>));
    shm = 2 * (size_t)CB * N * sizeof(cpx<R>);
    e = lds_opt_in(k_cols<R>, shm);
    if (e != hipSuccess) return e;
    if (D == 3) {   // axis 1: [N][N][nc]
        const long long tiles = ((long long)nc + CB - 1) / CB;
        k_cols<R><<<(unsigned)(N * tiles), 256, shm, s>>>(H, nc, CB, st, tw, -1);
    }
    if (D >= 2) {   // axis 0: [1][N][rows/N * nc]
        const long long inner = rows / N * nc;
        const long long tiles = (inner + CB - 1) / CB;
        k_cols<R><<<(unsigned)tiles, 256, shm, s>>>(H, inner, CB, st, tw, -1);
    }
    return hipGetLastError();
}

// Four-step spectrum of one long row into p->bins.
template <typename R>
static hipError_t long_row(FftPlan *p, const R *field, double pref, double dk, double Pnorm, hipStream_t st)
{
    const int N = p->N, N1 = p->N1, N2 = p->N2;
    cpx<R> *H = (cpx<R> *)p->H;
    k_real_to_cpx<R, R><<<(N + 255) / 256, 256, 0, st>>>(field, H, N);
    const int CB = batch_for(N1, sizeof(cpx<R>));
    size_t shm = 2 * (size_t)CB * N1 * sizeof(cpx<R>);
    hipError_t e = lds_opt_in(k_cols<R>, shm);
    if (e != hipSuccess) return e;
    k_cols<R><<<(unsigned)((N2 + CB - 1) / CB), 256, shm, st>>>(H, N2, CB, p->st1, (const cpx<R> *)p->tw1, -1);
    k_twiddle4<R><<<(N + 255) / 256, 256, 0, st>>>(H, N1, N2, (const cpx<R> *)p->tw, N);
    shm = 2 * (size_t)N2 * sizeof(cpx<R>);
    e = lds_opt_in(k_cols<R>, shm);
    if (e != hipSuccess) return e;
    k_cols<R><<<(unsigned)N1, 256, shm, st>>>(H, 1, 1, p->st2, (const cpx<R> *)p->tw2, -1);
    k_bin1d<R><<<(p->nBins + 255) / 256, 256, 0, st>>>(H, N1, N2, N, p->nBins, dk, pref, Pnorm, p->bins);
    return hipGetLastError();
}

void fft_destroy(FftPlan *p)
{
    if (!p) return;
    if (!p->owned) {
        delete p;
        return;
    }
    void *bufs[6] = {p->tw, p->tw1, p->tw2, p->H, p->binpart, p->bins};
    for (void *b : bufs) {
        if (!b) continue;
        if (p->pool_st) hipFreeAsync(b, p->pool_st);
        else hipFree(b);
    }
    delete p;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                \
            return CB_HIP_ERROR;                                                         \
        }                                                                                \
    } while (0)

// Shape of the transforms for s (no device work): factorisation, four-step
// split of long 1D rows, fused binning, bin and block counts.
static cb_status plan_shape(const cb_sim *s, FftPlan *p)
{
    const int N = s->p.nGrids, D = s->p.D;
    p->N = N;
    p->D = D;
    p->fp32 = s->fp32;
    const size_t csz0 = s->fp32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    bool ok = factorise(N, p->st);
    if (ok && batch_for(N, csz0) < 1) {
        ok = false;
        if (D == 1) {   // four-step: N1 N2 = N, both within one in-LDS transform
            int best = 0;
            for (int n2 = 2; n2 <= N / 2; ++n2) {
                if (N % n2) continue;
                const int n1 = N / n2;
                if (batch_for(n1, csz0) < 1 || batch_for(n2, csz0) < 1) continue;
                if (!best || std::abs(n2 - n1) < std::abs(best - N / best)) best = n2;
            }
            if (best) {
                p->N2 = best;
                p->N1 = N / best;
                ok = factorise(p->N1, p->st1) && factorise(p->N2, p->st2);
            }
        }
    }
    if (!ok) {
        set_error("FFT length not supported by the in-LDS transform");
        return CB_UNSUPPORTED;
    }
    p->rows = (D == 3) ? (long long)N * N : (D == 2 ? N : 1);
    p->fused = D >= 2 && !p->N1 && p2_supported(N);
    p->nBins = s->nkmax + 1;
    const long long modes = p->rows * (N / 2 + 1);
    p->binblocks = (int)std::max<long long>(1, std::min<long long>(256, (modes + 4095) / 4096));
    if (p->fused) p->fused_blocks = p2_bin_blocks_max();
    if (!p->N1) {   // every path but the four-step one bins with k_bin (the fused one its weights)
        const size_t shm = bin_shape(N, p->nBins).shm;
        if (shm > kLdsBlockMax) {
            set_error("too many spectrum bins for LDS");
            return CB_UNSUPPORTED;
        }
    }
    return CB_OK;
}

// Device buffers of a shaped plan (hipMalloc, or carved from arena), its
// twiddles (host staging in stage, or a temporary), and the fused path's
// per-bin weights and weighted k.
static hipError_t plan_buffers(FftPlan *p, Arena *arena, void *stage, double L, hipStream_t st)
{
    const int N = p->N, D = p->D, nc = N / 2 + 1;
    const size_t csz = p->fp32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    const size_t hbytes = csz * (p->N1 ? (size_t)N
                                       : p->fused ? (size_t)p->rows / N * nc * p2_pitch(N, p->fp32) : (size_t)p->rows * nc);
    size_t part = 3 * (size_t)p->nBins * p->binblocks;
    if (p->fused) part = std::max(part, (size_t)p->nBins * p->fused_blocks);   // the fused pass's partial rows
    // every buffer first, so that a scratch too small fails before any launch
    hipError_t e = dev_get(arena, &p->H, hbytes);
    if (e == hipSuccess) e = dev_get(arena, (void **)&p->binpart, sizeof(double) * part);
    if (e == hipSuccess) e = dev_get(arena, (void **)&p->bins, sizeof(double) * 3 * (size_t)p->nBins);
    if (e != hipSuccess) return e;
    auto tw = [&](int n, void **out) {
        return p->fp32 ? make_twiddles<float>(n, out, st, arena, stage) : make_twiddles<double>(n, out, st, arena, stage);
    };
    e = tw(N, &p->tw);
    if (e == hipSuccess && p->N1) e = tw(p->N1, &p->tw1);
    if (e == hipSuccess && p->N1) e = tw(p->N2, &p->tw2);
    if (e == hipSuccess && p->fused) {
        const int nB = p->nBins;
        const BinShape bs = bin_shape(N, nB);
        e = lds_opt_in(k_bin<double, true>, bs.shm);
        if (e != hipSuccess) return e;
        const double dk = 2.0 * 3.14159265358979323846 / L;
        k_bin<double, true><<<p->binblocks, 256, bs.shm, st>>>(nullptr, p->rows, N, D, nB, dk, 0.0, 0.0, bs.E, bs.K,
                                                               bs.RI, p->binpart);
        k_bin_reduce<<<(3 * nB + 63) / 64, 64 * BR_WAVES, 0, st>>>(p->binpart, p->binblocks, 3 * nB, p->bins);
        e = hipGetLastError();
    }
    return e;
}

static cb_status plan_for(cb_sim *s)
{
    if (s->fft) return CB_OK;
    FftPlan *p = new FftPlan();
    cb_status r = plan_shape(s, p);
    if (r != CB_OK) {
        delete p;
        return r;
    }
    Arena pool{};
    pool.pool = s->pool_st;
    p->pool_st = s->pool_st;
    hipError_t e = plan_buffers(p, s->pool_st ? &pool : nullptr, nullptr, s->p.L, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) {
        fft_destroy(p);
        set_error(std::string("FFT plan: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? CB_OUT_OF_MEMORY : CB_HIP_ERROR;
    }
    s->fft = p;
    return CB_OK;
}

// The spectrum of s's current phi with the plan p; h: host buffer of
// 3 nBins doubles.
static cb_status spectrum_run(cb_sim *s, FftPlan *p, double *h, double *kmean, double *kphys, double *delta,
                              int cap, int *nbins)
{
    const int D = s->p.D, N = s->p.nGrids, nB = p->nBins;
    const double L = s->p.L;
    const void *field = s->phi[s->iphi];
    const double PI = 3.14159265358979323846;
    const double dk = 2.0 * PI / L;
    const double pref = (D == 1) ? 1.0 / PI : (D == 2) ? 1.0 / (2.0 * PI) : 1.0 / (2.0 * PI * PI);
    const double Pnorm = s->dxD * s->dxD / std::pow(L, D);
    if (p->fused) {
        // rows r2c -> H [j][i1][i0] (3D) or [j][i0] (2D); (3D) along i1 in
        // place; along i0 fused with the binning
        const int f32 = s->fp32 ? 1 : 0;
        const long long nc = N / 2 + 1;
        const long long ld = p2_pitch(N, f32);
        HIP_TRY(p2_rows_r2c_t(f32, field, p->H, D == 3 ? N : 1, ld, N, p->tw, s->st));
        if (D == 3) HIP_TRY(p2_cols(f32, p->H, nc, ld, N, N, p->tw, -1, s->st));
        int nblk = 0;
        HIP_TRY(p2_rows_bin(f32, p->H, p->rows / N * nc, ld, N, p->tw, D, nB, dk, pref, Pnorm, p->binpart,
                            p->fused_blocks, &nblk, s->st));
        k_bin_reduce<<<(nB + 63) / 64, 64 * BR_WAVES, 0, s->st>>>(p->binpart, nblk, nB, p->bins);
        HIP_TRY(hipGetLastError());
    } else if (p->N1) {
        HIP_TRY(s->fp32 ? long_row<float>(p, (const float *)field, pref, dk, Pnorm, s->st)
                        : long_row<double>(p, (const double *)field, pref, dk, Pnorm, s->st));
    } else {
        hipError_t e;
        if (s->fp32) e = forward_r2c<float, float>(p->st, (const cpx<float> *)p->tw, (const float *)field, (cpx<float> *)p->H, D, s->st);
        else e = forward_r2c<double, double>(p->st, (const cpx<double> *)p->tw, (const double *)field, (cpx<double> *)p->H, D, s->st);
        HIP_TRY(e);
        const BinShape bs = bin_shape(N, nB);
        if (bs.shm > kLdsBlockMax ||
            (s->fp32 ? lds_opt_in(k_bin<float>, bs.shm) : lds_opt_in(k_bin<double>, bs.shm)) != hipSuccess) {
            set_error("too many spectrum bins for LDS");
            return CB_UNSUPPORTED;
        }
        if (s->fp32)
            k_bin<float><<<p->binblocks, 256, bs.shm, s->st>>>((const cpx<float> *)p->H, p->rows, N, D, nB, dk, pref, Pnorm,
                                                                bs.E, bs.K, bs.RI, p->binpart);
        else
            k_bin<double><<<p->binblocks, 256, bs.shm, s->st>>>((const cpx<double> *)p->H, p->rows, N, D, nB, dk, pref,
                                                                 Pnorm, bs.E, bs.K, bs.RI, p->binpart);
        HIP_TRY(hipGetLastError());
        k_bin_reduce<<<(3 * nB + 63) / 64, 64 * BR_WAVES, 0, s->st>>>(p->binpart, p->binblocks, 3 * nB, p->bins);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(h, p->bins, sizeof(double) * 3 * nB, hipMemcpyDeviceToHost, s->st));
    cb::DevScalars sc;
    HIP_TRY(hipMemcpyAsync(&sc, s->dsc, sizeof sc, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    int nout = 0;
    for (int b = 0; b < nB; ++b) {
        const double ws = h[(size_t)nB + b];
        if (ws > 0) {
            const double km = h[2 * (size_t)nB + b] / ws;
            if (nout < cap) {
                if (kmean) kmean[nout] = km;
                if (kphys) kphys[nout] = km / sc.a;
                if (delta) delta[nout] = h[(size_t)b] / ws;
            }
            ++nout;
        }
    }
    *nbins = nout;
    return CB_OK;
}

cb_status spectrum(cb_sim *s, double *kmean, double *kphys, double *delta, int cap, int *nbins)
{
    cb_status r = plan_for(s);
    if (r != CB_OK) return r;
    std::vector<double> h(3 * (size_t)s->fft->nBins);
    return spectrum_run(s, s->fft, h.data(), kmean, kphys, delta, cap, nbins);
}

// ---------------------------------------------------------------------------
// The spectrum with no allocation (cb_power_spectrum_noalloc).  The plan's
// device buffers (half-complex array, twiddles, bin partials and bins) are
// carved from `scratch`, a static device array of kNoAllocScratchBytes
// that each GPU holds for the life of the process; the host stages the
// twiddles and reads the bins back through static host buffers.  The device
// address of each GPU's copy comes from hipGetSymbolAddress on that GPU's
// first call (the array's name is only a host-side handle).  The scratch
// keeps the last call's contents; every buffer that must start at zero is
// cleared by a kernel on each call.  Two calls on one GPU whose work
// overlapped would race on it, so calls in one process run one at a time
// under a lock held until the call's stream has drained (also on the error
// paths).
// ---------------------------------------------------------------------------
static constexpr size_t kNoAllocScratchBytes = CB_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
constexpr int kMaxStageTw = 65536;          // longest twiddle table (1D N = 2^16)
constexpr int kMaxStageBins = 3 * 32769;    // 3 x bins of 1D N = 2^16
std::mutex g_noalloc_mu;
void *g_scratch_dev[kMaxDevices] = {};
cpx<double> g_stage_tw[kMaxStageTw];
double g_stage_bins[kMaxStageBins];
} // namespace

cb_status spectrum_noalloc(cb_sim *s, double *kmean, double *kphys, double *delta, int cap, int *nbins)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    const int dev = s->p.device;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device ordinal out of range for the no-allocation scratch");
        return CB_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(dev));
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) {
            set_error("no device address for the no-allocation scratch");
            return CB_HIP_ERROR;
        }
        g_scratch_dev[dev] = ptr;
    }
    FftPlan p;
    p.owned = false;
    cb_status r = plan_shape(s, &p);
    if (r != CB_OK) return r;
    if (p.N > kMaxStageTw || 3 * p.nBins > kMaxStageBins) {
        set_error("lattice too large for the no-allocation spectrum");
        return CB_UNSUPPORTED;
    }
    Arena arena{(char *)g_scratch_dev[dev], kNoAllocScratchBytes, 0};
    hipError_t e = plan_buffers(&p, &arena, g_stage_tw, s->p.L, s->st);
    if (e == hipErrorOutOfMemory) {   // raised before any launch
        set_error("lattice too large for the no-allocation spectrum: its work arrays exceed the " +
                  std::to_string(kNoAllocScratchBytes >> 20) + " MiB scratch");
        return CB_UNSUPPORTED;
    }
    if (e == hipSuccess) {
        r = spectrum_run(s, &p, g_stage_bins, kmean, kphys, delta, cap, nbins);
        if (r == CB_OK) return r;
    } else {
        set_error(std::string("no-allocation spectrum: ") + hipGetErrorString(e));
        r = CB_HIP_ERROR;
    }
    (void)hipStreamSynchronize(s->st);   // nothing of this call may still run on the scratch
    return r;
}

// Mode -3 field (GenerateThermalFluc, src/bubble.cpp:459-523), computed in
// double whatever the storage precision.
cb_status thermal_field(cb_sim *s, void *field, double T, double m, int velocity)
{
    const int N = s->p.nGrids, D = s->p.D;
    if (D < 2) return CB_OK;   // the reference never seeds a 1D thermal field
    Stages st;
    if (!factorise(N, st) || batch_for(N, sizeof(cpx<double>)) < 1) {
        set_error("FFT length not supported by the in-LDS transform");
        return CB_UNSUPPORTED;
    }
    const long long rows = (D == 3) ? (long long)N * N : N;
    const int nc = N / 2 + 1;
    void *tw = nullptr, *H = nullptr, *state = nullptr;
    auto cleanup = [&]() {
        sim_free(s, tw);
        sim_free(s, H);
        sim_free(s, state);
    };
    Arena pool{};
    pool.pool = s->pool_st;
    hipError_t e = make_twiddles<double>(N, &tw, s->st, s->pool_st ? &pool : nullptr);
    if (e == hipSuccess) e = sim_alloc(s, &H, sizeof(cpx<double>) * (size_t)rows * nc);
    if (e == hipSuccess) e = sim_alloc(s, &state, sizeof(unsigned long long) * 312 * (size_t)rows);
    if (e != hipSuccess) {
        cleanup();
        set_error(std::string("thermal IC: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? CB_OUT_OF_MEMORY : CB_HIP_ERROR;
    }
    // thermalSeed, src/bubble.cpp:134-142
    const char *spec = velocity ? "velocitythermalspec" : "thermalspec";
    unsigned long long off = 0;
    for (const char *c = spec; *c; ++c) off = off * 131 + (unsigned char)*c;
    const unsigned long long seed = (unsigned long long)s->p.thermal_base + off;
    const unsigned nb = (unsigned)((rows + 255) / 256);
    k_thermal_modes<<<nb, 256, 0, s->st>>>((cpx<double> *)H, (unsigned long long *)state, rows, N, D, s->p.L, T, m,
                                           velocity, seed, s->p.thermal_spectrum == 1 ? 1 : 0);
    k_hermitian<<<nb, 256, 0, s->st>>>((cpx<double> *)H, rows, N, D);
    const int CB = batch_for(N, sizeof(cpx<double>));
    const size_t shm = 2 * (size_t)CB * N * sizeof(cpx<double>);
    const double norm = 1.0 / std::sqrt(std::pow(s->p.L, D));
    if (p2_supported(N)) {
        e = p2_cols(0, H, 1, rows / N * nc, rows / N * nc, N, tw, +1, s->st);                 // axis 0
        if (e == hipSuccess && D == 3) e = p2_cols(0, H, N, nc, nc, N, tw, +1, s->st);   // axis 1
        if (e == hipSuccess) e = p2_rows_c2r(s->fp32, H, field, rows, N, tw, norm, 0.0, s->st);
        if (e == hipSuccess) e = hipStreamSynchronize(s->st);
        cleanup();
        if (e != hipSuccess) { set_error(std::string("thermal IC: ") + hipGetErrorString(e)); return CB_HIP_ERROR; }
        return CB_OK;
    }
    if (lds_opt_in(k_cols<double>, shm) != hipSuccess || lds_opt_in(k_rows_c2r<float, double>, shm) != hipSuccess ||
        lds_opt_in(k_rows_c2r<double, double>, shm) != hipSuccess) {
        cleanup();
        set_error("thermal IC: FFT length too long for LDS");
        return CB_UNSUPPORTED;
    }
    {   // axis 0 then (3D) axis 1, backward
        const long long inner = rows / N * nc;
        const long long tiles = (inner + CB - 1) / CB;
        k_cols<double><<<(unsigned)tiles, 256, shm, s->st>>>((cpx<double> *)H, inner, CB, st, (const cpx<double> *)tw, +1);
        if (D == 3) {
            const long long t2 = ((long long)nc + CB - 1) / CB;
            k_cols<double><<<(unsigned)(N * t2), 256, shm, s->st>>>((cpx<double> *)H, nc, CB, st, (const cpx<double> *)tw, +1);
        }
    }
    const int RB = CB;
    if (s->fp32)
        k_rows_c2r<float, double><<<(unsigned)((rows + RB - 1) / RB), 256, shm, s->st>>>((const cpx<double> *)H, (float *)field, rows, RB, st, (const cpx<double> *)tw, norm, 0.0);
    else
        k_rows_c2r<double, double><<<(unsigned)((rows + RB - 1) / RB), 256, shm, s->st>>>((const cpx<double> *)H, (double *)field, rows, RB, st, (const cpx<double> *)tw, norm, 0.0);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    cleanup();
    if (e != hipSuccess) { set_error(std::string("thermal IC: ") + hipGetErrorString(e)); return CB_HIP_ERROR; }
    return CB_OK;
}

} // namespace cb
