// cb_fft_p2.hip -- power-of-two FFTs (N = 8 .. 2048) for the spectrum and
// the thermal initial conditions, gfx950.
//
// Stockham autosort with the radices fixed at compile time: radix-8 stages,
// then one radix-4 or radix-2 stage.  A sequence is processed by N/8 threads;
// in each stage a thread loads the inputs of its 8/R butterflies from LDS
// into registers, applies the twiddles, then after a barrier writes the
// butterfly outputs back in place.  A 256-thread block transforms 2048/N
// sequences at a time (16 KiB of FP32 or 32 KiB of FP64 data), so several
// blocks share a CU and one block's loads overlap another's arithmetic.
// LDS positions carry one padding slot per 16 elements, which spreads the
// power-of-two strides of the Stockham writes over the banks.
//   rows r2c: two real rows per complex sequence (z = a + i b), split after
//             the transform: A_k = (Z_k + conj Z_{N-k}) / 2,
//             B_k = (Z_k - conj Z_{N-k}) / 2i;
//   rows c2r: two Hermitian rows per sequence (Z = A + i B), real and
//             imaginary parts of the backward transform;
//   columns : 2048/N adjacent columns of a strided axis, transposed into LDS;
//             the last forward pass of a spectrum bins the modes instead of
//             writing them back (ComputePowerSpectrum, src/bubble.cpp:668-709).
// The kernels are persistent: a grid of as many blocks as are resident at
// once walks the work items, split into 8 contiguous ranges, one per XCD
// (blocks are dispatched to the XCDs round-robin), so blocks that share cache
// lines (neighbouring columns) run at the same time behind the same L2.
#include "cb_fft.hpp"

namespace cb {

namespace {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
constexpr int nstages(int lg) { return lg / 3 + (lg % 3 ? 1 : 0); }
constexpr int stage_radix(int lg, int s) { return s < lg / 3 ? 8 : (lg % 3 == 2 ? 4 : 2); }

// LDS slot of element x of a sequence
__host__ __device__ constexpr int pidx(int x) { return x + (x >> 4); }
// sequences per block: one round of 256 threads at N/8 threads per sequence
template <int N>
constexpr int kSeq = (2048 / N) < 1 ? 1 : 2048 / N;
// LDS stride between sequences (rows / transposed columns)
template <int N>
constexpr int kRowStride = pidx(N) + 1;
template <typename R, int N>
constexpr int kColStride = pidx(N) + (sizeof(R) == 4 ? (N >= 256 ? 4 : 1) : (N >= 512 ? 2 : 1));
// twiddle table staged in LDS when it is small, else read through the caches
template <typename R, int N>
constexpr bool kTwLds = (size_t)N * 2 * sizeof(R) <= 8192;

// multiply by -i (forward) or +i (backward)
template <typename R>
__device__ inline cpx<R> rot(cpx<R> a, int sign)
{
    return sign < 0 ? cpx<R>{a.y, -a.x} : cpx<R>{-a.y, a.x};
}

template <typename R>
__device__ inline cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ inline cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }

template <typename R>
__device__ inline void dft2(cpx<R> *v)
{
    const cpx<R> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename R>
__device__ inline void dft4(cpx<R> &a0, cpx<R> &a1, cpx<R> &a2, cpx<R> &a3, int sign)
{
    const cpx<R> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = rot(csub(a1, a3), sign);
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <typename R>
__device__ inline void dft8(cpx<R> *v, int sign)
{
    cpx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3, sign);
    dft4(o0, o1, o2, o3, sign);
    const R c = (R)0.70710678118654752440;
    // W8^1 = c (1 - i) forward, c (1 + i) backward; W8^2 = -i / +i; W8^3 = c (-1 - i) / c (-1 + i)
    const cpx<R> w1o = sign < 0 ? cpx<R>{c * (o1.x + o1.y), c * (o1.y - o1.x)} : cpx<R>{c * (o1.x - o1.y), c * (o1.y + o1.x)};
    const cpx<R> w2o = rot(o2, sign);
    const cpx<R> w3o = sign < 0 ? cpx<R>{c * (o3.y - o3.x), -c * (o3.x + o3.y)} : cpx<R>{-c * (o3.x + o3.y), c * (o3.x - o3.y)};
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1o);
    v[5] = csub(e1, w1o);
    v[2] = cadd(e2, w2o);
    v[6] = csub(e2, w2o);
    v[3] = cadd(e3, w3o);
    v[7] = csub(e3, w3o);
}

// One Stockham stage on one sequence (in place, padded slots).
template <typename R, int N, int RADIX, int NS>
__device__ inline void p2_stage(cpx<R> *seq, int t, const cpx<R> *twl, int sign, bool active)
{
    constexpr int M = N / RADIX;
    constexpr int BPT = 8 / RADIX;
    constexpr int TPS = N / 8;
    constexpr int TSTEP = N / (NS * RADIX);
    cpx<R> v[8];
    if (active) {
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int j = t + b * TPS;
            const int k = j & (NS - 1);
#pragma unroll
            for (int r = 0; r < RADIX; ++r) {
                cpx<R> a = seq[pidx(j + r * M)];
                if (NS > 1 && r > 0) {
                    cpx<R> w = twl[r * k * TSTEP];
                    if (sign > 0) w.y = -w.y;
                    a = {fma(a.x, w.x, -a.y * w.y), fma(a.x, w.y, a.y * w.x)};
                }
                v[b * RADIX + r] = a;
            }
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int j = t + b * TPS;
            const int k = j & (NS - 1);
            cpx<R> *x = v + b * RADIX;
            if constexpr (RADIX == 8) dft8(x, sign);
            else if constexpr (RADIX == 4) dft4(x[0], x[1], x[2], x[3], sign);
            else dft2(x);
            const int dst = (j / NS) * NS * RADIX + k;
#pragma unroll
            for (int q = 0; q < RADIX; ++q) seq[pidx(dst + q * NS)] = x[q];
        }
    }
    __syncthreads();
}

template <typename R, int N, int S, int NS>
__device__ inline void p2_stages(cpx<R> *seq, int t, const cpx<R> *twl, int sign, bool active)
{
    constexpr int LG = ilog2c(N);
    if constexpr (S < nstages(LG)) {
        constexpr int RADIX = stage_radix(LG, S);
        p2_stage<R, N, RADIX, NS>(seq, t, twl, sign, active);
        p2_stages<R, N, S + 1, NS * RADIX>(seq, t, twl, sign, active);
    }
}

// Transform RND x kSeq<N> sequences at buf + s * bs, kSeq<N> at a time; all
// 256 threads call it.
template <typename R, int N, int RND = 1>
__device__ inline void p2_fft(cpx<R> *buf, int bs, const cpx<R> *twl, int sign)
{
    constexpr int TPS = N / 8;
    const int g = threadIdx.x / TPS, t = threadIdx.x - g * TPS;
    const bool active = g < kSeq<N>;
#pragma unroll 1
    for (int r = 0; r < RND; ++r)
        p2_stages<R, N, 0, 1>(buf + (size_t)((active ? g : 0) + r * kSeq<N>) * bs, t, twl, sign, active);
}

// LDS layout: [twiddles (when staged)] [data] [bins]
template <typename R, int N>
__device__ inline const cpx<R> *stage_twiddles(unsigned char *smem, const cpx<R> *tw)
{
    if constexpr (kTwLds<R, N>) {
        cpx<R> *twl = reinterpret_cast<cpx<R> *>(smem);
        for (int e = threadIdx.x; e < N; e += blockDim.x) twl[e] = tw[e];
        return twl;
    } else {
        return tw;
    }
}

template <typename R, int N>
__device__ inline cpx<R> *data_area(unsigned char *smem)
{
    return reinterpret_cast<cpx<R> *>(smem) + (kTwLds<R, N> ? N : 0);
}

template <typename R, int N>
constexpr size_t twiddle_bytes() { return kTwLds<R, N> ? (size_t)N * sizeof(cpx<R>) : 0; }

// Persistent walk over `items`: 8 contiguous ranges, one per XCD.
struct Sched {
    long long begin, end, stride;
};
__device__ inline Sched xcd_sched(long long items)
{
    const long long G = gridDim.x, b = blockIdx.x;
    if (G % 8 != 0 || items < G) return {b, items, G};
    const long long per = G / 8;
    const long long chunk = (items + 7) / 8;
    const long long lo = (b % 8) * chunk;
    const long long hi = lo + chunk < items ? lo + chunk : items;
    return {lo + b / 8, hi, per};
}

template <typename T>
struct Vec16 {
    T v[16 / sizeof(T)];
};

template <typename T, typename R, int N>
__global__ void __launch_bounds__(256)
k_p2_rows_r2c(const T *__restrict__ field, cpx<R> *__restrict__ H, long long rows, long long ldh,
              const cpx<R> *__restrict__ tw)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int nc = N / 2 + 1, NSQ = kSeq<N>, BS = kRowStride<N>;
    constexpr int VW = 16 / sizeof(T);
    const cpx<R> *twl = stage_twiddles<R, N>(smem_raw, tw);
    cpx<R> *buf = data_area<R, N>(smem_raw);
    const long long items = (rows + 2 * NSQ - 1) / (2 * NSQ);
    const Sched sc = xcd_sched(items);
    for (long long it = sc.begin; it < sc.end; it += sc.stride) {
        const long long r0 = it * 2 * NSQ;
        const long long left = rows - r0;
        const int nr = (int)(left < 2 * NSQ ? left : 2 * NSQ);
        const Vec16<T> *src = reinterpret_cast<const Vec16<T> *>(field + r0 * N);
        __syncthreads();   // the previous item's readers are done with buf
        for (int q = threadIdx.x; q < 2 * NSQ * N / VW; q += 256) {
            const int idx = q * VW;
            const int row = idx / N, e = idx - row * N;
            Vec16<T> x;
            if (row < nr) {
                x = src[q];
            } else {
#pragma unroll
                for (int u = 0; u < VW; ++u) x.v[u] = (T)0;
            }
#pragma unroll
            for (int u = 0; u < VW; ++u) {
                R *slot = reinterpret_cast<R *>(&buf[(row >> 1) * BS + pidx(e + u)]);
                slot[row & 1] = (R)x.v[u];
            }
        }
        __syncthreads();
        p2_fft<R, N>(buf, BS, twl, -1);
        for (int idx = threadIdx.x; idx < NSQ * nc; idx += 256) {
            const int s = idx / nc, k = idx - s * nc;
            const cpx<R> z = buf[s * BS + pidx(k)], zn = buf[s * BS + pidx((N - k) & (N - 1))];
            const cpx<R> A = {(R)0.5 * (z.x + zn.x), (R)0.5 * (z.y - zn.y)};
            const cpx<R> B = {(R)0.5 * (z.y + zn.y), (R)0.5 * (zn.x - z.x)};
            const int ra = 2 * s, rb = 2 * s + 1;
            if (ra < nr) H[(r0 + ra) * ldh + k] = A;
            if (rb < nr) H[(r0 + rb) * ldh + k] = B;
        }
    }
}

template <typename T, typename R, int N>
__global__ void __launch_bounds__(256)
k_p2_rows_c2r(const cpx<R> *__restrict__ H, T *__restrict__ field, long long rows, const cpx<R> *__restrict__ tw,
              double norm, double shift)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int nc = N / 2 + 1, NSQ = kSeq<N>, BS = kRowStride<N>;
    const cpx<R> *twl = stage_twiddles<R, N>(smem_raw, tw);
    cpx<R> *buf = data_area<R, N>(smem_raw);
    const long long items = (rows + 2 * NSQ - 1) / (2 * NSQ);
    const Sched sc = xcd_sched(items);
    for (long long it = sc.begin; it < sc.end; it += sc.stride) {
        const long long r0 = it * 2 * NSQ;
        const long long left = rows - r0;
        const int nr = (int)(left < 2 * NSQ ? left : 2 * NSQ);
        auto ext = [&](int row, int k) -> cpx<R> {
            if (row >= nr) return {(R)0, (R)0};
            const cpx<R> *h = H + (r0 + row) * nc;
            if (k == 0) return {h[0].x, (R)0};
            if (2 * k < N) return h[k];
            if (2 * k == N) return {h[k].x, (R)0};
            const cpx<R> c = h[N - k];
            return {c.x, -c.y};
        };
        __syncthreads();
        for (int idx = threadIdx.x; idx < NSQ * N; idx += 256) {
            const int s = idx / N, k = idx - s * N;
            const cpx<R> A = ext(2 * s, k), B = ext(2 * s + 1, k);
            buf[s * BS + pidx(k)] = {A.x - B.y, A.y + B.x};
        }
        __syncthreads();
        p2_fft<R, N>(buf, BS, twl, +1);
        T *dst = field + r0 * N;
        for (int idx = threadIdx.x; idx < 2 * NSQ * N; idx += 256) {
            const int row = idx / N, e = idx - row * N;
            if (row >= nr) continue;
            const cpx<R> z = buf[(row >> 1) * BS + pidx(e)];
            const R v = (row & 1) ? z.y : z.x;
            dst[idx] = (T)((double)v * norm + shift);
        }
    }
}

// Columns per block: RND rounds of kSeq<N> sequences (about 32 KiB of data
// per block keeps 3-4 blocks resident per CU with 64-byte row segments;
// 128-byte segments at 64-70 KiB per block measured slower: 1.19 against
// 0.89 ms for the 512^3 FP32 spectrum).
template <typename R, int N>
constexpr int kColRounds = sizeof(R) == 4 ? 2 : 1;
constexpr size_t kLdsBlockMax = 160 * 1024;   // gfx950: LDS a block may use

// In-place transform along the middle axis of [outer][N][inner], columns
// c < ncols (the rest of a row is padding, left as it is).
template <typename R, int N>
__global__ void __launch_bounds__(256)
k_p2_cols(cpx<R> *H, long long outer, long long inner, long long ncols, const cpx<R> *__restrict__ tw, int sign)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int RND = kColRounds<R, N>;
    constexpr int CB = kSeq<N> * RND, BS = kColStride<R, N>;
    const cpx<R> *twl = stage_twiddles<R, N>(smem_raw, tw);
    cpx<R> *buf = data_area<R, N>(smem_raw);
    const long long tiles = (ncols + CB - 1) / CB;
    const Sched sc = xcd_sched(outer * tiles);
    for (long long it = sc.begin; it < sc.end; it += sc.stride) {
        const long long o = it / tiles;
        const long long c0 = (it - o * tiles) * CB;
        const int ncol = (int)((ncols - c0) < CB ? (ncols - c0) : CB);
        cpx<R> *base = H + o * (long long)N * inner + c0;
        __syncthreads();
        // the loads of a batch are all issued before any is waited for (a
        // plain load-store loop keeps one load per thread in flight)
        constexpr int LIT = (N * CB + 255) / 256, LB = LIT < 8 ? LIT : 8;
        for (int r0 = 0; r0 < LIT; r0 += LB) {
            cpx<R> v[LB];
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int idx = threadIdx.x + (r0 + u) * 256;
                const int e = idx / CB, cc = idx - e * CB;
                v[u] = (idx < N * CB && cc < ncol) ? base[(long long)e * inner + cc] : cpx<R>{(R)0, (R)0};
            }
#pragma unroll
            for (int u = 0; u < LB; ++u) {
                const int idx = threadIdx.x + (r0 + u) * 256;
                const int e = idx / CB, cc = idx - e * CB;
                if (idx < N * CB) buf[cc * BS + pidx(e)] = v[u];
            }
        }
        __syncthreads();
        p2_fft<R, N, RND>(buf, BS, twl, sign);
        for (int idx = threadIdx.x; idx < N * CB; idx += 256) {
            const int e = idx / CB, cc = idx - e * CB;
            if (cc < ncol) base[(long long)e * inner + cc] = buf[cc * BS + pidx(e)];
        }
    }
}

// ---------------------------------------------------------------------------
// The spectrum's forward transforms.  The first pass transposes as it goes,
// so that the last one reads whole rows:
//   D = 3: field [i0][i1][i2] --r2c along i2--> H [j][i1][i0]
//          --along i1 (in place, strided)--> --along i0, binned--> bins
//   D = 2: field [i0][i1] --r2c along i1--> H [j][i0] --along i0, binned-->
// (j < N/2 + 1 the half-complex index).  The transposing pass reads whole
// rows and writes element j of the row (i0, i1) to H[(j M1 + i1) N + i0]
// (M1 = N for D = 3, 1 for D = 2): its stores are the strided side, row
// segments of 64 or 128 bytes (consecutive i0 of one i1).

// Rows per item of a transposing pass: RND rounds of kSeq<N> sequences.
template <typename R, int N>
constexpr int kRowRounds = (sizeof(R) == 4 && N >= 16) ? 2 : 1;

// Load CB = RND kSeq<N> complex rows of length N and pitch ld (nr of them
// valid) into buf with 16-byte loads, all issued before any is waited for.
template <typename R, int N, int CB>
__device__ inline void load_rows(cpx<R> *buf, const cpx<R> *__restrict__ src, long long ld, int nr)
{
    constexpr int BS = kRowStride<N>;
    constexpr int VW = 16 / sizeof(cpx<R>);         // elements per load
    constexpr int LIT = CB * N / VW / 256;          // loads per thread
    static_assert(LIT * 256 * VW == CB * N, "whole loads per thread");
    Vec16<cpx<R>> v[LIT];
#pragma unroll
    for (int u = 0; u < LIT; ++u) {
        const int q = threadIdx.x + u * 256;
        const int row = (q * VW) / N, e = q * VW - row * N;
        if (row < nr) {
            v[u] = *reinterpret_cast<const Vec16<cpx<R>> *>(src + row * ld + e);
        } else {
#pragma unroll
            for (int w = 0; w < VW; ++w) v[u].v[w] = cpx<R>{(R)0, (R)0};
        }
    }
#pragma unroll
    for (int u = 0; u < LIT; ++u) {
        const int idx = (threadIdx.x + u * 256) * VW;
        const int row = idx / N, e = idx - row * N;
#pragma unroll
        for (int w = 0; w < VW; ++w) buf[row * BS + pidx(e + w)] = v[u].v[w];
    }
}

// real rows -> half-complex, transposed: row p = i1 N + i0 of the output
// order (i0 fastest) is the field's row i0 M1 + i1; element k < N/2 + 1 of
// it goes to out[(k M1 + i1) ld + i0]; rows = M1 N
template <typename T, typename R, int N>
__global__ void __launch_bounds__(256)
k_p2_rows_r2c_t(const T *__restrict__ field, cpx<R> *__restrict__ out, long long rows, long long M1, long long ld,
                const cpx<R> *__restrict__ tw)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int RND = kRowRounds<R, N>;
    constexpr int nc = N / 2 + 1, NSQ = kSeq<N> * RND, RR = 2 * NSQ, BS = kRowStride<N>;
    constexpr int VW = 16 / sizeof(T);
    const cpx<R> *twl = stage_twiddles<R, N>(smem_raw, tw);
    cpx<R> *buf = data_area<R, N>(smem_raw);
    const long long items = (rows + RR - 1) / RR;
    const Sched sc = xcd_sched(items);
    for (long long it = sc.begin; it < sc.end; it += sc.stride) {
        const long long r0 = it * RR;
        const long long left = rows - r0;
        const int nr = (int)(left < RR ? left : RR);
        __syncthreads();   // the previous item's readers are done with buf
        constexpr int LIT = (RR * N / VW + 255) / 256;
        Vec16<T> x[LIT];
#pragma unroll
        for (int u = 0; u < LIT; ++u) {
            const int q = threadIdx.x + u * 256;
            const int row = q * VW / N, e = q * VW - row * N;
            if (q < RR * N / VW && row < nr) {
                const long long pr = r0 + row, i1 = pr / N, i0 = pr - i1 * N;
                x[u] = *reinterpret_cast<const Vec16<T> *>(field + (i0 * M1 + i1) * N + e);
            } else {
#pragma unroll
                for (int w = 0; w < VW; ++w) x[u].v[w] = (T)0;
            }
        }
#pragma unroll
        for (int u = 0; u < LIT; ++u) {
            const int q = threadIdx.x + u * 256;
            if (q >= RR * N / VW) continue;
            const int idx = q * VW;
            const int row = idx / N, e = idx - row * N;
#pragma unroll
            for (int w = 0; w < VW; ++w) {
                R *slot = reinterpret_cast<R *>(&buf[(row >> 1) * BS + pidx(e + w)]);
                slot[row & 1] = (R)x[u].v[w];
            }
        }
        __syncthreads();
        p2_fft<R, N, RND>(buf, BS, twl, -1);
        // z = a + i b transformed: A_k = (Z_k + conj Z_{N-k}) / 2 (row 2s),
        // B_k = (Z_k - conj Z_{N-k}) / 2i (row 2s + 1); rows fastest in the
        // store order, so each k's rows form one segment
        for (int idx = threadIdx.x; idx < nc * RR; idx += 256) {
            const int k = idx / RR, row = idx - k * RR, sq = row >> 1;
            const cpx<R> z = buf[sq * BS + pidx(k)], zn = buf[sq * BS + pidx((N - k) & (N - 1))];
            const cpx<R> v = (row & 1) ? cpx<R>{(R)0.5 * (z.y + zn.y), (R)0.5 * (zn.x - z.x)}
                                       : cpx<R>{(R)0.5 * (z.x + zn.x), (R)0.5 * (z.y - zn.y)};
            const long long pr = r0 + row, i1 = pr / N;
            if (row < nr) out[(k * M1 + i1) * ld + (pr - i1 * N)] = v;
        }
    }
}

// Radial binning of the last pass.  Row r of its input is the line of
// modes (s_0 = -N/2+1 .. N/2, fixed s_1, j) along axis 0: D = 3, r = j N + i1
// (s_1 = i1, or i1 - N above N/2); D = 2, r = j (s_1 = 0).
struct BinArgs {
    int D, nBins;
    double dk, pref, Pnorm;
    double *binpart;   // [grid][nBins] power partials, one row per block
};

// complex rows of pitch ld, forward, binned (ComputePowerSpectrum,
// src/bubble.cpp:668-709); writes one row of nBins power partials per block
template <typename R, int N>
__global__ void __launch_bounds__(256)
k_p2_rows_bin(const cpx<R> *__restrict__ H, long long rows, long long ld, const cpx<R> *__restrict__ tw, BinArgs ba)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int RND = kRowRounds<R, N>;
    constexpr int CB = kSeq<N> * RND, BS = kRowStride<N>, nc = N / 2 + 1;
    const cpx<R> *twl = stage_twiddles<R, N>(smem_raw, tw);
    cpx<R> *buf = data_area<R, N>(smem_raw);
    // the block's histogram, the per-thread run records and the bin range
    // of each row
    double *hist = reinterpret_cast<double *>(buf + CB * BS);
    double *rec_fs = hist + ba.nBins;
    int *rec_fk = reinterpret_cast<int *>(rec_fs + 256);
    int *rec_lk = rec_fk + 256;
    int *rec_one = rec_lk + 256;
    int *line_lo = rec_one + 256;          // [CB] bin at a = 0
    int *line_hi = line_lo + CB;           // [CB] highest bin binned, -1: none
    // run totals [CB][SPAN] by (row, bin - lowest bin), in the data area
    // (free once an item's values are read)
    constexpr int SPAN = N / 2 + 1;        // bins a row can meet
    static_assert((size_t)CB * SPAN * sizeof(double) <= (size_t)CB * BS * sizeof(cpx<R>), "run table fits");
    double *runtab = reinterpret_cast<double *>(buf);
    for (int b = threadIdx.x; b < ba.nBins; b += 256) hist[b] = 0.0;
    const long long items = (rows + CB - 1) / CB;
    const Sched sc = xcd_sched(items);
    for (long long it = sc.begin; it < sc.end; it += sc.stride) {
        const long long r0 = it * CB;
        const long long left = rows - r0;
        const int ncol = (int)(left < CB ? left : CB);
        __syncthreads();
        load_rows<R, N, CB>(buf, H + r0 * ld, ld, ncol);
        __syncthreads();
        p2_fft<R, N, RND>(buf, BS, twl, -1);
        {
            // Binning in a fixed order, no atomics.  The modes of a row at
            // s_0 = +a and -a (elements a and N - a) have the same |k|, so
            // the same bin and weight: thread t takes A consecutive values of
            // a (0 <= a < N/2) in one row, both elements of each, and the
            // last thread of a row also a = N/2.  Along a row the bin never
            // decreases with a and grows by at most one per step, so the row
            // meets every bin from its lowest to its highest once, as one
            // contiguous run of a.  Each thread sums the runs in its
            // a-range; a run that goes on into the next threads' ranges is
            // completed by the thread where it starts, from the next threads'
            // first partial sums (in order of a).  The run totals go to a
            // table indexed by (row, bin - lowest bin of the row), in the
            // data area (its values are read by then); the thread that owns a
            // bin in the block's histogram then adds that bin's totals row by
            // row, in a fixed order.  Bins are integer square
            // roots, found once per thread and then stepped with integer
            // compares.  The weight kc^D is a product of D factors in double
            // for double data (within 2 ulp of the reference's pow(kc, D));
            // for float data |k|^D is formed in float (a few float roundings,
            // far inside the error of the float transform).
            constexpr int TPC = 256 / CB;          // threads per column
            constexpr int A = N / (2 * TPC);       // values of a per thread
            static_assert(A >= 1 && TPC * A * 2 == N, "a-ranges tile the half column");
            const int t = threadIdx.x;
            const int cc = t / TPC, kt = t - cc * TPC, a0 = kt * A;
            const bool lastk = kt == TPC - 1;
            const long long q = r0 + cc;
            const int j = (int)(ba.D == 3 ? q / N : q);
            const int i1 = (int)(ba.D == 3 ? q % N : 0);
            const int s1 = (i1 <= N / 2) ? i1 : i1 - N;
            const bool colok = cc < ncol;
            const int C = s1 * s1 + j * j;
            const double w = (j == 0 || j == N / 2) ? 1.0 : 2.0;
            const double wc = sizeof(R) == 4 ? w * ba.pref * ba.Pnorm * (ba.D == 3 ? ba.dk * ba.dk * ba.dk : ba.dk * ba.dk)
                                             : w * ba.pref * ba.Pnorm;
            int key[A + 1];
            double val[A + 1];
            int b;
            {
                const int m = C + a0 * a0;
                b = (int)__builtin_amdgcn_sqrtf((float)m);
                b -= b * b > m ? 1 : 0;
                b += (b + 1) * (b + 1) <= m ? 1 : 0;
            }
            const int bfirst = b;
#pragma unroll
            for (int u = 0; u <= A; ++u) {
                const int a = a0 + u;
                const int m = C + a * a;
                if (u > 0) b += (b + 1) * (b + 1) <= m ? 1 : 0;
                const bool in = u < A || lastk;
                const cpx<R> f0 = buf[cc * BS + pidx(a)];
                const cpx<R> f1 = buf[cc * BS + pidx((N - a) & (N - 1))];
                const bool pair = a > 0 && a < N / 2;
                double v;
                if constexpr (sizeof(R) == 4) {
                    // |f|^2 |k|^D in float, the weight in double
                    const float p = fmaf(f0.x, f0.x, f0.y * f0.y) + (pair ? fmaf(f1.x, f1.x, f1.y * f1.y) : 0.0f);
                    const float mf = (float)m;
                    v = wc * (double)((ba.D == 3 ? mf * __builtin_amdgcn_sqrtf(mf) : mf) * p);
                } else {
                    double p = f0.x * f0.x + f0.y * f0.y;
                    if (pair) p += f1.x * f1.x + f1.y * f1.y;
                    const double kc = sqrt((double)m) * ba.dk;
                    v = wc * (ba.D == 3 ? kc * kc * kc : kc * kc) * p;
                }
                val[u] = in ? v : 0.0;
                key[u] = (in && colok && b < ba.nBins) ? b : -1;
            }
            // the row's bin range: lowest at a = 0, highest at a = N/2
            // (none binned when lowest > highest)
            if (kt == 0) line_lo[cc] = bfirst;
            if (lastk) line_hi[cc] = colok ? min(b, ba.nBins - 1) : -1;
            // runs of equal keys: val[u] becomes the sum of the run from u to
            // its end in this thread's range (static indices only)
#pragma unroll
            for (int u = A - 1; u >= 0; --u)
                if (key[u] >= 0 && key[u + 1] == key[u]) val[u] += val[u + 1];
            const bool starts = key[0] >= 0, ends = key[A - 1] >= 0;
            bool one = starts;
#pragma unroll
            for (int u = 1; u < A; ++u) one = one && key[u] == key[0];
            rec_fk[t] = starts ? key[0] : -1;
            rec_fs[t] = starts ? val[0] : 0.0;
            rec_lk[t] = ends ? key[A - 1] : -1;
            rec_one[t] = one ? 1 : 0;
            __syncthreads();   // records and bin ranges written; the data area is free
            const int tcol0 = cc * TPC;
            // the first run is owned by a previous thread when it goes on from there
            const bool cont = starts && kt > 0 && rec_lk[t - 1] == key[0];
            // the last run's part in the following threads' ranges: the
            // next four threads' records read at once, further ones (runs
            // longer than 4 A + 1 values of a) one by one
            double ext = 0.0;
            if (ends && !lastk && !(one && cont)) {
                constexpr int PF = 4;
                int fk[PF], on[PF];
                double fs[PF];
#pragma unroll
                for (int d = 0; d < PF; ++d) {
                    const int u2 = min(t + 1 + d, tcol0 + TPC - 1);
                    fk[d] = rec_fk[u2];
                    fs[d] = rec_fs[u2];
                    on[d] = rec_one[u2];
                }
                const int L = key[A - 1];
                bool alive = true;
#pragma unroll
                for (int d = 0; d < PF; ++d) {
                    alive = alive && kt + 1 + d < TPC && fk[d] == L;
                    ext += alive ? fs[d] : 0.0;
                    alive = alive && on[d];
                }
                for (int u2 = t + 1 + PF; alive && u2 < tcol0 + TPC && rec_fk[u2] == L; ++u2) {
                    ext += rec_fs[u2];
                    alive = rec_one[u2] != 0;
                }
            }
            const int lo = line_lo[cc];
#pragma unroll
            for (int u = 0; u <= A; ++u) {
                const bool head = key[u] >= 0 && (u == 0 || key[u - 1] != key[u]);
                if (!head || (u == 0 && cont)) continue;
                const bool last = u < A && ends && key[u] == key[A - 1];
                runtab[cc * SPAN + key[u] - lo] = last ? val[u] + ext : val[u];
            }
            __syncthreads();   // run table complete
            // each bin's totals added row by row (rows that miss the bin add
            // nothing; the reads are issued together, no branches)
            const int bpt = bins_per_thread(ba.nBins, 256);
            constexpr int GC = CB < 8 ? CB : 8;    // rows read per batch
            for (int bb = t * bpt; bb < t * bpt + bpt && bb < ba.nBins; ++bb) {
                double acc = hist[bb];
                for (int g0 = 0; g0 < CB; g0 += GC) {
                    int rlo[GC];
                    double tv[GC];
#pragma unroll
                    for (int g = 0; g < GC; ++g) {
                        rlo[g] = line_lo[g0 + g];
                        tv[g] = bb <= line_hi[g0 + g] ? 1.0 : 0.0;
                    }
#pragma unroll
                    for (int g = 0; g < GC; ++g) {
                        const bool hit = bb >= rlo[g] && tv[g] != 0.0;
                        const double x = runtab[(g0 + g) * SPAN + (hit ? bb - rlo[g] : 0)];
                        tv[g] = hit ? x : 0.0;
                    }
#pragma unroll
                    for (int g = 0; g < GC; ++g) acc += tv[g];
                }
                hist[bb] = acc;
            }
        }
    }
    // one row of partials per block, summed in block order afterwards
    __syncthreads();   // a block may have had no item
    double *dst = ba.binpart + (long long)blockIdx.x * ba.nBins;
    for (int b = threadIdx.x; b < ba.nBins; b += 256) dst[b] = hist[b];
}

int current_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    return cus;
}

// Blocks of a persistent launch: as many as are resident at once (at most
// `cap`), a multiple of 8 (one share per XCD), at most `items`.
template <typename K>
unsigned persistent_grid(K kernel, size_t shm, long long items, long long cap)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, shm) != hipSuccess || per_cu < 1) per_cu = 1;
    long long g = (long long)per_cu * current_cus();
    if (g > cap) g = cap;
    if (g > items) g = items;
    if (g >= 8) g -= g % 8;
    return (unsigned)(g < 1 ? 1 : g);
}

template <typename R, int N>
constexpr size_t rows_shm() { return twiddle_bytes<R, N>() + (size_t)kSeq<N> * kRowStride<N> * sizeof(cpx<R>); }
template <typename R, int N>
constexpr int kCols = kSeq<N> * kColRounds<R, N>;
template <typename R, int N>
constexpr size_t cols_shm() { return twiddle_bytes<R, N>() + (size_t)kCols<R, N> * kColStride<R, N> * sizeof(cpx<R>); }
template <typename R, int N>
constexpr int kRowsT = kSeq<N> * kRowRounds<R, N>;   // sequences per item of a row pass
template <typename R, int N>
constexpr size_t rows_t_shm() { return twiddle_bytes<R, N>() + (size_t)kRowsT<R, N> * kRowStride<N> * sizeof(cpx<R>); }

constexpr long long kNoCap = 1LL << 40;

// Kernels launched with more than 64 KiB of dynamic LDS ask for it first.
template <typename K>
hipError_t allow_lds(K kernel, size_t shm)
{
    if (shm <= 65536) return hipSuccess;
    if (shm > kLdsBlockMax) return hipErrorInvalidValue;
    return hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
}

template <typename R, int N>
hipError_t rows_r2c_n(const void *field, void *H, long long rows, long long ldh, const void *tw, hipStream_t st)
{
    const size_t shm = rows_shm<R, N>();
    const long long items = (rows + 2 * kSeq<N> - 1) / (2 * kSeq<N>);
    const unsigned grid = persistent_grid(k_p2_rows_r2c<R, R, N>, shm, items, kNoCap);
    k_p2_rows_r2c<R, R, N><<<grid, 256, shm, st>>>((const R *)field, (cpx<R> *)H, rows, ldh, (const cpx<R> *)tw);
    return hipGetLastError();
}

template <typename T, int N>
hipError_t rows_c2r_n(const void *H, void *field, long long rows, const void *tw, double norm, double shift,
                      hipStream_t st)
{
    const size_t shm = rows_shm<double, N>();
    const long long items = (rows + 2 * kSeq<N> - 1) / (2 * kSeq<N>);
    const unsigned grid = persistent_grid(k_p2_rows_c2r<T, double, N>, shm, items, kNoCap);
    k_p2_rows_c2r<T, double, N><<<grid, 256, shm, st>>>((const cpx<double> *)H, (T *)field, rows,
                                                        (const cpx<double> *)tw, norm, shift);
    return hipGetLastError();
}

template <typename R, int N>
hipError_t cols_n(void *H, long long outer, long long inner, long long ncols, const void *tw, int sign, hipStream_t st)
{
    constexpr int CB = kCols<R, N>;
    const size_t shm = cols_shm<R, N>();
    hipError_t e = allow_lds(k_p2_cols<R, N>, shm);
    if (e != hipSuccess) return e;
    const long long items = outer * ((ncols + CB - 1) / CB);
    const unsigned grid = persistent_grid(k_p2_cols<R, N>, shm, items, kNoCap);
    k_p2_cols<R, N><<<grid, 256, shm, st>>>((cpx<R> *)H, outer, inner, ncols, (const cpx<R> *)tw, sign);
    return hipGetLastError();
}

template <typename R, int N>
hipError_t rows_r2c_t_n(const void *field, void *out, long long M1, long long ld, const void *tw, hipStream_t st)
{
    const size_t shm = rows_t_shm<R, N>();
    hipError_t e = allow_lds(k_p2_rows_r2c_t<R, R, N>, shm);
    if (e != hipSuccess) return e;
    const long long rows = M1 * N;
    const long long items = (rows + 2 * kRowsT<R, N> - 1) / (2 * kRowsT<R, N>);
    const unsigned grid = persistent_grid(k_p2_rows_r2c_t<R, R, N>, shm, items, kNoCap);
    k_p2_rows_r2c_t<R, R, N><<<grid, 256, shm, st>>>((const R *)field, (cpx<R> *)out, rows, M1, ld, (const cpx<R> *)tw);
    return hipGetLastError();
}

template <typename R, int N>
hipError_t rows_bin_n(const void *H, long long rows, long long ld, const void *tw, BinArgs ba, int cap_blocks,
                      int *nblk, hipStream_t st)
{
    constexpr size_t CB = kRowsT<R, N>;
    const size_t shm = rows_t_shm<R, N>() + sizeof(double) * ((size_t)ba.nBins + 256) + sizeof(int) * (3 * 256 + 2 * CB);
    hipError_t e = allow_lds(k_p2_rows_bin<R, N>, shm);
    if (e != hipSuccess) return e;
    const long long items = (rows + (long long)CB - 1) / (long long)CB;
    const unsigned grid = persistent_grid(k_p2_rows_bin<R, N>, shm, items, cap_blocks);
    *nblk = (int)grid;
    k_p2_rows_bin<R, N><<<grid, 256, shm, st>>>((const cpx<R> *)H, rows, ld, (const cpx<R> *)tw, ba);
    return hipGetLastError();
}

#define CB_P2_SWITCH(N_, CALL)                    \
    switch (N_) {                                 \
    case 8: { constexpr int NN = 8; CALL; }       \
    case 16: { constexpr int NN = 16; CALL; }     \
    case 32: { constexpr int NN = 32; CALL; }     \
    case 64: { constexpr int NN = 64; CALL; }     \
    case 128: { constexpr int NN = 128; CALL; }   \
    case 256: { constexpr int NN = 256; CALL; }   \
    case 512: { constexpr int NN = 512; CALL; }   \
    case 1024: { constexpr int NN = 1024; CALL; } \
    case 2048: { constexpr int NN = 2048; CALL; } \
    default: return hipErrorInvalidValue;         \
    }

} // namespace

bool p2_supported(int N)
{
    return N >= 8 && N <= 2048 && (N & (N - 1)) == 0;
}

int p2_bin_blocks_max() { return 8 * current_cus(); }

hipError_t p2_rows_r2c(int fp32, const void *field, void *H, long long rows, long long ldh, int N, const void *tw,
                       hipStream_t st)
{
    if (fp32) { CB_P2_SWITCH(N, return (rows_r2c_n<float, NN>(field, H, rows, ldh, tw, st))) }
    CB_P2_SWITCH(N, return (rows_r2c_n<double, NN>(field, H, rows, ldh, tw, st)))
}

int p2_pitch(int N, int fp32) { return N >= 64 ? N + (fp32 ? 16 : 8) : N; }

hipError_t p2_cols(int fp32, void *H, long long outer, long long inner, long long ncols, int N, const void *tw,
                   int sign, hipStream_t st)
{
    if (fp32) { CB_P2_SWITCH(N, return (cols_n<float, NN>(H, outer, inner, ncols, tw, sign, st))) }
    CB_P2_SWITCH(N, return (cols_n<double, NN>(H, outer, inner, ncols, tw, sign, st)))
}

hipError_t p2_rows_r2c_t(int fp32, const void *field, void *out, long long M1, long long ld, int N, const void *tw,
                         hipStream_t st)
{
    if (fp32) { CB_P2_SWITCH(N, return (rows_r2c_t_n<float, NN>(field, out, M1, ld, tw, st))) }
    CB_P2_SWITCH(N, return (rows_r2c_t_n<double, NN>(field, out, M1, ld, tw, st)))
}

hipError_t p2_rows_bin(int fp32, const void *H, long long rows, long long ld, int N, const void *tw, int D, int nBins,
                       double dk, double pref, double Pnorm, double *binpart, int cap_blocks, int *nblk,
                       hipStream_t st)
{
    const BinArgs ba{D, nBins, dk, pref, Pnorm, binpart};
    if (fp32) { CB_P2_SWITCH(N, return (rows_bin_n<float, NN>(H, rows, ld, tw, ba, cap_blocks, nblk, st))) }
    CB_P2_SWITCH(N, return (rows_bin_n<double, NN>(H, rows, ld, tw, ba, cap_blocks, nblk, st)))
}

hipError_t p2_rows_c2r(int field_fp32, const void *H, void *field, long long rows, int N, const void *tw,
                       double norm, double shift, hipStream_t st)
{
    if (field_fp32) { CB_P2_SWITCH(N, return (rows_c2r_n<float, NN>(H, field, rows, tw, norm, shift, st))) }
    CB_P2_SWITCH(N, return (rows_c2r_n<double, NN>(H, field, rows, tw, norm, shift, st)))
}

} // namespace cb
