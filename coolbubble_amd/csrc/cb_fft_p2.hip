// cb_fft_p2.hip -- power-of-two FFTs (N = 8 .. 2048) for the spectrum and
// the thermal initial conditions, gfx950.
//
// Stockham autosort with the radices fixed at compile time: radix-8 stages,
// then one radix-4 or radix-2 stage.  A sequence is processed by N/8 threads;
// in each stage a thread loads the inputs of its 8/R butterflies from LDS
// into registers, applies the twiddles (from an LDS copy of the table), then
// after a barrier writes the butterfly outputs back in place.  A 256-thread
// block runs 2048/N sequences at a time.
//   rows r2c: two real rows per complex sequence (z = a + i b), split after
//             the transform: A_k = (Z_k + conj Z_{N-k}) / 2,
//             B_k = (Z_k - conj Z_{N-k}) / 2i;
//   rows c2r: two Hermitian rows per sequence (Z = A + i B), real and
//             imaginary parts of the backward transform;
//   columns : CB adjacent columns of a strided axis, transposed into LDS
//             with one element of padding per sequence.
#include "cb_fft.hpp"

namespace cb {

namespace {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
constexpr int nstages(int lg) { return lg / 3 + (lg % 3 ? 1 : 0); }
constexpr int stage_radix(int lg, int s) { return s < lg / 3 ? 8 : (lg % 3 == 2 ? 4 : 2); }

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

// One Stockham stage on one sequence (contiguous, in place).
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
                cpx<R> a = seq[j + r * M];
                if (NS > 1 && r > 0) {
                    cpx<R> w = twl[r * k * TSTEP];
                    if (sign > 0) w.y = -w.y;
                    a = cmul(a, w);
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
            for (int q = 0; q < RADIX; ++q) seq[dst + q * NS] = x[q];
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

// Transform nseq sequences at buf + s * bs; all 256 threads of the block call it.
template <typename R, int N>
__device__ inline void p2_fft(cpx<R> *buf, int nseq, int bs, const cpx<R> *twl, int sign)
{
    constexpr int TPS = N / 8;
    constexpr int SEQ = 256 / TPS;
    const int g = threadIdx.x / TPS, t = threadIdx.x - g * TPS;
    for (int s0 = 0; s0 < nseq; s0 += SEQ) {
        const int s = s0 + g;
        const bool active = s < nseq;
        p2_stages<R, N, 0, 1>(buf + (size_t)(active ? s : 0) * bs, t, twl, sign, active);
    }
}

template <typename R, int N>
__device__ inline cpx<R> *load_twiddles(unsigned char *smem, const cpx<R> *tw)
{
    cpx<R> *twl = reinterpret_cast<cpx<R> *>(smem);
    for (int e = threadIdx.x; e < N; e += blockDim.x) twl[e] = tw[e];
    return twl;
}

template <typename T, typename R, int N>
__global__ void __launch_bounds__(256)
k_p2_rows_r2c(const T *__restrict__ field, cpx<R> *__restrict__ H, long long rows, int nseq,
              const cpx<R> *__restrict__ tw)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int nc = N / 2 + 1;
    cpx<R> *twl = load_twiddles<R, N>(smem_raw, tw);
    cpx<R> *buf = twl + N;
    const long long r0 = (long long)blockIdx.x * 2 * nseq;
    const long long left = rows - r0;
    const int nr = (int)(left < 2 * nseq ? left : 2 * nseq);
    const T *src = field + r0 * N;
    for (int idx = threadIdx.x; idx < 2 * nseq * N; idx += blockDim.x) {
        const int row = idx / N, e = idx - row * N;
        const R v = (row < nr) ? (R)src[idx] : (R)0;
        R *slot = reinterpret_cast<R *>(&buf[(row >> 1) * N + e]);
        slot[row & 1] = v;
    }
    __syncthreads();
    p2_fft<R, N>(buf, nseq, N, twl, -1);
    for (int idx = threadIdx.x; idx < nseq * nc; idx += blockDim.x) {
        const int s = idx / nc, k = idx - s * nc;
        const cpx<R> z = buf[s * N + k], zn = buf[s * N + ((N - k) & (N - 1))];
        const cpx<R> A = {(R)0.5 * (z.x + zn.x), (R)0.5 * (z.y - zn.y)};
        const cpx<R> B = {(R)0.5 * (z.y + zn.y), (R)0.5 * (zn.x - z.x)};
        const int ra = 2 * s, rb = 2 * s + 1;
        if (ra < nr) H[(r0 + ra) * nc + k] = A;
        if (rb < nr) H[(r0 + rb) * nc + k] = B;
    }
}

template <typename T, typename R, int N>
__global__ void __launch_bounds__(256)
k_p2_rows_c2r(const cpx<R> *__restrict__ H, T *__restrict__ field, long long rows, int nseq,
              const cpx<R> *__restrict__ tw, double norm, double shift)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int nc = N / 2 + 1;
    cpx<R> *twl = load_twiddles<R, N>(smem_raw, tw);
    cpx<R> *buf = twl + N;
    const long long r0 = (long long)blockIdx.x * 2 * nseq;
    const long long left = rows - r0;
    const int nr = (int)(left < 2 * nseq ? left : 2 * nseq);
    auto ext = [&](int row, int k) -> cpx<R> {
        if (row >= nr) return {(R)0, (R)0};
        const cpx<R> *h = H + (r0 + row) * nc;
        if (k == 0) return {h[0].x, (R)0};
        if (2 * k < N) return h[k];
        if (2 * k == N) return {h[k].x, (R)0};
        const cpx<R> c = h[N - k];
        return {c.x, -c.y};
    };
    for (int idx = threadIdx.x; idx < nseq * N; idx += blockDim.x) {
        const int s = idx / N, k = idx - s * N;
        const cpx<R> A = ext(2 * s, k), B = ext(2 * s + 1, k);
        buf[idx] = {A.x - B.y, A.y + B.x};
    }
    __syncthreads();
    p2_fft<R, N>(buf, nseq, N, twl, +1);
    T *dst = field + r0 * N;
    for (int idx = threadIdx.x; idx < 2 * nseq * N; idx += blockDim.x) {
        const int row = idx / N, e = idx - row * N;
        if (row >= nr) continue;
        const cpx<R> z = buf[(row >> 1) * N + e];
        const R v = (row & 1) ? z.y : z.x;
        dst[idx] = (T)((double)v * norm + shift);
    }
}

template <typename R, int N>
__global__ void __launch_bounds__(256)
k_p2_cols(cpx<R> *H, long long inner, int CB, const cpx<R> *__restrict__ tw, int sign)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int bs = N + 1;
    cpx<R> *twl = load_twiddles<R, N>(smem_raw, tw);
    cpx<R> *buf = twl + N;
    const long long tiles = (inner + CB - 1) / CB;
    const long long o = blockIdx.x / tiles;
    const long long c0 = (blockIdx.x - o * tiles) * CB;
    const int ncol = (int)((inner - c0) < CB ? (inner - c0) : CB);
    cpx<R> *base = H + o * (long long)N * inner + c0;
    for (int idx = threadIdx.x; idx < N * CB; idx += blockDim.x) {
        const int e = idx / CB, cc = idx - e * CB;
        buf[cc * bs + e] = (cc < ncol) ? base[(long long)e * inner + cc] : cpx<R>{(R)0, (R)0};
    }
    __syncthreads();
    p2_fft<R, N>(buf, CB, bs, twl, sign);
    for (int idx = threadIdx.x; idx < N * CB; idx += blockDim.x) {
        const int e = idx / CB, cc = idx - e * CB;
        if (cc < ncol) base[(long long)e * inner + cc] = buf[cc * bs + e];
    }
}

// sequences per block for an LDS budget (table of N entries + data)
int seqs_for(int N, size_t csz, int pad)
{
    const size_t avail = kLdsBudget - (size_t)N * csz;
    int n = (int)(avail / ((size_t)(N + pad) * csz));
    if (n > 16) n = 16;
    return n;
}

template <typename R, int N>
void rows_r2c_n(const void *field, void *H, long long rows, const void *tw, hipStream_t st)
{
    const int nseq = seqs_for(N, sizeof(cpx<R>), 0);
    const size_t shm = (size_t)(N + nseq * N) * sizeof(cpx<R>);
    const long long blocks = (rows + 2 * nseq - 1) / (2 * nseq);
    k_p2_rows_r2c<R, R, N><<<(unsigned)blocks, 256, shm, st>>>((const R *)field, (cpx<R> *)H, rows, nseq,
                                                              (const cpx<R> *)tw);
}

template <typename T, int N>
void rows_c2r_n(const void *H, void *field, long long rows, const void *tw, double norm, double shift,
                hipStream_t st)
{
    const int nseq = seqs_for(N, sizeof(cpx<double>), 0);
    const size_t shm = (size_t)(N + nseq * N) * sizeof(cpx<double>);
    const long long blocks = (rows + 2 * nseq - 1) / (2 * nseq);
    k_p2_rows_c2r<T, double, N><<<(unsigned)blocks, 256, shm, st>>>((const cpx<double> *)H, (T *)field, rows, nseq,
                                                                    (const cpx<double> *)tw, norm, shift);
}

template <typename R, int N>
void cols_n(void *H, long long outer, long long inner, const void *tw, int sign, hipStream_t st)
{
    const int CB = seqs_for(N, sizeof(cpx<R>), 1);
    const size_t shm = (size_t)(N + CB * (N + 1)) * sizeof(cpx<R>);
    const long long tiles = (inner + CB - 1) / CB;
    k_p2_cols<R, N><<<(unsigned)(outer * tiles), 256, shm, st>>>((cpx<R> *)H, inner, CB, (const cpx<R> *)tw, sign);
}

#define CB_P2_SWITCH(N_, CALL)                  \
    switch (N_) {                               \
    case 8: { constexpr int NN = 8; CALL; break; }       \
    case 16: { constexpr int NN = 16; CALL; break; }     \
    case 32: { constexpr int NN = 32; CALL; break; }     \
    case 64: { constexpr int NN = 64; CALL; break; }     \
    case 128: { constexpr int NN = 128; CALL; break; }   \
    case 256: { constexpr int NN = 256; CALL; break; }   \
    case 512: { constexpr int NN = 512; CALL; break; }   \
    case 1024: { constexpr int NN = 1024; CALL; break; } \
    case 2048: { constexpr int NN = 2048; CALL; break; } \
    default: return hipErrorInvalidValue;       \
    }

} // namespace

bool p2_supported(int N, size_t csz)
{
    return N >= 8 && N <= 2048 && (N & (N - 1)) == 0 && seqs_for(N, csz, 1) >= 1;
}

hipError_t p2_rows_r2c(int fp32, const void *field, void *H, long long rows, int N, const void *tw, hipStream_t st)
{
    if (fp32) { CB_P2_SWITCH(N, (rows_r2c_n<float, NN>(field, H, rows, tw, st))) }
    else { CB_P2_SWITCH(N, (rows_r2c_n<double, NN>(field, H, rows, tw, st))) }
    return hipGetLastError();
}

hipError_t p2_cols(int fp32, void *H, long long outer, long long inner, int N, const void *tw, int sign,
                   hipStream_t st)
{
    if (fp32) { CB_P2_SWITCH(N, (cols_n<float, NN>(H, outer, inner, tw, sign, st))) }
    else { CB_P2_SWITCH(N, (cols_n<double, NN>(H, outer, inner, tw, sign, st))) }
    return hipGetLastError();
}

hipError_t p2_rows_c2r(int field_fp32, const void *H, void *field, long long rows, int N, const void *tw,
                       double norm, double shift, hipStream_t st)
{
    if (field_fp32) { CB_P2_SWITCH(N, (rows_c2r_n<float, NN>(H, field, rows, tw, norm, shift, st))) }
    else { CB_P2_SWITCH(N, (rows_c2r_n<double, NN>(H, field, rows, tw, norm, shift, st))) }
    return hipGetLastError();
}

} // namespace cb
