// cb_fft.hpp -- types shared by the FFT kernels (internal).
#pragma once

#include "cb_common.hpp"

namespace cb {

template <typename R>
struct cpx {
    R x, y;
};

static constexpr int kMaxStages = 40;
static constexpr size_t kLdsBudget = 65536;
// LDS a block may use on gfx950 (the binning histograms go up to it: 3 doubles
// per bin, so up to 6826 bins, e.g. 2D N up to ~9600)
static constexpr size_t kLdsBlockMax = 160 * 1024;

struct Stages {
    int n;
    int ns;
    int radix[kMaxStages];
};

template <typename R>
__device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Integer square roots of m < 2^52: floor(sqrt(m)) and ceil(sqrt(m)), 0 for
// m <= 0.  A mode's bin is (int)sqrt((double)n2) with n2 an integer below
// 2^52, which is exactly isqrt_floor(n2): the double square root is correctly
// rounded and cannot reach k + 1 from below (k + 1)^2.  So the modes of bin b
// on a line of squared offset C are those with b^2 <= C + a^2 < (b + 1)^2.
__device__ inline long long isqrt_floor(long long m)
{
    if (m <= 0) return 0;
    long long r = (long long)sqrt((double)m);
    while (r * r > m) --r;
    while ((r + 1) * (r + 1) <= m) ++r;
    return r;
}
__device__ inline long long isqrt_ceil(long long m)
{
    if (m <= 0) return 0;
    const long long r = isqrt_floor(m);
    return r * r == m ? r : r + 1;
}

// Bins per thread when a block's threads own contiguous bin ranges (each
// bin's LDS slot is written by its owner only: no atomics, a fixed order).
__host__ __device__ constexpr int bins_per_thread(int nBins, int threads)
{
    return (nBins + threads - 1) / threads;
}


// Power-of-two lengths 8..2048: compile-time radix-8/4/2 transforms
// (cb_fft_p2.hip).  tw = exp(-2 pi i e / N), e < N, in the transform's
// precision.  Each returns hipErrorInvalidValue for an unsupported N.
bool p2_supported(int N);
// real rows (T = R) -> half-complex rows of stride ldh (>= N/2 + 1), forward
hipError_t p2_rows_r2c(int fp32, const void *field, void *H, long long rows, long long ldh, int N, const void *tw,
                       hipStream_t st);
// in-place transform along the middle axis of [outer][N][inner], columns
// c < ncols
hipError_t p2_cols(int fp32, void *H, long long outer, long long inner, long long ncols, int N, const void *tw,
                   int sign, hipStream_t st);
// row pitch of the spectrum's transposed array: N padded by 128 bytes off a
// power of two (rows a power-of-two apart in memory load slowly; rows stay
// 128-byte aligned)
int p2_pitch(int N, int fp32);
// The spectrum's first forward pass (D = 2 / 3): real rows of length N ->
// half-complex, transposed: the field's row i0 M1 + i1 (M1 = N for D = 3,
// 1 for D = 2) goes to out[(j M1 + i1) ld + i0], j < N/2 + 1.
hipError_t p2_rows_r2c_t(int fp32, const void *field, void *out, long long M1, long long ld, int N, const void *tw,
                         hipStream_t st);
// last forward pass of a D = 2 / 3 spectrum: transform the complex rows of H
// (pitch ld; row r = j N + i1 for D = 3, r = j for D = 2, along axis 0) and
// bin the modes instead of storing them; writes *nblk rows of nBins power
// partials to binpart, one per block of the launch (at most cap_blocks rows).
hipError_t p2_rows_bin(int fp32, const void *H, long long rows, long long ld, int N, const void *tw, int D, int nBins,
                       double dk, double pref, double Pnorm, double *binpart, int cap_blocks, int *nblk,
                       hipStream_t st);
// largest block count p2_rows_bin uses
int p2_bin_blocks_max();
// half-complex rows (double) -> real rows (float or double), backward,
// field = x * norm + shift
hipError_t p2_rows_c2r(int field_fp32, const void *H, void *field, long long rows, int N, const void *tw,
                       double norm, double shift, hipStream_t st);

} // namespace cb
