// cb_probe.hip -- device-to-device copy bandwidth probe (measurement helper):
// a plain 16-byte-vector grid-stride copy, the yardstick for the stepper's
// HBM roofline fraction and for scaling the HBM byte counters.
#include <string>

#include "../../include/coolbubble.h"
#include "cb_sim.hpp"

namespace {

__global__ void __launch_bounds__(256) k_copy16(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

} // namespace

extern "C" cb_status cb_bandwidth_probe(int device, long long bytes, int reps, double *gbps)
{
    if (bytes < 16 || reps < 1 || !gbps) { cb::set_error("invalid argument"); return CB_INVALID_ARG; }
    if (hipSetDevice(device) != hipSuccess) { cb::set_error("hipSetDevice failed"); return CB_NO_DEVICE; }
    const long long n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, (size_t)n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, (size_t)n * 16);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, (size_t)n * 16, st);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const unsigned grid = 256 * 8 * 4;
    if (e == hipSuccess) {
        k_copy16<<<grid, 256, 0, st>>>((const float4 *)a, (float4 *)b, n);
        e = hipEventRecord(e0, st);
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
            k_copy16<<<grid, 256, 0, st>>>((const float4 *)a, (float4 *)b, n);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
    }
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (e != hipSuccess) { cb::set_error(std::string("bandwidth probe: ") + hipGetErrorString(e)); return CB_HIP_ERROR; }
    *gbps = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;   // read + write bytes
    return CB_OK;
}
