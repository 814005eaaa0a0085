// cb_probe.hip -- device-to-device copy bandwidth probe (measurement helper):
// a plain 16-byte-vector copy, the yardstick for the stepper's HBM roofline
// fraction and for scaling the HBM byte counters.  One vector per thread
// over a grid that covers the buffer, stores non-temporal: of the copy
// shapes tried on MI355X (grid-stride, 1-8 vectors per thread, 256-1024
// threads, non-temporal loads and stores) the fastest, ~6.2 TB/s read +
// write against 4.6 TB/s for a grid-stride loop over 8192 blocks.
#include <string>

#include "../../include/coolbubble.h"
#include "cb_sim.hpp"

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_copy16(const f4v *__restrict__ src, f4v *__restrict__ dst, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(src[i], dst + i);
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
    const long long blocks = (n + 255) / 256;
    if (e == hipSuccess && blocks > 0x7fffffffLL) e = hipErrorInvalidValue;
    const unsigned grid = (unsigned)blocks;
    if (e == hipSuccess) {
        k_copy16<<<grid, 256, 0, st>>>((const f4v *)a, (f4v *)b, n);
        e = hipEventRecord(e0, st);
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
            k_copy16<<<grid, 256, 0, st>>>((const f4v *)a, (f4v *)b, n);
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
