// cb_tile_common.hpp -- small device helpers of the tiled step kernels.
#pragma once

#include "cb_common.hpp"

namespace cb {

template <typename T, int VW>
struct alignas(sizeof(T) * VW) Vec {
    T v[VW];
};

// Wait for this wave's LDS traffic, then the workgroup barrier.  The kernels
// communicate through LDS only (global data is read from buffers nobody
// writes during the launch), so no global-memory fence is needed.
__device__ inline void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Wait until every vector memory load (direct-to-LDS loads included) and
// store of this wave has completed: s_waitcnt vmcnt(0), expcnt and lgkmcnt
// left at their maxima.  Written as the builtin so that the compiler's own
// wait insertion sees it.
__device__ inline void vm_wait_all()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);
}

template <typename T, int VW>
__device__ inline void store_vec(T *p, const Vec<T, VW> &v, bool nt)
{
    if (nt) {
        typedef T nvec __attribute__((ext_vector_type(VW)));
        nvec x;
#pragma unroll
        for (int e = 0; e < VW; ++e) x[e] = v.v[e];
        __builtin_nontemporal_store(x, reinterpret_cast<nvec *>(p));
    } else {
        *reinterpret_cast<Vec<T, VW> *>(p) = v;
    }
}

} // namespace cb
