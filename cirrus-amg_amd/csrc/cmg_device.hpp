// cmg_device.hpp -- device helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "cmg_kernels.hpp"

namespace cmg {
namespace k {

// LDS index of a cell of the 10^3 halo'd tile (local coords -1..8)
__device__ __forceinline__ int SX(int i, int j, int k) { return (i + 1) * 100 + (j + 1) * 10 + (k + 1); }
__device__ __forceinline__ int cidx(int i, int j, int k) { return (i << 6) | (j << 3) | k; }

__device__ __forceinline__ bool stopped(const double* S) { return S && S[S_DONE] != 0.0; }

// the list entry of this workgroup, or -2 when the solve has stopped; the entry's
// load is issued before the stop test so that the two round trips overlap
__device__ __forceinline__ int list_tile(const int32_t* list, const double* S)
{
    const int t = list[blockIdx.x];
    return stopped(S) ? -2 : t;
}

// wave64 + workgroup reduction of one double per thread; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < int(blockDim.x >> 6); ++i) s += red[i];
    return s;
}

__device__ __forceinline__ float tile_h(const TopoView& tp, int t) { return tp.h0 / float(1 << tp.tlevel[t]); }

// Grid-stride loops over the tiles of a list, 128 threads x float4 per tile,
// two tiles per 256-thread workgroup step.  Launch lists hold Morton-consecutive
// tiles 8 entries apart (entry b is meant for a one-tile workgroup on XCD b % 8),
// so step w takes entries 16 (w / 8) + w % 8 and the one 8 after it: two Morton
// neighbours of the XCD that a workgroup w runs on.
#define CMG_TILE_LOOP(list, n)                                                                      \
    for (int w_ = blockIdx.x; 16 * (w_ >> 3) < (n); w_ += gridDim.x)                               \
        if (const int pos = 16 * (w_ >> 3) + (w_ & 7) + 8 * int(threadIdx.x >> 7); pos < (n))       \
            if (const int t_ = (list)[pos]; t_ >= 0)

__device__ __forceinline__ size_t quad_index(int t) { return size_t(t) * 128 + (threadIdx.x & 127); }

}  // namespace k
}  // namespace cmg
