// cmg_tiles.hip -- device-resident exchange with the reference's own tile storage, and the
// per-tile region-of-interest flags.
//
// The reference keeps every tile as one PoissonTile<float,15> device allocation
// (src/PoissonTile.h:9-175): mData[15][736] floats (the 512 cell values of a channel
// first, then node values), mNeighbors[6], mCellType[512], ...; 44,880 bytes.  A caller
// that holds such tiles hands over, per tile of this grid's all-tiles order, the
// tile's device address (HATileInfo::mTilePtr of dAllTiles, src/HAGrid.h:654-685):
// a device array of N addresses, or a base address and a byte stride for tiles laid
// out in one slab.  One channel (or the cell types) moves per call, device to device.
//
// MarkRegionOfInterestWithChannelMinAndMax128Kernel (src/PoissonGridKernels.cuh:11-54):
// per tile, min and max of the 512 values of a channel, and a flag from a predicate on
// the pair; tiles outside the selected types get false.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "cmg_device.hpp"

namespace cmg {

constexpr size_t kTileImageBytes = 44880;          // sizeof(PoissonTile<float, 15>)
constexpr size_t kTileChannelFloats = 736;         // PoissonTile::CHNLSIZE
constexpr size_t kTileTypeOffset = 15 * 736 * 4 + 6 * 32;   // mCellType after mData and mNeighbors

namespace k {

// the address of tile t: from the pointer array, or base + t * stride
__device__ __forceinline__ const uint8_t* tile_addr(const uint64_t* ptrs, const uint8_t* base, size_t stride, int t)
{
    return ptrs ? reinterpret_cast<const uint8_t*>(ptrs[t]) : base + size_t(t) * stride;
}

// 128 threads a tile, one float4 (four cells) a thread; a tile whose address is null or
// not 16-byte aligned is skipped and counted in *bad
template <bool GATHER>
__global__ __launch_bounds__(128) void tile_channel_copy(const uint64_t* __restrict__ ptrs, uint8_t* base, size_t stride, int n,
                                                         int tile_channel, float* __restrict__ X, unsigned int* __restrict__ bad)
{
    const int t = blockIdx.x;
    if (t >= n) return;
    uint8_t* a = const_cast<uint8_t*>(tile_addr(ptrs, base, stride, t));
    if (!a || (reinterpret_cast<uintptr_t>(a) & 15)) {
        if (threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    float4* img = reinterpret_cast<float4*>(a + size_t(tile_channel) * kTileChannelFloats * sizeof(float));
    float4* own = reinterpret_cast<float4*>(X) + size_t(t) * 128;
    if (GATHER) own[threadIdx.x] = img[threadIdx.x];
    else img[threadIdx.x] = own[threadIdx.x];
}

template <bool GATHER>
__global__ __launch_bounds__(128) void tile_types_copy(const uint64_t* __restrict__ ptrs, uint8_t* base, size_t stride, int n,
                                                       uint8_t* __restrict__ T, unsigned int* __restrict__ bad)
{
    const int t = blockIdx.x;
    if (t >= n) return;
    uint8_t* a = const_cast<uint8_t*>(tile_addr(ptrs, base, stride, t));
    if (!a || (reinterpret_cast<uintptr_t>(a) & 15)) {
        if (threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    uchar4* img = reinterpret_cast<uchar4*>(a + kTileTypeOffset);
    uchar4* own = reinterpret_cast<uchar4*>(T) + size_t(t) * 128;
    if (GATHER) own[threadIdx.x] = img[threadIdx.x];
    else img[threadIdx.x] = own[threadIdx.x];
}

__device__ __forceinline__ float min_cub(float a, float b) { return a < b ? a : b; }   // CubMin
__device__ __forceinline__ float max_cub(float a, float b) { return a > b ? a : b; }   // CubMax

// one tile per 128-thread workgroup: per-thread min / max of a float4 (fminf / fmaxf, as
// the reference's min / max on floats), then the workgroup's reductions
__global__ __launch_bounds__(128) void roi_flags(TopoView tp, const float* __restrict__ X, int subtree_level, int launch_types,
                                                 int pred, float a, float b, int calc_locked, uint8_t* __restrict__ interest,
                                                 uint8_t* __restrict__ locked, float* __restrict__ tmin, float* __restrict__ tmax)
{
    __shared__ float red[2][2];
    const int t = blockIdx.x;
    const int lvl = tp.tlevel[t];
    uint8_t tt = tp.ttype[t];
    if (lvl == subtree_level && tt == NONLEAF) tt = LEAF;   // HATileInfo::subtreeType (src/HATileInfo.h:79-84)
    if (!(tt & launch_types)) {
        if (threadIdx.x == 0) {
            interest[t] = 0;
            if (calc_locked) locked[t] = 0;
        }
        return;
    }
    const float4 v = reinterpret_cast<const float4*>(X)[size_t(t) * 128 + threadIdx.x];
    float lo = fminf(fminf(v.x, v.y), fminf(v.z, v.w));
    float hi = fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w));
    for (int o = 32; o > 0; o >>= 1) {
        lo = min_cub(lo, __shfl_xor(lo, o, 64));
        hi = max_cub(hi, __shfl_xor(hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][0] = lo;
        red[threadIdx.x >> 6][1] = hi;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    lo = min_cub(red[0][0], red[1][0]);
    hi = max_cub(red[0][1], red[1][1]);
    bool f;
    switch (pred) {
    case CMG_ROI_MAX_GT: f = hi > a; break;
    case CMG_ROI_MIN_LT: f = lo < a; break;
    case CMG_ROI_RANGE_MEETS: f = lo <= b && hi >= a; break;
    default: f = (hi - lo) > a; break;   // CMG_ROI_SPAN_GT
    }
    interest[t] = f ? 1 : 0;
    if (calc_locked) locked[t] = f ? 1 : 0;
    if (tmin) tmin[t] = lo;
    if (tmax) tmax[t] = hi;
}

}  // namespace k

namespace {

void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// launches one of the copies and returns the number of tiles whose address was unusable
template <typename Launch>
unsigned int run_counted(DevGrid& d, Launch launch)
{
    if (!d.acount) throw std::logic_error("grid has no device counter");
    check_hip(hipMemsetAsync(d.acount, 0, sizeof(unsigned int), d.stream), "hipMemsetAsync");
    launch();
    check_hip(hipGetLastError(), "tile copy launch");
    unsigned int bad = 0;
    check_hip(hipMemcpyAsync(&bad, d.acount, sizeof(bad), hipMemcpyDeviceToHost, d.stream), "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
    return bad;
}

}  // namespace

unsigned int tile_channel_io(DevGrid& d, bool gather, const uint64_t* ptrs, void* base, size_t stride, int tile_channel,
                             float* X)
{
    if (d.N == 0) return 0;
    return run_counted(d, [&] {
        if (gather)
            k::tile_channel_copy<true><<<d.N, 128, 0, d.stream>>>(ptrs, static_cast<uint8_t*>(base), stride, d.N, tile_channel, X,
                                                                 d.acount);
        else
            k::tile_channel_copy<false><<<d.N, 128, 0, d.stream>>>(ptrs, static_cast<uint8_t*>(base), stride, d.N, tile_channel, X,
                                                                  d.acount);
    });
}

unsigned int tile_types_io(DevGrid& d, bool gather, const uint64_t* ptrs, void* base, size_t stride)
{
    if (d.N == 0) return 0;
    return run_counted(d, [&] {
        if (gather)
            k::tile_types_copy<true><<<d.N, 128, 0, d.stream>>>(ptrs, static_cast<uint8_t*>(base), stride, d.N, d.ctype, d.acount);
        else
            k::tile_types_copy<false><<<d.N, 128, 0, d.stream>>>(ptrs, static_cast<uint8_t*>(base), stride, d.N, d.ctype, d.acount);
    });
}

void launch_roi_flags(const DevGrid& d, const float* X, int subtree_level, int launch_types, int pred, float a, float b,
                      bool calc_locked, uint8_t* interest, uint8_t* locked, float* tmin, float* tmax, hipStream_t st)
{
    if (d.N == 0) return;
    k::roi_flags<<<d.N, 128, 0, st>>>(view(d), X, subtree_level, launch_types, pred, a, b, calc_locked ? 1 : 0, interest, locked,
                                      tmin, tmax);
    check_hip(hipGetLastError(), "roi_flags launch");
}

}  // namespace cmg
