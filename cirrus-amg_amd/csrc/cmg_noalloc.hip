// cmg_noalloc.hip -- cmg_op_norm without any allocation, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc).
//
// cmg_op_norm (NormSync, src/PoissonGrid.cu:322-375) allocates a weight buffer per
// call and sums the per-tile partials on the host.  This variant keeps both
// partial arrays and the two sums in one file-scope device array, `scratch`, of a
// fixed size, and sums the partials on the device in the same index order the
// host loop of cmg_op_norm uses, so the two entry points return the same bits.
//
// The array exists once per GPU for the life of the process.  Every call on a GPU
// overwrites it: calls that overlap on one GPU (two host threads, or two streams)
// race and must be serialised by the caller.  Nothing in it has to start at zero:
// every partial a call reads was written by the same call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "cmg_device.hpp"

namespace cmg {

// the most LEAF-list entries (workgroups of the per-tile partial kernel) one call takes
constexpr int kNoAllocMaxEntries = 65536;
constexpr int kScratchDoubles = 2 * kNoAllocMaxEntries + 8;

}  // namespace cmg

// [0, max): per-entry |x|^p sums; [max, 2 max): per-entry weights; [2 max, 2 max + 2): the two totals
static __device__ __align__(256) double scratch[cmg::kScratchDoubles];

namespace cmg {
namespace k {

// the host loop of cmg_op_norm, on the device: v = max or sum of vsum[i], w = sum of wsum[i],
// i ascending, one thread; the workgroup stages 4096 entries at a time in LDS so that the
// serial thread's loads do not wait on HBM
__global__ __launch_bounds__(1024) void norm_totals(const double* __restrict__ vsum, const double* __restrict__ wsum, int n,
                                                    int is_max, double* __restrict__ out)
{
    __shared__ double sv[4096], sw[4096];
    double v = 0.0, w = 0.0;
    for (int base = 0; base < n; base += 4096) {
        const int m = min(4096, n - base);
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
            sv[i] = vsum[base + i];
            sw[i] = wsum[base + i];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < m; ++i) {
                // std::max(v, a): a only when v < a (so a NaN partial never replaces v)
                v = is_max ? (v < sv[i] ? sv[i] : v) : v + sv[i];
                w = w + sw[i];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = v;
        out[1] = w;
    }
}

}  // namespace k

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the device that owns `stream`; looked up once per
// device with that device current (each GPU holds its own copy of the array)
double* scratch_on(hipStream_t stream)
{
    hipDevice_t dev = 0;
    hipError_t e = hipStreamGetDevice(stream, &dev);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipStreamGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        int cur = 0;
        e = hipGetDevice(&cur);
        if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess)
            throw std::runtime_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

}  // namespace

// cmg_op_norm's result from the same partials, with no allocation: the grid must be on the
// device and the channel allocated there already (both allocate on first use otherwise)
cmg_status op_norm_noalloc(cmg_grid* g, int order, int ch, int volume_weighted, int cell_types, double* out)
{
    if (!g || !g->topo || !out || ch < 0 || ch >= kChannels || order == 0 || order < -1 || (order == -1 && volume_weighted))
        return CMG_ERR_BAD_ARG;
    if (!g->topo->finalized() || !g->uploaded) {
        set_error("cmg_op_norm_noalloc: the grid is not on the device yet (a call that uploads it, e.g. cmg_op_norm, "
                  "allocates; make one first)");
        return CMG_ERR_STATE;
    }
    DevGrid& d = g->dev;
    if (!d.ch[ch]) {
        set_error("cmg_op_norm_noalloc: channel " + std::to_string(ch) + " is not allocated on the device");
        return CMG_ERR_STATE;
    }
    if (d.leafs.n > kNoAllocMaxEntries) {
        set_error("cmg_op_norm_noalloc: " + std::to_string(d.leafs.n) + " LEAF-list entries exceed the maximum of " +
                  std::to_string(kNoAllocMaxEntries) + " (use cmg_op_norm)");
        return CMG_ERR_BAD_ARG;
    }
    try {
        double* sc = scratch_on(d.stream);
        double* vsum = sc;
        double* wsum = sc + kNoAllocMaxEntries;
        double* tot = sc + 2 * kNoAllocMaxEntries;
        const int n = launch_power_sum(d, d.ch[ch], order, volume_weighted != 0, cell_types, vsum, wsum, d.stream);
        double res[2] = {0.0, 0.0};
        if (n) {
            k::norm_totals<<<1, 1024, 0, d.stream>>>(vsum, wsum, n, order == -1 ? 1 : 0, tot);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) throw std::runtime_error(std::string("norm_totals launch: ") + hipGetErrorString(e));
            e = hipMemcpyAsync(res, tot, sizeof(res), hipMemcpyDeviceToHost, d.stream);
            if (e != hipSuccess) throw std::runtime_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
        }
        hipError_t e = hipStreamSynchronize(d.stream);
        if (e != hipSuccess) throw std::runtime_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        const double vs = res[0], ws = res[1];
        if (order == -1) *out = vs;
        else if (ws == 0) *out = 0;
        else if (order == 1) *out = vs / ws;
        else if (order == 2) *out = std::sqrt(vs / ws);
        else *out = std::pow(vs / ws, 1.0 / order);
    } catch (const std::exception& ex) {
        set_error(ex.what());
        return CMG_ERR_HIP;
    }
    return CMG_OK;
}

}  // namespace cmg
