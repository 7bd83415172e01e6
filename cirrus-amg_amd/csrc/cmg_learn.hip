// cmg_learn.hip -- the learned tile smoother on gfx950: data gathering from a
// grid, FP32 training kernels (forward, loss, backward, weight gradients), the
// Adam update, and BF16 MFMA inference.
//
// Network (no reference counterpart): a 3-layer 3D conv net on one halo'd 10^3
// tile, 1 -> 16 -> 16 -> 1 channels, 3^3 kernels, zero "same" padding, ReLU
// after layers 1 and 2, no biases (so net(c x) = c net(x) for c > 0).  Weights
// are one flat array of 7776 floats: W1[16][27], W2[16 out][16 in][27],
// W3[16 in][27]; tap (di+1)*9 + (dj+1)*3 + (dk+1).  Volume index (i*10+j)*10+k.
// Input feature of a cell: r / D on INTERIOR cells of LEAF tiles of the tile's
// level (0 elsewhere); the output on the tile's own 8^3 cells is a correction
// of x.  Loss of a tile: sum over the interior 8^3 of (7-point Laplacian of
// (y - e))^2, mean over the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "cmg_device.hpp"
#include "cmg_kernels.hpp"

#pragma clang fp contract(off)

namespace cmg {
namespace k {

constexpr int NV = 1000, NC = 16, NW = 7776, OW2 = 432, OW3 = 7344;

__device__ __forceinline__ int vid(int i, int j, int k) { return (i * 10 + j) * 10 + k; }

// round to the nearest bfloat16 (ties to even) and back; NaN is kept quiet
__device__ __forceinline__ float bf16_round(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

__device__ __forceinline__ uint16_t bf16_bits(float f) { return uint16_t(__float_as_uint(bf16_round(f)) >> 16); }

// ---------------------------------------------------------------- gather ----
// one halo'd 10^3 volume per tile of `list`: mode 0 the input feature r / D,
// mode 1 the raw value; cells of the tile and of its same-level LEAF
// neighbours (INTERIOR only), 0 elsewhere (corners of the volume included)
__global__ __launch_bounds__(256) void gather_tiles(TopoView tp, const int32_t* __restrict__ tiles, int n,
                                                    const float* __restrict__ V, const float* __restrict__ D, int mode,
                                                    float* __restrict__ out)
{
    const int b = blockIdx.x;
    if (b >= n) return;
    const int t = tiles[b];
    for (int p = threadIdx.x; p < NV; p += blockDim.x) {
        const int i = p / 100 - 1, j = (p / 10) % 10 - 1, kk = p % 10 - 1;
        const int out_axes = (i < 0 || i > 7) + (j < 0 || j > 7) + (kk < 0 || kk > 7);
        float val = 0.0f;
        if (out_axes <= 1) {
            int tt = t, l[3] = {i, j, kk};
            if (out_axes == 1) {
                const int a = (i < 0 || i > 7) ? 0 : (j < 0 || j > 7) ? 1 : 2;
                const bool neg = l[a] < 0;
                tt = tp.nbr[size_t(t) * 6 + (neg ? a : a + 3)];
                l[a] = neg ? 7 : 0;
                if (tt >= 0 && !(tp.ttype[tt] & LEAF)) tt = -1;
            }
            if (tt >= 0) {
                const size_t idx = size_t(tt) * 512 + cidx(l[0], l[1], l[2]);
                if (tp.ctype[idx] & INTERIOR) {
                    const float v = V[idx];
                    if (mode == 0) {
                        const float d = D[idx];
                        val = d != 0 ? v / d : 0.0f;
                    } else {
                        val = v;
                    }
                }
            }
        }
        out[size_t(b) * NV + p] = val;
    }
}

// ---------------------------------------------------------------- training ----
// One workgroup (1024 threads) per chunk of tiles; x, a1, a2 and the loss
// buffers live in LDS.  BF16: weights and activations are rounded to bfloat16
// in the forward pass (FP32 accumulation); the backward pass is FP32.
struct TrainLds {
    float x[NV];
    float a1[NC * NV];
    float a2[NC * NV];
    float g[NV];    // y, then d = y - e, then dL/dy
    float s[NV];    // interior Laplacian of d
    double red[16];
};

template <bool BF16>
__global__ __launch_bounds__(1024) void net_train(const float* __restrict__ X, const float* __restrict__ E,
                                                  const int32_t* __restrict__ idx, int n, int per_wg,
                                                  const float* __restrict__ W, float* __restrict__ gpart,
                                                  double* __restrict__ lpart)
{
    extern __shared__ unsigned char smem[];
    TrainLds& s = *reinterpret_cast<TrainLds*>(smem);
    const int tid = threadIdx.x, nth = blockDim.x;
    // this thread's weight-gradient accumulators: weights tid, tid + nth, ...
    constexpr int WPT = (NW + 1023) / 1024;
    float gacc[WPT];
#pragma unroll
    for (int q = 0; q < WPT; ++q) gacc[q] = 0.0f;
    double lacc = 0;
    const int b0 = blockIdx.x * per_wg, b1 = min(n, b0 + per_wg);
    for (int bi = b0; bi < b1; ++bi) {
        const size_t b = idx ? size_t(idx[bi]) : size_t(bi);
        for (int p = tid; p < NV; p += nth) s.x[p] = X[b * NV + p];
        __syncthreads();
        // layer 1
        for (int o = tid; o < NC * NV; o += nth) {
            const int co = o / NV, p = o % NV, i = p / 100, j = (p / 10) % 10, kk = p % 10;
            float acc = 0.0f;
            for (int t = 0; t < 27; ++t) {
                const int a = i + t / 9 - 1, bb = j + (t / 3) % 3 - 1, c = kk + t % 3 - 1;
                if (a < 0 || a > 9 || bb < 0 || bb > 9 || c < 0 || c > 9) continue;
                float w = W[co * 27 + t], xv = s.x[vid(a, bb, c)];
                if (BF16) { w = bf16_round(w); xv = bf16_round(xv); }
                acc = fmaf(w, xv, acc);
            }
            acc = acc > 0 ? acc : 0.0f;
            s.a1[o] = BF16 ? bf16_round(acc) : acc;
        }
        __syncthreads();
        // layer 2
        for (int o = tid; o < NC * NV; o += nth) {
            const int co = o / NV, p = o % NV, i = p / 100, j = (p / 10) % 10, kk = p % 10;
            float acc = 0.0f;
            for (int t = 0; t < 27; ++t) {
                const int a = i + t / 9 - 1, bb = j + (t / 3) % 3 - 1, c = kk + t % 3 - 1;
                if (a < 0 || a > 9 || bb < 0 || bb > 9 || c < 0 || c > 9) continue;
                const int q = vid(a, bb, c);
                for (int ci = 0; ci < NC; ++ci) {
                    float w = W[OW2 + (co * NC + ci) * 27 + t];
                    if (BF16) w = bf16_round(w);
                    acc = fmaf(w, s.a1[ci * NV + q], acc);
                }
            }
            acc = acc > 0 ? acc : 0.0f;
            s.a2[o] = BF16 ? bf16_round(acc) : acc;
        }
        __syncthreads();
        // layer 3 and d = y - e
        for (int p = tid; p < NV; p += nth) {
            const int i = p / 100, j = (p / 10) % 10, kk = p % 10;
            float acc = 0.0f;
            for (int t = 0; t < 27; ++t) {
                const int a = i + t / 9 - 1, bb = j + (t / 3) % 3 - 1, c = kk + t % 3 - 1;
                if (a < 0 || a > 9 || bb < 0 || bb > 9 || c < 0 || c > 9) continue;
                const int q = vid(a, bb, c);
                for (int ci = 0; ci < NC; ++ci) {
                    float w = W[OW3 + ci * 27 + t];
                    if (BF16) w = bf16_round(w);
                    acc = fmaf(w, s.a2[ci * NV + q], acc);
                }
            }
            s.g[p] = acc - E[b * NV + p];
        }
        __syncthreads();
        // s = interior Laplacian of d, loss
        double lt = 0;
        for (int p = tid; p < NV; p += nth) {
            const int i = p / 100, j = (p / 10) % 10, kk = p % 10;
            float v = 0.0f;
            if (i >= 1 && i <= 8 && j >= 1 && j <= 8 && kk >= 1 && kk <= 8) {
                v = 6.0f * s.g[p] - s.g[p - 100] - s.g[p + 100] - s.g[p - 10] - s.g[p + 10] - s.g[p - 1] - s.g[p + 1];
                lt += double(v) * double(v);
            }
            s.s[p] = v;
        }
        lacc += lt;
        __syncthreads();
        // dL/dy = 2 Lap^T s / n
        const float sc = 2.0f / float(n);
        for (int p = tid; p < NV; p += nth) {
            const int i = p / 100, j = (p / 10) % 10, kk = p % 10;
            float v = 6.0f * s.s[p];
            if (i > 0) v -= s.s[p - 100];
            if (i < 9) v -= s.s[p + 100];
            if (j > 0) v -= s.s[p - 10];
            if (j < 9) v -= s.s[p + 10];
            if (kk > 0) v -= s.s[p - 1];
            if (kk < 9) v -= s.s[p + 1];
            s.g[p] = v * sc;
        }
        __syncthreads();
        // dL/dW3 (reads a2 before it is overwritten)
#pragma unroll
        for (int q = 0; q < WPT; ++q) {
            const int wi = tid + q * nth;
            if (wi < OW3 || wi >= NW) continue;
            const int ci = (wi - OW3) / 27, t = (wi - OW3) % 27;
            const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
            float acc = 0.0f;
            for (int p = 0; p < NV; ++p) {
                const int i = p / 100 + di, j = (p / 10) % 10 + dj, kk = p % 10 + dk;
                if (i < 0 || i > 9 || j < 0 || j > 9 || kk < 0 || kk > 9) continue;
                acc = fmaf(s.g[p], s.a2[ci * NV + vid(i, j, kk)], acc);
            }
            gacc[q] += acc;
        }
        __syncthreads();
        // dL/da2 masked by ReLU', in place of a2
        for (int o = tid; o < NC * NV; o += nth) {
            const int ci = o / NV, qv = o % NV, i = qv / 100, j = (qv / 10) % 10, kk = qv % 10;
            float acc = 0.0f;
            if (s.a2[o] > 0) {
                for (int t = 0; t < 27; ++t) {
                    const int a = i - (t / 9 - 1), bb = j - ((t / 3) % 3 - 1), c = kk - (t % 3 - 1);
                    if (a < 0 || a > 9 || bb < 0 || bb > 9 || c < 0 || c > 9) continue;
                    float w = W[OW3 + ci * 27 + t];
                    if (BF16) w = bf16_round(w);
                    acc = fmaf(w, s.g[vid(a, bb, c)], acc);
                }
            }
            s.a2[o] = acc;
        }
        __syncthreads();
        // dL/dW2
#pragma unroll
        for (int q = 0; q < WPT; ++q) {
            const int wi = tid + q * nth;
            if (wi < OW2 || wi >= OW3) continue;
            const int co = (wi - OW2) / (NC * 27), ci = ((wi - OW2) / 27) % NC, t = (wi - OW2) % 27;
            const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
            float acc = 0.0f;
            for (int p = 0; p < NV; ++p) {
                const int i = p / 100 + di, j = (p / 10) % 10 + dj, kk = p % 10 + dk;
                if (i < 0 || i > 9 || j < 0 || j > 9 || kk < 0 || kk > 9) continue;
                acc = fmaf(s.a2[co * NV + p], s.a1[ci * NV + vid(i, j, kk)], acc);
            }
            gacc[q] += acc;
        }
        __syncthreads();
        // dL/da1 masked by ReLU', in place of a1
        for (int o = tid; o < NC * NV; o += nth) {
            const int ci = o / NV, qv = o % NV, i = qv / 100, j = (qv / 10) % 10, kk = qv % 10;
            float acc = 0.0f;
            if (s.a1[o] > 0) {
                for (int t = 0; t < 27; ++t) {
                    const int a = i - (t / 9 - 1), bb = j - ((t / 3) % 3 - 1), c = kk - (t % 3 - 1);
                    if (a < 0 || a > 9 || bb < 0 || bb > 9 || c < 0 || c > 9) continue;
                    const int p = vid(a, bb, c);
                    for (int co = 0; co < NC; ++co) {
                        float w = W[OW2 + (co * NC + ci) * 27 + t];
                        if (BF16) w = bf16_round(w);
                        acc = fmaf(w, s.a2[co * NV + p], acc);
                    }
                }
            }
            s.a1[o] = acc;
        }
        __syncthreads();
        // dL/dW1
#pragma unroll
        for (int q = 0; q < WPT; ++q) {
            const int wi = tid + q * nth;
            if (wi >= OW2) continue;
            const int co = wi / 27, t = wi % 27;
            const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
            float acc = 0.0f;
            for (int p = 0; p < NV; ++p) {
                const int i = p / 100 + di, j = (p / 10) % 10 + dj, kk = p % 10 + dk;
                if (i < 0 || i > 9 || j < 0 || j > 9 || kk < 0 || kk > 9) continue;
                float xv = s.x[vid(i, j, kk)];
                if (BF16) xv = bf16_round(xv);
                acc = fmaf(s.a1[co * NV + p], xv, acc);
            }
            gacc[q] += acc;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < WPT; ++q) {
        const int wi = tid + q * nth;
        if (wi < NW) gpart[size_t(blockIdx.x) * NW + wi] = gacc[q];
    }
    const double lb = block_sum(lacc, s.red);
    if (tid == 0) lpart[blockIdx.x] = lb;
}

// grad[w] = sum over workgroups of the partials (fixed order, FP64); loss likewise
__global__ __launch_bounds__(256) void net_reduce(const float* __restrict__ gpart, const double* __restrict__ lpart, int nwg,
                                                  int n, float* __restrict__ grad, double* __restrict__ loss)
{
    const int wi = blockIdx.x * blockDim.x + threadIdx.x;
    if (wi < NW) {
        double acc = 0;
        for (int g = 0; g < nwg; ++g) acc += double(gpart[size_t(g) * NW + wi]);
        grad[wi] = float(acc);
    }
    if (wi == 0) {
        double l = 0;
        for (int g = 0; g < nwg; ++g) l += lpart[g];
        *loss = l / double(n);
    }
}

// Adam (Kingma and Ba, 2015), step t >= 1, FP32 state
__global__ __launch_bounds__(256) void adam(float* __restrict__ W, float* __restrict__ M, float* __restrict__ Vs,
                                            const float* __restrict__ G, int t, float lr, float b1, float b2, float eps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NW) return;
    const float c1 = 1.0f - powf(b1, float(t)), c2 = 1.0f - powf(b2, float(t));
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * Vs[i] + (1.0f - b2) * g * g;
    M[i] = m;
    Vs[i] = v;
    W[i] = W[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// -------------------------------------------------------------- inference ----
// BF16 MFMA forward pass on the tile's own 8^3 cells: y = net(x), x gathered
// from channel R / D as for training.  Persistent: each 256-thread workgroup
// (4 wave64s) holds the weights in registers and walks the tiles of a list.
// All three layers run on v_mfma_f32_16x16x{32,16}_bf16 with output channels
// (or taps) as the M side and 16 positions of the 10^3 volume as the N side:
//   layer 1: [16 co x 32 (27 taps + 5 zero)] x [32 x 16 positions]
//   layer 2: per tap pair, [16 co x 32 (2 taps x 16 ci)] x [32 x 16 positions]
//            (14 pairs, the 28th tap zero)
//   layer 3: U[tap][q] = sum_ci W3[ci][tap] a2[q][ci] for 9 taps at a time
//            ([16 (9 used) x 16 ci] x [16 x 16 positions]), then
//            y[p] = sum over taps of U[tap][p + off(tap)] in a fixed order.
// Activations are rounded to bfloat16 after each ReLU, products accumulate in
// FP32 (the same rounding points as the training forward pass in BF16 mode).
// The input of the next tile (r / D on 512 own and 384 face-halo cells, 3.5
// cells per thread) is loaded into registers while the current tile's layer 3
// is computed; the neighbours come from the tile's 32-byte halo record.
// LDS (71,488 B, two workgroups per CU).  Activations are stored by channel
// group so that the 16 positions one MFMA operand reads are contiguous (no bank
// conflicts): a1 as [2 halves of 8 ch][1000 positions][8 ch], a2 as
// [4 quarters of 4 ch][1000 positions][4 ch].
//   [0, 32000)        a1, bf16                         | U[9][1000] fp32
//   [0, 36000)                                          | (aliases a1)
//   [36000, 36032)    zero bytes (out-of-volume taps of layer 2)
//   [36032, 68032)    a2, bf16
//   [68032, 71488)    x on the zero-padded 12^3 volume, bf16 (edges and
//                     corners of the 10^3 volume and the padding stay 0)
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int IL_A1 = 0, IL_ZERO = 36000, IL_A2 = 36032, IL_X = 68032, IL_BYTES = 71488;
constexpr int IN_PT = 4;   // input cells per thread (896 = 512 own + 384 halo, 256 threads)

// registers of one tile's input: padded-volume position (-1: none), r, D, cell type
struct InferIn {
    int pos[IN_PT];
    float r[IN_PT], d[IN_PT];
    uint8_t ct[IN_PT];
};

__device__ __forceinline__ void infer_issue(InferIn& in, const TopoView& tp, int t, const float* __restrict__ R,
                                            const float* __restrict__ D)
{
#pragma unroll
    for (int m = 0; m < IN_PT; ++m) {
        in.pos[m] = -1;
        in.r[m] = 0.0f;
        in.d[m] = 0.0f;
        in.ct[m] = 0;
    }
    if (t < 0) return;
    const int4* hr = reinterpret_cast<const int4*>(tp.halo + size_t(t) * 8);
    const int4 ha = hr[0], hb = hr[1];
    const int nb[6] = {ha.x, ha.y, ha.z, ha.w, hb.x, hb.y};
    const uint32_t ty03 = uint32_t(hb.z), ty45 = uint32_t(hb.w);
#pragma unroll
    for (int m = 0; m < IN_PT; ++m) {
        const int e = threadIdx.x + 256 * m;
        if (e >= 896) continue;
        int tt, l0, l1, l2, p0, p1, p2;   // source tile, its cell, padded position
        if (e < 512) {
            tt = t;
            l0 = e >> 6; l1 = (e >> 3) & 7; l2 = e & 7;
            p0 = l0 + 2; p1 = l1 + 2; p2 = l2 + 2;
        } else {
            const int k = e - 512, f = k >> 6, fi = k & 63, axis = f >> 1;
            const bool neg = !(f & 1);
            const int slot = neg ? axis : axis + 3;
            int nt = nb[0];
            nt = slot == 1 ? nb[1] : nt;
            nt = slot == 2 ? nb[2] : nt;
            nt = slot == 3 ? nb[3] : nt;
            nt = slot == 4 ? nb[4] : nt;
            nt = slot == 5 ? nb[5] : nt;
            const uint32_t nty = (slot < 4 ? ty03 >> (8 * slot) : ty45 >> (8 * (slot - 4))) & 0xffu;
            tt = (nt >= 0 && (nty & LEAF)) ? nt : -1;
            const int a = fi >> 3, bq = fi & 7, lo = neg ? 7 : 0, po = neg ? 1 : 10;
            if (axis == 0) { l0 = lo; l1 = a; l2 = bq; p0 = po; p1 = a + 2; p2 = bq + 2; }
            else if (axis == 1) { l0 = a; l1 = lo; l2 = bq; p0 = a + 2; p1 = po; p2 = bq + 2; }
            else { l0 = a; l1 = bq; l2 = lo; p0 = a + 2; p1 = bq + 2; p2 = po; }
        }
        in.pos[m] = (p0 * 12 + p1) * 12 + p2;
        if (tt >= 0) {
            const size_t idx = size_t(tt) * 512 + cidx(l0, l1, l2);
            in.ct[m] = tp.ctype[idx];
            in.d[m] = D[idx];
            in.r[m] = R[idx];
        }
    }
}
constexpr int NBLK = (NV + 15) / 16;   // 63 blocks of 16 positions

__device__ __forceinline__ uint32_t pack_relu_bf16(float lo, float hi)
{
    return uint32_t(bf16_bits(lo > 0 ? lo : 0.0f)) | (uint32_t(bf16_bits(hi > 0 ? hi : 0.0f)) << 16);
}

__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// offset of tap t (0..26) in a volume of edge E
__device__ __forceinline__ int tap_off(int t, int E) { return (t / 9 - 1) * E * E + ((t / 3) % 3 - 1) * E + (t % 3 - 1); }

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void net_infer(TopoView tp, const int32_t* __restrict__ tiles, int n,
                                                 const float* __restrict__ R, const float* __restrict__ D,
                                                 const float* __restrict__ W, float* __restrict__ Y,
                                                 const double* __restrict__ S)
{
    extern __shared__ __align__(16) unsigned char lds[];
    if (stopped(S)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, grp = lane >> 4;
    // weights as MFMA A operands (row = col): layer 1 W1[co][tap 8 grp + j]
    s16x8 w1v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int tap = 8 * grp + j;
        w1v[j] = tap < 27 ? short(bf16_bits(W[col * 27 + tap])) : short(0);
    }
    // layer 2, pair m: tap 2m + (grp >> 1), input channel 8 (grp & 1) + j
    s16x8 w2v[14];
#pragma unroll
    for (int m = 0; m < 14; ++m) {
        const int tap = 2 * m + (grp >> 1);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int ci = 8 * (grp & 1) + j;
            w2v[m][j] = tap < 27 ? short(bf16_bits(W[OW2 + (col * NC + ci) * 27 + tap])) : short(0);
        }
    }
    // layer 3, group g: row = tap g*9 + col (col < 9), k = input channel 4 grp + j
    bf16x4 w3v[3];
#pragma unroll
    for (int g = 0; g < 3; ++g)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            w3v[g][j] = col < 9 ? short(bf16_bits(W[OW3 + (4 * grp + j) * 27 + g * 9 + col])) : short(0);
    // layer-1 tap offsets on the padded 12^3 volume (taps 27..31 are zero taps)
    int o1[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) o1[j] = 8 * grp + j < 27 ? tap_off(8 * grp + j, 12) : 0;
    if (tid < 8) reinterpret_cast<uint32_t*>(lds + IL_ZERO)[tid] = 0u;
    for (int e = tid; e < 1728 / 2; e += 256) reinterpret_cast<uint32_t*>(lds + IL_X)[e] = 0u;
    const uint16_t* xs = reinterpret_cast<const uint16_t*>(lds + IL_X);
    float* U = reinterpret_cast<float*>(lds);

    InferIn cur;
    infer_issue(cur, tp, blockIdx.x < n ? tiles[blockIdx.x] : -1, R, D);
    __syncthreads();
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const int t = tiles[b];
        // input on the zero-padded 12^3 volume: r / D on INTERIOR cells of the
        // tile and of its same-level LEAF face neighbours, 0 elsewhere
#pragma unroll
        for (int m = 0; m < IN_PT; ++m)
            if (cur.pos[m] >= 0) {
                const float val = (cur.ct[m] & INTERIOR) && cur.d[m] != 0 ? cur.r[m] / cur.d[m] : 0.0f;
                reinterpret_cast<uint16_t*>(lds + IL_X)[cur.pos[m]] = bf16_bits(val);
            }
        __syncthreads();
        const int bn = b + int(gridDim.x);
        if (t < 0) {
            infer_issue(cur, tp, bn < n ? tiles[bn] : -1, R, D);
            continue;
        }
        // layer 1
        for (int mb = wave; mb < NBLK; mb += 4) {
            const int p = mb * 16 + col;
            const bool ok = p < NV;
            const int pc = ok ? p : 0;
            const int c12 = ((pc / 100 + 1) * 12 + (pc / 10) % 10 + 1) * 12 + pc % 10 + 1;
            s16x8 xv;
#pragma unroll
            for (int j = 0; j < 8; ++j) xv[j] = (ok && 8 * grp + j < 27) ? short(xs[c12 + o1[j]]) : short(0);
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma32(w1v, xv, acc);
            if (ok) {
                uint2 v;
                v.x = pack_relu_bf16(acc[0], acc[1]);
                v.y = pack_relu_bf16(acc[2], acc[3]);
                *reinterpret_cast<uint2*>(lds + IL_A1 + (grp >> 1) * 16000 + p * 16 + (grp & 1) * 8) = v;
            }
        }
        __syncthreads();
        // layer 2: two position blocks per step (independent accumulators), the
        // next tap pair's operands read from LDS while this pair's MFMAs run
        for (int mb = wave; mb < NBLK; mb += 8) {
            const int p0 = mb * 16 + col, p1 = p0 + 64;
            const bool ok0 = p0 < NV, ok1 = p1 < NV;
            const int pi0 = p0 / 100, pj0 = (p0 / 10) % 10, pk0 = p0 % 10;
            const int pi1 = p1 / 100, pj1 = (p1 / 10) % 10, pk1 = p1 % 10;
            auto operand = [&](int m, int p, bool ok, int pi, int pj, int pk) {
                const int tap = 2 * m + (grp >> 1);
                const int di = tap / 9 - 1, dj = (tap / 3) % 3 - 1, dk = tap % 3 - 1;
                const bool in = ok && tap < 27 && unsigned(pi + di) < 10u && unsigned(pj + dj) < 10u && unsigned(pk + dk) < 10u;
                const int addr = in ? IL_A1 + (grp & 1) * 16000 + (p + di * 100 + dj * 10 + dk) * 16 : IL_ZERO;
                return *reinterpret_cast<const s16x8*>(lds + addr);
            };
            f32x4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
            s16x8 b0 = operand(0, p0, ok0, pi0, pj0, pk0), b1 = operand(0, p1, ok1, pi1, pj1, pk1);
#pragma unroll
            for (int m = 0; m < 14; ++m) {
                s16x8 n0 = b0, n1 = b1;
                if (m + 1 < 14) {
                    n0 = operand(m + 1, p0, ok0, pi0, pj0, pk0);
                    n1 = operand(m + 1, p1, ok1, pi1, pj1, pk1);
                }
                acc0 = mfma32(w2v[m], b0, acc0);
                acc1 = mfma32(w2v[m], b1, acc1);
                b0 = n0;
                b1 = n1;
            }
            if (ok0) {
                uint2 v;
                v.x = pack_relu_bf16(acc0[0], acc0[1]);
                v.y = pack_relu_bf16(acc0[2], acc0[3]);
                *reinterpret_cast<uint2*>(lds + IL_A2 + grp * 8000 + p0 * 8) = v;
            }
            if (ok1) {
                uint2 v;
                v.x = pack_relu_bf16(acc1[0], acc1[1]);
                v.y = pack_relu_bf16(acc1[2], acc1[3]);
                *reinterpret_cast<uint2*>(lds + IL_A2 + grp * 8000 + p1 * 8) = v;
            }
        }
        __syncthreads();
        // the next tile's input loads fly while layer 3 runs (issued here, after
        // layer 2, whose register needs are the kernel's peak)
        infer_issue(cur, tp, bn < n ? tiles[bn] : -1, R, D);
        // layer 3 in three groups of 9 taps (di = g - 1)
        float y0 = 0.0f, y1 = 0.0f;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            for (int mb = wave; mb < NBLK; mb += 4) {
                const int p = mb * 16 + col;
                const bool ok = p < NV;
                bf16x4 av = {0, 0, 0, 0};
                if (ok) av = *reinterpret_cast<const bf16x4*>(lds + IL_A2 + grp * 8000 + p * 8);
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w3v[g], av, acc, 0, 0, 0);
                if (ok) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (4 * grp + e < 9) U[(4 * grp + e) * NV + p] = acc[e];
                }
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int v = tid + 256 * h;
                const int q = ((v >> 6) + 1 + (g - 1)) * 100 + (((v >> 3) & 7) + 1) * 10 + (v & 7) + 1;
                float acc = h ? y1 : y0;
#pragma unroll
                for (int s9 = 0; s9 < 9; ++s9) acc += U[s9 * NV + q + (s9 / 3 - 1) * 10 + (s9 % 3 - 1)];
                if (h) y1 = acc; else y0 = acc;
            }
            __syncthreads();
        }
        const size_t base = size_t(t) * 512;
        Y[base + tid] = (tp.ctype[base + tid] & INTERIOR) ? y0 : 0.0f;
        Y[base + tid + 256] = (tp.ctype[base + tid + 256] & INTERIOR) ? y1 : 0.0f;
    }
}

// in-bounds check of the MFMA operand layouts: one wave64, C = A B for a
// 16 x K bf16 A and a K x 16 bf16 B (row-major bit patterns), K = 16 or 32
__global__ __launch_bounds__(64) void mfma_probe(int K, const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                 float* __restrict__ C)
{
    const int lane = threadIdx.x, row = lane & 15, grp = lane >> 4;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (K == 32) {
        s16x8 av, bv;
        for (int e = 0; e < 8; ++e) {
            av[e] = short(A[row * 32 + 8 * grp + e]);
            bv[e] = short(B[(8 * grp + e) * 16 + row]);
        }
        acc = mfma32(av, bv, acc);
    } else {
        bf16x4 av, bv;
        for (int e = 0; e < 4; ++e) {
            av[e] = short(A[row * 16 + 4 * grp + e]);
            bv[e] = short(B[(4 * grp + e) * 16 + row]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    for (int e = 0; e < 4; ++e) C[(4 * grp + e) * 16 + row] = acc[e];
}
}  // namespace k

// ---------------------------------------------------------------- launchers ---

void launch_gather_tiles(const DevGrid& d, const int32_t* tiles, int n, const float* V, const float* D, int mode, float* out,
                         hipStream_t st)
{
    if (n <= 0) return;
    k::gather_tiles<<<n, 256, 0, st>>>(view(d), tiles, n, V, D, mode, out);
}

int launch_net_train(const float* X, const float* E, const int32_t* idx, int n, const float* W, float* gpart, double* lpart,
                     int max_wg, bool bf16, hipStream_t st)
{
    if (n <= 0) return 0;
    const int nwg = std::max(1, std::min(max_wg, n));
    const int per = (n + nwg - 1) / nwg;
    const int used = (n + per - 1) / per;
    const size_t lds = sizeof(k::TrainLds);
    static bool attr = [] {
        (void)hipFuncSetAttribute((const void*)k::net_train<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(sizeof(k::TrainLds)));
        (void)hipFuncSetAttribute((const void*)k::net_train<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(sizeof(k::TrainLds)));
        return true;
    }();
    (void)attr;
    if (bf16) k::net_train<true><<<used, 1024, lds, st>>>(X, E, idx, n, per, W, gpart, lpart);
    else k::net_train<false><<<used, 1024, lds, st>>>(X, E, idx, n, per, W, gpart, lpart);
    return used;
}

void launch_net_reduce(const float* gpart, const double* lpart, int nwg, int n, float* grad, double* loss, hipStream_t st)
{
    k::net_reduce<<<(k::NW + 255) / 256, 256, 0, st>>>(gpart, lpart, nwg, n, grad, loss);
}

void launch_adam(float* W, float* M, float* V, const float* G, int t, float lr, float b1, float b2, float eps, hipStream_t st)
{
    k::adam<<<(k::NW + 255) / 256, 256, 0, st>>>(W, M, V, G, t, lr, b1, b2, eps);
}

void launch_net_infer(const DevGrid& d, const LaunchList& l, const float* R, const float* D, const float* W, float* Y,
                      const double* S, hipStream_t st)
{
    if (l.n <= 0) return;
    static bool attr = [] {
        (void)hipFuncSetAttribute((const void*)k::net_infer, hipFuncAttributeMaxDynamicSharedMemorySize, k::IL_BYTES);
        return true;
    }();
    (void)attr;
    // persistent: two workgroups per CU (68 KB of LDS each)
    const int grid = std::max(1, std::min(l.n, d.wg_cap / 4));
    k::net_infer<<<grid, 256, k::IL_BYTES, st>>>(view(d), l.d, l.n, R, D, W, Y, S);
}

void launch_mfma_probe(int K, const uint16_t* A, const uint16_t* B, float* C, hipStream_t st)
{
    k::mfma_probe<<<1, 64, 0, st>>>(K, A, B, C);
}

}  // namespace cmg
