// cmg_project.hip -- the pressure-projection wrappers around the solve, on gfx950:
// the divergence that forms the right-hand side and the gradient that corrects
// the face velocities (src/AMGSolver.cu:634-835, src/PoissonGrid.cu:523-586,
// cirrus_cutcell/FluidEuler.cu:531-552).
//
// Face velocities live on the negative faces of each cell: channel u0 + a holds
// the velocity through the face between a cell and its -a neighbour.
// Arithmetic is written out (fp-contract off) to match the CPU restatement.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cmg_device.hpp"
#include "cmg_kernels.hpp"

#pragma clang fp contract(off)

namespace cmg {
namespace k {

// AccumulateFacesToParentsOneStepKernel (src/PoissonGrid.cu:523-569) on LEAF tiles,
// coefficient 1/4, overwrite, every cell type: the parent face of each 2x2 block
// of child faces is a quarter of their sum, summed (jj, kk) over the face axes
__global__ __launch_bounds__(64) void faces_to_parents(TopoView tp, const int32_t* __restrict__ list, float* __restrict__ U0,
                                                       float* __restrict__ U1, float* __restrict__ U2)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const int pt = tp.parent[t];
    if (pt < 0) return;
    const int v = threadIdx.x, oc = tp.toct[t];
    const int fl[3] = {v / 16 * 2, (v / 4) % 4 * 2, v % 4 * 2};
    const size_t pidx =
        size_t(pt) * 512 + cidx(4 * (oc & 1) + (fl[0] >> 1), 4 * ((oc >> 1) & 1) + (fl[1] >> 1), 4 * ((oc >> 2) & 1) + (fl[2] >> 1));
    float* const U[3] = {U0, U1, U2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float sum = 0;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                int f1[3] = {fl[0], fl[1], fl[2]};
                f1[(a + 1) % 3] += jj;
                f1[(a + 2) % 3] += kk;
                sum += U[a][size_t(t) * 512 + cidx(f1[0], f1[1], f1[2])];
            }
        U[a][pidx] = sum * 0.25f;
    }
}

// AMGVolumeWeightedDivergence[WithoutCoeff]OnLeafs (src/AMGSolver.cu:750-835), last step,
// on LEAF tiles: faces in the order x-, x+, y-, y+, z-, z+; the + face is the
// same-level neighbour cell's negative face (0 when the neighbour tile is absent)
template <bool COEF>
__global__ __launch_bounds__(512) void divergence(TopoView tp, const int32_t* __restrict__ list, const float* __restrict__ U0,
                                                  const float* __restrict__ U1, const float* __restrict__ U2, Coef C,
                                                  float* __restrict__ B)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const int v = threadIdx.x;
    const size_t idx = size_t(t) * 512 + v;
    const float h = tile_h(tp, t);
    const float* const U[3] = {U0, U1, U2};
    const int l[3] = {v >> 6, (v >> 3) & 7, v & 7};
    const int stride[3] = {64, 8, 1};
    float sum = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float u0 = U[a][idx];
        size_t ni = idx + stride[a];
        bool present = true;
        if (l[a] == 7) {
            const int nt = tp.nbr[size_t(t) * 6 + a + 3];
            present = nt >= 0;
            ni = size_t(nt) * 512 + (v - 7 * stride[a]);
        }
        const float u1 = present ? U[a][ni] : 0.0f;
        if (COEF) {
            const float k0 = C.c[a][idx];
            const float k1 = present ? C.c[a][ni] : 0.0f;
            sum += (u0 * k0) * h;
            sum += (-u1 * k1) * h;
        } else {
            sum += (-u0 * h) * h;
            sum += (u1 * h) * h;
        }
    }
    B[idx] = (tp.ctype[idx] & INTERIOR) ? sum : 0.0f;
}

// AMGAddGradientToFace128Kernel (src/AMGSolver.cu:634-678) / the face-weighted
// twin (:690-736) on LEAF tiles: u_a += (p - p_down) / h where the fluid ratio
// -c_a / h is positive (weighted: times the ratio); p through the operator's
// loader (INTERIOR mask, ghost extrapolation, x = 0 outside the domain)
__global__ __launch_bounds__(512) void add_gradient(TopoView tp, const int32_t* __restrict__ list, const float* __restrict__ X,
                                                    Coef C, float* __restrict__ U0, float* __restrict__ U1,
                                                    float* __restrict__ U2, int weighted)
{
    __shared__ float xs[1000];
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const int v = threadIdx.x;
    const size_t base = size_t(t) * 512;
    const int l[3] = {v >> 6, (v >> 3) & 7, v & 7};
    xs[SX(l[0], l[1], l[2])] = tp.ctype[base + v] == INTERIOR ? X[base + v] : 0.0f;
    __syncthreads();
    // the three negative halo faces (the + side is not read)
    if (v < 192) {
        const int axis = v >> 6, fi = v & 63;
        const int p = fi >> 3, q = fi & 7;
        int fl[3], nl[3];
        if (axis == 0) { fl[0] = -1; fl[1] = p; fl[2] = q; }
        else if (axis == 1) { fl[0] = p; fl[1] = -1; fl[2] = q; }
        else { fl[0] = p; fl[1] = q; fl[2] = -1; }
        nl[0] = fl[0]; nl[1] = fl[1]; nl[2] = fl[2];
        nl[axis] = 7;
        int cl[3] = {fl[0], fl[1], fl[2]};
        cl[axis] += 1;
        const int nt = tp.nbr[size_t(t) * 6 + axis];
        float hv = 0.0f;
        if (nt >= 0) {
            const size_t ni = size_t(nt) * 512 + cidx(nl[0], nl[1], nl[2]);
            if (tp.ttype[nt] & GHOST) {
                float avg = 0;
#pragma unroll
                for (int di = 0; di < 2; ++di)
#pragma unroll
                    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                        for (int dk = 0; dk < 2; ++dk) avg += xs[SX(cl[0] ^ di, cl[1] ^ dj, cl[2] ^ dk)];
                avg /= 8;
                hv = float(double(xs[SX(cl[0], cl[1], cl[2])]) + 0.5 * double(X[ni] - avg));
            } else {
                hv = tp.ctype[ni] == INTERIOR ? X[ni] : 0.0f;
            }
        }
        xs[SX(fl[0], fl[1], fl[2])] = hv;
    }
    __syncthreads();
    const float h = tile_h(tp, t);
    float* const U[3] = {U0, U1, U2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int d[3] = {l[0], l[1], l[2]};
        d[a] -= 1;
        const float pu = xs[SX(l[0], l[1], l[2])], pd = xs[SX(d[0], d[1], d[2])];
        const float ratio = -C.c[a][base + v] / h;
        float u = U[a][base + v];
        if (weighted) u = u + (pu - pd) / h * ratio;
        else if (ratio > 0) u = u + (pu - pd) / h;
        U[a][base + v] = u;
    }
}

// FluidEuler::mixFluidAndSolidVelocityOnFaces (cirrus_cutcell/FluidEuler.cu:531-552) on
// LEAF tiles, the solid face velocity given as channels
__global__ __launch_bounds__(512) void mix_velocity(TopoView tp, const int32_t* __restrict__ list, Coef C,
                                                    const float* __restrict__ F0, const float* __restrict__ F1,
                                                    const float* __restrict__ F2, const float* __restrict__ S0,
                                                    const float* __restrict__ S1, const float* __restrict__ S2,
                                                    float* __restrict__ M0, float* __restrict__ M1, float* __restrict__ M2)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const size_t idx = size_t(t) * 512 + threadIdx.x;
    const float h = tile_h(tp, t);
    const float* const F[3] = {F0, F1, F2};
    const float* const Sv[3] = {S0, S1, S2};
    float* const M[3] = {M0, M1, M2};
    float out[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float fr = -C.c[a][idx] / h;
        const float sr = 1 - fr;
        float mix = 0;
        if (fr > 0) mix += fr * F[a][idx];
        if (sr > 0) mix += sr * Sv[a][idx];
        out[a] = mix;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) M[a][idx] = out[a];
}

}  // namespace k

void launch_faces_to_parents(const DevGrid& d, float* U0, float* U1, float* U2, hipStream_t st)
{
    if (d.leafs.n <= 0) return;
    k::faces_to_parents<<<d.leafs.n, 64, 0, st>>>(view(d), d.leafs.d, U0, U1, U2);
}

void launch_divergence(const DevGrid& d, const float* U0, const float* U1, const float* U2, const Coef* C, float* B,
                       hipStream_t st)
{
    if (d.leafs.n <= 0) return;
    if (C) k::divergence<true><<<d.leafs.n, 512, 0, st>>>(view(d), d.leafs.d, U0, U1, U2, *C, B);
    else k::divergence<false><<<d.leafs.n, 512, 0, st>>>(view(d), d.leafs.d, U0, U1, U2, Coef{}, B);
}

void launch_add_gradient(const DevGrid& d, const float* X, Coef C, float* U0, float* U1, float* U2, bool weighted,
                         hipStream_t st)
{
    if (d.leafs.n <= 0) return;
    k::add_gradient<<<d.leafs.n, 512, 0, st>>>(view(d), d.leafs.d, X, C, U0, U1, U2, int(weighted));
}

void launch_mix_velocity(const DevGrid& d, Coef C, const float* const F[3], const float* const S[3], float* const M[3],
                         hipStream_t st)
{
    if (d.leafs.n <= 0) return;
    k::mix_velocity<<<d.leafs.n, 512, 0, st>>>(view(d), d.leafs.d, C, F[0], F[1], F[2], S[0], S[1], S[2], M[0], M[1], M[2]);
}

}  // namespace cmg
