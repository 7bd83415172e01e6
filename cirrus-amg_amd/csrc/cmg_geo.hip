// cmg_geo.hip -- gfx950 kernels of the two geometric multigrid solvers:
// GMGSolver (src/GMGSolver.cu) and CMGSolver (src/CMGSolver.cu).
//
// Both operators come from cell and tile types alone:
//   GMG  face coefficient (1/dh)(1/h): dh = h between LEAF cells, 1.5 h
//        between a LEAF and a GHOST cell, else 0; 0 next to a NEUMANN cell;
//        the cell sum is scaled by h^3 (GMGSolver.cu:3-67)
//   CMG  face coefficient h, 0 next to a NEUMANN cell; GHOST neighbours are
//        extrapolated as in the AMG loader (CMGSolver.cu:3-170)
// One 8^3 tile per 512-thread workgroup; x (unmasked), cell types and the
// face neighbours' subtree types are staged in LDS.  Arithmetic is written out
// (fp-contract off, explicit fmaf) so that results match the CPU restatement
// bit for bit:  sum = fmaf(c, x0 - x1, sum) over axis 0..2, side -,+.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cmg_device.hpp"
#include "cmg_kernels.hpp"

#pragma clang fp contract(off)

namespace cmg {
namespace k {

constexpr int GEO_GMG = 0, GEO_CMG = 1;

struct GeoLds {
    float x[1000];
    uint8_t ct[1000];
    uint8_t ftt[6];   // subtree type behind face f = 2 * axis + (side > 0)
    uint8_t own;
};

// HATileInfo::subtreeType (src/HATileInfo.h:79-84); subtree_level < 0: the own type
__device__ __forceinline__ uint8_t subtree_type(uint8_t type, int level, int subtree_level)
{
    return (level == subtree_level && type == NONLEAF) ? uint8_t(LEAF) : type;
}

// LoadGMGLaplacianTileData (GMGSolver.cu:81-133) / LoadCMGLaplacianTileData (CMGSolver.cu:62-170)
template <int MODE>
__device__ __forceinline__ void geo_load(const TopoView& tp, int t, int subtree_level, const float* __restrict__ X, GeoLds& s)
{
    const int v = threadIdx.x;
    const size_t base = size_t(t) * 512;
    s.x[SX(v >> 6, (v >> 3) & 7, v & 7)] = X[base + v];
    s.ct[SX(v >> 6, (v >> 3) & 7, v & 7)] = tp.ctype[base + v];
    if (v < 7) {
        const uint8_t own = subtree_type(tp.ttype[t], tp.tlevel[t], subtree_level);
        if (v == 6) {
            s.own = own;
        } else {
            const int axis = v >> 1, pos = v & 1;
            const int nt = tp.nbr[size_t(t) * 6 + (pos ? axis + 3 : axis)];
            s.ftt[v] = nt < 0 ? own : subtree_type(tp.ttype[nt], tp.tlevel[nt], subtree_level);
        }
    }
    __syncthreads();
    if (v < 384) {
        const int f = v >> 6, fi = v & 63;
        const int axis = f >> 1, sgn = (f & 1) ? 1 : -1;
        const int p = fi >> 3, q = fi & 7;
        int fl[3], nl[3];
        const int axk = sgn < 0 ? -1 : 8;
        if (axis == 0) { fl[0] = axk; fl[1] = p; fl[2] = q; }
        else if (axis == 1) { fl[0] = p; fl[1] = axk; fl[2] = q; }
        else { fl[0] = p; fl[1] = q; fl[2] = axk; }
        nl[0] = fl[0]; nl[1] = fl[1]; nl[2] = fl[2];
        nl[axis] = sgn < 0 ? 7 : 0;
        const int nt = tp.nbr[size_t(t) * 6 + (sgn < 0 ? axis : axis + 3)];
        float xv = 0.0f;
        uint8_t cv = DIRICHLET;
        if (nt >= 0) {
            const size_t ni = size_t(nt) * 512 + cidx(nl[0], nl[1], nl[2]);
            cv = tp.ctype[ni];
            const float xn = X[ni];
            if (MODE == GEO_CMG && s.ftt[f] == GHOST) {
                int cl[3] = {fl[0], fl[1], fl[2]};
                cl[axis] -= sgn;
                float avg = 0;
#pragma unroll
                for (int di = 0; di < 2; ++di)
#pragma unroll
                    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                        for (int dk = 0; dk < 2; ++dk) avg += s.x[SX(cl[0] ^ di, cl[1] ^ dj, cl[2] ^ dk)];
                avg /= 8;
                xv = float(double(s.x[SX(cl[0], cl[1], cl[2])]) + 0.5 * double(xn - avg));
            } else {
                xv = xn;
            }
        }
        s.x[SX(fl[0], fl[1], fl[2])] = xv;
        s.ct[SX(fl[0], fl[1], fl[2])] = cv;
    }
    __syncthreads();
}

// NegativeLaplacianCoeff (GMGSolver.cu:3-30 / CMGSolver.cu:3-7)
template <int MODE>
__device__ __forceinline__ float geo_coeff(float h, float one_over_h, uint8_t t0, uint8_t t1, uint8_t c0, uint8_t c1)
{
    const bool neu = (c0 & NEUMANN) || (c1 & NEUMANN);
    if (MODE == GEO_CMG) return neu ? 0.0f : h;
    const bool both = (t0 & LEAF) && (t1 & LEAF);
    const bool lg = ((t0 & LEAF) && (t1 & GHOST)) || ((t0 & GHOST) && (t1 & LEAF));
    const float odh = both ? one_over_h : lg ? (2 * one_over_h / 3) : 0.0f;
    return neu ? 0.0f : odh * one_over_h;
}

// GMGLaplacianTileData::negativeLap (GMGSolver.cu:45-66) / CMG (CMGSolver.cu:23-45) for cell v
template <int MODE, bool DIAG>
__device__ __forceinline__ float geo_neg_lap(const GeoLds& s, float h, int v)
{
    const float one_over_h = 1 / h;
    const int l[3] = {v >> 6, (v >> 3) & 7, v & 7};
    const int c = SX(l[0], l[1], l[2]);
    const uint8_t t0 = s.own, c0 = s.ct[c];
    const float x0 = s.x[c];
    const int step[3] = {100, 10, 1};
    float sum = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int sg = -1; sg <= 1; sg += 2) {
            const int n = c + sg * step[a];
            const bool inside = sg < 0 ? l[a] > 0 : l[a] < 7;
            const uint8_t t1 = inside ? t0 : s.ftt[2 * a + (sg > 0)];
            const float cf = geo_coeff<MODE>(h, one_over_h, t0, t1, c0, s.ct[n]);
            if (DIAG) sum = sum + cf;
            else sum = fmaf(cf, x0 - s.x[n], sum);
        }
    if (!(c0 & INTERIOR)) return 0.0f;
    return MODE == GEO_GMG ? sum * h * h * h : sum;
}

// NegativeLaplacianSameLevel128Kernel (GMGSolver.cu:135-161) / CMG twin (CMGSolver.cu:172-201)
template <int MODE>
__global__ __launch_bounds__(512) void geo_apply(TopoView tp, const int32_t* __restrict__ list, int subtree_level,
                                                 const float* __restrict__ X, float* __restrict__ Y, int diag,
                                                 const double* __restrict__ S)
{
    __shared__ GeoLds s;
    if (stopped(S)) return;
    const int t = list[blockIdx.x];
    if (t < 0) return;
    geo_load<MODE>(tp, t, subtree_level, X, s);
    const float h = tile_h(tp, t);
    const int v = threadIdx.x;
    Y[size_t(t) * 512 + v] = diag ? geo_neg_lap<MODE, true>(s, h, v) : geo_neg_lap<MODE, false>(s, h, v);
}

// one colour of NegativeLaplacianAndGaussSeidelSameLevel128Kernel (GMGSolver.cu:173-206 / CMGSolver.cu:228-264)
template <int MODE>
__global__ __launch_bounds__(512) void geo_gs(TopoView tp, const int32_t* __restrict__ list, int level, float* __restrict__ X,
                                              const float* __restrict__ B, int color, const double* __restrict__ S)
{
    __shared__ GeoLds s;
    if (stopped(S)) return;
    const int t = list[blockIdx.x];
    if (t < 0) return;
    geo_load<MODE>(tp, t, level, X, s);
    const float h = tile_h(tp, t);
    const int v = threadIdx.x;
    if ((((v >> 6) + ((v >> 3) & 7) + (v & 7)) & 1) != color) return;
    const size_t idx = size_t(t) * 512 + v;
    const int c = SX(v >> 6, (v >> 3) & 7, v & 7);
    const float b = B[idx];
    const float ax = geo_neg_lap<MODE, false>(s, h, v);
    const float d = geo_neg_lap<MODE, true>(s, h, v);
    const float delta = (s.ct[c] & INTERIOR) ? (b - ax) / d : 0.0f;
    X[idx] = s.x[c] + delta;
}

// ResidualAndRestrict128Kernel (GMGSolver.cu:215-275) / CMG twin (CMGSolver.cu:273-330):
// parent b = sum8 r * one_over_alpha; a GHOST tile (GMG) adds sum8 r to its parent
template <int MODE>
__global__ __launch_bounds__(512) void geo_restrict(TopoView tp, const int32_t* __restrict__ list, int level,
                                                    const float* __restrict__ X, float* __restrict__ B, float one_over_alpha,
                                                    const double* __restrict__ S)
{
    __shared__ GeoLds s;
    __shared__ float res[512];
    if (stopped(S)) return;
    const int t = list[blockIdx.x];
    if (t < 0) return;
    geo_load<MODE>(tp, t, level, X, s);
    const float h = tile_h(tp, t);
    const int v = threadIdx.x;
    res[v] = B[size_t(t) * 512 + v] - geo_neg_lap<MODE, false>(s, h, v);
    __syncthreads();
    if (v >= 64) return;
    const int pt = tp.parent[t];
    if (pt < 0) return;
    const int oc = tp.toct[t];
    const int fl0 = v / 16 * 2, fl1 = (v / 4) % 4 * 2, fl2 = v % 4 * 2;
    float sum = 0;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) sum += res[cidx(fl0 + ii, fl1 + jj, fl2 + kk)];
    const size_t pidx =
        size_t(pt) * 512 + cidx(4 * (oc & 1) + (fl0 >> 1), 4 * ((oc >> 1) & 1) + (fl1 >> 1), 4 * ((oc >> 2) & 1) + (fl2 >> 1));
    if (s.own & GHOST) B[pidx] = B[pidx] + sum;
    else B[pidx] = sum * one_over_alpha;
}

// AccumulateToParentsOneStep(GHOST, coeff, additive, all cell types) (src/PoissonGrid.cu:408-465)
__global__ __launch_bounds__(64) void accumulate_add(TopoView tp, const int32_t* __restrict__ list, float* __restrict__ Y,
                                                     float coeff, const double* __restrict__ S)
{
    if (stopped(S)) return;
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const int pt = tp.parent[t];
    if (pt < 0) return;
    const int v = threadIdx.x, oc = tp.toct[t];
    const int fl0 = v / 16 * 2, fl1 = (v / 4) % 4 * 2, fl2 = v % 4 * 2;
    float sum = 0;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) sum += Y[size_t(t) * 512 + cidx(fl0 + ii, fl1 + jj, fl2 + kk)];
    const size_t pidx =
        size_t(pt) * 512 + cidx(4 * (oc & 1) + (fl0 >> 1), 4 * ((oc >> 1) & 1) + (fl1 >> 1), 4 * ((oc >> 2) & 1) + (fl2 >> 1));
    Y[pidx] = Y[pidx] + sum * coeff;
}

// r = b - Ax on INTERIOR cells of LEAF tiles, Ax kept elsewhere (GMGSolver.cu:421-431)
__global__ __launch_bounds__(256) void residual_interior_v(TopoView tp, const int32_t* __restrict__ list, int n,
                                                           const float* __restrict__ B, float* __restrict__ AX)
{
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        const float4 b = reinterpret_cast<const float4*>(B)[q];
        float4 a = reinterpret_cast<const float4*>(AX)[q];
        if (ct.x & INTERIOR) a.x = b.x - a.x;
        if (ct.y & INTERIOR) a.y = b.y - a.y;
        if (ct.z & INTERIOR) a.z = b.z - a.z;
        if (ct.w & INTERIOR) a.w = b.w - a.w;
        reinterpret_cast<float4*>(AX)[q] = a;
    }
}

// damped-Jacobi update x += omega r / D on INTERIOR cells, x = 0 where D = 0 (GMGSolver.cu:448-466)
__global__ __launch_bounds__(256) void djacobi_update_v(TopoView tp, const int32_t* __restrict__ list, int n,
                                                        float* __restrict__ X, const float* __restrict__ R,
                                                        const float* __restrict__ D, float omega)
{
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        const float4 r = reinterpret_cast<const float4*>(R)[q];
        const float4 d = reinterpret_cast<const float4*>(D)[q];
        float4 x = reinterpret_cast<const float4*>(X)[q];
        if (ct.x & INTERIOR) x.x = d.x == 0 ? 0.0f : x.x + omega * r.x / d.x;
        if (ct.y & INTERIOR) x.y = d.y == 0 ? 0.0f : x.y + omega * r.y / d.y;
        if (ct.z & INTERIOR) x.z = d.z == 0 ? 0.0f : x.z + omega * r.z / d.z;
        if (ct.w & INTERIOR) x.w = d.w == 0 ? 0.0f : x.w + omega * r.w / d.w;
        reinterpret_cast<float4*>(X)[q] = x;
    }
}

}  // namespace k

// ---------------------------------------------------------------- launchers ---

static inline int geo_vector_grid(const DevGrid& d, int n) { return std::max(1, std::min((n + 1) / 2, d.wg_cap)); }

void launch_geo_apply(const DevGrid& d, int mode, const LaunchList& l, int subtree_level, const float* X, float* Y, bool diag,
                      const double* S, hipStream_t st)
{
    if (l.n <= 0) return;
    if (mode == k::GEO_GMG) k::geo_apply<k::GEO_GMG><<<l.n, 512, 0, st>>>(view(d), l.d, subtree_level, X, Y, int(diag), S);
    else k::geo_apply<k::GEO_CMG><<<l.n, 512, 0, st>>>(view(d), l.d, subtree_level, X, Y, int(diag), S);
}

void launch_geo_gs(const DevGrid& d, int mode, int level, float* X, const float* B, int color, const double* S, hipStream_t st)
{
    const LaunchList& l = d.subtree[level];
    if (l.n <= 0) return;
    if (mode == k::GEO_GMG) k::geo_gs<k::GEO_GMG><<<l.n, 512, 0, st>>>(view(d), l.d, level, X, B, color, S);
    else k::geo_gs<k::GEO_CMG><<<l.n, 512, 0, st>>>(view(d), l.d, level, X, B, color, S);
}

void launch_geo_restrict(const DevGrid& d, int mode, int level, const float* X, float* B, float one_over_alpha,
                         const double* S, hipStream_t st)
{
    // GMG restricts from every tile of the level (GHOST included), CMG from LEAF|NONLEAF tiles
    const LaunchList& l = mode == k::GEO_GMG ? d.prolong[level] : d.subtree[level];
    if (l.n <= 0) return;
    if (mode == k::GEO_GMG) k::geo_restrict<k::GEO_GMG><<<l.n, 512, 0, st>>>(view(d), l.d, level, X, B, one_over_alpha, S);
    else k::geo_restrict<k::GEO_CMG><<<l.n, 512, 0, st>>>(view(d), l.d, level, X, B, one_over_alpha, S);
}

void launch_accumulate_ghosts_add(const DevGrid& d, float* Y, float coeff, const double* S, hipStream_t st)
{
    if (d.ghosts.n <= 0) return;
    k::accumulate_add<<<d.ghosts.n, 64, 0, st>>>(view(d), d.ghosts.d, Y, coeff, S);
}

void launch_residual_interior(const DevGrid& d, const float* B, float* AX, hipStream_t st)
{
    if (d.leafs.n <= 0) return;
    k::residual_interior_v<<<geo_vector_grid(d, d.leafs.n), 256, 0, st>>>(view(d), d.leafs.d, d.leafs.n, B, AX);
}

void launch_djacobi_update(const DevGrid& d, float* X, const float* R, const float* D, float omega, hipStream_t st)
{
    if (d.leafs.n <= 0) return;
    k::djacobi_update_v<<<geo_vector_grid(d, d.leafs.n), 256, 0, st>>>(view(d), d.leafs.d, d.leafs.n, X, R, D, omega);
}

}  // namespace cmg
