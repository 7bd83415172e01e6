// cmg_assemble.hip -- cut-cell operator assembly on the device: solid and wall
// cell types, ghost cell types, negative-face coefficients (open faces, corner
// FaceFluidRatio or 4x4 supersampled open fraction from an analytic SDF), the
// algebraic coarsening of face coefficients, diagonals and NONLEAF types, and
// the closed-cell fix.  Each kernel restates one step of the host assembly
// (CoarsenTypesAndAMGCoeffs, src/AMGSolver.cu:12-318;
// CreateAMGLaplacianSystemWithSolidCutOnNodeSDF, cirrus_cutcell/MeshCutCell.cu:416-543)
// with the same float operations in the same order, so the device operator is
// the host operator bit for bit.  Sphere and torus SDFs only.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "cmg_device.hpp"
#include "cmg_kernels.hpp"

#pragma clang fp contract(off)

namespace cmg {
namespace k {

struct SdfParams {
    float p[5];
};

__device__ __forceinline__ float sdf_d(int kind, const SdfParams& s, float x, float y, float z)
{
    if (kind == CMG_SDF_SPHERE) {
        const float dx = x - s.p[0], dy = y - s.p[1], dz = z - s.p[2];
        return sqrtf(dx * dx + dy * dy + dz * dz) - s.p[3];
    }
    const float dx = x - s.p[0], dy = y - s.p[1], dz = z - s.p[2];
    const float q = sqrtf(dx * dx + dy * dy) - s.p[3];
    return sqrtf(q * q + dz * dz) - s.p[4];
}

__device__ __forceinline__ void gcell(const int32_t* tcoord, int t, int v, int g[3])
{
    g[0] = tcoord[3 * t] * 8 + (v >> 6);
    g[1] = tcoord[3 * t + 1] * 8 + ((v >> 3) & 7);
    g[2] = tcoord[3 * t + 2] * 8 + (v & 7);
}

__device__ __forceinline__ float level_h(const TopoView& tp, int level) { return tp.h0 / float(1 << level); }

// same-level voxel of local (i, j, k), one axis possibly out of the tile; tile -1 if none
__device__ __forceinline__ void same_level(const TopoView& tp, int t, int i, int j, int k, int& nt, int& nv)
{
    int ax = -1, sg = 0;
    if (i < 0) { ax = 0; sg = -1; } else if (i > 7) { ax = 0; sg = 1; }
    else if (j < 0) { ax = 1; sg = -1; } else if (j > 7) { ax = 1; sg = 1; }
    else if (k < 0) { ax = 2; sg = -1; } else if (k > 7) { ax = 2; sg = 1; }
    nt = ax < 0 ? t : tp.nbr[size_t(t) * 6 + (sg < 0 ? ax : ax + 3)];
    nv = cidx(i & 7, j & 7, k & 7);
}

__device__ int corners_inside(int kind, const SdfParams& s, const int g[3], float h)
{
    float mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = float(g[a]) * h;
        mx[a] = float(g[a] + 1) * h;
    }
    int n = 0;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk)
                if (sdf_d(kind, s, di ? mx[0] : mn[0], dj ? mx[1] : mn[1], dk ? mx[2] : mn[2]) < 0) ++n;
    return n;
}

// NegativeLaplacianCoeff (src/AMGSolver.cu:12-24)
__device__ __forceinline__ float face_valid_h(float h, uint8_t tt0, uint8_t tt1, uint8_t ct0, uint8_t ct1)
{
    const bool both_leafs = (tt0 & LEAF) && (tt1 & LEAF);
    const bool leaf_ghost = ((tt0 & LEAF) && (tt1 & GHOST)) || ((tt0 & GHOST) && (tt1 & LEAF));
    const bool has_neumann = (ct0 & NEUMANN) || (ct1 & NEUMANN);
    const bool has_interior = (ct0 & INTERIOR) || (ct1 & INTERIOR);
    return ((both_leafs || leaf_ghost) && !has_neumann && has_interior) ? h : 0.0f;
}

__device__ __forceinline__ float frac_inside(float a, float b)
{
    if (a < 0.0 && b < 0.0) return 0.0f;
    if (a < 0.0 && b >= 0.0) return b / (b - a);
    if (a >= 0.0 && b < 0.0) return a / (a - b);
    return 1.0f;
}

// FaceFluidRatio (cirrus_cutcell/MeshCutCell.cu:305-408), double intermediates
// where the reference mixes in double literals
__device__ float face_fluid_ratio(float p0, float p1, float p2, float p3)
{
    const bool s0 = p0 < 0, s1 = p1 < 0, s2 = p2 < 0, s3 = p3 < 0;
    auto half_prod = [](float a, float b) { return float(0.5 * double(a) * double(b)); };
    auto half_sum = [](float a, float b) { return float(0.5 * (double(a) + double(b))); };
    auto one_minus = [](float a) { return float(1.0 - double(a)); };
    auto two_tri = [](float a, float b, float c, float d) {
        return float(0.5 * double(a) * double(b) + 0.5 * double(c) * double(d));
    };
    auto corner_out = [&](float a, float b) { return float(1.0 - 0.5 * double(one_minus(a)) * double(one_minus(b))); };
    float r;
    const int key = (int(s0) << 3) | (int(s1) << 2) | (int(s2) << 1) | int(s3);
    switch (key) {
    case 0b1111: r = 0; break;
    case 0b1110: r = half_prod(frac_inside(p3, p2), frac_inside(p3, p0)); break;
    case 0b1101: r = half_prod(frac_inside(p2, p1), frac_inside(p2, p3)); break;
    case 0b1100: r = half_sum(frac_inside(p2, p1), frac_inside(p3, p0)); break;
    case 0b1011: r = half_prod(frac_inside(p1, p0), frac_inside(p1, p2)); break;
    case 0b1010: r = two_tri(frac_inside(p1, p0), frac_inside(p1, p2), frac_inside(p3, p2), frac_inside(p3, p0)); break;
    case 0b1001: r = half_sum(frac_inside(p1, p0), frac_inside(p2, p3)); break;
    case 0b1000: r = corner_out(frac_inside(p0, p1), frac_inside(p0, p3)); break;
    case 0b0111: r = half_prod(frac_inside(p0, p1), frac_inside(p0, p3)); break;
    case 0b0110: r = half_sum(frac_inside(p0, p1), frac_inside(p3, p2)); break;
    case 0b0101: r = two_tri(frac_inside(p0, p1), frac_inside(p0, p3), frac_inside(p2, p1), frac_inside(p2, p3)); break;
    case 0b0100: r = corner_out(frac_inside(p1, p0), frac_inside(p1, p2)); break;
    case 0b0011: r = half_sum(frac_inside(p0, p3), frac_inside(p1, p2)); break;
    case 0b0010: r = corner_out(frac_inside(p2, p1), frac_inside(p2, p3)); break;
    case 0b0001: r = corner_out(frac_inside(p3, p0), frac_inside(p3, p2)); break;
    default: r = 1; break;
    }
    if (r < 0.1 && r != 0) r = 0.1f;
    return r;
}

__device__ float supersampled_ratio(int kind, const SdfParams& s, const int g[3], int axis, float h)
{
    const int b = (axis + 1) % 3, c = (axis + 2) % 3;
    float ctr[3];
    ctr[axis] = float(g[axis]) * h;
    ctr[b] = (float(g[b]) + 0.5f) * h;
    ctr[c] = (float(g[c]) + 0.5f) * h;
    if (sdf_d(kind, s, ctr[0], ctr[1], ctr[2]) > 0.6f * h) return 1.0f;
    int open = 0;
    for (int si = 0; si < 4; ++si)
        for (int q = 0; q < 4; ++q) {
            float p[3];
            p[axis] = float(g[axis]) * h;
            p[b] = (float(g[b]) + (0.125f + 0.25f * float(si))) * h;
            p[c] = (float(g[c]) + (0.125f + 0.25f * float(q))) * h;
            if (sdf_d(kind, s, p[0], p[1], p[2]) >= 0.0f) ++open;
        }
    float r = float(open) / 16.0f;
    if (r < 0.1 && r != 0) r = 0.1f;
    return r;
}

__device__ float corner_rule_ratio(int kind, const SdfParams& s, const int g[3], int axis, float h)
{
    const float kO[3][4][3] = {
        {{-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, -0.5f}},
        {{-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, -0.5f}},
        {{-0.5f, -0.5f, -0.5f}, {-0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}}};
    float cc[3];
    for (int d = 0; d < 3; ++d) cc[d] = float(g[d] + 0.5) * h;
    float phi[4];
    for (int q = 0; q < 4; ++q) {
        float p[3];
        for (int d = 0; d < 3; ++d) p[d] = cc[d] + kO[axis][q][d] * h;
        phi[q] = sdf_d(kind, s, p[0], p[1], p[2]);
    }
    return face_fluid_ratio(phi[0], phi[1], phi[2], phi[3]);
}

// one workgroup of 256 threads per tile of `list`, 2 cells per thread
#define ASM_TILE_CELLS(list)                          \
    const int t = (list)[blockIdx.x];                 \
    if (t < 0) return;                                \
    for (int v = threadIdx.x; v < 512; v += 256)

// cells whose 8 corners are inside the SDF become NEUMANN (LEAF tiles)
__global__ __launch_bounds__(256) void asm_solid(TopoView tp, const int32_t* __restrict__ list, const int32_t* __restrict__ tcoord,
                                                 int kind, SdfParams s, uint8_t* __restrict__ ctype)
{
    ASM_TILE_CELLS(list)
    {
        if (!(tp.ttype[t] & LEAF)) return;
        const float h = level_h(tp, tp.tlevel[t]);
        int g[3];
        gcell(tcoord, t, v, g);
        const float c = sdf_d(kind, s, (float(g[0]) + 0.5f) * h, (float(g[1]) + 0.5f) * h, (float(g[2]) + 0.5f) * h);
        if (c > 0.9f * h) continue;
        if (corners_inside(kind, s, g, h) == 8) ctype[size_t(t) * 512 + v] = NEUMANN;
    }
}

// INTERIOR cells of LEAF tiles next to the domain boundary at level min(chk, l):
// DIRICHLET on the y+ side when ytop, NEUMANN otherwise (FluidEuler.cu:1325-1391)
__global__ __launch_bounds__(256) void asm_walls(TopoView tp, const int32_t* __restrict__ list, const int32_t* __restrict__ tcoord,
                                                 int chk, int ytop, uint8_t* __restrict__ ctype)
{
    ASM_TILE_CELLS(list)
    {
        if (!(tp.ttype[t] & LEAF)) return;
        uint8_t& ty = ctype[size_t(t) * 512 + v];
        if (ty != INTERIOR) continue;
        const int l = tp.tlevel[t];
        const int cl = min(chk, l), diff = l - cl;
        int A = t;   // ancestor tile at level cl
        for (int s = 0; s < diff && A >= 0; ++s) A = tp.parent[A];
        if (A < 0) continue;
        int g[3];
        gcell(tcoord, t, v, g);
        for (int a = 0; a < 3; ++a) g[a] >>= diff;
        bool found = false;
        int baxis = -1, boff = 0;
        for (int a = 0; a < 3 && !found; ++a)
            for (int off = -1; off <= 1 && !found; off += 2) {
                const int n = g[a] + off;
                // the neighbour cell lies in A or in A's face neighbour along (a, off)
                const bool inside = (n >> 3) == tcoord[3 * A + a];
                const bool exists = inside || tp.nbr[size_t(A) * 6 + (off < 0 ? a : a + 3)] >= 0;
                if (!exists) {
                    found = true;
                    baxis = a;
                    boff = off;
                }
            }
        if (found) ty = (ytop && baxis == 1 && boff == 1) ? DIRICHLET : NEUMANN;
    }
}

// GHOST cell types from the parent cell (AMGSolver.cu:266-276)
__global__ __launch_bounds__(256) void asm_ghost_types(TopoView tp, const int32_t* __restrict__ list, uint8_t* __restrict__ ctype)
{
    ASM_TILE_CELLS(list)
    {
        if (!(tp.ttype[t] & GHOST)) return;
        const int p = tp.parent[t];
        const int oc = tp.toct[t];
        const int i = v >> 6, j = (v >> 3) & 7, kk = v & 7;
        const int pv = cidx(4 * (oc & 1) + (i >> 1), 4 * ((oc >> 1) & 1) + (j >> 1), 4 * ((oc >> 2) & 1) + (kk >> 1));
        ctype[size_t(t) * 512 + v] = p >= 0 ? ctype[size_t(p) * 512 + pv] : uint8_t(DIRICHLET);
    }
}

// negative-face coefficients of LEAF/GHOST cells; NONLEAF faces start at 0
__global__ __launch_bounds__(256) void asm_faces(TopoView tp, const int32_t* __restrict__ list, const int32_t* __restrict__ tcoord,
                                                 int face_rule, int kind, SdfParams s, const uint8_t* __restrict__ ctype,
                                                 float* __restrict__ C0, float* __restrict__ C1, float* __restrict__ C2)
{
    ASM_TILE_CELLS(list)
    {
        const size_t idx = size_t(t) * 512 + v;
        const uint8_t tt0 = tp.ttype[t];
        float* C[3] = {C0, C1, C2};
        if (tt0 == NONLEAF) {
            C0[idx] = 0;
            C1[idx] = 0;
            C2[idx] = 0;
            continue;
        }
        const float h = level_h(tp, tp.tlevel[t]);
        const uint8_t ct0 = ctype[idx];
        const int i = v >> 6, j = (v >> 3) & 7, kk = v & 7;
        for (int a = 0; a < 3; ++a) {
            int nt, nv;
            same_level(tp, t, i - (a == 0), j - (a == 1), kk - (a == 2), nt, nv);
            uint8_t tt1, ct1;
            if (nt < 0) { tt1 = tt0; ct1 = DIRICHLET; }
            else { tt1 = tp.ttype[nt]; ct1 = ctype[size_t(nt) * 512 + nv]; }
            float coeff = face_valid_h(h, tt0, tt1, ct0, ct1);
            if (coeff != 0.0f && face_rule >= 0) {
                int g[3];
                gcell(tcoord, t, v, g);
                coeff = h * (face_rule == 0 ? corner_rule_ratio(kind, s, g, a, h) : supersampled_ratio(kind, s, g, a, h));
            }
            C[a][idx] = -coeff;
        }
    }
}

// CoarsenOffDiagCoefficientsOneStepKernel (src/AMGSolver.cu:27-78): the tiles of
// one level add R times the sum of their 4 fine faces into the parent's face
__global__ __launch_bounds__(64) void asm_coarsen(TopoView tp, const int32_t* __restrict__ list, float R,
                                                  float* __restrict__ C0, float* __restrict__ C1, float* __restrict__ C2)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const int p = tp.parent[t];
    if (p < 0) return;
    const int oc = tp.toct[t];
    const int q = threadIdx.x;
    const int fl[3] = {q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2};
    const int pv = cidx(4 * (oc & 1) + (fl[0] >> 1), 4 * ((oc >> 1) & 1) + (fl[1] >> 1), 4 * ((oc >> 2) & 1) + (fl[2] >> 1));
    float* C[3] = {C0, C1, C2};
    for (int a = 0; a < 3; ++a) {
        float sum = 0;
        for (int jj = 0; jj < 2; ++jj)
            for (int kk = 0; kk < 2; ++kk) {
                int f1[3] = {fl[0], fl[1], fl[2]};
                f1[(a + 1) % 3] += jj;
                f1[(a + 2) % 3] += kk;
                sum += C[a][size_t(t) * 512 + cidx(f1[0], f1[1], f1[2])];
            }
        C[a][size_t(p) * 512 + pv] += sum * R;
    }
}

// PrepareLaplacianSystemFromLeafAndGhostCellTypesAndFaceCoeffs (src/AMGSolver.cu:89-254):
// diagonals of one level; NONLEAF diagonals and types from the children
__global__ __launch_bounds__(256) void asm_diag(TopoView tp, const int32_t* __restrict__ list, float R, uint8_t* __restrict__ ctype,
                                                const float* __restrict__ C0, const float* __restrict__ C1,
                                                const float* __restrict__ C2, float* __restrict__ C3)
{
    ASM_TILE_CELLS(list)
    {
        const uint8_t tt = tp.ttype[t];
        const float h = level_h(tp, tp.tlevel[t]);
        const size_t idx = size_t(t) * 512 + v;
        const int i = v >> 6, j = (v >> 3) & 7, kk = v & 7;
        const float* C[3] = {C0, C1, C2};
        if (tt & (LEAF | GHOST)) {
            float diag = 0;
            if (ctype[idx] & INTERIOR) {
                for (int a = 0; a < 3; ++a)
                    for (int s = -1; s <= 1; s += 2) {
                        const float c0v = C[a][idx];
                        int nt, nv;
                        same_level(tp, t, i + (a == 0) * s, j + (a == 1) * s, kk + (a == 2) * s, nt, nv);
                        const float c1v = nt < 0 ? -h : C[a][size_t(nt) * 512 + nv];
                        diag -= (s == -1 ? c0v : c1v);
                    }
            }
            C3[idx] = diag;
        } else {
            int interior_cnt = 0;
            uint8_t cty[2][2][2];
            float cof[3][2][2][2];
            float diag_sum = 0;
            const int oct = (i >> 2) | ((j >> 2) << 1) | ((kk >> 2) << 2);
            const int ct = tp.child[size_t(t) * 8 + oct];
            for (int ci = 0; ci < 2; ++ci)
                for (int cj = 0; cj < 2; ++cj)
                    for (int ck = 0; ck < 2; ++ck) {
                        if (ct >= 0) {
                            const size_t cidx2 = size_t(ct) * 512 + cidx(2 * (i & 3) + ci, 2 * (j & 3) + cj, 2 * (kk & 3) + ck);
                            diag_sum += C3[cidx2];
                            interior_cnt += ctype[cidx2] == INTERIOR;
                            cty[ci][cj][ck] = ctype[cidx2];
                            for (int a = 0; a < 3; ++a) cof[a][ci][cj][ck] = C[a][cidx2];
                        } else {
                            cty[ci][cj][ck] = DIRICHLET;
                            for (int a = 0; a < 3; ++a) cof[a][ci][cj][ck] = 0;
                        }
                    }
            for (int ci = 0; ci < 2; ++ci)
                for (int cj = 0; cj < 2; ++cj)
                    for (int ck = 0; ck < 2; ++ck) {
                        if (cty[0][cj][ck] == INTERIOR && cty[1][cj][ck] == INTERIOR) diag_sum += cof[0][1][cj][ck];
                        if (cty[ci][0][ck] == INTERIOR && cty[ci][1][ck] == INTERIOR) diag_sum += cof[1][ci][1][ck];
                        if (cty[ci][cj][0] == INTERIOR && cty[ci][cj][1] == INTERIOR) diag_sum += cof[2][ci][cj][1];
                    }
            ctype[idx] = interior_cnt > 0 ? INTERIOR : DIRICHLET;
            C3[idx] = diag_sum * R;
        }
    }
}

// closed cells: a LEAF INTERIOR cell with a zero diagonal becomes NEUMANN; count[0] += 1
__global__ __launch_bounds__(256) void asm_closed(TopoView tp, const int32_t* __restrict__ list, const float* __restrict__ C3,
                                                  uint8_t* __restrict__ ctype, unsigned int* __restrict__ count)
{
    ASM_TILE_CELLS(list)
    {
        if (!(tp.ttype[t] & LEAF)) return;
        const size_t idx = size_t(t) * 512 + v;
        if (ctype[idx] == INTERIOR && C3[idx] == 0.0f) {
            ctype[idx] = NEUMANN;
            atomicAdd(count, 1u);
        }
    }
}

// NONLEAF types from their children (after the closed-cell fix)
__global__ __launch_bounds__(256) void asm_nonleaf_types(TopoView tp, const int32_t* __restrict__ list, uint8_t* __restrict__ ctype)
{
    ASM_TILE_CELLS(list)
    {
        if (tp.ttype[t] & (LEAF | GHOST)) return;
        const int i = v >> 6, j = (v >> 3) & 7, kk = v & 7;
        const int oct = (i >> 2) | ((j >> 2) << 1) | ((kk >> 2) << 2);
        const int ct = tp.child[size_t(t) * 8 + oct];
        int cnt = 0;
        if (ct >= 0)
            for (int ci = 0; ci < 2; ++ci)
                for (int cj = 0; cj < 2; ++cj)
                    for (int ck = 0; ck < 2; ++ck)
                        cnt += ctype[size_t(ct) * 512 + cidx(2 * (i & 3) + ci, 2 * (j & 3) + cj, 2 * (kk & 3) + ck)] == INTERIOR;
        ctype[size_t(t) * 512 + v] = cnt > 0 ? INTERIOR : DIRICHLET;
    }
}

}  // namespace k

// ---------------------------------------------------------------- drivers ----

static void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

static k::SdfParams sdf_params(const float* sdf, int kind)
{
    k::SdfParams s{};
    const int n = kind == CMG_SDF_TORUS ? 5 : 4;
    for (int i = 0; i < n; ++i) s.p[i] = sdf[i];
    return s;
}

void launch_assemble_solid(const DevGrid& d, int kind, const float* sdf, hipStream_t st)
{
    const LaunchList& l = d.all;
    if (l.n <= 0) return;
    k::asm_solid<<<l.n, 256, 0, st>>>(view(d), l.d, d.tcoord, kind, sdf_params(sdf, kind), d.ctype);
}

void launch_assemble_walls(const DevGrid& d, int chk, int ytop, hipStream_t st)
{
    const LaunchList& l = d.all;
    if (l.n <= 0) return;
    k::asm_walls<<<l.n, 256, 0, st>>>(view(d), l.d, d.tcoord, chk, ytop, d.ctype);
}

int launch_assemble_operator(const DevGrid& d, int c0, float R, int face_rule, int kind, const float* sdf, unsigned int* count,
                             hipStream_t st)
{
    const TopoView tp = view(d);
    const k::SdfParams s = sdf ? sdf_params(sdf, kind) : k::SdfParams{};
    float *C0 = dev_channel(const_cast<DevGrid&>(d), c0), *C1 = dev_channel(const_cast<DevGrid&>(d), c0 + 1),
          *C2 = dev_channel(const_cast<DevGrid&>(d), c0 + 2), *C3 = dev_channel(const_cast<DevGrid&>(d), c0 + 3);
    const LaunchList& all = d.all;
    k::asm_ghost_types<<<all.n, 256, 0, st>>>(tp, all.d, d.ctype);
    k::asm_faces<<<all.n, 256, 0, st>>>(tp, all.d, d.tcoord, face_rule, kind, s, d.ctype, C0, C1, C2);
    for (int l = d.max_level; l > 0; --l) {
        const LaunchList& ll = d.prolong[l];
        if (ll.n > 0) k::asm_coarsen<<<ll.n, 64, 0, st>>>(tp, ll.d, R, C0, C1, C2);
    }
    for (int l = d.max_level; l >= 0; --l) {
        const LaunchList& ll = d.prolong[l];
        if (ll.n > 0) k::asm_diag<<<ll.n, 256, 0, st>>>(tp, ll.d, R, d.ctype, C0, C1, C2, C3);
    }
    hip_ok(hipMemsetAsync(count, 0, sizeof(unsigned int), st), "assembly counter");
    k::asm_closed<<<all.n, 256, 0, st>>>(tp, all.d, C3, d.ctype, count);
    unsigned int h = 0;
    hip_ok(hipMemcpyAsync(&h, count, sizeof(h), hipMemcpyDeviceToHost, st), "assembly counter");
    hip_ok(hipStreamSynchronize(st), "assembly");
    if (h) {
        k::asm_ghost_types<<<all.n, 256, 0, st>>>(tp, all.d, d.ctype);
        for (int l = d.max_level; l >= 0; --l) {
            const LaunchList& ll = d.prolong[l];
            if (ll.n > 0) k::asm_nonleaf_types<<<ll.n, 256, 0, st>>>(tp, ll.d, d.ctype);
        }
    }
    return int(h);
}

}  // namespace cmg
