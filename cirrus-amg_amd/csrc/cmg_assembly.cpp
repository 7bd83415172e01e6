// cmg_assembly.cpp -- cell typing and AMG operator assembly on the host (C++17).
//
// Produces the reference operator bit for bit:
//   face coefficients of LEAF/GHOST cells   src/AMGSolver.cu:278-315, SolverTests.cu:1627-1709
//   algebraically consistent coarsening    src/AMGSolver.cu:27-78, :89-110
//   diagonal + NONLEAF types               src/AMGSolver.cu:112-254
// Neighbour / parent / child voxels are resolved through the finalized tables
// instead of per-voxel hash probes.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "cmg_host.hpp"

namespace cmg {

std::vector<float>& HostCells::channel(int c, size_t n)
{
    auto& v = ch.at(c);
    if (v.size() != n) v.assign(n, 0.0f);
    return v;
}

namespace {

struct Vox {
    int t;
    int v;
};

// same-level voxel at local (i,j,k) possibly one cell outside the tile
inline Vox same_level(const Topology& topo, int t, int i, int j, int k)
{
    int ax = -1, sg = 0;
    if (i < 0) { ax = 0; sg = -1; } else if (i > 7) { ax = 0; sg = 1; }
    else if (j < 0) { ax = 1; sg = -1; } else if (j > 7) { ax = 1; sg = 1; }
    else if (k < 0) { ax = 2; sg = -1; } else if (k > 7) { ax = 2; sg = 1; }
    if (ax < 0) return {t, cell_index(i, j, k)};
    int nt = topo.nbr[size_t(t) * 6 + (sg < 0 ? ax : ax + 3)];
    return {nt, cell_index(i & 7, j & 7, k & 7)};
}

inline Vox parent_voxel(const Topology& topo, int t, int v)
{
    const Coord& c = topo.t_coord[t];
    int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
    return {topo.parent[t], cell_index(4 * (c[0] & 1) + (i >> 1), 4 * (c[1] & 1) + (j >> 1), 4 * (c[2] & 1) + (k >> 1))};
}

inline Vox child_voxel(const Topology& topo, int t, int v, int di, int dj, int dk)
{
    int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
    int oct = (i >> 2) | ((j >> 2) << 1) | ((k >> 2) << 2);
    return {topo.child[size_t(t) * 8 + oct], cell_index(2 * (i & 3) + di, 2 * (j & 3) + dj, 2 * (k & 3) + dk)};
}

inline void global_cell(const Topology& topo, int t, int v, int g[3])
{
    const Coord& c = topo.t_coord[t];
    g[0] = c[0] * 8 + (v >> 6);
    g[1] = c[1] * 8 + ((v >> 3) & 7);
    g[2] = c[2] * 8 + (v & 7);
}

int corners_inside_cell(int kind, const float* prm, const int g[3], float h)
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
                if (sdf_eval(kind, prm, di ? mx[0] : mn[0], dj ? mx[1] : mn[1], dk ? mx[2] : mn[2]) < 0) ++n;
    return n;
}

// NegativeLaplacianCoeff (src/AMGSolver.cu:12-24)
inline float face_valid_h(float h, uint8_t tt0, uint8_t tt1, uint8_t ct0, uint8_t ct1)
{
    bool both_leafs = (tt0 & LEAF) && (tt1 & LEAF);
    bool leaf_ghost = ((tt0 & LEAF) && (tt1 & GHOST)) || ((tt0 & GHOST) && (tt1 & LEAF));
    bool has_neumann = (ct0 & NEUMANN) || (ct1 & NEUMANN);
    bool has_interior = (ct0 & INTERIOR) || (ct1 & INTERIOR);
    return ((both_leafs || leaf_ghost) && !has_neumann && has_interior) ? h : 0.0f;
}

inline float frac_inside(float a, float b)
{
    if (a < 0.0 && b < 0.0) return 0.0f;
    if (a < 0.0 && b >= 0.0) return b / (b - a);
    if (a >= 0.0 && b < 0.0) return a / (a - b);
    return 1.0f;
}

float supersampled_ratio(int kind, const float* prm, const int g[3], int axis, float h)
{
    const int b = (axis + 1) % 3, c = (axis + 2) % 3;
    // the face is a square of side h; every sample lies within 0.54h of its centre
    float ctr[3];
    ctr[axis] = float(g[axis]) * h;
    ctr[b] = (float(g[b]) + 0.5f) * h;
    ctr[c] = (float(g[c]) + 0.5f) * h;
    if (sdf_eval(kind, prm, ctr[0], ctr[1], ctr[2]) > 0.6f * h) return 1.0f;
    int open = 0;
    for (int s = 0; s < 4; ++s)
        for (int q = 0; q < 4; ++q) {
            float p[3];
            p[axis] = float(g[axis]) * h;
            p[b] = (float(g[b]) + (0.125f + 0.25f * float(s))) * h;
            p[c] = (float(g[c]) + (0.125f + 0.25f * float(q))) * h;
            if (sdf_eval(kind, prm, p[0], p[1], p[2]) >= 0.0f) ++open;
        }
    float r = float(open) / 16.0f;
    if (r < 0.1 && r != 0) r = 0.1f;
    return r;
}

float corner_rule_ratio(int kind, const float* prm, const int g[3], int axis, float h)
{
    static const float kO[3][4][3] = {
        {{-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f}, {-0.5f, 0.5f, -0.5f}},
        {{-0.5f, -0.5f, -0.5f}, {-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, -0.5f}},
        {{-0.5f, -0.5f, -0.5f}, {-0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}}};
    float cc[3];
    for (int d = 0; d < 3; ++d) cc[d] = float(g[d] + 0.5) * h;
    float phi[4];
    for (int q = 0; q < 4; ++q) {
        float p[3];
        for (int d = 0; d < 3; ++d) p[d] = cc[d] + kO[axis][q][d] * h;
        phi[q] = sdf_eval(kind, prm, p[0], p[1], p[2]);
    }
    return face_fluid_ratio(phi[0], phi[1], phi[2], phi[3]);
}

}  // namespace

// FaceFluidRatio (cirrus_cutcell/MeshCutCell.cu:305-408), double intermediates
// where the reference mixes double literals in
float face_fluid_ratio(float p0, float p1, float p2, float p3)
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
    const int key = (s0 << 3) | (s1 << 2) | (s2 << 1) | int(s3);
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

void set_leaf_types_preset(const Topology& topo, HostCells& cells, const cmg_target& t)
{
    static const float kTestSphere[4] = {0.5f, 0.5f, 0.5f, 0.25f};
    const int N = topo.num_tiles();
    cells.type.assign(size_t(N) * kCells, 0);
    const bool solid = t.kind == CMG_TARGET_SPHERE_SOLID || t.kind == CMG_TARGET_STAR_SOLID;
    const int kind = t.kind == CMG_TARGET_SPHERE_SOLID ? CMG_SDF_SPHERE : CMG_SDF_STAR;
#pragma omp parallel for schedule(dynamic, 16)
    for (int tt = 0; tt < N; ++tt) {
        if (!(topo.t_type[tt] & LEAF)) continue;
        const float h = topo.h(topo.t_level[tt]);
        for (int v = 0; v < kCells; ++v) {
            uint8_t ty = INTERIOR;
            if (solid) {
                int g[3];
                global_cell(topo, tt, v, g);
                ty = corners_inside_cell(kind, kTestSphere, g, h) == 8 ? NEUMANN : INTERIOR;
            }
            cells.type[size_t(tt) * kCells + v] = ty;
        }
    }
}

void set_wall_types(const Topology& topo, HostCells& cells, int chk_level, bool dirichlet_ytop)
{
    const int N = topo.num_tiles();
#pragma omp parallel for schedule(dynamic, 16)
    for (int tt = 0; tt < N; ++tt) {
        if (!(topo.t_type[tt] & LEAF)) continue;
        const int l = topo.t_level[tt];
        const int cl = std::min(chk_level, l);
        const int diff = l - cl;
        for (int v = 0; v < kCells; ++v) {
            uint8_t& ty = cells.type[size_t(tt) * kCells + v];
            if (ty != INTERIOR) continue;
            int g[3];
            global_cell(topo, tt, v, g);
            for (int a = 0; a < 3; ++a) g[a] >>= diff;
            bool found = false;
            int baxis = -1, boff = 0;
            for (int a = 0; a < 3 && !found; ++a)
                for (int off = -1; off <= 1 && !found; off += 2) {
                    int n[3] = {g[0], g[1], g[2]};
                    n[a] += off;
                    if (!topo.find(cl, Coord{n[0] >> 3, n[1] >> 3, n[2] >> 3})) {
                        found = true;
                        baxis = a;
                        boff = off;
                    }
                }
            if (found) ty = (dirichlet_ytop && baxis == 1 && boff == 1) ? DIRICHLET : NEUMANN;
        }
    }
}

void set_solid_types(const Topology& topo, HostCells& cells, int sdf_kind, const float* sdf)
{
    const int N = topo.num_tiles();
#pragma omp parallel for schedule(dynamic, 16)
    for (int tt = 0; tt < N; ++tt) {
        if (!(topo.t_type[tt] & LEAF)) continue;
        const float h = topo.h(topo.t_level[tt]);
        for (int v = 0; v < kCells; ++v) {
            int g[3];
            global_cell(topo, tt, v, g);
            // far outside the solid: no corner can be inside (|grad sdf| <= 1)
            float c = sdf_eval(sdf_kind, sdf, (float(g[0]) + 0.5f) * h, (float(g[1]) + 0.5f) * h, (float(g[2]) + 0.5f) * h);
            if (c > 0.9f * h) continue;
            if (corners_inside_cell(sdf_kind, sdf, g, h) == 8) cells.type[size_t(tt) * kCells + v] = NEUMANN;
        }
    }
}

void assemble(const Topology& topo, HostCells& cells, int c0, float R, int face_rule, int sdf_kind, const float* sdf)
{
    if (c0 < 0 || c0 + 3 >= kChannels) throw std::invalid_argument("coefficient channel out of range");
    const int N = topo.num_tiles();
    const size_t NC = size_t(N) * kCells;
    if (cells.type.size() != NC) throw std::invalid_argument("cell types not set");
    float* C[4];
    for (int a = 0; a < 4; ++a) C[a] = cells.channel(c0 + a, NC).data();
    uint8_t* T = cells.type.data();

    // GHOST cell types from the parent cell (AMGSolver.cu:266-276)
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & GHOST)) continue;
        for (int v = 0; v < kCells; ++v) {
            Vox p = parent_voxel(topo, t, v);
            T[size_t(t) * kCells + v] = p.t >= 0 ? T[size_t(p.t) * kCells + p.v] : DIRICHLET;
        }
    }
    // negative-face coefficients of LEAF/GHOST cells; NONLEAF faces start at 0
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < N; ++t) {
        const uint8_t tt0 = topo.t_type[t];
        const float h = topo.h(topo.t_level[t]);
        for (int v = 0; v < kCells; ++v) {
            const size_t idx = size_t(t) * kCells + v;
            if (tt0 == NONLEAF) {
                for (int a = 0; a < 3; ++a) C[a][idx] = 0;
                continue;
            }
            const uint8_t ct0 = T[idx];
            const int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
            for (int a = 0; a < 3; ++a) {
                Vox n = same_level(topo, t, i - (a == 0), j - (a == 1), k - (a == 2));
                uint8_t tt1, ct1;
                if (n.t < 0) { tt1 = tt0; ct1 = DIRICHLET; }
                else { tt1 = topo.t_type[n.t]; ct1 = T[size_t(n.t) * kCells + n.v]; }
                float coeff = face_valid_h(h, tt0, tt1, ct0, ct1);
                if (coeff != 0.0f && face_rule >= 0) {
                    int g[3];
                    global_cell(topo, t, v, g);
                    coeff = h * (face_rule == 0 ? corner_rule_ratio(sdf_kind, sdf, g, a, h)
                                                : supersampled_ratio(sdf_kind, sdf, g, a, h));
                }
                C[a][idx] = -coeff;
            }
        }
    }
    // (a) coarsen negative-face coefficients level by level, finest first
    for (int lvl = topo.max_level(); lvl > 0; --lvl) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int t = topo.level_start[lvl]; t < topo.level_start[lvl + 1]; ++t) {
            const int p = topo.parent[t];
            if (p < 0) continue;
            const Coord& c = topo.t_coord[t];
            for (int q = 0; q < 64; ++q) {
                const int fl[3] = {q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2};
                const int pv = cell_index(4 * (c[0] & 1) + (fl[0] >> 1), 4 * (c[1] & 1) + (fl[1] >> 1), 4 * (c[2] & 1) + (fl[2] >> 1));
                for (int a = 0; a < 3; ++a) {
                    float sum = 0;
                    for (int jj = 0; jj < 2; ++jj)
                        for (int kk = 0; kk < 2; ++kk) {
                            int f1[3] = {fl[0], fl[1], fl[2]};
                            f1[(a + 1) % 3] += jj;
                            f1[(a + 2) % 3] += kk;
                            sum += C[a][size_t(t) * kCells + cell_index(f1[0], f1[1], f1[2])];
                        }
                    C[a][size_t(p) * kCells + pv] += sum * R;
                }
            }
        }
    }
    // (b) diagonals (and NONLEAF types), finest level first
    for (int lvl = topo.num_levels() - 1; lvl >= 0; --lvl) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int t = topo.level_start[lvl]; t < topo.level_start[lvl + 1]; ++t) {
            const uint8_t tt = topo.t_type[t];
            const float h = topo.h(lvl);
            for (int v = 0; v < kCells; ++v) {
                const size_t idx = size_t(t) * kCells + v;
                const int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
                if (tt & (LEAF | GHOST)) {
                    float diag = 0;
                    if (T[idx] & INTERIOR) {
                        for (int a = 0; a < 3; ++a)
                            for (int s = -1; s <= 1; s += 2) {
                                float c0v = C[a][idx];
                                Vox n = same_level(topo, t, i + (a == 0) * s, j + (a == 1) * s, k + (a == 2) * s);
                                float c1v = n.t < 0 ? -h : C[a][size_t(n.t) * kCells + n.v];
                                diag -= (s == -1 ? c0v : c1v);
                            }
                    }
                    C[3][idx] = diag;
                } else {
                    int interior_cnt = 0;
                    uint8_t cty[2][2][2];
                    float cof[3][2][2][2];
                    float diag_sum = 0;
                    for (int ci = 0; ci < 2; ++ci)
                        for (int cj = 0; cj < 2; ++cj)
                            for (int ck = 0; ck < 2; ++ck) {
                                Vox cv = child_voxel(topo, t, v, ci, cj, ck);
                                if (cv.t >= 0) {
                                    const size_t cidx = size_t(cv.t) * kCells + cv.v;
                                    diag_sum += C[3][cidx];
                                    interior_cnt += T[cidx] == INTERIOR;
                                    cty[ci][cj][ck] = T[cidx];
                                    for (int a = 0; a < 3; ++a) cof[a][ci][cj][ck] = C[a][cidx];
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
                    T[idx] = interior_cnt > 0 ? INTERIOR : DIRICHLET;
                    C[3][idx] = diag_sum * R;
                }
            }
        }
    }
    // (c) closed cells: a LEAF INTERIOR cell with all six faces closed has a zero
    // diagonal (the reference asserts diag > 0, AMGSolver.cu:151); it is
    // disconnected from the fluid and becomes NEUMANN.  Ghost types follow their
    // parents and NONLEAF types are re-derived; coefficients do not change.
    int closed = 0;
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & LEAF)) continue;
        for (int v = 0; v < kCells; ++v) {
            const size_t idx = size_t(t) * kCells + v;
            if (T[idx] == INTERIOR && C[3][idx] == 0.0f) { T[idx] = NEUMANN; ++closed; }
        }
    }
    if (!closed) return;
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & GHOST)) continue;
        for (int v = 0; v < kCells; ++v) {
            Vox p = parent_voxel(topo, t, v);
            T[size_t(t) * kCells + v] = p.t >= 0 ? T[size_t(p.t) * kCells + p.v] : DIRICHLET;
        }
    }
    for (int lvl = topo.num_levels() - 1; lvl >= 0; --lvl)
        for (int t = topo.level_start[lvl]; t < topo.level_start[lvl + 1]; ++t) {
            if (topo.t_type[t] & (LEAF | GHOST)) continue;
            for (int v = 0; v < kCells; ++v) {
                int cnt = 0;
                for (int ci = 0; ci < 2; ++ci)
                    for (int cj = 0; cj < 2; ++cj)
                        for (int ck = 0; ck < 2; ++ck) {
                            Vox cv = child_voxel(topo, t, v, ci, cj, ck);
                            if (cv.t >= 0 && T[size_t(cv.t) * kCells + cv.v] == INTERIOR) ++cnt;
                        }
                T[size_t(t) * kCells + v] = cnt > 0 ? INTERIOR : DIRICHLET;
            }
        }
}

// ---------------------------------------------------------- synthetic inputs ---

namespace {
uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float uniform_pm1(uint64_t& s) { return float(uint32_t(splitmix64(s) >> 40)) * (1.0f / 8388608.0f) - 1.0f; }
float normal01(uint64_t& s)
{
    const uint32_t a = uint32_t(splitmix64(s) >> 40), b = uint32_t(splitmix64(s) >> 40);
    const double u1 = (double(a) + 1.0) / 16777216.0, u2 = double(b) / 16777216.0;
    return float(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
}
}  // namespace

void rhs_uniform(const Topology& topo, HostCells& cells, int ch, uint64_t seed, bool zero_mean)
{
    const int N = topo.num_tiles();
    auto& B = cells.channel(ch, size_t(N) * kCells);
    std::fill(B.begin(), B.end(), 0.0f);
    uint64_t st = seed;
    double sum = 0, cnt = 0;
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & LEAF)) continue;
        for (int v = 0; v < kCells; ++v) {
            const size_t idx = size_t(t) * kCells + v;
            if (!(cells.type[idx] & INTERIOR)) continue;
            B[idx] = uniform_pm1(st);
            sum += B[idx];
            cnt += 1;
        }
    }
    if (!zero_mean || cnt == 0) return;
    const double mean = sum / cnt;
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & LEAF)) continue;
        for (int v = 0; v < kCells; ++v) {
            const size_t idx = size_t(t) * kCells + v;
            if (cells.type[idx] & INTERIOR) B[idx] = float(double(B[idx]) - mean);
        }
    }
}

void rhs_divergence(const Topology& topo, HostCells& cells, int ch, int u0, uint64_t seed)
{
    const int N = topo.num_tiles();
    const size_t NC = size_t(N) * kCells;
    float* U[3];
    for (int a = 0; a < 3; ++a) {
        auto& u = cells.channel(u0 + a, NC);
        std::fill(u.begin(), u.end(), 0.0f);
        U[a] = u.data();
    }
    uint64_t st = seed;
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & LEAF)) continue;
        for (int v = 0; v < kCells; ++v)
            for (int a = 0; a < 3; ++a) U[a][size_t(t) * kCells + v] = normal01(st);
    }
    // ghost faces = parent faces (PropagateToChildren); the parent of a ghost is a LEAF
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & GHOST)) continue;
        for (int v = 0; v < kCells; ++v) {
            Vox p = parent_voxel(topo, t, v);
            for (int a = 0; a < 3; ++a) U[a][size_t(t) * kCells + v] = p.t >= 0 ? U[a][size_t(p.t) * kCells + p.v] : 0.0f;
        }
    }
    // NONLEAF faces = 1/4 of the 4 child faces on the parent face (AccumulateFacesToParentsOneStep)
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & LEAF)) continue;
        const int p = topo.parent[t];
        if (p < 0) continue;
        const Coord& c = topo.t_coord[t];
        for (int q = 0; q < 64; ++q) {
            const int fl[3] = {q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2};
            const int pv = cell_index(4 * (c[0] & 1) + (fl[0] >> 1), 4 * (c[1] & 1) + (fl[1] >> 1), 4 * (c[2] & 1) + (fl[2] >> 1));
            for (int a = 0; a < 3; ++a) {
                float sum = 0;
                for (int jj = 0; jj < 2; ++jj)
                    for (int kk = 0; kk < 2; ++kk) {
                        int f1[3] = {fl[0], fl[1], fl[2]};
                        f1[(a + 1) % 3] += jj;
                        f1[(a + 2) % 3] += kk;
                        sum += U[a][size_t(t) * kCells + cell_index(f1[0], f1[1], f1[2])];
                    }
                U[a][size_t(p) * kCells + pv] = sum * 0.25f;
            }
        }
    }
    auto& B = cells.channel(ch, NC);
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < N; ++t) {
        if (!(topo.t_type[t] & LEAF)) {
            std::fill(B.begin() + size_t(t) * kCells, B.begin() + size_t(t + 1) * kCells, 0.0f);
            continue;
        }
        const float h = topo.h(topo.t_level[t]);
        for (int v = 0; v < kCells; ++v) {
            const size_t idx = size_t(t) * kCells + v;
            const int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
            float sum = 0;
            for (int a = 0; a < 3; ++a)
                for (int sg = -1; sg <= 1; sg += 2) {
                    const float uu0 = U[a][idx];
                    Vox n = same_level(topo, t, i + (a == 0) * sg, j + (a == 1) * sg, k + (a == 2) * sg);
                    const float uu1 = n.t < 0 ? 0.0f : U[a][size_t(n.t) * kCells + n.v];
                    sum += sg < 0 ? (-uu0 * h) * h : (uu1 * h) * h;
                }
            B[idx] = (cells.type[idx] & INTERIOR) ? sum : 0.0f;
        }
    }
}

}  // namespace cmg
