/*
 * og_geo.c -- oracle restatement of the two geometric multigrid solvers of the
 * reference: GMGSolver (src/GMGSolver.cu) and CMGSolver (src/CMGSolver.cu).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Both discretise -lap(x) from cell and tile types alone (no stored
 * coefficients):
 *   GMG  face coefficient (1/dh)(1/h), dh = h between two LEAF cells and
 *        1.5 h between a LEAF and a GHOST cell, 0 otherwise or next to a
 *        NEUMANN cell; the cell sum is scaled by h^3 (GMGSolver.cu:3-67).
 *        The full operator evaluates LEAF and GHOST tiles and adds the GHOST
 *        results into their parents (GMGSolver.cu:167-171).
 *   CMG  face coefficient h (0 next to a NEUMANN cell); a GHOST neighbour is
 *        extrapolated as in the AMG loader (CMGSolver.cu:3-170); prolongation
 *        updates LEAF tiles only (CMGSolver.cu:372-374).
 * Arithmetic contract shared with the HIP kernels:
 *   sum = fmaf(c, x0 - x1, sum) over axis 0..2, side -,+ (diagonal: sum + c)
 *   GMG  result = ((sum * h) * h) * h ; CMG result = sum ; 0 off INTERIOR
 *   GS   x = x + (b - Ax) / D
 *   ghost (CMG) vg = (float)((double)x_c + 0.5 * (double)(V1 - avg8))
 */
#include "og_internal.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
    float x[10][10][10];
    uint8_t tt[10][10][10];
    uint8_t ct[10][10][10];
} geo_lds;

#define GX(s, i, j, k) ((s)->x[(i) + 1][(j) + 1][(k) + 1])
#define GT(s, i, j, k) ((s)->tt[(i) + 1][(j) + 1][(k) + 1])
#define GC(s, i, j, k) ((s)->ct[(i) + 1][(j) + 1][(k) + 1])

static void face_coord(int axis, int sgn, int vi, int* f)
{
    int axk = sgn == -1 ? -1 : 8;
    int i = vi / 8, j = vi % 8;
    if (axis == 0) { f[0] = axk; f[1] = i; f[2] = j; }
    else if (axis == 1) { f[0] = i; f[1] = axk; f[2] = j; }
    else { f[0] = i; f[1] = j; f[2] = axk; }
}

/* LoadGMGLaplacianTileData (GMGSolver.cu:81-133) and LoadCMGLaplacianTileData
 * (CMGSolver.cu:62-170): x is not masked by the cell type */
static void geo_load(const og_grid* g, int mode, int t, int subtree_level, int xch, geo_lds* s)
{
    const float* X = g->ch[xch];
    size_t base = (size_t)t * 512;
    uint8_t own = og_subtree_type(g->t_type[t], g->t_level[t], subtree_level);
    for (int v = 0; v < 512; v++) {
        int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
        GX(s, i, j, k) = X[base + v];
        GT(s, i, j, k) = own;
        GC(s, i, j, k) = g->ctype[base + v];
    }
    for (int a = 0; a < 3; a++)
        for (int sg = -1; sg <= 1; sg += 2)
            for (int fi = 0; fi < 64; fi++) {
                int fl[3], cl[3];
                face_coord(a, sg, fi, fl);
                cl[0] = fl[0]; cl[1] = fl[1]; cl[2] = fl[2];
                cl[a] -= sg;
                int nl[3] = { fl[0], fl[1], fl[2] };
                nl[a] = sg == -1 ? 7 : 0;
                int nt = g->nbr[6 * t + (sg == -1 ? a : a + 3)];
                size_t ni = (size_t)nt * 512 + og_vi(nl[0], nl[1], nl[2]);
                if (nt < 0) {
                    GX(s, fl[0], fl[1], fl[2]) = 0.0f;
                    GT(s, fl[0], fl[1], fl[2]) = own;
                    GC(s, fl[0], fl[1], fl[2]) = OG_DIRICHLET;
                    continue;
                }
                uint8_t nty = og_subtree_type(g->t_type[nt], g->t_level[nt], subtree_level);
                GT(s, fl[0], fl[1], fl[2]) = nty;
                GC(s, fl[0], fl[1], fl[2]) = g->ctype[ni];
                if (mode == OG_GEO_CMG && nty == OG_GHOST) {
                    float avg = 0;
                    for (int di = 0; di < 2; di++)
                        for (int dj = 0; dj < 2; dj++)
                            for (int dk = 0; dk < 2; dk++)
                                avg += GX(s, cl[0] ^ di, cl[1] ^ dj, cl[2] ^ dk);
                    avg /= 8;
                    GX(s, fl[0], fl[1], fl[2]) = (float)((double)GX(s, cl[0], cl[1], cl[2]) + 0.5 * (double)(X[ni] - avg));
                } else {
                    GX(s, fl[0], fl[1], fl[2]) = X[ni];
                }
            }
}

/* NegativeLaplacianCoeff (GMGSolver.cu:3-30 / CMGSolver.cu:3-7) */
static float geo_coeff(int mode, float h, float one_over_h, uint8_t t0, uint8_t t1, uint8_t c0, uint8_t c1)
{
    int neu = (c0 & OG_NEUMANN) || (c1 & OG_NEUMANN);
    if (mode == OG_GEO_CMG) return neu ? 0.0f : h;
    int both = (t0 & OG_LEAF) && (t1 & OG_LEAF);
    int lg = ((t0 & OG_LEAF) && (t1 & OG_GHOST)) || ((t0 & OG_GHOST) && (t1 & OG_LEAF));
    float odh = both ? one_over_h : lg ? (2 * one_over_h / 3) : 0.0f;
    return neu ? 0.0f : odh * one_over_h;
}

/* GMGLaplacianTileData::negativeLap (GMGSolver.cu:45-66) / CMGLaplacianTileData::negativeLap (CMGSolver.cu:23-45) */
static float geo_neg_lap(const geo_lds* s, int mode, float h, int i, int j, int k, int diag)
{
    const float one_over_h = 1 / h;
    uint8_t t0 = GT(s, i, j, k), c0 = GC(s, i, j, k);
    float x0 = GX(s, i, j, k);
    float sum = 0;
    const int l[3] = { i, j, k };
    for (int a = 0; a < 3; a++)
        for (int sg = -1; sg <= 1; sg += 2) {
            int n[3] = { l[0], l[1], l[2] };
            n[a] += sg;
            float c = geo_coeff(mode, h, one_over_h, t0, GT(s, n[0], n[1], n[2]), c0, GC(s, n[0], n[1], n[2]));
            if (diag) sum = sum + c;
            else sum = fmaf(c, x0 - GX(s, n[0], n[1], n[2]), sum);
        }
    if (!(c0 & OG_INTERIOR)) return 0.0f;
    return mode == OG_GEO_GMG ? sum * h * h * h : sum;
}

/* NegativeLaplacianSameLevel128Kernel (GMGSolver.cu:135-161) /
 * ConservativeNegativeLaplacianSameLevel128Kernel (CMGSolver.cu:172-201) on
 * every tile whose subtree type at `subtree_level` (-1: its own type) is in
 * `types` */
void og_geo_neg_lap_tiles(og_grid* g, int mode, int subtree_level, int types, int x, int y, int diag)
{
    #pragma omp parallel
    {
        geo_lds* s = (geo_lds*)malloc(sizeof(geo_lds));
        #pragma omp for schedule(dynamic, 8)
        for (int t = 0; t < g->N; t++) {
            if (!(og_subtree_type(g->t_type[t], g->t_level[t], subtree_level) & types)) continue;
            if (subtree_level >= 0 && g->t_level[t] != subtree_level) continue;
            geo_load(g, mode, t, subtree_level, x, s);
            float h = og_h(g, g->t_level[t]);
            for (int v = 0; v < 512; v++)
                g->ch[y][(size_t)t * 512 + v] = geo_neg_lap(s, mode, h, v >> 6, (v >> 3) & 7, v & 7, diag);
        }
        free(s);
    }
}

/* AccumulateToParentsOneStep(GHOST, coeff, additive, all cell types)
 * (PoissonGrid.cu:408-465) */
static void accumulate_ghosts_add(og_grid* g, int ch, float coeff)
{
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_GHOST)) continue;
        int lvl = g->t_level[t];
        const int32_t* c = &g->t_coord[3 * t];
        for (int q = 0; q < 64; q++) {
            int fl[3] = { q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2 };
            int off, p = og_voxel(g, lvl - 1, (c[0] * 8 + fl[0]) >> 1, (c[1] * 8 + fl[1]) >> 1, (c[2] * 8 + fl[2]) >> 1, &off);
            if (p < 0) continue;
            float sum = 0;
            for (int ii = 0; ii < 2; ii++)
                for (int jj = 0; jj < 2; jj++)
                    for (int kk = 0; kk < 2; kk++)
                        sum += g->ch[ch][(size_t)t * 512 + og_vi(fl[0] + ii, fl[1] + jj, fl[2] + kk)];
            g->ch[ch][(size_t)p * 512 + off] += sum * coeff;
        }
    }
}

/* FullNegativeLaplacian (GMGSolver.cu:167-171) / ConservativeFullNegativeLaplacian
 * (CMGSolver.cu:209-225) */
void og_geo_full_neg_lap(og_grid* g, int mode, int x, int y, int diag)
{
    og_propagate_ghosts(g, x);
    if (mode == OG_GEO_GMG) {
        og_geo_neg_lap_tiles(g, mode, -1, OG_LEAF | OG_GHOST, x, y, diag);
        accumulate_ghosts_add(g, y, 1.0f);
    } else {
        og_accumulate_leafs(g, x);
        og_geo_neg_lap_tiles(g, mode, -1, OG_LEAF, x, y, 0);
    }
}

/* one colour of NegativeLaplacianAndGaussSeidelSameLevel128Kernel
 * (GMGSolver.cu:173-206) / its CMG twin (CMGSolver.cu:228-264) on the LEAF
 * subtree tiles of `level` */
void og_geo_gs_color(og_grid* g, int mode, int level, int x, int rhs, int color)
{
    float* live = g->ch[x];
    float* snap = (float*)malloc(sizeof(float) * (size_t)g->N * 512);
    memcpy(snap, live, sizeof(float) * (size_t)g->N * 512);
    g->ch[x] = snap;
    #pragma omp parallel
    {
        geo_lds* s = (geo_lds*)malloc(sizeof(geo_lds));
        #pragma omp for schedule(dynamic, 8)
        for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
            if (!(og_subtree_type(g->t_type[t], level, level) & OG_LEAF)) continue;
            geo_load(g, mode, t, level, x, s);
            float h = og_h(g, level);
            for (int v = 0; v < 512; v++) {
                int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
                if ((i + j + k) % 2 != color) continue;
                size_t idx = (size_t)t * 512 + v;
                uint8_t ct = GC(s, i, j, k);
                float b = g->ch[rhs][idx];
                float Ax = geo_neg_lap(s, mode, h, i, j, k, 0);
                float D = geo_neg_lap(s, mode, h, i, j, k, 1);
                float delta = (ct & OG_INTERIOR) ? (b - Ax) / D : 0.0f;
                live[idx] = GX(s, i, j, k) + delta;
            }
        }
        free(s);
    }
    g->ch[x] = live;
    free(snap);
}

static void geo_gauss_seidel(og_grid* g, int mode, int iters, int order, int level, int x, int rhs)
{
    for (int it = 0; it < iters; it++) {
        og_geo_gs_color(g, mode, level, x, rhs, order);
        og_geo_gs_color(g, mode, level, x, rhs, 1 - order);
    }
}

/* ResidualAndRestrict128Kernel (GMGSolver.cu:215-275, LEAF|NONLEAF|GHOST tiles:
 * GHOST tiles add their residual sums to the parent) / its CMG twin
 * (CMGSolver.cu:273-330, LEAF|NONLEAF tiles) */
void og_geo_residual_restrict(og_grid* g, int mode, int level, int x, int rhs, float one_over_alpha)
{
    int types = mode == OG_GEO_GMG ? (OG_LEAF | OG_GHOST) : OG_LEAF;
    #pragma omp parallel
    {
        geo_lds* s = (geo_lds*)malloc(sizeof(geo_lds));
        float res[512];
        #pragma omp for schedule(dynamic, 8)
        for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
            uint8_t st = og_subtree_type(g->t_type[t], level, level);
            if (!(st & types)) continue;
            geo_load(g, mode, t, level, x, s);
            float h = og_h(g, level);
            for (int v = 0; v < 512; v++)
                res[v] = g->ch[rhs][(size_t)t * 512 + v] - geo_neg_lap(s, mode, h, v >> 6, (v >> 3) & 7, v & 7, 0);
            const int32_t* c = &g->t_coord[3 * t];
            for (int q = 0; q < 64; q++) {
                int fl[3] = { q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2 };
                int off, p = og_voxel(g, level - 1, (c[0] * 8 + fl[0]) >> 1, (c[1] * 8 + fl[1]) >> 1, (c[2] * 8 + fl[2]) >> 1, &off);
                if (p < 0) continue;
                float sum = 0;
                for (int ii = 0; ii < 2; ii++)
                    for (int jj = 0; jj < 2; jj++)
                        for (int kk = 0; kk < 2; kk++) sum += res[og_vi(fl[0] + ii, fl[1] + jj, fl[2] + kk)];
                float* dst = &g->ch[rhs][(size_t)p * 512 + off];
                if (st & OG_GHOST) *dst += sum;
                else *dst = sum * one_over_alpha;
            }
        }
        free(s);
    }
}

/* ProlongateAndUpdate128Kernel (GMGSolver.cu:277-307, LEAF|GHOST) /
 * ConservativeProlongateAndUpdate128Kernel (CMGSolver.cu:338-369, LEAF only) */
void og_geo_prolong(og_grid* g, int mode, int level, int x, float P)
{
    if (mode == OG_GEO_GMG) {
        og_prolong_add(g, level, x, P);
        return;
    }
    #pragma omp parallel for schedule(dynamic, 8)
    for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
        if (!(og_subtree_type(g->t_type[t], level, level) & OG_LEAF)) continue;
        const int32_t* c = &g->t_coord[3 * t];
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            int gi = c[0] * 8 + (v >> 6), gj = c[1] * 8 + ((v >> 3) & 7), gk = c[2] * 8 + (v & 7);
            int off, p = og_voxel(g, level - 1, gi >> 1, gj >> 1, gk >> 1, &off);
            float cx = p < 0 ? 0.0f : g->ch[x][(size_t)p * 512 + off];
            g->ch[x][idx] = (g->ctype[idx] & OG_INTERIOR) ? g->ch[x][idx] + P * cx : 0.0f;
        }
    }
}

/* GMGSolver::VCycle (GMGSolver.cu:310-387) / CMGSolver::VCycle (CMGSolver.cu:372-424) */
void og_geo_vcycle(og_grid* g, int mode, int x, int f, int rhs, int level_iters, int coarsest_iters,
                   float one_over_alpha, float P)
{
    for (int t = 0; t < g->N; t++) {
        size_t b = (size_t)t * 512;
        memset(&g->ch[x][b], 0, 512 * sizeof(float));
        if (g->t_type[t] & OG_GHOST) memset(&g->ch[rhs][b], 0, 512 * sizeof(float));
        else memcpy(&g->ch[rhs][b], &g->ch[f][b], 512 * sizeof(float));
    }
    for (int l = g->max_level; l > 0; l--) {
        geo_gauss_seidel(g, mode, level_iters, 0, l, x, rhs);
        og_geo_residual_restrict(g, mode, l, x, rhs, one_over_alpha);
    }
    geo_gauss_seidel(g, mode, coarsest_iters / 2, 0, 0, x, rhs);
    geo_gauss_seidel(g, mode, coarsest_iters / 2, 1, 0, x, rhs);
    for (int l = 1; l <= g->max_level; l++) {
        og_geo_prolong(g, mode, l, x, P);
        geo_gauss_seidel(g, mode, level_iters, 1, l, x, rhs);
    }
}

/* GMGSolver::solve (GMGSolver.cu:472-607) / CMGSolver::solve (CMGSolver.cu:428-601),
 * sync_stride = 1; channels x=0, r=1, p=2, Ap=3, z=4 */
og_solve_out og_geo_pcg(og_grid* g, int mode, int level_iters, int coarsest_iters, float one_over_alpha, float Pc,
                        int max_iters, double tol, int pure_neumann)
{
    enum { X = 0, R = 1, P = 2, AP = 3, Z = 4 };
    og_solve_out out = { 0, 0.0 };
    for (int t = 0; t < g->N; t++)
        if (g->t_type[t] & OG_LEAF) memset(&g->ch[X][(size_t)t * 512], 0, 512 * sizeof(float));
    double rhs2 = og_dot(g, R, R);
    if (rhs2 == 0) return out;
    double thr2 = tol * tol * rhs2;
    if (thr2 < 2.2250738585072014e-308) thr2 = 2.2250738585072014e-308;
    if (pure_neumann) og_recenter(g, R);
    og_geo_vcycle(g, mode, Z, R, AP, level_iters, coarsest_iters, one_over_alpha, Pc);
    for (int t = 0; t < g->N; t++)
        if (g->t_type[t] & OG_LEAF) memcpy(&g->ch[P][(size_t)t * 512], &g->ch[Z][(size_t)t * 512], 512 * sizeof(float));
    double gamma = og_dot(g, Z, R);
    double rr = rhs2;
    int i;
    for (i = 0; i < max_iters; i++) {
        og_geo_full_neg_lap(g, mode, P, AP, 0);
        double fp = og_dot(g, P, AP);
        float a = (float)(gamma / fp);
        for (int t = 0; t < g->N; t++) {
            if (!(g->t_type[t] & OG_LEAF)) continue;
            for (int v = 0; v < 512; v++) {
                size_t idx = (size_t)t * 512 + v;
                if (!(g->ctype[idx] & OG_INTERIOR)) continue;
                g->ch[X][idx] = fmaf(a, g->ch[P][idx], g->ch[X][idx]);
                g->ch[R][idx] = fmaf(-a, g->ch[AP][idx], g->ch[R][idx]);
            }
        }
        rr = og_dot(g, R, R);
        if (rr < thr2) { out.iters = i + 1; out.rel_res = sqrt(rr / rhs2); return out; }
        if (pure_neumann) og_recenter(g, R);
        og_geo_vcycle(g, mode, Z, R, AP, level_iters, coarsest_iters, one_over_alpha, Pc);
        double gamma_old = gamma;
        gamma = og_dot(g, R, Z);
        float b = (float)(gamma / gamma_old);
        for (int t = 0; t < g->N; t++) {
            if (!(g->t_type[t] & OG_LEAF)) continue;
            for (int v = 0; v < 512; v++) {
                size_t idx = (size_t)t * 512 + v;
                if (!(g->ctype[idx] & OG_INTERIOR)) continue;
                g->ch[P][idx] = fmaf(b, g->ch[P][idx], g->ch[Z][idx]);
            }
        }
    }
    rr = og_dot(g, R, R);
    out.iters = i;
    out.rel_res = sqrt(rr / rhs2);
    return out;
}

/* GMGSolver::dampedJacobiSolve (GMGSolver.cu:390-470): x = 0 (channel 0), b in
 * channel 1, residual in channel 2, diagonal in channel 3 */
og_solve_out og_gmg_damped_jacobi(og_grid* g, int max_iters, double tol, float omega)
{
    enum { X = 0, B = 1, AX = 2, D = 3 };
    og_solve_out out = { 0, 0.0 };
    memset(g->ch[X], 0, sizeof(float) * (size_t)g->N * 512);
    double bb = 0;
    for (size_t idx = 0; idx < (size_t)g->N * 512; idx++) bb += (double)g->ch[B][idx] * (double)g->ch[B][idx];
    double rhs_norm = sqrt(bb);
    double thr = tol * rhs_norm;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    double res = rhs_norm;
    int i;
    for (i = 0; i < max_iters; i++) {
        og_geo_full_neg_lap(g, OG_GEO_GMG, X, AX, 0);
        for (int t = 0; t < g->N; t++) {
            if (!(g->t_type[t] & OG_LEAF)) continue;
            for (int v = 0; v < 512; v++) {
                size_t idx = (size_t)t * 512 + v;
                if (g->ctype[idx] & OG_INTERIOR) g->ch[AX][idx] = g->ch[B][idx] - g->ch[AX][idx];
            }
        }
        res = sqrt(og_dot(g, AX, AX));
        if (res < thr) break;
        og_geo_full_neg_lap(g, OG_GEO_GMG, X, D, 1);
        for (int t = 0; t < g->N; t++) {
            if (!(g->t_type[t] & OG_LEAF)) continue;
            for (int v = 0; v < 512; v++) {
                size_t idx = (size_t)t * 512 + v;
                if (!(g->ctype[idx] & OG_INTERIOR)) continue;
                float d = g->ch[D][idx];
                if (d == 0) g->ch[X][idx] = 0;
                else g->ch[X][idx] = g->ch[X][idx] + omega * g->ch[AX][idx] / d;
            }
        }
    }
    out.iters = i;
    out.rel_res = res / rhs_norm;
    return out;
}
