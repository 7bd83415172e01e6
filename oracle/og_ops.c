/*
 * og_ops.c -- oracle restatement of the matrix-free AMG operator, smoothers,
 * inter-level transfers, FAS cycles and PCG.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows src/AMGSolver.cu:320-1549 and src/PoissonGrid.cu:156-465,889-921.
 * The reference launches one block per tile; every kernel restated here reads
 * only values that no other block of the same launch writes, so a sequential
 * loop over tiles (parallel under OpenMP) reproduces it exactly.
 *
 * Arithmetic contract shared with the HIP kernels (bit-exact comparisons):
 *   apply:     s = x0*D; then s = fmaf(off, xn, s) for axis 0..2, side -,+
 *   ghost:     vg = (float)((double)x_c + 0.5 * (double)(V1 - avg8)),
 *              avg8 = (sum of the aligned 2^3 block in di,dj,dk order) / 8
 *   RB-SOR:    x = (float)((double)x + (double)((b - Ax)/D) * omega)
 *   Jacobi:    x = x + w * ((b - Ax)/D)
 *   restrict:  b_c = (sum_8 r) * R ; x_c = (sum_8 x) / 8   (ii,jj,kk order)
 *   prolong:   x = x + P*(cx - cx0)
 *   axpy:      x = fmaf(a, p, x); r = fmaf(-a, Ap, r); p = fmaf(b, p, z)
 * Dot products use exact float*float products accumulated in double.
 */
#include "og_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float x[10][10][10];
    float diag[8][8][8];
    float off0[9][8][8];
    float off1[8][9][8];
    float off2[8][8][9];
} og_lds;

static inline float* offp(og_lds* s, int a, int i, int j, int k)
{
    if (a == 0) return &s->off0[i][j][k];
    if (a == 1) return &s->off1[i][j][k];
    return &s->off2[i][j][k];
}
#define XV(s, i, j, k) ((s)->x[(i) + 1][(j) + 1][(k) + 1])

/* FaceNeighborLocalCoord (AMGSolver.cu:384-389) */
static void face_nb_coord(int axis, int sgn, int vi, int* f)
{
    int axk = sgn == -1 ? -1 : 8;
    int i = vi / 8, j = vi % 8;
    if (axis == 0) { f[0] = axk; f[1] = i; f[2] = j; }
    else if (axis == 1) { f[0] = i; f[1] = axk; f[2] = j; }
    else { f[0] = i; f[1] = j; f[2] = axk; }
}

/* LoadAMGLaplacianTileData (AMGSolver.cu:395-543); ghost values are read from
 * channel gch (the reference reads them from the x channel itself) */
static void load_tile(const og_grid* g, int t, int xch, int gch, int c0, og_lds* s)
{
    const float* X = g->ch[xch];
    const uint8_t* T = g->ctype;
    size_t base = (size_t)t * 512;
    float h = og_h(g, g->t_level[t]);
    for (int v = 0; v < 512; v++) {
        int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
        XV(s, i, j, k) = T[base + v] == OG_INTERIOR ? X[base + v] : 0.0f;
        s->off0[i][j][k] = g->ch[c0][base + v];
        s->off1[i][j][k] = g->ch[c0 + 1][base + v];
        s->off2[i][j][k] = g->ch[c0 + 2][base + v];
        s->diag[i][j][k] = g->ch[c0 + 3][base + v];
    }
    for (int a = 0; a < 3; a++)
        for (int sg = -1; sg <= 1; sg += 2)
            for (int fi = 0; fi < 64; fi++) {
                int fl[3], cl[3];
                face_nb_coord(a, sg, fi, fl);
                cl[0] = fl[0]; cl[1] = fl[1]; cl[2] = fl[2];
                cl[a] -= sg;
                float avg = 0;
                for (int di = 0; di < 2; di++)
                    for (int dj = 0; dj < 2; dj++)
                        for (int dk = 0; dk < 2; dk++)
                            avg += XV(s, cl[0] ^ di, cl[1] ^ dj, cl[2] ^ dk);
                avg /= 8;
                int nl[3] = { fl[0], fl[1], fl[2] };
                nl[a] = sg == -1 ? 7 : 0;
                int nt = g->nbr[6 * t + (sg == -1 ? a : a + 3)];
                int nv = og_vi(nl[0], nl[1], nl[2]);
                float* xh = &XV(s, fl[0], fl[1], fl[2]);
                if (nt < 0) {
                    *xh = 0;
                    if (sg == 1) *offp(s, a, fl[0], fl[1], fl[2]) = h;
                } else if (g->t_type[nt] & OG_GHOST) {
                    float V1 = g->ch[gch][(size_t)nt * 512 + nv];
                    float V0 = avg;
                    *xh = (float)((double)XV(s, cl[0], cl[1], cl[2]) + 0.5 * (double)(V1 - V0));
                    if (sg == 1) *offp(s, a, fl[0], fl[1], fl[2]) = g->ch[c0 + a][(size_t)nt * 512 + nv];
                } else {
                    size_t ni = (size_t)nt * 512 + nv;
                    *xh = T[ni] == OG_INTERIOR ? X[ni] : 0.0f;
                    if (sg == 1) *offp(s, a, fl[0], fl[1], fl[2]) = g->ch[c0 + a][ni];
                }
            }
}

/* AMGLaplacianTileData::negativeLap (AMGSolver.cu:345-382) */
static float neg_lap(og_lds* s, int i, int j, int k)
{
    float d = s->diag[i][j][k];
    if (d == 0) return 0;
    float sum = XV(s, i, j, k) * d;
    const int l[3] = { i, j, k };
    for (int a = 0; a < 3; a++)
        for (int sg = -1; sg <= 1; sg += 2) {
            int n[3] = { l[0], l[1], l[2] }, u[3] = { l[0], l[1], l[2] };
            n[a] += sg;
            u[a] += sg == -1 ? 0 : 1;
            sum = fmaf(*offp(s, a, u[0], u[1], u[2]), XV(s, n[0], n[1], n[2]), sum);
        }
    return sum;
}

/* PropagateToChildren(GHOST, subtree, coarse first, all cell types)
 * (PoissonGrid.cu:377-405): ghost cell = parent cell value */
void og_propagate_ghosts(og_grid* g, int ch)
{
    for (int lvl = 0; lvl < g->num_levels; lvl++)
        for (int t = g->lstart[lvl]; t < g->lstart[lvl + 1]; t++) {
            if (!(g->t_type[t] & OG_GHOST)) continue;
            const int32_t* c = &g->t_coord[3 * t];
            for (int v = 0; v < 512; v++) {
                int gi = c[0] * 8 + (v >> 6), gj = c[1] * 8 + ((v >> 3) & 7), gk = c[2] * 8 + (v & 7);
                int off, p = og_voxel(g, lvl - 1, gi >> 1, gj >> 1, gk >> 1, &off);
                g->ch[ch][(size_t)t * 512 + v] = p >= 0 ? g->ch[ch][(size_t)p * 512 + off] : 0.0f;
            }
        }
}

/* AccumulateToParentsOneStep(LEAF, 1/8, overwrite, all cell types)
 * (PoissonGrid.cu:408-465) */
void og_accumulate_leafs(og_grid* g, int ch)
{
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
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
            g->ch[ch][(size_t)p * 512 + off] = sum * 0.125f;
        }
    }
}

void og_neg_lap_leafs(og_grid* g, int x, int c0, int y)
{
    #pragma omp parallel
    {
        og_lds* s = (og_lds*)malloc(sizeof(og_lds));
        #pragma omp for schedule(dynamic, 8)
        for (int t = 0; t < g->N; t++) {
            if (!(g->t_type[t] & OG_LEAF)) continue;
            load_tile(g, t, x, x, c0, s);
            for (int v = 0; v < 512; v++)
                g->ch[y][(size_t)t * 512 + v] = g->ctype[(size_t)t * 512 + v] == OG_INTERIOR ? neg_lap(s, v >> 6, (v >> 3) & 7, v & 7) : 0.0f;
        }
        free(s);
    }
}

void og_full_neg_lap(og_grid* g, int x, int c0, int y)
{
    og_propagate_ghosts(g, x);
    og_accumulate_leafs(g, x);
    og_neg_lap_leafs(g, x, c0, y);
}

static float* snapshot(const og_grid* g, int ch)
{
    float* s = (float*)malloc(sizeof(float) * (size_t)g->N * 512);
    memcpy(s, g->ch[ch], sizeof(float) * (size_t)g->N * 512);
    return s;
}

/* one colour of NegativeLaplacianAndGaussSeidelSameLevelAMG128Kernel
 * (AMGSolver.cu:838-877) on subtree-LEAF tiles of `level`.  Updated cells only
 * read opposite-colour values, so reading a snapshot is exact. */
void og_gs_color(og_grid* g, int level, int x, int c0, int rhs, int color, double omega)
{
    float* live = g->ch[x];
    float* snap = snapshot(g, x);
    g->ch[x] = snap;
    #pragma omp parallel
    {
        og_lds* s = (og_lds*)malloc(sizeof(og_lds));
        #pragma omp for schedule(dynamic, 8)
        for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
            if (!(og_subtree_type(g->t_type[t], level, level) & OG_LEAF)) continue;
            load_tile(g, t, x, x, c0, s);
            for (int v = 0; v < 512; v++) {
                int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
                if ((i + j + k) % 2 != color) continue;
                size_t idx = (size_t)t * 512 + v;
                uint8_t ct = g->ctype[idx];
                float b = g->ch[rhs][idx];
                float Ax = neg_lap(s, i, j, k);
                float D = s->diag[i][j][k];
                float delta = (ct & OG_INTERIOR) ? (b - Ax) / D : 0.0f;
                live[idx] = (float)((double)XV(s, i, j, k) + (double)delta * omega);
            }
        }
        free(s);
    }
    g->ch[x] = live;
    free(snap);
}

/* GaussSeidelAMG (AMGSolver.cu:887-894) */
void og_gauss_seidel(og_grid* g, int iters, int order, int level, int x, int c0, int rhs, double omega)
{
    for (int it = 0; it < iters; it++) {
        og_gs_color(g, level, x, c0, rhs, order, omega);
        og_gs_color(g, level, x, c0, rhs, 1 - order, omega);
    }
}

void og_jacobi(og_grid* g, int level, int x_in, int x_out, int ghost_ch, int c0, int rhs, float omega)
{
    #pragma omp parallel
    {
        og_lds* s = (og_lds*)malloc(sizeof(og_lds));
        #pragma omp for schedule(dynamic, 8)
        for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
            if (!(og_subtree_type(g->t_type[t], level, level) & OG_LEAF)) continue;
            load_tile(g, t, x_in, ghost_ch, c0, s);
            for (int v = 0; v < 512; v++) {
                int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
                size_t idx = (size_t)t * 512 + v;
                uint8_t ct = g->ctype[idx];
                float b = g->ch[rhs][idx];
                float Ax = neg_lap(s, i, j, k);
                float D = s->diag[i][j][k];
                float delta = (ct & OG_INTERIOR) ? (b - Ax) / D : 0.0f;
                g->ch[x_out][idx] = XV(s, i, j, k) + omega * delta;
            }
        }
        free(s);
    }
}

/* ResidualAndRestrictAMG128Kernel (AMGSolver.cu:896-964), launched on
 * subtree-LEAF tiles of `level` (LEAF and NONLEAF, never GHOST) */
void og_residual_restrict(og_grid* g, int level, int x, int c0, int rhs, float Rr)
{
    #pragma omp parallel
    {
        og_lds* s = (og_lds*)malloc(sizeof(og_lds));
        float res[512];
        #pragma omp for schedule(dynamic, 8)
        for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
            if (!(og_subtree_type(g->t_type[t], level, level) & OG_LEAF)) continue;
            load_tile(g, t, x, x, c0, s);
            for (int v = 0; v < 512; v++)
                res[v] = g->ch[rhs][(size_t)t * 512 + v] - neg_lap(s, v >> 6, (v >> 3) & 7, v & 7);
            const int32_t* c = &g->t_coord[3 * t];
            for (int q = 0; q < 64; q++) {
                int fl[3] = { q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2 };
                int off, p = og_voxel(g, level - 1, (c[0] * 8 + fl[0]) >> 1, (c[1] * 8 + fl[1]) >> 1, (c[2] * 8 + fl[2]) >> 1, &off);
                if (p < 0) continue;
                float rs = 0, xs = 0;
                for (int ii = 0; ii < 2; ii++)
                    for (int jj = 0; jj < 2; jj++)
                        for (int kk = 0; kk < 2; kk++) {
                            rs += res[og_vi(fl[0] + ii, fl[1] + jj, fl[2] + kk)];
                            xs += XV(s, fl[0] + ii, fl[1] + jj, fl[2] + kk);
                        }
                g->ch[rhs][(size_t)p * 512 + off] = rs * Rr;
                g->ch[x][(size_t)p * 512 + off] = xs / 8;
            }
        }
        free(s);
    }
}

/* FASSaveSolutionAndUpdateSource128Kernel (AMGSolver.cu:1176-1214) on NONLEAF tiles */
void og_fas_save(og_grid* g, int level, int x, int c0, int rhs, int x0)
{
    #pragma omp parallel
    {
        og_lds* s = (og_lds*)malloc(sizeof(og_lds));
        #pragma omp for schedule(dynamic, 8)
        for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
            if (!(g->t_type[t] & OG_NONLEAF)) continue;
            load_tile(g, t, x, x, c0, s);
            for (int v = 0; v < 512; v++) {
                size_t idx = (size_t)t * 512 + v;
                if (g->ctype[idx] & OG_INTERIOR) {
                    g->ch[rhs][idx] = g->ch[rhs][idx] + neg_lap(s, v >> 6, (v >> 3) & 7, v & 7);
                    g->ch[x0][idx] = g->ch[x][idx];
                } else {
                    g->ch[rhs][idx] = 0;
                }
            }
        }
        free(s);
    }
}

/* FASProlongateAndUpdateAMG128Kernel (AMGSolver.cu:1216-1255) on LEAF|GHOST
 * subtree tiles of `level` */
void og_fas_prolong(og_grid* g, int level, int x, int x0, float P)
{
    #pragma omp parallel for schedule(dynamic, 8)
    for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
        uint8_t ft = og_subtree_type(g->t_type[t], level, level);
        if (!(ft & (OG_LEAF | OG_GHOST))) continue;
        const int32_t* c = &g->t_coord[3 * t];
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            int gi = c[0] * 8 + (v >> 6), gj = c[1] * 8 + ((v >> 3) & 7), gk = c[2] * 8 + (v & 7);
            int off, p = og_voxel(g, level - 1, gi >> 1, gj >> 1, gk >> 1, &off);
            float cx = p < 0 ? 0.0f : g->ch[x][(size_t)p * 512 + off];
            float cx0 = p < 0 ? 0.0f : g->ch[x0][(size_t)p * 512 + off];
            float fx = g->ch[x][idx];
            uint8_t ct = g->ctype[idx];
            if (ft & OG_GHOST) g->ch[x][idx] = (ct & OG_INTERIOR) ? cx : 0.0f;
            else g->ch[x][idx] = (ct & OG_INTERIOR) ? fx + P * (cx - cx0) : 0.0f;
        }
    }
}

/* RestrictSourceToCoarse (AMGSolver.cu:1295-1307) =
 * AccumulateToParentsOneStepKernel (PoissonGrid.cu:408-451) over the LEAF
 * subtree tiles of `level`, INTERIOR cells, overwriting the parent */
void og_restrict_source(og_grid* g, int level, int ch, float coeff)
{
    #pragma omp parallel for schedule(dynamic, 8)
    for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
        if (!(og_subtree_type(g->t_type[t], level, level) & OG_LEAF)) continue;
        const int32_t* c = &g->t_coord[3 * t];
        for (int q = 0; q < 64; q++) {
            int fl[3] = { q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2 };
            int off, p = og_voxel(g, level - 1, (c[0] * 8 + fl[0]) >> 1, (c[1] * 8 + fl[1]) >> 1, (c[2] * 8 + fl[2]) >> 1, &off);
            if (p < 0) continue;
            float sum = 0;
            for (int ii = 0; ii < 2; ii++)
                for (int jj = 0; jj < 2; jj++)
                    for (int kk = 0; kk < 2; kk++) {
                        size_t idx = (size_t)t * 512 + og_vi(fl[0] + ii, fl[1] + jj, fl[2] + kk);
                        sum += (g->ctype[idx] & OG_INTERIOR) ? g->ch[ch][idx] : 0.0f;
                    }
            g->ch[ch][(size_t)p * 512 + off] = sum * coeff;
        }
    }
}

/* ProlongateAndUpdateAMG128Kernel (AMGSolver.cu:1022-1053) on LEAF|GHOST
 * subtree tiles of `level`: GHOST x = parent x, others x += P * parent x */
void og_prolong_add(og_grid* g, int level, int x, float P)
{
    #pragma omp parallel for schedule(dynamic, 8)
    for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
        uint8_t ft = og_subtree_type(g->t_type[t], level, level);
        if (!(ft & (OG_LEAF | OG_GHOST))) continue;
        const int32_t* c = &g->t_coord[3 * t];
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            int gi = c[0] * 8 + (v >> 6), gj = c[1] * 8 + ((v >> 3) & 7), gk = c[2] * 8 + (v & 7);
            int off, p = og_voxel(g, level - 1, gi >> 1, gj >> 1, gk >> 1, &off);
            float cx = p < 0 ? 0.0f : g->ch[x][(size_t)p * 512 + off];
            float fx = g->ch[x][idx];
            uint8_t ct = g->ctype[idx];
            if (ft & OG_GHOST) g->ch[x][idx] = (ct & OG_INTERIOR) ? cx : 0.0f;
            else g->ch[x][idx] = (ct & OG_INTERIOR) ? fx + P * cx : 0.0f;
        }
    }
}

/* DeductFluxOnAbsLeafs128Kernel (AMGSolver.cu:966-993) on LEAF tiles of
 * `level`: rhs = f - negativeLapOuterNonLeaf(x) on INTERIOR cells, else 0.
 * negativeLapOuterNonLeaf (AMGSolver.cu:383-403) keeps only the terms whose
 * cell lies in a NONLEAF tile; the own tile is LEAF, so only halo cells of
 * NONLEAF face neighbours contribute. */
void og_deduct_flux(og_grid* g, int level, int x, int c0, int f, int rhs)
{
    #pragma omp parallel
    {
        og_lds* s = (og_lds*)malloc(sizeof(og_lds));
        #pragma omp for schedule(dynamic, 8)
        for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++) {
            if (!(g->t_type[t] & OG_LEAF)) continue;
            load_tile(g, t, x, x, c0, s);
            /* keep only halo values of NONLEAF neighbours */
            for (int i = -1; i <= 8; i++)
                for (int j = -1; j <= 8; j++)
                    for (int k = -1; k <= 8; k++) {
                        int out = (i < 0) + (i > 7) + (j < 0) + (j > 7) + (k < 0) + (k > 7);
                        if (out != 1) { XV(s, i, j, k) = 0.0f; continue; }
                        int a = (i < 0 || i > 7) ? 0 : (j < 0 || j > 7) ? 1 : 2;
                        int l = a == 0 ? i : a == 1 ? j : k;
                        int nt = g->nbr[6 * t + (l < 0 ? a : a + 3)];
                        if (nt < 0 || g->t_type[nt] != OG_NONLEAF) XV(s, i, j, k) = 0.0f;
                    }
            for (int v = 0; v < 512; v++) {
                size_t idx = (size_t)t * 512 + v;
                if (!(g->ctype[idx] & OG_INTERIOR)) { g->ch[rhs][idx] = 0.0f; continue; }
                int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
                float d = s->diag[i][j][k];
                float sum = 0.0f;
                if (d != 0) {
                    sum = 0.0f * d;
                    const int l[3] = { i, j, k };
                    for (int a = 0; a < 3; a++)
                        for (int sg = -1; sg <= 1; sg += 2) {
                            int n[3] = { l[0], l[1], l[2] }, u[3] = { l[0], l[1], l[2] };
                            n[a] += sg;
                            u[a] += sg == -1 ? 0 : 1;
                            sum = fmaf(*offp(s, a, u[0], u[1], u[2]), XV(s, n[0], n[1], n[2]), sum);
                        }
                }
                g->ch[rhs][idx] = g->ch[f][idx] - sum;
            }
        }
        free(s);
    }
}

/* SetLevelZeroAbsTiles128Kernel (AMGSolver.cu:995-1016) */
void og_set_zero_tiles(og_grid* g, int level, int tile_types, int x)
{
    for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++)
        if (g->t_type[t] & tile_types) memset(&g->ch[x][(size_t)t * 512], 0, 512 * sizeof(float));
}

/* AMGAddGradientToFace (AMGSolver.cu:634-688), weighted = 0: u += (p - p_down) / h
 * on faces whose fluid ratio -c/h is positive; AMGAddFaceWeightedGradientToFace
 * (:690-748), weighted = 1: u += (p - p_down) / h * (-c / h).  On LEAF tiles; x is
 * read through the operator's loader after ghosts take their parents' values and
 * NONLEAF cells the mean of their children (x keeps those values). */
void og_add_gradient(og_grid* g, int x, int c0, int u0, int weighted)
{
    og_propagate_ghosts(g, x);
    og_accumulate_leafs(g, x);
    #pragma omp parallel
    {
        og_lds* s = (og_lds*)malloc(sizeof(og_lds));
        #pragma omp for schedule(dynamic, 8)
        for (int t = 0; t < g->N; t++) {
            if (!(g->t_type[t] & OG_LEAF)) continue;
            load_tile(g, t, x, x, c0, s);
            float h = og_h(g, g->t_level[t]);
            for (int v = 0; v < 512; v++) {
                size_t idx = (size_t)t * 512 + v;
                int l[3] = { v >> 6, (v >> 3) & 7, v & 7 };
                for (int a = 0; a < 3; a++) {
                    int d[3] = { l[0], l[1], l[2] };
                    d[a] -= 1;
                    float pu = XV(s, l[0], l[1], l[2]), pd = XV(s, d[0], d[1], d[2]);
                    float ratio = -*offp(s, a, l[0], l[1], l[2]) / h;
                    float* u = &g->ch[u0 + a][idx];
                    if (weighted) *u = *u + (pu - pd) / h * ratio;
                    else if (ratio > 0) *u = *u + (pu - pd) / h;
                }
            }
        }
        free(s);
    }
}

/* FluidEuler::mixFluidAndSolidVelocityOnFaces (cirrus_cutcell/FluidEuler.cu:531-552)
 * with the solid face velocity given in channels s0..s0+2: on LEAF tiles
 * u_mix = r u_fluid (r > 0) + (1 - r) u_solid (1 - r > 0), r = -c/h */
void og_mix_velocity(og_grid* g, int c0, int uf0, int us0, int um0)
{
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        float h = og_h(g, g->t_level[t]);
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            for (int a = 0; a < 3; a++) {
                float fr = -g->ch[c0 + a][idx] / h;
                float sr = 1 - fr;
                float mix = 0;
                if (fr > 0) mix += fr * g->ch[uf0 + a][idx];
                if (sr > 0) mix += sr * g->ch[us0 + a][idx];
                g->ch[um0 + a][idx] = mix;
            }
        }
    }
}

double og_dot(og_grid* g, int a, int b)
{
    double s = 0;
    #pragma omp parallel for reduction(+:s) schedule(static)
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        double ts = 0;
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            ts += (double)g->ch[a][idx] * (double)g->ch[b][idx];
        }
        s += ts;
    }
    return s;
}

/* MarkRegionOfInterestWithChannelMinAndMax128Kernel (src/PoissonGridKernels.cuh:11-54):
 * for every tile whose subtree type (NONLEAF at subtree_level counts as LEAF,
 * src/HATileInfo.h:79-84) is in launch_types, min and max of the 512 values of ch and
 * the flag pred(min, max): 0 max > a, 1 min < a, 2 [min, max] meets [a, b], 3 max - min > a
 * (the reference takes a device lambda; its callers use max > 0, src/MarkerParticles.cu:49-61);
 * other tiles get flag 0 and keep tmin / tmax */
void og_roi_flags(const og_grid* g, int ch, int subtree_level, int launch_types, int pred, float a, float b,
                  uint8_t* interest, float* tmin, float* tmax)
{
    for (int t = 0; t < g->N; t++) {
        uint8_t tt = g->t_type[t];
        if (g->t_level[t] == subtree_level && tt == OG_NONLEAF) tt = OG_LEAF;
        if (!(tt & launch_types)) { interest[t] = 0; continue; }
        const float* x = g->ch[ch] + (size_t)t * 512;
        float lo = x[0], hi = x[0];
        for (int v = 1; v < 512; v++) {
            if (x[v] < lo) lo = x[v];
            if (x[v] > hi) hi = x[v];
        }
        int f;
        switch (pred) {
        case 0: f = hi > a; break;
        case 1: f = lo < a; break;
        case 2: f = lo <= b && hi >= a; break;
        default: f = (hi - lo) > a; break;
        }
        interest[t] = (uint8_t)f;
        tmin[t] = lo;
        tmax[t] = hi;
    }
}

/* NormSync (PoissonGrid.cu:322-375) with ChannelPowerSumKernel128 (:223-319):
 * over LEAF tiles, cells whose type is in cell_types, |x|; order -1 = max,
 * else (sum |x|^p w / sum w)^(1/p), w = h^3 when volume weighted; FP64 */
double og_norm(og_grid* g, int order, int ch, int volume_weighted, int cell_types)
{
    double vs = 0, ws = 0;
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        double w = 1;
        if (volume_weighted) { double h = og_h(g, g->t_level[t]); w = h * h * h; }
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            if (!(g->ctype[idx] & cell_types)) continue;
            double a = fabs((double)g->ch[ch][idx]);
            if (order == -1) { if (a > vs) vs = a; continue; }
            vs += (order == 1 ? a : order == 2 ? a * a : pow(a, order)) * w;
            ws += w;
        }
    }
    if (order == -1) return vs;
    if (ws == 0) return 0;
    if (order == 1) return vs / ws;
    if (order == 2) return sqrt(vs / ws);
    return pow(vs / ws, 1.0 / order);
}

/* ReCenterLeafCells (PoissonGrid.cu:889-921): subtract the pointwise mean over
 * INTERIOR cells of LEAF tiles */
void og_recenter(og_grid* g, int ch)
{
    double sum = 0, cnt = 0;
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            if (g->ctype[idx] & OG_INTERIOR) { sum += g->ch[ch][idx]; cnt += 1; }
        }
    }
    double mean = sum / cnt;
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            if (g->ctype[idx] & OG_INTERIOR) g->ch[ch][idx] = (float)((double)g->ch[ch][idx] - mean);
        }
    }
}

/* ---- coarsest-level FP64 dense solve --------------------------------------- */

/* dense FP64 solve of the level-0 operator restricted to its INTERIOR cells
 * (columns probed from the same tile operator); FAS: x = A0^-1 b */
static void coarsest_direct(og_grid* g, int x, int c0, int rhs)
{
    int lo = g->lstart[0], hi = g->lstart[1];
    int n = 0;
    int* map = (int*)malloc(sizeof(int) * (size_t)(hi - lo) * 512);
    for (int t = lo; t < hi; t++)
        for (int v = 0; v < 512; v++)
            map[(t - lo) * 512 + v] = (g->ctype[(size_t)t * 512 + v] & OG_INTERIOR) && (og_subtree_type(g->t_type[t], 0, 0) & OG_LEAF) ? n++ : -1;
    if (n == 0) { free(map); return; }
    double* A = (double*)calloc((size_t)n * n, sizeof(double));
    double* b = (double*)malloc(sizeof(double) * n);
    /* probe column by column with unit vectors in a scratch channel */
    float* save = snapshot(g, x);
    og_lds* s = (og_lds*)malloc(sizeof(og_lds));
    int col = 0;
    for (int t = lo; t < hi; t++)
        for (int v = 0; v < 512; v++) {
            if (map[(t - lo) * 512 + v] < 0) continue;
            for (int tt = lo; tt < hi; tt++) memset(&g->ch[x][(size_t)tt * 512], 0, 512 * sizeof(float));
            g->ch[x][(size_t)t * 512 + v] = 1.0f;
            for (int tt = lo; tt < hi; tt++) {
                load_tile(g, tt, x, x, c0, s);
                for (int w = 0; w < 512; w++) {
                    int r = map[(tt - lo) * 512 + w];
                    if (r >= 0) A[(size_t)r * n + col] = (double)neg_lap(s, w >> 6, (w >> 3) & 7, w & 7);
                }
            }
            col++;
        }
    memcpy(g->ch[x], save, sizeof(float) * (size_t)g->N * 512);
    free(save); free(s);
    for (int t = lo; t < hi; t++)
        for (int v = 0; v < 512; v++) {
            int r = map[(t - lo) * 512 + v];
            if (r >= 0) b[r] = g->ch[rhs][(size_t)t * 512 + v];
        }
    /* LU with partial pivoting */
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++) if (fabs(A[(size_t)i * n + k]) > fabs(A[(size_t)piv * n + k])) piv = i;
        if (piv != k) {
            for (int j = 0; j < n; j++) { double tmp = A[(size_t)k * n + j]; A[(size_t)k * n + j] = A[(size_t)piv * n + j]; A[(size_t)piv * n + j] = tmp; }
            double tmp = b[k]; b[k] = b[piv]; b[piv] = tmp;
        }
        double d = A[(size_t)k * n + k];
        if (d == 0) continue;
        for (int i = k + 1; i < n; i++) {
            double f = A[(size_t)i * n + k] / d;
            if (f == 0) continue;
            for (int j = k; j < n; j++) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double acc = b[i];
        for (int j = i + 1; j < n; j++) acc -= A[(size_t)i * n + j] * b[j];
        b[i] = A[(size_t)i * n + i] != 0 ? acc / A[(size_t)i * n + i] : 0;
    }
    for (int t = lo; t < hi; t++)
        for (int v = 0; v < 512; v++) {
            int r = map[(t - lo) * 512 + v];
            if (og_subtree_type(g->t_type[t], 0, 0) & OG_LEAF)
                g->ch[x][(size_t)t * 512 + v] = r >= 0 ? (float)b[r] : 0.0f;
        }
    free(A); free(b); free(map);
}

/* ---- cycles ----------------------------------------------------------------- */

static void copy_level(og_grid* g, int level, int src, int dst)
{
    for (int t = g->lstart[level]; t < g->lstart[level + 1]; t++)
        if (og_subtree_type(g->t_type[t], level, level) & OG_LEAF)
            memcpy(&g->ch[dst][(size_t)t * 512], &g->ch[src][(size_t)t * 512], 512 * sizeof(float));
}

static void smooth(og_grid* g, const og_cycle* cyc, int level, int iters, int order, int x, int c0, int rhs)
{
    if (cyc->smoother == 0) {
        og_gauss_seidel(g, iters, order, level, x, c0, rhs, cyc->omega);
        return;
    }
    int cur = x, other = cyc->tmp_ch;
    for (int it = 0; it < iters; it++) {
        og_jacobi(g, level, cur, other, x, c0, rhs, (float)cyc->omega);
        int tmp = cur; cur = other; other = tmp;
    }
    if (cur != x) copy_level(g, level, cur, x);
}

/* FASMuCycleStep (AMGSolver.cu:1257-1277) with the smoother / coarsest choice */
void og_fas_cycle_step(og_grid* g, const og_cycle* cyc, int level, int x, int rhs, int x0, int c0)
{
    if (level == 0) {
        if (cyc->coarsest_direct) coarsest_direct(g, x, c0, rhs);
        else if (cyc->smoother == 0) {
            og_gauss_seidel(g, cyc->coarsest_iters / 2, 0, 0, x, c0, rhs, cyc->omega);
            og_gauss_seidel(g, cyc->coarsest_iters / 2, 1, 0, x, c0, rhs, cyc->omega);
        } else smooth(g, cyc, 0, cyc->coarsest_iters, 0, x, c0, rhs);
        return;
    }
    smooth(g, cyc, level, cyc->level_iters, 0, x, c0, rhs);
    og_residual_restrict(g, level, x, c0, rhs, cyc->R_restrict);
    og_fas_save(g, level - 1, x, c0, rhs, x0);
    for (int i = 0; i < cyc->mu; i++) og_fas_cycle_step(g, cyc, level - 1, x, rhs, x0, c0);
    og_fas_prolong(g, level, x, x0, cyc->P);
    smooth(g, cyc, level, cyc->level_iters, 1, x, c0, rhs);
}

void og_fas_cycle(og_grid* g, const og_cycle* cyc, int x, int rhs, int x0, int c0)
{
    memset(g->ch[x], 0, sizeof(float) * (size_t)g->N * 512);
    og_fas_cycle_step(g, cyc, g->max_level, x, rhs, x0, c0);
}

static void coarsest_solve(og_grid* g, const og_cycle* cyc, int x, int c0, int rhs)
{
    if (cyc->coarsest_direct) coarsest_direct(g, x, c0, rhs);
    else if (cyc->smoother == 0) {
        og_gauss_seidel(g, cyc->coarsest_iters / 2, 0, 0, x, c0, rhs, cyc->omega);
        og_gauss_seidel(g, cyc->coarsest_iters / 2, 1, 0, x, c0, rhs, cyc->omega);
    } else smooth(g, cyc, 0, cyc->coarsest_iters, 0, x, c0, rhs);
}

/* x = 0 on every tile; rhs = f on LEAF|NONLEAF tiles, 0 on GHOST tiles
 * (the initial launch of VCycle / muCycle, AMGSolver.cu:1062-1078) */
static void nonfas_init(og_grid* g, int x, int f, int rhs)
{
    for (int t = 0; t < g->N; t++) {
        size_t b = (size_t)t * 512;
        memset(&g->ch[x][b], 0, 512 * sizeof(float));
        if (g->t_type[t] & OG_GHOST) memset(&g->ch[rhs][b], 0, 512 * sizeof(float));
        else memcpy(&g->ch[rhs][b], &g->ch[f][b], 512 * sizeof(float));
    }
}

void og_vcycle(og_grid* g, const og_cycle* cyc, int x, int f, int rhs, int c0)
{
    nonfas_init(g, x, f, rhs);
    for (int l = g->max_level; l > 0; l--) {
        smooth(g, cyc, l, cyc->level_iters, 0, x, c0, rhs);
        og_residual_restrict(g, l, x, c0, rhs, cyc->R_restrict);
    }
    coarsest_solve(g, cyc, x, c0, rhs);
    for (int l = 1; l <= g->max_level; l++) {
        og_prolong_add(g, l, x, cyc->P);
        smooth(g, cyc, l, cyc->level_iters, 1, x, c0, rhs);
    }
}

static void mu_step(og_grid* g, const og_cycle* cyc, int level, int x, int f, int rhs, int c0)
{
    if (level == 0) {
        coarsest_solve(g, cyc, x, c0, rhs);
        return;
    }
    smooth(g, cyc, level, cyc->level_iters, 0, x, c0, rhs);
    og_residual_restrict(g, level, x, c0, rhs, cyc->R_restrict);
    og_deduct_flux(g, level - 1, x, c0, f, rhs);
    og_set_zero_tiles(g, level - 1, OG_NONLEAF, x);
    for (int i = 0; i < cyc->mu; i++) mu_step(g, cyc, level - 1, x, f, rhs, c0);
    og_prolong_add(g, level, x, cyc->P);
    smooth(g, cyc, level, cyc->level_iters, 1, x, c0, rhs);
}

void og_mu_cycle(og_grid* g, const og_cycle* cyc, int x, int f, int rhs, int c0)
{
    nonfas_init(g, x, f, rhs);
    mu_step(g, cyc, g->max_level, x, f, rhs, c0);
}

void og_fas_fcycle(og_grid* g, const og_cycle* cyc, int x, int rhs, int x0, int c0)
{
    memset(g->ch[x], 0, sizeof(float) * (size_t)g->N * 512);
    for (int l = g->max_level; l > 0; l--) og_restrict_source(g, l, rhs, cyc->R_restrict);
    coarsest_solve(g, cyc, x, c0, rhs);
    for (int l = 1; l <= g->max_level; l++) {
        og_prolong_add(g, l, x, cyc->P);
        og_fas_cycle_step(g, cyc, l, x, rhs, x0, c0);
    }
}

void og_precondition(og_grid* g, const og_cycle* cyc, int z, int r, int s, int c0)
{
    switch (cyc->kind) {
    case 1: og_fas_fcycle(g, cyc, z, r, s, c0); break;
    case 2: og_vcycle(g, cyc, z, r, s, c0); break;
    case 3: og_mu_cycle(g, cyc, z, r, s, c0); break;
    default: og_fas_cycle(g, cyc, z, r, s, c0); break;
    }
}

og_solve_out og_fas_mu_solve(og_grid* g, const og_cycle* cyc, int c0, int max_iters, double tol)
{
    enum { X = 0, B = 1, X0 = 2 };
    og_solve_out out = { 0, 0.0 };
    memset(g->ch[X], 0, sizeof(float) * (size_t)g->N * 512);
    double rhs_norm = sqrt(og_dot(g, B, B));
    double thr = tol * rhs_norm;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    double res = rhs_norm;
    int i;
    for (i = 0; i < max_iters; i++) {
        og_fas_cycle_step(g, cyc, g->max_level, X, B, X0, c0);
        og_full_neg_lap(g, X, c0, X0);
        for (int t = 0; t < g->N; t++) {
            if (!(g->t_type[t] & OG_LEAF)) continue;
            for (int v = 0; v < 512; v++) {
                size_t idx = (size_t)t * 512 + v;
                g->ch[X0][idx] = g->ch[B][idx] - g->ch[X0][idx];
            }
        }
        res = sqrt(og_dot(g, X0, X0));
        if (res < thr) break;
    }
    out.iters = i;
    out.rel_res = res / rhs_norm;
    return out;
}

/* AMGSolver::solve (AMGSolver.cu:1399-1549), sync_stride = 1.  flexible: flexible
 * CG, beta = z.(r - r_old) / gamma_old = -alpha z.Ap / gamma_old (Polak-Ribiere;
 * the reference uses Fletcher-Reeves, AMGSolver.cu:1520); the cycle's scratch
 * is then channel 7 so that Ap survives it */
og_solve_out og_pcg_ex(og_grid* g, const og_cycle* cyc, int c0, int max_iters, double tol, int pure_neumann, int flexible)
{
    enum { X = 0, R = 1, P = 2, AP = 3, Z = 4 };
    og_solve_out out = { 0, 0.0 };
    for (int t = 0; t < g->N; t++)
        if (g->t_type[t] & OG_LEAF) memset(&g->ch[X][(size_t)t * 512], 0, 512 * sizeof(float));
    double rhs2 = og_dot(g, R, R);
    if (rhs2 == 0) return out;
    double thr2 = tol * tol * rhs2;
    if (thr2 < 2.2250738585072014e-308) thr2 = 2.2250738585072014e-308;
    const int SCR = flexible ? 7 : AP;
    if (pure_neumann) og_recenter(g, R);
    og_precondition(g, cyc, Z, R, SCR, c0);
    for (int t = 0; t < g->N; t++)
        if (g->t_type[t] & OG_LEAF) memcpy(&g->ch[P][(size_t)t * 512], &g->ch[Z][(size_t)t * 512], 512 * sizeof(float));
    double gamma = og_dot(g, Z, R);
    double rr = rhs2;
    int i;
    for (i = 0; i < max_iters; i++) {
        og_full_neg_lap(g, P, c0, AP);
        double fp = og_dot(g, P, AP);
        double alpha = gamma / fp;
        float a = (float)alpha;
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
        og_precondition(g, cyc, Z, R, SCR, c0);
        double gamma_old = gamma;
        double beta;
        if (flexible) {
            gamma = og_dot(g, Z, R);
            beta = -alpha * og_dot(g, Z, AP) / gamma_old;
        } else {
            gamma = og_dot(g, R, Z);
            beta = gamma / gamma_old;
        }
        float b = (float)beta;
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

og_solve_out og_pcg(og_grid* g, const og_cycle* cyc, int c0, int max_iters, double tol, int pure_neumann)
{
    return og_pcg_ex(g, cyc, c0, max_iters, tol, pure_neumann, 0);
}
