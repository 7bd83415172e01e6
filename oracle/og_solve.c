/*
 * og_solve.c -- FP64 assembled oracle of the composite LEAF operator, Krylov
 * solvers, seeded input generators.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The CSR rows restate AMGFullNegativeLaplacianOnLeafs (src/AMGSolver.cu:589-632)
 * symbolically: PropagateToChildren makes a ghost cell equal its parent cell,
 * AccumulateToParentsOneStep makes a NONLEAF cell the mean of its 8 children,
 * and LoadAMGLaplacianTileData (AMGSolver.cu:419-543) fills the halo by the
 * same-level / ghost-extrapolation / empty rules.
 */
#include "og_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- seeded generators ------------------------------------------------------ */

uint64_t og_splitmix64(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* top 24 bits -> [0,1) exactly representable, mapped to [-1,1) */
float og_uniform_pm1(uint64_t* s)
{
    uint32_t m = (uint32_t)(og_splitmix64(s) >> 40);
    return (float)m * (1.0f / 8388608.0f) - 1.0f;
}

/* Box-Muller on two 24-bit uniforms in (0,1], cosine branch, float */
float og_normal(uint64_t* s)
{
    uint32_t a = (uint32_t)(og_splitmix64(s) >> 40), b = (uint32_t)(og_splitmix64(s) >> 40);
    double u1 = ((double)a + 1.0) / 16777216.0, u2 = (double)b / 16777216.0;
    return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

void og_rhs_uniform(og_grid* g, int ch, uint64_t seed, int zero_mean)
{
    uint64_t s = seed;
    double sum = 0, cnt = 0;
    float* B = g->ch[ch];
    memset(B, 0, sizeof(float) * (size_t)g->N * 512);
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            if (!(g->ctype[idx] & OG_INTERIOR)) continue;
            B[idx] = og_uniform_pm1(&s);
            sum += B[idx]; cnt += 1;
        }
    }
    if (!zero_mean || cnt == 0) return;
    double mean = sum / cnt;
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            if (g->ctype[idx] & OG_INTERIOR) B[idx] = (float)((double)B[idx] - mean);
        }
    }
}

/* AccumulateFacesToParentsOneStep(LEAF, 1/4, overwrite, all cell types)
 * (PoissonGrid.cu:523-586): the parent face of each 2x2 block of LEAF child
 * faces = 1/4 of their sum, summed (jj, kk) over the rotated face axes */
static void faces_to_parents(og_grid* g, int u0)
{
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        int lvl = g->t_level[t];
        const int32_t* c = &g->t_coord[3 * t];
        for (int q = 0; q < 64; q++) {
            int fl[3] = { q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2 };
            int off, p = og_voxel(g, lvl - 1, (c[0] * 8 + fl[0]) >> 1, (c[1] * 8 + fl[1]) >> 1, (c[2] * 8 + fl[2]) >> 1, &off);
            if (p < 0) continue;
            for (int a = 0; a < 3; a++) {
                float sum = 0;
                for (int jj = 0; jj < 2; jj++)
                    for (int kk = 0; kk < 2; kk++) {
                        int f1[3] = { fl[0], fl[1], fl[2] };
                        f1[(a + 1) % 3] += jj;
                        f1[(a + 2) % 3] += kk;
                        sum += g->ch[u0 + a][(size_t)t * 512 + og_vi(f1[0], f1[1], f1[2])];
                    }
                g->ch[u0 + a][(size_t)p * 512 + off] = sum * 0.25f;
            }
        }
    }
}

/* AMGVolumeWeightedDivergenceWithoutCoeffOnLeafs (AMGSolver.cu:750-788), c0 < 0:
 *   b = sum over faces of (-u0 h h on the - side, u1 h h on the + side);
 * AMGVolumeWeightedDivergenceOnLeafs (AMGSolver.cu:790-835), c0 >= 0:
 *   b = sum of (u0 c0 h on the - side, -u1 c1 h on the + side), c from the face's cell;
 * ghost faces = parent faces, NONLEAF faces = 1/4 sum of the 4 child faces;
 * faces in the order x-, x+, y-, y+, z-, z+; b = 0 off INTERIOR cells */
void og_divergence(og_grid* g, int u0, int ch, int c0)
{
    for (int a = 0; a < 3; a++) og_propagate_ghosts(g, u0 + a);
    faces_to_parents(g, u0);
    float* B = g->ch[ch];
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        int lvl = g->t_level[t];
        float h = og_h(g, lvl);
        const int32_t* c = &g->t_coord[3 * t];
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            int gi[3] = { c[0] * 8 + (v >> 6), c[1] * 8 + ((v >> 3) & 7), c[2] * 8 + (v & 7) };
            float sum = 0;
            for (int a = 0; a < 3; a++)
                for (int sg = -1; sg <= 1; sg += 2) {
                    float uu0 = g->ch[u0 + a][idx];
                    int n[3] = { gi[0], gi[1], gi[2] };
                    n[a] += sg;
                    int off, nt = og_voxel(g, lvl, n[0], n[1], n[2], &off);
                    float uu1 = nt < 0 ? 0.0f : g->ch[u0 + a][(size_t)nt * 512 + off];
                    if (c0 < 0) {
                        sum += sg == -1 ? (-uu0 * h) * h : (uu1 * h) * h;
                    } else {
                        float k0 = g->ch[c0 + a][idx];
                        float k1 = nt < 0 ? 0.0f : g->ch[c0 + a][(size_t)nt * 512 + off];
                        sum += sg == -1 ? (uu0 * k0) * h : (-uu1 * k1) * h;
                    }
                }
            B[idx] = (g->ctype[idx] & OG_INTERIOR) ? sum : 0.0f;
        }
    }
}

/* b = volume-weighted divergence of i.i.d. N(0,1) negative-face velocities
 * written to channels u0..u0+2 on LEAF tiles (0 elsewhere) */
void og_rhs_divergence(og_grid* g, int ch, int u0, uint64_t seed)
{
    uint64_t s = seed;
    for (int a = 0; a < 3; a++) memset(g->ch[u0 + a], 0, sizeof(float) * (size_t)g->N * 512);
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        for (int v = 0; v < 512; v++)
            for (int a = 0; a < 3; a++) g->ch[u0 + a][(size_t)t * 512 + v] = og_normal(&s);
    }
    memset(g->ch[ch], 0, sizeof(float) * (size_t)g->N * 512);
    og_divergence(g, u0, ch, -1);
}

/* ---- FP64 CSR of the composite operator ------------------------------------- */

typedef struct { int32_t col; double val; } og_ent;
typedef struct { og_ent* e; int n, cap; } og_row;

static void row_add(og_row* r, int32_t col, double val)
{
    if (col < 0 || val == 0.0) return;
    for (int i = 0; i < r->n; i++)
        if (r->e[i].col == col) { r->e[i].val += val; return; }
    if (r->n == r->cap) { r->cap = r->cap ? 2 * r->cap : 16; r->e = (og_ent*)realloc(r->e, sizeof(og_ent) * r->cap); }
    r->e[r->n].col = col; r->e[r->n].val = val; r->n++;
}

static int cmp_ent(const void* a, const void* b)
{
    int32_t x = ((const og_ent*)a)->col, y = ((const og_ent*)b)->col;
    return (x > y) - (x < y);
}

og_csr* og_assemble_csr(og_grid* g, int c0)
{
    size_t NC = (size_t)g->N * 512;
    int32_t* uid = (int32_t*)malloc(sizeof(int32_t) * NC);
    int n = 0;
    for (int t = 0; t < g->N; t++)
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            uid[idx] = ((g->t_type[t] & OG_LEAF) && g->ctype[idx] == OG_INTERIOR) ? n++ : -1;
        }
    og_csr* A = (og_csr*)calloc(1, sizeof(og_csr));
    A->n = n;
    A->cell = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    for (size_t idx = 0; idx < NC; idx++) if (uid[idx] >= 0) A->cell[uid[idx]] = (int32_t)idx;
    og_row* rows = (og_row*)calloc((size_t)n + 1, sizeof(og_row));
    const float* C[4] = { g->ch[c0], g->ch[c0 + 1], g->ch[c0 + 2], g->ch[c0 + 3] };

    #pragma omp parallel for schedule(dynamic, 64)
    for (int r = 0; r < n; r++) {
        og_row* row = &rows[r];
        size_t idx = (size_t)A->cell[r];
        int t = (int)(idx / 512), v = (int)(idx % 512);
        int lvl = g->t_level[t];
        float h = og_h(g, lvl);
        int l[3] = { v >> 6, (v >> 3) & 7, v & 7 };
        double D = C[3][idx];
        if (D == 0) continue;
        row_add(row, r, D);
        for (int a = 0; a < 3; a++)
            for (int sg = -1; sg <= 1; sg += 2) {
                int nl[3] = { l[0], l[1], l[2] };
                nl[a] += sg;
                int inside = nl[a] >= 0 && nl[a] < 8;
                double off;
                if (sg == -1) off = C[a][idx];
                else if (inside) off = C[a][(size_t)t * 512 + og_vi(nl[0], nl[1], nl[2])];
                else {
                    int nt = g->nbr[6 * t + a + 3];
                    if (nt < 0) off = h;  /* x = 0 there */
                    else { int m[3] = { nl[0], nl[1], nl[2] }; m[a] = 0; off = C[a][(size_t)nt * 512 + og_vi(m[0], m[1], m[2])]; }
                }
                if (inside) {
                    size_t ni = (size_t)t * 512 + og_vi(nl[0], nl[1], nl[2]);
                    if (g->ctype[ni] == OG_INTERIOR) row_add(row, uid[ni], off);
                    continue;
                }
                int nt = g->nbr[6 * t + (sg == -1 ? a : a + 3)];
                if (nt < 0) continue;
                int m[3] = { nl[0], nl[1], nl[2] };
                m[a] = sg == -1 ? 7 : 0;
                size_t ni = (size_t)nt * 512 + og_vi(m[0], m[1], m[2]);
                if (g->t_type[nt] & OG_GHOST) {
                    /* vg = x_F + 0.5 * (x_parent - mean of F's aligned 2^3 block) */
                    row_add(row, r, off);
                    int pg[3], off_p;
                    const int32_t* cc = &g->t_coord[3 * nt];
                    pg[0] = (cc[0] * 8 + m[0]) >> 1; pg[1] = (cc[1] * 8 + m[1]) >> 1; pg[2] = (cc[2] * 8 + m[2]) >> 1;
                    int pt = og_voxel(g, lvl - 1, pg[0], pg[1], pg[2], &off_p);
                    if (pt >= 0) {
                        size_t pi = (size_t)pt * 512 + off_p;
                        if (g->t_type[pt] & OG_LEAF && g->ctype[pi] == OG_INTERIOR) row_add(row, uid[pi], 0.5 * off);
                    }
                    for (int di = 0; di < 2; di++)
                        for (int dj = 0; dj < 2; dj++)
                            for (int dk = 0; dk < 2; dk++) {
                                size_t bi = (size_t)t * 512 + og_vi(l[0] ^ di, l[1] ^ dj, l[2] ^ dk);
                                if (g->ctype[bi] == OG_INTERIOR) row_add(row, uid[bi], -0.5 * off / 8.0);
                            }
                } else if (g->t_type[nt] & OG_NONLEAF) {
                    if (g->ctype[ni] != OG_INTERIOR) continue;
                    /* NONLEAF cell = mean of its 8 children (LEAF tile under 2:1 grading) */
                    const int32_t* cc = &g->t_coord[3 * nt];
                    int gg[3] = { cc[0] * 8 + m[0], cc[1] * 8 + m[1], cc[2] * 8 + m[2] };
                    for (int di = 0; di < 2; di++)
                        for (int dj = 0; dj < 2; dj++)
                            for (int dk = 0; dk < 2; dk++) {
                                int off_c, ct = og_voxel(g, lvl + 1, 2 * gg[0] + di, 2 * gg[1] + dj, 2 * gg[2] + dk, &off_c);
                                if (ct < 0) continue;
                                size_t ci = (size_t)ct * 512 + off_c;
                                if ((g->t_type[ct] & OG_LEAF) && g->ctype[ci] == OG_INTERIOR) row_add(row, uid[ci], off / 8.0);
                            }
                } else {
                    if (g->ctype[ni] == OG_INTERIOR) row_add(row, uid[ni], off);
                }
            }
        qsort(row->e, row->n, sizeof(og_ent), cmp_ent);
    }
    A->rowptr = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    A->rowptr[0] = 0;
    for (int r = 0; r < n; r++) A->rowptr[r + 1] = A->rowptr[r] + rows[r].n;
    A->nnz = A->rowptr[n];
    A->col = (int32_t*)malloc(sizeof(int32_t) * (A->nnz + 1));
    A->val = (double*)malloc(sizeof(double) * (A->nnz + 1));
    for (int r = 0; r < n; r++) {
        for (int i = 0; i < rows[r].n; i++) { A->col[A->rowptr[r] + i] = rows[r].e[i].col; A->val[A->rowptr[r] + i] = rows[r].e[i].val; }
        free(rows[r].e);
    }
    free(rows); free(uid);
    return A;
}

void og_csr_free(og_csr* A)
{
    if (!A) return;
    free(A->rowptr); free(A->col); free(A->val); free(A->cell); free(A);
}

void og_csr_matvec(const og_csr* A, const double* x, double* y)
{
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < A->n; r++) {
        double s = 0;
        for (int64_t k = A->rowptr[r]; k < A->rowptr[r + 1]; k++) s += A->val[k] * x[A->col[k]];
        y[r] = s;
    }
}

double og_csr_asym(const og_csr* A)
{
    double mx = 0, amax = 0;
    for (int r = 0; r < A->n; r++)
        for (int64_t k = A->rowptr[r]; k < A->rowptr[r + 1]; k++) {
            int c = A->col[k];
            double v = A->val[k], vt = 0;
            for (int64_t q = A->rowptr[c]; q < A->rowptr[c + 1]; q++) if (A->col[q] == r) { vt = A->val[q]; break; }
            if (fabs(v - vt) > mx) mx = fabs(v - vt);
            if (fabs(v) > amax) amax = fabs(v);
        }
    return amax > 0 ? mx / amax : 0;
}

static double vdot(int n, const double* a, const double* b)
{
    double s = 0;
    #pragma omp parallel for reduction(+:s) schedule(static)
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int og_bicgstab(const og_csr* A, const double* b, double* x, double tol, int max_iters, double* rel_res, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int n = A->n;
    double *r = malloc(sizeof(double) * n), *r0 = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n);
    double *v = malloc(sizeof(double) * n), *s = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
    og_csr_matvec(A, x, r);
    for (int i = 0; i < n; i++) { r[i] = b[i] - r[i]; r0[i] = r[i]; p[i] = 0; v[i] = 0; }
    double bn = sqrt(vdot(n, b, b));
    if (bn == 0) bn = 1;
    double rho = 1, alpha = 1, omega = 1;
    int it;
    double res = sqrt(vdot(n, r, r)) / bn;
    for (it = 0; it < max_iters && res > tol; it++) {
        double rho1 = vdot(n, r0, r);
        double beta = (rho1 / rho) * (alpha / omega);
        rho = rho1;
        #pragma omp parallel for
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        og_csr_matvec(A, p, v);
        alpha = rho / vdot(n, r0, v);
        #pragma omp parallel for
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        og_csr_matvec(A, s, t);
        double tt = vdot(n, t, t);
        omega = tt > 0 ? vdot(n, t, s) / tt : 0;
        #pragma omp parallel for
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i] + omega * s[i]; r[i] = s[i] - omega * t[i]; }
        res = sqrt(vdot(n, r, r)) / bn;
    }
    /* true residual */
    og_csr_matvec(A, x, r);
    for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
    *rel_res = sqrt(vdot(n, r, r)) / bn;
    free(r); free(r0); free(p); free(v); free(s); free(t);
    return it;
}

int og_cg(const og_csr* A, const double* b, double* x, double tol, int max_iters, double* rel_res, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int n = A->n;
    double *r = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
    og_csr_matvec(A, x, r);
    #pragma omp parallel for
    for (int i = 0; i < n; i++) { r[i] = b[i] - r[i]; p[i] = r[i]; }
    double bn2 = vdot(n, b, b);
    if (bn2 == 0) bn2 = 1;
    double rr = vdot(n, r, r);
    int it;
    for (it = 0; it < max_iters && rr > tol * tol * bn2; it++) {
        og_csr_matvec(A, p, q);
        double alpha = rr / vdot(n, p, q);
        #pragma omp parallel for
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        double rr1 = vdot(n, r, r);
        double beta = rr1 / rr;
        rr = rr1;
        #pragma omp parallel for
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
    }
    *rel_res = sqrt(rr / bn2);
    free(r); free(p); free(q);
    return it;
}

int og_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
