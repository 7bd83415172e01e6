/*
 * og_assembly.c -- oracle restatement of the AMG operator assembly.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows src/AMGSolver.cu:12-318 (NegativeLaplacianCoeff,
 * CoarsenOffDiagCoefficientsOneStepKernel, PrepareLaplacianSystemFrom...,
 * CoarsenTypesAndAMGCoeffs), tests/SolverTests.cu:1595-1714
 * (CreateLaplacianSystemWithSolidCut) and cirrus_cutcell/MeshCutCell.cu:262-408
 * (FracInside, FaceFluidRatio).
 */
#include "og_internal.h"

#include <stdlib.h>
#include <string.h>

/* NegativeLaplacianCoeff (src/AMGSolver.cu:12-24) */
static float neg_lap_coeff(float h, uint8_t tt0, uint8_t tt1, uint8_t ct0, uint8_t ct1)
{
    int both_leafs = (tt0 & OG_LEAF) && (tt1 & OG_LEAF);
    int leaf_ghost = ((tt0 & OG_LEAF) && (tt1 & OG_GHOST)) || ((tt0 & OG_GHOST) && (tt1 & OG_LEAF));
    int has_neumann = (ct0 & OG_NEUMANN) || (ct1 & OG_NEUMANN);
    int has_interior = (ct0 & OG_INTERIOR) || (ct1 & OG_INTERIOR);
    return ((both_leafs || leaf_ghost) && !has_neumann && has_interior) ? h : 0.0f;
}

/* FracInside (cirrus_cutcell/MeshCutCell.cu:262-272) */
static float frac_inside(float a, float b)
{
    if (a < 0.0 && b < 0.0) return 0.0f;
    if (a < 0.0 && b >= 0.0) return b / (b - a);
    if (a >= 0.0 && b < 0.0) return a / (a - b);
    return 1.0f;
}

#define HALF_PROD(e1, e2) ((float)(0.5 * (double)(e1) * (double)(e2)))
#define HALF_SUM(e1, e2) ((float)(0.5 * ((double)(e1) + (double)(e2))))
#define ONE_MINUS(f) ((float)(1.0 - (double)(f)))

/* FaceFluidRatio (cirrus_cutcell/MeshCutCell.cu:305-408): 16-case table over the
 * signs of the four face-corner SDF values, double intermediates as in the
 * reference's double literals, with the 0.1 floor for non-zero ratios */
float og_face_fluid_ratio(float phi0, float phi1, float phi2, float phi3)
{
    float ret;
    int s0 = phi0 < 0, s1 = phi1 < 0, s2 = phi2 < 0, s3 = phi3 < 0;
    if (s0 && s1 && s2 && s3) ret = 0;
    else if (s0 && s1 && s2 && !s3) ret = HALF_PROD(frac_inside(phi3, phi2), frac_inside(phi3, phi0));
    else if (s0 && s1 && !s2 && s3) ret = HALF_PROD(frac_inside(phi2, phi1), frac_inside(phi2, phi3));
    else if (s0 && s1 && !s2 && !s3) ret = HALF_SUM(frac_inside(phi2, phi1), frac_inside(phi3, phi0));
    else if (s0 && !s1 && s2 && s3) ret = HALF_PROD(frac_inside(phi1, phi0), frac_inside(phi1, phi2));
    else if (s0 && !s1 && s2 && !s3) {
        float e1 = frac_inside(phi1, phi0), e2 = frac_inside(phi1, phi2);
        float e3 = frac_inside(phi3, phi2), e4 = frac_inside(phi3, phi0);
        ret = (float)(0.5 * (double)e1 * (double)e2 + 0.5 * (double)e3 * (double)e4);
    }
    else if (s0 && !s1 && !s2 && s3) ret = HALF_SUM(frac_inside(phi1, phi0), frac_inside(phi2, phi3));
    else if (s0 && !s1 && !s2 && !s3) {
        float e1 = ONE_MINUS(frac_inside(phi0, phi1)), e2 = ONE_MINUS(frac_inside(phi0, phi3));
        ret = (float)(1.0 - 0.5 * (double)e1 * (double)e2);
    }
    else if (!s0 && s1 && s2 && s3) ret = HALF_PROD(frac_inside(phi0, phi1), frac_inside(phi0, phi3));
    else if (!s0 && s1 && s2 && !s3) ret = HALF_SUM(frac_inside(phi0, phi1), frac_inside(phi3, phi2));
    else if (!s0 && s1 && !s2 && s3) {
        float e1 = frac_inside(phi0, phi1), e2 = frac_inside(phi0, phi3);
        float e3 = frac_inside(phi2, phi1), e4 = frac_inside(phi2, phi3);
        ret = (float)(0.5 * (double)e1 * (double)e2 + 0.5 * (double)e3 * (double)e4);
    }
    else if (!s0 && s1 && !s2 && !s3) {
        float e1 = ONE_MINUS(frac_inside(phi1, phi0)), e2 = ONE_MINUS(frac_inside(phi1, phi2));
        ret = (float)(1.0 - 0.5 * (double)e1 * (double)e2);
    }
    else if (!s0 && !s1 && s2 && s3) ret = HALF_SUM(frac_inside(phi0, phi3), frac_inside(phi1, phi2));
    else if (!s0 && !s1 && s2 && !s3) {
        float e1 = ONE_MINUS(frac_inside(phi2, phi1)), e2 = ONE_MINUS(frac_inside(phi2, phi3));
        ret = (float)(1.0 - 0.5 * (double)e1 * (double)e2);
    }
    else if (!s0 && !s1 && !s2 && s3) {
        float e1 = ONE_MINUS(frac_inside(phi3, phi0)), e2 = ONE_MINUS(frac_inside(phi3, phi2));
        ret = (float)(1.0 - 0.5 * (double)e1 * (double)e2);
    }
    else ret = 1;
    if (ret < 0.1 && ret != 0) ret = 0.1f;
    return ret;
}

/* 4x4 supersampled open fraction of the negative face of cell gi along axis:
 * samples at sub-face centres, open where the SDF is >= 0; same 0.1 floor.
 * The config's rule (no reference counterpart; the reference's corner rule is
 * og_face_fluid_ratio). */
float og_face_supersampled(int kind, const float* prm, const int* gi, int axis, float h)
{
    int b = (axis + 1) % 3, c = (axis + 2) % 3;
    int open = 0;
    for (int s = 0; s < 4; s++)
        for (int t = 0; t < 4; t++) {
            float p[3];
            p[axis] = (float)gi[axis] * h;
            p[b] = ((float)gi[b] + (0.125f + 0.25f * (float)s)) * h;
            p[c] = ((float)gi[c] + (0.125f + 0.25f * (float)t)) * h;
            if (og_sdf(kind, prm, p[0], p[1], p[2]) >= 0.0f) open++;
        }
    float r = (float)open / 16.0f;
    if (r < 0.1 && r != 0) r = 0.1f;
    return r;
}

/* set GHOST cell types from the parent cell (AMGSolver.cu:266-276) */
static void ghost_types_from_parents(og_grid* g)
{
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_GHOST)) continue;
        int l = g->t_level[t];
        const int32_t* c = &g->t_coord[3 * t];
        for (int v = 0; v < 512; v++) {
            int gi = c[0] * 8 + (v >> 6), gj = c[1] * 8 + ((v >> 3) & 7), gk = c[2] * 8 + (v & 7);
            int off, p = og_voxel(g, l - 1, gi >> 1, gj >> 1, gk >> 1, &off);
            g->ctype[(size_t)t * 512 + v] = p >= 0 ? g->ctype[(size_t)p * 512 + off] : OG_DIRICHLET;
        }
    }
}

/* negative-face coefficients of LEAF/GHOST cells (AMGSolver.cu:278-315 and
 * SolverTests.cu:1627-1709); face_rule -1 = binary, 0 = corner rule, 1 = supersampled */
static void face_coeffs(og_grid* g, int c0, int face_rule, int sdf_kind, const float* sdf)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < g->N; t++) {
        uint8_t tt0 = g->t_type[t];
        int l = g->t_level[t];
        float h = og_h(g, l);
        const int32_t* c = &g->t_coord[3 * t];
        for (int v = 0; v < 512; v++) {
            if (tt0 == OG_NONLEAF) {
                for (int a = 0; a < 3; a++) g->ch[c0 + a][(size_t)t * 512 + v] = 0;
                continue;
            }
            uint8_t ct0 = g->ctype[(size_t)t * 512 + v];
            int gi[3] = { c[0] * 8 + (v >> 6), c[1] * 8 + ((v >> 3) & 7), c[2] * 8 + (v & 7) };
            for (int a = 0; a < 3; a++) {
                int n[3] = { gi[0], gi[1], gi[2] };
                n[a] -= 1;
                int off, nt = og_voxel(g, l, n[0], n[1], n[2], &off);
                uint8_t tt1, ct1;
                if (nt < 0) { tt1 = tt0; ct1 = OG_DIRICHLET; }
                else { tt1 = g->t_type[nt]; ct1 = g->ctype[(size_t)nt * 512 + off]; }
                float coeff = neg_lap_coeff(h, tt0, tt1, ct0, ct1);
                if (coeff != 0.0f && face_rule >= 0) {
                    if (face_rule == 0) {
                        /* cell centre (g+0.5)*h, corners at +-0.5h (SolverTests.cu:1664-1688) */
                        float cc[3], q[4][3];
                        for (int d = 0; d < 3; d++) cc[d] = (float)(gi[d] + 0.5) * h;
                        static const float O[3][4][3] = {
                            { {-0.5f,-0.5f,-0.5f}, {-0.5f,-0.5f,0.5f}, {-0.5f,0.5f,0.5f}, {-0.5f,0.5f,-0.5f} },
                            { {-0.5f,-0.5f,-0.5f}, {-0.5f,-0.5f,0.5f}, {0.5f,-0.5f,0.5f}, {0.5f,-0.5f,-0.5f} },
                            { {-0.5f,-0.5f,-0.5f}, {-0.5f,0.5f,-0.5f}, {0.5f,0.5f,-0.5f}, {0.5f,-0.5f,-0.5f} } };
                        float phi[4];
                        for (int k = 0; k < 4; k++) {
                            for (int d = 0; d < 3; d++) q[k][d] = cc[d] + O[a][k][d] * h;
                            phi[k] = og_sdf(sdf_kind, sdf, q[k][0], q[k][1], q[k][2]);
                        }
                        coeff = h * og_face_fluid_ratio(phi[0], phi[1], phi[2], phi[3]);
                    } else {
                        coeff = h * og_face_supersampled(sdf_kind, sdf, gi, a, h);
                    }
                }
                g->ch[c0 + a][(size_t)t * 512 + v] = -coeff;
            }
        }
    }
}

void og_coarsen_types_and_coeffs(og_grid* g, int c0, float R)
{
    ghost_types_from_parents(g);
    face_coeffs(g, c0, -1, 0, NULL);
    og_prepare_system(g, c0, R);
}

void og_cutcell_coeffs(og_grid* g, int sdf_kind, const float* sdf, int face_rule, int c0, float R)
{
    og_set_solid_types(g, sdf_kind, sdf);
    ghost_types_from_parents(g);
    face_coeffs(g, c0, face_rule, sdf_kind, sdf);
    og_prepare_system(g, c0, R);
}

/* PrepareLaplacianSystemFromLeafAndGhostCellTypesAndFaceCoeffs (AMGSolver.cu:89-254).
 * (a) levels max..1: each tile adds R * (sum of the 4 fine faces on the parent
 *     cell's negative face) into the parent's coefficient.  The reference masks
 *     the fine values by cell type; LEAF/GHOST types always pass and NONLEAF
 *     types are INTERIOR/DIRICHLET once assembled, so the mask passes all cells.
 * (b) levels finest..0: LEAF/GHOST INTERIOR diagonal = -sum of face terms
 *     (+ side from the neighbour, -h when it is missing); NONLEAF diagonal =
 *     R * (sum of child diagonals + twice the faces between INTERIOR children),
 *     NONLEAF type = INTERIOR iff any child is INTERIOR.
 * (c) closed cells: a LEAF INTERIOR cell whose six faces are all closed ends with
 *     a zero diagonal, which the reference rejects (CUDA_ASSERT diag > 0,
 *     AMGSolver.cu:151).  Such a cell is disconnected from the fluid; it is
 *     retyped NEUMANN, its ghost children follow, and NONLEAF types are
 *     re-derived bottom-up.  Coefficients and diagonals do not change. */
void og_prepare_system(og_grid* g, int c0, float R)
{
    for (int i = g->max_level; i > 0; i--) {
        for (int t = g->lstart[i]; t < g->lstart[i + 1]; t++) {
            const int32_t* c = &g->t_coord[3 * t];
            for (int q = 0; q < 64; q++) {
                int fl[3] = { q / 16 * 2, (q / 4) % 4 * 2, q % 4 * 2 };
                int fg[3] = { c[0] * 8 + fl[0], c[1] * 8 + fl[1], c[2] * 8 + fl[2] };
                int coff, ct = og_voxel(g, i - 1, fg[0] >> 1, fg[1] >> 1, fg[2] >> 1, &coff);
                if (ct < 0) continue;
                for (int a = 0; a < 3; a++) {
                    float sum = 0;
                    for (int jj = 0; jj < 2; jj++)
                        for (int kk = 0; kk < 2; kk++) {
                            int f1[3] = { fl[0], fl[1], fl[2] };
                            f1[(a + 1) % 3] += jj;
                            f1[(a + 2) % 3] += kk;
                            sum += g->ch[c0 + a][(size_t)t * 512 + og_vi(f1[0], f1[1], f1[2])];
                        }
                    float* dst = &g->ch[c0 + a][(size_t)ct * 512 + coff];
                    *dst += sum * R;
                }
            }
        }
    }
    for (int lvl = g->num_levels - 1; lvl >= 0; lvl--) {
        #pragma omp parallel for schedule(dynamic, 16)
        for (int t = g->lstart[lvl]; t < g->lstart[lvl + 1]; t++) {
            uint8_t tt = g->t_type[t];
            float h = og_h(g, lvl);
            const int32_t* c = &g->t_coord[3 * t];
            for (int v = 0; v < 512; v++) {
                int gi[3] = { c[0] * 8 + (v >> 6), c[1] * 8 + ((v >> 3) & 7), c[2] * 8 + (v & 7) };
                size_t idx = (size_t)t * 512 + v;
                if (tt & (OG_LEAF | OG_GHOST)) {
                    float diag = 0;
                    if (g->ctype[idx] & OG_INTERIOR) {
                        for (int a = 0; a < 3; a++)
                            for (int s = -1; s <= 1; s += 2) {
                                float c0v = g->ch[c0 + a][idx];
                                int n[3] = { gi[0], gi[1], gi[2] };
                                n[a] += s;
                                int off, nt = og_voxel(g, lvl, n[0], n[1], n[2], &off);
                                float c1v = nt < 0 ? -h : g->ch[c0 + a][(size_t)nt * 512 + off];
                                float face = s == -1 ? c0v : c1v;
                                diag -= face;
                            }
                    }
                    g->ch[c0 + 3][idx] = diag;
                } else {
                    int interior_cnt = 0;
                    uint8_t cty[2][2][2];
                    float cof[3][2][2][2];
                    float diag_sum = 0;
                    for (int ci = 0; ci < 2; ci++)
                        for (int cj = 0; cj < 2; cj++)
                            for (int ck = 0; ck < 2; ck++) {
                                int off, ct = og_voxel(g, lvl + 1, gi[0] * 2 + ci, gi[1] * 2 + cj, gi[2] * 2 + ck, &off);
                                if (ct >= 0) {
                                    size_t cidx = (size_t)ct * 512 + off;
                                    diag_sum += g->ch[c0 + 3][cidx];
                                    interior_cnt += g->ctype[cidx] == OG_INTERIOR;
                                    cty[ci][cj][ck] = g->ctype[cidx];
                                    for (int a = 0; a < 3; a++) cof[a][ci][cj][ck] = g->ch[c0 + a][cidx];
                                } else {
                                    cty[ci][cj][ck] = OG_DIRICHLET;
                                    for (int a = 0; a < 3; a++) cof[a][ci][cj][ck] = 0;
                                }
                            }
                    for (int ci = 0; ci < 2; ci++)
                        for (int cj = 0; cj < 2; cj++)
                            for (int ck = 0; ck < 2; ck++) {
                                if (cty[0][cj][ck] == OG_INTERIOR && cty[1][cj][ck] == OG_INTERIOR) diag_sum += cof[0][1][cj][ck];
                                if (cty[ci][0][ck] == OG_INTERIOR && cty[ci][1][ck] == OG_INTERIOR) diag_sum += cof[1][ci][1][ck];
                                if (cty[ci][cj][0] == OG_INTERIOR && cty[ci][cj][1] == OG_INTERIOR) diag_sum += cof[2][ci][cj][1];
                            }
                    g->ctype[idx] = interior_cnt > 0 ? OG_INTERIOR : OG_DIRICHLET;
                    g->ch[c0 + 3][idx] = diag_sum * R;
                }
            }
        }
    }
    /* (c) closed cells */
    int closed = 0;
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_LEAF)) continue;
        for (int v = 0; v < 512; v++) {
            size_t idx = (size_t)t * 512 + v;
            if (g->ctype[idx] == OG_INTERIOR && g->ch[c0 + 3][idx] == 0.0f) { g->ctype[idx] = OG_NEUMANN; closed++; }
        }
    }
    if (!closed) return;
    for (int t = 0; t < g->N; t++) {
        if (!(g->t_type[t] & OG_GHOST)) continue;
        int l = g->t_level[t];
        const int32_t* c = &g->t_coord[3 * t];
        for (int v = 0; v < 512; v++) {
            int gi = c[0] * 8 + (v >> 6), gj = c[1] * 8 + ((v >> 3) & 7), gk = c[2] * 8 + (v & 7);
            int off, p = og_voxel(g, l - 1, gi >> 1, gj >> 1, gk >> 1, &off);
            g->ctype[(size_t)t * 512 + v] = p >= 0 ? g->ctype[(size_t)p * 512 + off] : OG_DIRICHLET;
        }
    }
    for (int lvl = g->num_levels - 1; lvl >= 0; lvl--)
        for (int t = g->lstart[lvl]; t < g->lstart[lvl + 1]; t++) {
            if (g->t_type[t] & (OG_LEAF | OG_GHOST)) continue;
            const int32_t* c = &g->t_coord[3 * t];
            for (int v = 0; v < 512; v++) {
                int gi[3] = { c[0] * 8 + (v >> 6), c[1] * 8 + ((v >> 3) & 7), c[2] * 8 + (v & 7) };
                int cnt = 0;
                for (int ci = 0; ci < 8; ci++) {
                    int off, ct = og_voxel(g, lvl + 1, gi[0] * 2 + (ci >> 2), gi[1] * 2 + ((ci >> 1) & 1), gi[2] * 2 + (ci & 1), &off);
                    if (ct >= 0 && g->ctype[(size_t)ct * 512 + off] == OG_INTERIOR) cnt++;
                }
                g->ctype[(size_t)t * 512 + v] = cnt > 0 ? OG_INTERIOR : OG_DIRICHLET;
            }
        }
}
