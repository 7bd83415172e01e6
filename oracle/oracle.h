/*
 * oracle.h -- CPU restatement of the octree multigrid Poisson path of
 * wang-mengdi/Cirrus-amg.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: it is the checker that tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg compare the HIP path against.
 *
 * Every function restates one reference function; the citation next to each
 * declaration names the reference file:line it follows.  Floating-point
 * semantics are spelled out exactly (operation order, explicit fmaf where the
 * device uses a fused multiply-add, float vs double intermediates) so that the
 * per-kernel comparisons against the device can be bit-exact.
 *
 * Data model (the reference's HADeviceGrid + PoissonTile, src/HAGrid.h:356-1326,
 * src/PoissonTile.h:9-175):
 *   - per-level open-addressing hash tables of tiles; a tile is an 8^3 block of
 *     cells at level l with spacing h0 / 2^l;
 *   - "all tiles" order = concatenation of the per-level compressed arrays,
 *     each compressed array being the non-empty hash slots in slot order;
 *   - per-tile cell data: 15 float channels of 512 cells and 512 cell-type
 *     bytes, local offset (i<<6)|(j<<3)|k with i the x index.
 */
#ifndef CMG_ORACLE_H
#define CMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OG_LEAF = 1, OG_NONLEAF = 2, OG_GHOST = 4 };          /* src/HATileInfo.h:10 */
enum { OG_DIRICHLET = 1, OG_INTERIOR = 2, OG_NEUMANN = 4 };  /* src/PoissonTile.h:7 */
enum { OG_MAX_LEVELS = 12, OG_NUM_CHANNELS = 15 };

/* level-target / cell-type presets */
enum {
    OG_TARGET_UNIFORM = 0,      /* tests/TestGrids.cuh:33-41  */
    OG_TARGET_CENTER_POINT = 1, /* tests/TestGrids.cuh:43-56  */
    OG_TARGET_SPHERE_EMPTY = 2, /* tests/TestGrids.cuh:58-73  */
    OG_TARGET_SPHERE_SOLID = 3, /* tests/TestGrids.cuh:75-93  */
    OG_TARGET_STAR_EMPTY = 4,   /* tests/TestGrids.cuh:95-116 */
    OG_TARGET_STAR_SOLID = 5,   /* tests/TestGrids.cuh:118-142 */
    OG_TARGET_SDF_BANDS = 6     /* narrow bands around an analytic SDF (sphere/torus) */
};

enum { OG_SDF_SPHERE = 0, OG_SDF_TORUS = 1, OG_SDF_STAR = 2 };

typedef struct og_target {
    int kind;
    int min_level, max_level;
    /* OG_TARGET_SDF_BANDS: sdf kind + params, and band[l] = |d| bound that
     * puts a tile at level l (levels min_level+1..max_level) */
    int sdf_kind;
    float sdf[8];
    float band[OG_MAX_LEVELS];
} og_target;

typedef struct og_grid og_grid;

/* ---- topology: src/HAGrid.h, src/HAAccessor.h --------------------------- */
og_grid* og_create(float h0, int num_levels, const uint32_t* log2_hash);     /* HAGrid.h:374-399 */
void og_free(og_grid* g);
uint32_t og_hash(int32_t x, int32_t y, int32_t z, uint32_t log2);           /* HAAccessor.h:217-226 */
int og_set_tile(og_grid* g, int level, int i, int j, int k, uint8_t type);  /* HAGrid.h:526-548 */
void og_compress(og_grid* g);                                              /* HAGrid.h:634-653 */
void og_spawn_ghosts(og_grid* g);                                          /* HAGrid.h:856-889 */
int og_refine_step(og_grid* g, const og_target* t, int* counts);           /* HAGrid.h:1088-1171 */
void og_iterative_refine(og_grid* g, const og_target* t);                  /* HAGrid.h:1173-1182 */
/* CreateTestGrid topology: one LEAF root, rebuild, iterativeRefine (tests/TestGrids.cuh:144-156) */
og_grid* og_create_test_grid(const og_target* t, const uint32_t* log2_hash, int num_levels, float h0);
/* builds the all-tile arrays and neighbour/parent/child tables (HAGrid.h:654-685, PoissonGrid.cu:139-152) */
void og_finalize(og_grid* g);

int og_num_tiles(const og_grid* g);
int og_max_level(const og_grid* g);
void og_level_range(const og_grid* g, int32_t* starts /* num_levels+1 */);
void og_get_tiles(const og_grid* g, int32_t* level, int32_t* coord /* N*3 */, uint8_t* type);
void og_get_slots(const og_grid* g, int32_t* slot);
void og_get_neighbors(const og_grid* g, int32_t* nbr /* N*6, x-,y-,z-,x+,y+,z+ */);
void og_get_parents(const og_grid* g, int32_t* parent);
void og_get_children(const og_grid* g, int32_t* child /* N*8 */);
float* og_channel(og_grid* g, int c);       /* N*512 floats */
uint8_t* og_cell_types(og_grid* g);         /* N*512 bytes  */
float og_h(const og_grid* g, int level);

/* ---- cell types ---------------------------------------------------------- */
void og_set_leaf_types_preset(og_grid* g, const og_target* t);              /* TestGrids.cuh:161-167 */
/* wall typing of tests/SolverTests.cu:1880-1897 (QueryEffectiveBoundaryDirection :746-782):
 * INTERIOR leaf cells whose chk_level ancestor touches the box become NEUMANN, or
 * DIRICHLET when the first empty direction found is +y and dirichlet_ytop is set */
void og_set_wall_types(og_grid* g, int chk_level, int dirichlet_ytop);
/* NEUMANN where all 8 cell corners are inside the SDF (SolverTests.cu:1602-1610) */
void og_set_solid_types(og_grid* g, int sdf_kind, const float* sdf);

/* ---- operator assembly: src/AMGSolver.cu:12-318, SolverTests.cu:1595-1714 -- */
void og_coarsen_types_and_coeffs(og_grid* g, int c0, float R);              /* AMGSolver.cu:256-318 */
/* face rule: 0 = reference corner rule FaceFluidRatio (MeshCutCell.cu:305-408),
 *            1 = 4x4 supersampled open fraction with the same 0.1 floor */
void og_cutcell_coeffs(og_grid* g, int sdf_kind, const float* sdf, int face_rule, int c0, float R);
void og_prepare_system(og_grid* g, int c0, float R);                        /* AMGSolver.cu:89-254 */
float og_face_fluid_ratio(float p0, float p1, float p2, float p3);          /* MeshCutCell.cu:305-408 */
float og_sdf(int kind, const float* prm, float x, float y, float z);
/* 4x4 supersampled open fraction of the negative face of cell gi (global cell coords) along axis */
float og_face_supersampled(int kind, const float* prm, const int* gi, int axis, float h);

/* ---- matrix-free operator and multigrid pieces: src/AMGSolver.cu ---------- */
void og_propagate_ghosts(og_grid* g, int ch);                  /* PoissonGrid.cu:377-405 (GHOST, subtree, coarse first) */
void og_accumulate_leafs(og_grid* g, int ch);                  /* PoissonGrid.cu:408-465 (LEAF -> parent, 1/8) */
/* NegativeLaplacianSameLevelAMG128 on absolute LEAF tiles (AMGSolver.cu:545-587) */
void og_neg_lap_leafs(og_grid* g, int x, int c0, int y);
/* AMGFullNegativeLaplacianOnLeafs (AMGSolver.cu:589-632) = propagate + accumulate + apply */
void og_full_neg_lap(og_grid* g, int x, int c0, int y);
/* one red-black colour sweep on subtree-LEAF tiles of `level` (AMGSolver.cu:838-885) */
void og_gs_color(og_grid* g, int level, int x, int c0, int rhs, int color, double omega);
void og_gauss_seidel(og_grid* g, int iters, int order, int level, int x, int c0, int rhs, double omega); /* :887-894 */
/* damped Jacobi sweep on subtree-LEAF tiles of `level`: x_out = x_in + w (b - A x_in)/D
 * (same halo rules as the GS kernel; ghost values always read from channel `ghost_ch`) */
void og_jacobi(og_grid* g, int level, int x_in, int x_out, int ghost_ch, int c0, int rhs, float omega);
void og_residual_restrict(og_grid* g, int level, int x, int c0, int rhs, float Rr);   /* :896-964 */
void og_fas_save(og_grid* g, int level, int x, int c0, int rhs, int x0);              /* :1176-1214 */
void og_fas_prolong(og_grid* g, int level, int x, int x0, float P);                   /* :1216-1255 */
double og_dot(og_grid* g, int a, int b);                       /* PoissonGrid.cu:156-214, LEAF tiles */
void og_recenter(og_grid* g, int ch);                          /* PoissonGrid.cu:889-921 */
double og_norm(og_grid* g, int order, int ch, int volume_weighted, int cell_types); /* PoissonGrid.cu:322-375 */
void og_roi_flags(const og_grid* g, int ch, int subtree_level, int launch_types, int pred, float a, float b,
                  uint8_t* interest, float* tmin, float* tmax);                  /* PoissonGridKernels.cuh:11-54 */
/* projection wrappers: divergence of face velocities u0..u0+2 into ch (c0 < 0:
 * AMGVolumeWeightedDivergenceWithoutCoeffOnLeafs, AMGSolver.cu:750-788; c0 >= 0:
 * AMGVolumeWeightedDivergenceOnLeafs, :790-835), gradient of x added to faces
 * (AMGAddGradientToFace :634-688 / AMGAddFaceWeightedGradientToFace :690-748),
 * fluid/solid face velocity mix (FluidEuler.cu:531-552, solid velocity from channels) */
void og_divergence(og_grid* g, int u0, int ch, int c0);
void og_add_gradient(og_grid* g, int x, int c0, int u0, int weighted);
void og_mix_velocity(og_grid* g, int c0, int uf0, int us0, int um0);

/* multigrid cycle options */
typedef struct og_cycle {
    int mu;                /* 2 = reference W-cycle (FASMuCycle), 1 = V-cycle */
    int smoother;          /* 0 = red-black SOR (omega double), 1 = damped Jacobi */
    double omega;          /* 1.5 (reference) or 2/3 */
    int level_iters;       /* reference: GS iterations per visit (2) ; Jacobi: sweeps per visit */
    int coarsest_iters;    /* reference: 10 => 5+5 RB-GS ; Jacobi: sweeps at level 0 */
    int coarsest_direct;   /* 1 = FP64 dense solve at level 0 */
    float R_restrict, P;
    int tmp_ch;            /* scratch channel for Jacobi ping-pong */
    int kind;              /* preconditioner: 0 FAS mu-cycle, 1 FAS F-cycle, 2 VCycle, 3 muCycle */
} og_cycle;

/* RestrictSourceToCoarse (AMGSolver.cu:1295-1307): parent = coeff * sum of the
 * INTERIOR children of the LEAF/NONLEAF tiles of `level` */
void og_restrict_source(og_grid* g, int level, int ch, float coeff);
/* ProlongateAndUpdateAMG128Kernel (AMGSolver.cu:1022-1057): x += P * parent x */
void og_prolong_add(og_grid* g, int level, int x, float P);
/* DeductFluxOnAbsLeafs128Kernel (AMGSolver.cu:966-993) */
void og_deduct_flux(og_grid* g, int level, int x, int c0, int f, int rhs);
/* SetLevelZeroAbsTiles128Kernel (AMGSolver.cu:995-1016) */
void og_set_zero_tiles(og_grid* g, int level, int tile_types, int x);
/* AMGSolver::VCycle (AMGSolver.cu:1059-1128), non-FAS; f is kept, rhs is scratch */
void og_vcycle(og_grid* g, const og_cycle* cyc, int x, int f, int rhs, int c0);
/* AMGSolver::muCycle / muCycleStep (AMGSolver.cu:1130-1174), non-FAS */
void og_mu_cycle(og_grid* g, const og_cycle* cyc, int x, int f, int rhs, int c0);
/* AMGSolver::FASFCycle (AMGSolver.cu:1358-1387) */
void og_fas_fcycle(og_grid* g, const og_cycle* cyc, int x, int rhs, int x0, int c0);
/* the preconditioner named by cyc->kind: z ~ A^-1 r with scratch channel s */
void og_precondition(og_grid* g, const og_cycle* cyc, int z, int r, int s, int c0);

/* FASMuCycle (AMGSolver.cu:1279-1292): x = 0 on all tiles, then FASMuCycleStep(maxLevel) */
void og_fas_cycle(og_grid* g, const og_cycle* cyc, int x, int rhs, int x0, int c0);
void og_fas_cycle_step(og_grid* g, const og_cycle* cyc, int level, int x, int rhs, int x0, int c0);

typedef struct og_solve_out { int iters; double rel_res; } og_solve_out;
/* AMGSolver::solve (AMGSolver.cu:1399-1549); channels x=0, r=1, p=2, Ap=3, z=4 */
og_solve_out og_pcg(og_grid* g, const og_cycle* cyc, int c0, int max_iters, double tol, int pure_neumann);
/* og_pcg with flexible CG (flexible = 1): beta = -alpha z.Ap / gamma_old */
og_solve_out og_pcg_ex(og_grid* g, const og_cycle* cyc, int c0, int max_iters, double tol, int pure_neumann, int flexible);
/* AMGSolver::FASMuCycleSolve (AMGSolver.cu:1311-1356): x = 0 (channel 0), b in
 * channel 1, residual and FAS x0 in channel 2; returns the reference's
 * (loop index at exit, |r| / |b|) */
og_solve_out og_fas_mu_solve(og_grid* g, const og_cycle* cyc, int c0, int max_iters, double tol);

/* ---- geometric multigrid solvers (src/GMGSolver.cu, src/CMGSolver.cu) ------ */
enum { OG_GEO_GMG = 0, OG_GEO_CMG = 1 };
/* the family's operator on every tile whose subtree type at subtree_level (-1:
 * own type) is in `types` (and at that level when subtree_level >= 0) */
void og_geo_neg_lap_tiles(og_grid* g, int mode, int subtree_level, int types, int x, int y, int diag);
/* FullNegativeLaplacian (GMGSolver.cu:167-171) / ConservativeFullNegativeLaplacian (CMGSolver.cu:209-225) */
void og_geo_full_neg_lap(og_grid* g, int mode, int x, int y, int diag);
void og_geo_gs_color(og_grid* g, int mode, int level, int x, int rhs, int color);          /* GMG :173-206, CMG :228-264 */
void og_geo_residual_restrict(og_grid* g, int mode, int level, int x, int rhs, float one_over_alpha); /* :215-275 / :273-330 */
void og_geo_prolong(og_grid* g, int mode, int level, int x, float P);                       /* :277-307 / :338-369 */
void og_geo_vcycle(og_grid* g, int mode, int x, int f, int rhs, int level_iters, int coarsest_iters,
                   float one_over_alpha, float P);                                          /* :310-387 / :372-424 */
og_solve_out og_geo_pcg(og_grid* g, int mode, int level_iters, int coarsest_iters, float one_over_alpha, float P,
                        int max_iters, double tol, int pure_neumann);                       /* :472-607 / :428-601 */
og_solve_out og_gmg_damped_jacobi(og_grid* g, int max_iters, double tol, float omega);     /* GMGSolver.cu:390-470 */

/* ---- learned tile smoother (og_learn.c; no reference counterpart) ------------ */
enum { OG_NET_W = 7776 };   /* 16*27 + 16*16*27 + 16*27 weights, no biases */
void og_net_forward(const double* w, const double* x, double* y);          /* one 10^3 tile */
double og_net_loss_grad(const double* w, int n, const float* x, const float* e, double* grad);
void og_adam_step(double* w, double* m, double* v, const double* g, int n, int t, double lr, double b1, double b2,
                  double eps);

/* ---- FP64 assembled oracle of the composite LEAF operator ----------------- */
typedef struct og_csr {
    int n;                 /* number of unknowns = LEAF INTERIOR cells            */
    int64_t nnz;
    int64_t* rowptr; int32_t* col; double* val;
    int32_t* cell;         /* unknown -> tile*512+cell                             */
} og_csr;
og_csr* og_assemble_csr(og_grid* g, int c0);   /* rows of AMGFullNegativeLaplacianOnLeafs */
void og_csr_free(og_csr* A);
void og_csr_matvec(const og_csr* A, const double* x, double* y);
/* BiCGSTAB in FP64 on the CSR; returns iterations, writes rel residual */
int og_bicgstab(const og_csr* A, const double* b, double* x, double tol, int max_iters, double* rel_res, int nthreads);
/* plain threaded CG in FP64 (the north star's "assembled CSR + threaded CG" baseline) */
int og_cg(const og_csr* A, const double* b, double* x, double tol, int max_iters, double* rel_res, int nthreads);
double og_csr_asym(const og_csr* A);           /* max |A_ij - A_ji| / max |A_ij| */

/* ---- seeded inputs shared with the product and the tests ------------------ */
uint64_t og_splitmix64(uint64_t* state);
float og_uniform_pm1(uint64_t* state);         /* 24-bit mantissa uniform in [-1,1) */
float og_normal(uint64_t* state);              /* Box-Muller, float */
/* b on LEAF INTERIOR cells: i.i.d. U[-1,1) (seed), then mean removed in double */
void og_rhs_uniform(og_grid* g, int ch, uint64_t seed, int zero_mean);
/* b = volume-weighted divergence of i.i.d. N(0,1) negative-face velocities on LEAF
 * cells (AMGVolumeWeightedDivergenceWithoutCoeffOnLeafs, AMGSolver.cu:750-788);
 * u channels u0..u0+2 are written */
void og_rhs_divergence(og_grid* g, int ch, int u0, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
