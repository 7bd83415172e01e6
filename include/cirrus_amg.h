/*
 * cirrus_amg.h -- C ABI of the MI355X-native octree multigrid Poisson solve.
 *
 * Drop-in boundary for the pressure-projection solve of wang-mengdi/Cirrus-amg:
 * the grid construction of HADeviceGrid (src/HAGrid.h), the operator assembly
 * of AMGSolver / the cut-cell builder (src/AMGSolver.cu, tests/SolverTests.cu,
 * cirrus_cutcell/MeshCutCell.cu), the matrix-free kernels, and the
 * AMGSolver::solve(grid, ...) entry point (src/AMGSolver.h:62) that
 * FluidEuler::project calls (cirrus_cutcell/FluidEuler.cu:702-706).
 *
 * Plain C types only.  All per-cell arrays are tile-major in the grid's
 * all-tiles order (levels 0..L concatenated, each level in hash-slot order,
 * exactly HADeviceGrid::dAllTiles), 512 cells per tile, local cell offset
 * (i<<6)|(j<<3)|k with i along x (src/HAAccessor.h:32-42).
 *
 * Channels mirror PoissonTile<float,15> (src/PoissonTile.h:13-18,129):
 * 0 = x, 1 = b / r, 2 = p, 3 = Ap (FAS x0), 4 = z, and the operator occupies
 * four consecutive channels c0..c0+3 (3 negative-face coefficients, diagonal).
 *
 * Threading: a grid handle and its solvers are used from one host thread; all
 * device work of a handle runs on the handle's own HIP stream.
 */
#ifndef CIRRUS_AMG_H
#define CIRRUS_AMG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cmg_status {
    CMG_OK = 0,
    CMG_ERR_BAD_ARG = 1,        /* invalid argument / channel / level */
    CMG_ERR_STATE = 2,          /* call out of order (e.g. solve before assembly) */
    CMG_ERR_HIP = 3,            /* a HIP runtime call failed */
    CMG_ERR_NONFINITE = 4,      /* NaN/Inf met in a reduction */
    CMG_ERR_BREAKDOWN = 5,      /* CG breakdown (p.Ap == 0) */
    CMG_ERR_NOT_CONVERGED = 6,  /* max_iters reached above tolerance */
    CMG_ERR_HASH_FULL = 7       /* a level's hash table is full */
} cmg_status;

enum { CMG_LEAF = 1, CMG_NONLEAF = 2, CMG_GHOST = 4 };            /* src/HATileInfo.h:10 */
enum { CMG_DIRICHLET = 1, CMG_INTERIOR = 2, CMG_NEUMANN = 4 };    /* src/PoissonTile.h:7 */

typedef struct cmg_grid cmg_grid;
typedef struct cmg_solver cmg_solver;
typedef struct cmg_trainer cmg_trainer;

/* ---------------------------------------------------------------- grid ---- */

/* HADeviceGrid(h0, levels_log2_hash) (src/HAGrid.h:374-399) */
cmg_status cmg_grid_create(float h0, int num_levels, const uint32_t* log2_hash, cmg_grid** out);
/* ~HADeviceGrid (src/HAGrid.h:453-465) */
void cmg_grid_destroy(cmg_grid* g);
/* setTileHost(level, b_ijk, Tile(), type) (src/HAGrid.h:526-548) */
cmg_status cmg_grid_set_tile(cmg_grid* g, int level, int i, int j, int k, uint8_t type);
/* compressHost + spawnGhostTiles, i.e. rebuild() (src/HAGrid.h:634-653,856-897) */
cmg_status cmg_grid_rebuild(cmg_grid* g);

/* level-target presets for iterativeRefine (tests/TestGrids.cuh:33-142 and the
 * narrow-band SDF targets used by the benchmark configurations) */
typedef enum cmg_target_kind {
    CMG_TARGET_UNIFORM = 0, CMG_TARGET_CENTER_POINT = 1, CMG_TARGET_SPHERE_EMPTY = 2,
    CMG_TARGET_SPHERE_SOLID = 3, CMG_TARGET_STAR_EMPTY = 4, CMG_TARGET_STAR_SOLID = 5,
    CMG_TARGET_SDF_BANDS = 6, CMG_TARGET_CALLBACK = 7
} cmg_target_kind;
enum { CMG_SDF_SPHERE = 0, CMG_SDF_TORUS = 1, CMG_SDF_STAR = 2 };
typedef int (*cmg_level_target_fn)(void* user, int level, const int32_t coord[3]);
typedef struct cmg_target {
    int kind;                 /* cmg_target_kind */
    int min_level, max_level;
    int sdf_kind;             /* CMG_SDF_*; sphere (cx,cy,cz,r), torus (cx,cy,cz,R,r) */
    float sdf[8];
    float band[12];           /* SDF_BANDS: tile goes to level l when |d|lower bound < band[l] */
    cmg_level_target_fn fn;   /* CALLBACK: evaluated on each LEAF tile */
    void* user;
} cmg_target;

/* iterativeRefine(level_target) (src/HAGrid.h:1173-1182; refineStep :1088-1171) */
cmg_status cmg_grid_iterative_refine(cmg_grid* g, const cmg_target* target);
/* freeze the topology: all-tiles arrays, CalculateNeighborTiles
 * (src/PoissonGrid.cu:139-152), parent/child tables, device upload */
cmg_status cmg_grid_finalize(cmg_grid* g);

/* The same three construction steps on the device, leaving the grid exactly as
 * the host forms above do (same hash slots, tile order and tables).  Refine
 * flags, 2:1 grading through ghost children, child and ghost insertion into the
 * level hash tables and slot-order compression run as kernels; the reference
 * drives these steps from the host with a device flag pass per level
 * (refineStep, src/HAGrid.h:1088-1171; spawnGhostTiles, :856-889).  Star and
 * callback level targets are evaluated on the host, one level at a time. */
cmg_status cmg_grid_rebuild_device(cmg_grid* g);
cmg_status cmg_grid_iterative_refine_device(cmg_grid* g, const cmg_target* target);
cmg_status cmg_grid_finalize_device(cmg_grid* g);
/* cmg_grid_rebuild_device, cmg_grid_iterative_refine_device and
 * cmg_grid_finalize_device in one call over one device copy of the tables
 * (CreateTestGrid's sequence after setTileHost, tests/TestGrids.cuh:144-169) */
cmg_status cmg_grid_build_device(cmg_grid* g, const cmg_target* target);

int cmg_grid_num_tiles(const cmg_grid* g);
int cmg_grid_max_level(const cmg_grid* g);
/* hNumTiles prefix sums: starts[l] .. starts[l+1] are the tiles of level l */
cmg_status cmg_grid_level_starts(const cmg_grid* g, int32_t* starts /* num_levels+1 */);
/* HATileInfo per tile in all-tiles order: level, coord (3), type */
cmg_status cmg_grid_get_tiles(const cmg_grid* g, int32_t* level, int32_t* coord, uint8_t* type);
/* hash slot of each tile (defines the order; src/HAGrid.h:638-642) */
cmg_status cmg_grid_get_slots(const cmg_grid* g, int32_t* slot);
/* PoissonTile::mNeighbors as tile indices, order x-,y-,z-,x+,y+,z+, -1 = empty */
cmg_status cmg_grid_get_neighbors(const cmg_grid* g, int32_t* nbr);
cmg_status cmg_grid_get_parents(const cmg_grid* g, int32_t* parent);
cmg_status cmg_grid_get_children(const cmg_grid* g, int32_t* child /* 8 per tile */);

/* ------------------------------------------------------- cell data --------- */

/* PoissonTile::mCellType of every tile (N*512 bytes); the host copy is authoritative
 * and mirrored to the device */
cmg_status cmg_grid_set_cell_types(cmg_grid* g, const uint8_t* types);
cmg_status cmg_grid_get_cell_types(cmg_grid* g, uint8_t* types);
/* host-side result of the last assembly for a channel (no device needed) */
cmg_status cmg_grid_get_host_channel(cmg_grid* g, int channel, float* host);
/* GridCase::type on LEAF tiles (tests/TestGrids.cuh:161-167) */
cmg_status cmg_grid_set_leaf_types_preset(cmg_grid* g, const cmg_target* target);
/* boundary typing of tests/SolverTests.cu:1880-1897 */
cmg_status cmg_grid_set_wall_types(cmg_grid* g, int chk_level, int dirichlet_ytop);
/* one float channel of every tile (N*512 floats), host <-> device */
cmg_status cmg_grid_set_channel(cmg_grid* g, int channel, const float* host);
cmg_status cmg_grid_get_channel(cmg_grid* g, int channel, float* host);
cmg_status cmg_grid_fill_channel(cmg_grid* g, int channel, float value);
/* device-side copy of a whole channel (e.g. to restore a right-hand side) */
cmg_status cmg_grid_copy_channel(cmg_grid* g, int src, int dst);

/* ------------------------------------------------------- device-resident exchange */
/* One channel of every tile (N*512 floats, tile-major, the layout of every per-cell
 * array here) from / to device memory of the caller, device to device; no host copy.
 * The caller's writes to d_src must be complete before the call (e.g. its stream
 * synchronised); the copy has completed when the call returns.  No reference
 * counterpart as an entry point: in the reference the solver reads the tiles'
 * channels in place (PoissonTile::mData, src/PoissonTile.h:35). */
cmg_status cmg_grid_set_channel_device(cmg_grid* g, int channel, const float* d_src);
cmg_status cmg_grid_get_channel_device(cmg_grid* g, int channel, float* d_dst);
/* The reference's own device tiles.  Each tile of this grid's all-tiles order is a
 * PoissonTile<float,15> device image (src/PoissonTile.h:9-175, 44,880 bytes:
 * mData[15][736] floats, mNeighbors[6], mCellType[512], ...) at d_tile_addrs[t] (a
 * device array of N 64-bit addresses: HATileInfo::mTilePtr of dAllTiles,
 * src/HAGrid.h:654-685), or, with d_tile_addrs == NULL, at d_base + t * stride (tiles in
 * one slab; stride a multiple of 16 and at least 44,880).  Gather copies cells 0..511 of
 * tile channel `tile_channel` (0..14) into this grid's `channel`; scatter writes this
 * grid's `channel` into the tiles' `tile_channel` (node entries 512..735 untouched).
 * The types calls move mCellType.  Tiles must be 16-byte aligned; a null or unaligned
 * address returns CMG_ERR_BAD_ARG after the other tiles were copied.  This is how
 * FluidEuler::project (cirrus_cutcell/FluidEuler.cu:688-752) hands its velocities,
 * types and right-hand side to the solve and takes the pressure back without a host
 * round trip. */
cmg_status cmg_grid_gather_tile_channel(cmg_grid* g, const uint64_t* d_tile_addrs, const void* d_base, size_t stride,
                                        int tile_channel, int channel);
cmg_status cmg_grid_scatter_tile_channel(cmg_grid* g, int channel, const uint64_t* d_tile_addrs, void* d_base,
                                         size_t stride, int tile_channel);
cmg_status cmg_grid_gather_tile_types(cmg_grid* g, const uint64_t* d_tile_addrs, const void* d_base, size_t stride);
cmg_status cmg_grid_scatter_tile_types(cmg_grid* g, const uint64_t* d_tile_addrs, void* d_base, size_t stride);

/* ------------------------------------------------------- region of interest */
/* MarkRegionOfInterestWithChannelMinAndMax128Kernel (src/PoissonGridKernels.cuh:11-54)
 * over every tile: a tile whose subtree type (NONLEAF at subtree_level counts as LEAF,
 * src/HATileInfo.h:79-84; -1 = none) is in launch_types gets the min and max of the
 * 512 values of `channel` and the flag pred(min, max); other tiles get 0.  The
 * reference passes the predicate as a device lambda; here it is one of the forms
 * below (its callers use CMG_ROI_MAX_GT with a = 0, src/MarkerParticles.cu:49-61).
 * calc_locked != 0 also sets the locked-refine flags (mIsLockedRefine) to the same
 * value; otherwise they are left as they are.  Flags stay on the device (one byte a
 * tile, all-tiles order) until read. */
enum { CMG_ROI_MAX_GT = 0,        /* max > a */
       CMG_ROI_MIN_LT = 1,        /* min < a */
       CMG_ROI_RANGE_MEETS = 2,   /* [min, max] meets [a, b] */
       CMG_ROI_SPAN_GT = 3 };     /* max - min > a */
cmg_status cmg_mark_region_of_interest(cmg_grid* g, int subtree_level, uint8_t launch_types, int channel, int pred,
                                       float a, float b, int calc_locked);
/* host copies of the flags and of the tile minima / maxima of the last marking (N each;
 * any pointer may be NULL; min / max of unselected tiles are left 0) */
cmg_status cmg_grid_get_roi(cmg_grid* g, uint8_t* interest, uint8_t* locked, float* tile_min, float* tile_max);

/* ------------------------------------------------------- assembly ---------- */

/* AMGSolver::prepareTypesAndCoeffs -> CoarsenTypesAndAMGCoeffs (src/AMGSolver.cu:256-318,1394-1397) */
cmg_status cmg_prepare_types_and_coeffs(cmg_grid* g, int coeff_channel, float R_matrix);
/* CreateLaplacianSystemWithSolidCut with an analytic SDF (tests/SolverTests.cu:1595-1714);
 * cells with all 8 corners inside the SDF become NEUMANN; face_rule -1 = open faces
 * (CoarsenTypesAndAMGCoeffs), 0 = corner FaceFluidRatio (cirrus_cutcell/MeshCutCell.cu:305-408),
 * 1 = 4x4 supersampled open fraction (same 0.1 floor) */
cmg_status cmg_assemble_cutcell(cmg_grid* g, int sdf_kind, const float* sdf, int face_rule,
                                int coeff_channel, float R_matrix);

/* ------------------------------------------------------- inputs ------------ */

/* seeded right-hand sides on LEAF INTERIOR cells (SplitMix64 streams in all-tiles
 * order): i.i.d. U[-1,1) with the mean removed, or the volume-weighted divergence of
 * i.i.d. N(0,1) negative-face velocities written to channels u0..u0+2
 * (AMGVolumeWeightedDivergenceWithoutCoeffOnLeafs, src/AMGSolver.cu:750-788).
 * No reference counterpart for the generators themselves. */
cmg_status cmg_rhs_uniform(cmg_grid* g, int channel, uint64_t seed, int zero_mean);
cmg_status cmg_rhs_divergence(cmg_grid* g, int channel, int u0, uint64_t seed);

/* The same assembly on the device, starting from the grid's current cell types:
 * (wall_chk_level >= 0) the wall typing of cmg_grid_set_wall_types, (sdf != NULL)
 * solid cells as cmg_assemble_cutcell, then ghost types, faces (face_rule as
 * cmg_assemble_cutcell; -1 also without an SDF, = cmg_prepare_types_and_coeffs),
 * coarsening, diagonals, NONLEAF types and the closed-cell fix; bit-identical to
 * the host assembly.  Sphere and torus SDFs.  The cell types are copied back to
 * the host; the coefficients stay on the device.  No reference counterpart as an
 * entry point (the reference assembles on the device in
 * CreateAMGLaplacianSystemWithSolidCutOnNodeSDF, cirrus_cutcell/MeshCutCell.cu:416-543). */
cmg_status cmg_assemble_cutcell_device(cmg_grid* g, int sdf_kind, const float* sdf, int face_rule, int coeff_channel,
                                      float R_matrix, int wall_chk_level, int dirichlet_ytop);

/* ------------------------------------------------------- HAGR blobs -------- */
/* HADeviceGrid::dumpBinaryBlob (src/HAGrid.h:687-758) for PoissonTile<float, 15>:
 * the tiles whose type is in tile_types (CMG_LEAF | CMG_NONLEAF | CMG_GHOST) up to
 * max_level (-1: all), in the all-tiles order, as a 'HAGR' version-1 blob; a tile
 * record carries channels 0..14 (cells 0..511; the node entries 512..735 are 0),
 * the cell types and the face neighbours' coordinates, levels and types.  With
 * out == NULL, *size receives the blob's size; otherwise *size is the buffer's
 * capacity on entry and the bytes written on return. */
cmg_status cmg_grid_dump_blob(cmg_grid* g, uint8_t tile_types, int max_level, uint8_t* out, size_t* size);
/* HADeviceGrid::loadBinaryBlob (src/HAGrid.h:760-853): a new finalized grid with the
 * blob's tiles (its ghosts are tiles of the blob; none is spawned), cell types and
 * every channel that holds a nonzero value.  CMG_ERR_BAD_ARG for a bad magic,
 * version, tile or coordinate size, or a truncated buffer. */
cmg_status cmg_grid_load_blob(const uint8_t* data, size_t size, cmg_grid** out);

/* ------------------------------------------------------- projection -------- */
/* Face velocities: channel u + a holds the velocity through each cell's negative
 * a-face.  Divergence on LEAF tiles into `channel` after ghost faces take their
 * parents' values and NONLEAF faces a quarter of their four child faces (left in
 * u): coeff_channel < 0 = AMGVolumeWeightedDivergenceWithoutCoeffOnLeafs
 * (src/AMGSolver.cu:750-788), else AMGVolumeWeightedDivergenceOnLeafs (:790-835). */
cmg_status cmg_divergence(cmg_grid* g, int u0, int channel, int coeff_channel);
/* u += grad(x) on LEAF faces: weighted = 0 AMGAddGradientToFace (src/AMGSolver.cu:634-688,
 * open faces only), 1 AMGAddFaceWeightedGradientToFace (:690-748, times the fluid ratio);
 * x keeps its propagated ghost / accumulated parent values */
cmg_status cmg_add_gradient(cmg_grid* g, int x, int coeff_channel, int u0, int weighted);
/* FluidEuler::mixFluidAndSolidVelocityOnFaces (cirrus_cutcell/FluidEuler.cu:531-552) on LEAF
 * tiles; the solid face-centre velocity is read from channels u_solid..+2 (the reference
 * evaluates its scene function there) */
cmg_status cmg_mix_velocity(cmg_grid* g, int coeff_channel, int u_fluid, int u_solid, int u_mix);

/* ------------------------------------------------------- solver ------------ */

/* preconditioner cycles of AMGSolver: FAS mu-cycle (FASMuCycle, src/AMGSolver.cu:1257-1292)
 * as W (mu = 2, the reference default) or V (mu = 1); FAS F-cycle (FASFCycle, :1358-1387);
 * the non-FAS VCycle (:1059-1128) and muCycle (:1130-1174) */
typedef enum {
    CMG_CYCLE_FAS_W = 0, CMG_CYCLE_FAS_V = 1, CMG_CYCLE_FAS_F = 2, CMG_CYCLE_VCYCLE = 3, CMG_CYCLE_MU = 4
} cmg_cycle_kind;
typedef enum { CMG_SMOOTH_RBGS = 0, CMG_SMOOTH_JACOBI = 1 } cmg_smoother_kind;
typedef enum { CMG_PREC_FP32 = 0, CMG_PREC_FP64_REFINE = 1 } cmg_precision;

typedef struct cmg_solve_opts {
    /* the eight parameters of AMGSolver::solve (src/AMGSolver.h:62) */
    int verbose;
    int max_iters;
    double relative_tolerance;
    int level_iters;          /* smoothing sweeps (RBGS iterations) per level visit */
    int coarsest_iters;       /* RBGS: iterations split 1/2 + 1/2 ; Jacobi: sweeps */
    int sync_stride;          /* convergence checked every sync_stride iterations on the device */
    int is_pure_neumann;
    /* extensions */
    int cycle;                /* cmg_cycle_kind: FAS_W = reference (mu = 2) */
    int smoother;             /* cmg_smoother_kind */
    double omega;             /* RBGS 1.5 (reference) ; Jacobi 2/3 */
    int coarsest_direct;      /* 1: FP64 dense solve of the level-0 system */
    int precision;            /* cmg_precision */
    double outer_tolerance;   /* FP64 refinement: stop at this true relative residual */
    double inner_tolerance;   /* FP64 refinement: relative tolerance of each inner FP32 solve */
    int warm_start;           /* 1: start from channel x (FP32) instead of 0 */
    int use_graph;            /* 1: capture one PCG iteration in a hipGraph */
    int mu;                   /* coarse-visit repetitions (mu_cycle_repeat_times); 0: 1 for FAS_V, else 2 */
    int flexible;             /* 1: flexible CG (Polak-Ribiere beta) for a varying preconditioner */
    int learned;              /* 1: the learned tile smoother replaces the finest pre-smoothing;
                               * 2: it runs before the usual finest pre-sweeps (FAS cycles;
                               * weights from cmg_solver_set_net; use with flexible) */
    int cycle_coef_fp16;      /* 1: the cycle's smoothing sweeps, residual restrictions and FAS
                               * saves read FP16 copies of the coefficients divided by the tile
                               * spacing (made at each solve; LEAF diagonals are then read, not
                               * derived); the PCG operator and the FP64 level-0 solve keep FP32
                               * coefficients, so the solution and its tolerance are those of
                               * the FP32 operator */
    int no_true_residual;     /* 1: an FP32 solve skips the FP64 recompute of ||b - A x|| / ||b||
                               * after the iterations (stats.true_rel_res = -1); a diagnostic of
                               * this library that AMGSolver::solve does not compute.  0 (the
                               * default): computed */
} cmg_solve_opts;

typedef struct cmg_solve_stats {
    int iters;                /* PCG iterations (summed over refinement steps) */
    int outer_iters;          /* FP64 refinement steps (0 in FP32 mode) */
    double rel_res;           /* recursive relative residual, as AMGSolver::solve returns */
    double true_rel_res;      /* ||b - A x|| / ||b|| recomputed in FP64 on the device (-1 when
                               * no_true_residual skipped it) */
    double ms;                /* device time of the solve (hipEvents on the handle's stream) */
} cmg_solve_stats;

void cmg_solve_opts_default(cmg_solve_opts* o);   /* FluidEuler::project settings */

/* solver families: AMGSolver (assembled coefficients), GMGSolver (src/GMGSolver.h:14-50),
 * CMGSolver (src/CMGSolver.h:12-47) */
enum { CMG_SOLVER_AMG = 0, CMG_SOLVER_GMG = 1, CMG_SOLVER_CMG = 2 };

/* AMGSolver(coeff_channel, R_matrix, R_restrict, P) (src/AMGSolver.h:32-42).
 * A solver holds device buffers sized for the grid as finalized when it was made;
 * once the grid is finalized again (cmg_grid_finalize, _finalize_device,
 * _build_device) its calls return CMG_ERR_STATE and a new solver is needed. */
cmg_status cmg_solver_create(cmg_grid* g, int coeff_channel, float R_matrix, float R_restrict,
                             float P, cmg_solver** out);
/* GMGSolver(one_over_alpha, prolong_coeff) / CMGSolver(one_over_alpha, prolong_coeff)
 * (src/GMGSolver.h:18-26, src/CMGSolver.h:16-24): family = CMG_SOLVER_GMG or CMG_SOLVER_CMG.
 * Their operators come from cell and tile types alone (no coefficient channels).
 * cmg_solve on such a solver runs GMGSolver::solve / CMGSolver::solve
 * (src/GMGSolver.cu:472-607, src/CMGSolver.cu:428-601): PCG with their V-cycle
 * of red-black Gauss-Seidel (level_iters, coarsest_iters); cycle, smoother,
 * omega, coarsest_direct and FP64 refinement do not apply (CMG_ERR_BAD_ARG for
 * precision != FP32), and true_rel_res repeats rel_res. */
cmg_status cmg_geo_solver_create(cmg_grid* g, int family, float one_over_alpha, float prolong_coeff, cmg_solver** out);
void cmg_solver_destroy(cmg_solver* s);
/* AMGSolver::solve (src/AMGSolver.cu:1399-1549): solves -lap(x) = b with b in
 * channel 1 on LEAF INTERIOR cells; x lands in channel 0; channels 1-4 of
 * every tile are clobbered (1 ends as the residual). */
cmg_status cmg_solve(cmg_solver* s, const cmg_solve_opts* opts, cmg_solve_stats* stats);
/* AMGSolver::FASMuCycleSolve (src/AMGSolver.cu:1311-1356): FAS mu-cycles as a
 * stand-alone iteration from x = 0 (channel 0) on b (channel 1, kept), residual
 * in channel 2, until |r| < relative_tolerance |b| or max_iters cycles.  Uses
 * opts' cycle (mu), smoother, omega, level_iters, coarsest_iters /
 * coarsest_direct.  stats->iters is the reference's return value: the loop
 * index at exit (cycles run - 1 when converged). */
cmg_status cmg_mg_solve(cmg_solver* s, const cmg_solve_opts* opts, cmg_solve_stats* stats);
/* GMGSolver::dampedJacobiSolve (src/GMGSolver.cu:390-470) on a GMG solver: x = 0
 * (channel 0), b (channel 1, kept), residual channel 2, diagonal channel 3,
 * x += omega r / D until |r| < relative_tolerance |b| (|b| over every tile);
 * opts' max_iters, relative_tolerance and omega are used; stats->iters is the
 * loop index at exit, as the reference returns it. */
cmg_status cmg_gmg_damped_jacobi_solve(cmg_solver* s, const cmg_solve_opts* opts, cmg_solve_stats* stats);
/* y = A x with the solver's operator on LEAF tiles (AMG: AMGFullNegativeLaplacianOnLeafs;
 * GMG: FullNegativeLaplacian, src/GMGSolver.cu:167-171; CMG: ConservativeFullNegativeLaplacian,
 * src/CMGSolver.cu:209-225); x keeps the propagated ghost / accumulated parent values */
cmg_status cmg_solver_apply(cmg_solver* s, int x, int y);
/* the FP64 solution of the last FP64-refinement solve (N*512 doubles) */
cmg_status cmg_solver_get_x64(cmg_solver* s, double* host);
/* levels (bit l: level l) whose face coefficients the solver's stencil kernels read as
 * exact codes (forms bit 6 of cmg_set_kernel_forms), as built by the last solve, cycle or
 * multigrid solve; 0 before the first, after the coefficient channels or cell types were
 * written since, or with bit 6 set.  wide (may be NULL): the subset with ten-bit codes
 * (256 to 1023 distinct values; the operator reads the FP32 channels on those levels).
 * No reference counterpart. */
cmg_status cmg_solver_coded_levels(cmg_solver* s, uint32_t* mask, uint32_t* wide);

/* ------------------------------------------------------- learned smoother -- */
/* No reference counterpart.  A 3-layer 3D conv net on one halo'd 10^3 tile
 * (1 -> 16 -> 16 -> 1 channels, 3^3 kernels, zero padding, ReLU after layers 1
 * and 2, no biases); 7776 weights W1[16][27], W2[16 out][16 in][27], W3[16][27],
 * tap (di+1)*9 + (dj+1)*3 + (dk+1), volume index (i*10+j)*10+k.  Input: r / D on
 * the INTERIOR cells of the tile and its same-level LEAF neighbours; output: a
 * correction of x on the tile's own cells.  Inference runs on BF16 MFMA with FP32
 * accumulation.  Loss of a tile: the interior 7-point Laplacian of (y - e),
 * squared and summed; a batch takes the mean. */
enum { CMG_NET_WEIGHTS = 7776, CMG_NET_VOLUME = 1000 };
cmg_status cmg_solver_set_net(cmg_solver* s, const float* weights);
/* y = net(r / D) on the finest-level LEAF tiles (channels r, diag, y) */
cmg_status cmg_solver_net_infer(cmg_solver* s, int r, int diag, int y);
/* over the last solve with opts.learned: finest LEAF tiles judged by the learned
 * step and tiles that kept the net's correction (the rest took the Jacobi sweep) */
cmg_status cmg_solver_learned_counts(cmg_solver* s, uint64_t* judged, uint64_t* kept);
/* training: tiles resident on the device, FP32 master weights (seeded initial
 * values), Adam (beta 0.9 / 0.999, eps 1e-8), batches drawn from a seeded
 * per-epoch permutation; bf16 = 1 rounds weights and activations to bfloat16 in
 * the forward pass (FP32 accumulation) */
cmg_status cmg_trainer_create(int capacity_tiles, uint64_t seed, cmg_trainer** out);
void cmg_trainer_destroy(cmg_trainer* t);
/* append the finest-level LEAF tiles of g: input r / D from channels r, diag;
 * target from channel e; *appended = tiles added (capacity permitting) */
cmg_status cmg_trainer_collect(cmg_trainer* t, cmg_grid* g, int r, int diag, int e, int* appended);
cmg_status cmg_trainer_add_tiles(cmg_trainer* t, const float* x, const float* e, int n);
int cmg_trainer_size(const cmg_trainer* t);
cmg_status cmg_trainer_get_weights(cmg_trainer* t, float* w);
cmg_status cmg_trainer_set_weights(cmg_trainer* t, const float* w);
/* mean loss and its gradient over tiles [first, first + n), no update */
cmg_status cmg_trainer_loss_grad(cmg_trainer* t, int first, int n, int bf16, double* loss, float* grad);
/* one Adam step on the next batch; *loss = the batch loss before the step */
cmg_status cmg_trainer_step(cmg_trainer* t, int batch, float lr, int bf16, double* loss);
/* C = A B for a 16 x k bfloat16 A and a k x 16 bfloat16 B (row-major bit
 * patterns, k = 16 or 32) on one MFMA instruction: checks the operand layouts
 * the inference kernel uses */
cmg_status cmg_mfma_probe(int k, const uint16_t* A, const uint16_t* B, float* C);

/* ------------------------------------------------------- kernels ----------- */
/* Individual device operations on channels; each restates one reference function. */

/* AMGFullNegativeLaplacianOnLeafs (src/AMGSolver.cu:589-632); only y on LEAF tiles is defined */
cmg_status cmg_op_full_neg_lap(cmg_grid* g, int x, int coeff_channel, int y);
/* PropagateToChildren(GHOST, subtree) (src/PoissonGrid.cu:377-405) */
cmg_status cmg_op_propagate_ghosts(cmg_grid* g, int channel);
/* AccumulateToParentsOneStep(LEAF, 1/8) (src/PoissonGrid.cu:408-465) */
cmg_status cmg_op_accumulate_leafs(cmg_grid* g, int channel);
/* one colour of NegativeLaplacianAndGaussSeidelSameLevelAMG128Kernel (src/AMGSolver.cu:838-885) */
cmg_status cmg_op_gs_color(cmg_grid* g, int level, int x, int coeff_channel, int rhs, int color, double omega);
/* damped-Jacobi sweep on the same tiles and halo (no reference counterpart; GMGSolver::dampedJacobiSolve
 * src/GMGSolver.cu:390-470 is the closest) */
cmg_status cmg_op_jacobi(cmg_grid* g, int level, int x_in, int x_out, int ghost_channel, int coeff_channel,
                         int rhs, float omega);
/* ResidualAndRestrictAMG128Kernel (src/AMGSolver.cu:896-964) */
cmg_status cmg_op_residual_restrict(cmg_grid* g, int level, int x, int coeff_channel, int rhs, float R_restrict);
/* FASSaveSolutionAndUpdateSource128Kernel (src/AMGSolver.cu:1176-1214) */
cmg_status cmg_op_fas_save(cmg_grid* g, int level, int x, int coeff_channel, int rhs, int x0);
/* FASProlongateAndUpdateAMG128Kernel (src/AMGSolver.cu:1216-1255) */
cmg_status cmg_op_fas_prolong(cmg_grid* g, int level, int x, int x0, float P);
/* Dot over LEAF tiles (src/PoissonGrid.cu:156-214), FP64 */
cmg_status cmg_op_dot(cmg_grid* g, int a, int b, double* out);
/* NormSync (src/PoissonGrid.cu:322-375) over LEAF tiles: order -1 = max |x| (not volume
 * weighted), order p >= 1 = (sum |x|^p w / sum w)^(1/p) over the cells whose type is in
 * cell_types, w = 1 or the cell volume h^3; sums in FP64 */
cmg_status cmg_op_norm(cmg_grid* g, int order, int channel, int volume_weighted, int cell_types, double* out);
/* cmg_op_norm with no allocation of any kind.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and the same result bits as
 * cmg_op_norm.  Its partial sums live in a fixed device array of 1 MiB that each GPU
 * holds until the process ends; at most 65,536 LEAF-list entries (a larger grid returns
 * CMG_ERR_BAD_ARG).  The grid must already be on the device and `channel` allocated
 * there (any earlier device call on the grid and channel does it; CMG_ERR_STATE
 * otherwise).  Calls that overlap on one GPU share the array and race. */
cmg_status cmg_op_norm_noalloc(cmg_grid* g, int order, int channel, int volume_weighted, int cell_types, double* out);
/* the solver's preconditioner cycle on channels (x, rhs, x0): AMG FAS kinds use rhs as
 * the right-hand side and x0 as FAS scratch; the non-FAS kinds and the GMG / CMG
 * V-cycles keep rhs (f) and use x0 as the working right-hand side */
cmg_status cmg_op_cycle(cmg_solver* s, const cmg_solve_opts* opts, int x, int rhs, int x0);

/* ------------------------------------------------------- measurement ------- */
/* Launches one kernel `reps` times back to back on the grid's stream between two
 * hipEvents and returns the average duration in ms, plus the number of tiles one
 * launch processes.  Kernels: 0 = operator on LEAF tiles (+ p.Ap partials),
 * 1 = damped-Jacobi sweep on the finest level, 2 = residual + restriction on the
 * finest level, 3 = PCG x/r update (+ r.r partials), 4 = FAS prolongation on the
 * finest level, 5 = PCG p update, 6 = r.z partials, 7 = first Jacobi sweep from
 * zero on the finest level, 8 = copy of one vector on LEAF tiles, 9 = learned
 * smoother inference on the finest level (needs cmg_solver_set_net; CMG_ERR_STATE
 * otherwise), 10 = flat copy of 1 GiB into a second 1 GiB buffer (the on-box HBM
 * copy ceiling; allocates both buffers for the call; tiles = 2^28 / 512),
 * 11 = the damped-Jacobi sweep of kernel 1 reading the FP16 coefficient copies
 * of cycle_coef_fp16 (made by the call), 12 / 13 / 14 = kernels 1 / 0 / 2 reading
 * the coded coefficients of forms bit 6 (made by the call; CMG_ERR_BAD_ARG when
 * bit 6 is set or the finest level is not coded).
 * Overwrites the solver's work channels (0-4) and scalars.  No reference counterpart (the reference times whole solves
 * with CPUTimer, cirrus_cutcell/FluidEuler.cu:705-707). */
cmg_status cmg_bench_kernel(cmg_solver* s, int kernel, int reps, double* avg_ms, int* tiles);

/* Forms of the stencil kernels (operator, Jacobi sweep, residual + restriction),
 * process-wide.  persistent_mask: bit 0 operator, bit 1 Jacobi, bit 2 residual +
 * restriction; a set bit selects a persistent pipelined kernel (each workgroup
 * walks several tiles, prefetching the next tile's loads), a clear bit one tile
 * per workgroup; bit 3 set keeps the stored diagonal of LEAF cells (by default a
 * solve derives it from the face coefficients, one load fewer per cell, after
 * checking that every stored LEAF diagonal equals the derived one bit for bit);
 * bit 4 set keeps the preconditioner's boundary work in launches of its own (by
 * default PCG iterations with FAS mu-cycles and damped Jacobi fuse the first
 * finest-level sweep from x = 0 into the x/r update, sum r.z over each level's
 * LEAF tiles in that level's last post-sweep, which regroups the FP64 partial
 * sums of r.z, do a coarse level's FAS save in its first pre-sweep and the FAS
 * prolongation to the next finer level in its last post-sweep); bit 5 set keeps
 * the prolongation into the finest level out of the finest post-sweeps (by
 * default a PCG solve's finest post-sweeps compute it on the fly; the cycle's
 * output on GHOST tiles of the finest level, which PCG never reads, then keeps 0);
 * bit 6 set keeps the FP32 coefficient channels in every stencil kernel (by default
 * an AMG solve, cycle or multigrid solve first codes, per level, the face
 * coefficients of every cell as indices into a table of that level's distinct
 * FP32 values: one-byte codes when there are at most 255 of them, ten-bit codes
 * when there are at most 1023 (and the cell types are 0, DIRICHLET, INTERIOR or
 * NEUMANN); the one-tile Jacobi, Gauss-Seidel, residual-restriction and FAS-save
 * kernels, and the one-tile operator on one-byte-coded levels, read the cell's
 * type and three codes as one 32-bit word: the values, and so every
 * result, are those of the FP32 channels; the codes are rebuilt after the
 * coefficient channels or the cell types are written; the persistent forms of
 * bits 0-2 read the FP32 channels).
 * stencil_threads: 128, 256 or 512 threads per tile, or 0 for the library's
 * defaults (CMG_STENCIL_THREADS when set).  Stencil results are
 * bit-identical across forms; PCG scalars differ only by the grouping of FP64
 * partial sums.  Default: mask 0, 256 threads (the Jacobi sweep on FP16 coefficient
 * copies: 128 until a thread count is set here or by CMG_STENCIL_THREADS).  No reference counterpart (the reference
 * kernels have one fixed form, src/AMGSolver.cu:589-632, 838-964).  Returns
 * CMG_ERR_BAD_ARG for an invalid mask or thread count. */
cmg_status cmg_set_kernel_forms(int persistent_mask, int stencil_threads);   /* mask 0..127 */
cmg_status cmg_get_kernel_forms(int* persistent_mask, int* stencil_threads);

/* hipDeviceSynchronize on the current device (timing brackets) */
cmg_status cmg_device_synchronize(void);

/* ------------------------------------------------------- misc -------------- */
const char* cmg_status_string(cmg_status st);
/* last error message of this thread */
const char* cmg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
