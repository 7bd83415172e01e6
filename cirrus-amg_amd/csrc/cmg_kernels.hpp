// cmg_kernels.hpp -- kernel-side views and launcher declarations.
#pragma once

#include "cmg_internal.hpp"

namespace cmg {

// read-only topology view passed by value to every kernel
struct TopoView {
    const uint8_t* ttype;
    const uint8_t* toct;
    const int32_t* tlevel;
    const int32_t* nbr;
    const int32_t* parent;
    const int32_t* child;
    const uint8_t* ctype;
    const int32_t* halo;    // 8 per tile: face neighbours, packed neighbour tile types
    const int32_t* ghalo;   // 8 per tile: parents of the face neighbours, packed octants
    float h0;
    int derive;   // LEAF-tile diagonals are derived from the face coefficients (verified equal)
};

// the four operator channels c0..c0+3; q (optional): FP16 copies of c[a] / h
// (h = the tile's spacing) that the cycle kernels (damped Jacobi, residual and
// restriction, FAS save) read in place of c when set.
// w, lut, coded, wide (optional, exact): on a level whose bit is set in `coded`, every
// cell's face coefficients c[0..2] are FP32 values of that level's table
// lut[level * kWideCodes + code].  One-byte codes (bit clear in `wide`): codes 0..254,
// entry 255 holds the level's tile spacing, and w holds per cell its type (byte 0) and
// the codes of c[0], c[1], c[2] (bytes 1..3).  Ten-bit codes (bit set in `wide`, for
// levels with 256 to 1023 distinct values): codes 0..1022, entry 1023 the spacing, and w
// holds the type as 0..3 (none, DIRICHLET, INTERIOR, NEUMANN; bits 0-1) and the codes in
// bits 2-11, 12-21 and 22-31.  The kernels that read them compute the same values as
// from c (the table holds the FP32 bits).  zw (optional, with w): per tile 128 words, the
// words of the cells across its -z face (the z-neighbour's k = 7 plane) then across its +z
// face (k = 0), 64 each in (i, j) order, so that a tile reads them contiguously instead of
// one 4-byte word per 32-byte sector (the tile's own cells where it has no z-neighbour).
struct Coef {
    const float* c[4];
    const uint16_t* q[4];
    const uint32_t* w;
    const float* lut;
    uint32_t coded;
    uint32_t wide;
    const uint32_t* zw;
};
constexpr int kCodes = 256;        // table entries of a one-byte-coded level
constexpr int kWideCodes = 1024;   // table entries of a ten-bit-coded level; the table's stride per level

inline TopoView view(const DevGrid& d)
{
    return TopoView{d.ttype, d.toct, d.tlevel, d.nbr, d.parent, d.child, d.ctype, d.halo, d.ghalo, d.h0, d.derive};
}

enum FinishKind { FIN_SUM = 0, FIN_GAMMA0 = 1, FIN_ALPHA = 2, FIN_RR = 3, FIN_BETA = 4, FIN_MEAN = 5, FIN_BETA_FLEX = 6,
                  FIN_RHS2 = 7 };

// launchers that write per-workgroup partials return how many they wrote
int launch_apply(const DevGrid& d, const float* X, Coef C, float* Y, double* partial, const double* S, hipStream_t st);
void launch_residual64(const DevGrid& d, const double* X, Coef C, const double* B, double* R64, float* R32, double* partial,
                       hipStream_t st);
void launch_gs_color(const DevGrid& d, int level, float* X, Coef C, const float* B, int color, double omega, const double* S,
                     hipStream_t st);
// partial: one b.x_out sum per list entry (one-tile form); returns their number
// X0: the level's FAS save on its NONLEAF tiles in the same pass (B updated in place, x -> X0)
// push: instead, the FAS prolongation to level + 1 in the same pass (a last post-sweep that
// writes the cycle's x; X0 is then the FAS-saved x it reads; P the prolongation weight)
// pull 1 / 2: a post-sweep that pulls the prolongation into this level in (the first one:
// own, same-level and GHOST neighbour cells; the later ones: GHOST neighbour cells), so that
// fas_prolong_v need not run; X0 is the FAS-saved x, P the prolongation weight.  The level's
// GHOST tiles then keep their x (the cycle's output there is not the prolongated one).
int launch_jacobi(const DevGrid& d, int level, const float* Xin, float* Xout, const float* G, Coef C, const float* B, float w,
                  const double* S, hipStream_t st, double* partial = nullptr, float* X0 = nullptr, bool push = false,
                  float P = 0.0f, int pull = 0);
// the prolongation to level + 1 can ride on level's last Jacobi sweep (every tile of level + 1
// has a LEAF or NONLEAF parent, one-tile Jacobi form)
bool jacobi_push_available(const DevGrid& d, int level);
// Dq (optional): the diagonal as FP16 bits of D / h instead of D
void launch_jacobi_zero(const DevGrid& d, int level, float* Xout, const float* D, const float* B, float w, const double* S,
                        hipStream_t st, const uint16_t* Dq = nullptr);
void launch_residual_restrict(const DevGrid& d, int level, float* X, Coef C, float* B, float Rr, const double* S, hipStream_t st);
void launch_fas_save(const DevGrid& d, int level, const float* X, Coef C, float* B, float* X0, const double* S, hipStream_t st);
void launch_fas_prolong(const DevGrid& d, int level, float* X, const float* X0, float P, const double* S, hipStream_t st);
void launch_propagate_ghosts(const DevGrid& d, float* X, hipStream_t st);
void launch_accumulate_leafs(const DevGrid& d, float* X, hipStream_t st);
void launch_fill_all(const DevGrid& d, float* X, float value, const double* S, hipStream_t st);
void launch_fill_list(const DevGrid& d, const LaunchList& l, float* X, float value, hipStream_t st,
                      const double* S = nullptr);
void launch_copy_flat(const DevGrid& d, const float* A, float* B, size_t n, hipStream_t st);
// q[a] = fp16(c[a] / h) on every tile; counts non-finite results in *bad (set to 0 first)
void launch_coef_to_half(const DevGrid& d, Coef C, uint16_t* const q[4], unsigned int* bad, hipStream_t st);
void launch_copy_list(const DevGrid& d, const LaunchList& l, const float* A, float* B, const double* S, hipStream_t st);
int launch_dot(const DevGrid& d, const LaunchList& l, const float* A, const float* B, double* partial, const double* S,
               hipStream_t st);
// partial = sum A.B, partial2 = sum A.C per workgroup (one pass over A)
int launch_dot2(const DevGrid& d, const LaunchList& l, const float* A, const float* B, const float* C, double* partial,
                double* partial2, const double* S, hipStream_t st);
// X == nullptr: x += alpha p deferred to the p update (launch_update_p with X), and after a solve's
// last iteration to launch_x_pending (which runs only when it is pending)
int launch_update_xr(const DevGrid& d, float* X, float* R, const float* P, const float* AP, double* partial, double* S,
                     hipStream_t st, float* Z0 = nullptr, const float* D = nullptr, float w = 0.0f,
                     const uint16_t* Dq = nullptr);
void launch_update_p(const DevGrid& d, float* P, const float* Z, double* S, hipStream_t st, float* X = nullptr);
void launch_x_pending(const DevGrid& d, float* X, const float* P, const double* S, hipStream_t st);
void launch_sum_count(const DevGrid& d, const float* X, double* psum, double* pcnt, const double* S, hipStream_t st);
void launch_sub_mean(const DevGrid& d, float* X, const double* S, hipStream_t st);
void launch_dot64(const LaunchList& l, const double* A, const double* B, double* partial, hipStream_t st);
void launch_axpy(const DevGrid& d, const LaunchList& l, float* Y, const float* X, float a, hipStream_t st);
void launch_update_x64(const DevGrid& d, double* X64, const float* E, hipStream_t st);
void launch_to_double(const DevGrid& d, const LaunchList& l, const float* A, double* B, bool masked, hipStream_t st);
void launch_to_float(const LaunchList& l, const double* A, float* B, hipStream_t st);
// sig (FIN_RR only; host-mapped coherent memory, may be null): receives S_POLL * 8 + S_DONE,
// the count of the solve's FIN_RR launches and the stop code, as one 64-bit word
void launch_finish(const double* partial, const double* partial2, int n, double* S, int kind, hipStream_t st,
                   unsigned long long* sig = nullptr);
// non-FAS cycles and FASFCycle / FASMuCycleSolve
void launch_restrict_source(const DevGrid& d, int level, float* X, float coeff, const double* S, hipStream_t st);
// leaf_only: LEAF|NONLEAF tiles of the level (CMG), else GHOST tiles too
void launch_prolong_add(const DevGrid& d, int level, float* X, float P, const double* S, hipStream_t st,
                        bool leaf_only = false);

// geometric solvers (cmg_geo.hip); mode 0 = GMG, 1 = CMG
void launch_geo_apply(const DevGrid& d, int mode, const LaunchList& l, int subtree_level, const float* X, float* Y, bool diag,
                      const double* S, hipStream_t st);
void launch_geo_gs(const DevGrid& d, int mode, int level, float* X, const float* B, int color, const double* S, hipStream_t st);
void launch_geo_restrict(const DevGrid& d, int mode, int level, const float* X, float* B, float one_over_alpha,
                         const double* S, hipStream_t st);
void launch_accumulate_ghosts_add(const DevGrid& d, float* Y, float coeff, const double* S, hipStream_t st);
void launch_residual_interior(const DevGrid& d, const float* B, float* AX, hipStream_t st);
void launch_djacobi_update(const DevGrid& d, float* X, const float* R, const float* D, float omega, hipStream_t st);
void launch_deduct_flux(const DevGrid& d, int level, const float* X, Coef C, const float* F, float* RHS, const double* S,
                        hipStream_t st);
void launch_nonfas_init(const DevGrid& d, float* X, const float* F, float* RHS, const double* S, hipStream_t st);
void launch_sub_from(const DevGrid& d, const LaunchList& l, const float* B, float* Y, hipStream_t st);

// per-LEAF-tile partials of NormSync (src/PoissonGrid.cu:223-375); returns the partial count
int launch_power_sum(const DevGrid& d, const float* X, int order, bool volume_weighted, int cell_types, double* vsum,
                     double* wsum, hipStream_t st);

// projection wrappers (cmg_project.hip)
void launch_faces_to_parents(const DevGrid& d, float* U0, float* U1, float* U2, hipStream_t st);
void launch_divergence(const DevGrid& d, const float* U0, const float* U1, const float* U2, const Coef* C, float* B,
                       hipStream_t st);
void launch_add_gradient(const DevGrid& d, const float* X, Coef C, float* U0, float* U1, float* U2, bool weighted,
                         hipStream_t st);
void launch_mix_velocity(const DevGrid& d, Coef C, const float* const F[3], const float* const S[3], float* const M[3],
                         hipStream_t st);

// number of LEAF cells whose stored diagonal differs (bitwise) from the one the
// face coefficients give (derived = 0 - sum of the six faces, -h for a missing
// + neighbour, 0 off INTERIOR); partials reduced by the caller
int launch_diag_check(const DevGrid& d, Coef C, double* partial, hipStream_t st);

// learned tile smoother (cmg_learn.hip)
void launch_gather_tiles(const DevGrid& d, const int32_t* tiles, int n, const float* V, const float* D, int mode, float* out,
                         hipStream_t st);
int launch_net_train(const float* X, const float* E, const int32_t* idx, int n, const float* W, float* gpart, double* lpart,
                     int max_wg, bool bf16, hipStream_t st);
void launch_net_reduce(const float* gpart, const double* lpart, int nwg, int n, float* grad, double* loss, hipStream_t st);
void launch_adam(float* W, float* M, float* V, const float* G, int t, float lr, float b1, float b2, float eps, hipStream_t st);
void launch_net_infer(const DevGrid& d, const LaunchList& l, const float* R, const float* D, const float* W, float* Y,
                      const double* S, hipStream_t st);
void launch_mfma_probe(int K, const uint16_t* A, const uint16_t* B, float* C, hipStream_t st);
// learned pre-smoothing acceptance on the finest level: keep the net's correction
// (in Y) on a tile when it lowers the tile's residual, else the first Jacobi sweep
// count (optional): [0] += tiles judged, [1] += tiles that kept the net's correction
void launch_learned_accept(const DevGrid& d, int level, const float* Y, float* X, Coef C, const float* B, float w,
                           const double* S, hipStream_t st, unsigned int* count = nullptr);

// device cut-cell assembly (same operations as the host assembly, bit for bit)
void launch_assemble_solid(const DevGrid& d, int kind, const float* sdf, hipStream_t st);
void launch_assemble_walls(const DevGrid& d, int chk, int ytop, hipStream_t st);
// returns the number of closed cells retyped NEUMANN (synchronises the stream)
int launch_assemble_operator(const DevGrid& d, int c0, float R, int face_rule, int kind, const float* sdf, unsigned int* count,
                             hipStream_t st);

// stencil kernel forms (persistence mask bit 0 apply / 1 Jacobi / 2 restrict, bit 3 stored
// diagonals, bit 4 unfused boundary work, bit 5 no pulled prolongation, bit 6 no coefficient
// codes; threads per tile, 0 for the defaults)
bool set_kernel_forms(int mask, int threads);
// the thread counts the operator, the compact-form kernels and small level launches use now
// (they follow stencil_threads only once a count was chosen): part of a captured graph's identity
int launched_threads_key();
void get_kernel_forms(int* mask, int* threads);
bool derive_allowed();             // forms bit 3 clear
bool fuse_first_sweep_allowed();   // forms bit 4 clear
bool fuse_rz_allowed();            // forms bit 4 clear and the one-tile Jacobi form
bool fuse_fas_save_allowed();      // forms bit 4 clear and the one-tile Jacobi form
bool pull_prolong_allowed();       // forms bits 4 and 5 clear and the one-tile Jacobi form
bool coef_codes_allowed();         // forms bit 6 clear and the one-tile forms of apply, Jacobi and restrict

// exact coefficient codes (Coef::w / lut / coded / wide): ws is a workspace of
// coef_code_words() uint32 words, lut kMaxLevels * kWideCodes floats, W one word per cell.
// Enqueues the build; once the stream has reached it, the kMaxLevels words at
// ws + coef_code_result_offset() hold per level 1 (one-byte codes), 2 (ten-bit codes) or
// 0 (not coded).
size_t coef_code_words();
// Coef::zw from the cell words W (N * 128 words), for every tile
void launch_code_zfaces(const DevGrid& d, const uint32_t* W, uint32_t* ZW, hipStream_t st);
size_t coef_code_result_offset();
void launch_coef_codes(const DevGrid& d, Coef C, uint32_t* ws, float* lut, uint32_t* W, hipStream_t st);

// reference tile images (PoissonTile<float,15>) <-> this grid, device to device (cmg_tiles.hip):
// tile t at ptrs[t] (device array of N addresses) or at base + t * stride; returns the number
// of tiles skipped for a null or unaligned address; synchronises the grid's stream
unsigned int tile_channel_io(DevGrid& d, bool gather, const uint64_t* ptrs, void* base, size_t stride, int tile_channel,
                             float* X);
unsigned int tile_types_io(DevGrid& d, bool gather, const uint64_t* ptrs, void* base, size_t stride);
// MarkRegionOfInterestWithChannelMinAndMax128Kernel (src/PoissonGridKernels.cuh:11-54) on every tile
void launch_roi_flags(const DevGrid& d, const float* X, int subtree_level, int launch_types, int pred, float a, float b,
                      bool calc_locked, uint8_t* interest, uint8_t* locked, float* tmin, float* tmax, hipStream_t st);

// levels 0 and 1 fit one workgroup (level 0 one tile without neighbours, level 1 its 8
// children) and forms bit 4 is clear: FASMuCycleStep at level 1 can run as
// launch_cycle_bottom (phase 3; or phase 1, the coarsest solve, phase 2)
bool cycle_bottom_available(const DevGrid& d);
void launch_cycle_bottom(const DevGrid& d, int phase, float* X, float* X0, float* B, Coef C, float Rr, float P, bool rbgs,
                         int iters, double omega, int mu, int coarsest_iters, bool fas_save1, bool push, double* rz,
                         const double* S, hipStream_t st);
// level 0 is one tile without neighbours (and forms bit 4 is clear): its smoothing
// iterations can run in one workgroup (launch_coarsest_iterate)
bool coarsest_iterate_available(const DevGrid& d);
// rbgs: `iters` red-black iterations from colour `order`, then `iters2` from 1 - order
// (w unused); else `iters` damped-Jacobi sweeps with weight w (omega, iters2, order unused)
void launch_coarsest_iterate(const DevGrid& d, float* X, Coef C, const float* B, bool rbgs, int iters, int iters2, int order,
                             double omega, float w, const double* S, hipStream_t st);
void launch_coarse_direct(const double* inv, const int32_t* cells, int n, const int32_t* zero_cells, int nz, float* X,
                          const float* B, const double* S, hipStream_t st);

}  // namespace cmg
