// cmg_internal.hpp -- state shared by the C ABI and the device code.
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

#include "cmg_host.hpp"

namespace cmg {

// scalar slots of the solver's device scalar block
enum Scalar : int {
    S_DONE = 0,      // 0 running, 1 converged, 2 breakdown, 3 non-finite
    S_ITER = 1,      // completed PCG iterations
    S_GAMMA = 2, S_GAMMA_OLD = 3, S_ALPHA = 4, S_BETA = 5, S_FP = 6,
    S_RR = 7, S_RHS2 = 8, S_THR2 = 9, S_MEAN = 10, S_TMP = 11,
    S_MAXIT = 12,
    S_STRIDE = 13,   // convergence tested when S_ITER % S_STRIDE == 0; 0 = never
    S_XPEND = 14,    // 1: the last x/r update's x += alpha p waits for its p update
    S_POLL = 15,     // stopping-test reductions (FIN_RR) run in this solve, skipped ones included
    S_COUNT = 16
};

// a launch list: tile ids in a spatially coherent order, padded for XCD
// round-robin dispatch; entries < 0 are skipped by the kernels
struct LaunchList {
    int32_t* d = nullptr;
    int n = 0;        // workgroups to launch
    int tiles = 0;    // real tiles in the list
    int4* meta = nullptr;   // prolongation lists: per entry (tile, parent, octant | type << 8, 0)
    // per entry 7 int4: (tile, level, type | octant << 8, parent), the tile's 32-byte halo record,
    // its ghost-parent record and its children (-1 tile on padding entries); read by the one-tile
    // stencil kernels
    int4* rec = nullptr;
    std::vector<int32_t> order;   // the entries on the host (-1: padding)
};

struct DevGrid {
    int N = 0;
    int num_levels = 0;
    int max_level = -1;
    float h0 = 0;
    std::vector<int32_t> level_start;
    hipStream_t stream = nullptr;
    // topology
    uint8_t* ttype = nullptr;     // N
    uint8_t* toct = nullptr;      // N, bit a = coord[a] & 1
    int32_t* tlevel = nullptr;    // N
    int32_t* nbr = nullptr;       // 6N
    int32_t* parent = nullptr;    // N
    int32_t* child = nullptr;     // 8N
    int32_t* tcoord = nullptr;    // 3N tile coordinates (device assembly)
    unsigned int* acount = nullptr;   // device assembly: closed cells found
    int32_t* halo = nullptr;      // 8N: per tile the 6 face neighbours, their tile types packed in bytes, own type in byte 7.3
    int32_t* ghalo = nullptr;     // 8N: per tile the parents of its face neighbours (-1: none), octants packed 3 bits a face
    uint8_t* ctype = nullptr;     // 512N
    float* ch[kChannels] = {};
    // launch lists
    LaunchList leafs;                   // absolute LEAF tiles
    LaunchList all;                     // every tile
    LaunchList subtree[kMaxLevels];     // LEAF | NONLEAF at the level
    LaunchList nonleaf[kMaxLevels];
    LaunchList leafat[kMaxLevels];      // LEAF tiles at the level
    LaunchList prolong[kMaxLevels];     // LEAF | NONLEAF | GHOST at the level
    // every tile of the level has a LEAF or NONLEAF parent, so the FAS prolongation to the
    // level can be done by the parents' last post-sweep (each parent writes its children)
    bool push_ok[kMaxLevels] = {};
    bool root_alone = false;   // level 0 is one LEAF or NONLEAF tile without neighbour tiles
    bool bottom_ok = false;    // root_alone, and level 1 is its 8 children, LEAF or NONLEAF
    int bottom_root = -1;      // the root; by octant its children: tile, type, subtree[1] list entry,
    int bottom_tile[8] = {};   // and their children (8 per child, -1 none; bit 8 o + c of
    int bottom_type[8] = {};   // bottom_ghost: that grandchild is GHOST)
    int bottom_entry[8] = {};
    int bottom_child[64] = {};
    unsigned long long bottom_ghost = 0;
    LaunchList ghosts;                  // all GHOST tiles
    LaunchList leafghost;               // all LEAF and GHOST tiles
    LaunchList cyclezero;               // GHOST tiles and LEAF tiles below the finest level
    LaunchList cyclezero_pull;          // the same without the finest GHOST tiles (a cycle that pulls its
                                        // finest prolongation never writes them)
    LaunchList finest_ghosts;           // GHOST tiles of the finest level
    int max_list = 0;
    int wg_cap = 2048;                  // persistent grid cap: 8 workgroups per CU
    int derive = 0;                     // set by a solve whose operator passed the diagonal check
    // reduction scratch for the op-level entry points
    double* partial = nullptr;
    double* scal = nullptr;
    // the scalar block of launches outside a solve: its stop flag is never set
    double* idle = nullptr;
    // region-of-interest flags and the tile minima / maxima they came from (N each, on first use)
    uint8_t* roi = nullptr;
    uint8_t* roi_locked = nullptr;
    float* roi_min = nullptr;
    float* roi_max = nullptr;
};

void dev_upload(DevGrid& d, const Topology& topo);
void dev_free(DevGrid& d);
float* dev_channel(DevGrid& d, int c);   // allocates (zeroed) on first use

// octree construction on the device (cmg_build.hip); each leaves the host topology exactly as
// Topology::compress + spawn_ghosts, iterative_refine and finalize would
void device_rebuild(Topology& topo);
void device_iterative_refine(Topology& topo, const cmg_target& t);
void device_finalize(Topology& topo);
// device_rebuild, device_iterative_refine and device_finalize over one device copy of the tables
void device_build(Topology& topo, const cmg_target& t);

}  // namespace cmg

struct cmg_grid {
    std::unique_ptr<cmg::Topology> topo;
    cmg::HostCells cells;
    cmg::DevGrid dev;
    bool uploaded = false;
    bool types_on_device = false;
    int assembled_c0 = -1;
    int generation = 0;   // bumped by every (re)finalization: solvers made before it are stale
    // bumped by every entry point that may write the channel / the cell types; a solver keeps what
    // it derives from its operator (FP16 copies, the diagonal check, the level-0 inverse) while
    // the generations of its coefficient channels and of the types are unchanged
    uint64_t ch_gen[cmg::kChannels] = {};
    uint64_t types_gen = 0;
    bool host_types_stale = false;   // the device types were written on the device (the host copy follows on demand)
};

struct cmg_solver;
struct cmg_trainer;

namespace cmg {
inline void touch_channel(cmg_grid* g, int c)
{
    if (g && c >= 0 && c < kChannels) ++g->ch_gen[c];
}
inline void touch_channels(cmg_grid* g, int c, int n)
{
    for (int i = 0; i < n; ++i) touch_channel(g, c + i);
}
inline void touch_types(cmg_grid* g)
{
    if (g) ++g->types_gen;
}
inline void touch_all(cmg_grid* g)
{
    touch_channels(g, 0, kChannels);
    touch_types(g);
}
// the grid a solver runs on
cmg_grid* solver_grid(const cmg_solver* s);
void set_error(const std::string& msg);
// uploads topology, cell types and host-assembled channels on first device use
void ensure_device(cmg_grid* g);
// device entry points used by the C ABI (cmg_device.hip)
cmg_status op_full_neg_lap(cmg_grid* g, int x, int c0, int y);
cmg_status op_propagate_ghosts(cmg_grid* g, int ch);
cmg_status op_accumulate_leafs(cmg_grid* g, int ch);
cmg_status op_gs_color(cmg_grid* g, int level, int x, int c0, int rhs, int color, double omega);
cmg_status op_jacobi(cmg_grid* g, int level, int x_in, int x_out, int gch, int c0, int rhs, float omega);
cmg_status op_residual_restrict(cmg_grid* g, int level, int x, int c0, int rhs, float Rr);
cmg_status op_fas_save(cmg_grid* g, int level, int x, int c0, int rhs, int x0);
cmg_status op_fas_prolong(cmg_grid* g, int level, int x, int x0, float P);
cmg_status op_dot(cmg_grid* g, int a, int b, double* out);
cmg_status op_norm(cmg_grid* g, int order, int ch, int volume_weighted, int cell_types, double* out);
// op_norm with no allocation (cmg_noalloc.hip)
cmg_status op_norm_noalloc(cmg_grid* g, int order, int ch, int volume_weighted, int cell_types, double* out);
cmg_status op_divergence(cmg_grid* g, int u0, int ch, int c0);
cmg_status op_add_gradient(cmg_grid* g, int x, int c0, int u0, int weighted);
cmg_status op_mix_velocity(cmg_grid* g, int c0, int uf0, int us0, int um0);
cmg_status solver_create(cmg_grid* g, int c0, float Rm, float Rr, float P, cmg_solver** out);
void solver_destroy(cmg_solver* s);
cmg_status solver_solve(cmg_solver* s, const cmg_solve_opts* o, cmg_solve_stats* st);
cmg_status solver_cycle(cmg_solver* s, const cmg_solve_opts* o, int x, int rhs, int x0);
cmg_status solver_mg_solve(cmg_solver* s, const cmg_solve_opts* o, cmg_solve_stats* st);
cmg_status solver_create_geo(cmg_grid* g, int family, float one_over_alpha, float P, cmg_solver** out);
cmg_status solver_damped_jacobi(cmg_solver* s, const cmg_solve_opts* o, cmg_solve_stats* st);
cmg_status solver_apply(cmg_solver* s, int x, int y);
// learned tile smoother
constexpr int kNetWeights = 7776, kNetVolume = 1000;
cmg_status solver_set_net(cmg_solver* s, const float* w);
cmg_status solver_net_infer(cmg_solver* s, int r, int dch, int y);
cmg_status solver_learned_counts(cmg_solver* s, uint64_t* judged, uint64_t* kept);
cmg_status mfma_probe(int K, const uint16_t* A, const uint16_t* B, float* C);
cmg_status trainer_create(int capacity, uint64_t seed, cmg_trainer** out);
void trainer_destroy(cmg_trainer* t);
int trainer_size(const cmg_trainer* t);
cmg_status trainer_collect(cmg_trainer* t, cmg_grid* g, int r_ch, int d_ch, int e_ch, int* appended);
cmg_status trainer_add_tiles(cmg_trainer* t, const float* x, const float* e, int n);
cmg_status trainer_weights(cmg_trainer* t, float* get, const float* set);
cmg_status trainer_loss_grad(cmg_trainer* t, int first, int n, int bf16, double* loss, float* grad);
cmg_status trainer_step(cmg_trainer* t, int batch, float lr, int bf16, double* loss);
cmg_status solver_get_x64(cmg_solver* s, double* host);
cmg_status solver_coded_levels(cmg_solver* s, uint32_t* mask, uint32_t* wide);
cmg_status bench_kernel(cmg_solver* s, int kernel, int reps, double* avg_ms, int* tiles);
cmg_status op_copy(cmg_grid* g, int src, int dst);
}  // namespace cmg
