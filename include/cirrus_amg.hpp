// cirrus_amg.hpp -- header-only C++17 mirror of the reference's grid and solver
// classes over the C ABI of cirrus_amg.h.
//
// A caller written against wang-mengdi/Cirrus-amg keeps its call sites:
//
//   HADeviceGrid grid(h, {16, 16, 16, 16, 16, 16, 18, 16});      // tests/SolverTests.cu:250
//   grid.setTileHost(0, {0, 0, 0}, LEAF);  grid.rebuild();          // src/HAGrid.h:526-548,856-897
//   grid.iterativeRefine(target);                                   // src/HAGrid.h:1173-1182
//   AMGSolver solver(11, 0.5, 1, 1);                                // src/AMGSolver.h:32-42
//   solver.prepareTypesAndCoeffs(grid);                             // src/AMGSolver.cu:1394-1397
//   auto [iters, err] = solver.solve(grid, false, 100, 1e-6, 2, 10, 1, false);   // src/AMGSolver.h:62
//
// Differences a caller sees: tiles are addressed by index in the all-tiles order
// instead of HATileInfo pointers, per-cell data moves through whole-channel host
// arrays (setChannel/getChannel) instead of device lambdas, and failures throw
// cirrus_amg::Error (carrying the cmg_status) where the reference traps.
#ifndef CIRRUS_AMG_HPP
#define CIRRUS_AMG_HPP

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "cirrus_amg.h"

namespace cirrus_amg {

class Error : public std::runtime_error {
public:
    Error(cmg_status st, const std::string& where)
        : std::runtime_error(where + ": " + cmg_status_string(st) + " -- " + cmg_last_error()), status(st)
    {
    }
    cmg_status status;
};

inline void check(cmg_status st, const char* where)
{
    if (st != CMG_OK) throw Error(st, where);
}

enum : uint8_t { LEAF = CMG_LEAF, NONLEAF = CMG_NONLEAF, GHOST = CMG_GHOST };
enum : uint8_t { DIRICHLET = CMG_DIRICHLET, INTERIOR = CMG_INTERIOR, NEUMANN = CMG_NEUMANN };

// Tile::SIZE and the channel map of PoissonTile<float,15> (src/PoissonTile.h:13-18)
struct Tile {
    static constexpr int SIZE = 512;
    static constexpr int x_channel = 0, b_channel = 1, r_channel = 1, p_channel = 2, Ap_channel = 3, z_channel = 4;
};

// HADeviceGrid<Tile> (src/HAGrid.h:356-1326)
class HADeviceGrid {
public:
    HADeviceGrid(float h0, const std::vector<uint32_t>& log2_hash)
    {
        cmg_grid* g = nullptr;
        check(cmg_grid_create(h0, int(log2_hash.size()), log2_hash.data(), &g), "HADeviceGrid");
        h_.reset(g);
    }

    cmg_grid* handle() const { return h_.get(); }

    // setTileHost(level, b_ijk, Tile(), type) (src/HAGrid.h:526-548)
    void setTileHost(int level, const std::array<int32_t, 3>& b_ijk, uint8_t type)
    {
        check(cmg_grid_set_tile(h_.get(), level, b_ijk[0], b_ijk[1], b_ijk[2], type), "setTileHost");
    }
    // compressHost + spawnGhostTiles (src/HAGrid.h:634-653,856-897); on_device: the same
    // steps as kernels (cmg_grid_rebuild_device), with the same result
    void rebuild(bool on_device = false)
    {
        check(on_device ? cmg_grid_rebuild_device(h_.get()) : cmg_grid_rebuild(h_.get()), "rebuild");
    }

    // iterativeRefine(levelTarget) with a host callback evaluated per LEAF tile
    // (src/HAGrid.h:1173-1182); finalize() freezes the topology afterwards
    void iterativeRefine(const std::function<int(int, const std::array<int32_t, 3>&)>& level_target,
                         bool on_device = false)
    {
        cmg_target t{};
        t.kind = CMG_TARGET_CALLBACK;
        t.fn = &HADeviceGrid::trampoline;
        auto fn = level_target;
        t.user = &fn;
        iterativeRefine(t, on_device);
    }
    void iterativeRefine(const cmg_target& preset, bool on_device = false)
    {
        check(on_device ? cmg_grid_iterative_refine_device(h_.get(), &preset) : cmg_grid_iterative_refine(h_.get(), &preset),
              "iterativeRefine");
    }

    // rebuild + iterativeRefine + finalize on the device over one copy of the tables
    void buildOnDevice(const cmg_target& preset) { check(cmg_grid_build_device(h_.get(), &preset), "buildOnDevice"); }

    // all-tiles arrays, CalculateNeighborTiles (src/PoissonGrid.cu:139-152), parent/child links
    void finalize(bool on_device = false)
    {
        check(on_device ? cmg_grid_finalize_device(h_.get()) : cmg_grid_finalize(h_.get()), "finalize");
    }

    int numTotalTiles() const { return cmg_grid_num_tiles(h_.get()); }
    int maxLevel() const { return cmg_grid_max_level(h_.get()); }
    int numTotalLeafTiles() const
    {
        int n = 0;
        for (uint8_t t : tileTypes()) n += (t & LEAF) ? 1 : 0;
        return n;
    }
    std::vector<uint8_t> tileTypes() const
    {
        const size_t n = size_t(numTotalTiles());
        std::vector<int32_t> level(n), coord(3 * n);
        std::vector<uint8_t> type(n);
        check(cmg_grid_get_tiles(h_.get(), level.data(), coord.data(), type.data()), "tiles");
        return type;
    }

    std::vector<float> getChannel(int c) const
    {
        std::vector<float> v(size_t(numTotalTiles()) * Tile::SIZE);
        check(cmg_grid_get_channel(h_.get(), c, v.data()), "getChannel");
        return v;
    }
    void setChannel(int c, const std::vector<float>& v)
    {
        if (v.size() != size_t(numTotalTiles()) * Tile::SIZE) throw std::invalid_argument("setChannel: size");
        check(cmg_grid_set_channel(h_.get(), c, v.data()), "setChannel");
    }
    std::vector<uint8_t> getCellTypes() const
    {
        std::vector<uint8_t> v(size_t(numTotalTiles()) * Tile::SIZE);
        check(cmg_grid_get_cell_types(h_.get(), v.data()), "getCellTypes");
        return v;
    }
    void setCellTypes(const std::vector<uint8_t>& v)
    {
        if (v.size() != size_t(numTotalTiles()) * Tile::SIZE) throw std::invalid_argument("setCellTypes: size");
        check(cmg_grid_set_cell_types(h_.get(), v.data()), "setCellTypes");
    }

private:
    struct Del {
        void operator()(cmg_grid* g) const { cmg_grid_destroy(g); }
    };
    static int trampoline(void* user, int level, const int32_t coord[3])
    {
        auto* fn = static_cast<std::function<int(int, const std::array<int32_t, 3>&)>*>(user);
        return (*fn)(level, {coord[0], coord[1], coord[2]});
    }
    std::unique_ptr<cmg_grid, Del> h_;
};

// AMGSolver (src/AMGSolver.h:19-74).  The public tuning members are the
// reference's; the extra members select this library's extensions.
class AMGSolver {
public:
    float R_matrix_coeff = 1, R_restrict_coeff = 1, prolong_coeff = 1;
    int mu_cycle_repeat_times = 2;   // 2 = FAS W-cycle, 1 = FAS V-cycle
    double omega = 1.5;              // Gauss-Seidel over-relaxation
    int coeff_channel;
    cmg_solve_opts extra;            // smoother, direct coarsest solve, precision, warm start ...
    cmg_solve_stats last{};

    AMGSolver(int _coeff_channel, float _R_matrix_coeff = 1, float _R_restrict_coeff = 1, float _P_coeff = 1)
        : R_matrix_coeff(_R_matrix_coeff), R_restrict_coeff(_R_restrict_coeff), prolong_coeff(_P_coeff),
          coeff_channel(_coeff_channel)
    {
        cmg_solve_opts_default(&extra);
    }

    // prepareTypesAndCoeffs (src/AMGSolver.cu:1394-1397)
    void prepareTypesAndCoeffs(HADeviceGrid& grid)
    {
        check(cmg_prepare_types_and_coeffs(grid.handle(), coeff_channel, R_matrix_coeff), "prepareTypesAndCoeffs");
    }

    // solve -lap(x) = b (src/AMGSolver.cu:1399-1549); b (channel 1) is modified;
    // returns (iterations, relative residual), (max_iters, -1) when sync_stride <= 0
    std::tuple<int, double> solve(HADeviceGrid& grid, bool verbose, int max_iters, double relative_tolerance,
                                  int level_iters, int coarsest_iters, int sync_stride, bool is_pure_neumann)
    {
        cmg_solver* s = solver_for(grid);
        cmg_solve_opts o = extra;
        o.verbose = verbose;
        o.max_iters = max_iters;
        o.relative_tolerance = relative_tolerance;
        o.level_iters = level_iters;
        o.coarsest_iters = coarsest_iters;
        o.sync_stride = sync_stride;
        o.is_pure_neumann = is_pure_neumann;
        o.cycle = mu_cycle_repeat_times == 1 ? CMG_CYCLE_FAS_V : CMG_CYCLE_FAS_W;
        o.omega = omega;
        const cmg_status st = cmg_solve(s, &o, &last);
        if (st != CMG_OK && st != CMG_ERR_NOT_CONVERGED) throw Error(st, "AMGSolver::solve");
        return {last.iters, last.rel_res};
    }

private:
    struct Del {
        void operator()(cmg_solver* s) const { cmg_solver_destroy(s); }
    };
    cmg_solver* solver_for(HADeviceGrid& grid)
    {
        if (!s_ || grid_ != grid.handle()) {
            cmg_solver* s = nullptr;
            check(cmg_solver_create(grid.handle(), coeff_channel, R_matrix_coeff, R_restrict_coeff, prolong_coeff, &s),
                  "AMGSolver");
            s_.reset(s);
            grid_ = grid.handle();
        }
        return s_.get();
    }
    std::unique_ptr<cmg_solver, Del> s_;
    cmg_grid* grid_ = nullptr;
};

}  // namespace cirrus_amg

#endif
