// C++ caller of the library through include/cirrus_amg.hpp, written the way the
// reference's solver tests drive AMGSolver (tests/SolverTests.cu:1087-1101,
// FluidEuler::project cirrus_cutcell/FluidEuler.cu:702-706).
//
//   test_amg_wrapper cpu   topology + host assembly through the C++ mirror
//   test_amg_wrapper gpu   adds the solves (needs a GPU)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cirrus_amg.hpp"

using namespace cirrus_amg;

#define EXPECT(cond)                                                              \
    do {                                                                          \
        if (!(cond)) {                                                            \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            return 1;                                                             \
        }                                                                         \
    } while (0)

// 64^3 uniform grid: 8^3 tiles at level 3 under one root tile (h0 = 1/8); on_device:
// the construction steps run as kernels
static void build_uniform(HADeviceGrid& grid, bool on_device = false)
{
    grid.setTileHost(0, {0, 0, 0}, LEAF);
    grid.rebuild(on_device);
    grid.iterativeRefine([](int level, const std::array<int32_t, 3>&) { return level < 3 ? level + 1 : 3; }, on_device);
    grid.finalize(on_device);
    // every cell of every tile INTERIOR; the box faces are Dirichlet through the
    // empty neighbours (src/AMGSolver.cu:301-304)
    grid.setCellTypes(std::vector<uint8_t>(size_t(grid.numTotalTiles()) * Tile::SIZE, INTERIOR));
}

static int cpu_part()
{
    HADeviceGrid grid(0.125f, {16, 16, 16, 16});
    build_uniform(grid);
    EXPECT(grid.numTotalTiles() == 585);
    EXPECT(grid.numTotalLeafTiles() == 512);
    EXPECT(grid.maxLevel() == 3);
    AMGSolver solver(11, 0.5, 1, 1);
    solver.prepareTypesAndCoeffs(grid);
    // finest diagonal of an interior cell = 6 h (h = 1/64) times R_matrix^0
    std::vector<float> diag(size_t(grid.numTotalTiles()) * Tile::SIZE);
    check(cmg_grid_get_host_channel(grid.handle(), 14, diag.data()), "host channel");
    const auto types = grid.tileTypes();
    int checked = 0;
    for (int t = 0; t < grid.numTotalTiles(); ++t) {
        if (!(types[t] & LEAF)) continue;
        const float d = diag[size_t(t) * Tile::SIZE + ((3 << 6) | (3 << 3) | 3)];
        EXPECT(std::fabs(d - 6.0f / 64.0f) < 1e-7f);
        ++checked;
    }
    EXPECT(checked == 512);
    // errors surface as exceptions carrying the status
    bool thrown = false;
    try {
        grid.setChannel(0, std::vector<float>(3));
    } catch (const std::invalid_argument&) {
        thrown = true;
    }
    EXPECT(thrown);
    std::printf("cpu part OK\n");
    return 0;
}

static int gpu_part()
{
    HADeviceGrid grid(0.125f, {16, 16, 16, 16});
    build_uniform(grid);
    {
        // the device construction gives the same tiles in the same order
        HADeviceGrid dgrid(0.125f, {16, 16, 16, 16});
        build_uniform(dgrid, true);
        EXPECT(dgrid.numTotalTiles() == grid.numTotalTiles());
        std::vector<int32_t> sa(size_t(grid.numTotalTiles())), sb(sa.size());
        check(cmg_grid_get_slots(grid.handle(), sa.data()), "slots");
        check(cmg_grid_get_slots(dgrid.handle(), sb.data()), "slots");
        EXPECT(sa == sb);
        EXPECT(dgrid.tileTypes() == grid.tileTypes());
    }
    AMGSolver solver(11, 0.5, 1, 1);
    solver.prepareTypesAndCoeffs(grid);
    check(cmg_rhs_uniform(grid.handle(), Tile::b_channel, 0, 1), "rhs");
    const std::vector<float> b = grid.getChannel(Tile::b_channel);
    // FluidEuler::project's call
    auto [iters, err] = solver.solve(grid, false, 100, 1e-6, 2, 10, 1, false);
    std::printf("reference settings: %d iterations, rel res %.3e, true rel res %.3e\n", iters, err,
                solver.last.true_rel_res);
    EXPECT(iters > 0 && iters < 100);
    EXPECT(err < 1e-6);
    EXPECT(solver.last.true_rel_res < 1e-5);
    const std::vector<float> x = grid.getChannel(Tile::x_channel);
    // V-cycle form, same answer
    solver.mu_cycle_repeat_times = 1;
    grid.setChannel(Tile::b_channel, b);
    auto [iters_v, err_v] = solver.solve(grid, false, 100, 1e-6, 2, 10, 1, false);
    EXPECT(err_v < 1e-6);
    const std::vector<float> xv = grid.getChannel(Tile::x_channel);
    double num = 0, den = 0;
    for (size_t i = 0; i < x.size(); ++i) {
        num += double(x[i] - xv[i]) * double(x[i] - xv[i]);
        den += double(x[i]) * double(x[i]);
    }
    EXPECT(std::sqrt(num / den) < 1e-5);
    // no synchronised dot: (max_iters, -1)
    grid.setChannel(Tile::b_channel, b);
    auto [iters_n, err_n] = solver.solve(grid, false, 6, 1e-6, 2, 10, -1, false);
    EXPECT(iters_n == 6 && err_n == -1.0);
    std::printf("gpu part OK (%d W-cycle, %d V-cycle iterations)\n", iters, iters_v);
    return 0;
}

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    try {
        if (int rc = cpu_part()) return rc;
        if (gpu)
            if (int rc = gpu_part()) return rc;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 2;
    }
    return 0;
}
