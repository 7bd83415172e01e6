// cmg_host.hpp -- host-side (C++17) octree topology and operator assembly.
//
// The topology reproduces HADeviceGrid's construction exactly (same hash,
// same linear probing, same insertion order, slot-order compression), so the
// all-tiles order, the neighbour tables and the level assignment are the
// reference's (src/HAGrid.h, src/HAAccessor.h).  Assembly produces the
// reference AMG operator (src/AMGSolver.cu:12-318) on the host with OpenMP.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "cirrus_amg.h"

namespace cmg {

constexpr int kMaxLevels = 12;
constexpr int kCells = 512;
constexpr int kChannels = 15;
// int4 words of a stencil workgroup's tile record: (tile, level, type | octant << 8, parent),
// halo record (2), ghost-parent record (2), children (2)
constexpr int kRecWords = 7;

enum : uint8_t { LEAF = 1, NONLEAF = 2, GHOST = 4 };
enum : uint8_t { DIRICHLET = 1, INTERIOR = 2, NEUMANN = 4 };

using Coord = std::array<int32_t, 3>;

inline int cell_index(int i, int j, int k) { return (i << 6) | (j << 3) | k; }

// one hash slot of a level table: the tile identity plus its refine status
struct Slot {
    Coord c{0, 0, 0};
    uint8_t type = 0;
    bool used = false;
    bool refine = false;   // REFINE_FLAG of mStatus
};

class Topology {
public:
    Topology(float h0, int num_levels, const uint32_t* log2_hash);

    float h0() const { return h0_; }
    int num_levels() const { return num_levels_; }
    int max_level() const { return max_level_; }
    float h(int level) const { return h0_ / float(1 << level); }
    uint32_t log2_hash(int level) const { return log2_[level]; }

    // HATileAccessor::hash / tileIdx (src/HAAccessor.h:217-246)
    static uint32_t hash(const Coord& c, uint32_t log2);
    int probe(int level, const Coord& c) const;          // slot index, -1 on bad level, throws when full
    const Slot* find(int level, const Coord& c) const;   // nullptr when absent
    Slot* find(int level, const Coord& c);

    int set_tile(int level, const Coord& c, uint8_t type);   // setTileHost
    void compress();                                         // compressHost
    void spawn_ghosts();                                     // spawnGhostTiles
    int refine_step(const std::function<int(int, const Coord&)>& target);   // refineStep
    void iterative_refine(const std::function<int(int, const Coord&)>& target);

    // frozen all-tiles view (after finalize)
    void finalize();
    bool finalized() const { return finalized_; }
    int num_tiles() const { return int(t_level.size()); }
    int tile_at(int level, const Coord& c) const;         // -1 if absent

    std::vector<int32_t> t_level, t_slot;
    std::vector<Coord> t_coord;
    std::vector<uint8_t> t_type;
    std::vector<int32_t> nbr;      // 6 per tile, x-,y-,z-,x+,y+,z+
    std::vector<int32_t> parent;   // tile at level-1 containing the tile
    std::vector<int32_t> child;    // 8 per tile, octant bit0 = x
    std::vector<int32_t> level_start;

private:
    friend class DeviceBuilder;   // device construction reads and writes the tables (cmg_build.hip)
    float h0_;
    int num_levels_;
    int max_level_ = -1;
    bool finalized_ = false;
    std::array<uint32_t, kMaxLevels> log2_{};
    std::array<std::vector<Slot>, kMaxLevels> table_;
    std::array<std::vector<int32_t>, kMaxLevels> compressed_;   // slot indices in slot order
    std::array<std::vector<int32_t>, kMaxLevels> touched_;      // slots claimed since the last compression
    std::array<std::vector<int32_t>, kMaxLevels> slot_tile_;    // slot -> tile index
};

// analytic SDFs shared by targets and assembly (float, no contraction)
float sdf_eval(int kind, const float* prm, float x, float y, float z);
std::function<int(int, const Coord&)> make_target(const Topology& topo, const cmg_target& t);

// host copy of the per-cell data needed for assembly
struct HostCells {
    std::vector<uint8_t> type;                  // N*512
    std::array<std::vector<float>, kChannels> ch;   // lazily sized N*512
    std::vector<float>& channel(int c, size_t n);
};

void set_leaf_types_preset(const Topology& topo, HostCells& cells, const cmg_target& t);
void set_wall_types(const Topology& topo, HostCells& cells, int chk_level, bool dirichlet_ytop);
void set_solid_types(const Topology& topo, HostCells& cells, int sdf_kind, const float* sdf);
// face_rule: -1 binary (CoarsenTypesAndAMGCoeffs), 0 corner FaceFluidRatio, 1 supersampled
void assemble(const Topology& topo, HostCells& cells, int c0, float R, int face_rule,
              int sdf_kind, const float* sdf);
float face_fluid_ratio(float p0, float p1, float p2, float p3);

// seeded synthetic inputs (SplitMix64; 24-bit uniforms; Box-Muller normals)
void rhs_uniform(const Topology& topo, HostCells& cells, int ch, uint64_t seed, bool zero_mean);
// b = volume-weighted divergence of i.i.d. N(0,1) negative-face velocities
// (AMGVolumeWeightedDivergenceWithoutCoeffOnLeafs, src/AMGSolver.cu:750-788)
void rhs_divergence(const Topology& topo, HostCells& cells, int ch, int u0, uint64_t seed);

}  // namespace cmg
