// cmg_topology.cpp -- octree tile topology on the host (C++17).
//
// Construction order is the reference's (src/HAGrid.h:526-897,1088-1182), so
// hash-slot positions, the compressed per-level arrays and therefore the
// all-tiles order coincide with HADeviceGrid's dAllTiles.
#include <cmath>
#include <stdexcept>

#include "cmg_host.hpp"

namespace cmg {

namespace {
// neighbour offsets in HACoordAccessor::neighborOffset order (src/HAAccessor.h:165-171)
constexpr int kNbOrder[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

Coord child_coord(const Coord& c, int octant)
{
    return {2 * c[0] + (octant & 1), 2 * c[1] + ((octant >> 1) & 1), 2 * c[2] + ((octant >> 2) & 1)};
}
}  // namespace

Topology::Topology(float h0, int num_levels, const uint32_t* log2_hash) : h0_(h0), num_levels_(num_levels)
{
    if (num_levels <= 0 || num_levels > kMaxLevels) throw std::invalid_argument("num_levels out of range");
    for (int l = 0; l < num_levels; ++l) {
        if (log2_hash[l] < 1 || log2_hash[l] > 26) throw std::invalid_argument("log2 hash size out of range");
        log2_[l] = log2_hash[l];
        table_[l].assign(size_t(1) << log2_hash[l], Slot{});
    }
}

uint32_t Topology::hash(const Coord& c, uint32_t log2)
{
    constexpr uint32_t kFactors[3] = {1u, 2654435761u, 805459861u};
    uint32_t r = 0;
    for (int a = 0; a < 3; ++a) r ^= uint32_t(c[a]) * kFactors[a];
    return r & ((1u << log2) - 1u);
}

int Topology::probe(int level, const Coord& c) const
{
    if (level < 0 || level >= num_levels_) return -1;
    const auto& tab = table_[level];
    const uint32_t mask = uint32_t(tab.size()) - 1u;
    const uint32_t start = hash(c, log2_[level]);
    uint32_t k = start;
    do {
        const Slot& s = tab[k];
        if (!s.used || s.c == c) return int(k);
        k = (k + 1) & mask;
    } while (k != start);
    throw std::runtime_error("hash table full at level " + std::to_string(level));
}

const Slot* Topology::find(int level, const Coord& c) const
{
    int k = probe(level, c);
    if (k < 0) return nullptr;
    const Slot& s = table_[level][k];
    return s.used ? &s : nullptr;
}

Slot* Topology::find(int level, const Coord& c)
{
    return const_cast<Slot*>(static_cast<const Topology*>(this)->find(level, c));
}

int Topology::set_tile(int level, const Coord& c, uint8_t type)
{
    int k = probe(level, c);
    if (k < 0) throw std::invalid_argument("level out of range");
    Slot& s = table_[level][k];
    if (!s.used) {
        s.used = true;
        s.c = c;
        s.refine = false;
        touched_[level].push_back(int32_t(k));
    }
    s.type = type;
    finalized_ = false;
    return k;
}

void Topology::compress()
{
    max_level_ = -1;
    for (int l = 0; l < num_levels_; ++l) {
        touched_[l].clear();
        auto& out = compressed_[l];
        out.clear();
        for (size_t k = 0; k < table_[l].size(); ++k)
            if (table_[l][k].used) out.push_back(int32_t(k));
        if (!out.empty()) max_level_ = l;
    }
}

void Topology::spawn_ghosts()
{
    for (int level = max_level_; level > 0; --level) {
        // the compressed array holds copies taken at the last compression
        for (int32_t k : compressed_[level]) {
            const Slot info = table_[level][k];
            if (!(info.type & (LEAF | NONLEAF))) continue;
            for (const auto& o : kNbOrder) {
                Coord n{info.c[0] + o[0], info.c[1] + o[1], info.c[2] + o[2]};
                if (find(level, n)) continue;
                const Slot* p = find(level - 1, Coord{n[0] >> 1, n[1] >> 1, n[2] >> 1});
                if (p && (p->type & LEAF)) set_tile(level, n, GHOST);
            }
        }
    }
    compress();
}

int Topology::refine_step(const std::function<int(int, const Coord&)>& target)
{
    int total = 0;
    for (int i = max_level_; i >= 0; --i) {
        const auto& arr = compressed_[i];
        if (arr.empty()) continue;
        std::vector<char> flag(arr.size(), 0);
        // leaf flags: target above the level, or a flagged ghost child (2:1 grading)
        for (size_t j = 0; j < arr.size(); ++j) {
            Slot& e = table_[i][arr[j]];
            if (!(e.type & LEAF)) continue;
            bool r = target(i, e.c) > i;
            if (i < max_level_)
                for (int o = 0; o < 8; ++o) {
                    const Slot* c = find(i + 1, child_coord(e.c, o));
                    if (c && (c->type & GHOST) && c->refine) r = true;
                }
            if (i + 1 >= num_levels_) r = false;
            e.refine = r;
            flag[j] = r;
        }
        // ghost flags: any same-level LEAF neighbour flagged
        for (size_t j = 0; j < arr.size(); ++j) {
            Slot& e = table_[i][arr[j]];
            if (!(e.type & GHOST)) continue;
            bool f = false;
            for (const auto& o : kNbOrder) {
                const Slot* n = find(i, Coord{e.c[0] + o[0], e.c[1] + o[1], e.c[2] + o[2]});
                if (n && (n->type & LEAF) && n->refine) f = true;
            }
            e.refine = f;
        }
        for (size_t j = 0; j < arr.size(); ++j) {
            if (!flag[j]) continue;
            ++total;
            Slot& e = table_[i][arr[j]];
            e.type = NONLEAF;
            const Coord pc = e.c;
            const bool pstatus = e.refine;
            for (int o = 0; o < 8; ++o) {
                int k = set_tile(i + 1, child_coord(pc, o), LEAF);
                table_[i + 1][k].refine = pstatus;   // the child tile copies its parent's status
            }
        }
    }
    compress();
    return total;
}

void Topology::iterative_refine(const std::function<int(int, const Coord&)>& target)
{
    for (;;) {
        int n = refine_step(target);
        spawn_ghosts();
        if (n == 0) break;
    }
}

int Topology::tile_at(int level, const Coord& c) const
{
    if (level < 0 || level >= num_levels_) return -1;
    int k = probe(level, c);
    if (k < 0 || !table_[level][k].used) return -1;
    return slot_tile_[level][k];
}

void Topology::finalize()
{
    compress();
    t_level.clear(); t_slot.clear(); t_coord.clear(); t_type.clear();
    level_start.assign(num_levels_ + 1, 0);
    for (int l = 0; l < num_levels_; ++l) {
        level_start[l] = int32_t(t_level.size());
        slot_tile_[l].assign(table_[l].size(), -1);
        for (int32_t k : compressed_[l]) {
            slot_tile_[l][k] = int32_t(t_level.size());
            t_level.push_back(l);
            t_slot.push_back(k);
            t_coord.push_back(table_[l][k].c);
            t_type.push_back(table_[l][k].type);
        }
    }
    level_start[num_levels_] = int32_t(t_level.size());
    const int N = num_tiles();
    nbr.assign(size_t(N) * 6, -1);
    parent.assign(N, -1);
    child.assign(size_t(N) * 8, -1);
    for (int t = 0; t < N; ++t) {
        const int l = t_level[t];
        const Coord& c = t_coord[t];
        for (int a = 0; a < 3; ++a) {
            Coord m = c, p = c;
            m[a] -= 1;
            p[a] += 1;
            nbr[size_t(t) * 6 + a] = tile_at(l, m);
            nbr[size_t(t) * 6 + 3 + a] = tile_at(l, p);
        }
        parent[t] = tile_at(l - 1, Coord{c[0] >> 1, c[1] >> 1, c[2] >> 1});
        for (int o = 0; o < 8; ++o) child[size_t(t) * 8 + o] = tile_at(l + 1, child_coord(c, o));
    }
    finalized_ = true;
}

// ---------------------------------------------------------------- targets ---

float sdf_eval(int kind, const float* p, float x, float y, float z)
{
    switch (kind) {
    case CMG_SDF_SPHERE: {
        float dx = x - p[0], dy = y - p[1], dz = z - p[2];
        return std::sqrt(dx * dx + dy * dy + dz * dz) - p[3];
    }
    case CMG_SDF_TORUS: {
        float dx = x - p[0], dy = y - p[1], dz = z - p[2];
        float q = std::sqrt(dx * dx + dy * dy) - p[3];
        return std::sqrt(q * q + dz * dz) - p[4];
    }
    default: {   // star of tests/TestGrids.cuh:97-107
        float sx = x - 0.5f, sy = y - 0.5f, sz = z - 0.5f;
        float r = std::sqrt(sx * sx + sy * sy + sz * sz);
        if (r == 0) return -3.402823466e+38f;
        float theta = std::acos(sz / r);
        float phi = std::atan2(sy, sx);
        return r - float(0.237 + 0.079 * double(std::cos(6 * theta)) * double(std::cos(6 * phi)));
    }
    }
}

namespace {
constexpr float kTestSphere[4] = {0.5f, 0.5f, 0.5f, 0.25f};

int corners_inside(int kind, const float* prm, const float mn[3], const float mx[3], float iso)
{
    int n = 0;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk)
                if (sdf_eval(kind, prm, di ? mx[0] : mn[0], dj ? mx[1] : mn[1], dk ? mx[2] : mn[2]) - iso < 0) ++n;
    return n;
}
}  // namespace

std::function<int(int, const Coord&)> make_target(const Topology& topo, const cmg_target& t)
{
    const cmg_target tc = t;
    const float h0 = topo.h0();
    auto bbox = [h0](int level, const Coord& b, float mn[3], float mx[3]) {
        float h = h0 / float(1 << level);
        for (int a = 0; a < 3; ++a) {
            mn[a] = (float(b[a]) * 8.0f) * h;
            mx[a] = (float(b[a] + 1) * 8.0f) * h;
        }
    };
    switch (t.kind) {
    case CMG_TARGET_UNIFORM:
        return [tc](int, const Coord&) { return tc.max_level; };
    case CMG_TARGET_CENTER_POINT:
        return [tc, bbox](int level, const Coord& b) {
            float mn[3], mx[3];
            bbox(level, b, mn, mx);
            bool in = true;
            for (int a = 0; a < 3; ++a) {
                const float eps = float(1e-6);
                float lo = mn[a] - eps, hi = mx[a] + eps;
                in = in && (0.5f > lo && 0.5f < hi);
            }
            return in ? tc.max_level : tc.min_level;
        };
    case CMG_TARGET_SPHERE_EMPTY:
    case CMG_TARGET_STAR_EMPTY:
        return [tc, bbox](int level, const Coord& b) {
            float mn[3], mx[3];
            bbox(level, b, mn, mx);
            int kind = tc.kind == CMG_TARGET_SPHERE_EMPTY ? CMG_SDF_SPHERE : CMG_SDF_STAR;
            int n = corners_inside(kind, kTestSphere, mn, mx, 0.0f);
            return (n == 0 || n == 8) ? tc.min_level : tc.max_level;
        };
    case CMG_TARGET_SPHERE_SOLID:
    case CMG_TARGET_STAR_SOLID:
        return [tc, bbox](int level, const Coord& b) {
            float mn[3], mx[3];
            bbox(level, b, mn, mx);
            int kind = tc.kind == CMG_TARGET_SPHERE_SOLID ? CMG_SDF_SPHERE : CMG_SDF_STAR;
            int n0 = corners_inside(kind, kTestSphere, mn, mx, 0.0f);
            int n1 = corners_inside(kind, kTestSphere, mn, mx, 0.01f);
            return (n0 == 8 || n1 == 0) ? tc.min_level : tc.max_level;
        };
    case CMG_TARGET_SDF_BANDS:
        return [tc, h0](int level, const Coord& b) {
            float w = 8.0f * (h0 / float(1 << level));
            float cx = (float(b[0]) + 0.5f) * w, cy = (float(b[1]) + 0.5f) * w, cz = (float(b[2]) + 0.5f) * w;
            float d = std::fabs(sdf_eval(tc.sdf_kind, tc.sdf, cx, cy, cz)) - 0.8660254f * w;
            for (int l = tc.max_level; l > tc.min_level; --l)
                if (d < tc.band[l]) return l;
            return tc.min_level;
        };
    case CMG_TARGET_CALLBACK:
        return [tc](int level, const Coord& b) {
            int32_t c[3] = {b[0], b[1], b[2]};
            return tc.fn(tc.user, level, c);
        };
    }
    throw std::invalid_argument("unknown level target kind");
}

}  // namespace cmg
