// cmg_build.hip -- octree construction on the device: refine flags with 2:1
// grading, child and ghost insertion into the per-level hash tables, slot-order
// compression, and the all-tiles, neighbour, parent and child tables.
//
// The host construction (cmg_topology.cpp) inserts keys one at a time with
// linear probing, in the reference's order (refineStep, src/HAGrid.h:1088-1171;
// spawnGhostTiles, :856-889; setTileHost, :526-548; compressHost, :634-653).
// Here every new key of one insertion round carries its rank in that
// sequential order as a priority, and the round is one lock-free pass of
// ordered linear probing: a key passes slots that hold keys of higher priority
// (or keys from earlier rounds), and displaces a key of lower priority, which
// it then carries on from the next slot.  The final layout is the unique one
// in which every slot on a key's probe path before its own holds a key of
// higher priority, and sequential insertion in rank order produces exactly
// that layout; so hash slots, compressed arrays and tile ids are those of the
// host construction, bit for bit.
//
// Ghost spawning can propose one key from several tiles (sequentially, the
// first proposal inserts it and the later ones find it); a lock-free hash set
// keeps, for every proposed key, the proposal of lowest rank before insertion.
//
// Level targets are evaluated on the device for the uniform, centre-point,
// test-sphere and sphere/torus narrow-band kinds; the star SDF (whose acos /
// atan2 / cos the device libraries do not round like the host's) and user
// callbacks are evaluated on the host for the tiles of one level at a time.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "cmg_internal.hpp"

#pragma clang fp contract(off)

namespace cmg {

namespace {

constexpr unsigned long long kEmpty = ~0ull;
constexpr uint32_t kEmpty32 = 0xffffffffu;
constexpr int kBlock = 256;

// meta word of a slot: bit 0 used, bits 1..3 tile type, bit 4 REFINE_FLAG
__host__ __device__ __forceinline__ int make_meta(int type, int refine) { return 1 | (type << 1) | (refine << 4); }
__device__ __forceinline__ int m_type(int m) { return (m >> 1) & 7; }
__device__ __forceinline__ int m_refine(int m) { return (m >> 4) & 1; }

struct Tabs {
    int4* slot[kMaxLevels];    // (x, y, z, meta) per hash slot
    uint32_t mask[kMaxLevels];
    int num_levels;
};

struct SlotTiles {
    int32_t* p[kMaxLevels];    // slot -> tile index
};

struct TargetDev {
    int kind, min_level, max_level, sdf_kind;
    float sdf[8];
    float band[12];
    float h0;
};

// HATileAccessor::hash (src/HAAccessor.h:217-226)
__device__ __forceinline__ uint32_t khash(int x, int y, int z)
{
    return uint32_t(x) ^ (uint32_t(y) * 2654435761u) ^ (uint32_t(z) * 805459861u);
}

// slot of (x, y, z) at a level, -1 when absent; a full table sets err (the host probe throws there)
__device__ int dfind(const Tabs& T, int level, int x, int y, int z, int* err)
{
    if (level < 0 || level >= T.num_levels) return -1;
    const uint32_t mask = T.mask[level];
    const int4* tab = T.slot[level];
    uint32_t k = khash(x, y, z) & mask;
    for (uint32_t it = 0; it <= mask; ++it) {
        const int4 s = tab[k];
        if (!(s.w & 1)) return -1;
        if (s.x == x && s.y == y && s.z == z) return int(k);
        k = (k + 1) & mask;
    }
    *err = 1;
    return -1;
}

// ------------------------------------------------------------ level targets

__device__ float sdf_t(int kind, const float* p, float x, float y, float z)
{
    if (kind == CMG_SDF_SPHERE) {
        const float dx = x - p[0], dy = y - p[1], dz = z - p[2];
        return sqrtf(dx * dx + dy * dy + dz * dz) - p[3];
    }
    const float dx = x - p[0], dy = y - p[1], dz = z - p[2];
    const float q = sqrtf(dx * dx + dy * dy) - p[3];
    return sqrtf(q * q + dz * dz) - p[4];
}

__device__ int corners_in(const float* p, const float mn[3], const float mx[3], float iso)
{
    int n = 0;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk)
                if (sdf_t(CMG_SDF_SPHERE, p, di ? mx[0] : mn[0], dj ? mx[1] : mn[1], dk ? mx[2] : mn[2]) - iso < 0) ++n;
    return n;
}

// the level-target presets of make_target (cmg_topology.cpp) that need no star SDF
__device__ int eval_target(const TargetDev& t, int level, int bx, int by, int bz)
{
    const int b[3] = {bx, by, bz};
    if (t.kind == CMG_TARGET_UNIFORM) return t.max_level;
    if (t.kind == CMG_TARGET_SDF_BANDS) {
        const float w = 8.0f * (t.h0 / float(1 << level));
        const float cx = (float(bx) + 0.5f) * w, cy = (float(by) + 0.5f) * w, cz = (float(bz) + 0.5f) * w;
        const float d = fabsf(sdf_t(t.sdf_kind, t.sdf, cx, cy, cz)) - 0.8660254f * w;
        for (int l = t.max_level; l > t.min_level; --l)
            if (d < t.band[l]) return l;
        return t.min_level;
    }
    const float h = t.h0 / float(1 << level);
    float mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = (float(b[a]) * 8.0f) * h;
        mx[a] = (float(b[a] + 1) * 8.0f) * h;
    }
    if (t.kind == CMG_TARGET_CENTER_POINT) {
        bool in = true;
        for (int a = 0; a < 3; ++a) {
            const float eps = float(1e-6);
            const float lo = mn[a] - eps, hi = mx[a] + eps;
            in = in && (0.5f > lo && 0.5f < hi);
        }
        return in ? t.max_level : t.min_level;
    }
    const float sphere[4] = {0.5f, 0.5f, 0.5f, 0.25f};   // CreateTestGrid's sphere (tests/TestGrids.cuh:33-95)
    if (t.kind == CMG_TARGET_SPHERE_EMPTY) {
        const int n = corners_in(sphere, mn, mx, 0.0f);
        return (n == 0 || n == 8) ? t.min_level : t.max_level;
    }
    const int n0 = corners_in(sphere, mn, mx, 0.0f);   // CMG_TARGET_SPHERE_SOLID
    const int n1 = corners_in(sphere, mn, mx, 0.01f);
    return (n0 == 8 || n1 == 0) ? t.min_level : t.max_level;
}

}  // namespace

namespace k {

__global__ void b_gather(const int4* slot, const int32_t* comp, int n, int4* out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = slot[comp[j]];
}

// MarkRefineFlagOnLeafsWithLevelTargetHelperKernel (src/HAGrid.h:279-322)
// tgt: host-evaluated targets per compressed entry, or nullptr to evaluate t here
__global__ void b_leaf_flags(Tabs T, int i, const int32_t* comp, int n, const int32_t* tgt, TargetDev t, int max_level,
                             int32_t* flag, int* err)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int4* tab = T.slot[i];
    const int k = comp[j];
    const int4 s = tab[k];
    if (!(m_type(s.w) & LEAF)) {
        flag[j] = 0;
        return;
    }
    bool r = (tgt ? tgt[j] : eval_target(t, i, s.x, s.y, s.z)) > i;
    if (i < max_level)
        for (int o = 0; o < 8; ++o) {
            const int q = dfind(T, i + 1, 2 * s.x + (o & 1), 2 * s.y + ((o >> 1) & 1), 2 * s.z + ((o >> 2) & 1), err);
            if (q >= 0) {
                const int m = T.slot[i + 1][q].w;
                if ((m_type(m) & GHOST) && m_refine(m)) r = true;
            }
        }
    if (i + 1 >= T.num_levels) r = false;
    tab[k].w = (s.w & ~16) | (r ? 16 : 0);
    flag[j] = r ? 1 : 0;
}

// ghost REFINE_FLAG: any same-level LEAF neighbour flagged (src/HAGrid.h:1116-1134)
__global__ void b_ghost_flags(Tabs T, int i, const int32_t* comp, int n, int* err)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int4* tab = T.slot[i];
    const int k = comp[j];
    const int4 s = tab[k];
    if (!(m_type(s.w) & GHOST)) return;
    const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    bool f = false;
    for (int o = 0; o < 6; ++o) {
        const int q = dfind(T, i, s.x + off[o][0], s.y + off[o][1], s.z + off[o][2], err);
        if (q >= 0) {
            const int m = tab[q].w;
            if ((m_type(m) & LEAF) && m_refine(m)) f = true;
        }
    }
    tab[k].w = (s.w & ~16) | (f ? 16 : 0);
}

// a flagged LEAF becomes NONLEAF and proposes its eight children (src/HAGrid.h:1140-1164); children that
// exist (ghosts) become LEAF in place, the others are candidates of rank 8 * (flagged index) + octant + 1
__global__ void b_children(Tabs T, int i, const int32_t* comp, int n, const int32_t* flag, const int32_t* pos, int4* cand,
                           uint8_t* valid, unsigned int* count, int* err)
{
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const int j = int(idx >> 3), o = int(idx & 7);
    if (j >= n || !flag[j]) return;
    const int k = comp[j];
    const int4 s = T.slot[i][k];
    const int refine = m_refine(s.w);
    if (o == 0) T.slot[i][k].w = make_meta(NONLEAF, refine);
    const int c = pos[j] * 8 + o;
    const int cx = 2 * s.x + (o & 1), cy = 2 * s.y + ((o >> 1) & 1), cz = 2 * s.z + ((o >> 2) & 1);
    const int meta = make_meta(LEAF, refine);   // the child tile copies its parent's status
    const int q = dfind(T, i + 1, cx, cy, cz, err);
    if (q >= 0) {
        T.slot[i + 1][q].w = meta;
        valid[c] = 0;
        return;
    }
    cand[c] = make_int4(cx, cy, cz, meta);
    valid[c] = 1;
    atomicAdd(count, 1u);
}

// spawnGhostTiles proposals at level L (src/HAGrid.h:856-889): an empty face neighbour of a LEAF or
// NONLEAF tile whose parent is a LEAF; rank 6 * (compressed index) + neighbour + 1
__global__ void b_ghost_cand(Tabs T, int L, const int32_t* comp, int n, int4* cand, uint8_t* valid, int* err)
{
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(n) * 6) return;
    const int j = int(idx / 6), o = int(idx % 6);
    const int4 s = T.slot[L][comp[j]];
    valid[idx] = 0;
    if (!(m_type(s.w) & (LEAF | NONLEAF))) return;
    const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    const int nx = s.x + off[o][0], ny = s.y + off[o][1], nz = s.z + off[o][2];
    if (dfind(T, L, nx, ny, nz, err) >= 0) return;
    const int q = dfind(T, L - 1, nx >> 1, ny >> 1, nz >> 1, err);
    if (q < 0 || !(m_type(T.slot[L - 1][q].w) & LEAF)) return;
    cand[idx] = make_int4(nx, ny, nz, make_meta(GHOST, 0));
    valid[idx] = 1;
}

__global__ void b_set_init(uint32_t* dword, uint32_t* drank, uint32_t dsize, uint32_t* dslot, int nc)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < dsize) {
        dword[k] = kEmpty32;
        drank[k] = kEmpty32;
    }
    if (k < uint32_t(nc)) dslot[k] = kEmpty32;
}

// lock-free set of proposed keys; drank[slot] ends as the lowest rank proposing the key
__global__ void b_dedupe(const int4* cand, const uint8_t* valid, int n, uint32_t* dword, uint32_t* drank, uint32_t dmask,
                         uint32_t* dslot, int* err)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || !valid[c]) return;
    const int4 key = cand[c];
    uint32_t k = khash(key.x, key.y, key.z) & dmask;
    for (uint32_t it = 0; it <= dmask; ++it) {
        uint32_t cur = __hip_atomic_load(&dword[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == kEmpty32) {
            const uint32_t prev = atomicCAS(&dword[k], kEmpty32, uint32_t(c));
            cur = prev == kEmpty32 ? uint32_t(c) : prev;
        }
        const int4 o = cand[cur];
        if (o.x == key.x && o.y == key.y && o.z == key.z) {
            atomicMin(&drank[k], uint32_t(c) + 1u);
            dslot[c] = k;
            return;
        }
        k = (k + 1) & dmask;
    }
    *err = 3;
}

__global__ void b_representative(uint8_t* valid, int n, const uint32_t* dslot, const uint32_t* drank, uint32_t dmask,
                                 unsigned int* count)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || !valid[c]) return;
    const uint32_t k = dslot[c];
    if (k > dmask || drank[k] != uint32_t(c) + 1u) {   // k > dmask only after a failed b_dedupe (err set)
        valid[c] = 0;
        return;
    }
    atomicAdd(count, 1u);
}

// ordered linear probing: word = 0 for a key of an earlier round, ~0 for an empty slot, else
// (rank << 32 | candidate); a smaller word has the higher priority
// The round's new keys (*fresh) must fit the table next to its used slots (*used),
// else nothing is inserted and err = 1.  limit (optional): only candidates below
// per * *limit belong to the round.
__global__ void b_insert(unsigned long long* word, uint32_t mask, const int4* cand, const uint8_t* valid, int n,
                         const int32_t* limit, int per, const unsigned int* fresh, const int32_t* used, int* err)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if ((unsigned long long)(uint32_t(*used)) + *fresh > mask + 1ull) {
        if (c == 0) *err = 1;
        return;
    }
    if (c >= n || (limit && c >= per * *limit) || !valid[c]) return;
    unsigned long long mine = ((unsigned long long)(uint32_t(c) + 1u) << 32) | uint32_t(c);
    const int4 key = cand[c];
    uint32_t k = khash(key.x, key.y, key.z) & mask;
    const unsigned long long cap = 8ull * (mask + 1ull) + 1024ull;
    for (unsigned long long it = 0; it < cap; ++it) {
        const unsigned long long cur = __hip_atomic_load(&word[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur < mine) {
            k = (k + 1) & mask;
            continue;
        }
        const unsigned long long prev = atomicCAS(&word[k], cur, mine);
        if (prev != cur) continue;
        if (cur == kEmpty) return;
        mine = cur;
        k = (k + 1) & mask;
    }
    *err = 2;
}

// also books the round: used += fresh when it fitted, fresh = 0 for the next round
__global__ void b_materialize(unsigned long long* word, int4* slot, uint32_t size, const int4* cand, unsigned int* fresh,
                              int32_t* used)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        const unsigned long long need = (unsigned long long)(uint32_t(*used)) + *fresh;
        if (need <= size) *used = int32_t(need);
        *fresh = 0;
    }
    if (k >= size) return;
    const unsigned long long w = word[k];
    if (w == 0 || w == kEmpty) return;
    slot[k] = cand[uint32_t(w)];
    word[k] = 0;
}

__global__ void b_upload(const int32_t* rec, int n, int4* slot, unsigned long long* word)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int k = rec[5 * j];
    slot[k] = make_int4(rec[5 * j + 1], rec[5 * j + 2], rec[5 * j + 3], rec[5 * j + 4]);
    word[k] = 0;
}

__global__ void b_scatter(const int4* slot, const int32_t* pos, uint32_t size, int32_t* comp)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < size && (slot[k].w & 1)) comp[pos[k]] = int32_t(k);
}

// exclusive scan of 0/1 flags: 1024 items a workgroup, then the workgroup sums, then the offsets
constexpr int kScanItems = 4, kScanTile = kBlock * kScanItems;

__device__ int block_inclusive(int v, int* s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const int t = threadIdx.x >= unsigned(off) ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    const int r = s[threadIdx.x];
    __syncthreads();
    return r;
}

// total (one-tile scans only): the scan's total, written here instead of by b_scan_sums
// slots (optional): the flags are the slots' used bits instead of in[]
__global__ void __launch_bounds__(kBlock) b_scan_tiles(const int32_t* in, const int4* slots, int n, int32_t* out,
                                                       int32_t* tsum, int32_t* total)
{
    __shared__ int s[kBlock];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems], sum = 0;
    for (int e = 0; e < kScanItems; ++e) {
        v[e] = base + e < n ? (slots ? (slots[base + e].w & 1) : in[base + e]) : 0;
        sum += v[e];
    }
    const int incl = block_inclusive(sum, s);
    int run = incl - sum;
    for (int e = 0; e < kScanItems; ++e) {
        if (base + e < n) out[base + e] = run;
        run += v[e];
    }
    if (threadIdx.x == kBlock - 1) {
        if (total) total[0] = incl;
        else tsum[blockIdx.x] = incl;
    }
}

__global__ void __launch_bounds__(kBlock) b_scan_sums(int32_t* tsum, int nt, int32_t* total)
{
    __shared__ int s[kBlock];
    int carry = 0;
    for (int b0 = 0; b0 < nt; b0 += kBlock) {
        const int i = b0 + threadIdx.x;
        const int v = i < nt ? tsum[i] : 0;
        const int incl = block_inclusive(v, s);
        if (i < nt) tsum[i] = carry + incl - v;
        s[threadIdx.x] = incl;
        __syncthreads();
        carry += s[kBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[0] = carry;
}

__global__ void b_scan_add(int32_t* out, int n, const int32_t* tsum)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tsum[i / kScanTile];
}

// CalculateNeighborTiles (src/PoissonGrid.cu:139-152) plus parent and child tables
__global__ void b_final_level(const int4* slot, const int32_t* comp, int n, int level, int start, int32_t* t_level,
                              int32_t* t_slot, int32_t* t_coord, uint8_t* t_type, int32_t* slot_tile)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int t = start + j, k = comp[j];
    const int4 s = slot[k];
    t_level[t] = level;
    t_slot[t] = k;
    t_coord[3 * t] = s.x;
    t_coord[3 * t + 1] = s.y;
    t_coord[3 * t + 2] = s.z;
    t_type[t] = uint8_t(m_type(s.w));
    slot_tile[k] = t;
}

__global__ void b_final_links(Tabs T, SlotTiles st, int N, const int32_t* t_level, const int32_t* t_coord, int32_t* nbr,
                              int32_t* parent, int32_t* child, int* err)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int l = t_level[t];
    const int c[3] = {t_coord[3 * t], t_coord[3 * t + 1], t_coord[3 * t + 2]};
    auto at = [&](int lv, int x, int y, int z) {
        const int q = dfind(T, lv, x, y, z, err);
        return q < 0 ? -1 : st.p[lv][q];
    };
    for (int a = 0; a < 3; ++a) {
        int m[3] = {c[0], c[1], c[2]}, p[3] = {c[0], c[1], c[2]};
        m[a] -= 1;
        p[a] += 1;
        nbr[size_t(t) * 6 + a] = at(l, m[0], m[1], m[2]);
        nbr[size_t(t) * 6 + 3 + a] = at(l, p[0], p[1], p[2]);
    }
    parent[t] = at(l - 1, c[0] >> 1, c[1] >> 1, c[2] >> 1);
    for (int o = 0; o < 8; ++o)
        child[size_t(t) * 8 + o] = at(l + 1, 2 * c[0] + (o & 1), 2 * c[1] + ((o >> 1) & 1), 2 * c[2] + ((o >> 2) & 1));
}

}  // namespace k

namespace {

void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string("device octree construction: ") + what + ": " +
                                                  hipGetErrorString(e));
}

unsigned grid_of(long n) { return unsigned((n + kBlock - 1) / kBlock); }

// device buffer that grows on demand
struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    template <class T>
    T* need(size_t n)
    {
        const size_t b = std::max<size_t>(n, 1) * sizeof(T);
        if (b > bytes) {
            if (p) hip_ok(hipFree(p), "free");
            p = nullptr;
            size_t grown = 4096;   // power-of-two growth: few reallocations as a build proceeds
            while (grown < b) grown <<= 1;
            hip_ok(hipMalloc(&p, grown), "scratch allocation");
            bytes = grown;
        }
        return static_cast<T*>(p);
    }
    ~Buf()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

// Owns the device copy of a Topology's hash tables for the duration of one call.
// Counts stay on the device between kernels (scan totals, new-key counts, used
// slots); the host reads them back once per compression, together with the
// error flag and the refined-tile counts.
class DeviceBuilder {
public:
    explicit DeviceBuilder(Topology& topo) : topo_(topo), L_(topo.num_levels_)
    {
        try {
            hip_ok(hipStreamCreate(&st_), "stream");
            hip_ok(hipMalloc(&small_, sizeof(Small)), "counters");
            hip_ok(hipHostMalloc(&pinned_, sizeof(Small), hipHostMallocDefault), "pinned counters");
            // every level's slots, then every level's insertion words, then the compressed arrays
            size_t slots = 0;
            for (int l = 0; l < L_; ++l) slots += topo_.table_[l].size();
            hip_ok(hipMalloc(&block_, slots * (sizeof(int4) + sizeof(unsigned long long) + sizeof(int32_t))), "tables");
            int4* sp = static_cast<int4*>(block_);
            unsigned long long* wp = reinterpret_cast<unsigned long long*>(sp + slots);
            int32_t* cp = reinterpret_cast<int32_t*>(wp + slots);
            for (int l = 0; l < L_; ++l) {
                const size_t size = topo_.table_[l].size();
                slot_[l] = sp;
                word_[l] = wp;
                comp_[l] = cp;
                sp += size;
                wp += size;
                cp += size;
            }
            hip_ok(hipMemsetAsync(slot_[0], 0, slots * sizeof(int4), st_), "clear slots");
            hip_ok(hipMemsetAsync(word_[0], 0xff, slots * sizeof(unsigned long long), st_), "clear words");
            hip_ok(hipMemsetAsync(small_, 0, sizeof(Small), st_), "clear counters");
        } catch (...) {
            release();   // a throwing constructor gets no destructor call
            throw;
        }
    }
    ~DeviceBuilder() { release(); }

    // the host tables as they stand: the used slots (those of the last compression plus the
    // ones claimed since), the compressed arrays of the last compression, the max level
    void upload()
    {
        std::vector<int32_t> rec, used(kMaxLevels, 0);
        for (int l = 0; l < L_; ++l) {
            const auto& tab = topo_.table_[l];
            rec.clear();
            for (const auto* list : {&topo_.compressed_[l], &topo_.touched_[l]})
                for (int32_t k : *list) {
                    const Slot& s = tab[k];
                    rec.insert(rec.end(), {k, s.c[0], s.c[1], s.c[2], make_meta(s.type, s.refine ? 1 : 0)});
                }
            used[l] = int(rec.size() / 5);
            if (used[l]) {
                int32_t* d = recbuf_.need<int32_t>(rec.size());
                hip_ok(hipMemcpyAsync(d, rec.data(), rec.size() * sizeof(int32_t), hipMemcpyHostToDevice, st_), "upload");
                k::b_upload<<<grid_of(used[l]), kBlock, 0, st_>>>(d, used[l], slot_[l], word_[l]);
                hip_ok(hipStreamSynchronize(st_), "upload");   // rec is reused for the next level
            }
            const auto& comp = topo_.compressed_[l];
            cnt_[l] = int(comp.size());
            if (cnt_[l]) {
                hip_ok(hipMemcpyAsync(comp_[l], comp.data(), comp.size() * sizeof(int32_t), hipMemcpyHostToDevice, st_),
                       "upload");
                hip_ok(hipStreamSynchronize(st_), "upload");
            }
        }
        hip_ok(hipMemcpyAsync(small_->used, used.data(), kMaxLevels * sizeof(int32_t), hipMemcpyHostToDevice, st_),
               "upload");
        hip_ok(hipStreamSynchronize(st_), "upload");
        max_level_ = topo_.max_level_;
        for (int l = 0; l < L_; ++l) dirty_[l] = true;   // the host's compressed arrays may be stale
    }

    int refine_step(const cmg_target& t)
    {
        const bool on_device = target_on_device(t);
        std::function<int(int, const Coord&)> host_target;
        if (!on_device) host_target = make_target(topo_, t);
        TargetDev td{};
        if (on_device) {
            td.kind = t.kind;
            td.min_level = t.min_level;
            td.max_level = t.max_level;
            td.sdf_kind = t.sdf_kind;
            std::memcpy(td.sdf, t.sdf, sizeof(td.sdf));
            std::memcpy(td.band, t.band, sizeof(td.band));
            td.h0 = topo_.h0_;
        }
        const Tabs T = tabs();
        hip_ok(hipMemsetAsync(small_->refined, 0, sizeof(small_->refined), st_), "refined counts");
        for (int i = max_level_; i >= 0; --i) {
            const int n = cnt_[i];
            if (!n) continue;
            int32_t* tgt = on_device ? nullptr : tgt_.need<int32_t>(n);
            if (!on_device) {
                std::vector<int4> rec(n);
                int4* d = gather_.need<int4>(n);
                k::b_gather<<<grid_of(n), kBlock, 0, st_>>>(slot_[i], comp_[i], n, d);
                hip_ok(hipMemcpyAsync(rec.data(), d, n * sizeof(int4), hipMemcpyDeviceToHost, st_), "gather");
                hip_ok(hipStreamSynchronize(st_), "gather");
                std::vector<int32_t> h(n, 0);
                for (int j = 0; j < n; ++j)
                    if ((rec[j].w >> 1) & LEAF) h[j] = host_target(i, Coord{rec[j].x, rec[j].y, rec[j].z});
                hip_ok(hipMemcpyAsync(tgt, h.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st_), "targets");
                hip_ok(hipStreamSynchronize(st_), "targets");
            }
            int32_t* flag = flag_.need<int32_t>(n);
            k::b_leaf_flags<<<grid_of(n), kBlock, 0, st_>>>(T, i, comp_[i], n, on_device ? nullptr : tgt, td, max_level_,
                                                             flag, err());
            k::b_ghost_flags<<<grid_of(n), kBlock, 0, st_>>>(T, i, comp_[i], n, err());
            if (i + 1 >= L_) continue;   // no level above: every flag is clear
            int32_t* pos = pos_.need<int32_t>(n);
            scan(flag, n, pos, small_->refined + i);
            // children of the flagged tiles: candidate c = 8 (flagged index) + octant, at most 8 n
            const int nc = 8 * n;
            int4* cand = cand_.need<int4>(nc);
            uint8_t* valid = valid_.need<uint8_t>(nc);
            k::b_children<<<grid_of(long(n) * 8), kBlock, 0, st_>>>(T, i, comp_[i], n, flag, pos, cand, valid,
                                                                     small_->fresh + (i + 1), err());
            insert(i + 1, cand, valid, nc, small_->refined + i, 8);   // entries below 8 x (flagged tiles)
        }
        compress();
        int total = 0;
        for (int l = 0; l < L_; ++l) total += host_.refined[l];
        return total;
    }

    void spawn_ghosts()
    {
        const Tabs T = tabs();
        for (int L = max_level_; L > 0; --L) {
            const int n = cnt_[L];
            if (!n) continue;
            const int nc = 6 * n;
            int4* cand = cand_.need<int4>(nc);
            uint8_t* valid = valid_.need<uint8_t>(nc);
            k::b_ghost_cand<<<grid_of(nc), kBlock, 0, st_>>>(T, L, comp_[L], n, cand, valid, err());
            uint32_t dsize = 1024;
            while (dsize < 2u * uint32_t(nc)) dsize <<= 1;
            uint32_t* dword = dword_.need<uint32_t>(dsize);
            uint32_t* drank = drank_.need<uint32_t>(dsize);
            uint32_t* dslot = dslot_.need<uint32_t>(nc);
            k::b_set_init<<<grid_of(std::max<long>(dsize, nc)), kBlock, 0, st_>>>(dword, drank, dsize, dslot, nc);
            k::b_dedupe<<<grid_of(nc), kBlock, 0, st_>>>(cand, valid, nc, dword, drank, dsize - 1, dslot, err());
            k::b_representative<<<grid_of(nc), kBlock, 0, st_>>>(valid, nc, dslot, drank, dsize - 1, small_->fresh + L);
            insert(L, cand, valid, nc, nullptr, 0);
        }
        compress();
    }

    // compressHost: per level the used slots in slot order (levels with no insertion round
    // since the last compression are already compressed); one read-back of every count
    void compress()
    {
        for (int l = 0; l < L_; ++l) {
            if (!dirty_[l]) continue;
            dirty_[l] = false;
            const uint32_t size = uint32_t(topo_.table_[l].size());
            int32_t* pos = pos_.need<int32_t>(size);
            scan(nullptr, int(size), pos, small_->count + l, slot_[l]);
            k::b_scatter<<<grid_of(size), kBlock, 0, st_>>>(slot_[l], pos, size, comp_[l]);
        }
        read_back();
        max_level_ = -1;
        for (int l = 0; l < L_; ++l) {
            cnt_[l] = host_.count[l];
            if (cnt_[l]) max_level_ = l;
        }
    }

    // used slots and compressed arrays back into the host tables
    void download()
    {
        for (int l = 0; l < L_; ++l) {
            const int n = cnt_[l];
            auto& comp = topo_.compressed_[l];
            comp.resize(n);
            topo_.touched_[l].clear();
            if (!n) continue;
            std::vector<int4> rec(n);
            int4* d = gather_.need<int4>(n);
            k::b_gather<<<grid_of(n), kBlock, 0, st_>>>(slot_[l], comp_[l], n, d);
            hip_ok(hipMemcpyAsync(rec.data(), d, n * sizeof(int4), hipMemcpyDeviceToHost, st_), "download");
            hip_ok(hipMemcpyAsync(comp.data(), comp_[l], n * sizeof(int32_t), hipMemcpyDeviceToHost, st_), "download");
            hip_ok(hipStreamSynchronize(st_), "download");
            auto& tab = topo_.table_[l];
            for (int j = 0; j < n; ++j) {
                Slot& s = tab[comp[j]];
                s.c = Coord{rec[j].x, rec[j].y, rec[j].z};
                s.type = uint8_t((rec[j].w >> 1) & 7);
                s.used = true;
                s.refine = ((rec[j].w >> 4) & 1) != 0;
            }
        }
        topo_.max_level_ = max_level_;
        topo_.finalized_ = false;
    }

    // Topology::finalize with the tile arrays and link tables computed here
    void finalize()
    {
        compress();
        download();
        std::vector<int32_t> start(L_ + 1, 0);
        for (int l = 0; l < L_; ++l) start[l + 1] = start[l] + cnt_[l];
        const int N = start[L_];
        Buf tl, ts, tc, tt, nb, pa, ch;
        int32_t* t_level = tl.need<int32_t>(N);
        int32_t* t_slot = ts.need<int32_t>(N);
        int32_t* t_coord = tc.need<int32_t>(size_t(N) * 3);
        uint8_t* t_type = tt.need<uint8_t>(N);
        int32_t* nbr = nb.need<int32_t>(size_t(N) * 6);
        int32_t* par = pa.need<int32_t>(N);
        int32_t* chi = ch.need<int32_t>(size_t(N) * 8);
        std::vector<Buf> stile(L_);
        SlotTiles st{};
        for (int l = 0; l < L_; ++l) {
            const size_t size = topo_.table_[l].size();
            st.p[l] = stile[l].need<int32_t>(size);
            hip_ok(hipMemsetAsync(st.p[l], 0xff, size * sizeof(int32_t), st_), "slot map");
            if (cnt_[l])
                k::b_final_level<<<grid_of(cnt_[l]), kBlock, 0, st_>>>(slot_[l], comp_[l], cnt_[l], l, start[l], t_level,
                                                                       t_slot, t_coord, t_type, st.p[l]);
        }
        if (N) k::b_final_links<<<grid_of(N), kBlock, 0, st_>>>(tabs(), st, N, t_level, t_coord, nbr, par, chi, err());
        Topology& T = topo_;
        T.t_level.resize(N);
        T.t_slot.resize(N);
        T.t_coord.resize(N);
        T.t_type.resize(N);
        T.nbr.resize(size_t(N) * 6);
        T.parent.resize(N);
        T.child.resize(size_t(N) * 8);
        if (N) {
            auto d2h = [&](void* h, const void* d, size_t b) {
                hip_ok(hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, st_), "download tables");
            };
            d2h(T.t_level.data(), t_level, N * sizeof(int32_t));
            d2h(T.t_slot.data(), t_slot, N * sizeof(int32_t));
            static_assert(sizeof(Coord) == 3 * sizeof(int32_t), "Coord layout");
            d2h(T.t_coord.data(), t_coord, size_t(N) * 3 * sizeof(int32_t));
            d2h(T.t_type.data(), t_type, N);
            d2h(T.nbr.data(), nbr, size_t(N) * 6 * sizeof(int32_t));
            d2h(T.parent.data(), par, N * sizeof(int32_t));
            d2h(T.child.data(), chi, size_t(N) * 8 * sizeof(int32_t));
        }
        read_back();
        T.level_start = start;
        for (int l = 0; l < L_; ++l) {
            T.slot_tile_[l].assign(T.table_[l].size(), -1);
            for (int t = start[l]; t < start[l + 1]; ++t) T.slot_tile_[l][T.t_slot[t]] = t;
        }
        T.finalized_ = true;
    }

    static bool target_on_device(const cmg_target& t)
    {
        switch (t.kind) {
        case CMG_TARGET_UNIFORM:
        case CMG_TARGET_CENTER_POINT:
        case CMG_TARGET_SPHERE_EMPTY:
        case CMG_TARGET_SPHERE_SOLID:
            return true;
        case CMG_TARGET_SDF_BANDS:
            return t.sdf_kind == CMG_SDF_SPHERE || t.sdf_kind == CMG_SDF_TORUS;
        default:
            return false;
        }
    }

private:
    // device-resident counters, read back in one copy
    struct Small {
        int err;                           // 1: a level table is full; else an insertion did not finish
        int pad[15];
        int32_t count[kMaxLevels];         // compressed count per level (scan totals)
        int32_t used[kMaxLevels];          // used slots per level
        int32_t refined[kMaxLevels];       // flagged LEAF tiles per level in this refine pass
        unsigned int fresh[kMaxLevels];    // new keys of the level's insertion round (0 between rounds)
    };

    void release()
    {
        if (st_) (void)hipStreamSynchronize(st_);
        if (block_) (void)hipFree(block_);
        block_ = nullptr;
        for (int l = 0; l < L_; ++l) {
            slot_[l] = nullptr;
            word_[l] = nullptr;
            comp_[l] = nullptr;
        }
        if (small_) (void)hipFree(small_);
        small_ = nullptr;
        if (pinned_) (void)hipHostFree(pinned_);
        pinned_ = nullptr;
        if (st_) (void)hipStreamDestroy(st_);
        st_ = nullptr;
    }

    Tabs tabs() const
    {
        Tabs T{};
        for (int l = 0; l < L_; ++l) {
            T.slot[l] = slot_[l];
            T.mask[l] = uint32_t(topo_.table_[l].size()) - 1u;
        }
        T.num_levels = L_;
        return T;
    }
    int* err() { return &small_->err; }

    // every device counter to the host (one synchronisation); throws on the error flag
    void read_back()
    {
        hip_ok(hipMemcpyAsync(pinned_, small_, sizeof(Small), hipMemcpyDeviceToHost, st_), "counters");
        hip_ok(hipStreamSynchronize(st_), "counters");
        host_ = *pinned_;
        if (host_.err == 1) throw std::runtime_error("hash table full (device octree construction)");
        if (host_.err)
            throw std::runtime_error("device octree construction: insertion did not finish (code " +
                                     std::to_string(host_.err) + ")");
    }

    // exclusive prefix sums of n 0/1 flags (or of the used bits of n slots) into pos; the
    // total goes to *total (device)
    void scan(const int32_t* flag, int n, int32_t* pos, int32_t* total, const int4* slots = nullptr)
    {
        const int nt = (n + k::kScanTile - 1) / k::kScanTile;
        if (nt == 1) {   // one workgroup does it all
            k::b_scan_tiles<<<1, kBlock, 0, st_>>>(flag, slots, n, pos, nullptr, total);
            return;
        }
        int32_t* tsum = tsum_.need<int32_t>(nt);
        k::b_scan_tiles<<<nt, kBlock, 0, st_>>>(flag, slots, n, pos, tsum, nullptr);
        k::b_scan_sums<<<1, kBlock, 0, st_>>>(tsum, nt, total);
        k::b_scan_add<<<grid_of(n), kBlock, 0, st_>>>(pos, n, tsum);
    }

    // one round of ordered insertion of the valid candidates (at most nc; with limit, those
    // below per * *limit) into a level table; the room check (small_->fresh[level] new keys
    // next to small_->used[level]) and its bookkeeping run on the device
    void insert(int level, const int4* cand, const uint8_t* valid, int nc, const int32_t* limit, int per)
    {
        dirty_[level] = true;
        const size_t size = topo_.table_[level].size();
        k::b_insert<<<grid_of(nc), kBlock, 0, st_>>>(word_[level], uint32_t(size) - 1u, cand, valid, nc, limit, per,
                                                     small_->fresh + level, small_->used + level, err());
        k::b_materialize<<<grid_of(long(size)), kBlock, 0, st_>>>(word_[level], slot_[level], uint32_t(size), cand,
                                                                  small_->fresh + level, small_->used + level);
    }

    Topology& topo_;
    int L_;
    hipStream_t st_ = nullptr;
    Small* small_ = nullptr;
    Small* pinned_ = nullptr;   // host-pinned landing buffer of the counters
    Small host_{};
    void* block_ = nullptr;     // the level tables in one allocation
    int4* slot_[kMaxLevels] = {};
    unsigned long long* word_[kMaxLevels] = {};
    int32_t* comp_[kMaxLevels] = {};
    int cnt_[kMaxLevels] = {};
    bool dirty_[kMaxLevels] = {};   // an insertion round ran since the level's last compression
    int max_level_ = -1;
    Buf recbuf_, tgt_, gather_, flag_, pos_, tsum_, cand_, valid_, dword_, drank_, dslot_;
};

void device_rebuild(Topology& topo)
{
    DeviceBuilder b(topo);
    b.upload();
    b.compress();
    b.spawn_ghosts();
    b.download();
}

void device_iterative_refine(Topology& topo, const cmg_target& t)
{
    DeviceBuilder b(topo);
    b.upload();
    for (;;) {
        const int n = b.refine_step(t);
        b.spawn_ghosts();
        if (n == 0) break;
    }
    b.download();
}

void device_build(Topology& topo, const cmg_target& t)
{
    DeviceBuilder b(topo);
    b.upload();
    b.compress();
    b.spawn_ghosts();
    for (;;) {
        const int n = b.refine_step(t);
        b.spawn_ghosts();
        if (n == 0) break;
    }
    b.finalize();
}

void device_finalize(Topology& topo)
{
    DeviceBuilder b(topo);
    b.upload();
    b.finalize();
}

}  // namespace cmg
