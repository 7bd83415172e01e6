// cmg_solver.cpp -- device grid upload, launch lists, kernel-level operations and
// the PCG / FAS multigrid driver (AMGSolver::solve, src/AMGSolver.cu:1399-1549).
//
// Everything of one grid runs on the grid's own stream.  One PCG iteration
// (operator, reductions, scalar recurrence, the whole multigrid cycle) is
// captured into a hipGraph and replayed; every kernel of the iteration exits
// at once when the device-side convergence flag is set, so replays past
// convergence are no-ops and the host only polls the flag.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "cmg_kernels.hpp"

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) throw std::runtime_error(std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

struct cmg_solver {
    cmg_grid* g = nullptr;
    int gen = 0;                   // the grid's generation this solver's buffers were sized for
    int family = CMG_SOLVER_AMG;   // AMGSolver, GMGSolver or CMGSolver
    int c0 = 11;
    float Rm = 0.5f, Rr = 1.0f, P = 1.0f;
    double* S = nullptr;
    double* hS = nullptr;          // pinned ring of scalar snapshots
    // host-mapped coherent word the PCG iteration's stopping test publishes: tests run so far
    // times 8 plus the stop code; sig is its device address
    unsigned long long* hsig = nullptr;
    unsigned long long* sig = nullptr;
    double* partial = nullptr;
    double* partial2 = nullptr;
    float* tmp = nullptr;          // Jacobi ping-pong buffer
    float* scratch = nullptr;      // cycle scratch of flexible CG (keeps Ap)
    float* net_w = nullptr;        // learned tile smoother weights (7776 floats)
    unsigned int* lcount = nullptr;   // learned steps of the last solve: tiles judged, tiles kept
    uint16_t* q16 = nullptr;          // FP16 copies of the four coefficient channels / h (cycle_coef_fp16)
    unsigned int* qbad = nullptr;     // workgroups whose FP16 copy has a value that does not fit
    // exact per-level codes of the face coefficients (Coef::w / lut): cell words, tables,
    // build workspace, and the levels coded by the last build
    uint32_t* cw = nullptr;
    float* clut = nullptr;
    uint32_t* cws = nullptr;
    uint32_t coded = 0, wide = 0;   // levels with codes; those with ten-bit codes
    uint32_t* czw = nullptr;        // per tile, the words across its z faces (Coef::zw)
    double *x64 = nullptr, *b64 = nullptr, *r64 = nullptr;
    // level-0 dense inverse
    double* inv0 = nullptr;
    int32_t* cells0 = nullptr;
    int32_t* zero0 = nullptr;
    int n0 = 0, nz0 = 0;
    bool direct_ready = false;
    bool direct_pure_neumann = false;
    // generations of the cell types and of channels c0..c0+3 that each piece of derived data
    // was computed from; it is recomputed as soon as one of them moves
    struct Stamp {
        uint64_t t = ~0ull, c[4] = {~0ull, ~0ull, ~0ull, ~0ull};
        bool operator==(const Stamp& o) const
        {
            return t == o.t && c[0] == o.c[0] && c[1] == o.c[1] && c[2] == o.c[2] && c[3] == o.c[3];
        }
    };
    Stamp direct_stamp, q16_stamp, derive_stamp, code_stamp;
    bool derive_ok = false;
    // captured iteration
    hipGraphExec_t gexec = nullptr;
    std::string gkey;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace cmg {

cmg_grid* solver_grid(const cmg_solver* s) { return s ? s->g : nullptr; }

namespace {

template <typename T>
T* dmalloc(size_t n)
{
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T*>(p);
}

uint64_t spread3(uint32_t x)
{
    uint64_t v = x & 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// per tile: the kRecWords-int4 record of the one-tile stencil kernels (see LaunchList::rec)
using TileRecords = std::vector<int4>;

// Morton order (coordinates scaled to the finest level), then an XCD-aware
// permutation: workgroup b runs on XCD b % 8.  The Morton sequence is cut into blocks
// of about 512 tiles (a multiple of 8 blocks) and block k goes to XCD k % 8, which
// walks its blocks in order: inside a block, face neighbours share the XCD's L2 (halo
// reuse), and at any moment the 8 XCDs stream 8 adjacent blocks, so each array is read
// from a narrow window of memory (DRAM locality).  (A 1024^3-equivalent finest-level
// sweep on MI355X: 512-tile blocks 68 us, one contiguous chunk per XCD 75.5 us, no
// XCD grouping 90 us; blocks of 256 to 4096 tiles all 69-71 us.)

// tiles a Morton block holds (CMG_XCD_BLOCK, 64..65536, default 512)
int xcd_block_tiles()
{
    static const int v = [] {
        const char* e = getenv("CMG_XCD_BLOCK");
        const int b = e ? atoi(e) : 512;
        return b >= 64 && b <= 65536 ? b : 512;
    }();
    return v;
}

LaunchList make_list(const Topology& topo, std::vector<int32_t> tiles, const TileRecords& trec, bool with_meta = false)
{
    LaunchList l;
    l.tiles = int(tiles.size());
    if (tiles.empty()) return l;
    const int L = topo.max_level();
    std::vector<std::pair<uint64_t, int32_t>> key(tiles.size());
    for (size_t i = 0; i < tiles.size(); ++i) {
        const int t = tiles[i];
        const int sh = L - topo.t_level[t];
        const Coord& c = topo.t_coord[t];
        // z in the lowest bit of each triple: z-neighbours are the closest in launch order, so the
        // strided z-faces of a tile's halo (one 32-byte sector a cell) are the most likely to be
        // in L2, and x-faces (contiguous) take the long Morton distances.  (MI355X, finest sweep
        // of the 1024^3-equivalent torus workload: 4.2 % less fetched from HBM than with x in the
        // lowest bit, at the same time.)
        uint64_t m = spread3(uint32_t(c[2]) << sh) | (spread3(uint32_t(c[1]) << sh) << 1) | (spread3(uint32_t(c[0]) << sh) << 2);
        key[i] = {m, t};
    }
    std::sort(key.begin(), key.end());
    std::vector<int32_t> order;
    const int n = int(key.size());
    constexpr int kXcd = 8;
    const int kBlock = xcd_block_tiles();
    if (n < 64) {
        for (auto& k : key) order.push_back(k.second);
    } else {
        // rounds of 8 blocks; the block size spreads n over whole rounds, so that padding
        // (-1 entries) stays below 8 entries a block
        const int rounds = (n + kXcd * kBlock - 1) / (kXcd * kBlock);
        const int bs = (n + kXcd * rounds - 1) / (kXcd * rounds);
        order.assign(size_t(bs) * kXcd * rounds, -1);
        for (int b = 0; b < int(order.size()); ++b) {
            const int x = b % kXcd, kk = b / kXcd;            // the kk-th workgroup of XCD x
            const int pos = (x + kXcd * (kk / bs)) * bs + kk % bs;
            order[b] = pos < n ? key[pos].second : -1;
        }
    }
    l.n = int(order.size());
    l.order = order;
    l.d = dmalloc<int32_t>(order.size());
    HIPCHK(hipMemcpy(l.d, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    {
        std::vector<int4> rec(order.size() * kRecWords, make_int4(0, 0, 0, 0));
        for (size_t i = 0; i < order.size(); ++i) {
            const int t = order[i];
            if (t < 0) {
                rec[i * kRecWords] = make_int4(-1, 0, 0, -1);
                continue;
            }
            for (int k = 0; k < kRecWords; ++k) rec[i * kRecWords + k] = trec[size_t(t) * kRecWords + k];
        }
        l.rec = dmalloc<int4>(rec.size());
        HIPCHK(hipMemcpy(l.rec, rec.data(), rec.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    if (with_meta) {
        // a tile's parent, octant and type next to it, so that one load gives a
        // prolongation workgroup all it needs to address the parent
        std::vector<int4> meta(order.size(), make_int4(-1, -1, 0, 0));
        for (size_t i = 0; i < order.size(); ++i) {
            const int t = order[i];
            if (t < 0) continue;
            const Coord& c = topo.t_coord[t];
            const int oct = (c[0] & 1) | ((c[1] & 1) << 1) | ((c[2] & 1) << 2);
            meta[i] = make_int4(t, topo.parent[t], oct | (int(topo.t_type[t]) << 8), 0);
        }
        l.meta = dmalloc<int4>(meta.size());
        HIPCHK(hipMemcpy(l.meta, meta.data(), meta.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    return l;
}

void free_list(LaunchList& l)
{
    if (l.d) (void)hipFree(l.d);
    if (l.meta) (void)hipFree(l.meta);
    if (l.rec) (void)hipFree(l.rec);
    l = LaunchList{};
}

Coef coef(DevGrid& d, int c0)
{
    Coef c{};
    for (int a = 0; a < 4; ++a) c.c[a] = dev_channel(d, c0 + a);
    return c;
}

bool valid_channel(int c) { return c >= 0 && c < kChannels; }

}  // namespace

float* dev_channel(DevGrid& d, int c)
{
    if (!d.ch[c]) {
        d.ch[c] = dmalloc<float>(size_t(d.N) * kCells);
        HIPCHK(hipMemsetAsync(d.ch[c], 0, size_t(d.N) * kCells * sizeof(float), d.stream));
        HIPCHK(hipStreamSynchronize(d.stream));
    }
    return d.ch[c];
}

void dev_free(DevGrid& d)
{
    for (void* p : {(void*)d.ttype, (void*)d.toct, (void*)d.tlevel, (void*)d.nbr, (void*)d.parent, (void*)d.child,
                    (void*)d.halo, (void*)d.ghalo, (void*)d.tcoord, (void*)d.acount,
                    (void*)d.ctype, (void*)d.partial, (void*)d.scal, (void*)d.idle, (void*)d.roi, (void*)d.roi_locked,
                    (void*)d.roi_min, (void*)d.roi_max})
        if (p) (void)hipFree(p);
    for (auto& c : d.ch)
        if (c) (void)hipFree(c), c = nullptr;
    free_list(d.leafs);
    free_list(d.all);
    free_list(d.ghosts);
    free_list(d.leafghost);
    free_list(d.cyclezero);
    free_list(d.cyclezero_pull);
    free_list(d.finest_ghosts);
    for (int l = 0; l < kMaxLevels; ++l) {
        free_list(d.subtree[l]);
        free_list(d.nonleaf[l]);
        free_list(d.prolong[l]);
        free_list(d.leafat[l]);
    }
    if (d.stream) (void)hipStreamDestroy(d.stream);
    d = DevGrid{};
}

void dev_upload(DevGrid& d, const Topology& topo)
{
    dev_free(d);
    HIPCHK(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    const int N = topo.num_tiles();
    d.N = N;
    d.num_levels = topo.num_levels();
    d.max_level = topo.max_level();
    d.h0 = topo.h0();
    d.level_start = topo.level_start;
    std::vector<uint8_t> oct(N);
    TileRecords trec;
    for (int t = 0; t < N; ++t)
        oct[t] = uint8_t((topo.t_coord[t][0] & 1) | ((topo.t_coord[t][1] & 1) << 1) | ((topo.t_coord[t][2] & 1) << 2));
    auto up = [&](auto*& dst, const auto& src) {
        using E = typename std::remove_reference<decltype(src)>::type::value_type;
        dst = reinterpret_cast<std::remove_reference_t<decltype(dst)>>(dmalloc<E>(src.size()));
        if (!src.empty()) HIPCHK(hipMemcpy((void*)dst, src.data(), src.size() * sizeof(E), hipMemcpyHostToDevice));
    };
    up(d.ttype, topo.t_type);
    up(d.toct, oct);
    up(d.tlevel, topo.t_level);
    up(d.nbr, topo.nbr);
    {
        std::vector<int32_t> tc(size_t(N) * 3);
        for (int t = 0; t < N; ++t)
            for (int a = 0; a < 3; ++a) tc[size_t(t) * 3 + a] = topo.t_coord[t][a];
        up(d.tcoord, tc);
        d.acount = dmalloc<unsigned int>(1);
    }
    {
        // halo records: one 32-byte read gives a stencil workgroup everything it
        // needs to address its halo (no dependent per-face lookups)
        std::vector<int32_t> halo(size_t(N) * 8, 0), ghalo(size_t(N) * 8, 0);
        for (int t = 0; t < N; ++t) {
            uint32_t ty03 = 0, ty45 = 0, goct = 0;
            for (int f = 0; f < 6; ++f) {
                const int nb = topo.nbr[size_t(t) * 6 + f];
                const uint32_t ty = nb >= 0 ? topo.t_type[nb] : 0u;
                halo[size_t(t) * 8 + f] = nb;
                if (f < 4) ty03 |= ty << (8 * f);
                else ty45 |= ty << (8 * (f - 4));
                // every face neighbour's parent and octant: GHOST values of the operator, the
                // pulled prolongation of the cycle's finest post-sweeps
                ghalo[size_t(t) * 8 + f] = nb >= 0 ? topo.parent[nb] : -1;
                if (nb >= 0) goct |= uint32_t(oct[nb] & 7) << (3 * f);
            }
            halo[size_t(t) * 8 + 6] = int32_t(ty03);
            halo[size_t(t) * 8 + 7] = int32_t(ty45 | (uint32_t(topo.t_type[t]) << 24));
            ghalo[size_t(t) * 8 + 6] = int32_t(goct);
        }
        up(d.halo, halo);
        up(d.ghalo, ghalo);
        trec.assign(size_t(N) * kRecWords, make_int4(0, 0, 0, 0));
        for (int t = 0; t < N; ++t) {
            const int32_t* hh = &halo[size_t(t) * 8];
            const int32_t* gg = &ghalo[size_t(t) * 8];
            const int32_t* cc = &topo.child[size_t(t) * 8];
            int4* r = &trec[size_t(t) * kRecWords];
            r[0] = make_int4(t, topo.t_level[t], int(topo.t_type[t]) | (int(oct[t]) << 8), topo.parent[t]);
            r[1] = make_int4(hh[0], hh[1], hh[2], hh[3]);
            r[2] = make_int4(hh[4], hh[5], hh[6], hh[7]);
            r[3] = make_int4(gg[0], gg[1], gg[2], gg[3]);
            r[4] = make_int4(gg[4], gg[5], gg[6], gg[7]);
            r[5] = make_int4(cc[0], cc[1], cc[2], cc[3]);
            r[6] = make_int4(cc[4], cc[5], cc[6], cc[7]);
        }
    }
    up(d.parent, topo.parent);
    up(d.child, topo.child);
    d.ctype = dmalloc<uint8_t>(size_t(N) * kCells);
    HIPCHK(hipMemset(d.ctype, 0, size_t(N) * kCells));
    std::vector<int32_t> leafs, all, ghosts, leafghost, cyclezero, cyclezero_pull, finest_ghosts;
    std::vector<int32_t> sub[kMaxLevels], nl[kMaxLevels], pr[kMaxLevels], la[kMaxLevels];
    for (int t = 0; t < N; ++t) {
        const uint8_t ty = topo.t_type[t];
        const int l = topo.t_level[t];
        all.push_back(t);
        if (ty & LEAF) leafs.push_back(t);
        if (ty & GHOST) ghosts.push_back(t);
        if (ty & (LEAF | GHOST)) leafghost.push_back(t);
        if (ty & (LEAF | NONLEAF)) sub[l].push_back(t);
        if (ty & NONLEAF) nl[l].push_back(t);
        if (ty & LEAF) la[l].push_back(t);
        if ((ty & GHOST) || ((ty & LEAF) && l < topo.max_level())) cyclezero.push_back(t);
        if (((ty & GHOST) || (ty & LEAF)) && l < topo.max_level()) cyclezero_pull.push_back(t);
        if ((ty & GHOST) && l == topo.max_level()) finest_ghosts.push_back(t);
        pr[l].push_back(t);
    }
    d.leafs = make_list(topo, leafs, trec);
    d.all = make_list(topo, all, trec);
    d.ghosts = make_list(topo, ghosts, trec);
    d.leafghost = make_list(topo, leafghost, trec);
    d.cyclezero = make_list(topo, cyclezero, trec);
    d.cyclezero_pull = make_list(topo, cyclezero_pull, trec);
    d.finest_ghosts = make_list(topo, finest_ghosts, trec);
    d.max_list = std::max(d.leafs.n, d.all.n);
    for (int l = 0; l < topo.num_levels(); ++l) {
        d.subtree[l] = make_list(topo, sub[l], trec);
        d.nonleaf[l] = make_list(topo, nl[l], trec);
        d.prolong[l] = make_list(topo, pr[l], trec, true);
        d.leafat[l] = make_list(topo, la[l], trec);
    }
    // The prolongation rides on the parents' last post-sweep only into levels of at most
    // 16384 tiles: there a launch of its own is latency-bound.  Into larger levels the
    // streaming prolongation kernel is faster than the sweep with its children's traffic
    // (MI355X, 1024^3-equivalent torus workload: 60 us fused against 18.7 + 36.4 into the
    // finest level, 18.1 against 8.5 + 12.3 into the next).
    for (int l = 0; l < topo.num_levels(); ++l) {
        bool ok = l >= 2 && !pr[l].empty() && pr[l].size() <= 16384;
        for (int t : pr[l]) {
            const int p = topo.parent[t];
            ok = ok && p >= 0 && (topo.t_type[p] & (LEAF | NONLEAF)) && topo.t_level[p] == l - 1;
        }
        d.push_ok[l] = ok;
    }
    if (sub[0].size() == 1) {
        const int t = sub[0][0];
        bool alone = true;
        for (int f = 0; f < 6; ++f) alone = alone && topo.nbr[size_t(t) * 6 + f] < 0;
        d.root_alone = alone;
        if (alone && topo.num_levels() > 1 && sub[1].size() == 8 && pr[1].size() == 8) {
            bool kids = true;
            d.bottom_root = t;
            d.bottom_ghost = 0;
            for (int o = 0; o < 8; ++o) {
                const int c = topo.child[size_t(t) * 8 + o];
                kids = kids && c >= 0 && topo.parent[c] == t && (topo.t_type[c] & (LEAF | NONLEAF));
                int e = -1;
                for (int i = 0; i < int(d.subtree[1].order.size()); ++i)
                    if (d.subtree[1].order[i] == c) e = i;
                d.bottom_entry[o] = e;
                d.bottom_tile[o] = c;
                d.bottom_type[o] = c >= 0 ? topo.t_type[c] : 0;
                kids = kids && e >= 0;
                for (int q = 0; q < 8; ++q) {
                    const int gc = c >= 0 ? topo.child[size_t(c) * 8 + q] : -1;
                    d.bottom_child[8 * o + q] = gc;
                    if (gc >= 0 && (topo.t_type[gc] & GHOST)) d.bottom_ghost |= 1ull << (8 * o + q);
                }
            }
            d.bottom_ok = kids;
        }
    }
    {
        int dev = 0, cus = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        d.wg_cap = std::max(8, 8 * cus);
    }
    d.partial = dmalloc<double>(size_t(d.max_list) + 1);
    d.scal = dmalloc<double>(S_COUNT);
    HIPCHK(hipMemset(d.scal, 0, S_COUNT * sizeof(double)));
    d.idle = dmalloc<double>(S_COUNT);
    HIPCHK(hipMemset(d.idle, 0, S_COUNT * sizeof(double)));
    HIPCHK(hipDeviceSynchronize());
}

// ------------------------------------------------------------- single ops ---

// argument errors found while enqueueing (std::invalid_argument) are CMG_ERR_BAD_ARG,
// other logic errors CMG_ERR_STATE, everything else (HIP failures) CMG_ERR_HIP
#define CMG_TRY(...)                                         \
    try {                                                    \
        __VA_ARGS__;                                         \
    } catch (const std::invalid_argument& ex) {              \
        set_error(ex.what());                                \
        return CMG_ERR_BAD_ARG;                              \
    } catch (const std::logic_error& ex) {                   \
        set_error(ex.what());                                \
        return CMG_ERR_STATE;                                \
    } catch (const std::exception& ex) {                     \
        set_error(ex.what());                                \
        return CMG_ERR_HIP;                                  \
    }

static cmg_status check_grid(cmg_grid* g)
{
    if (!g || !g->topo) return CMG_ERR_BAD_ARG;
    if (!g->topo->finalized()) {
        set_error("grid not finalized");
        return CMG_ERR_STATE;
    }
    try {
        ensure_device(g);
    } catch (const std::exception& ex) {
        set_error(ex.what());
        return CMG_ERR_HIP;
    }
    return CMG_OK;
}

cmg_status op_full_neg_lap(cmg_grid* g, int x, int c0, int y)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_channel(x) || !valid_channel(y) || c0 < 0 || c0 + 3 >= kChannels || x == y) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        launch_apply(d, dev_channel(d, x), coef(d, c0), dev_channel(d, y), nullptr, nullptr, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

cmg_status op_propagate_ghosts(cmg_grid* g, int ch)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_channel(ch)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        launch_propagate_ghosts(d, dev_channel(d, ch), d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

cmg_status op_accumulate_leafs(cmg_grid* g, int ch)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_channel(ch)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        launch_accumulate_leafs(d, dev_channel(d, ch), d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

static bool valid_level(const cmg_grid* g, int level) { return level >= 0 && level < g->dev.num_levels; }

cmg_status op_gs_color(cmg_grid* g, int level, int x, int c0, int rhs, int color, double omega)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_level(g, level) || !valid_channel(x) || !valid_channel(rhs) || c0 < 0 || c0 + 3 >= kChannels || (color & ~1))
        return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        launch_gs_color(d, level, dev_channel(d, x), coef(d, c0), dev_channel(d, rhs), color, omega, nullptr, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

cmg_status op_jacobi(cmg_grid* g, int level, int x_in, int x_out, int gch, int c0, int rhs, float omega)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_level(g, level) || !valid_channel(x_in) || !valid_channel(x_out) || !valid_channel(gch) || !valid_channel(rhs) ||
        c0 < 0 || c0 + 3 >= kChannels || x_in == x_out)
        return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        launch_jacobi(d, level, dev_channel(d, x_in), dev_channel(d, x_out), dev_channel(d, gch), coef(d, c0),
                      dev_channel(d, rhs), omega, nullptr, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

cmg_status op_residual_restrict(cmg_grid* g, int level, int x, int c0, int rhs, float Rr)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_level(g, level) || level == 0 || !valid_channel(x) || !valid_channel(rhs) || c0 < 0 || c0 + 3 >= kChannels)
        return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        launch_residual_restrict(d, level, dev_channel(d, x), coef(d, c0), dev_channel(d, rhs), Rr, nullptr, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

cmg_status op_fas_save(cmg_grid* g, int level, int x, int c0, int rhs, int x0)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_level(g, level) || !valid_channel(x) || !valid_channel(rhs) || !valid_channel(x0) || c0 < 0 || c0 + 3 >= kChannels)
        return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        launch_fas_save(d, level, dev_channel(d, x), coef(d, c0), dev_channel(d, rhs), dev_channel(d, x0), nullptr, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

cmg_status op_fas_prolong(cmg_grid* g, int level, int x, int x0, float P)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_level(g, level) || level == 0 || !valid_channel(x) || !valid_channel(x0)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        launch_fas_prolong(d, level, dev_channel(d, x), dev_channel(d, x0), P, nullptr, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

// ---- projection wrappers --------------------------------------------------------

static bool valid_vec(int c) { return c >= 0 && c + 2 < kChannels; }

// AMGVolumeWeightedDivergence[WithoutCoeff]OnLeafs (src/AMGSolver.cu:750-835):
// ghost faces take their parents' values, NONLEAF faces a quarter of their four
// child faces (both stay in u), then b on LEAF tiles; c0 < 0: without coefficients
cmg_status op_divergence(cmg_grid* g, int u0, int ch, int c0)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_vec(u0) || !valid_channel(ch) || (c0 >= 0 && c0 + 3 >= kChannels) || (ch >= u0 && ch <= u0 + 2))
        return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        float* U[3] = {dev_channel(d, u0), dev_channel(d, u0 + 1), dev_channel(d, u0 + 2)};
        for (int a = 0; a < 3; ++a) launch_propagate_ghosts(d, U[a], d.stream);
        launch_faces_to_parents(d, U[0], U[1], U[2], d.stream);
        Coef C{};
        if (c0 >= 0) C = coef(d, c0);
        launch_divergence(d, U[0], U[1], U[2], c0 >= 0 ? &C : nullptr, dev_channel(d, ch), d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

// AMGAddGradientToFace / AMGAddFaceWeightedGradientToFace (src/AMGSolver.cu:634-748)
cmg_status op_add_gradient(cmg_grid* g, int x, int c0, int u0, int weighted)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_channel(x) || c0 < 0 || c0 + 3 >= kChannels || !valid_vec(u0) || (x >= u0 && x <= u0 + 2))
        return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        float* X = dev_channel(d, x);
        launch_propagate_ghosts(d, X, d.stream);
        launch_accumulate_leafs(d, X, d.stream);
        launch_add_gradient(d, X, coef(d, c0), dev_channel(d, u0), dev_channel(d, u0 + 1), dev_channel(d, u0 + 2),
                            weighted != 0, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

// FluidEuler::mixFluidAndSolidVelocityOnFaces (cirrus_cutcell/FluidEuler.cu:531-552)
cmg_status op_mix_velocity(cmg_grid* g, int c0, int uf0, int us0, int um0)
{
    if (auto st = check_grid(g)) return st;
    if (c0 < 0 || c0 + 3 >= kChannels || !valid_vec(uf0) || !valid_vec(us0) || !valid_vec(um0)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        const float* F[3] = {dev_channel(d, uf0), dev_channel(d, uf0 + 1), dev_channel(d, uf0 + 2)};
        const float* Sv[3] = {dev_channel(d, us0), dev_channel(d, us0 + 1), dev_channel(d, us0 + 2)};
        float* M[3] = {dev_channel(d, um0), dev_channel(d, um0 + 1), dev_channel(d, um0 + 2)};
        launch_mix_velocity(d, coef(d, c0), F, Sv, M, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

// NormSync (src/PoissonGrid.cu:322-375): order -1 = max |x|; order p > 0 =
// (sum |x|^p w / sum w)^(1/p) over LEAF cells whose type is in cell_types, w = 1
// or the cell volume; tile partials are summed on the host in tile order
cmg_status op_norm(cmg_grid* g, int order, int ch, int volume_weighted, int cell_types, double* out)
{
    if (auto st = check_grid(g)) return st;
    if (!out || !valid_channel(ch) || order == 0 || order < -1 || (order == -1 && volume_weighted)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        double* w = dmalloc<double>(size_t(d.max_list) + 1);
        std::vector<double> hv, hw;
        try {
            const int n = launch_power_sum(d, dev_channel(d, ch), order, volume_weighted != 0, cell_types, d.partial, w, d.stream);
            hv.resize(size_t(n));
            hw.resize(size_t(n));
            if (n) {
                HIPCHK(hipMemcpyAsync(hv.data(), d.partial, n * sizeof(double), hipMemcpyDeviceToHost, d.stream));
                HIPCHK(hipMemcpyAsync(hw.data(), w, n * sizeof(double), hipMemcpyDeviceToHost, d.stream));
            }
            HIPCHK(hipStreamSynchronize(d.stream));
        } catch (...) {
            (void)hipFree(w);
            throw;
        }
        (void)hipFree(w);
        double vs = 0, ws = 0;
        for (size_t i = 0; i < hv.size(); ++i) {
            vs = order == -1 ? std::max(vs, hv[i]) : vs + hv[i];
            ws += hw[i];
        }
        if (order == -1) *out = vs;
        else if (ws == 0) *out = 0;
        else if (order == 1) *out = vs / ws;
        else if (order == 2) *out = std::sqrt(vs / ws);
        else *out = std::pow(vs / ws, 1.0 / order);
    });
    return CMG_OK;
}

cmg_status op_dot(cmg_grid* g, int a, int b, double* out)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_channel(a) || !valid_channel(b) || !out) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        auto& d = g->dev;
        const int np = launch_dot(d, d.leafs, dev_channel(d, a), dev_channel(d, b), d.partial, nullptr, d.stream);
        launch_finish(d.partial, nullptr, np, d.scal, FIN_SUM, d.stream);
        double hs[S_COUNT];
        HIPCHK(hipMemcpyAsync(hs, d.scal, sizeof(hs), hipMemcpyDeviceToHost, d.stream));
        HIPCHK(hipStreamSynchronize(d.stream));
        *out = d.leafs.n ? hs[S_TMP] : 0.0;
    });
    return CMG_OK;
}

// ---------------------------------------------------------------- solver ----

namespace {

enum { CH_X = 0, CH_R = 1, CH_P = 2, CH_AP = 3, CH_Z = 4 };

// the generations the solver's operator data has now
cmg_solver::Stamp stamp_now(const cmg_solver* s)
{
    cmg_solver::Stamp st;
    st.t = s->g->types_gen;
    for (int a = 0; a < 4; ++a) st.c[a] = s->g->ch_gen[s->c0 + a];
    return st;
}

// exact per-level codes of the face coefficients for the one-tile stencil kernels (Coef::w):
// rebuilt whenever the coefficient channels or the cell types have been written since the
// last build; a level whose face coefficients take at most kCodes - 1 distinct values gets
// one-byte codes, one with at most kWideCodes - 1 ten-bit codes, any other keeps the FP32
// channels
void prepare_coef_codes(cmg_solver* s)
{
    if (s->family != CMG_SOLVER_AMG || !coef_codes_allowed()) return;
    DevGrid& d = s->g->dev;
    const cmg_solver::Stamp stamp = stamp_now(s);
    if (s->cw && s->code_stamp == stamp) return;
    s->code_stamp = cmg_solver::Stamp{};
    s->coded = 0;
    s->wide = 0;
    if (!s->cw) {
        s->cw = dmalloc<uint32_t>(size_t(d.N) * kCells);
        s->clut = dmalloc<float>(size_t(kMaxLevels) * kWideCodes);
        s->cws = dmalloc<uint32_t>(coef_code_words());
        s->czw = dmalloc<uint32_t>(size_t(d.N) * 128);
    }
    launch_coef_codes(d, coef(d, s->c0), s->cws, s->clut, s->cw, d.stream);
    launch_code_zfaces(d, s->cw, s->czw, d.stream);
    uint32_t ok[kMaxLevels];
    HIPCHK(hipMemcpyAsync(ok, s->cws + coef_code_result_offset(), sizeof(ok), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    for (int l = 0; l <= d.max_level && l < kMaxLevels; ++l) {
        if (ok[l]) s->coded |= 1u << l;
        if (ok[l] == 2) s->wide |= 1u << l;
    }
    s->code_stamp = stamp;
}

// the solver's operator for the stencil kernels: the FP32 channels, and the codes of the
// levels that have them when they are current (built from the channels as they are now)
Coef solver_coef(const cmg_solver* s)
{
    DevGrid& d = s->g->dev;
    Coef c = coef(d, s->c0);
    if (s->family == CMG_SOLVER_AMG && s->cw && s->coded && coef_codes_allowed() && s->code_stamp == stamp_now(s)) {
        c.w = s->cw;
        c.lut = s->clut;
        c.coded = s->coded;
        c.wide = s->wide;
        c.zw = s->czw;
    }
    return c;
}

// dense FP64 inverse of the level-0 operator on its INTERIOR cells, probed
// cell by cell with the same halo rules as the device stencil (level 0 has no
// ghosts); Gauss-Jordan with partial pivoting.  For a pure-Neumann system the
// constant null space is removed by adding mean(D)/n * 1 1^T.
void build_direct(cmg_solver* s, bool pure_neumann)
{
    DevGrid& d = s->g->dev;
    const int lo = d.level_start[0], hi = d.level_start[1];
    const int nt = hi - lo;
    const Topology& topo = *s->g->topo;
    const cmg_solver::Stamp stamp = stamp_now(s);
    // level 0 is the first block of tiles and its face neighbours are level-0 tiles, so only
    // those tiles' types and coefficients are read back
    if (lo != 0) throw std::logic_error("level 0 does not start the tile order");
    const size_t n0cells = size_t(hi) * kCells;
    std::vector<float> C[4];
    std::vector<uint8_t> T(n0cells);
    HIPCHK(hipMemcpy(T.data(), d.ctype, T.size(), hipMemcpyDeviceToHost));
    for (int a = 0; a < 4; ++a) {
        C[a].resize(n0cells);
        HIPCHK(hipMemcpy(C[a].data(), dev_channel(d, s->c0 + a), n0cells * sizeof(float), hipMemcpyDeviceToHost));
    }
    std::vector<int> map(size_t(nt) * kCells, -1);
    std::vector<int32_t> cells, zeros;
    for (int t = lo; t < hi; ++t)
        for (int v = 0; v < kCells; ++v) {
            const size_t idx = size_t(t) * kCells + v;
            if (T[idx] & INTERIOR) {
                map[size_t(t - lo) * kCells + v] = int(cells.size());
                cells.push_back(int32_t(idx));
            } else {
                zeros.push_back(int32_t(idx));
            }
        }
    const int n = int(cells.size());
    if (n > 4096) throw std::runtime_error("level-0 system too large for the dense direct solve");
    std::vector<double> A(size_t(n) * n, 0.0);
    auto unknown = [&](int t, int v) -> int {
        if (t < lo || t >= hi) return -1;
        return T[size_t(t) * kCells + v] == INTERIOR ? map[size_t(t - lo) * kCells + v] : -1;
    };
    for (int r = 0; r < n; ++r) {
        const int t = cells[r] / kCells, v = cells[r] % kCells;
        const double D = C[3][cells[r]];
        if (D == 0) continue;
        A[size_t(r) * n + r] += D;
        const int l[3] = {v >> 6, (v >> 3) & 7, v & 7};
        for (int a = 0; a < 3; ++a)
            for (int sg = -1; sg <= 1; sg += 2) {
                int nl[3] = {l[0], l[1], l[2]};
                nl[a] += sg;
                int tt = t;
                if (nl[a] < 0 || nl[a] > 7) {
                    tt = topo.nbr[size_t(t) * 6 + (sg < 0 ? a : a + 3)];
                    nl[a] &= 7;
                }
                const int nv = cell_index(nl[0], nl[1], nl[2]);
                double off;
                if (sg < 0) off = C[a][cells[r]];
                else if (tt < 0) off = 0;   // empty: x = 0
                else off = C[a][size_t(tt) * kCells + nv];
                if (tt < 0) continue;
                const int c = unknown(tt, nv);
                if (c >= 0) A[size_t(r) * n + c] += off;
            }
    }
    if (pure_neumann && n > 0) {
        double md = 0;
        for (int r = 0; r < n; ++r) md += A[size_t(r) * n + r];
        md /= double(n) * double(n);
        for (auto& a : A) a += md;
    }
    // Gauss-Jordan inverse
    std::vector<double> inv(size_t(n) * n, 0.0);
    for (int i = 0; i < n; ++i) inv[size_t(i) * n + i] = 1.0;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[size_t(i) * n + k]) > std::fabs(A[size_t(piv) * n + k])) piv = i;
        if (A[size_t(piv) * n + k] == 0) throw std::runtime_error("singular level-0 operator");
        if (piv != k)
            for (int j = 0; j < n; ++j) {
                std::swap(A[size_t(k) * n + j], A[size_t(piv) * n + j]);
                std::swap(inv[size_t(k) * n + j], inv[size_t(piv) * n + j]);
            }
        const double dk = 1.0 / A[size_t(k) * n + k];
        for (int j = 0; j < n; ++j) {
            A[size_t(k) * n + j] *= dk;
            inv[size_t(k) * n + j] *= dk;
        }
        // rows are independent once row k is scaled
#pragma omp parallel for schedule(static) if (n >= 256)
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const double f = A[size_t(i) * n + k];
            if (f == 0) continue;
            for (int j = 0; j < n; ++j) {
                A[size_t(i) * n + j] -= f * A[size_t(k) * n + j];
                inv[size_t(i) * n + j] -= f * inv[size_t(k) * n + j];
            }
        }
    }
    if (s->inv0) (void)hipFree(s->inv0);
    if (s->cells0) (void)hipFree(s->cells0);
    if (s->zero0) (void)hipFree(s->zero0);
    s->n0 = n;
    s->nz0 = int(zeros.size());
    s->inv0 = dmalloc<double>(inv.size());
    s->cells0 = dmalloc<int32_t>(cells.size());
    s->zero0 = dmalloc<int32_t>(zeros.size());
    if (n) {
        HIPCHK(hipMemcpy(s->inv0, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(s->cells0, cells.data(), cells.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    if (!zeros.empty()) HIPCHK(hipMemcpy(s->zero0, zeros.data(), zeros.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    s->direct_ready = true;
    s->direct_pure_neumann = pure_neumann;
    s->direct_stamp = stamp;
    // a captured iteration holds the old inverse's address and the old cell counts
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec), s->gexec = nullptr;
    s->gkey.clear();
}

// the level-0 inverse is (re)built when missing, when the null-space handling changes, or when
// the cell types or the coefficient channels have been written since it was built
void ensure_direct(cmg_solver* s, bool pure_neumann)
{
    if (!s->direct_ready || s->direct_pure_neumann != pure_neumann || !(s->direct_stamp == stamp_now(s)))
        build_direct(s, pure_neumann);
}

struct CycleCtx {
    cmg_solver* s;
    const cmg_solve_opts* o;
    float *X, *B, *X0;
    Coef C;
    const double* S;
    bool zero_done = false;   // the finest level's first sweep from x = 0 is already in s->tmp
    double* rz = nullptr;     // the last finest-level post-sweep writes its b.x partials here
};

// slice of the r.z partials written by a level's last post-sweep: one per list entry,
// the finest level first
int rz_offset(const DevGrid& d, int level)
{
    int off = 0;
    for (int l = d.max_level; l > level; --l) off += d.subtree[l].n;
    return off;
}

// `x_zero`: every x of the cycle is still zero (first sweep at the finest level);
// `fas_save`: the level's FAS save is still due and rides on the first sweep;
// `push`: the FAS prolongation to level + 1 comes right after these sweeps and may ride
// on the last one.  Returns whether it did.
// `pull`: the FAS prolongation into this level has not been written; the sweeps compute it
// on the fly (the first one for every cell it reads, the later ones for GHOST neighbours)
bool smooth(const CycleCtx& c, int level, int iters, int order, bool x_zero = false, bool fas_save = false,
            bool push = false, bool pull = false)
{
    DevGrid& d = c.s->g->dev;
    hipStream_t st = d.stream;
    if (c.o->smoother == CMG_SMOOTH_RBGS) {
        for (int it = 0; it < iters; ++it) {
            launch_gs_color(d, level, c.X, c.C, c.B, order, c.o->omega, c.S, st);
            launch_gs_color(d, level, c.X, c.C, c.B, 1 - order, c.o->omega, c.S, st);
        }
        return false;
    }
    float* cur = c.X;
    float* other = c.s->tmp;
    // an even number of ping-pong sweeps ends in X, where the finer level's x is; a single
    // sweep would also carry the FAS save
    push = push && iters >= 2 && iters % 2 == 0 && !x_zero && jacobi_push_available(d, level);
    for (int it = 0; it < iters; ++it) {
        if (it == 0 && x_zero) {
            if (!(c.zero_done && level == d.max_level))
                launch_jacobi_zero(d, level, other, c.C.c[3], c.B, float(c.o->omega), c.S, st, c.C.q[3]);
        }
        else {
            // the level's last post-sweep sums b.x = r.z over its LEAF tiles into its slice of c.rz
            const bool last = c.rz && order == 1 && it == iters - 1;
            const bool pu = push && it == iters - 1;
            launch_jacobi(d, level, cur, other, c.X, c.C, c.B, float(c.o->omega), c.S, st,
                          last ? c.rz + rz_offset(d, level) : nullptr,
                          (fas_save && it == 0) || pu || pull ? c.X0 : nullptr, pu, c.s->P, pull ? (it == 0 ? 1 : 2) : 0);
        }
        std::swap(cur, other);
    }
    if (cur != c.X) launch_copy_list(d, d.subtree[level], cur, c.X, c.S, st);
    return push;
}

// repetitions of the coarse visit (AMGSolver::mu_cycle_repeat_times)
int cycle_mu(const cmg_solve_opts* o)
{
    if (o->mu > 0) return o->mu;
    return o->cycle == CMG_CYCLE_FAS_V ? 1 : 2;
}

void coarsest(const CycleCtx& c)
{
    DevGrid& d = c.s->g->dev;
    if (c.o->coarsest_direct) {
        launch_coarse_direct(c.s->inv0, c.s->cells0, c.s->n0, c.s->zero0, c.s->nz0, c.X, c.B, c.S, d.stream);
    } else if (coarsest_iterate_available(d)) {
        // every pass in one workgroup on the LDS copy of the root tile
        const bool rbgs = c.o->smoother == CMG_SMOOTH_RBGS;
        const int it = rbgs ? c.o->coarsest_iters / 2 : c.o->coarsest_iters;
        launch_coarsest_iterate(d, c.X, c.C, c.B, rbgs, it, rbgs ? it : 0, 0, c.o->omega, float(c.o->omega), c.S, d.stream);
    } else if (c.o->smoother == CMG_SMOOTH_RBGS) {
        smooth(c, 0, c.o->coarsest_iters / 2, 0);
        smooth(c, 0, c.o->coarsest_iters / 2, 1);
    } else {
        smooth(c, 0, c.o->coarsest_iters, 0);
    }
}

// FASMuCycleStep (src/AMGSolver.cu:1257-1277); x_zero: x is still 0 everywhere
// `fas_save`: the FAS save of this level (after the restriction from level + 1)
// has not run yet; the first pre-sweep does it
// `push_up`: the caller prolongs to level + 1 next; the last post-sweep may do it
// (returns whether it did)
bool cycle_step(const CycleCtx& c, int level, bool x_zero, bool fas_save = false, bool push_up = false)
{
    DevGrid& d = c.s->g->dev;
    hipStream_t st = d.stream;
    if (level == 0) {
        if (fas_save) launch_fas_save(d, 0, c.X, c.C, c.B, c.X0, c.S, st);
        coarsest(c);
        return false;
    }
    if (level == 1 && level < d.max_level && !x_zero && cycle_bottom_available(d)) {
        // levels 1 and 0 in one workgroup (the coarsest direct solve keeps its own launch)
        const bool rbgs = c.o->smoother == CMG_SMOOTH_RBGS;
        const int n = c.o->level_iters;
        // (the final x is written in place whatever the sweep count, so any count can prolong)
        const bool push = push_up && !rbgs && n >= 1 && jacobi_push_available(d, 1);
        double* rz = c.rz && !rbgs && n >= 1 ? c.rz + rz_offset(d, 1) : nullptr;
        const bool fas1 = fas_save && !rbgs;
        if (c.o->coarsest_direct) {
            launch_cycle_bottom(d, 1, c.X, c.X0, c.B, c.C, c.s->Rr, c.s->P, rbgs, n, c.o->omega, 0, 0, fas1, false, nullptr,
                                c.S, st);
            for (int i = 0; i < cycle_mu(c.o); ++i) coarsest(c);
            launch_cycle_bottom(d, 2, c.X, c.X0, c.B, c.C, c.s->Rr, c.s->P, rbgs, n, c.o->omega, 0, 0, false, push, rz, c.S,
                                st);
        } else {
            launch_cycle_bottom(d, 3, c.X, c.X0, c.B, c.C, c.s->Rr, c.s->P, rbgs, n, c.o->omega, cycle_mu(c.o),
                                c.o->coarsest_iters, fas1, push, rz, c.S, st);
        }
        return push;
    }
    if (x_zero && level == d.max_level && c.o->learned && c.s->net_w) {
        // the learned tile smoother replaces the pre-smoothing of the finest level
        // (mode 2: runs before it); a tile whose residual it does not lower takes
        // the first Jacobi sweep instead
        launch_net_infer(d, d.subtree[level], c.B, c.C.c[3], c.s->net_w, c.s->tmp, c.S, st);
        launch_learned_accept(d, level, c.s->tmp, c.X, c.C, c.B, float(c.o->omega), c.S, st, c.s->lcount);
        if (c.o->learned == 2) smooth(c, level, c.o->level_iters, 0, false);
    } else {
        smooth(c, level, c.o->level_iters, 0, x_zero, fas_save);
    }
    launch_residual_restrict(d, level, c.X, c.C, c.B, c.s->Rr, c.S, st);
    // damped Jacobi: the coarser level's FAS save rides on its first pre-sweep
    const bool fused = c.o->smoother == CMG_SMOOTH_JACOBI && c.o->level_iters >= 1 && fuse_fas_save_allowed();
    if (!fused) launch_fas_save(d, level - 1, c.X, c.C, c.B, c.X0, c.S, st);
    // In a PCG solve (the r.z fusion is on) the cycle's output is read on LEAF tiles only, so
    // the prolongation into the finest level is left to its post-sweeps, which compute it on
    // the fly: no launch of its own and no sweep of the coarser level carrying it; the
    // finest GHOST tiles then keep their zeros.
    const bool pull = c.rz && level == d.max_level && c.o->smoother == CMG_SMOOTH_JACOBI && c.o->level_iters >= 1 &&
                      pull_prolong_allowed();
    const int mu = cycle_mu(c.o);
    bool pushed = false;
    for (int i = 0; i < mu; ++i) pushed = cycle_step(c, level - 1, false, fused && i == 0, !pull && i == mu - 1);
    if (!pushed && !pull) launch_fas_prolong(d, level, c.X, c.X0, c.s->P, c.S, st);
    return smooth(c, level, c.o->level_iters, 1, false, false, push_up, pull);
}

// AMGSolver::FASFCycle (src/AMGSolver.cu:1358-1387)
void fas_fcycle(const CycleCtx& c)
{
    DevGrid& d = c.s->g->dev;
    hipStream_t st = d.stream;
    launch_fill_all(d, c.X, 0.0f, c.S, st);
    for (int l = d.max_level; l > 0; --l) launch_restrict_source(d, l, c.B, c.s->Rr, c.S, st);
    coarsest(c);
    for (int l = 1; l <= d.max_level; ++l) {
        launch_prolong_add(d, l, c.X, c.s->P, c.S, st);
        cycle_step(c, l, false);
    }
}

// AMGSolver::VCycle (src/AMGSolver.cu:1059-1128): non-FAS; c.X0 holds f, c.B is the working rhs
void vcycle(const CycleCtx& c)
{
    DevGrid& d = c.s->g->dev;
    hipStream_t st = d.stream;
    launch_nonfas_init(d, c.X, c.X0, c.B, c.S, st);
    for (int l = d.max_level; l > 0; --l) {
        smooth(c, l, c.o->level_iters, 0, l == d.max_level);
        launch_residual_restrict(d, l, c.X, c.C, c.B, c.s->Rr, c.S, st);
    }
    coarsest(c);
    for (int l = 1; l <= d.max_level; ++l) {
        launch_prolong_add(d, l, c.X, c.s->P, c.S, st);
        smooth(c, l, c.o->level_iters, 1);
    }
}

// AMGSolver::muCycleStep (src/AMGSolver.cu:1130-1152): non-FAS with flux correction
void mu_step(const CycleCtx& c, int level, bool x_zero)
{
    DevGrid& d = c.s->g->dev;
    hipStream_t st = d.stream;
    if (level == 0) {
        coarsest(c);
        return;
    }
    smooth(c, level, c.o->level_iters, 0, x_zero);
    launch_residual_restrict(d, level, c.X, c.C, c.B, c.s->Rr, c.S, st);
    launch_deduct_flux(d, level - 1, c.X, c.C, c.X0, c.B, c.S, st);
    launch_fill_list(d, d.nonleaf[level - 1], c.X, 0.0f, st, c.S);
    const int mu = cycle_mu(c.o);
    for (int i = 0; i < mu; ++i) mu_step(c, level - 1, false);
    launch_prolong_add(d, level, c.X, c.s->P, c.S, st);
    smooth(c, level, c.o->level_iters, 1);
}

// ---- geometric solvers (GMGSolver, CMGSolver) ------------------------------

int geo_mode(const cmg_solver* s) { return s->family == CMG_SOLVER_GMG ? 0 : 1; }

void geo_gauss_seidel(cmg_solver* s, int iters, int order, int level, float* X, const float* B, const double* S)
{
    DevGrid& d = s->g->dev;
    for (int it = 0; it < iters; ++it) {
        launch_geo_gs(d, geo_mode(s), level, X, B, order, S, d.stream);
        launch_geo_gs(d, geo_mode(s), level, X, B, 1 - order, S, d.stream);
    }
}

// GMGSolver::VCycle (src/GMGSolver.cu:310-387) / CMGSolver::VCycle (src/CMGSolver.cu:372-424)
void geo_vcycle(cmg_solver* s, const cmg_solve_opts* o, float* X, const float* F, float* B, const double* S)
{
    DevGrid& d = s->g->dev;
    hipStream_t st = d.stream;
    const int mode = geo_mode(s);
    launch_nonfas_init(d, X, F, B, S, st);
    for (int l = d.max_level; l > 0; --l) {
        geo_gauss_seidel(s, o->level_iters, 0, l, X, B, S);
        launch_geo_restrict(d, mode, l, X, B, s->Rr, S, st);
    }
    geo_gauss_seidel(s, o->coarsest_iters / 2, 0, 0, X, B, S);
    geo_gauss_seidel(s, o->coarsest_iters / 2, 1, 0, X, B, S);
    for (int l = 1; l <= d.max_level; ++l) {
        // GMG prolongs into LEAF|NONLEAF|GHOST tiles, CMG into LEAF|NONLEAF only
        launch_prolong_add(d, l, X, s->P, S, st, mode == 1);
        geo_gauss_seidel(s, o->level_iters, 1, l, X, B, S);
    }
}

// FullNegativeLaplacian (src/GMGSolver.cu:167-171) / ConservativeFullNegativeLaplacian
// (src/CMGSolver.cu:209-225): y = A x on LEAF tiles (GMG also fills GHOST tiles
// and adds them into their parents); the propagated ghost and accumulated
// NONLEAF values stay in x
void geo_full_apply(cmg_solver* s, float* X, float* Y, bool diag, const double* S)
{
    DevGrid& d = s->g->dev;
    hipStream_t st = d.stream;
    launch_propagate_ghosts(d, X, st);
    if (s->family == CMG_SOLVER_GMG) {
        launch_geo_apply(d, 0, d.leafghost, -1, X, Y, diag, S, st);
        launch_accumulate_ghosts_add(d, Y, 1.0f, S, st);
    } else {
        launch_accumulate_leafs(d, X, st);
        launch_geo_apply(d, 1, d.leafs, -1, X, Y, false, S, st);
    }
}

// y = A x with the solver's operator; returns the number of x.y partials when `partial`
int enqueue_apply(cmg_solver* s, float* X, float* Y, double* partial, const double* S)
{
    DevGrid& d = s->g->dev;
    if (s->family == CMG_SOLVER_AMG) return launch_apply(d, X, solver_coef(s), Y, partial, S, d.stream);
    geo_full_apply(s, X, Y, false, S);
    return partial ? launch_dot(d, d.leafs, X, Y, partial, S, d.stream) : 0;
}

// the preconditioner z ~ A^-1 r of the options' cycle kind.  FAS kinds: x, rhs
// = r (its NONLEAF tiles become coarse right-hand sides), x0 = FAS scratch.
// Non-FAS kinds: x, f = r (kept), x0 = working right-hand side.
void enqueue_cycle_ptr(cmg_solver* s, const cmg_solve_opts* o, float* Xp, float* RHSp, float* X0p, const double* S,
                       bool zero_done = false, double* rz = nullptr)
{
    DevGrid& d = s->g->dev;
    if (s->family != CMG_SOLVER_AMG) {
        geo_vcycle(s, o, Xp, RHSp, X0p, S);
        return;
    }
    Coef C = solver_coef(s);
    if (o->cycle_coef_fp16) {
        if (!s->q16) throw std::logic_error("cycle_coef_fp16: coefficient copies not prepared");
        const size_t n = size_t(d.N) * kCells;
        for (int a = 0; a < 4; ++a) C.q[a] = s->q16 + size_t(a) * n;
    }
    switch (o->cycle) {
    case CMG_CYCLE_VCYCLE:
    case CMG_CYCLE_MU: {
        const CycleCtx c{s, o, Xp, X0p, RHSp, C, S};
        if (o->cycle == CMG_CYCLE_VCYCLE) {
            vcycle(c);
        } else {
            launch_nonfas_init(d, c.X, c.X0, c.B, S, d.stream);
            mu_step(c, d.max_level, true);
        }
        break;
    }
    case CMG_CYCLE_FAS_F: {
        const CycleCtx c{s, o, Xp, RHSp, X0p, C, S};
        fas_fcycle(c);
        break;
    }
    default: {
        // FASMuCycle (src/AMGSolver.cu:1279-1292): x = 0 on every tile, then the step at the finest level.
        // With damped Jacobi the finest LEAF tiles are written before they are read
        // (the first sweep goes to the ping-pong buffer) and every NONLEAF tile
        // takes its x from the restriction, so only GHOST tiles and the coarser LEAF
        // tiles need the zero.
        const CycleCtx c{s, o, Xp, RHSp, X0p, C, S, zero_done, rz};
        // a cycle that pulls its finest prolongation (cycle_step) leaves the finest GHOST tiles
        // at the zeros the solve wrote there once
        const bool pull = rz && o->smoother == CMG_SMOOTH_JACOBI && o->level_iters >= 1 && pull_prolong_allowed();
        if (o->smoother == CMG_SMOOTH_JACOBI && o->level_iters >= 1 && d.max_level >= 1)
            launch_fill_list(d, pull ? d.cyclezero_pull : d.cyclezero, c.X, 0.0f, d.stream, S);
        else
            launch_fill_all(d, c.X, 0.0f, S, d.stream);
        cycle_step(c, d.max_level, true);
    }
    }
}

void enqueue_cycle(cmg_solver* s, const cmg_solve_opts* o, int x, int rhs, int x0, const double* S)
{
    DevGrid& d = s->g->dev;
    enqueue_cycle_ptr(s, o, dev_channel(d, x), dev_channel(d, rhs), dev_channel(d, x0), S);
}

// the preconditioner of one PCG step, z ~ A^-1 r; flexible CG keeps Ap, so
// the cycle's scratch is then a solver-private buffer
void enqueue_precondition(cmg_solver* s, const cmg_solve_opts* o, const double* S, bool zero_done = false,
                          double* rz = nullptr)
{
    DevGrid& d = s->g->dev;
    if (!o->flexible) {
        enqueue_cycle_ptr(s, o, dev_channel(d, CH_Z), dev_channel(d, CH_R), dev_channel(d, CH_AP), S, zero_done, rz);
        return;
    }
    if (!s->scratch) {
        s->scratch = dmalloc<float>(size_t(d.N) * kCells);
        HIPCHK(hipMemset(s->scratch, 0, size_t(d.N) * kCells * sizeof(float)));
    }
    enqueue_cycle_ptr(s, o, dev_channel(d, CH_Z), dev_channel(d, CH_R), s->scratch, S, zero_done);
}

// the preconditioner's first finest-level sweep from x = 0 can ride on the PCG
// x/r update: FAS mu-cycles with damped Jacobi whose input r is final there
bool first_sweep_fusable(const cmg_solver* s, const cmg_solve_opts* o)
{
    return s->family == CMG_SOLVER_AMG && (o->cycle == CMG_CYCLE_FAS_W || o->cycle == CMG_CYCLE_FAS_V) &&
           o->smoother == CMG_SMOOTH_JACOBI && o->level_iters >= 1 && !o->learned && !o->is_pure_neumann &&
           s->g->dev.max_level >= 1 && fuse_first_sweep_allowed();
}

// r.z can be summed by each level's last post-sweep over its LEAF tiles (whose b
// is r and whose output is z, final after that level's post-smoothing), plus a
// dot over the LEAF tiles of level 0 when there are any
bool rz_fusable(const cmg_solver* s, const cmg_solve_opts* o)
{
    return s->family == CMG_SOLVER_AMG && (o->cycle == CMG_CYCLE_FAS_W || o->cycle == CMG_CYCLE_FAS_V) &&
           o->smoother == CMG_SMOOTH_JACOBI && o->level_iters >= 1 && !o->flexible && !o->learned &&
           s->g->dev.max_level >= 1 && fuse_rz_allowed();
}

void enqueue_recenter(cmg_solver* s, float* R)
{
    DevGrid& d = s->g->dev;
    launch_sum_count(d, R, s->partial, s->partial2, s->S, d.stream);
    launch_finish(s->partial, s->partial2, d.leafs.n, s->S, FIN_MEAN, d.stream);
    launch_sub_mean(d, R, s->S, d.stream);
}

// one PCG iteration (AMGSolver.cu:1446-1537)
void enqueue_iteration(cmg_solver* s, const cmg_solve_opts* o)
{
    DevGrid& d = s->g->dev;
    hipStream_t st = d.stream;
    float *X = dev_channel(d, CH_X), *R = dev_channel(d, CH_R), *P = dev_channel(d, CH_P), *AP = dev_channel(d, CH_AP),
          *Z = dev_channel(d, CH_Z);
    const int n1 = enqueue_apply(s, P, AP, s->partial, s->S);
    launch_finish(s->partial, nullptr, n1, s->S, FIN_ALPHA, st);
    const bool fuse = first_sweep_fusable(s, o);
    const uint16_t* dq = o->cycle_coef_fp16 ? s->q16 + size_t(3) * size_t(d.N) * kCells : nullptr;
    // x += alpha p waits for the p update, which reads p anyway (4 B a cell less over the two)
    const int n2 = fuse ? launch_update_xr(d, nullptr, R, P, AP, s->partial, s->S, st, s->tmp, coef(d, s->c0).c[3],
                                           float(o->omega), dq)
                        : launch_update_xr(d, nullptr, R, P, AP, s->partial, s->S, st);
    launch_finish(s->partial, nullptr, n2, s->S, FIN_RR, st, s->sig);
    if (o->is_pure_neumann) enqueue_recenter(s, R);
    const bool frz = rz_fusable(s, o);
    enqueue_precondition(s, o, s->S, fuse, frz ? s->partial : nullptr);
    if (frz) {
        // every level's last post-sweep wrote its LEAF tiles' r.z; level 0 (the direct solve) has
        // LEAF tiles only on one-level grids
        const int nl = rz_offset(d, 0);
        const int n0 = launch_dot(d, d.leafat[0], R, Z, s->partial + nl, s->S, st);
        launch_finish(s->partial, nullptr, nl + n0, s->S, FIN_BETA, st);
    } else if (o->flexible) {
        // flexible CG: beta = z.(r - r_old) / gamma_old (Polak-Ribiere; AMGSolver.cu:1520 uses Fletcher-Reeves)
        const int n3 = launch_dot2(d, d.leafs, Z, R, AP, s->partial, s->partial2, s->S, st);
        launch_finish(s->partial, s->partial2, n3, s->S, FIN_BETA_FLEX, st);
    } else {
        const int n3 = launch_dot(d, d.leafs, R, Z, s->partial, s->S, st);
        launch_finish(s->partial, nullptr, n3, s->S, FIN_BETA, st);
    }
    launch_update_p(d, P, Z, s->S, st, X);
}

// code_forms: the coded levels (bits 0-15) and those with ten-bit codes (bits 16-31)
std::string graph_key(const cmg_solve_opts* o, int derive, uint32_t code_forms)
{
    std::ostringstream k;
    k << o->cycle << ',' << cycle_mu(o) << ',' << o->flexible << ',' << o->learned << ',' << o->smoother << ',' << o->omega << ',' << o->level_iters << ',' << o->coarsest_iters << ','
      << o->coarsest_direct << ',' << o->is_pure_neumann << ',' << o->cycle_coef_fp16;
    int mask = 0, threads = 0;
    get_kernel_forms(&mask, &threads);
    k << ',' << mask << ',' << threads << ',' << launched_threads_key() << ',' << derive << ',' << code_forms;
    return k.str();
}

void ensure_buffers(cmg_solver* s)
{
    DevGrid& d = s->g->dev;
    if (!s->tmp) {
        s->tmp = dmalloc<float>(size_t(d.N) * kCells);
        HIPCHK(hipMemset(s->tmp, 0, size_t(d.N) * kCells * sizeof(float)));
    }
    for (int c : {CH_X, CH_R, CH_P, CH_AP, CH_Z}) dev_channel(d, c);
    if (s->family == CMG_SOLVER_AMG)
        for (int a = 0; a < 4; ++a) dev_channel(d, s->c0 + a);
}

void ensure_fp64(cmg_solver* s)
{
    DevGrid& d = s->g->dev;
    const size_t n = size_t(d.N) * kCells;
    if (!s->x64) {
        s->x64 = dmalloc<double>(n);
        s->b64 = dmalloc<double>(n);
        s->r64 = dmalloc<double>(n);
        HIPCHK(hipMemset(s->x64, 0, n * sizeof(double)));
        HIPCHK(hipMemset(s->b64, 0, n * sizeof(double)));
        HIPCHK(hipMemset(s->r64, 0, n * sizeof(double)));
    }
}

void set_scalars(cmg_solver* s, double tol2, int max_iters, int stride = 0)
{
    double h[S_COUNT] = {};
    h[S_THR2] = tol2;   // relative; the FIN_RHS2 reduction turns it absolute
    h[S_MAXIT] = max_iters;
    h[S_STRIDE] = stride > 0 ? stride : 0;
    DevGrid& d = s->g->dev;
    HIPCHK(hipMemcpyAsync(s->S, h, sizeof(h), hipMemcpyHostToDevice, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));   // h lives on this stack frame
}

void read_scalars(cmg_solver* s, double* h)
{
    DevGrid& d = s->g->dev;
    HIPCHK(hipMemcpyAsync(s->hS, s->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    std::memcpy(h, s->hS, S_COUNT * sizeof(double));
}

// waits until the solve's stopping test has run `tests` times (the count the finish kernel
// publishes in host-mapped memory) and returns its stop code, or -1 if the stream drained first
static int wait_stop_test(cmg_solver* s, int tests)
{
    const unsigned long long want = static_cast<unsigned long long>(tests);
    for (unsigned int spin = 1;; ++spin) {
        unsigned long long w = __atomic_load_n(s->hsig, __ATOMIC_ACQUIRE);
        if ((w >> 3) >= want) return int(w & 7u);
        if ((spin & 255u) == 0) {
            const hipError_t e = hipStreamQuery(s->g->dev.stream);
            if (e == hipSuccess) {
                w = __atomic_load_n(s->hsig, __ATOMIC_ACQUIRE);
                return (w >> 3) >= want ? int(w & 7u) : -1;
            }
            if (e != hipErrorNotReady) HIPCHK(e);
        }
    }
}

// one FP32 PCG solve on channels 0-4; returns the device status code
// stride > 0: the stopping test runs after iterations stride, 2 stride, ...
// (AMGSolver.cu:1487-1497); stride <= 0: exactly max_iters iterations, no test
int pcg32(cmg_solver* s, const cmg_solve_opts* o, double tol, int max_iters, bool warm, int stride, int* iters,
          double* rel_res)
{
    DevGrid& d = s->g->dev;
    hipStream_t st = d.stream;
    float *X = dev_channel(d, CH_X), *R = dev_channel(d, CH_R), *P = dev_channel(d, CH_P), *AP = dev_channel(d, CH_AP),
          *Z = dev_channel(d, CH_Z);
    set_scalars(s, tol * tol, max_iters, stride);
    double h[S_COUNT];
    *iters = 0;
    *rel_res = 0;
    // |b|^2 fixes the tolerance (AMGSolver.cu:1410-1425 with x0 = 0)
    const int nb = launch_dot(d, d.leafs, R, R, s->partial, nullptr, st);
    if (!warm) {
        // on the device, with no host round trip: a zero or non-finite b stops the solve
        // through the stop flag before any cycle runs
        launch_finish(s->partial, nullptr, nb, s->S, FIN_RHS2, st);
        launch_fill_list(d, d.leafs, X, 0.0f, st);
    } else {
        launch_finish(s->partial, nullptr, nb, s->S, FIN_SUM, st);
        read_scalars(s, h);
        const double rhs2 = h[S_TMP];
        if (rhs2 == 0) return 1;
        if (!std::isfinite(rhs2)) return 3;
        // r = b - A x (warm start; the reference always starts from 0, AMGSolver.cu:1407);
        // the stopping test stays relative to |b|
        enqueue_apply(s, X, AP, nullptr, nullptr);
        launch_axpy(d, d.leafs, R, AP, -1.0f, st);
        double thr2 = tol * tol * rhs2;
        thr2 = std::max(thr2, 2.2250738585072014e-308);
        // a warm start may already satisfy the tolerance: no iteration then
        const int nr = launch_dot(d, d.leafs, R, R, s->partial, nullptr, st);
        launch_finish(s->partial, nullptr, nr, s->S, FIN_SUM, st);
        read_scalars(s, h);
        if (h[S_TMP] < thr2) {
            *rel_res = std::sqrt(h[S_TMP] / rhs2);
            return 1;
        }
        double init[S_COUNT] = {};
        init[S_THR2] = thr2;
        init[S_MAXIT] = max_iters;
        init[S_RHS2] = rhs2;
        init[S_STRIDE] = stride > 0 ? stride : 0;
        HIPCHK(hipMemcpyAsync(s->S, init, sizeof(init), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    if (o->is_pure_neumann) enqueue_recenter(s, R);
    enqueue_precondition(s, o, s->S);
    // the iterations' cycles pull their finest prolongation and never write the finest GHOST
    // tiles of z: they are zeroed once here, after this first cycle wrote them
    if (rz_fusable(s, o) && pull_prolong_allowed() && d.max_level >= 1) launch_fill_list(d, d.finest_ghosts, Z, 0.0f, st, s->S);
    launch_copy_list(d, d.leafs, Z, P, s->S, st);
    const int ng = launch_dot(d, d.leafs, Z, R, s->partial, s->S, st);
    launch_finish(s->partial, nullptr, ng, s->S, FIN_GAMMA0, st);
    if (max_iters <= 0) {
        read_scalars(s, h);
        if (h[S_DONE] == 1 && h[S_RHS2] == 0) return 1;   // zero b
        if (h[S_DONE] == 3) return 3;
        *rel_res = 1.0;
        return 4;
    }
    // capture (or reuse) the iteration graph
    if (o->use_graph) {
        const Coef sc = solver_coef(s);
        const std::string key = graph_key(o, d.derive, sc.coded | (sc.wide << 16));
        if (!s->gexec || s->gkey != key) {
            if (s->gexec) (void)hipGraphExecDestroy(s->gexec), s->gexec = nullptr;
            hipGraph_t graph;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            enqueue_iteration(s, o);
            HIPCHK(hipStreamEndCapture(st, &graph));
            HIPCHK(hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            s->gkey = key;
        }
    }
    // Iteration k + 1 is launched as soon as iteration k's stopping test (the r.r reduction,
    // early in the iteration; every stride-th iteration tests, and breakdowns stop at any) has
    // published that the solve goes on: the device never waits for the host, and no iteration
    // runs that the test has ruled out (AMGSolver.cu:1487-1497 reads the residual norm back
    // every sync_stride iterations).
    __atomic_store_n(s->hsig, 0ull, __ATOMIC_RELEASE);
    int launched = 0;
    while (launched < max_iters) {
        if (o->use_graph) HIPCHK(hipGraphLaunch(s->gexec, st));
        else enqueue_iteration(s, o);
        ++launched;
        int stop = wait_stop_test(s, launched);
        if (stop < 0) {
            // the stream drained without the publication: the scalars decide
            read_scalars(s, h);
            stop = int(h[S_DONE]);
        }
        if (stop != 0) break;
    }
    // the last iteration's x += alpha p when its p update did not run (the solve stopped after
    // the x/r update)
    launch_x_pending(d, X, P, s->S, st);
    read_scalars(s, h);
    *iters = int(h[S_ITER]);
    const double rhs2 = h[S_RHS2];
    if (!(rhs2 > 0) || !std::isfinite(rhs2)) return rhs2 == 0 ? 1 : 3;   // zero b: x = 0; non-finite b
    *rel_res = std::sqrt(h[S_RR] / rhs2);
    const int done = int(h[S_DONE]);
    if (stride <= 0) {
        // no synchronised dot: the reference returns (max_iters, -1)
        *rel_res = -1;
        return done == 4 ? 1 : done;
    }
    if (done == 1) return 1;
    if (done == 2 || done == 3) return done;
    return 4;
}

cmg_status status_of(int done)
{
    switch (done) {
    case 1: return CMG_OK;
    case 2: return CMG_ERR_BREAKDOWN;
    case 3: return CMG_ERR_NONFINITE;
    default: return CMG_ERR_NOT_CONVERGED;
    }
}

// squared FP64 residual norm ||b - A x||^2 (x from channel 0 when !from_x64)
double residual_norm2(cmg_solver* s, bool from_x64)
{
    DevGrid& d = s->g->dev;
    hipStream_t st = d.stream;
    if (!from_x64) launch_to_double(d, d.leafs, dev_channel(d, CH_X), s->x64, true, st);
    launch_residual64(d, s->x64, coef(d, s->c0), s->b64, s->r64, s->tmp, s->partial, st);
    launch_finish(s->partial, nullptr, d.leafs.n, s->S, FIN_SUM, st);
    double h[S_COUNT];
    read_scalars(s, h);
    return h[S_TMP];
}

double b_norm2(cmg_solver* s)
{
    DevGrid& d = s->g->dev;
    launch_dot64(d.leafs, s->b64, s->b64, s->partial, d.stream);
    launch_finish(s->partial, nullptr, d.leafs.n, s->S, FIN_SUM, d.stream);
    double h[S_COUNT];
    read_scalars(s, h);
    return h[S_TMP];
}

}  // namespace

cmg_status solver_create(cmg_grid* g, int c0, float Rm, float Rr, float P, cmg_solver** out)
{
    if (auto st = check_grid(g)) return st;
    if (!out || c0 < 0 || c0 + 3 >= kChannels) return CMG_ERR_BAD_ARG;
    auto* s = new cmg_solver;
    s->g = g;
    s->gen = g->generation;
    s->c0 = c0;
    s->Rm = Rm;
    s->Rr = Rr;
    s->P = P;
    try {
        s->S = dmalloc<double>(S_COUNT);
        HIPCHK(hipMemset(s->S, 0, S_COUNT * sizeof(double)));
        HIPCHK(hipHostMalloc((void**)&s->hS, S_COUNT * sizeof(double), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&s->hsig, sizeof(unsigned long long), hipHostMallocMapped | hipHostMallocCoherent));
        *s->hsig = 0;
        HIPCHK(hipHostGetDevicePointer((void**)&s->sig, s->hsig, 0));
        // room for a fused r.z: one partial per finest-level list entry plus a vector grid
        const size_t np = size_t(g->dev.max_list) + size_t(g->dev.wg_cap) + 64;
        s->partial = dmalloc<double>(np);
        s->partial2 = dmalloc<double>(np);
        HIPCHK(hipEventCreate(&s->e0));
        HIPCHK(hipEventCreate(&s->e1));
    } catch (const std::exception& ex) {
        set_error(ex.what());
        solver_destroy(s);
        return CMG_ERR_HIP;
    }
    *out = s;
    return CMG_OK;
}

cmg_status solver_create_geo(cmg_grid* g, int family, float one_over_alpha, float P, cmg_solver** out)
{
    if (family != CMG_SOLVER_GMG && family != CMG_SOLVER_CMG) return CMG_ERR_BAD_ARG;
    cmg_status st = solver_create(g, 11, 1.0f, one_over_alpha, P, out);
    if (st == CMG_OK) (*out)->family = family;
    return st;
}

// a solver's buffers are sized for the grid as it was when the solver was made; a grid
// finalized again since (new tiles, new device arrays) needs a new solver
static cmg_status check_solver(const cmg_solver* s)
{
    if (!s || !s->g) return CMG_ERR_BAD_ARG;
    if (s->gen != s->g->generation) {
        set_error("the grid was finalized again after this solver was created: create a new solver");
        return CMG_ERR_STATE;
    }
    return CMG_OK;
}

cmg_status solver_apply(cmg_solver* s, int x, int y)
{
    if (!s || !valid_channel(x) || !valid_channel(y) || x == y) return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    CMG_TRY({
        ensure_buffers(s);
        DevGrid& d = s->g->dev;
        enqueue_apply(s, dev_channel(d, x), dev_channel(d, y), nullptr, nullptr);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

// GMGSolver::dampedJacobiSolve (src/GMGSolver.cu:390-470)
cmg_status solver_damped_jacobi(cmg_solver* s, const cmg_solve_opts* o, cmg_solve_stats* st)
{
    if (!s || !o || !st || s->family != CMG_SOLVER_GMG || o->max_iters < 0 || o->cycle_coef_fp16) return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    *st = cmg_solve_stats{};
    bool converged = false;
    CMG_TRY({
        DevGrid& d = s->g->dev;
        hipStream_t stream = d.stream;
        ensure_buffers(s);
        float *X = dev_channel(d, CH_X), *B = dev_channel(d, CH_R), *AX = dev_channel(d, CH_P), *D = dev_channel(d, CH_AP);
        set_scalars(s, 0.0, o->max_iters);
        double h[S_COUNT];
        HIPCHK(hipEventRecord(s->e0, stream));
        launch_fill_all(d, X, 0.0f, nullptr, stream);
        int nb = launch_dot(d, d.all, B, B, s->partial, nullptr, stream);   // |b| over every tile
        launch_finish(s->partial, nullptr, nb, s->S, FIN_SUM, stream);
        read_scalars(s, h);
        const double rhs_norm = std::sqrt(h[S_TMP]);
        const double thr = std::max(o->relative_tolerance * rhs_norm, 2.2250738585072014e-308);
        double res = rhs_norm;
        int i = 0;
        for (i = 0; i < o->max_iters; ++i) {
            geo_full_apply(s, X, AX, false, nullptr);
            launch_residual_interior(d, B, AX, stream);
            nb = launch_dot(d, d.leafs, AX, AX, s->partial, nullptr, stream);
            launch_finish(s->partial, nullptr, nb, s->S, FIN_SUM, stream);
            read_scalars(s, h);
            res = std::sqrt(h[S_TMP]);
            if (!std::isfinite(res)) break;
            if (res < thr) {
                converged = true;
                break;
            }
            geo_full_apply(s, X, D, true, nullptr);
            launch_djacobi_update(d, X, AX, D, float(o->omega), stream);
        }
        HIPCHK(hipEventRecord(s->e1, stream));
        HIPCHK(hipEventSynchronize(s->e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
        st->ms = ms;
        st->iters = i;
        st->rel_res = rhs_norm > 0 ? res / rhs_norm : 0.0;
        st->true_rel_res = st->rel_res;
        if (!std::isfinite(res)) return CMG_ERR_NONFINITE;
    });
    return converged ? CMG_OK : CMG_ERR_NOT_CONVERGED;
}

void solver_destroy(cmg_solver* s)
{
    if (!s) return;
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    for (void* p : {(void*)s->S, (void*)s->partial, (void*)s->partial2, (void*)s->tmp, (void*)s->scratch, (void*)s->net_w,
                    (void*)s->lcount, (void*)s->q16, (void*)s->qbad, (void*)s->cw, (void*)s->clut, (void*)s->cws, (void*)s->czw,
                    (void*)s->x64, (void*)s->b64,
                    (void*)s->r64, (void*)s->inv0, (void*)s->cells0, (void*)s->zero0})
        if (p) (void)hipFree(p);
    if (s->hS) (void)hipHostFree(s->hS);
    if (s->hsig) (void)hipHostFree(s->hsig);
    if (s->e0) (void)hipEventDestroy(s->e0);
    if (s->e1) (void)hipEventDestroy(s->e1);
    delete s;
}

static cmg_status check_opts(const cmg_solve_opts* o)
{
    if (!o || o->max_iters < 0 || !(o->relative_tolerance >= 0) || o->level_iters < 0 || o->coarsest_iters < 0)
        return CMG_ERR_BAD_ARG;
    if (o->cycle < CMG_CYCLE_FAS_W || o->cycle > CMG_CYCLE_MU || o->mu < 0) return CMG_ERR_BAD_ARG;
    if (o->smoother != CMG_SMOOTH_RBGS && o->smoother != CMG_SMOOTH_JACOBI) return CMG_ERR_BAD_ARG;
    if (o->precision != CMG_PREC_FP32 && o->precision != CMG_PREC_FP64_REFINE) return CMG_ERR_BAD_ARG;
    if (o->learned < 0 || o->learned > 2) return CMG_ERR_BAD_ARG;
    if (o->learned && (o->cycle == CMG_CYCLE_VCYCLE || o->cycle == CMG_CYCLE_MU)) return CMG_ERR_BAD_ARG;
    if (o->cycle_coef_fp16 < 0 || o->cycle_coef_fp16 > 1) return CMG_ERR_BAD_ARG;
    return CMG_OK;
}

void prepare_coef_fp16(cmg_solver* s);

cmg_status solver_cycle(cmg_solver* s, const cmg_solve_opts* o, int x, int rhs, int x0)
{
    if (!s) return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    if (auto st = check_opts(o)) return st;
    if (o->cycle_coef_fp16 && s->family != CMG_SOLVER_AMG) return CMG_ERR_BAD_ARG;
    if (!valid_channel(x) || !valid_channel(rhs) || !valid_channel(x0)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        ensure_buffers(s);
        if (s->family == CMG_SOLVER_AMG && o->coarsest_direct) ensure_direct(s, o->is_pure_neumann != 0);
        if (s->family == CMG_SOLVER_AMG && o->cycle_coef_fp16) prepare_coef_fp16(s);
        if (s->family == CMG_SOLVER_AMG) prepare_coef_codes(s);
        enqueue_cycle(s, o, x, rhs, x0, nullptr);
        HIPCHK(hipStreamSynchronize(s->g->dev.stream));
    });
    return CMG_OK;
}

// FP16 copies of the coefficient channels divided by the tile spacing, for the
// cycle kernels; rebuilt whenever the coefficient channels or the types have been written
// since the last copy
void prepare_coef_fp16(cmg_solver* s)
{
    DevGrid& d = s->g->dev;
    const size_t n = size_t(d.N) * kCells;
    const cmg_solver::Stamp stamp = stamp_now(s);
    if (s->q16 && s->q16_stamp == stamp) return;
    s->q16_stamp = cmg_solver::Stamp{};
    if (!s->q16) {
        s->q16 = dmalloc<uint16_t>(4 * n);
        s->qbad = dmalloc<unsigned int>(1);
    }
    uint16_t* const q[4] = {s->q16, s->q16 + n, s->q16 + 2 * n, s->q16 + 3 * n};
    HIPCHK(hipMemsetAsync(s->qbad, 0, sizeof(unsigned int), d.stream));
    launch_coef_to_half(d, coef(d, s->c0), q, s->qbad, d.stream);
    unsigned int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, s->qbad, sizeof(bad), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    if (bad)
        throw std::invalid_argument("cycle_coef_fp16: a coefficient divided by the tile spacing is outside the FP16 range "
                                    "(or a nonzero diagonal below its smallest normal)");
    s->q16_stamp = stamp;
}

// the stencils may derive LEAF diagonals from the face coefficients (one load
// fewer per cell) when every stored LEAF diagonal equals the derived one bitwise
bool diagonal_derivable(cmg_solver* s)
{
    DevGrid& d = s->g->dev;
    const cmg_solver::Stamp stamp = stamp_now(s);
    if (s->derive_stamp == stamp) return s->derive_ok;
    set_scalars(s, 0.0, 0);
    const int n = launch_diag_check(d, coef(d, s->c0), s->partial, d.stream);
    launch_finish(s->partial, nullptr, n, s->S, FIN_SUM, d.stream);
    double h[S_COUNT];
    read_scalars(s, h);
    s->derive_ok = h[S_TMP] == 0.0;
    s->derive_stamp = stamp;
    return s->derive_ok;
}

struct DeriveScope {
    DevGrid& d;
    DeriveScope(DevGrid& dd, bool on) : d(dd) { d.derive = on ? 1 : 0; }
    ~DeriveScope() { d.derive = 0; }
};

cmg_status solver_solve(cmg_solver* s, const cmg_solve_opts* o, cmg_solve_stats* st)
{
    if (!s || !st) return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    if (auto e = check_opts(o)) return e;
    if (s->family != CMG_SOLVER_AMG && o->precision != CMG_PREC_FP32) return CMG_ERR_BAD_ARG;
    if (s->family != CMG_SOLVER_AMG && o->cycle_coef_fp16) return CMG_ERR_BAD_ARG;
    if (o->learned && (!s->net_w || s->family != CMG_SOLVER_AMG)) {
        set_error("learned smoothing needs the net's weights (cmg_solver_set_net) on an AMG solver");
        return CMG_ERR_STATE;
    }
    *st = cmg_solve_stats{};
    int done = 1;
    CMG_TRY({
        DevGrid& d = s->g->dev;
        hipStream_t stream = d.stream;
        ensure_buffers(s);
        const bool amg = s->family == CMG_SOLVER_AMG;
        if (amg) ensure_fp64(s);
        if (o->learned) HIPCHK(hipMemset(s->lcount, 0, 2 * sizeof(unsigned int)));   // counts of this solve
        if (amg && o->coarsest_direct) ensure_direct(s, o->is_pure_neumann != 0);
        if (amg && o->cycle_coef_fp16) prepare_coef_fp16(s);
        if (amg) prepare_coef_codes(s);
        DeriveScope derive(d, amg && derive_allowed() && diagonal_derivable(s));
        // keep b in FP64 for the true residual / refinement (not needed when an FP32 solve skips it)
        if (amg && !(o->precision == CMG_PREC_FP32 && o->no_true_residual))
            launch_to_double(d, d.leafs, dev_channel(d, CH_R), s->b64, true, stream);
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipEventRecord(s->e0, stream));
        if (o->precision == CMG_PREC_FP32) {
            int iters = 0;
            double rel = 0;
            done = pcg32(s, o, o->relative_tolerance, o->max_iters, o->warm_start != 0, o->sync_stride, &iters, &rel);
            st->iters = iters;
            st->rel_res = rel;
            HIPCHK(hipEventRecord(s->e1, stream));
            HIPCHK(hipEventSynchronize(s->e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
            st->ms = ms;
            // true residual from the FP32 solution (outside the timed region)
            if (amg && o->no_true_residual) {
                st->true_rel_res = -1;
            } else if (amg) {
                const double rn2 = residual_norm2(s, false);
                const double bn2 = b_norm2(s);
                st->true_rel_res = bn2 > 0 ? std::sqrt(rn2 / bn2) : 0.0;
            } else {
                st->true_rel_res = st->rel_res;
            }
        } else {
            // FP64 iterative refinement around FP32 PCG
            const size_t n = size_t(d.N) * kCells;
            if (o->warm_start) launch_to_double(d, d.leafs, dev_channel(d, CH_X), s->x64, true, stream);
            else HIPCHK(hipMemsetAsync(s->x64, 0, n * sizeof(double), stream));
            const double bn2 = b_norm2(s);
            double h[S_COUNT];
            double rn2 = residual_norm2(s, true);
            // the inner solves read their right-hand side from channel 1
            launch_to_float(d.leafs, s->r64, dev_channel(d, CH_R), stream);
            int total = 0, outer = 0;
            done = 1;
            const double otol = o->outer_tolerance > 0 ? o->outer_tolerance : o->relative_tolerance;
            const double itol = o->inner_tolerance > 0 ? o->inner_tolerance : 1e-4;
            while (bn2 > 0 && std::sqrt(rn2 / bn2) >= otol && outer < 50 && total < o->max_iters) {
                int iters = 0;
                double rel = 0;
                int dd = pcg32(s, o, itol, o->max_iters - total, false, std::max(1, o->sync_stride), &iters, &rel);
                total += iters;
                ++outer;
                if (dd == 2 || dd == 3) { done = dd; break; }
                launch_update_x64(d, s->x64, dev_channel(d, CH_X), stream);
                launch_residual64(d, s->x64, coef(d, s->c0), s->b64, s->r64, dev_channel(d, CH_R), s->partial, stream);
                launch_finish(s->partial, nullptr, d.leafs.n, s->S, FIN_SUM, stream);
                read_scalars(s, h);
                rn2 = h[S_TMP];
                (void)n;
                if (!std::isfinite(rn2)) { done = 3; break; }
            }
            if (done == 1 && bn2 > 0 && std::sqrt(rn2 / bn2) >= otol) done = 4;
            launch_to_float(d.leafs, s->x64, dev_channel(d, CH_X), stream);
            HIPCHK(hipEventRecord(s->e1, stream));
            HIPCHK(hipEventSynchronize(s->e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
            st->ms = ms;
            st->iters = total;
            st->outer_iters = outer;
            st->rel_res = bn2 > 0 ? std::sqrt(rn2 / bn2) : 0.0;
            st->true_rel_res = st->rel_res;
        }
    });
    return status_of(done);
}

// AMGSolver::FASMuCycleSolve (src/AMGSolver.cu:1311-1356)
cmg_status solver_mg_solve(cmg_solver* s, const cmg_solve_opts* o, cmg_solve_stats* st)
{
    if (!s || !st) return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    if (auto e = check_opts(o)) return e;
    if (o->cycle != CMG_CYCLE_FAS_W && o->cycle != CMG_CYCLE_FAS_V) return CMG_ERR_BAD_ARG;
    if (o->cycle_coef_fp16) return CMG_ERR_BAD_ARG;   // a stand-alone cycle solve would converge to the FP16 operator
    if (s->family != CMG_SOLVER_AMG) return CMG_ERR_BAD_ARG;
    *st = cmg_solve_stats{};
    bool converged = false;
    CMG_TRY({
        DevGrid& d = s->g->dev;
        hipStream_t stream = d.stream;
        ensure_buffers(s);
        if (o->coarsest_direct) ensure_direct(s, o->is_pure_neumann != 0);
        prepare_coef_codes(s);
        float *X = dev_channel(d, CH_X), *B = dev_channel(d, CH_R), *X0 = dev_channel(d, CH_P);
        const Coef C = solver_coef(s);
        set_scalars(s, 0.0, o->max_iters);
        double h[S_COUNT];
        HIPCHK(hipEventRecord(s->e0, stream));
        launch_fill_all(d, X, 0.0f, nullptr, stream);
        int nb = launch_dot(d, d.leafs, B, B, s->partial, nullptr, stream);
        launch_finish(s->partial, nullptr, nb, s->S, FIN_SUM, stream);
        read_scalars(s, h);
        const double rhs_norm = std::sqrt(h[S_TMP]);
        const double thr = std::max(o->relative_tolerance * rhs_norm, 2.2250738585072014e-308);
        double res = rhs_norm;
        const CycleCtx c{s, o, X, B, X0, C, nullptr};
        int i = 0;
        for (i = 0; i < o->max_iters; ++i) {
            cycle_step(c, d.max_level, i == 0);
            // AMGFullNegativeLaplacianOnLeafs(x -> x0) leaves the propagated ghost
            // and accumulated NONLEAF values in x, as the next cycle sees them
            launch_propagate_ghosts(d, X, stream);
            launch_accumulate_leafs(d, X, stream);
            launch_apply(d, X, C, X0, nullptr, nullptr, stream);
            launch_sub_from(d, d.leafs, B, X0, stream);
            nb = launch_dot(d, d.leafs, X0, X0, s->partial, nullptr, stream);
            launch_finish(s->partial, nullptr, nb, s->S, FIN_SUM, stream);
            read_scalars(s, h);
            res = std::sqrt(h[S_TMP]);
            if (!std::isfinite(res)) break;
            if (res < thr) {
                converged = true;
                break;
            }
        }
        HIPCHK(hipEventRecord(s->e1, stream));
        HIPCHK(hipEventSynchronize(s->e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
        st->ms = ms;
        st->iters = i;
        st->rel_res = rhs_norm > 0 ? res / rhs_norm : 0.0;
        st->true_rel_res = st->rel_res;
        if (!std::isfinite(res)) return CMG_ERR_NONFINITE;
    });
    return converged ? CMG_OK : CMG_ERR_NOT_CONVERGED;
}

// ---- learned tile smoother -----------------------------------------------------

cmg_status solver_set_net(cmg_solver* s, const float* w)
{
    if (!s || !w) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        if (!s->net_w) s->net_w = dmalloc<float>(kNetWeights);
        if (!s->lcount) {
            s->lcount = dmalloc<unsigned int>(2);
            HIPCHK(hipMemset(s->lcount, 0, 2 * sizeof(unsigned int)));
        }
        HIPCHK(hipMemcpy(s->net_w, w, kNetWeights * sizeof(float), hipMemcpyHostToDevice));
        if (s->gexec) (void)hipGraphExecDestroy(s->gexec), s->gexec = nullptr;   // the graph holds the old pointer
    });
    return CMG_OK;
}

// y = net(r / D) on the finest-level LEAF tiles (BF16 MFMA), channels r, D, y
cmg_status solver_net_infer(cmg_solver* s, int r, int dch, int y)
{
    if (!s || !s->net_w || !valid_channel(r) || !valid_channel(dch) || !valid_channel(y) || y == r || y == dch)
        return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    CMG_TRY({
        DevGrid& d = s->g->dev;
        launch_net_infer(d, d.subtree[d.max_level], dev_channel(d, r), dev_channel(d, dch), s->net_w, dev_channel(d, y),
                         nullptr, d.stream);
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

cmg_status solver_learned_counts(cmg_solver* s, uint64_t* judged, uint64_t* kept)
{
    if (!s || !judged || !kept) return CMG_ERR_BAD_ARG;
    if (!s->lcount) return CMG_ERR_STATE;
    CMG_TRY({
        unsigned int h[2] = {0, 0};
        HIPCHK(hipStreamSynchronize(s->g->dev.stream));
        HIPCHK(hipMemcpy(h, s->lcount, sizeof(h), hipMemcpyDeviceToHost));
        *judged = h[0];
        *kept = h[1];
    });
    return CMG_OK;
}

cmg_status mfma_probe(int K, const uint16_t* A, const uint16_t* B, float* C)
{
    if (!A || !B || !C || (K != 16 && K != 32)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        uint16_t *dA = dmalloc<uint16_t>(16 * K), *dB = dmalloc<uint16_t>(16 * K);
        float* dC = dmalloc<float>(256);
        HIPCHK(hipMemcpy(dA, A, 32 * K, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dB, B, 32 * K, hipMemcpyHostToDevice));
        launch_mfma_probe(K, dA, dB, dC, nullptr);
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(C, dC, 1024, hipMemcpyDeviceToHost));
        (void)hipFree(dA);
        (void)hipFree(dB);
        (void)hipFree(dC);
    });
    return CMG_OK;
}

}  // namespace cmg

// training state: tiles (input volume, target volume) resident in HBM, the FP32
// master weights, Adam moments, a seeded per-epoch permutation of the tiles
struct cmg_trainer {
    int capacity = 0, size = 0;
    float *X = nullptr, *E = nullptr;
    float *W = nullptr, *M = nullptr, *V = nullptr, *G = nullptr;
    float* gpart = nullptr;
    double *lpart = nullptr, *loss = nullptr;
    int32_t* idx = nullptr;
    int max_wg = 0;
    int step = 0;
    uint64_t rng = 0;
    std::vector<int32_t> perm;
    size_t perm_pos = 0;
    hipStream_t stream = nullptr;
};

namespace cmg {

namespace {

uint64_t splitmix(uint64_t& x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

cmg_status trainer_create(int capacity, uint64_t seed, cmg_trainer** out)
{
    if (capacity <= 0 || !out) return CMG_ERR_BAD_ARG;
    auto* t = new cmg_trainer;
    try {
        t->capacity = capacity;
        t->X = dmalloc<float>(size_t(capacity) * kNetVolume);
        t->E = dmalloc<float>(size_t(capacity) * kNetVolume);
        t->W = dmalloc<float>(kNetWeights);
        t->M = dmalloc<float>(kNetWeights);
        t->V = dmalloc<float>(kNetWeights);
        t->G = dmalloc<float>(kNetWeights);
        int dev = 0, cus = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        t->max_wg = std::max(1, cus);
        t->gpart = dmalloc<float>(size_t(t->max_wg) * kNetWeights);
        t->lpart = dmalloc<double>(size_t(t->max_wg));
        t->loss = dmalloc<double>(1);
        t->idx = dmalloc<int32_t>(size_t(capacity));
        HIPCHK(hipMemset(t->M, 0, kNetWeights * sizeof(float)));
        HIPCHK(hipMemset(t->V, 0, kNetWeights * sizeof(float)));
        HIPCHK(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
        // He-style initial weights from the seed: N(0, 2 / fan_in) by Box-Muller
        std::vector<float> w(kNetWeights);
        t->rng = seed;
        for (int i = 0; i < kNetWeights; ++i) {
            const int fan_in = i < 432 ? 27 : 432;
            const double u1 = (double(splitmix(t->rng) >> 11) + 1.0) * 0x1.0p-53, u2 = double(splitmix(t->rng) >> 11) * 0x1.0p-53;
            const double nrm = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            w[i] = float(nrm * std::sqrt(2.0 / fan_in) * (i >= 7344 ? 0.1 : 1.0));
        }
        HIPCHK(hipMemcpy(t->W, w.data(), kNetWeights * sizeof(float), hipMemcpyHostToDevice));
    } catch (const std::exception& ex) {
        set_error(ex.what());
        trainer_destroy(t);
        return CMG_ERR_HIP;
    }
    *out = t;
    return CMG_OK;
}

int trainer_size(const cmg_trainer* t) { return t ? t->size : 0; }

void trainer_destroy(cmg_trainer* t)
{
    if (!t) return;
    for (void* p : {(void*)t->X, (void*)t->E, (void*)t->W, (void*)t->M, (void*)t->V, (void*)t->G, (void*)t->gpart,
                    (void*)t->lpart, (void*)t->loss, (void*)t->idx})
        if (p) (void)hipFree(p);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

// the finest-level LEAF tiles of a grid: input r / D (channels r_ch, d_ch), target e_ch
cmg_status trainer_collect(cmg_trainer* t, cmg_grid* g, int r_ch, int d_ch, int e_ch, int* appended)
{
    if (!t) return CMG_ERR_BAD_ARG;
    if (auto st = check_grid(g)) return st;
    if (!valid_channel(r_ch) || !valid_channel(d_ch) || !valid_channel(e_ch)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        DevGrid& d = g->dev;
        const Topology& topo = *g->topo;
        std::vector<int32_t> tiles;
        for (int i = topo.level_start[d.max_level]; i < topo.level_start[d.max_level + 1]; ++i)
            if (topo.t_type[i] & LEAF) tiles.push_back(i);
        const int n = std::min(int(tiles.size()), t->capacity - t->size);
        if (n > 0) {
            int32_t* dt = dmalloc<int32_t>(size_t(n));
            HIPCHK(hipMemcpy(dt, tiles.data(), size_t(n) * sizeof(int32_t), hipMemcpyHostToDevice));
            launch_gather_tiles(d, dt, n, dev_channel(d, r_ch), dev_channel(d, d_ch), 0, t->X + size_t(t->size) * kNetVolume,
                                d.stream);
            launch_gather_tiles(d, dt, n, dev_channel(d, e_ch), nullptr, 1, t->E + size_t(t->size) * kNetVolume, d.stream);
            HIPCHK(hipStreamSynchronize(d.stream));
            (void)hipFree(dt);
            t->size += n;
        }
        if (appended) *appended = std::max(n, 0);
    });
    return CMG_OK;
}

cmg_status trainer_add_tiles(cmg_trainer* t, const float* x, const float* e, int n)
{
    if (!t || !x || !e || n < 0 || t->size + n > t->capacity) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        HIPCHK(hipMemcpy(t->X + size_t(t->size) * kNetVolume, x, size_t(n) * kNetVolume * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(t->E + size_t(t->size) * kNetVolume, e, size_t(n) * kNetVolume * sizeof(float), hipMemcpyHostToDevice));
        t->size += n;
    });
    return CMG_OK;
}

cmg_status trainer_weights(cmg_trainer* t, float* get, const float* set)
{
    if (!t || (!get && !set)) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        if (set) HIPCHK(hipMemcpy(t->W, set, kNetWeights * sizeof(float), hipMemcpyHostToDevice));
        if (get) HIPCHK(hipMemcpy(get, t->W, kNetWeights * sizeof(float), hipMemcpyDeviceToHost));
    });
    return CMG_OK;
}

// mean loss and gradient over tiles [first, first + n)
cmg_status trainer_loss_grad(cmg_trainer* t, int first, int n, int bf16, double* loss, float* grad)
{
    if (!t || first < 0 || n <= 0 || first + n > t->size || !loss) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        const int used = launch_net_train(t->X + size_t(first) * kNetVolume, t->E + size_t(first) * kNetVolume, nullptr, n,
                                          t->W, t->gpart, t->lpart, t->max_wg, bf16 != 0, t->stream);
        launch_net_reduce(t->gpart, t->lpart, used, n, t->G, t->loss, t->stream);
        HIPCHK(hipMemcpyAsync(loss, t->loss, sizeof(double), hipMemcpyDeviceToHost, t->stream));
        if (grad) HIPCHK(hipMemcpyAsync(grad, t->G, kNetWeights * sizeof(float), hipMemcpyDeviceToHost, t->stream));
        HIPCHK(hipStreamSynchronize(t->stream));
    });
    return CMG_OK;
}

// one Adam step on the next `batch` tiles of a seeded per-epoch permutation;
// *loss = the batch loss before the step
cmg_status trainer_step(cmg_trainer* t, int batch, float lr, int bf16, double* loss)
{
    if (!t || batch <= 0 || t->size <= 0) return CMG_ERR_BAD_ARG;
    CMG_TRY({
        batch = std::min(batch, t->size);
        if (t->perm.size() != size_t(t->size) || t->perm_pos + batch > t->perm.size()) {
            t->perm.resize(size_t(t->size));
            for (int i = 0; i < t->size; ++i) t->perm[i] = i;
            for (int i = t->size - 1; i > 0; --i) std::swap(t->perm[i], t->perm[splitmix(t->rng) % uint64_t(i + 1)]);
            t->perm_pos = 0;
        }
        HIPCHK(hipMemcpyAsync(t->idx, t->perm.data() + t->perm_pos, size_t(batch) * sizeof(int32_t), hipMemcpyHostToDevice,
                              t->stream));
        t->perm_pos += batch;
        const int used = launch_net_train(t->X, t->E, t->idx, batch, t->W, t->gpart, t->lpart, t->max_wg, bf16 != 0, t->stream);
        launch_net_reduce(t->gpart, t->lpart, used, batch, t->G, t->loss, t->stream);
        launch_adam(t->W, t->M, t->V, t->G, ++t->step, lr, 0.9f, 0.999f, 1e-8f, t->stream);
        if (loss) HIPCHK(hipMemcpyAsync(loss, t->loss, sizeof(double), hipMemcpyDeviceToHost, t->stream));
        HIPCHK(hipStreamSynchronize(t->stream));
    });
    return CMG_OK;
}

cmg_status op_copy(cmg_grid* g, int src, int dst)
{
    if (auto st = check_grid(g)) return st;
    if (!valid_channel(src) || !valid_channel(dst)) return CMG_ERR_BAD_ARG;
    if (src == dst) return CMG_OK;
    CMG_TRY({
        auto& d = g->dev;
        HIPCHK(hipMemcpyAsync(dev_channel(d, dst), dev_channel(d, src), size_t(d.N) * kCells * sizeof(float),
                              hipMemcpyDeviceToDevice, d.stream));
        HIPCHK(hipStreamSynchronize(d.stream));
    });
    return CMG_OK;
}

cmg_status bench_kernel(cmg_solver* s, int kernel, int reps, double* avg_ms, int* tiles)
{
    if (!s || !avg_ms || reps <= 0 || kernel < 0 || kernel > 14 || s->family != CMG_SOLVER_AMG) return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    if (kernel == 9 && !s->net_w) return CMG_ERR_STATE;
    CMG_TRY({
        ensure_buffers(s);
        DevGrid& d = s->g->dev;
        hipStream_t st = d.stream;
        const int L = d.max_level;
        float *X = dev_channel(d, CH_X), *R = dev_channel(d, CH_R), *Pp = dev_channel(d, CH_P), *AP = dev_channel(d, CH_AP),
              *Z = dev_channel(d, CH_Z);
        Coef C = coef(d, s->c0);
        Coef Cq = C;   // kernel 11: the cycle's FP16 coefficient copies
        if (kernel == 11) {
            prepare_coef_fp16(s);
            const size_t n = size_t(d.N) * kCells;
            for (int a = 0; a < 4; ++a) Cq.q[a] = s->q16 + size_t(a) * n;
        }
        // kernels 12-14: the coded coefficients (the finest level must be coded), with LEAF
        // diagonals derived from the faces when a solve would derive them
        Coef Ck = C;
        if (kernel >= 12) {
            if (!coef_codes_allowed()) throw std::invalid_argument("bench_kernel: the coded coefficients are switched off");
            prepare_coef_codes(s);
            Ck = solver_coef(s);
            if (!((Ck.coded >> L) & 1u)) throw std::invalid_argument("bench_kernel: the finest level is not coded");
        }
        DeriveScope derive(d, kernel >= 12 && derive_allowed() && diagonal_derivable(s));
        // live scalars for the kernels that read alpha / beta: not stopped, alpha = beta = 1/2
        double h[S_COUNT] = {};
        h[S_ALPHA] = 0.5;
        h[S_BETA] = 0.5;
        HIPCHK(hipMemcpy(s->S, h, sizeof(h), hipMemcpyHostToDevice));
        const double* S = s->S;
        int nt = 0;
        // kernel 10: flat copy of 1 GiB into another 1 GiB (4x the 256 MB last-level cache)
        const size_t nflat = kernel == 10 ? (size_t(1) << 28) : 0;
        float* fa = nflat ? dmalloc<float>(nflat) : nullptr;
        float* fb = nflat ? dmalloc<float>(nflat) : nullptr;
        if (nflat) HIPCHK(hipMemset(fa, 0, nflat * sizeof(float)));
        auto one = [&]() {
            switch (kernel) {
            case 0: launch_apply(d, Pp, C, AP, s->partial, S, st); nt = d.leafs.tiles; break;
            case 1: launch_jacobi(d, L, Z, s->tmp, Z, C, R, 2.0f / 3.0f, S, st); nt = d.subtree[L].tiles; break;
            case 2: launch_residual_restrict(d, L, Z, C, AP, 1.0f, S, st); nt = d.subtree[L].tiles; break;
            case 3: launch_update_xr(d, X, Pp, Z, AP, s->partial, s->S, st); nt = d.leafs.tiles; break;
            case 4: launch_fas_prolong(d, L, s->tmp, AP, 1.0f, S, st); nt = d.prolong[L].tiles; break;
            case 5: launch_update_p(d, Pp, Z, s->S, st); nt = d.leafs.tiles; break;
            case 6: launch_dot(d, d.leafs, R, Z, s->partial, S, st); nt = d.leafs.tiles; break;
            case 7: launch_jacobi_zero(d, L, s->tmp, C.c[3], R, 2.0f / 3.0f, S, st); nt = d.subtree[L].tiles; break;
            case 8: launch_copy_list(d, d.leafs, Z, s->tmp, S, st); nt = d.leafs.tiles; break;
            case 9: launch_net_infer(d, d.subtree[L], R, C.c[3], s->net_w, s->tmp, S, st); nt = d.subtree[L].tiles; break;
            case 10: launch_copy_flat(d, fa, fb, nflat, st); nt = int(nflat / 512); break;
            case 11: launch_jacobi(d, L, Z, s->tmp, Z, Cq, R, 2.0f / 3.0f, S, st); nt = d.subtree[L].tiles; break;
            case 12: launch_jacobi(d, L, Z, s->tmp, Z, Ck, R, 2.0f / 3.0f, S, st); nt = d.subtree[L].tiles; break;
            case 13: launch_apply(d, Pp, Ck, AP, s->partial, S, st); nt = d.leafs.tiles; break;
            case 14: launch_residual_restrict(d, L, Z, Ck, AP, 1.0f, S, st); nt = d.subtree[L].tiles; break;
            }
        };
        one();
        HIPCHK(hipEventRecord(s->e0, st));
        for (int i = 0; i < reps; ++i) one();
        HIPCHK(hipEventRecord(s->e1, st));
        HIPCHK(hipEventSynchronize(s->e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, s->e0, s->e1));
        *avg_ms = double(ms) / reps;
        if (tiles) *tiles = nt;
        if (fa) (void)hipFree(fa);
        if (fb) (void)hipFree(fb);
    });
    return CMG_OK;
}

// levels whose face coefficients the stencil kernels read as codes now (bit l: level l)
cmg_status solver_coded_levels(cmg_solver* s, uint32_t* mask, uint32_t* wide)
{
    if (!s || !mask) return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    const Coef c = solver_coef(s);
    *mask = c.coded;
    if (wide) *wide = c.wide;
    return CMG_OK;
}

cmg_status solver_get_x64(cmg_solver* s, double* host)
{
    if (!s || !host) return CMG_ERR_BAD_ARG;
    if (auto e = check_solver(s)) return e;
    if (!s->x64) return CMG_ERR_STATE;
    CMG_TRY({
        HIPCHK(hipMemcpy(host, s->x64, size_t(s->g->dev.N) * kCells * sizeof(double), hipMemcpyDeviceToHost));
    });
    return CMG_OK;
}

}  // namespace cmg
