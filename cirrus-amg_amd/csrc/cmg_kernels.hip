// cmg_kernels.hip -- gfx950 kernels of the octree multigrid Poisson solve.
//
// One 8^3 tile per 512-thread workgroup (8 wave64s, one x-plane per wave),
// thread v <-> cell v = (i<<6)|(j<<3)|k.  The 10^3 halo'd tile and the three
// negative-face coefficient arrays (+ the + side face from the neighbour)
// are staged in LDS.  Workgroups walk Morton-ordered launch lists so that a
// tile's face neighbours are processed by nearby workgroups on the same XCD
// and their halo reads hit that XCD's L2.
//
// Arithmetic is written out exactly (fp-contract off, explicit fmaf) so that
// each kernel reproduces the CPU restatement bit for bit:
//   apply   s = x0*D ; s = fmaf(off, xn, s) for axis 0..2, side -,+
//   ghost   vg = (float)((double)x_c + 0.5*(double)(V1 - avg8))
//   RB-SOR  x = (float)((double)x + (double)((b - Ax)/D) * omega)
//   Jacobi  x = x + w*((b - Ax)/D)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <type_traits>

#include "cmg_device.hpp"
#include "cmg_kernels.hpp"

#pragma clang fp contract(off)

namespace cmg {
namespace k {


// Coefficient forms of the one-tile stencil kernels (template argument Q):
//   0: the FP32 channels;
//   1: the FP16 copies (c / h); the coefficient planes are then staged as those FP16 bits
//      (7.5 KB a tile instead of 10.9, so more tiles fit on a CU);
//   2: the exact per-level codes (Coef::w, Coef::lut): a cell's type and three face codes
//      in one 32-bit word; the planes are staged as the codes (1.7 KB a tile) and read
//      through the level's table, which is staged in LDS (entry kCodes - 1 holds the
//      tile spacing, the plane value of a face without neighbour tile); every value
//      computed is that of form 0;
//   3: the same codes, the planes staged as their FP32 values (the LDS layout of form 0
//      plus the table, so that one kernel can take forms 0 and 3 tile by tile);
//   4: ten-bit codes (levels with 256 to 1023 distinct values), staged as 16-bit codes
//      and read through the level's table of kWideCodes entries.
template <typename T, int Q = 0>
struct Lds {
    T x[1000];
    // [a][0..511] own negative faces, [a][512+f] neighbour face on the + side
    std::conditional_t<Q == 1 || Q == 4, uint16_t, std::conditional_t<Q == 2, uint8_t, float>> off[3][576];
    uint8_t empty[6];    // face f = 2 axis + (side > 0) has no neighbour tile
    float lut[Q == 4 ? kWideCodes : Q >= 2 ? kCodes : 1];   // forms 2-4: the tile level's code table
};

// plane a's staged coefficient at index i as FP32 (h: the tile spacing)
template <typename T, int LQ>
__device__ __forceinline__ float staged_coef(const Lds<T, LQ>& s, int a, int i, float h)
{
    if constexpr (LQ == 1) return float(__builtin_bit_cast(_Float16, s.off[a][i])) * h;
    else if constexpr (LQ == 2 || LQ == 4) return s.lut[s.off[a][i]];
    else return s.off[a][i];
}

// the entries of a coded level's table (the last one holds the tile spacing)
template <int Q>
constexpr int code_table() { return Q == 4 ? kWideCodes : kCodes; }

// a cell word's code of the face coefficient on axis a: bits 8a+8.. (forms 2, 3) or 10a+2.. (form 4)
template <int Q>
__device__ __forceinline__ uint32_t word_code(uint32_t w, int a)
{
    if constexpr (Q == 4) return (w >> (10 * a + 2)) & 0x3ffu;
    else return (w >> (8 * a + 8)) & 0xffu;
}

// a cell word's cell type: byte 0 (forms 2, 3), or 0..3 in bits 0-1 for 0, DIRICHLET, INTERIOR, NEUMANN (form 4)
template <int Q>
__device__ __forceinline__ uint8_t word_type(uint32_t w)
{
    if constexpr (Q == 4) return uint8_t((0x4210u >> (4 * (w & 3u))) & 0xfu);
    else return uint8_t(w);
}

// a staged coefficient as FP32: FP32 as is, FP16 bits of c / h times h (exact: h is a
// power of two times h0)
__device__ __forceinline__ float lds_coef(float v, float) { return v; }
__device__ __forceinline__ float lds_coef(uint16_t v, float h) { return float(__builtin_bit_cast(_Float16, v)) * h; }

// diagonal of a LEAF INTERIOR cell from its face coefficients, in the order the
// assembly sums them (CoarsenTypesAndAMGCoeffs, src/AMGSolver.cu:129-160):
// d = 0 - c-(x) - c+(x) - c-(y) - c+(y) - c-(z) - c+(z), c+ = -h without neighbour
template <typename T, int LQ>
__device__ __forceinline__ float derived_diag(const Lds<T, LQ>& s, int v, const float c[3], float h)
{
    static_assert(LQ != 1, "the FP16 form reads every diagonal");
    const int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
    // one plane read per face (a face without neighbour tile reads its staged h, unused)
    const float t0 = staged_coef(s, 0, i < 7 ? v + 64 : 512 + (j * 8 + k), h);
    const float t1 = staged_coef(s, 1, j < 7 ? v + 8 : 512 + (i * 8 + k), h);
    const float t2 = staged_coef(s, 2, k < 7 ? v + 1 : 512 + (i * 8 + j), h);
    const float cp0 = (i == 7 && s.empty[1]) ? -h : t0;
    const float cp1 = (j == 7 && s.empty[3]) ? -h : t1;
    const float cp2 = (k == 7 && s.empty[5]) ? -h : t2;
    float d = 0.0f;
    d = d - c[0];
    d = d - cp0;
    d = d - c[1];
    d = d - cp1;
    d = d - c[2];
    d = d - cp2;
    return d;
}



// Per-thread state of the cells a thread owns: with NT threads per tile each
// thread owns CPT = 512/NT cells, v = own_cell<NT>(c) = CPT*tid + c (consecutive
// along k) for the one-tile-per-workgroup kernels (load_tile), v = tid + c*NT
// (CPT x-planes apart) for the persistent pipelined forms (issue / stage_lds).
template <typename T, int NT>
struct Own {
    static constexpr int CPT = 512 / NT;
    T x[CPT];          // x masked to INTERIOR cells
    float c[CPT][3];   // negative-face coefficients
    float d[CPT];      // diagonal
    uint8_t ct[CPT];   // cell type
};

// a pointer argument as a workgroup-uniform value (kept in SGPRs, selectable
// without a memory access)
template <typename P>
__device__ __forceinline__ const P* uniform_ptr(const P* p)
{
    const uint64_t u = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(u));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(u >> 32));
    return reinterpret_cast<const P*>((uint64_t(hi) << 32) | lo);
}

// the cells of a thread in the one-tile kernels: CPT = 512/NT consecutive cells, so each
// per-cell array of a tile is read (and written) with one vector access per thread
template <int NT>
__device__ __forceinline__ int own_cell(int c, int tid)
{
    return (512 / NT) * tid + c;
}
template <int NT>
__device__ __forceinline__ int own_cell(int c)
{
    return own_cell<NT>(c, int(threadIdx.x));
}

template <int N, typename E>
struct alignas(N * sizeof(E)) Pack {
    E v[N];
};

// N consecutive elements at p (aligned to N elements) in one access
template <int N, typename E>
__device__ __forceinline__ void vload(const E* __restrict__ p, E (&out)[N])
{
    const Pack<N, E> pk = *reinterpret_cast<const Pack<N, E>*>(p);
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = pk.v[i];
}

template <int N, typename E>
__device__ __forceinline__ void vstore(E* __restrict__ p, const E (&in)[N])
{
    Pack<N, E> pk;
#pragma unroll
    for (int i = 0; i < N; ++i) pk.v[i] = in[i];
    *reinterpret_cast<Pack<N, E>*>(p) = pk;
}

// four consecutive FP16 coefficients (quad q of the tile-major layout), back to FP32
__device__ __forceinline__ float4 coef_q4(const uint16_t* q, size_t quad, float h)
{
    const ushort4 u = reinterpret_cast<const ushort4*>(q)[quad];
    return make_float4(float(__builtin_bit_cast(_Float16, u.x)) * h, float(__builtin_bit_cast(_Float16, u.y)) * h,
                       float(__builtin_bit_cast(_Float16, u.z)) * h, float(__builtin_bit_cast(_Float16, u.w)) * h);
}

// a coefficient stored as FP16 bits of c / h, back to FP32 (h is a power of two
// times h0, so the product restores the scale exactly)
__device__ __forceinline__ float coef_q(const uint16_t* q, size_t i, float h)
{
    return float(__builtin_bit_cast(_Float16, q[i])) * h;
}

// element s (0..5, thread-varying) of (a.x, a.y, a.z, a.w, b.x, b.y) by selects,
// with no private array (which the compiler would place in LDS)
__device__ __forceinline__ int pick6(int4 a, int4 b, int s)
{
    int r = a.x;
    r = s == 1 ? a.y : r;
    r = s == 2 ? a.z : r;
    r = s == 3 ? a.w : r;
    r = s == 4 ? b.x : r;
    r = s == 5 ? b.y : r;
    return r;
}

// element s (0..7, thread-varying) of (a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w) by selects
__device__ __forceinline__ int pick8(int4 a, int4 b, int s)
{
    int r = pick6(a, b, s);
    r = s == 6 ? b.z : r;
    r = s == 7 ? b.w : r;
    return r;
}

// A one-tile stencil workgroup's record (launch-list order, 7 x 16 bytes, read with
// scalar loads right after the launch): the tile, its level, type | octant << 8 and
// parent; its halo record (face neighbours, their tile types packed in bytes, own
// type in byte 7.3); the parents of the face neighbours and their octants (the
// operator's fresh GHOST values, the cycle's pulled prolongation); its children by
// octant (for a sweep that prolongs into them).  One round trip gives the workgroup every address it needs:
// the tile's own loads and its halo loads follow at once.
struct TileRec {
    int t, level, meta, parent;
    int4 ha, hb, ga, gb;
    int4 ca, cb;   // children by octant (CHILD records only)
    int tid;       // the thread's index among the threads working on the tile
};

// t = -1: padding entry, -2: the solve has stopped.  S is never null here (launches
// outside a solve pass the grid's idle block, whose flag is never set), so the stop
// flag is read with no branch and its load overlaps the record's.
template <bool FRESH, bool STOP = true, bool CHILD = false>
__device__ __forceinline__ TileRec rec_tile(const int4* __restrict__ rec, const double* __restrict__ S)
{
    const int4* r = rec + kRecWords * size_t(blockIdx.x);
    const int4 r0 = r[0];
    TileRec x;
    x.ha = r[1];
    x.hb = r[2];
    if (FRESH) {
        x.ga = r[3];
        x.gb = r[4];
    } else {
        x.ga = make_int4(-1, -1, -1, -1);
        x.gb = make_int4(-1, -1, 0, 0);
    }
    if (CHILD) {
        x.ca = r[5];
        x.cb = r[6];
    } else {
        x.ca = x.cb = make_int4(-1, -1, -1, -1);
    }
    x.t = (STOP && S[S_DONE] != 0.0) ? -2 : r0.x;
    x.tid = threadIdx.x;
    x.level = r0.y;
    x.meta = r0.z;
    x.parent = r0.w;
    return x;
}

__device__ __forceinline__ float rec_h(const TopoView& tp, const TileRec& r) { return tp.h0 / float(1 << r.level); }

// LoadAMGLaplacianTileData (src/AMGSolver.cu:395-543).
// FRESH: the halo is evaluated the way AMGFullNegativeLaplacianOnLeafs sees it
//   after PropagateToChildren + AccumulateToParentsOneStep: a GHOST cell is its
//   parent cell and a NONLEAF cell is the mean of its 8 children.
// !FRESH: ghost values come from G, NONLEAF values from X (cycle state).
// Q: the face coefficients and diagonals come from the FP16 copies Qc (c / h).
// All global loads (own cells and the 384 halo cells) are issued before the
// first barrier; only the ghost extrapolation waits for the tile's own values.
// `extra` is called right after the halo loads are issued and before any loaded
// value is used: loads a kernel needs later (not for the stencil) issued there
// are in flight together with the tile's.
struct NoExtraLoads {
    __device__ void operator()() const {}
};
// Prolongation pulled into a sweep (the cycle's finest level after the coarse correction,
// in place of a fas_prolong_v launch): mode 1 (the first post-sweep) reads every own and
// same-level neighbour cell as x + P (x_parent - X0_parent) on INTERIOR cells (0 elsewhere)
// and every GHOST neighbour cell as x_parent on INTERIOR cells (0 elsewhere); mode 2 (the
// later post-sweeps) reads GHOST neighbour cells that way and the rest as they are.  Parent
// values come from G (the cycle's x, whose coarser levels hold the corrected solution) and
// X0; the parents and octants of face neighbours come with the tile record.  The values are
// those fas_prolong_v writes, computed with the same operations.
template <int M>
struct Pull {
    static constexpr int mode = M;
    const float* X0;
    float P;
};
// Q 2: W and L are the coded form's cell words and tables (Coef::w, Coef::lut); the LDS
// may be a form-2 one for a form-0 load (a kernel that takes either form per tile).
template <bool FRESH, typename T, int NT, int Q = 0, typename Extra = NoExtraLoads, typename PullT = Pull<0>, int LQ>
__device__ __forceinline__ void load_tile(const TopoView& tp, const TileRec& r, float h, const T* __restrict__ X,
                                          const T* __restrict__ G, const float* const* C, Lds<T, LQ>& s, Own<T, NT>& o,
                                          const uint16_t* const* Qc = nullptr, const Extra& extra = Extra{},
                                          const PullT& pull = PullT{nullptr, 0.0f}, const uint32_t* __restrict__ W = nullptr,
                                          const float* __restrict__ L = nullptr, const uint32_t* __restrict__ ZW = nullptr)
{
    constexpr int PM = PullT::mode;
    constexpr bool K = Q >= 2;   // coded coefficients (2, 4: staged as codes, 3: as values)
    static_assert((Q == 1) == (LQ == 1) && (K ? LQ == Q : (LQ != 2 && LQ != 4)), "the LDS layout follows the coefficient form");
    static_assert(PM == 0 || (!FRESH && sizeof(T) == 4), "a pulled prolongation is a cycle (FP32) form");
    constexpr int CPT = 512 / NT;
    constexpr int HPT = (384 + NT - 1) / NT;
    const int tid = r.tid;
    const size_t base = size_t(r.t) * 512;
    // the tile's halo record: face neighbours, their tile types, the tile's own
    // type and (FRESH) the parents and octants of GHOST neighbours
    const int4 ha = r.ha, hb = r.hb;
    const uint32_t ty03 = uint32_t(hb.z), ty45 = uint32_t(hb.w);
    const bool dv = Q != 1 && tp.derive && ((ty45 >> 24) & LEAF);   // the FP16 form reads every diagonal
    // K: the level's code table goes out first (its wait then covers no other load)
    constexpr int TAB = code_table<Q>();
    constexpr int LPT = NT >= TAB ? 1 : TAB / NT;
    float lutv[K ? LPT : 1];
    const bool lut_lane = NT <= TAB || tid < TAB;
    if constexpr (K) {
        if (lut_lane) vload<LPT>(L + size_t(r.level) * kWideCodes + size_t(LPT) * tid, lutv);
    }
    // the coefficient planes as values (a runtime-indexed C[axis] would be a load)
    const float* const cp0 = uniform_ptr(C[0]);
    const float* const cp1 = uniform_ptr(C[1]);
    const float* const cp2 = uniform_ptr(C[2]);
    const uint16_t* const qp0 = Q == 1 ? uniform_ptr(Qc[0]) : nullptr;
    const uint16_t* const qp1 = Q == 1 ? uniform_ptr(Qc[1]) : nullptr;
    const uint16_t* const qp2 = Q == 1 ? uniform_ptr(Qc[2]) : nullptr;
    const int4 ga = r.ga, gb = r.gb;
    const uint32_t goct = uint32_t(gb.z);
    constexpr int PC = PM == 1 ? CPT : 1;
    T own_xr[PC], own_px[PC], own_p0[PC];   // PM 1: own x and parent values, combined after the halo loads
    uint32_t ow[K ? CPT : 1];                // K: own cell words, decoded once the table is in LDS
    {
        // own cells: one vector load per array
        const size_t v0 = base + size_t(CPT) * tid;
        T xr[CPT];
        if constexpr (K) {
            vload<CPT>(W + v0, ow);
#pragma unroll
            for (int c = 0; c < CPT; ++c) o.ct[c] = word_type<Q>(ow[c]);
        } else {
            vload<CPT>(tp.ctype + v0, o.ct);
        }
        vload<CPT>(X + v0, xr);
        if constexpr (K) {
            float d3[CPT];
            if (!dv) vload<CPT>(C[3] + v0, d3);
#pragma unroll
            for (int c = 0; c < CPT; ++c) o.d[c] = dv ? 0.0f : d3[c];
        } else if constexpr (Q == 1) {
            uint16_t q[4][CPT];
            vload<CPT>(qp0 + v0, q[0]);
            vload<CPT>(qp1 + v0, q[1]);
            vload<CPT>(qp2 + v0, q[2]);
            if (!dv) vload<CPT>(Qc[3] + v0, q[3]);
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
#pragma unroll
                for (int a = 0; a < 3; ++a) o.c[c][a] = float(__builtin_bit_cast(_Float16, q[a][c])) * h;
                o.d[c] = dv ? 0.0f : float(__builtin_bit_cast(_Float16, q[3][c])) * h;
            }
        } else {
            float cf[4][CPT];
            vload<CPT>(cp0 + v0, cf[0]);
            vload<CPT>(cp1 + v0, cf[1]);
            vload<CPT>(cp2 + v0, cf[2]);
            if (!dv) vload<CPT>(C[3] + v0, cf[3]);
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
#pragma unroll
                for (int a = 0; a < 3; ++a) o.c[c][a] = cf[a][c];
                o.d[c] = dv ? 0.0f : cf[3][c];
            }
        }
        if constexpr (PM == 1) {
            const int pt = r.parent, poc = (r.meta >> 8) & 7;
            const size_t pb = size_t(pt >= 0 ? pt : r.t) * 512;
            // a thread's cells run along k in one row: their parents are CPT / 2 consecutive
            // cells (aligned), read as one vector per array
            constexpr int PP = CPT >= 2 ? CPT / 2 : 1;
            const int v = own_cell<NT>(0, tid);
            const size_t pc = pb + cidx(4 * (poc & 1) + ((v >> 6) >> 1), 4 * ((poc >> 1) & 1) + (((v >> 3) & 7) >> 1),
                                        4 * ((poc >> 2) & 1) + ((v & 7) >> 1));
            T ppx[PP], pp0[PP];
            vload<PP>(G + pc, ppx);
            vload<PP>(pull.X0 + pc, pp0);
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
                own_px[c] = ppx[CPT >= 2 ? c >> 1 : 0];
                own_p0[c] = pp0[CPT >= 2 ? c >> 1 : 0];
            }
#pragma unroll
            for (int c = 0; c < CPT; ++c) own_xr[c] = xr[c];
        } else {
#pragma unroll
            for (int c = 0; c < CPT; ++c) o.x[c] = o.ct[c] == INTERIOR ? xr[c] : T(0);
        }
    }
    // halo, phase 1: every load is issued, unconditionally and with no loaded value used yet
    // (a use, or a register copy at a branch join, makes the wave wait for all loads in
    // flight, the tile's own included).  Cells without a neighbour tile read a cell of the
    // tile itself instead, and the kind decides in phase 2 what the loaded values mean.
    T hraw[HPT];
    constexpr int PH = PM == 1 ? HPT : 1;
    T hpx[PH], hp0[PH];    // PM 1: a same-level neighbour cell's parent values
    uint32_t hcb[HPT];     // the neighbour's + face coefficient: FP16 bits (Q) or FP32 bits
    uint32_t hct[HPT];     // a same-level neighbour cell's type
    int hkind[HPT];        // 0 none, 1 same-level cell (INTERIOR mask), 2 GHOST, 3 GHOST without parent (0),
                           // 4 FRESH NONLEAF (the value is computed below)
    int hni[HPT];          // the neighbour cell (kind 4)
#pragma unroll
    for (int m = 0; m < HPT; ++m) {
        const int k = tid + m * NT;
        const bool in = k < 384;   // false only at 256 and 512 threads a tile, for whole waves
        const int kk = in ? k : 0;
        const int f = kk >> 6, fi = kk & 63;
        const int axis = f >> 1, sgn = (f & 1) ? 1 : -1;
        const int p = fi >> 3, q = fi & 7;
        int nl[3];
        if (axis == 0) { nl[0] = sgn < 0 ? 7 : 0; nl[1] = p; nl[2] = q; }
        else if (axis == 1) { nl[0] = p; nl[1] = sgn < 0 ? 7 : 0; nl[2] = q; }
        else { nl[0] = p; nl[1] = q; nl[2] = sgn < 0 ? 7 : 0; }
        const int slot = sgn < 0 ? axis : axis + 3;
        const int nt = pick6(ha, hb, slot);
        if (in && fi == 0) s.empty[f] = nt < 0;
        const bool ok = in && nt >= 0;
        const int cl = cidx(nl[0], nl[1], nl[2]);
        const size_t ni = size_t(ok ? nt : r.t) * 512 + cl;
        const uint32_t nty = ok ? (slot < 4 ? ty03 >> (8 * slot) : ty45 >> (8 * (slot - 4))) & 0xffu : 0u;
        // a - face's coefficient is the cell's own: the + face reads the neighbour's, a - face
        // re-reads the tile's own first 64 (cached) so that the load is unconditional
        const size_t ci = sgn > 0 ? ni : base + fi;
        // K: the neighbour cell's word, its type and its + face code in one load (z faces from
        // the tile's contiguous copy of those words when there is one: the same word)
        if constexpr (K) hcb[m] = (ZW && axis == 2) ? ZW[base / 4 + (f - 4) * 64 + fi] : W[ni];
        else if constexpr (Q == 1) hcb[m] = (axis == 0 ? qp0 : axis == 1 ? qp1 : qp2)[ci];
        else hcb[m] = __float_as_uint((axis == 0 ? cp0 : axis == 1 ? cp1 : cp2)[ci]);
        const bool ghost = nty & GHOST;
        size_t xi = ni;
        const T* xs = X;
        int kind = ok ? 1 : 0;
        // PM: the face neighbour's parent cell (the record holds every neighbour's parent)
        const int npt = PM ? pick6(ga, gb, slot) : -1;
        const int noc = PM ? int((goct >> (3 * slot)) & 7u) : 0;
        const size_t npc = size_t(npt >= 0 ? npt : r.t) * 512 +
                           cidx(4 * (noc & 1) + (nl[0] >> 1), 4 * ((noc >> 1) & 1) + (nl[1] >> 1), 4 * ((noc >> 2) & 1) + (nl[2] >> 1));
        if constexpr (PM == 1) {
            hpx[m] = G[npc];
            hp0[m] = pull.X0[npc];
        }
        if (ghost && PM) {
            // the GHOST cell as the prolongation leaves it: its parent's value, masked below by its type
            kind = npt >= 0 ? 2 : 3;
            xi = npc;
            xs = G;
        } else if (ghost) {
            if (FRESH) {
                const int pt = pick6(ga, gb, slot);
                const int oc = int((goct >> (3 * slot)) & 7u);
                kind = pt >= 0 ? 2 : 3;
                xi = pt >= 0 ? size_t(pt) * 512 + cidx(4 * (oc & 1) + (nl[0] >> 1), 4 * ((oc >> 1) & 1) + (nl[1] >> 1),
                                                       4 * ((oc >> 2) & 1) + (nl[2] >> 1))
                             : ni;
            } else {
                kind = 2;
                xs = G;
            }
        } else if (FRESH && (nty & NONLEAF)) {
            kind = 4;
        }
        hraw[m] = xs[xi];
        // the type is needed for same-level and NONLEAF neighbours (PM: GHOST ones too); other
        // cells re-read the tile's own (K: it came with the cell word)
        if constexpr (!K) hct[m] = tp.ctype[(kind == 1 || kind == 4 || (PM && kind == 2)) ? ni : base + fi];
        hkind[m] = kind;
        hni[m] = int(ni - size_t(ok ? nt : r.t) * 512) | ((ok ? nt : r.t) << 9);
    }
    // a halo cell's type, read once its load is complete
    auto htype = [&](int m) -> uint32_t {
        if constexpr (K) return word_type<Q>(hcb[m]);
        else return hct[m];
    };
    extra();
    if constexpr (PM == 1) {
        const bool hp = r.parent >= 0;
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const T c1 = hp ? own_px[c] : T(0), c0 = hp ? own_p0[c] : T(0);
            o.x[c] = (o.ct[c] & INTERIOR) ? own_xr[c] + pull.P * (c1 - c0) : T(0);
        }
    }
    if (FRESH) {
        // NONLEAF neighbours (coarse LEAF tiles next to finer ones only): the mean of the 8
        // children of an INTERIOR cell, dependent loads
#pragma unroll
        for (int m = 0; m < HPT; ++m) {
            if (hkind[m] != 4) continue;
            const int nt = hni[m] >> 9, cl = hni[m] & 511;
            const size_t ni = size_t(nt) * 512 + cl;
            T val = T(0);
            if (htype(m) == INTERIOR) {
                const int n0 = cl >> 6, n1 = (cl >> 3) & 7, n2 = cl & 7;
                const int oct = (n0 >> 2) | ((n1 >> 2) << 1) | ((n2 >> 2) << 2);
                const int ctile = tp.child[size_t(nt) * 8 + oct];
                if (ctile >= 0 && (tp.ttype[ctile] & LEAF)) {
                    T sum = 0;
                    const size_t cb = size_t(ctile) * 512;
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                            for (int k2 = 0; k2 < 2; ++k2)
                                sum += X[cb + cidx(2 * (n0 & 3) + ii, 2 * (n1 & 3) + jj, 2 * (n2 & 3) + k2)];
                    val = sum * T(0.125);
                } else {
                    val = X[ni];
                }
            }
            hraw[m] = val;
        }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int v = own_cell<NT>(c, tid);
        s.x[SX(v >> 6, (v >> 3) & 7, v & 7)] = o.x[c];
    }
    if constexpr (K) {
        // the table now; form 2 stages the own codes too, form 3 the values once the table is visible
        if (lut_lane) vstore<LPT>(&s.lut[LPT * tid], lutv);
        if constexpr (Q == 2 || Q == 4) {
            using CodeT = std::conditional_t<Q == 4, uint16_t, uint8_t>;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                CodeT oc[CPT];
#pragma unroll
                for (int c = 0; c < CPT; ++c) oc[c] = CodeT(word_code<Q>(ow[c], a));
                vstore<CPT>(&s.off[a][CPT * tid], oc);
            }
        }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if constexpr (Q == 1) {
                // back to the FP16 bits: c / h is exact, and so is its conversion
                const float ih = 1.0f / h;
                uint16_t oc[CPT];
#pragma unroll
                for (int c = 0; c < CPT; ++c) oc[c] = __builtin_bit_cast(uint16_t, _Float16(o.c[c][a] * ih));
                vstore<CPT>(&s.off[a][CPT * tid], oc);
            } else {
                float oc[CPT];
#pragma unroll
                for (int c = 0; c < CPT; ++c) oc[c] = o.c[c][a];
                vstore<CPT>(&s.off[a][CPT * tid], oc);
            }
        }
    }
    __syncthreads();
    if constexpr (K) {
        // own face coefficients from the codes of the cell word; form 3 stages
        // them (phase 2 below reads no coefficient plane, and the planes are read after the
        // next barrier)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float oc[CPT];
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
                oc[c] = s.lut[word_code<Q>(ow[c], a)];
                o.c[c][a] = oc[c];
            }
            if constexpr (Q == 3) vstore<CPT>(&s.off[a][CPT * tid], oc);
        }
    }
#pragma unroll
    for (int m = 0; m < HPT; ++m) {
        const int k = tid + m * NT;
        if (k >= 384) continue;
        const int f = k >> 6, fi = k & 63;
        const int axis = f >> 1, sgn = (f & 1) ? 1 : -1;
        const int p = fi >> 3, q = fi & 7;
        int fl[3];
        const int axk = sgn < 0 ? -1 : 8;
        if (axis == 0) { fl[0] = axk; fl[1] = p; fl[2] = q; }
        else if (axis == 1) { fl[0] = p; fl[1] = axk; fl[2] = q; }
        else { fl[0] = p; fl[1] = q; fl[2] = axk; }
        // halo, phase 2: the loaded values
        T hv = hkind[m] == 0 || hkind[m] == 3 || (hkind[m] == 1 && htype(m) != INTERIOR) ? T(0) : hraw[m];
        if constexpr (PM != 0) {
            if (hkind[m] == 2 && !(htype(m) & INTERIOR)) hv = T(0);
        }
        if constexpr (PM == 1) {
            if (hkind[m] == 1 && htype(m) == INTERIOR) {
                const int npt = pick6(ga, gb, (f >> 1) + ((f & 1) ? 3 : 0));
                const T c1 = npt >= 0 ? hpx[m] : T(0), c0 = npt >= 0 ? hp0[m] : T(0);
                hv = hraw[m] + pull.P * (c1 - c0);
            }
        }
        if (hkind[m] == 2 || hkind[m] == 3) {
            int cl[3] = {fl[0], fl[1], fl[2]};
            cl[axis] -= sgn;
            T avg = 0;
#pragma unroll
            for (int di = 0; di < 2; ++di)
#pragma unroll
                for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                    for (int dk = 0; dk < 2; ++dk) avg += s.x[SX(cl[0] ^ di, cl[1] ^ dj, cl[2] ^ dk)];
            avg /= 8;
            if constexpr (sizeof(T) == 4)
                hv = float(double(s.x[SX(cl[0], cl[1], cl[2])]) + 0.5 * double(hv - avg));
            else
                hv = s.x[SX(cl[0], cl[1], cl[2])] + 0.5 * (hv - avg);
        }
        s.x[SX(fl[0], fl[1], fl[2])] = hv;
        if (sgn > 0) {
            // no neighbour tile: h (the halo value there is 0); FP16 1.0 times h in the Q form
            if constexpr (Q == 1) s.off[axis][512 + fi] = hkind[m] != 0 ? uint16_t(hcb[m]) : uint16_t(0x3C00);
            else if constexpr (Q == 2) s.off[axis][512 + fi] = hkind[m] != 0 ? uint8_t(word_code<Q>(hcb[m], axis)) : uint8_t(kCodes - 1);
            else if constexpr (Q == 4) s.off[axis][512 + fi] = hkind[m] != 0 ? uint16_t(word_code<Q>(hcb[m], axis)) : uint16_t(kWideCodes - 1);
            else if constexpr (Q == 3) s.off[axis][512 + fi] = hkind[m] != 0 ? s.lut[word_code<Q>(hcb[m], axis)] : h;
            else s.off[axis][512 + fi] = hkind[m] != 0 ? __uint_as_float(hcb[m]) : h;
        }
    }
    __syncthreads();
    if constexpr (Q != 1) {
        if (dv) {
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
                const int v = own_cell<NT>(c, tid);
                const float dd = derived_diag(s, v, o.c[c], h);
                o.d[c] = (o.ct[c] & INTERIOR) ? dd : 0.0f;
            }
        }
    }
}

// AMGLaplacianTileData::negativeLap (src/AMGSolver.cu:345-382) for cell v
// (h: the tile spacing, needed only to read FP16 staged coefficients)
template <typename T, int LQ>
__device__ __forceinline__ T neg_lap(const Lds<T, LQ>& s, int v, T x0, const float c[3], float d, float h = 0.0f)
{
    if (d == 0) return T(0);
    const int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
    const float cp0 = staged_coef(s, 0, i < 7 ? v + 64 : 512 + (j * 8 + k), h);
    const float cp1 = staged_coef(s, 1, j < 7 ? v + 8 : 512 + (i * 8 + k), h);
    const float cp2 = staged_coef(s, 2, k < 7 ? v + 1 : 512 + (i * 8 + j), h);
    T r = x0 * T(d);
    r = fma(T(c[0]), s.x[SX(i - 1, j, k)], r);
    r = fma(T(cp0), s.x[SX(i + 1, j, k)], r);
    r = fma(T(c[1]), s.x[SX(i, j - 1, k)], r);
    r = fma(T(cp1), s.x[SX(i, j + 1, k)], r);
    r = fma(T(c[2]), s.x[SX(i, j, k - 1)], r);
    r = fma(T(cp2), s.x[SX(i, j, k + 1)], r);
    return r;
}


// ------------------------------------------------- pipelined tile loops ----
// The hot stencil kernels run persistent workgroups that walk their share of a
// launch list (entries blockIdx.x, +gridDim.x, ...).  While one tile is staged
// in LDS and computed, all global loads of the workgroup's next tile are
// already in flight in a second register stage.

// global loads of one tile, held in registers until the tile is processed
template <typename T, int NT, bool WB>
struct Stage {
    static constexpr int CPT = 512 / NT;
    static constexpr int HPT = (384 + NT - 1) / NT;
    int t;
    T xr[CPT];
    uint8_t ct[CPT];
    float c[CPT][3];
    float d[CPT];
    float b[WB ? CPT : 1];
    T hraw[HPT];
    float hcp[HPT];
    bool hghost[HPT];
};

template <bool FRESH, typename T, int NT, bool WB>
__device__ __forceinline__ void issue(Stage<T, NT, WB>& st, const TopoView& tp, int t, const T* __restrict__ X,
                                      const T* __restrict__ G, const float* const* C, const float* __restrict__ B)
{
    constexpr int CPT = 512 / NT;
    constexpr int HPT = (384 + NT - 1) / NT;
    st.t = t;
    if (t < 0) return;
    const int tid = threadIdx.x;
    const size_t base = size_t(t) * 512;
    const float h = tile_h(tp, t);
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int v = tid + c * NT;
        st.ct[c] = tp.ctype[base + v];
        st.xr[c] = X[base + v];
#pragma unroll
        for (int a = 0; a < 3; ++a) st.c[c][a] = C[a][base + v];
        st.d[c] = C[3][base + v];
        if (WB) st.b[c] = B[base + v];
    }
#pragma unroll
    for (int m = 0; m < HPT; ++m) {
        const int k = tid + m * NT;
        st.hraw[m] = T(0);
        st.hcp[m] = h;
        st.hghost[m] = false;
        if (k >= 384) continue;
        const int f = k >> 6, fi = k & 63;
        const int axis = f >> 1, sgn = (f & 1) ? 1 : -1;
        const int p = fi >> 3, q = fi & 7;
        int nl[3];
        if (axis == 0) { nl[0] = sgn < 0 ? 7 : 0; nl[1] = p; nl[2] = q; }
        else if (axis == 1) { nl[0] = p; nl[1] = sgn < 0 ? 7 : 0; nl[2] = q; }
        else { nl[0] = p; nl[1] = q; nl[2] = sgn < 0 ? 7 : 0; }
        const int nt = tp.nbr[size_t(t) * 6 + (sgn < 0 ? axis : axis + 3)];
        if (nt < 0) continue;
        const size_t ni = size_t(nt) * 512 + cidx(nl[0], nl[1], nl[2]);
        const uint8_t nty = tp.ttype[nt];
        if (sgn > 0) st.hcp[m] = C[axis][ni];
        if (nty & GHOST) {
            st.hghost[m] = true;
            if (FRESH) {
                const int pt = tp.parent[nt];
                const int oc = tp.toct[nt];
                st.hraw[m] = pt < 0 ? T(0)
                                    : X[size_t(pt) * 512 + cidx(4 * (oc & 1) + (nl[0] >> 1), 4 * ((oc >> 1) & 1) + (nl[1] >> 1),
                                                                 4 * ((oc >> 2) & 1) + (nl[2] >> 1))];
            } else {
                st.hraw[m] = G[ni];
            }
        } else if (FRESH && (nty & NONLEAF)) {
            if (tp.ctype[ni] == INTERIOR) {
                const int oct = (nl[0] >> 2) | ((nl[1] >> 2) << 1) | ((nl[2] >> 2) << 2);
                const int ctile = tp.child[size_t(nt) * 8 + oct];
                if (ctile >= 0 && (tp.ttype[ctile] & LEAF)) {
                    T sum = 0;
                    const size_t cb = size_t(ctile) * 512;
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                            for (int k2 = 0; k2 < 2; ++k2)
                                sum += X[cb + cidx(2 * (nl[0] & 3) + ii, 2 * (nl[1] & 3) + jj, 2 * (nl[2] & 3) + k2)];
                    st.hraw[m] = sum * T(0.125);
                } else {
                    st.hraw[m] = X[ni];
                }
            }
        } else {
            const T xv = X[ni];
            st.hraw[m] = tp.ctype[ni] == INTERIOR ? xv : T(0);
        }
    }
}

// stage -> LDS (own cells, then the halo with the ghost extrapolation); the
// caller guarantees the previous tile's LDS reads are finished
template <typename T, int NT, bool WB>
__device__ __forceinline__ void stage_lds(const Stage<T, NT, WB>& st, Lds<T>& s, Own<T, NT>& o)
{
    constexpr int CPT = 512 / NT;
    constexpr int HPT = (384 + NT - 1) / NT;
    const int tid = threadIdx.x;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int v = tid + c * NT;
        o.ct[c] = st.ct[c];
        o.x[c] = st.ct[c] == INTERIOR ? st.xr[c] : T(0);
        o.d[c] = st.d[c];
        s.x[SX(v >> 6, (v >> 3) & 7, v & 7)] = o.x[c];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o.c[c][a] = st.c[c][a];
            s.off[a][v] = st.c[c][a];
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < HPT; ++m) {
        const int k = tid + m * NT;
        if (k >= 384) continue;
        const int f = k >> 6, fi = k & 63;
        const int axis = f >> 1, sgn = (f & 1) ? 1 : -1;
        const int p = fi >> 3, q = fi & 7;
        int fl[3];
        const int axk = sgn < 0 ? -1 : 8;
        if (axis == 0) { fl[0] = axk; fl[1] = p; fl[2] = q; }
        else if (axis == 1) { fl[0] = p; fl[1] = axk; fl[2] = q; }
        else { fl[0] = p; fl[1] = q; fl[2] = axk; }
        T hv = st.hraw[m];
        if (st.hghost[m]) {
            int cl[3] = {fl[0], fl[1], fl[2]};
            cl[axis] -= sgn;
            T avg = 0;
#pragma unroll
            for (int di = 0; di < 2; ++di)
#pragma unroll
                for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                    for (int dk = 0; dk < 2; ++dk) avg += s.x[SX(cl[0] ^ di, cl[1] ^ dj, cl[2] ^ dk)];
            avg /= 8;
            if constexpr (sizeof(T) == 4)
                hv = float(double(s.x[SX(cl[0], cl[1], cl[2])]) + 0.5 * double(st.hraw[m] - avg));
            else
                hv = s.x[SX(cl[0], cl[1], cl[2])] + 0.5 * (st.hraw[m] - avg);
        }
        s.x[SX(fl[0], fl[1], fl[2])] = hv;
        if (sgn > 0) s.off[axis][512 + fi] = st.hcp[m];
    }
    __syncthreads();
}

// y = A x on LEAF tiles (persistent, pipelined); optional x.y partial per workgroup
template <bool DOT, int NT>
__global__ __launch_bounds__(NT) void apply_p(TopoView tp, const int32_t* __restrict__ list, int n, const float* __restrict__ X,
                                              Coef C, float* __restrict__ Y, double* __restrict__ partial,
                                              const double* __restrict__ S)
{
    __shared__ Lds<float> s;
    __shared__ double red[8];
    if (stopped(S)) return;
    int i = blockIdx.x;
    Stage<float, NT, false> cur, nxt;
    issue<true>(cur, tp, list[i], X, X, C.c, nullptr);
    double acc = 0;
    for (;;) {
        const int in = i + gridDim.x;
        if (in < n) issue<true>(nxt, tp, list[in], X, X, C.c, nullptr);
        if (cur.t >= 0) {
            Own<float, NT> o;
            stage_lds(cur, s, o);
#pragma unroll
            for (int c = 0; c < 512 / NT; ++c) {
                const int v = threadIdx.x + c * NT;
                const float y = o.ct[c] == INTERIOR ? neg_lap(s, v, o.x[c], o.c[c], o.d[c]) : 0.0f;
                Y[size_t(cur.t) * 512 + v] = y;
                if (DOT) acc += double(cur.xr[c]) * double(y);
            }
            __syncthreads();
        }
        if (in >= n) break;
        cur = nxt;
        i = in;
    }
    if (DOT) {
        double b = block_sum(acc, red);
        if (threadIdx.x == 0) partial[blockIdx.x] = b;
    }
}

// damped Jacobi sweep x_out = x_in + w (b - A x_in) / D (persistent, pipelined)
template <int NT>
__global__ __launch_bounds__(NT) void jacobi_p(TopoView tp, const int32_t* __restrict__ list, int n, const float* __restrict__ Xin,
                                               float* __restrict__ Xout, const float* __restrict__ G, Coef C,
                                               const float* __restrict__ B, float w, const double* __restrict__ S)
{
    __shared__ Lds<float> s;
    if (stopped(S)) return;
    int i = blockIdx.x;
    Stage<float, NT, true> cur, nxt;
    issue<false>(cur, tp, list[i], Xin, G, C.c, B);
    for (;;) {
        const int in = i + gridDim.x;
        if (in < n) issue<false>(nxt, tp, list[in], Xin, G, C.c, B);
        if (cur.t >= 0) {
            Own<float, NT> o;
            stage_lds(cur, s, o);
#pragma unroll
            for (int c = 0; c < 512 / NT; ++c) {
                const int v = threadIdx.x + c * NT;
                const float ax = neg_lap(s, v, o.x[c], o.c[c], o.d[c]);
                const float delta = (o.ct[c] & INTERIOR) ? (cur.b[c] - ax) / o.d[c] : 0.0f;
                Xout[size_t(cur.t) * 512 + v] = o.x[c] + w * delta;
            }
            __syncthreads();
        }
        if (in >= n) break;
        cur = nxt;
        i = in;
    }
}

// ResidualAndRestrictAMG128Kernel (src/AMGSolver.cu:896-964), persistent and pipelined
template <int NT>
__global__ __launch_bounds__(NT) void restrict_p(TopoView tp, const int32_t* __restrict__ list, int n, float* __restrict__ X,
                                                 Coef C, float* __restrict__ B, float Rr, const double* __restrict__ S)
{
    __shared__ Lds<float> s;
    __shared__ float res[512];
    if (stopped(S)) return;
    int i = blockIdx.x;
    Stage<float, NT, true> cur, nxt;
    issue<false>(cur, tp, list[i], X, X, C.c, B);
    for (;;) {
        const int in = i + gridDim.x;
        if (in < n) issue<false>(nxt, tp, list[in], X, X, C.c, B);
        if (cur.t >= 0) {
            Own<float, NT> o;
            stage_lds(cur, s, o);
#pragma unroll
            for (int c = 0; c < 512 / NT; ++c) {
                const int v = threadIdx.x + c * NT;
                res[v] = cur.b[c] - neg_lap(s, v, o.x[c], o.c[c], o.d[c]);
            }
            __syncthreads();
            const int v = threadIdx.x;
            const int pt = tp.parent[cur.t];
            if (v < 64 && pt >= 0) {
                const int oc = tp.toct[cur.t];
                const int fl0 = v / 16 * 2, fl1 = (v / 4) % 4 * 2, fl2 = v % 4 * 2;
                float rs = 0, xs = 0;
#pragma unroll
                for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                        for (int kk = 0; kk < 2; ++kk) {
                            rs += res[cidx(fl0 + ii, fl1 + jj, fl2 + kk)];
                            xs += s.x[SX(fl0 + ii, fl1 + jj, fl2 + kk)];
                        }
                const size_t pidx = size_t(pt) * 512 +
                                    cidx(4 * (oc & 1) + (fl0 >> 1), 4 * ((oc >> 1) & 1) + (fl1 >> 1), 4 * ((oc >> 2) & 1) + (fl2 >> 1));
                B[pidx] = rs * Rr;
                X[pidx] = xs / 8;
            }
            __syncthreads();
        }
        if (in >= n) break;
        cur = nxt;
        i = in;
    }
}

// ------------------------------------------------------ vector (float4) ----
// (CMG_TILE_LOOP walks the tiles of a list, 128 threads x float4 per tile)

// Z0: with the first damped-Jacobi sweep from x = 0 of level `zlevel` fused in,
// z0 = 0 + w (r / D) on INTERIOR cells (0 elsewhere) of the level's tiles, from
// the updated r (bit-identical to jacobi_zero_v on r)
// X == nullptr: x += alpha p is deferred to the p update of the same iteration, which reads p
// anyway (update_p_v with X; x_pending_v after the last iteration); S[S_XPEND] marks it pending
template <bool Z0>
__global__ __launch_bounds__(256) void update_xr_v(TopoView tp, const int32_t* __restrict__ list, int n, float* __restrict__ X,
                                                   float* __restrict__ R, const float* __restrict__ P,
                                                   const float* __restrict__ AP, double* __restrict__ partial,
                                                   double* __restrict__ S, int zlevel, float* __restrict__ Zo,
                                                   const float* __restrict__ D, float w, const uint16_t* __restrict__ Dq)
{
    __shared__ double red[8];
    if (stopped(S)) return;
    const float a = float(S[S_ALPHA]);
    const bool with_x = X != nullptr;
    if (!with_x && blockIdx.x == 0 && threadIdx.x == 0) S[S_XPEND] = 1.0;
    double acc = 0;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        float4 r = reinterpret_cast<const float4*>(R)[q];
        const float4 ap = reinterpret_cast<const float4*>(AP)[q];
        if (with_x) {
            float4 x = reinterpret_cast<const float4*>(X)[q];
            const float4 p = reinterpret_cast<const float4*>(P)[q];
            if (ct.x & INTERIOR) x.x = fmaf(a, p.x, x.x);
            if (ct.y & INTERIOR) x.y = fmaf(a, p.y, x.y);
            if (ct.z & INTERIOR) x.z = fmaf(a, p.z, x.z);
            if (ct.w & INTERIOR) x.w = fmaf(a, p.w, x.w);
            reinterpret_cast<float4*>(X)[q] = x;
        }
        if (ct.x & INTERIOR) r.x = fmaf(-a, ap.x, r.x);
        if (ct.y & INTERIOR) r.y = fmaf(-a, ap.y, r.y);
        if (ct.z & INTERIOR) r.z = fmaf(-a, ap.z, r.z);
        if (ct.w & INTERIOR) r.w = fmaf(-a, ap.w, r.w);
        reinterpret_cast<float4*>(R)[q] = r;
        acc += double(r.x) * double(r.x) + double(r.y) * double(r.y) + double(r.z) * double(r.z) + double(r.w) * double(r.w);
        if (Z0 && tp.tlevel[t_] == zlevel) {
            const float4 d = Dq ? coef_q4(Dq, q, tile_h(tp, t_)) : reinterpret_cast<const float4*>(D)[q];
            float4 z;
            z.x = 0.0f + w * ((ct.x & INTERIOR) ? (r.x - 0.0f) / d.x : 0.0f);
            z.y = 0.0f + w * ((ct.y & INTERIOR) ? (r.y - 0.0f) / d.y : 0.0f);
            z.z = 0.0f + w * ((ct.z & INTERIOR) ? (r.z - 0.0f) / d.z : 0.0f);
            z.w = 0.0f + w * ((ct.w & INTERIOR) ? (r.w - 0.0f) / d.w : 0.0f);
            reinterpret_cast<float4*>(Zo)[q] = z;
        }
    }
    double b = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// X: also the iteration's deferred x += alpha p (the p it reads is the iteration's), which
// clears S[S_XPEND]
__global__ __launch_bounds__(256) void update_p_v(TopoView tp, const int32_t* __restrict__ list, int n, float* __restrict__ P,
                                                  const float* __restrict__ Z, double* __restrict__ S, float* __restrict__ X)
{
    if (stopped(S)) return;
    const float b = float(S[S_BETA]);
    const float a = float(S[S_ALPHA]);
    if (X && blockIdx.x == 0 && threadIdx.x == 0) S[S_XPEND] = 0.0;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        float4 p = reinterpret_cast<const float4*>(P)[q];
        const float4 z = reinterpret_cast<const float4*>(Z)[q];
        if (X) {
            float4 x = reinterpret_cast<const float4*>(X)[q];
            if (ct.x & INTERIOR) x.x = fmaf(a, p.x, x.x);
            if (ct.y & INTERIOR) x.y = fmaf(a, p.y, x.y);
            if (ct.z & INTERIOR) x.z = fmaf(a, p.z, x.z);
            if (ct.w & INTERIOR) x.w = fmaf(a, p.w, x.w);
            reinterpret_cast<float4*>(X)[q] = x;
        }
        if (ct.x & INTERIOR) p.x = fmaf(b, p.x, z.x);
        if (ct.y & INTERIOR) p.y = fmaf(b, p.y, z.y);
        if (ct.z & INTERIOR) p.z = fmaf(b, p.z, z.z);
        if (ct.w & INTERIOR) p.w = fmaf(b, p.w, z.w);
        reinterpret_cast<float4*>(P)[q] = p;
    }
}

// the deferred x += alpha p of a solve's last iteration (its p update did not run), when pending
__global__ __launch_bounds__(256) void x_pending_v(TopoView tp, const int32_t* __restrict__ list, int n, float* __restrict__ X,
                                                   const float* __restrict__ P, const double* __restrict__ S)
{
    if (S[S_XPEND] == 0.0) return;
    const float a = float(S[S_ALPHA]);
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        float4 x = reinterpret_cast<const float4*>(X)[q];
        const float4 p = reinterpret_cast<const float4*>(P)[q];
        if (ct.x & INTERIOR) x.x = fmaf(a, p.x, x.x);
        if (ct.y & INTERIOR) x.y = fmaf(a, p.y, x.y);
        if (ct.z & INTERIOR) x.z = fmaf(a, p.z, x.z);
        if (ct.w & INTERIOR) x.w = fmaf(a, p.w, x.w);
        reinterpret_cast<float4*>(X)[q] = x;
    }
}

__global__ __launch_bounds__(256) void dot_v(const int32_t* __restrict__ list, int n, const float* __restrict__ A,
                                             const float* __restrict__ B, double* __restrict__ partial,
                                             const double* __restrict__ S)
{
    __shared__ double red[8];
    if (stopped(S)) return;
    double acc = 0;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const float4 a = reinterpret_cast<const float4*>(A)[q];
        const float4 b = reinterpret_cast<const float4*>(B)[q];
        acc += double(a.x) * double(b.x) + double(a.y) * double(b.y) + double(a.z) * double(b.z) + double(a.w) * double(b.w);
    }
    double b = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// two dot products sharing the first operand: partial = A.B, partial2 = A.C
__global__ __launch_bounds__(256) void dot2_v(const int32_t* __restrict__ list, int n, const float* __restrict__ A,
                                              const float* __restrict__ B, const float* __restrict__ Cv,
                                              double* __restrict__ partial, double* __restrict__ partial2,
                                              const double* __restrict__ S)
{
    __shared__ double red[8];
    if (stopped(S)) return;
    double acc = 0, acc2 = 0;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const float4 a = reinterpret_cast<const float4*>(A)[q];
        const float4 b = reinterpret_cast<const float4*>(B)[q];
        const float4 c = reinterpret_cast<const float4*>(Cv)[q];
        acc += double(a.x) * double(b.x) + double(a.y) * double(b.y) + double(a.z) * double(b.z) + double(a.w) * double(b.w);
        acc2 += double(a.x) * double(c.x) + double(a.y) * double(c.y) + double(a.z) * double(c.z) + double(a.w) * double(c.w);
    }
    double b = block_sum(acc, red);
    __syncthreads();
    double b2 = block_sum(acc2, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = b;
        partial2[blockIdx.x] = b2;
    }
}

// first damped-Jacobi sweep from x = 0: with every x (own, neighbours, ghosts)
// zero, A x = +0 and the sweep reduces to x = 0 + w (b / D), bit-identical to
// the general sweep; it reads b, D and the cell type only
__global__ __launch_bounds__(256) void jacobi_zero_v(TopoView tp, const int32_t* __restrict__ list, int n, float* __restrict__ Xout,
                                                     const float* __restrict__ D, const float* __restrict__ B, float w,
                                                     const double* __restrict__ S, const uint16_t* __restrict__ Dq)
{
    if (stopped(S)) return;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        const float4 b = reinterpret_cast<const float4*>(B)[q];
        const float4 d = Dq ? coef_q4(Dq, q, tile_h(tp, t_)) : reinterpret_cast<const float4*>(D)[q];
        float4 x;
        x.x = 0.0f + w * ((ct.x & INTERIOR) ? (b.x - 0.0f) / d.x : 0.0f);
        x.y = 0.0f + w * ((ct.y & INTERIOR) ? (b.y - 0.0f) / d.y : 0.0f);
        x.z = 0.0f + w * ((ct.z & INTERIOR) ? (b.z - 0.0f) / d.z : 0.0f);
        x.w = 0.0f + w * ((ct.w & INTERIOR) ? (b.w - 0.0f) / d.w : 0.0f);
        reinterpret_cast<float4*>(Xout)[q] = x;
    }
}

// flat 16-byte copy (the HBM copy ceiling of the measurement tools): one
// element per thread over a full grid, non-temporal (streaming) load and store;
// on MI355X this form copies 1 GiB at about 6.5 TB/s, grid-stride forms at 4.5-4.9
typedef float f32x4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy_flat(const f32x4v* __restrict__ A, f32x4v* __restrict__ B, size_t n4)
{
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(&A[i]), &B[i]);
}

// q[a] = FP16 bits of c[a] / h on every tile of the list, 128 threads x 4 cells a
// tile (float4 in, 4 halves out per channel); *bad counts workgroups that produced
// a non-finite value
// four coefficients as FP16 bits of c / h; `nf` is set by a value that does not fit: a
// non-finite result, or (DIAG) a nonzero diagonal that falls below FP16's smallest normal
// (2^-14), which would divide the Jacobi and Gauss-Seidel updates by 0 or by a few bits
template <bool DIAG = false>
__device__ __forceinline__ uint2 half4_of(float4 v, float h, bool& nf)
{
    const _Float16 a = _Float16(v.x / h), b = _Float16(v.y / h), c = _Float16(v.z / h), d = _Float16(v.w / h);
    nf = nf || !isfinite(float(a)) || !isfinite(float(b)) || !isfinite(float(c)) || !isfinite(float(d));
    if (DIAG) {
        constexpr float kMinNormal = 6.103515625e-05f;
        nf = nf || (v.x != 0.0f && fabsf(float(a)) < kMinNormal) || (v.y != 0.0f && fabsf(float(b)) < kMinNormal) ||
             (v.z != 0.0f && fabsf(float(c)) < kMinNormal) || (v.w != 0.0f && fabsf(float(d)) < kMinNormal);
    }
    const uint32_t lo = uint32_t(__builtin_bit_cast(uint16_t, a)) | (uint32_t(__builtin_bit_cast(uint16_t, b)) << 16);
    const uint32_t hi = uint32_t(__builtin_bit_cast(uint16_t, c)) | (uint32_t(__builtin_bit_cast(uint16_t, d)) << 16);
    return make_uint2(lo, hi);
}

__global__ __launch_bounds__(256) void coef_to_half(TopoView tp, const int32_t* __restrict__ list, int n, Coef C,
                                                    uint16_t* __restrict__ q0, uint16_t* __restrict__ q1,
                                                    uint16_t* __restrict__ q2, uint16_t* __restrict__ q3,
                                                    unsigned int* __restrict__ bad)
{
    bool nf = false;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const float h = tile_h(tp, t_);
        const float4 c0 = reinterpret_cast<const float4*>(C.c[0])[q], c1 = reinterpret_cast<const float4*>(C.c[1])[q];
        const float4 c2 = reinterpret_cast<const float4*>(C.c[2])[q], c3 = reinterpret_cast<const float4*>(C.c[3])[q];
        reinterpret_cast<uint2*>(q0)[q] = half4_of(c0, h, nf);
        reinterpret_cast<uint2*>(q1)[q] = half4_of(c1, h, nf);
        reinterpret_cast<uint2*>(q2)[q] = half4_of(c2, h, nf);
        reinterpret_cast<uint2*>(q3)[q] = half4_of<true>(c3, h, nf);
    }
    if (__syncthreads_or(nf) && threadIdx.x == 0) atomicAdd(bad, 1u);
}

// ------------------------------------------------- exact coefficient codes ----
// Per level, the distinct FP32 bit patterns of the three face coefficients are gathered in
// an open-addressing set of kCodeSlots keys (code_collect).  A level with at most
// kWideCodes - 1 of them numbers them in slot order and writes its table of the values,
// the level's spacing in entry kCodes - 1 (at most kCodes - 1 values: one-byte codes) or
// kWideCodes - 1 (ten-bit codes) (code_number); every cell of such a level then gets its
// word, type | code(c0) << 8 | code(c1) << 16 | code(c2) << 24 or, with ten-bit codes,
// type2 | code(c0) << 2 | code(c1) << 12 | code(c2) << 22 with type2 = 0..3 for the cell
// types 0, DIRICHLET, INTERIOR, NEUMANN (a level with another type value is not coded)
// (code_encode).  A table entry holds the coefficient's own bits, so a kernel that
// decodes computes exactly what it computes from the FP32 channels.
// Workspace (uint32): keys [kMaxLevels][kCodeSlots], slot codes [kMaxLevels][kCodeSlots],
// then per level the key count, a "not codable" flag and the result (1: coded).
constexpr int kCodeSlots = 4096;   // a quarter full at most
constexpr uint32_t kNoKey = 0xffffffffu;   // a NaN pattern; a coefficient with these bits leaves its level uncoded

__device__ __forceinline__ uint32_t code_slot(uint32_t key) { return (key * 2654435761u) >> 20; }

__device__ __forceinline__ void code_insert(uint32_t* __restrict__ set, unsigned int* count, unsigned int* bad,
                                            uint32_t key)
{
    if (key == kNoKey) {
        *bad = 1u;
        return;
    }
    uint32_t sl = code_slot(key);
    for (int p = 0; p < kCodeSlots; ++p, sl = (sl + 1) & (kCodeSlots - 1)) {
        // a slot goes from kNoKey to its key once; a stale kNoKey read is settled by the CAS
        const uint32_t cur = set[sl];
        if (cur == key) return;
        if (cur != kNoKey) continue;
        const uint32_t prev = atomicCAS(&set[sl], kNoKey, key);
        if (prev == kNoKey) {
            atomicAdd(count, 1u);
            return;
        }
        if (prev == key) return;
    }
    *bad = 1u;   // more distinct values than slots
}

// f(key) for this lane's key, with f called once per distinct key of the wave (by the first
// lane holding it; the lanes of the wave must all be active): most waves hold one to three
// distinct coefficient values
template <typename F>
__device__ __forceinline__ uint32_t per_distinct_key(uint32_t key, F f)
{
    uint32_t out = 0;
    unsigned long long pending = __ballot(1);
    while (pending) {
        const int first = __ffsll(pending) - 1;
        const uint32_t k0 = __builtin_amdgcn_readlane(key, first);
        uint32_t r = 0;
        if (int(__lane_id()) == first) r = f(k0);
        r = __builtin_amdgcn_readlane(r, first);
        const bool mine = key == k0;
        if (mine) out = r;
        pending &= ~__ballot(mine);
    }
    return out;
}

__global__ __launch_bounds__(256) void code_collect(TopoView tp, const int32_t* __restrict__ list, int n, Coef C,
                                                    uint32_t* __restrict__ ws)
{
    uint32_t* keys = ws;
    unsigned int* count = ws + 2 * kMaxLevels * kCodeSlots;
    unsigned int* bad = count + kMaxLevels;
    CMG_TILE_LOOP(list, n)
    {
        const int l = tp.tlevel[t_];
        // a level known to be uncodable needs no more work (a stale read only costs time)
        if (bad[l] || count[l] > unsigned(kWideCodes)) continue;
        const size_t q = quad_index(t_);
        auto ins = [&](uint32_t k) -> uint32_t {
            code_insert(keys + l * kCodeSlots, count + l, bad + l, k);
            return 0u;
        };
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint4 v = reinterpret_cast<const uint4*>(C.c[a])[q];
            per_distinct_key(v.x, ins);
            per_distinct_key(v.y, ins);
            per_distinct_key(v.z, ins);
            per_distinct_key(v.w, ins);
        }
    }
}

// one workgroup per level: codes 0..n-1 to the occupied slots in slot order, the table
// (kWideCodes entries, unused ones 0), and the level's result (1 one-byte, 2 ten-bit codes)
__global__ __launch_bounds__(256) void code_number(uint32_t* __restrict__ ws, float* __restrict__ lut, float h0)
{
    const int l = blockIdx.x;
    const uint32_t* keys = ws + l * kCodeSlots;
    uint32_t* slotcode = ws + kMaxLevels * kCodeSlots + l * kCodeSlots;
    const unsigned int* count = ws + 2 * kMaxLevels * kCodeSlots;
    const unsigned int* bad = count + kMaxLevels;
    unsigned int* ok = ws + 2 * kMaxLevels * kCodeSlots + 2 * kMaxLevels;
    const bool coded = !bad[l] && count[l] <= unsigned(kWideCodes - 1);
    const bool narrow = count[l] <= unsigned(kCodes - 1);
    // the last entry of the code range: the level's spacing, as the tile kernels compute it (rec_h)
    const int last = narrow ? kCodes - 1 : kWideCodes - 1;
    for (int e = threadIdx.x; e < kWideCodes; e += blockDim.x)
        lut[l * kWideCodes + e] = e == last ? h0 / float(1 << l) : 0.0f;
    __syncthreads();
    if (!coded) {
        if (threadIdx.x == 0) ok[l] = 0u;
        return;
    }
    __shared__ unsigned int part[256];
    constexpr int SPT = kCodeSlots / 256;   // slots a thread numbers, consecutive
    uint32_t k[SPT];
    unsigned int mine = 0;
#pragma unroll
    for (int e = 0; e < SPT; ++e) {
        k[e] = keys[threadIdx.x * SPT + e];
        mine += k[e] != kNoKey;
    }
    part[threadIdx.x] = mine;
    __syncthreads();
    unsigned int first = 0;   // occupied slots before this thread's (256 partial sums, read directly)
    for (int i = 0; i < int(threadIdx.x); ++i) first += part[i];
#pragma unroll
    for (int e = 0; e < SPT; ++e) {
        if (k[e] == kNoKey) continue;
        slotcode[threadIdx.x * SPT + e] = first;
        lut[l * kWideCodes + first] = __uint_as_float(k[e]);
        ++first;
    }
    if (threadIdx.x == 0) ok[l] = narrow ? 1u : 2u;
}

// the cell words of the coded levels; a key missing from its set (it cannot be, the
// collection inserted every key of a coded level) clears the level's result
__global__ __launch_bounds__(256) void code_encode(TopoView tp, const int32_t* __restrict__ list, int n, Coef C,
                                                   uint32_t* __restrict__ ws, uint32_t* __restrict__ W)
{
    const uint32_t* keys = ws;
    const uint32_t* slotcode = ws + kMaxLevels * kCodeSlots;
    unsigned int* ok = ws + 2 * kMaxLevels * kCodeSlots + 2 * kMaxLevels;
    CMG_TILE_LOOP(list, n)
    {
        const int l = tp.tlevel[t_];
        const unsigned int form = ok[l];
        if (!form) continue;
        const bool wide = form == 2u;
        const size_t q = quad_index(t_);
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        uint32_t w[4] = {ct.x, ct.y, ct.z, ct.w};
        bool missing = false;
        if (wide) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t t = w[e];
                missing |= t != 0u && t != DIRICHLET && t != INTERIOR && t != NEUMANN;
                w[e] = t == NEUMANN ? 3u : t;
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const uint4 v = reinterpret_cast<const uint4*>(C.c[a])[q];
            const uint32_t key[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                uint32_t sl = code_slot(key[e]);
                int p = 0;
                while (p < kCodeSlots && keys[l * kCodeSlots + sl] != key[e]) {
                    sl = (sl + 1) & (kCodeSlots - 1);
                    ++p;
                }
                if (p == kCodeSlots) missing = true;
                else w[e] |= slotcode[l * kCodeSlots + sl] << (wide ? 10 * a + 2 : 8 * a + 8);
            }
        }
        if (missing) ok[l] = 0u;
        reinterpret_cast<uint4*>(W)[q] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Coef::zw: thread f * 64 + (8 i + j) of a tile's workgroup copies the word of the cell across
// face f (0: -z, the z-neighbour's cell (i, j, 7); 1: +z, its cell (i, j, 0)), or the tile's own
// cell there when it has no z-neighbour (a word the stencil kernels then do not use)
__global__ __launch_bounds__(128) void code_zfaces(TopoView tp, const int32_t* __restrict__ list, int n,
                                                   const uint32_t* __restrict__ W, uint32_t* __restrict__ ZW)
{
    if (int(blockIdx.x) >= n) return;
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const int f = threadIdx.x >> 6, fi = threadIdx.x & 63;
    const int nt = tp.halo[size_t(t) * 8 + (f ? 5 : 2)];
    const int src = nt >= 0 ? nt : t;
    ZW[size_t(t) * 128 + threadIdx.x] = W[size_t(src) * 512 + cidx(fi >> 3, fi & 7, f ? 0 : 7)];
}

__global__ __launch_bounds__(256) void copy_v(const int32_t* __restrict__ list, int n, const float* __restrict__ A,
                                              float* __restrict__ B, const double* __restrict__ S)
{
    if (stopped(S)) return;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        reinterpret_cast<float4*>(B)[q] = reinterpret_cast<const float4*>(A)[q];
    }
}

__global__ __launch_bounds__(256) void fill_v(const int32_t* __restrict__ list, int n, float* __restrict__ X, float value,
                                              const double* __restrict__ S)
{
    if (stopped(S)) return;
    CMG_TILE_LOOP(list, n)
    {
        reinterpret_cast<float4*>(X)[quad_index(t_)] = make_float4(value, value, value, value);
    }
}

// FASProlongateAndUpdateAMG128Kernel (src/AMGSolver.cu:1216-1255), four cells per thread
// meta: the prolongation list's (tile, parent, octant | type << 8) records
__global__ __launch_bounds__(256) void fas_prolong_v(TopoView tp, const int4* __restrict__ meta, int n, float* __restrict__ X,
                                                     const float* __restrict__ X0, float P, const double* __restrict__ S)
{
    if (stopped(S)) return;
    for (int w_ = blockIdx.x; 16 * (w_ >> 3) < n; w_ += gridDim.x)   // the pairing of CMG_TILE_LOOP
    {
        const int pos = 16 * (w_ >> 3) + (w_ & 7) + 8 * int(threadIdx.x >> 7);
        if (pos >= n) continue;
        const int4 m = meta[pos];
        const int t_ = m.x;
        if (t_ < 0) continue;
        const size_t q = quad_index(t_);
        const int v0 = int(threadIdx.x & 127) * 4;   // cells v0..v0+3 share i, j; k = 0..3 or 4..7
        const int pt = m.y;
        const int oc = m.z & 7;
        const bool ghost = (m.z >> 8) & GHOST;
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        float4 x = reinterpret_cast<const float4*>(X)[q];
        const int i = v0 >> 6, j = (v0 >> 3) & 7, k0 = v0 & 7;
        const int pbase = cidx(4 * (oc & 1) + (i >> 1), 4 * ((oc >> 1) & 1) + (j >> 1), 4 * ((oc >> 2) & 1) + (k0 >> 1));
        float cx[2] = {0.0f, 0.0f}, cx0[2] = {0.0f, 0.0f};
        if (pt >= 0) {
            // pbase is even (k0 / 2 is 0 or 2), so the two parent cells are one aligned float2
            const size_t pb = size_t(pt) * 512 + pbase;
            const float2 a = *reinterpret_cast<const float2*>(X + pb), b = *reinterpret_cast<const float2*>(X0 + pb);
            cx[0] = a.x; cx[1] = a.y;
            cx0[0] = b.x; cx0[1] = b.y;
        }
        const uint8_t cts[4] = {ct.x, ct.y, ct.z, ct.w};
        float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const bool in = cts[e] & INTERIOR;
            const float c1 = cx[e >> 1], c0 = cx0[e >> 1];
            xs[e] = ghost ? (in ? c1 : 0.0f) : (in ? xs[e] + P * (c1 - c0) : 0.0f);
        }
        reinterpret_cast<float4*>(X)[q] = make_float4(xs[0], xs[1], xs[2], xs[3]);
    }
}

// ------------------------------------------------- one tile per workgroup ----
// y = A x on LEAF tiles, one tile per workgroup (AMGFullNegativeLaplacianOnLeafs); optional x.y partial
// K: tiles of the levels with one-byte codes (bit set in C.coded, clear in C.wide) read the
// coded coefficients (the same values)
template <bool DOT, int NT, bool K = false>
__global__ __launch_bounds__(NT) void apply_leaf(TopoView tp, const int4* __restrict__ rec, const float* __restrict__ X,
                                                 Coef C, float* __restrict__ Y, double* __restrict__ partial,
                                                 const double* __restrict__ S)
{
    __shared__ Lds<float, K ? 3 : 0> s;
    __shared__ double red[8];
    const TileRec r = rec_tile<true>(rec, S);
    const int t = r.t;   // -1: no tile, -2: the solve has stopped
    if (t < 0) {
        if (DOT && t == -1 && threadIdx.x == 0) partial[blockIdx.x] = 0;
        return;
    }
    Own<float, NT> o;
    if (K && (((C.coded & ~C.wide) >> r.level) & 1u))
        load_tile<true, float, NT, K ? 3 : 0>(tp, r, rec_h(tp, r), X, X, C.c, s, o, nullptr, NoExtraLoads{},
                                             Pull<0>{nullptr, 0.0f}, C.w, C.lut, C.zw);
    else
        load_tile<true, float, NT, 0>(tp, r, rec_h(tp, r), X, X, C.c, s, o);
    constexpr int CPT = 512 / NT;
    const size_t v0 = size_t(t) * 512 + size_t(CPT) * threadIdx.x;
    float xr[CPT], y[CPT];
    if (DOT) vload<CPT>(X + v0, xr);
    double acc = 0;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int v = own_cell<NT>(c);
        y[c] = o.ct[c] == INTERIOR ? neg_lap(s, v, o.x[c], o.c[c], o.d[c]) : 0.0f;
        if (DOT) acc += double(xr[c]) * double(y[c]);
    }
    vstore<CPT>(Y + v0, y);
    if (DOT) {
        double b = block_sum(acc, red);
        if (threadIdx.x == 0) partial[blockIdx.x] = b;
    }
}

// per tile: the number of cells whose stored diagonal is not bitwise the derived one
template <int NT>
__global__ __launch_bounds__(NT) void diag_check(TopoView tp, const int4* __restrict__ rec, Coef C,
                                                 double* __restrict__ partial)
{
    __shared__ Lds<float> s;
    __shared__ double red[8];
    const TileRec r = rec_tile<true, false>(rec, nullptr);
    double bad = 0;
    if (r.t >= 0) {
        Own<float, NT> o;
        tp.derive = 0;
        const float h = rec_h(tp, r);
        load_tile<true>(tp, r, h, C.c[0], C.c[0], C.c, s, o);
#pragma unroll
        for (int c = 0; c < 512 / NT; ++c) {
            const int v = own_cell<NT>(c);
            const float d = (o.ct[c] & INTERIOR) ? derived_diag(s, v, o.c[c], h) : 0.0f;
            bad += __float_as_uint(d) != __float_as_uint(o.d[c]) ? 1.0 : 0.0;
        }
    } else {
        __syncthreads();   // matches the barriers of load_tile
        __syncthreads();
    }
    const double b = block_sum(bad, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// Learned pre-smoothing step: the net's correction Y of a LEAF tile is kept when
// |b - A Y| < |b| on the tile's INTERIOR cells (x = 0 before the step, the same-
// level neighbours' Y in the halo, ghosts 0); otherwise the tile falls back to the
// plain first damped-Jacobi sweep x = w b / D.  NONLEAF tiles of the level take
// the Jacobi sweep.
template <int NT>
__global__ __launch_bounds__(NT) void learned_accept(TopoView tp, const int4* __restrict__ rec, const float* __restrict__ Y,
                                                     float* __restrict__ X, const float* __restrict__ Gz, Coef C,
                                                     const float* __restrict__ B, float w, const double* __restrict__ S,
                                                     unsigned int* __restrict__ count)
{
    __shared__ Lds<float> s;
    __shared__ double red[8];
    const TileRec r = rec_tile<false>(rec, S);
    const int t = r.t;   // < 0: no tile, or the solve has stopped
    if (t < 0) return;
    Own<float, NT> o;
    constexpr int CPT = 512 / NT;
    const size_t v0 = size_t(t) * 512 + size_t(CPT) * threadIdx.x;
    float b[CPT];
    vload<CPT>(B + v0, b);
    load_tile<false>(tp, r, rec_h(tp, r), Y, Gz, C.c, s, o);
    double rr = 0, bb = 0;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int v = own_cell<NT>(c);
        if (!(o.ct[c] & INTERIOR)) continue;
        const float r = b[c] - neg_lap(s, v, o.x[c], o.c[c], o.d[c]);
        rr += double(r) * double(r);
        bb += double(b[c]) * double(b[c]);
    }
    const double rs = block_sum(rr, red);
    __syncthreads();
    const double bs = block_sum(bb, red);
    __shared__ int keep;
    if (threadIdx.x == 0) {
        const bool leaf = r.meta & LEAF;
        keep = leaf && rs < bs;
        // tiles judged and tiles that kept the net's correction
        if (count && leaf) {
            atomicAdd(&count[0], 1u);
            if (keep) atomicAdd(&count[1], 1u);
        }
    }
    __syncthreads();
    float yv[CPT], xo[CPT];
    vload<CPT>(Y + v0, yv);
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const float jac = 0.0f + w * ((o.ct[c] & INTERIOR) ? (b[c] - 0.0f) / o.d[c] : 0.0f);
        xo[c] = keep ? ((o.ct[c] & INTERIOR) ? yv[c] : 0.0f) : jac;
    }
    vstore<CPT>(X + v0, xo);
}

// damped Jacobi sweep x_out = x_in + w (b - A x_in) / D on subtree-LEAF tiles of a level;
// DOT: also the tile's sum of b x_out (FP64, all cells) into partial[blockIdx.x]
// SAVE: the first sweep of a coarse level right after the restriction also does
// that level's FAS save on its NONLEAF tiles (B += A x on INTERIOR cells, 0
// elsewhere; X0 = x), with the A x the sweep computes anyway; the sweep then
// uses the updated B.  Same values as fas_save followed by the plain sweep.
// PUSH: the level's last post-sweep also does the FAS prolongation to the next finer
// level (FASProlongateAndUpdateAMG128Kernel, src/AMGSolver.cu:1216-1255) for the
// children of its tile, from the x it has just computed: a child's INTERIOR cells get
// x += P (x_parent - X0_parent) (a GHOST child: x = x_parent), its other cells 0.  Every
// tile has one parent, so each child is written by exactly one workgroup; the sweep
// reads no cell of the finer level.  Same values as the sweep followed by fas_prolong_v.
// EARLY: the children's loads go out with the tile's (else after the sweep, with fewer
// registers held across it)
// PM: the prolongation into this level pulled into the sweep (Pull<PM>, X0 and P read)
template <int NT, bool DOT = false, int Q = 0, bool SAVE = false, bool PUSH = false, bool EARLY = true, int PM = 0>
__device__ __forceinline__ double jacobi_tile(const TopoView& tp, const TileRec& r, Lds<float, Q>& s,
                                              const float* __restrict__ Xin, float* __restrict__ Xout,
                                              const float* __restrict__ G, const Coef& C, const float* __restrict__ B,
                                              float w, float* __restrict__ X0, float P, bool act)
{
    static_assert(!(SAVE && PUSH), "the FAS save rides on a first sweep, the prolongation on a last one");
    const int t = r.t, tid = r.tid;
    Own<float, NT> o;
    constexpr int CPT = 512 / NT;
    const size_t v0 = size_t(t) * 512 + size_t(CPT) * tid;
    float b[CPT], xo[CPT];
    vload<CPT>(B + v0, b);
    const float h = rec_h(tp, r);
    // PUSH: a thread's cells run along k; G4 of them lie in one child octant and cover
    // 2 x 2 rows of 2 G4 consecutive cells of that child.  The children's x and types
    // are loaded together with the tile's own loads (a missing child: the tile's own
    // cells, unused).
    constexpr int G4 = CPT < 4 ? CPT : 4;
    constexpr int NG = PUSH ? CPT / G4 : 1;
    constexpr int PR = PUSH ? 2 * G4 : 1;
    float pxf[NG][4][PR];
    uint8_t pcf[NG][4][PR];
    float px0[PUSH ? CPT : 1];
    int pch[NG];
    uint8_t pty[NG];
    auto push_loads = [&]() {
        if constexpr (PUSH && EARLY) {
            vload<CPT>(X0 + v0, px0);   // the FAS-saved x on NONLEAF tiles (unused for GHOST children)
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int v = own_cell<NT>(g * G4, tid);
                const int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
                const int ch = pick8(r.ca, r.cb, (i >> 2) | ((j >> 2) << 1) | ((k >> 2) << 2));
                const int cs = ch >= 0 ? ch : t;
                pch[g] = ch;
                pty[g] = tp.ttype[cs];
#pragma unroll
                for (int ab = 0; ab < 4; ++ab) {
                    const size_t fb = size_t(cs) * 512 + cidx(2 * (i & 3) + (ab >> 1), 2 * (j & 3) + (ab & 1), 2 * (k & 3));
                    vload<PR>(Xout + fb, pxf[g][ab]);
                    vload<PR>(tp.ctype + fb, pcf[g][ab]);
                }
            }
        }
    };
    static_assert(!(PM && (SAVE || PUSH)), "a pulled prolongation is a post-sweep of the finest level");
    load_tile<false, float, NT, Q>(tp, r, h, Xin, G, C.c, s, o, C.q, push_loads, Pull<PM>{X0, P}, C.w, C.lut, C.zw);
    const bool save = SAVE && (r.meta & NONLEAF);
    const bool dot_tile = DOT && (r.meta & LEAF);   // b.x is r.z on LEAF tiles only
    double acc = 0;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int v = own_cell<NT>(c, tid);
        const float ax = neg_lap(s, v, o.x[c], o.c[c], o.d[c], h);
        if (save) {
            const size_t idx = size_t(t) * 512 + v;
            if (o.ct[c] & INTERIOR) {
                b[c] = b[c] + ax;
                if (act) X0[idx] = o.x[c];
            } else {
                b[c] = 0.0f;
            }
        }
        const float dq = (b[c] - ax) / o.d[c];   // masked below (non-INTERIOR cells: d = 0)
        const float delta = (o.ct[c] & INTERIOR) ? dq : 0.0f;
        xo[c] = o.x[c] + w * delta;
        if (dot_tile) acc += double(b[c]) * double(xo[c]);
    }
    // B is the level's right-hand side, updated in place (one vector a thread)
    if (save && act) vstore<CPT>(const_cast<float*>(B) + v0, b);
    if (act) vstore<CPT>(Xout + v0, xo);
    if constexpr (PUSH) {
        if constexpr (!EARLY) {
            vload<CPT>(X0 + v0, px0);
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int v = own_cell<NT>(g * G4, tid);
                const int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
                const int ch = pick8(r.ca, r.cb, (i >> 2) | ((j >> 2) << 1) | ((k >> 2) << 2));
                const int cs = ch >= 0 ? ch : t;
                pch[g] = ch;
                pty[g] = tp.ttype[cs];
#pragma unroll
                for (int ab = 0; ab < 4; ++ab) {
                    const size_t fb = size_t(cs) * 512 + cidx(2 * (i & 3) + (ab >> 1), 2 * (j & 3) + (ab & 1), 2 * (k & 3));
                    vload<PR>(Xout + fb, pxf[g][ab]);
                    vload<PR>(tp.ctype + fb, pcf[g][ab]);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            if (pch[g] < 0 || !act) continue;
            const bool ghost = pty[g] & GHOST;
            const int v = own_cell<NT>(g * G4, tid);
            const int i = v >> 6, j = (v >> 3) & 7, k = v & 7;
#pragma unroll
            for (int ab = 0; ab < 4; ++ab) {
#pragma unroll
                for (int e = 0; e < PR; ++e) {
                    const bool in = pcf[g][ab][e] & INTERIOR;
                    const float c1 = xo[g * G4 + (e >> 1)], c0 = px0[g * G4 + (e >> 1)];
                    pxf[g][ab][e] = ghost ? (in ? c1 : 0.0f) : (in ? pxf[g][ab][e] + P * (c1 - c0) : 0.0f);
                }
                const size_t fb = size_t(pch[g]) * 512 + cidx(2 * (i & 3) + (ab >> 1), 2 * (j & 3) + (ab & 1), 2 * (k & 3));
                vstore<PR>(Xout + fb, pxf[g][ab]);
            }
        }
    }
    return acc;
}

template <int NT, bool DOT = false, int Q = 0, bool SAVE = false, bool PUSH = false, int PM = 0>
__global__ __launch_bounds__(NT) void jacobi(TopoView tp, const int4* __restrict__ rec, const float* __restrict__ Xin,
                                             float* __restrict__ Xout, const float* __restrict__ G, Coef C,
                                             const float* __restrict__ B, float w, const double* __restrict__ S,
                                             double* __restrict__ partial, float* __restrict__ X0 = nullptr,
                                             float P = 0.0f)
{
    __shared__ Lds<float, Q> s;
    __shared__ double red[DOT ? 8 : 1];
    const TileRec r = rec_tile<PM != 0, true, PUSH>(rec, S);
    if (r.t < 0) {   // -1: no tile, -2: the solve has stopped
        if (DOT && r.t == -1 && threadIdx.x == 0) partial[blockIdx.x] = 0;
        return;
    }
    const double acc = jacobi_tile<NT, DOT, Q, SAVE, PUSH, true, PM>(tp, r, s, Xin, Xout, G, C, B, w, X0, P, true);
    if (DOT) {
        const double sum = block_sum(acc, red);
        if (threadIdx.x == 0) partial[blockIdx.x] = sum;
    }
}

// ResidualAndRestrictAMG128Kernel (src/AMGSolver.cu:896-964): parent b = R*sum8 r, parent x = sum8 x / 8
template <int NT, int Q = 0>
__device__ __forceinline__ void residual_restrict_tile(const TopoView& tp, const TileRec& r, Lds<float, Q>& s,
                                                       float* __restrict__ res, float* __restrict__ X, const Coef& C,
                                                       float* __restrict__ B, float Rr, bool act)
{
    const int t = r.t, tid = r.tid;
    Own<float, NT> o;
    constexpr int CPT = 512 / NT;
    float b[CPT], rv[CPT];
    vload<CPT>(B + size_t(t) * 512 + size_t(CPT) * tid, b);
    const float h = rec_h(tp, r);
    load_tile<false, float, NT, Q>(tp, r, h, X, X, C.c, s, o, C.q, NoExtraLoads{}, Pull<0>{nullptr, 0.0f}, C.w, C.lut, C.zw);
#pragma unroll
    for (int c = 0; c < CPT; ++c) rv[c] = b[c] - neg_lap(s, own_cell<NT>(c, tid), o.x[c], o.c[c], o.d[c], h);
    vstore<CPT>(&res[CPT * tid], rv);
    __syncthreads();
    const int v = tid;
    const int pt = r.parent;
    if (v >= 64 || pt < 0 || !act) return;
    const int oc = (r.meta >> 8) & 7;
    const int fl0 = v / 16 * 2, fl1 = (v / 4) % 4 * 2, fl2 = v % 4 * 2;
    float rs = 0, xs = 0;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                rs += res[cidx(fl0 + ii, fl1 + jj, fl2 + kk)];
                xs += s.x[SX(fl0 + ii, fl1 + jj, fl2 + kk)];
            }
    const size_t pidx = size_t(pt) * 512 + cidx(4 * (oc & 1) + (fl0 >> 1), 4 * ((oc >> 1) & 1) + (fl1 >> 1), 4 * ((oc >> 2) & 1) + (fl2 >> 1));
    B[pidx] = rs * Rr;
    X[pidx] = xs / 8;
}

template <int NT, int Q = 0>
__global__ __launch_bounds__(NT) void residual_restrict(TopoView tp, const int4* __restrict__ rec, float* __restrict__ X,
                                                        Coef C, float* __restrict__ B, float Rr,
                                                        const double* __restrict__ S)
{
    __shared__ Lds<float, Q> s;
    __shared__ float res[512];
    const TileRec r = rec_tile<false>(rec, S);
    if (r.t < 0) return;   // no tile, or the solve has stopped
    residual_restrict_tile<NT, Q>(tp, r, s, res, X, C, B, Rr, true);
}

// r64 = b64 - A x64 on LEAF tiles in FP64 (FRESH halo), r32 = float(r64), partial |r64|^2
template <int NT>
__global__ __launch_bounds__(NT) void residual64(TopoView tp, const int4* __restrict__ rec, const double* __restrict__ X,
                                                 Coef C, const double* __restrict__ B, double* __restrict__ R64,
                                                 float* __restrict__ R32, double* __restrict__ partial)
{
    __shared__ Lds<double> s;
    __shared__ double red[8];
    const TileRec r = rec_tile<true, false>(rec, nullptr);
    const int t = r.t;
    if (t < 0) {
        if (threadIdx.x == 0) partial[blockIdx.x] = 0;
        return;
    }
    Own<double, NT> o;
    load_tile<true>(tp, r, rec_h(tp, r), X, X, C.c, s, o);
    constexpr int CPT = 512 / NT;
    const size_t v0 = size_t(t) * 512 + size_t(CPT) * threadIdx.x;
    double bv[CPT], rv[CPT];
    float r32[CPT];
    vload<CPT>(B + v0, bv);
    double acc = 0;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        rv[c] = 0;
        if (o.ct[c] == INTERIOR) rv[c] = bv[c] - neg_lap(s, own_cell<NT>(c), o.x[c], o.c[c], o.d[c]);
        r32[c] = float(rv[c]);
        acc += rv[c] * rv[c];
    }
    vstore<CPT>(R64 + v0, rv);
    vstore<CPT>(R32 + v0, r32);
    double b = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// ------------------------------------------------------------- smoothers ----
// one colour of red-black SOR on subtree-LEAF tiles of a level, in place
// (NegativeLaplacianAndGaussSeidelSameLevelAMG128Kernel, src/AMGSolver.cu:838-877)
template <int NT, int Q = 0>
__device__ __forceinline__ void gs_color_tile(const TopoView& tp, const TileRec& r, Lds<float, Q>& s, float* __restrict__ X,
                                              const Coef& C, const float* __restrict__ B, int color, double omega, bool act)
{
    const int t = r.t, tid = r.tid;
    Own<float, NT> o;
    constexpr int CPT = 512 / NT;
    float bv[CPT];
    vload<CPT>(B + size_t(t) * 512 + size_t(CPT) * tid, bv);
    const float h = rec_h(tp, r);
    load_tile<false, float, NT, Q>(tp, r, h, X, X, C.c, s, o, C.q, NoExtraLoads{}, Pull<0>{nullptr, 0.0f}, C.w, C.lut, C.zw);
    if (!act) return;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int v = own_cell<NT>(c, tid);
        if ((((v >> 6) + ((v >> 3) & 7) + (v & 7)) & 1) != color) continue;
        const size_t idx = size_t(t) * 512 + v;
        const float b = bv[c];
        const float ax = neg_lap(s, v, o.x[c], o.c[c], o.d[c], h);
        const float delta = (o.ct[c] & INTERIOR) ? (b - ax) / o.d[c] : 0.0f;
        X[idx] = float(double(o.x[c]) + double(delta) * omega);
    }
}

template <int NT, int Q = 0>
__global__ __launch_bounds__(NT) void gs_color(TopoView tp, const int4* __restrict__ rec, float* __restrict__ X, Coef C,
                                               const float* __restrict__ B, int color, double omega,
                                               const double* __restrict__ S)
{
    __shared__ Lds<float, Q> s;
    const TileRec r = rec_tile<false>(rec, S);
    if (r.t < 0) return;   // no tile, or the solve has stopped
    gs_color_tile<NT, Q>(tp, r, s, X, C, B, color, omega, true);
}

// The coarsest level's smoothing iterations on a level 0 of one tile without neighbour
// tiles (the root), in one workgroup: the tile is loaded once and every pass runs on its
// LDS copy (its halo is 0 throughout).  rbgs: `iters` iterations of colour `order` then
// 1 - order, followed by `iters2` iterations starting with 1 - order (coarsest(): half
// the coarsest iterations each way, src/AMGSolver.cu:1257-1263 with the red-black SOR of
// :838-877); else `iters` damped-Jacobi sweeps.  Same values as one launch per pass of
// gs_color or jacobi (each pass re-reads x masked to INTERIOR cells, which is what the
// LDS copy holds).
template <int NT, int Q = 0>
__device__ __forceinline__ void coarsest_tile(const TopoView& tp, const TileRec& r, Lds<float, Q>& s, float* __restrict__ X,
                                              const Coef& C, const float* __restrict__ B, int rbgs, int iters, int iters2,
                                              int order, double omega, float w, bool act)
{
    const int t = r.t, tid = r.tid;
    Own<float, NT> o;
    constexpr int CPT = 512 / NT;
    const size_t v0 = size_t(t) * 512 + size_t(CPT) * tid;
    float bv[CPT], xn[CPT];
    vload<CPT>(B + v0, bv);
    const float h = rec_h(tp, r);
    load_tile<false, float, NT, Q>(tp, r, h, X, X, C.c, s, o, C.q, NoExtraLoads{}, Pull<0>{nullptr, 0.0f}, C.w, C.lut, C.zw);
    const int passes = rbgs ? 2 * (iters + iters2) : iters;
    if (passes == 0) return;
    for (int p = 0; p < passes; ++p) {
        // colour of this pass: order, 1 - order, ... then from 1 - order after 2 iters passes
        const int color = rbgs ? ((p < 2 * iters ? order : 1 - order) ^ (p & 1)) : -1;
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const int v = own_cell<NT>(c, tid);
            if (rbgs && (((v >> 6) + ((v >> 3) & 7) + (v & 7)) & 1) != color) continue;
            const float ax = neg_lap(s, v, o.x[c], o.c[c], o.d[c], h);
            const float delta = (o.ct[c] & INTERIOR) ? (bv[c] - ax) / o.d[c] : 0.0f;
            xn[c] = rbgs ? float(double(o.x[c]) + double(delta) * omega) : o.x[c] + w * delta;
        }
        if (!rbgs) __syncthreads();   // every read of this sweep is done before x changes
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const int v = own_cell<NT>(c, tid);
            if (rbgs && (((v >> 6) + ((v >> 3) & 7) + (v & 7)) & 1) != color) continue;
            o.x[c] = o.ct[c] == INTERIOR ? xn[c] : 0.0f;
            s.x[SX(v >> 6, (v >> 3) & 7, v & 7)] = o.x[c];
        }
        __syncthreads();
    }
    // every cell was written by a pass (rbgs: both colours ran at least once)
    if (act) vstore<CPT>(X + v0, xn);
}

template <int NT, int Q = 0>
__global__ __launch_bounds__(NT) void coarsest_iterate(TopoView tp, const int4* __restrict__ rec, float* __restrict__ X,
                                                       Coef C, const float* __restrict__ B, int rbgs, int iters,
                                                       int iters2, int order, double omega, float w,
                                                       const double* __restrict__ S)
{
    __shared__ Lds<float, Q> s;
    const TileRec r = rec_tile<false>(rec, S);
    if (r.t < 0) return;   // no tile, or the solve has stopped
    coarsest_tile<NT, Q>(tp, r, s, X, C, B, rbgs, iters, iters2, order, omega, w, true);
}

// ------------------------------------------------------------- transfers ----
// FASSaveSolutionAndUpdateSource128Kernel (src/AMGSolver.cu:1176-1214) on NONLEAF tiles
template <int NT, int Q = 0>
__device__ __forceinline__ void fas_save_tile(const TopoView& tp, const TileRec& r, Lds<float, Q>& s,
                                              const float* __restrict__ X, const Coef& C, float* __restrict__ B,
                                              float* __restrict__ X0, bool act)
{
    const int t = r.t, tid = r.tid;
    Own<float, NT> o;
    constexpr int CPT = 512 / NT;
    const size_t v0 = size_t(t) * 512 + size_t(CPT) * tid;
    float bv[CPT];
    vload<CPT>(B + v0, bv);
    const float h = rec_h(tp, r);
    load_tile<false, float, NT, Q>(tp, r, h, X, X, C.c, s, o, C.q, NoExtraLoads{}, Pull<0>{nullptr, 0.0f}, C.w, C.lut, C.zw);
    if (!act) return;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        if (o.ct[c] & INTERIOR) {
            bv[c] = bv[c] + neg_lap(s, own_cell<NT>(c, tid), o.x[c], o.c[c], o.d[c], h);
            X0[v0 + c] = o.x[c];   // x itself on INTERIOR cells
        } else {
            bv[c] = 0;
        }
    }
    vstore<CPT>(B + v0, bv);
}

template <int NT, int Q = 0>
__global__ __launch_bounds__(NT) void fas_save(TopoView tp, const int4* __restrict__ rec, const float* __restrict__ X,
                                               Coef C, float* __restrict__ B, float* __restrict__ X0,
                                               const double* __restrict__ S)
{
    __shared__ Lds<float, Q> s;
    const TileRec r = rec_tile<false>(rec, S);
    if (r.t < 0) return;   // no tile, or the solve has stopped
    fas_save_tile<NT, Q>(tp, r, s, X, C, B, X0, true);
}

// a record read at a given address (not from the launch list position), thread index tid
__device__ __forceinline__ TileRec rec_tile_at(const int4* __restrict__ r, int tid)
{
    TileRec x;
    const int4 r0 = r[0];
    x.t = r0.x;
    x.level = r0.y;
    x.meta = r0.z;
    x.parent = r0.w;
    x.ha = r[1];
    x.hb = r[2];
    x.ga = make_int4(-1, -1, -1, -1);
    x.gb = make_int4(-1, -1, 0, 0);
    x.ca = r[5];
    x.cb = r[6];
    x.tid = tid;
    return x;
}

// The bottom of a FAS mu-cycle, FASMuCycleStep at level 1 with everything below it
// (src/AMGSolver.cu:1257-1277), in one workgroup of 1024 threads, for a level 0 of one
// tile without neighbours (the root) whose 8 children are the whole of level 1.  Level 1
// is then a dense 16^3 block: its x (masked to INTERIOR cells, with a zero halo at the
// domain boundary) and face coefficients stay in LDS for every pass, so a pass costs LDS
// reads and a barrier instead of a launch and a memory round trip.  Group g (128 threads,
// 4 cells each) holds the level-1 tile of octant g.  Every step does the arithmetic of
// the per-pass kernel it replaces (gs_color, jacobi, residual_restrict, fas_save,
// fas_prolong_v, jacobi's r.z partial and prolongation into level 2), so the results
// are bit-identical; the root's steps (FAS save, coarsest iterations) are the per-tile
// functions themselves, on 512 threads (the other 512 redo them without writing).
// phase 1: pre-smoothing, residual + restriction, the root's FAS save; phase 2: the
// prolongation to level 1 and its post-smoothing; phase 3: both with the coarsest
// iterations (mu times) in between.
struct BottomArgs {
    const int4* rec0;    // the root's record
    int root;            // the root tile
    int tile[8];         // the level-1 tile of each octant, its type, its list entry (r.z partials)
    int type[8];
    int entry[8];
    int child[64];       // their children (level 2) by octant, -1 none; bit c of ghost: child c is GHOST
    unsigned long long ghost;
    int nonleaf0;
    float *X, *X0, *B;
    Coef C;
    float Rr, P, w;
    double omega;
    int rbgs, iters, mu, coarsest_iters;
    int phase, fas_save1, push;
    double* rz;          // r.z partials of level 1's last post-sweep (Jacobi), or null
};

template <bool Q>
struct BottomLds {
    float x[18 * 18 * 18];                                    // level 1, masked, zero halo
    std::conditional_t<Q, uint16_t, float> c[3][4096];        // negative-face coefficients
    float res[4096];                                          // residuals for the restriction
    Lds<float, Q> root[2];                                    // the root's steps, two copies
    double red[16];
};

__device__ __forceinline__ int bx(int i, int j, int k) { return (i + 1) * 324 + (j + 1) * 18 + (k + 1); }
__device__ __forceinline__ int bc(int i, int j, int k) { return (i << 8) | (j << 4) | k; }

template <bool Q>
__device__ __forceinline__ float bottom_lap(const BottomLds<Q>& s, int i, int j, int k, float x0, const float c[3], float d,
                                            float h)
{
    if (d == 0) return 0.0f;
    // the + face coefficient is the + neighbour's own; h where the level ends (its x is 0 there)
    const float cp0 = i < 15 ? lds_coef(s.c[0][bc(i + 1, j, k)], h) : h;
    const float cp1 = j < 15 ? lds_coef(s.c[1][bc(i, j + 1, k)], h) : h;
    const float cp2 = k < 15 ? lds_coef(s.c[2][bc(i, j, k + 1)], h) : h;
    float r = x0 * d;
    r = fmaf(c[0], s.x[bx(i - 1, j, k)], r);
    r = fmaf(cp0, s.x[bx(i + 1, j, k)], r);
    r = fmaf(c[1], s.x[bx(i, j - 1, k)], r);
    r = fmaf(cp1, s.x[bx(i, j + 1, k)], r);
    r = fmaf(c[2], s.x[bx(i, j, k - 1)], r);
    r = fmaf(cp2, s.x[bx(i, j, k + 1)], r);
    return r;
}

template <bool Q>
__global__ __launch_bounds__(1024) void cycle_bottom(TopoView tp, BottomArgs a, const double* __restrict__ S)
{
    __shared__ BottomLds<Q> s;
    if (S[S_DONE] != 0.0) return;   // the solve has stopped (uniform)
    // the group is uniform over a wave (two waves a group)
    const int g = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 7)), tid = threadIdx.x & 127;
    const int4* r0p = a.rec0;
    const int root = a.root;
    const int t = a.tile[g];   // the level-1 tile of octant g (kernel arguments: no dependent load)
    const int tmeta = a.type[g];
    const float h = tp.h0 / 2.0f;
    const size_t v0 = size_t(t) * 512 + 4 * tid;
    const int li = tid >> 4, lj = (tid >> 1) & 7, lk0 = 4 * (tid & 1);   // the thread's cells in the tile
    const int gi = 8 * (g & 1) + li, gj = 8 * ((g >> 1) & 1) + lj, gk0 = 8 * ((g >> 2) & 1) + lk0;
    float x[4], b[4], cf[4][3], d[4];
    uint8_t ct[4];
    // ---- level 1 into registers and LDS
    {
        float xr[4];
        vload<4>(a.X + v0, xr);
        vload<4>(a.B + v0, b);
        vload<4>(tp.ctype + v0, ct);
        if constexpr (Q) {
            uint16_t q[4][4];
#pragma unroll
            for (int e = 0; e < 4; ++e) vload<4>(a.C.q[e] + v0, q[e]);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    cf[c][e] = float(__builtin_bit_cast(_Float16, q[e][c])) * h;
                    s.c[e][bc(gi, gj, gk0 + c)] = q[e][c];
                }
                d[c] = float(__builtin_bit_cast(_Float16, q[3][c])) * h;
            }
        } else {
            float q[4][4];
            const bool dv = tp.derive && (tmeta & LEAF);
#pragma unroll
            for (int e = 0; e < 3; ++e) vload<4>(a.C.c[e] + v0, q[e]);
            if (!dv) vload<4>(a.C.c[3] + v0, q[3]);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    cf[c][e] = q[e][c];
                    s.c[e][bc(gi, gj, gk0 + c)] = q[e][c];
                }
                d[c] = dv ? 0.0f : q[3][c];
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            x[c] = ct[c] == INTERIOR ? xr[c] : 0.0f;
            s.x[bx(gi, gj, gk0 + c)] = x[c];
        }
        // the zero halo of the 16^3 block
        for (int m = threadIdx.x; m < 18 * 18 * 18; m += 1024) {
            const int hi = m / 324, hj = (m / 18) % 18, hk = m % 18;
            if (hi == 0 || hi == 17 || hj == 0 || hj == 17 || hk == 0 || hk == 17) s.x[m] = 0.0f;
        }
        __syncthreads();
        if constexpr (!Q) {
            if (tp.derive && (tmeta & LEAF)) {
                // derived_diag: d = 0 - c-(x) - c+(x) - ..., c+ = -h where the level ends
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int k = gk0 + c;
                    const float cp0 = gi < 15 ? s.c[0][bc(gi + 1, gj, k)] : -h;
                    const float cp1 = gj < 15 ? s.c[1][bc(gi, gj + 1, k)] : -h;
                    const float cp2 = k < 15 ? s.c[2][bc(gi, gj, k + 1)] : -h;
                    float dd = 0.0f;
                    dd = dd - cf[c][0];
                    dd = dd - cp0;
                    dd = dd - cf[c][1];
                    dd = dd - cp1;
                    dd = dd - cf[c][2];
                    dd = dd - cp2;
                    d[c] = (ct[c] & INTERIOR) ? dd : 0.0f;
                }
            }
        }
    }
    const int n = a.iters;
    // one smoothing pass on level 1: colour `col` of red-black SOR (col >= 0) or a damped-Jacobi
    // sweep (col < 0, `save`: with the FAS save of NONLEAF tiles; `dot`: returns sum b x_out)
    auto pass = [&](int col, bool save, bool dot, float* xo) -> double {
        double acc = 0;
        float ax[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) ax[c] = bottom_lap(s, gi, gj, gk0 + c, x[c], cf[c], d[c], h);
        if (col >= 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (((li + lj + lk0 + c) & 1) != col) continue;
                const float delta = (ct[c] & INTERIOR) ? (b[c] - ax[c]) / d[c] : 0.0f;
                x[c] = float(double(x[c]) + double(delta) * a.omega);
                s.x[bx(gi, gj, gk0 + c)] = x[c];   // 0 on other cells already
            }
            __syncthreads();
            return 0;
        }
        const bool sv = save && (tmeta & NONLEAF);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (sv) {
                if (ct[c] & INTERIOR) {
                    b[c] = b[c] + ax[c];
                    a.X0[v0 + c] = x[c];
                } else {
                    b[c] = 0.0f;
                }
            }
            const float delta = (ct[c] & INTERIOR) ? (b[c] - ax[c]) / d[c] : 0.0f;
            xo[c] = x[c] + a.w * delta;
            if (dot && (tmeta & LEAF)) acc += double(b[c]) * double(xo[c]);
        }
        if (sv) vstore<4>(a.B + v0, b);
        __syncthreads();   // every read of this sweep is done before x changes
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            x[c] = ct[c] == INTERIOR ? xo[c] : 0.0f;
            s.x[bx(gi, gj, gk0 + c)] = x[c];
        }
        __syncthreads();
        return acc;
    };
    float xo[4];
    if (a.phase & 1) {
        // pre-smoothing (order 0)
        for (int it = 0; it < n; ++it) {
            if (a.rbgs) {
                pass(0, false, false, xo);
                pass(1, false, false, xo);
            } else {
                pass(-1, a.fas_save1 && it == 0, false, xo);
            }
        }
        // residual and restriction into the root (residual_restrict)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            s.res[g * 512 + 4 * tid + c] = b[c] - bottom_lap(s, gi, gj, gk0 + c, x[c], cf[c], d[c], h);
        __syncthreads();
        if (tid < 64) {
            const int fl0 = tid / 16 * 2, fl1 = (tid / 4) % 4 * 2, fl2 = tid % 4 * 2;
            float rs = 0, xs = 0;
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                    for (int kk = 0; kk < 2; ++kk) {
                        rs += s.res[g * 512 + cidx(fl0 + ii, fl1 + jj, fl2 + kk)];
                        xs += s.x[bx(8 * (g & 1) + fl0 + ii, 8 * ((g >> 1) & 1) + fl1 + jj, 8 * ((g >> 2) & 1) + fl2 + kk)];
                    }
            const size_t pidx = size_t(root) * 512 + cidx(4 * (g & 1) + (fl0 >> 1), 4 * ((g >> 1) & 1) + (fl1 >> 1),
                                                         4 * ((g >> 2) & 1) + (fl2 >> 1));
            a.B[pidx] = rs * a.Rr;
            a.X[pidx] = xs / 8;
        }
        // level 1's x after the pre-smoothing (read back by phase 2 when the coarsest solve
        // runs in between; the jacobi ping-pong buffer is not written)
        vstore<4>(a.X + v0, x);
        __syncthreads();
        if (a.nonleaf0) {
            TileRec r = rec_tile_at(r0p, int(threadIdx.x & 511));
            fas_save_tile<512, Q>(tp, r, s.root[threadIdx.x >> 9], a.X, a.C, a.B, a.X0, threadIdx.x < 512);
            __syncthreads();
        }
    }
    if (a.phase == 3) {
        const int it1 = a.rbgs ? a.coarsest_iters / 2 : a.coarsest_iters;
        for (int i = 0; i < a.mu; ++i) {
            TileRec r = rec_tile_at(r0p, int(threadIdx.x & 511));
            coarsest_tile<512, Q>(tp, r, s.root[threadIdx.x >> 9], a.X, a.C, a.B, a.rbgs, it1, a.rbgs ? it1 : 0, 0,
                                  a.omega, a.w, threadIdx.x < 512);
            __syncthreads();
        }
    }
    if (a.phase & 2) {
        // the prolongation to level 1 (fas_prolong_v; no GHOST tile on a complete level 1)
        {
            const size_t pb = size_t(root) * 512 + cidx(gi >> 1, gj >> 1, gk0 >> 1);
            const float2 xc = *reinterpret_cast<const float2*>(a.X + pb), x0c = *reinterpret_cast<const float2*>(a.X0 + pb);
            const float cx[2] = {xc.x, xc.y}, cx0[2] = {x0c.x, x0c.y};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                x[c] = (ct[c] & INTERIOR) ? x[c] + a.P * (cx[c >> 1] - cx0[c >> 1]) : 0.0f;
                s.x[bx(gi, gj, gk0 + c)] = x[c];
            }
            __syncthreads();
        }
        // post-smoothing (order 1); the last Jacobi sweep sums b.x = r.z and prolongs into level 2,
        // whose x and types are loaded before that sweep
        const int oc = (li >> 2) | ((lj >> 2) << 1) | ((lk0 >> 2) << 2);
        const int ch = a.push ? a.child[8 * g + oc] : -1;
        float pxf[4][8];
        uint8_t pcf[4][8];
        float px0[4];
        double acc = 0;
        for (int it = 0; it < n; ++it) {
            if (a.rbgs) {
                pass(1, false, false, xo);
                pass(0, false, false, xo);
            } else {
                if (it == n - 1 && ch >= 0) {
                    vload<4>(a.X0 + v0, px0);
#pragma unroll
                    for (int ab = 0; ab < 4; ++ab) {
                        const size_t fb = size_t(ch) * 512 + cidx(2 * (li & 3) + (ab >> 1), 2 * (lj & 3) + (ab & 1), 0);
                        vload<8>(a.X + fb, pxf[ab]);
                        vload<8>(tp.ctype + fb, pcf[ab]);
                    }
                }
                acc = pass(-1, false, a.rz && it == n - 1, xo);
            }
        }
        vstore<4>(a.X + v0, x);
        if (a.rz) {
            // the tile's b.x partial, summed as block_sum sums a 128-thread workgroup's
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if ((threadIdx.x & 63) == 0) s.red[threadIdx.x >> 6] = acc;
            __syncthreads();
            if (tid == 0) {
                double sum = 0;
                sum += s.red[2 * g];
                sum += s.red[2 * g + 1];
                a.rz[a.entry[g]] = sum;
            }
        }
        if (ch >= 0) {
            // jacobi's prolongation into the children (level 2) from the last sweep's x
            const bool ghost = (a.ghost >> (8 * g + oc)) & 1ull;
#pragma unroll
            for (int ab = 0; ab < 4; ++ab) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const bool in = pcf[ab][e] & INTERIOR;
                    const float c1 = xo[e >> 1], c0 = px0[e >> 1];
                    pxf[ab][e] = ghost ? (in ? c1 : 0.0f) : (in ? pxf[ab][e] + a.P * (c1 - c0) : 0.0f);
                }
                const size_t fb = size_t(ch) * 512 + cidx(2 * (li & 3) + (ab >> 1), 2 * (lj & 3) + (ab & 1), 0);
                vstore<8>(a.X + fb, pxf[ab]);
            }
        }
    }
}

// PropagateToChildren on GHOST tiles: ghost cell = parent cell (src/PoissonGrid.cu:377-405)
__global__ __launch_bounds__(512) void propagate_ghosts(TopoView tp, const int32_t* __restrict__ list, float* __restrict__ X)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const int v = threadIdx.x;
    const int pt = tp.parent[t];
    const int o = tp.toct[t];
    const int pv = cidx(4 * (o & 1) + ((v >> 6) >> 1), 4 * ((o >> 1) & 1) + (((v >> 3) & 7) >> 1), 4 * ((o >> 2) & 1) + ((v & 7) >> 1));
    X[size_t(t) * 512 + v] = pt < 0 ? 0.0f : X[size_t(pt) * 512 + pv];
}

// AccumulateToParentsOneStep(LEAF, 1/8) (src/PoissonGrid.cu:408-465)
__global__ __launch_bounds__(64) void accumulate_leafs(TopoView tp, const int32_t* __restrict__ list, float* __restrict__ X)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const int pt = tp.parent[t];
    if (pt < 0) return;
    const int v = threadIdx.x, o = tp.toct[t];
    const int fl0 = v / 16 * 2, fl1 = (v / 4) % 4 * 2, fl2 = v % 4 * 2;
    float sum = 0;
    for (int ii = 0; ii < 2; ++ii)
        for (int jj = 0; jj < 2; ++jj)
            for (int kk = 0; kk < 2; ++kk) sum += X[size_t(t) * 512 + cidx(fl0 + ii, fl1 + jj, fl2 + kk)];
    X[size_t(pt) * 512 + cidx(4 * (o & 1) + (fl0 >> 1), 4 * ((o >> 1) & 1) + (fl1 >> 1), 4 * ((o >> 2) & 1) + (fl2 >> 1))] = sum * 0.125f;
}

// ------------------------------------------------------------ PCG vectors ----
__global__ __launch_bounds__(512) void fill_all(float* __restrict__ X, size_t n4, float value, const double* __restrict__ S)
{
    if (stopped(S)) return;
    float4* p = reinterpret_cast<float4*>(X);
    const float4 v4 = make_float4(value, value, value, value);
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += size_t(gridDim.x) * blockDim.x) p[i] = v4;
}

// pure-Neumann recentring (ReCenterLeafCells, src/PoissonGrid.cu:889-921)
__global__ __launch_bounds__(512) void sum_count(TopoView tp, const int32_t* __restrict__ list, const float* __restrict__ X,
                                                 double* __restrict__ psum, double* __restrict__ pcnt,
                                                 const double* __restrict__ S)
{
    __shared__ double red[8];
    const int t = list[blockIdx.x];
    if (stopped(S)) return;
    double v = 0, c = 0;
    if (t >= 0) {
        const size_t idx = size_t(t) * 512 + threadIdx.x;
        if (tp.ctype[idx] & INTERIOR) { v = X[idx]; c = 1; }
    }
    double b = block_sum(v, red);
    __syncthreads();
    double bc = block_sum(c, red);
    if (threadIdx.x == 0) { psum[blockIdx.x] = b; pcnt[blockIdx.x] = bc; }
}

__global__ __launch_bounds__(512) void sub_mean(TopoView tp, const int32_t* __restrict__ list, float* __restrict__ X,
                                                const double* __restrict__ S)
{
    const int t = list_tile(list, S);   // < 0: no tile, or the solve has stopped
    if (t < 0) return;
    const size_t idx = size_t(t) * 512 + threadIdx.x;
    if (tp.ctype[idx] & INTERIOR) X[idx] = float(double(X[idx]) - S[S_MEAN]);
}

// partial[wg] = sum over the tile of a*b in FP64
__global__ __launch_bounds__(512) void dot64_list(const int32_t* __restrict__ list, const double* __restrict__ A,
                                                  const double* __restrict__ B, double* __restrict__ partial)
{
    __shared__ double red[8];
    const int t = list[blockIdx.x];
    double v = 0;
    if (t >= 0) {
        const size_t idx = size_t(t) * 512 + threadIdx.x;
        v = A[idx] * B[idx];
    }
    double b = block_sum(v, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// y = fmaf(a, x, y) on INTERIOR cells of the listed tiles
__global__ __launch_bounds__(512) void axpy_list(TopoView tp, const int32_t* __restrict__ list, float* __restrict__ Y,
                                                 const float* __restrict__ X, float a)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const size_t idx = size_t(t) * 512 + threadIdx.x;
    if (tp.ctype[idx] & INTERIOR) Y[idx] = fmaf(a, X[idx], Y[idx]);
}

// x64 += x32 on INTERIOR cells of LEAF tiles (FP64 solution update)
__global__ __launch_bounds__(512) void update_x64(TopoView tp, const int32_t* __restrict__ list, double* __restrict__ X64,
                                                  const float* __restrict__ E)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const size_t idx = size_t(t) * 512 + threadIdx.x;
    if (tp.ctype[idx] & INTERIOR) X64[idx] += double(E[idx]);
}

__global__ __launch_bounds__(512) void to_double(const int32_t* __restrict__ list, const float* __restrict__ A, double* __restrict__ B,
                                                 TopoView tp, int masked)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const size_t idx = size_t(t) * 512 + threadIdx.x;
    B[idx] = (!masked || (tp.ctype[idx] & INTERIOR)) ? double(A[idx]) : 0.0;
}

__global__ __launch_bounds__(512) void to_float(const int32_t* __restrict__ list, const double* __restrict__ A, float* __restrict__ B)
{
    const int t = list[blockIdx.x];
    if (t < 0) return;
    const size_t idx = size_t(t) * 512 + threadIdx.x;
    B[idx] = float(A[idx]);
}

// ------------------------------------------------------------ reductions ----
// deterministic single-workgroup sum of partials, then the scalar step of the
// PCG recurrence (alpha, convergence test, beta), replacing TernaryOnArray
// one thread's share of a finish reduction: terms threadIdx.x, threadIdx.x + blockDim.x, ...
// added in index order; U terms are loaded together (masked past n), so a thread's loads
// overlap in one round trip per U terms
template <int U, bool TWO>
__device__ __forceinline__ void finish_terms(const double* __restrict__ p, const double* __restrict__ p2, int n, double& v,
                                             double& v2)
{
    const int step = blockDim.x;
    for (int i = threadIdx.x; i < n; i += U * step) {
        double a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = i + u * step;
            a[u] = j < n ? p[j] : 0.0;
            if (TWO) b[u] = j < n ? p2[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (i + u * step < n) {
                v += a[u];
                if (TWO) v2 += b[u];
            }
        }
    }
}

// thread 0 of finish (1024 threads): the 16 wave sums, in wave order, into the scalars of `kind`
__device__ __forceinline__ void finish_apply(double* __restrict__ S, int kind, const double* red, const double* red2)
{
    double s = 0, s2 = 0;
    for (int i = 0; i < 16; ++i) { s += red[i]; s2 += red2[i]; }
    switch (kind) {
    case FIN_SUM: S[S_TMP] = s; break;
    case FIN_GAMMA0: S[S_GAMMA] = s; break;
    case FIN_ALPHA:
        S[S_FP] = s;
        if (!isfinite(s)) S[S_DONE] = 3;
        else if (s == 0.0) S[S_DONE] = 2;
        else S[S_ALPHA] = S[S_GAMMA] / s;
        break;
    case FIN_RR:
        S[S_RR] = s;
        S[S_ITER] += 1;
        if (!isfinite(s)) S[S_DONE] = 3;
        else if (S[S_STRIDE] > 0 && s < S[S_THR2] && fmod(S[S_ITER], S[S_STRIDE]) == 0.0) S[S_DONE] = 1;
        else if (S[S_ITER] >= S[S_MAXIT]) S[S_DONE] = 4;
        break;
    case FIN_BETA:
        S[S_GAMMA_OLD] = S[S_GAMMA];
        S[S_GAMMA] = s;
        S[S_BETA] = s / S[S_GAMMA_OLD];
        if (!isfinite(S[S_BETA])) S[S_DONE] = 3;
        break;
    case FIN_MEAN: S[S_MEAN] = s / s2; break;
    case FIN_RHS2:
        // |b|^2 turns the relative threshold tol^2 into tol^2 |b|^2; a zero b is solved by
        // x = 0, a non-finite one stops the solve
        S[S_TMP] = s;
        S[S_RHS2] = s;
        if (!isfinite(s)) S[S_DONE] = 3;
        else if (s == 0.0) S[S_DONE] = 1;
        else S[S_THR2] = fmax(S[S_THR2] * s, 2.2250738585072014e-308);
        break;
    case FIN_BETA_FLEX:
        // flexible CG (Polak-Ribiere): beta = z.(r - r_old) / gamma_old = -alpha z.Ap / gamma_old
        S[S_GAMMA_OLD] = S[S_GAMMA];
        S[S_GAMMA] = s;
        S[S_BETA] = -S[S_ALPHA] * s2 / S[S_GAMMA_OLD];
        if (!isfinite(S[S_BETA])) S[S_DONE] = 3;
        break;
    }
}

// FIN_RR with sig (host-mapped coherent memory): one 64-bit word, the count of FIN_RR launches
// of the solve (S_POLL) times 8 plus the stop code S_DONE (0-4), stored as a system-scope
// atomic: the host sees the outcome of each stopping test without a copy or a stream
// synchronisation (the word is all it reads there, so no release of other data is needed)
__global__ __launch_bounds__(1024) void finish(const double* __restrict__ partial, const double* __restrict__ partial2, int n,
                                               double* __restrict__ S, int kind, unsigned long long* __restrict__ sig)
{
    __shared__ double red[16];
    __shared__ double red2[16];
    // the stop flag is read first and tested only before the writes, so that its load
    // overlaps the partials' (a stopped solve's partials are stale but in bounds)
    const bool skip = kind != FIN_SUM && kind != FIN_MEAN && S[S_DONE] != 0.0;
    // each thread adds its strided partials in index order
    double v = 0, v2 = 0;
    if (partial2) finish_terms<16, true>(partial, partial2, n, v, v2);
    else finish_terms<40, false>(partial, nullptr, n, v, v2);
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
        v2 += __shfl_xor(v2, o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[w] = v; red2[w] = v2; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (kind == FIN_RR) {
        // the stopping test (skipped once the solve has stopped) and its outcome
        S[S_POLL] += 1;
        if (!skip) finish_apply(S, kind, red, red2);
        if (sig) {
            const unsigned long long word = (static_cast<unsigned long long>(S[S_POLL]) << 3) |
                                            static_cast<unsigned long long>(S[S_DONE]);
            __hip_atomic_store(sig, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        return;
    }
    if (!skip) finish_apply(S, kind, red, red2);
}

// level-0 FP64 dense solve: x = A0^-1 b on the level's INTERIOR cells.
// One wave per row: lanes stride the row (coalesced FP64 loads of the inverse),
// then a shuffle reduction; 4 rows per 256-thread workgroup.
__global__ __launch_bounds__(256) void coarse_direct(const double* __restrict__ inv, const int32_t* __restrict__ cells, int n,
                                                     const int32_t* __restrict__ zero_cells, int nz, float* __restrict__ X,
                                                     const float* __restrict__ B, const double* __restrict__ S)
{
    extern __shared__ double sb[];
    if (stopped(S)) return;
    for (int i = threadIdx.x; i < n; i += blockDim.x) sb[i] = double(B[cells[i]]);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r < n) {
        const double* row = inv + size_t(r) * n;
        double acc = 0;
        for (int c = lane; c < n; c += 64) acc = fma(row[c], sb[c], acc);
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) X[cells[r]] = float(acc);
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nz; i += gridDim.x * blockDim.x) X[zero_cells[i]] = 0.0f;
}

}  // namespace k

namespace k {

// ChannelPowerSumKernel128 (src/PoissonGrid.cu:223-319) on one tile per 128-thread
// workgroup: per tile the sum of |x|^order (order -1: max |x|) over the cells
// whose type is in `cell_types`, and their count, both times h^3 when volume
// weighted; FP64 throughout
__global__ __launch_bounds__(128) void power_sum(TopoView tp, const int32_t* __restrict__ list, const float* __restrict__ X,
                                                 int order, int volume_weighted, int cell_types, double* __restrict__ vsum,
                                                 double* __restrict__ wsum)
{
    __shared__ double red[2][2];
    const int t = list[blockIdx.x];
    double v = 0, c = 0;
    if (t >= 0) {
        const size_t q = size_t(t) * 128 + threadIdx.x;
        const float4 x = reinterpret_cast<const float4*>(X)[q];
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        const float xs[4] = {x.x, x.y, x.z, x.w};
        const uint8_t cs[4] = {ct.x, ct.y, ct.z, ct.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (!(cs[e] & cell_types)) continue;
            const double a = fabs(double(xs[e]));
            c += 1;
            if (order == -1) v = fmax(v, a);
            else if (order == 1) v += a;
            else if (order == 2) v += a * a;
            else v += pow(a, double(order));
        }
    }
    // wave64 reductions, then the two waves
    for (int o = 32; o > 0; o >>= 1) {
        const double vo = __shfl_xor(v, o, 64);
        v = order == -1 ? fmax(v, vo) : v + vo;
        c += __shfl_xor(c, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][0] = v;
        red[threadIdx.x >> 6][1] = c;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double bv = order == -1 ? fmax(red[0][0], red[1][0]) : red[0][0] + red[1][0];
    double bc = red[0][1] + red[1][1];
    if (t >= 0 && volume_weighted && order != -1) {
        const double h = double(tile_h(tp, t));
        bv *= h * h * h;
        bc *= h * h * h;
    }
    vsum[blockIdx.x] = bv;
    wsum[blockIdx.x] = bc;
}

// ------------------------------------------------- non-FAS transfers ----

// RestrictSourceToCoarse (src/AMGSolver.cu:1295-1307) = AccumulateToParentsOneStep
// over LEAF|NONLEAF tiles of a level, INTERIOR cells: parent = coeff * sum8
__global__ __launch_bounds__(64) void restrict_source(TopoView tp, const int32_t* __restrict__ list, float* __restrict__ X,
                                                      float coeff, const double* __restrict__ S)
{
    const int t = list_tile(list, S);   // < 0: no tile, or the solve has stopped
    if (t < 0) return;
    const int pt = tp.parent[t];
    if (pt < 0) return;
    const int v = threadIdx.x, oc = tp.toct[t];
    const int fl0 = v / 16 * 2, fl1 = (v / 4) % 4 * 2, fl2 = v % 4 * 2;
    float sum = 0;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                const size_t idx = size_t(t) * 512 + cidx(fl0 + ii, fl1 + jj, fl2 + kk);
                sum += (tp.ctype[idx] & INTERIOR) ? X[idx] : 0.0f;
            }
    X[size_t(pt) * 512 + cidx(4 * (oc & 1) + (fl0 >> 1), 4 * ((oc >> 1) & 1) + (fl1 >> 1), 4 * ((oc >> 2) & 1) + (fl2 >> 1))] =
        sum * coeff;
}

// ProlongateAndUpdateAMG128Kernel (src/AMGSolver.cu:1022-1053), four cells per thread:
// GHOST x = parent x, others x += P * parent x (non-INTERIOR cells -> 0)
__global__ __launch_bounds__(256) void prolong_add_v(TopoView tp, const int32_t* __restrict__ list, int n, float* __restrict__ X,
                                                     float P, const double* __restrict__ S)
{
    if (stopped(S)) return;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const int v0 = int(threadIdx.x & 127) * 4;
        const int pt = tp.parent[t_];
        const int oc = tp.toct[t_];
        const bool ghost = tp.ttype[t_] & GHOST;
        const uchar4 ct = reinterpret_cast<const uchar4*>(tp.ctype)[q];
        const float4 x = reinterpret_cast<const float4*>(X)[q];
        const int i = v0 >> 6, j = (v0 >> 3) & 7, k0 = v0 & 7;
        const int pbase = cidx(4 * (oc & 1) + (i >> 1), 4 * ((oc >> 1) & 1) + (j >> 1), 4 * ((oc >> 2) & 1) + (k0 >> 1));
        float cx[2] = {0.0f, 0.0f};
        if (pt >= 0) {
            const size_t pb = size_t(pt) * 512 + pbase;
            cx[0] = X[pb];
            cx[1] = X[pb + 1];
        }
        const uint8_t cts[4] = {ct.x, ct.y, ct.z, ct.w};
        float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const bool in = cts[e] & INTERIOR;
            const float c1 = cx[e >> 1];
            xs[e] = ghost ? (in ? c1 : 0.0f) : (in ? xs[e] + P * c1 : 0.0f);
        }
        reinterpret_cast<float4*>(X)[q] = make_float4(xs[0], xs[1], xs[2], xs[3]);
    }
}

// DeductFluxOnAbsLeafs128Kernel (src/AMGSolver.cu:966-993) on LEAF tiles of a
// level: rhs = f - negativeLapOuterNonLeaf(x) on INTERIOR cells, else 0.  Only
// halo cells in NONLEAF face neighbours carry x; every other term is off * 0,
// kept so that the FMA chain (and the sign of a zero) is the operator's.
__global__ __launch_bounds__(512) void deduct_flux(TopoView tp, const int32_t* __restrict__ list, const float* __restrict__ X,
                                                   Coef C, const float* __restrict__ F, float* __restrict__ RHS,
                                                   const double* __restrict__ S)
{
    const int t = list_tile(list, S);   // < 0: no tile, or the solve has stopped
    if (t < 0) return;
    const int v = threadIdx.x;
    const size_t idx = size_t(t) * 512 + v;
    if (!(tp.ctype[idx] & INTERIOR)) {
        RHS[idx] = 0.0f;
        return;
    }
    const float h = tile_h(tp, t);
    const int l[3] = {v >> 6, (v >> 3) & 7, v & 7};
    const int stride[3] = {64, 8, 1};
    const float d = C.c[3][idx];
    float sum = 0.0f;
    if (d != 0) {
        sum = 0.0f * d;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int sg = -1; sg <= 1; sg += 2) {
                float off, x1 = 0.0f;
                const bool inside = sg < 0 ? l[a] > 0 : l[a] < 7;
                if (sg < 0) off = C.c[a][idx];
                else if (inside) off = C.c[a][idx + stride[a]];
                else off = h;
                if (!inside) {
                    const int nt = tp.nbr[size_t(t) * 6 + (sg < 0 ? a : a + 3)];
                    if (nt >= 0) {
                        const size_t ni = size_t(nt) * 512 + (v + (sg < 0 ? 7 : -7) * stride[a]);
                        if (sg > 0) off = C.c[a][ni];
                        if (tp.ttype[nt] == NONLEAF) x1 = tp.ctype[ni] == INTERIOR ? X[ni] : 0.0f;
                    }
                }
                sum = fmaf(off, x1, sum);
            }
        }
    }
    RHS[idx] = F[idx] - sum;
}

// initial launch of VCycle / muCycle (src/AMGSolver.cu:1062-1078): x = 0 on every
// tile, rhs = f on LEAF|NONLEAF tiles and 0 on GHOST tiles
__global__ __launch_bounds__(256) void nonfas_init_v(TopoView tp, const int32_t* __restrict__ list, int n, float* __restrict__ X,
                                                     const float* __restrict__ F, float* __restrict__ RHS,
                                                     const double* __restrict__ S)
{
    if (stopped(S)) return;
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        reinterpret_cast<float4*>(X)[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        reinterpret_cast<float4*>(RHS)[q] =
            (tp.ttype[t_] & GHOST) ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : reinterpret_cast<const float4*>(F)[q];
    }
}

// Y = B - Y on the tiles of a list (the residual of FASMuCycleSolve, src/AMGSolver.cu:1340-1344)
__global__ __launch_bounds__(256) void sub_from_v(const int32_t* __restrict__ list, int n, const float* __restrict__ B,
                                                  float* __restrict__ Y)
{
    CMG_TILE_LOOP(list, n)
    {
        const size_t q = quad_index(t_);
        const float4 b = reinterpret_cast<const float4*>(B)[q];
        float4 y = reinterpret_cast<const float4*>(Y)[q];
        y.x = b.x - y.x; y.y = b.y - y.y; y.z = b.z - y.z; y.w = b.w - y.w;
        reinterpret_cast<float4*>(Y)[q] = y;
    }
}

}  // namespace k

// ---------------------------------------------------------------- launchers ---

static inline bool nonempty(const LaunchList& l) { return l.n > 0; }

// Kernel forms of the stencil families.  Threads per tile: 512 / 256 / 128
// (1 / 2 / 4 cells per thread).  Persistence mask: bit 0 apply, bit 1 Jacobi,
// bit 2 residual + restrict; a set bit selects the persistent pipelined kernel,
// a clear bit the one-tile-per-workgroup kernel.  Bit 3 keeps the stored LEAF
// diagonals even where they could be derived from the face coefficients.  Bit 4
// keeps the preconditioner's boundary work in launches of their own: the first
// finest-level Jacobi sweep (otherwise fused into the PCG x/r update) and r.z
// (otherwise summed over each level's LEAF tiles by that level's last
// post-sweep).  Defaults may be overridden by
// the environment (CMG_STENCIL_THREADS, CMG_PERSISTENT) or set_kernel_forms.
enum { PERSIST_APPLY = 1, PERSIST_JACOBI = 2, PERSIST_RESTRICT = 4 };
static std::atomic<int> g_threads{0}, g_mask{-1};
static std::atomic<bool> g_threads_chosen{false};   // by set_kernel_forms or CMG_STENCIL_THREADS

static int stencil_threads()
{
    int v = g_threads.load(std::memory_order_relaxed);
    if (v) return v;
    const char* e = getenv("CMG_STENCIL_THREADS");
    if (e) g_threads_chosen.store(true, std::memory_order_relaxed);
    v = e ? atoi(e) : 256;
    v = (v == 128 || v == 256 || v == 512) ? v : 256;
    g_threads.store(v, std::memory_order_relaxed);
    return v;
}

static int persistent_mask()
{
    int m = g_mask.load(std::memory_order_relaxed);
    if (m >= 0) return m;
    const char* e = getenv("CMG_PERSISTENT");
    // one tile per workgroup everywhere: faster than the persistent forms on MI355X at
    // every thread count (the persistent forms hold two tiles' loads in registers)
    m = e ? (atoi(e) & 127) : 0;
    g_mask.store(m, std::memory_order_relaxed);
    return m;
}

bool set_kernel_forms(int mask, int threads)
{
    if (mask < 0 || mask > 127 || (threads != 0 && threads != 128 && threads != 256 && threads != 512)) return false;
    g_mask.store(mask, std::memory_order_relaxed);
    // 0: the library's thread counts again (CMG_STENCIL_THREADS if set, else the per-kernel defaults)
    g_threads_chosen.store(threads != 0, std::memory_order_relaxed);
    g_threads.store(threads, std::memory_order_relaxed);
    return true;
}

// the one-tile kernels on FP16 coefficient copies or codes stage fewer bytes a tile and
// carry fewer bytes in a tile's latency chain; twice the tiles in flight per CU pays there
// (MI355X, finest level of the 1024^3 torus workload: FP16 Jacobi sweep 72.6 us at 128
// threads a tile against 77.8 at 256, 64 threads a tile, 90 VGPRs, 71.6 against 69.0;
// coded residual + restriction 50.0 against 59.4).  They run 128 unless a thread count was
// chosen.
static int q_jacobi_threads()
{
    const int v = stencil_threads();
    return g_threads_chosen.load(std::memory_order_relaxed) ? v : 128;
}

// the thread count of a one-tile stencil launch of coefficient form F (before the small-launch
// rule): the compact forms (FP16 copies, one-byte codes) run q_jacobi_threads; the FP32
// channels and the ten-bit codes, whose table makes a tile's LDS 11.5 KB, stencil_threads
// (MI355X, the 2,048 tiles of the torus workload's level 5 with ten-bit codes: the sweep
// carrying the prolongation to level 6 18.8 us at 128 threads a tile, 17.2 with FP32 at 256)
static int form_threads(int F) { return (F == 1 || F == 2) ? q_jacobi_threads() : stencil_threads(); }

// the operator on LEAF tiles likewise runs 128 threads a tile unless a thread count
// was chosen (106 us against 108 at 256 on the 1024^3 torus workload)
static int apply_threads() { return g_threads_chosen.load(std::memory_order_relaxed) ? stencil_threads() : 128; }

// a level launch of at most kSmallLaunch tiles is one round of workgroups at 512 threads a
// tile, and its time is one tile's latency chain, which fewer cells a thread shorten: such
// launches run 512 threads a tile unless a thread count was chosen
constexpr int kSmallLaunch = 1024;
static int level_threads(int def, int tiles)
{
    return (!g_threads_chosen.load(std::memory_order_relaxed) && tiles <= kSmallLaunch) ? 512 : def;
}

int launched_threads_key()
{
    // the chosen flag decides the small-launch counts as well
    return (g_threads_chosen.load(std::memory_order_relaxed) ? 1000000 : 0) + apply_threads() * 1000 + q_jacobi_threads();
}

void get_kernel_forms(int* mask, int* threads)
{
    if (mask) *mask = persistent_mask();
    if (threads) *threads = stencil_threads();
}

// persistent stencil kernels run exactly as many workgroups as can be resident
// (occupancy x CUs, a multiple of 8 so a workgroup stays on one XCD's share of
// the list); vector kernels use up to 8 workgroups per CU
static int resident_per_cu(const void* kern, int nt)
{
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_pair(kern, nt);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, nt, 0) != hipSuccess || nb < 1) nb = 1;
    cache[key] = nb;
    return nb;
}

static inline int persistent_grid(const DevGrid& d, int n, const void* kern, int nt)
{
    int g = resident_per_cu(kern, nt) * (d.wg_cap / 8);
    g = std::max(8, g - g % 8);
    return std::max(1, std::min(n, g));
}
// one two-tile step per workgroup over a full grid: on MI355X a full grid streams at
// 6.2-6.5 TB/s where a grid-stride loop over 8 workgroups per CU reaches 4.5-4.9
static inline int vector_grid(const DevGrid&, int n) { return std::max(8, 8 * ((n + 15) / 16)); }

#define CMG_NT_PERSISTENT(KERNEL, N, STREAM, ...)                                                                   \
    do {                                                                                                          \
        switch (stencil_threads()) {                                                                              \
        case 128: { const int g_ = persistent_grid(d, (N), (const void*)KERNEL<128>, 128);                       \
                    KERNEL<128><<<g_, 128, 0, (STREAM)>>>(__VA_ARGS__); } break;                                  \
        case 512: { const int g_ = persistent_grid(d, (N), (const void*)KERNEL<512>, 512);                       \
                    KERNEL<512><<<g_, 512, 0, (STREAM)>>>(__VA_ARGS__); } break;                                  \
        default:  { const int g_ = persistent_grid(d, (N), (const void*)KERNEL<256>, 256);                       \
                    KERNEL<256><<<g_, 256, 0, (STREAM)>>>(__VA_ARGS__); } break;                                  \
        }                                                                                                         \
    } while (0)

#define CMG_NT_DISPATCH(KERNEL, GRID, STREAM, ...)                                            \
    do {                                                                                      \
        switch (stencil_threads()) {                                                          \
        case 128: KERNEL<128><<<(GRID), 128, 0, (STREAM)>>>(__VA_ARGS__); break;             \
        case 512: KERNEL<512><<<(GRID), 512, 0, (STREAM)>>>(__VA_ARGS__); break;             \
        default: KERNEL<256><<<(GRID), 256, 0, (STREAM)>>>(__VA_ARGS__); break;             \
        }                                                                                     \
    } while (0)

// the same with coefficient form F (second template argument: 1 FP16 copies, 2 one-byte
// codes, 4 ten-bit codes), at the thread count of its form (or of a small level launch)
#define CMG_NT_DISPATCH_F(KERNEL, F, GRID, STREAM, ...)                                       \
    do {                                                                                      \
        switch (level_threads(form_threads(F), (GRID))) {                                     \
        case 128: KERNEL<128, F><<<(GRID), 128, 0, (STREAM)>>>(__VA_ARGS__); break;          \
        case 512: KERNEL<512, F><<<(GRID), 512, 0, (STREAM)>>>(__VA_ARGS__); break;          \
        default: KERNEL<256, F><<<(GRID), 256, 0, (STREAM)>>>(__VA_ARGS__); break;          \
        }                                                                                     \
    } while (0)

// every form of a one-tile stencil kernel KERNEL<NT, Q> on one level
#define CMG_NT_DISPATCH_ANY(KERNEL, FORM, GRID, STREAM, ...)                                  \
    do {                                                                                      \
        if ((FORM) == 1) CMG_NT_DISPATCH_F(KERNEL, 1, GRID, STREAM, __VA_ARGS__);            \
        else if ((FORM) == 2) CMG_NT_DISPATCH_F(KERNEL, 2, GRID, STREAM, __VA_ARGS__);       \
        else if ((FORM) == 4) CMG_NT_DISPATCH_F(KERNEL, 4, GRID, STREAM, __VA_ARGS__);       \
        else CMG_NT_DISPATCH_F(KERNEL, 0, GRID, STREAM, __VA_ARGS__);                        \
    } while (0)

// the coefficient form of a one-tile stencil launch on `level`: 1 the FP16 copies, 2 the
// one-byte codes, 4 the ten-bit codes (set for that level), 0 the FP32 channels
static int coef_form(const Coef& C, int level)
{
    if (C.q[0]) return 1;
    if (!C.w || level < 0 || level >= 32 || !((C.coded >> level) & 1u)) return 0;
    return ((C.wide >> level) & 1u) ? 4 : 2;
}

template <bool DOT, int NT>
static int apply_launch(const DevGrid& d, const float* X, Coef C, float* Y, double* partial, const double* S, hipStream_t st)
{
    const int n = d.leafs.n;
    const int g = persistent_grid(d, n, (const void*)k::apply_p<DOT, NT>, NT);
    k::apply_p<DOT, NT><<<g, NT, 0, st>>>(view(d), d.leafs.d, n, X, C, Y, partial, S);
    return g;
}

// the stop block a one-tile stencil launch reads: the solve's, or the grid's idle block
static const double* run_flag(const DevGrid& d, const double* S) { return S ? S : d.idle; }

int launch_apply(const DevGrid& d, const float* X, Coef C, float* Y, double* partial, const double* S, hipStream_t st)
{
    if (!nonempty(d.leafs)) return 0;
    if (!(persistent_mask() & PERSIST_APPLY)) {
        const TopoView tp = view(d);
        const double* Sr = run_flag(d, S);
        if (C.w && (C.coded & ~C.wide)) {
            // tiles of levels with one-byte codes decode their coefficients (the same values);
            // ten-bit-coded levels read the FP32 channels here
            double* const pp = partial;
            switch (level_threads(apply_threads(), d.leafs.n)) {
            case 128:
                if (pp) k::apply_leaf<true, 128, true><<<d.leafs.n, 128, 0, st>>>(tp, d.leafs.rec, X, C, Y, pp, Sr);
                else k::apply_leaf<false, 128, true><<<d.leafs.n, 128, 0, st>>>(tp, d.leafs.rec, X, C, Y, nullptr, Sr);
                break;
            case 512:
                if (pp) k::apply_leaf<true, 512, true><<<d.leafs.n, 512, 0, st>>>(tp, d.leafs.rec, X, C, Y, pp, Sr);
                else k::apply_leaf<false, 512, true><<<d.leafs.n, 512, 0, st>>>(tp, d.leafs.rec, X, C, Y, nullptr, Sr);
                break;
            default:
                if (pp) k::apply_leaf<true, 256, true><<<d.leafs.n, 256, 0, st>>>(tp, d.leafs.rec, X, C, Y, pp, Sr);
                else k::apply_leaf<false, 256, true><<<d.leafs.n, 256, 0, st>>>(tp, d.leafs.rec, X, C, Y, nullptr, Sr);
                break;
            }
        } else if (partial) {
            switch (level_threads(apply_threads(), d.leafs.n)) {
            case 128: k::apply_leaf<true, 128><<<d.leafs.n, 128, 0, st>>>(tp, d.leafs.rec, X, C, Y, partial, Sr); break;
            case 512: k::apply_leaf<true, 512><<<d.leafs.n, 512, 0, st>>>(tp, d.leafs.rec, X, C, Y, partial, Sr); break;
            default: k::apply_leaf<true, 256><<<d.leafs.n, 256, 0, st>>>(tp, d.leafs.rec, X, C, Y, partial, Sr); break;
            }
        } else {
            switch (level_threads(apply_threads(), d.leafs.n)) {
            case 128: k::apply_leaf<false, 128><<<d.leafs.n, 128, 0, st>>>(tp, d.leafs.rec, X, C, Y, nullptr, Sr); break;
            case 512: k::apply_leaf<false, 512><<<d.leafs.n, 512, 0, st>>>(tp, d.leafs.rec, X, C, Y, nullptr, Sr); break;
            default: k::apply_leaf<false, 256><<<d.leafs.n, 256, 0, st>>>(tp, d.leafs.rec, X, C, Y, nullptr, Sr); break;
            }
        }
        return d.leafs.n;
    }
    switch (stencil_threads()) {
    case 128: return partial ? apply_launch<true, 128>(d, X, C, Y, partial, S, st) : apply_launch<false, 128>(d, X, C, Y, nullptr, S, st);
    case 512: return partial ? apply_launch<true, 512>(d, X, C, Y, partial, S, st) : apply_launch<false, 512>(d, X, C, Y, nullptr, S, st);
    default: return partial ? apply_launch<true, 256>(d, X, C, Y, partial, S, st) : apply_launch<false, 256>(d, X, C, Y, nullptr, S, st);
    }
}

void launch_residual64(const DevGrid& d, const double* X, Coef C, const double* B, double* R64, float* R32, double* partial,
                       hipStream_t st)
{
    if (!nonempty(d.leafs)) return;
    CMG_NT_DISPATCH(k::residual64, d.leafs.n, st, view(d), d.leafs.rec, X, C, B, R64, R32, partial);
}

void launch_gs_color(const DevGrid& d, int level, float* X, Coef C, const float* B, int color, double omega, const double* S,
                     hipStream_t st)
{
    const LaunchList& l = d.subtree[level];
    if (!nonempty(l)) return;
    CMG_NT_DISPATCH_ANY(k::gs_color, coef_form(C, level), l.n, st, view(d), l.rec, X, C, B, color, omega, run_flag(d, S));
}

// one launch of the one-tile Jacobi form jacobi<NT, DOT, Q, SAVE, PUSH, PM> with nt threads a tile
template <bool DOT, int Q, bool SAVE, bool PUSH = false, int PM = 0>
static void jacobi_at(int nt, const LaunchList& l, hipStream_t st, const TopoView& tp, const float* Xin, float* Xout,
                      const float* G, Coef C, const float* B, float w, const double* S, double* partial, float* X0,
                      float P = 0.0f)
{
    switch (nt) {
    case 128:
        k::jacobi<128, DOT, Q, SAVE, PUSH, PM><<<l.n, 128, 0, st>>>(tp, l.rec, Xin, Xout, G, C, B, w, S, partial, X0, P);
        break;
    case 512:
        k::jacobi<512, DOT, Q, SAVE, PUSH, PM><<<l.n, 512, 0, st>>>(tp, l.rec, Xin, Xout, G, C, B, w, S, partial, X0, P);
        break;
    default:
        k::jacobi<256, DOT, Q, SAVE, PUSH, PM><<<l.n, 256, 0, st>>>(tp, l.rec, Xin, Xout, G, C, B, w, S, partial, X0, P);
        break;
    }
}

bool jacobi_push_available(const DevGrid& d, int level)
{
    return level + 1 <= d.max_level && d.push_ok[level + 1] && !(persistent_mask() & (16 | PERSIST_JACOBI));
}

// the one-tile Jacobi launch of coefficient form Q (arguments checked by launch_jacobi)
template <int Q>
static int jacobi_one_tile(const LaunchList& l, int nt, hipStream_t st, const TopoView& tp, const float* Xin, float* Xout,
                           const float* G, Coef C, const float* B, float w, const double* S, double* partial, float* X0,
                           bool push, float P, int pull)
{
    if (pull == 1) {
        if (partial) jacobi_at<true, Q, false, false, 1>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, partial, X0, P);
        else jacobi_at<false, Q, false, false, 1>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, nullptr, X0, P);
    } else if (pull == 2) {
        if (partial) jacobi_at<true, Q, false, false, 2>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, partial, X0, P);
        else jacobi_at<false, Q, false, false, 2>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, nullptr, X0, P);
    } else if (push) {
        if (partial) jacobi_at<true, Q, false, true>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, partial, X0, P);
        else jacobi_at<false, Q, false, true>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, nullptr, X0, P);
    } else if (partial) {
        jacobi_at<true, Q, false>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, partial, nullptr);
    } else if (X0) {
        jacobi_at<false, Q, true>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, nullptr, X0);
    } else {
        jacobi_at<false, Q, false>(nt, l, st, tp, Xin, Xout, G, C, B, w, S, nullptr, nullptr);
    }
    return partial ? l.n : 0;   // one partial per list entry
}

int launch_jacobi(const DevGrid& d, int level, const float* Xin, float* Xout, const float* G, Coef C, const float* B, float w,
                  const double* S, hipStream_t st, double* partial, float* X0, bool push, float P, int pull)
{
    if (X0 && partial && !push && !pull) throw std::invalid_argument("jacobi: the FAS-save form writes no partials");
    const LaunchList& l = d.subtree[level];
    if (!nonempty(l)) return 0;
    const int form = coef_form(C, level);
    const bool one_tile = !(persistent_mask() & PERSIST_JACOBI);
    if ((partial || X0 || form) && !one_tile)
        throw std::invalid_argument("jacobi: b.x partials, the FAS save and FP16 or coded coefficients need the one-tile form");
    if (pull && (!X0 || push || pull < 1 || pull > 2))
        throw std::invalid_argument("jacobi: the pulled prolongation is a one-tile post-sweep reading X0");
    if (push && (!X0 || !jacobi_push_available(d, level)))
        throw std::invalid_argument("jacobi: the prolongation cannot ride on this level's sweep");
    const TopoView tp = view(d);
    // pull: a post-sweep of a level whose prolongation it pulls in; push: the last post-sweep
    // with the prolongation to level + 1 (X0: the FAS-saved x either way); X0 alone: the first
    // coarse-level sweep with the level's FAS save
    if (!one_tile) {
        CMG_NT_PERSISTENT(k::jacobi_p, l.n, st, tp, l.d, l.n, Xin, Xout, G, C, B, w, S);
        return 0;
    }
    S = run_flag(d, S);
    const int nt = level_threads(form_threads(form), l.n);
    if (form == 1) return jacobi_one_tile<1>(l, nt, st, tp, Xin, Xout, G, C, B, w, S, partial, X0, push, P, pull);
    if (form == 2) return jacobi_one_tile<2>(l, nt, st, tp, Xin, Xout, G, C, B, w, S, partial, X0, push, P, pull);
    if (form == 4) return jacobi_one_tile<4>(l, nt, st, tp, Xin, Xout, G, C, B, w, S, partial, X0, push, P, pull);
    return jacobi_one_tile<0>(l, nt, st, tp, Xin, Xout, G, C, B, w, S, partial, X0, push, P, pull);
}

void launch_jacobi_zero(const DevGrid& d, int level, float* Xout, const float* D, const float* B, float w, const double* S,
                        hipStream_t st, const uint16_t* Dq)
{
    const LaunchList& l = d.subtree[level];
    if (!nonempty(l)) return;
    k::jacobi_zero_v<<<vector_grid(d, l.n), 256, 0, st>>>(view(d), l.d, l.n, Xout, D, B, w, S, Dq);
}

void launch_residual_restrict(const DevGrid& d, int level, float* X, Coef C, float* B, float Rr, const double* S, hipStream_t st)
{
    const LaunchList& l = d.subtree[level];
    if (!nonempty(l)) return;
    if (const int f = coef_form(C, level)) {
        if (persistent_mask() & PERSIST_RESTRICT)
            throw std::invalid_argument("residual_restrict: FP16 or coded coefficients need the one-tile form");
        CMG_NT_DISPATCH_ANY(k::residual_restrict, f, l.n, st, view(d), l.rec, X, C, B, Rr, run_flag(d, S));
        return;
    }
    if (!(persistent_mask() & PERSIST_RESTRICT))
        CMG_NT_DISPATCH_F(k::residual_restrict, 0, l.n, st, view(d), l.rec, X, C, B, Rr, run_flag(d, S));
    else CMG_NT_PERSISTENT(k::restrict_p, l.n, st, view(d), l.d, l.n, X, C, B, Rr, S);
}

void launch_fas_save(const DevGrid& d, int level, const float* X, Coef C, float* B, float* X0, const double* S, hipStream_t st)
{
    const LaunchList& l = d.nonleaf[level];
    if (!nonempty(l)) return;
    CMG_NT_DISPATCH_ANY(k::fas_save, coef_form(C, level), l.n, st, view(d), l.rec, X, C, B, X0, run_flag(d, S));
}

void launch_coef_to_half(const DevGrid& d, Coef C, uint16_t* const q[4], unsigned int* bad, hipStream_t st)
{
    if (!nonempty(d.all)) return;
    k::coef_to_half<<<vector_grid(d, d.all.n), 256, 0, st>>>(view(d), d.all.d, d.all.n, C, q[0], q[1], q[2], q[3], bad);
}

size_t coef_code_words() { return size_t(2) * kMaxLevels * k::kCodeSlots + 3 * kMaxLevels; }
size_t coef_code_result_offset() { return size_t(2) * kMaxLevels * k::kCodeSlots + 2 * kMaxLevels; }

void launch_coef_codes(const DevGrid& d, Coef C, uint32_t* ws, float* lut, uint32_t* W, hipStream_t st)
{
    // empty key sets (all bits set), zero counts, flags and results
    if (hipMemsetAsync(ws, 0xff, sizeof(uint32_t) * kMaxLevels * k::kCodeSlots, st) != hipSuccess ||
        hipMemsetAsync(ws + 2 * kMaxLevels * k::kCodeSlots, 0, sizeof(uint32_t) * 3 * kMaxLevels, st) != hipSuccess)
        throw std::runtime_error("coefficient codes: hipMemsetAsync failed");
    if (nonempty(d.all))
        k::code_collect<<<vector_grid(d, d.all.n), 256, 0, st>>>(view(d), d.all.d, d.all.n, C, ws);
    k::code_number<<<kMaxLevels, 256, 0, st>>>(ws, lut, d.h0);
    if (nonempty(d.all))
        k::code_encode<<<vector_grid(d, d.all.n), 256, 0, st>>>(view(d), d.all.d, d.all.n, C, ws, W);
}

void launch_code_zfaces(const DevGrid& d, const uint32_t* W, uint32_t* ZW, hipStream_t st)
{
    if (!nonempty(d.all)) return;
    k::code_zfaces<<<d.all.n, 128, 0, st>>>(view(d), d.all.d, d.all.n, W, ZW);
}

void launch_fas_prolong(const DevGrid& d, int level, float* X, const float* X0, float P, const double* S, hipStream_t st)
{
    const LaunchList& l = d.prolong[level];
    if (!nonempty(l)) return;
    if (!l.meta) throw std::logic_error("fas_prolong: the prolongation list has no parent records");
    k::fas_prolong_v<<<vector_grid(d, l.n), 256, 0, st>>>(view(d), l.meta, l.n, X, X0, P, S);
}

void launch_propagate_ghosts(const DevGrid& d, float* X, hipStream_t st)
{
    if (!nonempty(d.ghosts)) return;
    k::propagate_ghosts<<<d.ghosts.n, 512, 0, st>>>(view(d), d.ghosts.d, X);
}

void launch_accumulate_leafs(const DevGrid& d, float* X, hipStream_t st)
{
    if (!nonempty(d.leafs)) return;
    k::accumulate_leafs<<<d.leafs.n, 64, 0, st>>>(view(d), d.leafs.d, X);
}

void launch_fill_all(const DevGrid& d, float* X, float value, const double* S, hipStream_t st)
{
    const size_t n4 = size_t(d.N) * 128;
    if (!n4) return;
    int blocks = int(std::min<size_t>((n4 + 511) / 512, 8192));
    k::fill_all<<<blocks, 512, 0, st>>>(X, n4, value, S);
}

void launch_fill_list(const DevGrid& d, const LaunchList& l, float* X, float value, hipStream_t st, const double* S)
{
    if (!nonempty(l)) return;
    k::fill_v<<<vector_grid(d, l.n), 256, 0, st>>>(l.d, l.n, X, value, S);
}

void launch_copy_flat(const DevGrid& d, const float* A, float* B, size_t n, hipStream_t st)
{
    if (n % 4) throw std::invalid_argument("copy_flat: length must be a multiple of 4");
    const size_t n4 = n / 4, blocks = (n4 + 255) / 256;
    if (blocks == 0) return;
    if (blocks > 0x7fffffffu) throw std::invalid_argument("copy_flat: length too large");
    k::copy_flat<<<unsigned(blocks), 256, 0, st>>>(reinterpret_cast<const k::f32x4v*>(A), reinterpret_cast<k::f32x4v*>(B), n4);
}

void launch_copy_list(const DevGrid& d, const LaunchList& l, const float* A, float* B, const double* S, hipStream_t st)
{
    if (!nonempty(l)) return;
    k::copy_v<<<vector_grid(d, l.n), 256, 0, st>>>(l.d, l.n, A, B, S);
}

int launch_dot(const DevGrid& d, const LaunchList& l, const float* A, const float* B, double* partial, const double* S,
               hipStream_t st)
{
    if (!nonempty(l)) return 0;
    const int g = vector_grid(d, l.n);
    k::dot_v<<<g, 256, 0, st>>>(l.d, l.n, A, B, partial, S);
    return g;
}

int launch_dot2(const DevGrid& d, const LaunchList& l, const float* A, const float* B, const float* C, double* partial,
                double* partial2, const double* S, hipStream_t st)
{
    if (!nonempty(l)) return 0;
    const int g = vector_grid(d, l.n);
    k::dot2_v<<<g, 256, 0, st>>>(l.d, l.n, A, B, C, partial, partial2, S);
    return g;
}

int launch_update_xr(const DevGrid& d, float* X, float* R, const float* P, const float* AP, double* partial, double* S,
                     hipStream_t st, float* Z0, const float* D, float w, const uint16_t* Dq)
{
    if (!S) throw std::invalid_argument("update_xr needs the solver scalars (alpha)");
    if (Z0 && !D) throw std::invalid_argument("update_xr: the fused first sweep needs the diagonal");
    if (!nonempty(d.leafs)) return 0;
    const int g = vector_grid(d, d.leafs.n);
    if (Z0)
        k::update_xr_v<true><<<g, 256, 0, st>>>(view(d), d.leafs.d, d.leafs.n, X, R, P, AP, partial, S, d.max_level, Z0, D, w, Dq);
    else
        k::update_xr_v<false><<<g, 256, 0, st>>>(view(d), d.leafs.d, d.leafs.n, X, R, P, AP, partial, S, -1, nullptr, nullptr, 0.0f,
                                                 nullptr);
    return g;
}

void launch_update_p(const DevGrid& d, float* P, const float* Z, double* S, hipStream_t st, float* X)
{
    if (!S) throw std::invalid_argument("update_p needs the solver scalars (beta)");
    if (!nonempty(d.leafs)) return;
    k::update_p_v<<<vector_grid(d, d.leafs.n), 256, 0, st>>>(view(d), d.leafs.d, d.leafs.n, P, Z, S, X);
}

void launch_x_pending(const DevGrid& d, float* X, const float* P, const double* S, hipStream_t st)
{
    if (!S) throw std::invalid_argument("x_pending needs the solver scalars (alpha)");
    if (!nonempty(d.leafs)) return;
    k::x_pending_v<<<vector_grid(d, d.leafs.n), 256, 0, st>>>(view(d), d.leafs.d, d.leafs.n, X, P, S);
}

void launch_sum_count(const DevGrid& d, const float* X, double* psum, double* pcnt, const double* S, hipStream_t st)
{
    if (!nonempty(d.leafs)) return;
    k::sum_count<<<d.leafs.n, 512, 0, st>>>(view(d), d.leafs.d, X, psum, pcnt, S);
}

void launch_sub_mean(const DevGrid& d, float* X, const double* S, hipStream_t st)
{
    if (!S) throw std::invalid_argument("sub_mean needs the solver scalars (mean)");
    if (!nonempty(d.leafs)) return;
    k::sub_mean<<<d.leafs.n, 512, 0, st>>>(view(d), d.leafs.d, X, S);
}

void launch_dot64(const LaunchList& l, const double* A, const double* B, double* partial, hipStream_t st)
{
    if (!nonempty(l)) return;
    k::dot64_list<<<l.n, 512, 0, st>>>(l.d, A, B, partial);
}

void launch_axpy(const DevGrid& d, const LaunchList& l, float* Y, const float* X, float a, hipStream_t st)
{
    if (!nonempty(l)) return;
    k::axpy_list<<<l.n, 512, 0, st>>>(view(d), l.d, Y, X, a);
}

void launch_update_x64(const DevGrid& d, double* X64, const float* E, hipStream_t st)
{
    if (!nonempty(d.leafs)) return;
    k::update_x64<<<d.leafs.n, 512, 0, st>>>(view(d), d.leafs.d, X64, E);
}

void launch_to_double(const DevGrid& d, const LaunchList& l, const float* A, double* B, bool masked, hipStream_t st)
{
    if (!nonempty(l)) return;
    k::to_double<<<l.n, 512, 0, st>>>(l.d, A, B, view(d), masked ? 1 : 0);
}

void launch_to_float(const LaunchList& l, const double* A, float* B, hipStream_t st)
{
    if (!nonempty(l)) return;
    k::to_float<<<l.n, 512, 0, st>>>(l.d, A, B);
}

void launch_finish(const double* partial, const double* partial2, int n, double* S, int kind, hipStream_t st,
                   unsigned long long* sig)
{
    if (!S || !partial) throw std::invalid_argument("finish needs the partials and the solver scalars");
    k::finish<<<1, 1024, 0, st>>>(partial, partial2, n, S, kind, sig);
}

bool cycle_bottom_available(const DevGrid& d)
{
    return d.bottom_ok && !(persistent_mask() & 16);
}

void launch_cycle_bottom(const DevGrid& d, int phase, float* X, float* X0, float* B, Coef C, float Rr, float P, bool rbgs,
                         int iters, double omega, int mu, int coarsest_iters, bool fas_save1, bool push, double* rz,
                         const double* S, hipStream_t st)
{
    if (!cycle_bottom_available(d)) throw std::logic_error("cycle_bottom: levels 0 and 1 do not fit one workgroup");
    if (push && (rbgs || iters < 1 || !jacobi_push_available(d, 1)))
        throw std::invalid_argument("cycle_bottom: the prolongation to level 2 cannot ride on level 1's last sweep");
    if (rz && (rbgs || iters < 1)) throw std::invalid_argument("cycle_bottom: r.z partials come from a Jacobi sweep");
    k::BottomArgs a;
    a.rec0 = d.subtree[0].rec;
    a.root = d.bottom_root;
    a.ghost = d.bottom_ghost;
    for (int o = 0; o < 8; ++o) {
        a.tile[o] = d.bottom_tile[o];
        a.type[o] = d.bottom_type[o];
        a.entry[o] = d.bottom_entry[o];
    }
    for (int c = 0; c < 64; ++c) a.child[c] = d.bottom_child[c];
    a.nonleaf0 = d.nonleaf[0].tiles;
    a.X = X; a.X0 = X0; a.B = B; a.C = C;
    a.Rr = Rr; a.P = P; a.w = float(omega); a.omega = omega;
    a.rbgs = rbgs; a.iters = iters; a.mu = mu; a.coarsest_iters = coarsest_iters;
    a.phase = phase; a.fas_save1 = fas_save1; a.push = push; a.rz = rz;
    if (C.q[0]) k::cycle_bottom<true><<<1, 1024, 0, st>>>(view(d), a, run_flag(d, S));
    else k::cycle_bottom<false><<<1, 1024, 0, st>>>(view(d), a, run_flag(d, S));
}

bool coarsest_iterate_available(const DevGrid& d)
{
    return d.root_alone && !(persistent_mask() & 16);
}

void launch_coarsest_iterate(const DevGrid& d, float* X, Coef C, const float* B, bool rbgs, int iters, int iters2, int order,
                             double omega, float w, const double* S, hipStream_t st)
{
    if (!coarsest_iterate_available(d)) throw std::logic_error("coarsest_iterate: level 0 is not a single tile");
    const LaunchList& l = d.subtree[0];
    if (C.q[0])
        k::coarsest_iterate<256, true><<<1, 256, 0, st>>>(view(d), l.rec, X, C, B, rbgs, iters, iters2, order, omega, w,
                                                        run_flag(d, S));
    else
        k::coarsest_iterate<256, false><<<1, 256, 0, st>>>(view(d), l.rec, X, C, B, rbgs, iters, iters2, order, omega, w,
                                                         run_flag(d, S));
}

void launch_coarse_direct(const double* inv, const int32_t* cells, int n, const int32_t* zero_cells, int nz, float* X,
                          const float* B, const double* S, hipStream_t st)
{
    if (n == 0 && nz == 0) return;
    const int blocks = std::max(1, (n + 3) / 4);
    k::coarse_direct<<<blocks, 256, size_t(std::max(n, 1)) * sizeof(double), st>>>(inv, cells, n, zero_cells, nz, X, B, S);
}

void launch_restrict_source(const DevGrid& d, int level, float* X, float coeff, const double* S, hipStream_t st)
{
    const LaunchList& l = d.subtree[level];
    if (!nonempty(l)) return;
    k::restrict_source<<<l.n, 64, 0, st>>>(view(d), l.d, X, coeff, S);
}

void launch_prolong_add(const DevGrid& d, int level, float* X, float P, const double* S, hipStream_t st, bool leaf_only)
{
    const LaunchList& l = leaf_only ? d.subtree[level] : d.prolong[level];
    if (!nonempty(l)) return;
    k::prolong_add_v<<<vector_grid(d, l.n), 256, 0, st>>>(view(d), l.d, l.n, X, P, S);
}

void launch_deduct_flux(const DevGrid& d, int level, const float* X, Coef C, const float* F, float* RHS, const double* S,
                        hipStream_t st)
{
    const LaunchList& l = d.leafat[level];
    if (!nonempty(l)) return;
    k::deduct_flux<<<l.n, 512, 0, st>>>(view(d), l.d, X, C, F, RHS, S);
}

void launch_nonfas_init(const DevGrid& d, float* X, const float* F, float* RHS, const double* S, hipStream_t st)
{
    const LaunchList& l = d.all;
    if (!nonempty(l)) return;
    k::nonfas_init_v<<<vector_grid(d, l.n), 256, 0, st>>>(view(d), l.d, l.n, X, F, RHS, S);
}

void launch_sub_from(const DevGrid& d, const LaunchList& l, const float* B, float* Y, hipStream_t st)
{
    if (!nonempty(l)) return;
    k::sub_from_v<<<vector_grid(d, l.n), 256, 0, st>>>(l.d, l.n, B, Y);
}

int launch_power_sum(const DevGrid& d, const float* X, int order, bool volume_weighted, int cell_types, double* vsum,
                     double* wsum, hipStream_t st)
{
    if (!nonempty(d.leafs)) return 0;
    k::power_sum<<<d.leafs.n, 128, 0, st>>>(view(d), d.leafs.d, X, order, int(volume_weighted), cell_types, vsum, wsum);
    return d.leafs.n;
}

int launch_diag_check(const DevGrid& d, Coef C, double* partial, hipStream_t st)
{
    if (!nonempty(d.leafs)) return 0;
    k::diag_check<256><<<d.leafs.n, 256, 0, st>>>(view(d), d.leafs.rec, C, partial);
    return d.leafs.n;
}

bool derive_allowed() { return !(persistent_mask() & 8); }
bool fuse_first_sweep_allowed() { return !(persistent_mask() & 16); }
bool fuse_fas_save_allowed() { return !(persistent_mask() & (16 | PERSIST_JACOBI)); }
bool fuse_rz_allowed() { return !(persistent_mask() & (16 | PERSIST_JACOBI)); }
bool pull_prolong_allowed() { return !(persistent_mask() & (16 | 32 | PERSIST_JACOBI)); }

// the coded coefficients are read by the one-tile kernels only
bool coef_codes_allowed() { return !(persistent_mask() & (64 | PERSIST_APPLY | PERSIST_JACOBI | PERSIST_RESTRICT)); }

void launch_learned_accept(const DevGrid& d, int level, const float* Y, float* X, Coef C, const float* B, float w,
                           const double* S, hipStream_t st, unsigned int* count)
{
    const LaunchList& l = d.subtree[level];
    if (!nonempty(l)) return;
    // ghosts of the level are 0 at the start of a cycle: X is zero-filled there
    k::learned_accept<256><<<l.n, 256, 0, st>>>(view(d), l.rec, Y, X, X, C, B, w, run_flag(d, S), count);
}

}  // namespace cmg
