// cmg_capi.cpp -- extern "C" boundary (include/cirrus_amg.h).
#include <algorithm>
#include <array>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "cmg_internal.hpp"
#include "cmg_kernels.hpp"

namespace cmg {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace cmg

using namespace cmg;

#define CAPI_GUARD(...)                                    \
    try {                                                  \
        __VA_ARGS__;                                       \
    } catch (const std::invalid_argument& ex) {            \
        set_error(ex.what());                              \
        return CMG_ERR_BAD_ARG;                            \
    } catch (const std::exception& ex) {                   \
        set_error(ex.what());                              \
        std::string w = ex.what();                         \
        if (w.find("hash table full") != std::string::npos) return CMG_ERR_HASH_FULL; \
        return CMG_ERR_HIP;                                \
    }

namespace {

cmg_status need_topology(const cmg_grid* g)
{
    if (!g || !g->topo) return CMG_ERR_BAD_ARG;
    return CMG_OK;
}

cmg_status need_final(const cmg_grid* g)
{
    if (!g || !g->topo) return CMG_ERR_BAD_ARG;
    if (!g->topo->finalized()) {
        set_error("grid is not finalized (call cmg_grid_finalize)");
        return CMG_ERR_STATE;
    }
    return CMG_OK;
}

// host cell types -> device (only once the device copy exists)
void upload_types(cmg_grid* g)
{
    touch_types(g);
    if (!g->uploaded) return;
    auto& d = g->dev;
    if (hipMemcpy(d.ctype, g->cells.type.data(), g->cells.type.size(), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("hipMemcpy cell types failed");
    g->types_on_device = true;
}

void upload_channel(cmg_grid* g, int c)
{
    touch_channel(g, c);
    if (!g->uploaded) return;
    auto& d = g->dev;
    float* dst = dev_channel(d, c);
    const auto& src = g->cells.ch[c];
    if (hipMemcpy(dst, src.data(), src.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("hipMemcpy channel failed");
}

// a solver's own runs write its work channels 0..4 (x, r, p, Ap, z)
void touch_work(cmg_solver* s)
{
    if (s) touch_channels(solver_grid(s), 0, 5);
}

void ensure_types(cmg_grid* g)
{
    const size_t n = size_t(g->topo->num_tiles()) * kCells;
    if (g->cells.type.size() != n) g->cells.type.assign(n, 0);
    if (g->host_types_stale && g->uploaded) {
        if (hipMemcpy(g->cells.type.data(), g->dev.ctype, n, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy cell types failed");
        g->host_types_stale = false;
    }
}

}  // namespace

namespace cmg {
void ensure_device(cmg_grid* g)
{
    if (g->uploaded) return;
    dev_upload(g->dev, *g->topo);
    g->uploaded = true;
    touch_all(g);
    upload_types(g);
    for (int c = 0; c < kChannels; ++c)
        if (!g->cells.ch[c].empty()) upload_channel(g, c);
}
}  // namespace cmg

extern "C" {

cmg_status cmg_grid_create(float h0, int num_levels, const uint32_t* log2_hash, cmg_grid** out)
{
    if (!out || !log2_hash || !(h0 > 0)) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        auto* g = new cmg_grid;
        g->topo = std::make_unique<Topology>(h0, num_levels, log2_hash);
        *out = g;
    });
    return CMG_OK;
}

void cmg_grid_destroy(cmg_grid* g)
{
    if (!g) return;
    dev_free(g->dev);
    delete g;
}

cmg_status cmg_grid_set_tile(cmg_grid* g, int level, int i, int j, int k, uint8_t type)
{
    if (auto st = need_topology(g)) return st;
    if (type != CMG_LEAF && type != CMG_NONLEAF && type != CMG_GHOST) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        g->topo->set_tile(level, Coord{i, j, k}, type);
        g->uploaded = false;
    });
    return CMG_OK;
}

cmg_status cmg_grid_rebuild(cmg_grid* g)
{
    if (auto st = need_topology(g)) return st;
    CAPI_GUARD({
        g->topo->compress();
        g->topo->spawn_ghosts();
        g->uploaded = false;
    });
    return CMG_OK;
}

cmg_status cmg_grid_iterative_refine(cmg_grid* g, const cmg_target* target)
{
    if (auto st = need_topology(g)) return st;
    if (!target || (target->kind == CMG_TARGET_CALLBACK && !target->fn)) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        auto fn = make_target(*g->topo, *target);
        g->topo->iterative_refine(fn);
        g->uploaded = false;
    });
    return CMG_OK;
}

cmg_status cmg_grid_finalize(cmg_grid* g)
{
    if (auto st = need_topology(g)) return st;
    CAPI_GUARD({
        g->topo->finalize();
        dev_free(g->dev);
        ++g->generation;
        g->uploaded = false;
        g->cells = HostCells{};
        ensure_types(g);
        g->assembled_c0 = -1;
    });
    return CMG_OK;
}

cmg_status cmg_grid_rebuild_device(cmg_grid* g)
{
    if (auto st = need_topology(g)) return st;
    CAPI_GUARD({
        device_rebuild(*g->topo);
        g->uploaded = false;
    });
    return CMG_OK;
}

cmg_status cmg_grid_iterative_refine_device(cmg_grid* g, const cmg_target* target)
{
    if (auto st = need_topology(g)) return st;
    if (!target || (target->kind == CMG_TARGET_CALLBACK && !target->fn)) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        device_iterative_refine(*g->topo, *target);
        g->uploaded = false;
    });
    return CMG_OK;
}

cmg_status cmg_grid_build_device(cmg_grid* g, const cmg_target* target)
{
    if (auto st = need_topology(g)) return st;
    if (!target || (target->kind == CMG_TARGET_CALLBACK && !target->fn)) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        device_build(*g->topo, *target);
        dev_free(g->dev);
        ++g->generation;
        g->uploaded = false;
        g->cells = HostCells{};
        ensure_types(g);
        g->assembled_c0 = -1;
    });
    return CMG_OK;
}

cmg_status cmg_grid_finalize_device(cmg_grid* g)
{
    if (auto st = need_topology(g)) return st;
    CAPI_GUARD({
        device_finalize(*g->topo);
        dev_free(g->dev);
        ++g->generation;
        g->uploaded = false;
        g->cells = HostCells{};
        ensure_types(g);
        g->assembled_c0 = -1;
    });
    return CMG_OK;
}

int cmg_grid_num_tiles(const cmg_grid* g) { return (g && g->topo && g->topo->finalized()) ? g->topo->num_tiles() : 0; }
int cmg_grid_max_level(const cmg_grid* g) { return (g && g->topo) ? g->topo->max_level() : -1; }

cmg_status cmg_grid_level_starts(const cmg_grid* g, int32_t* starts)
{
    if (auto st = need_final(g)) return st;
    if (!starts) return CMG_ERR_BAD_ARG;
    std::memcpy(starts, g->topo->level_start.data(), g->topo->level_start.size() * sizeof(int32_t));
    return CMG_OK;
}

cmg_status cmg_grid_get_tiles(const cmg_grid* g, int32_t* level, int32_t* coord, uint8_t* type)
{
    if (auto st = need_final(g)) return st;
    const auto& T = *g->topo;
    for (int t = 0; t < T.num_tiles(); ++t) {
        if (level) level[t] = T.t_level[t];
        if (coord) for (int a = 0; a < 3; ++a) coord[3 * t + a] = T.t_coord[t][a];
        if (type) type[t] = T.t_type[t];
    }
    return CMG_OK;
}

cmg_status cmg_grid_get_slots(const cmg_grid* g, int32_t* slot)
{
    if (auto st = need_final(g)) return st;
    if (!slot) return CMG_ERR_BAD_ARG;
    std::memcpy(slot, g->topo->t_slot.data(), g->topo->t_slot.size() * sizeof(int32_t));
    return CMG_OK;
}

cmg_status cmg_grid_get_neighbors(const cmg_grid* g, int32_t* nbr)
{
    if (auto st = need_final(g)) return st;
    if (!nbr) return CMG_ERR_BAD_ARG;
    std::memcpy(nbr, g->topo->nbr.data(), g->topo->nbr.size() * sizeof(int32_t));
    return CMG_OK;
}

cmg_status cmg_grid_get_parents(const cmg_grid* g, int32_t* parent)
{
    if (auto st = need_final(g)) return st;
    if (!parent) return CMG_ERR_BAD_ARG;
    std::memcpy(parent, g->topo->parent.data(), g->topo->parent.size() * sizeof(int32_t));
    return CMG_OK;
}

cmg_status cmg_grid_get_children(const cmg_grid* g, int32_t* child)
{
    if (auto st = need_final(g)) return st;
    if (!child) return CMG_ERR_BAD_ARG;
    std::memcpy(child, g->topo->child.data(), g->topo->child.size() * sizeof(int32_t));
    return CMG_OK;
}

cmg_status cmg_grid_set_cell_types(cmg_grid* g, const uint8_t* types)
{
    if (auto st = need_final(g)) return st;
    if (!types) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_types(g);
        std::memcpy(g->cells.type.data(), types, g->cells.type.size());
        upload_types(g);
    });
    return CMG_OK;
}

cmg_status cmg_grid_get_cell_types(cmg_grid* g, uint8_t* types)
{
    if (auto st = need_final(g)) return st;
    if (!types) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_types(g);
        std::memcpy(types, g->cells.type.data(), g->cells.type.size());
    });
    return CMG_OK;
}

cmg_status cmg_grid_get_host_channel(cmg_grid* g, int channel, float* host)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels || !host) return CMG_ERR_BAD_ARG;
    const auto& c = g->cells.ch[channel];
    if (c.empty()) {
        set_error("channel has no host-side assembly data");
        return CMG_ERR_STATE;
    }
    std::memcpy(host, c.data(), c.size() * sizeof(float));
    return CMG_OK;
}

cmg_status cmg_grid_set_leaf_types_preset(cmg_grid* g, const cmg_target* target)
{
    if (auto st = need_final(g)) return st;
    if (!target) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        set_leaf_types_preset(*g->topo, g->cells, *target);
        g->host_types_stale = false;   // every type was just rewritten on the host
        upload_types(g);
    });
    return CMG_OK;
}

cmg_status cmg_grid_set_wall_types(cmg_grid* g, int chk_level, int dirichlet_ytop)
{
    if (auto st = need_final(g)) return st;
    if (chk_level < 0) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_types(g);
        set_wall_types(*g->topo, g->cells, chk_level, dirichlet_ytop != 0);
        upload_types(g);
    });
    return CMG_OK;
}

cmg_status cmg_grid_set_channel(cmg_grid* g, int channel, const float* host)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels || !host) return CMG_ERR_BAD_ARG;
    touch_channel(g, channel);
    CAPI_GUARD({
        ensure_device(g);
        float* dst = dev_channel(g->dev, channel);
        if (hipMemcpy(dst, host, size_t(g->dev.N) * kCells * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy set channel failed");
    });
    return CMG_OK;
}

cmg_status cmg_grid_get_channel(cmg_grid* g, int channel, float* host)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels || !host) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_device(g);
        float* src = dev_channel(g->dev, channel);
        if (hipMemcpy(host, src, size_t(g->dev.N) * kCells * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy get channel failed");
    });
    return CMG_OK;
}

cmg_status cmg_grid_fill_channel(cmg_grid* g, int channel, float value)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels) return CMG_ERR_BAD_ARG;
    touch_channel(g, channel);
    CAPI_GUARD({
        ensure_device(g);
        auto& d = g->dev;
        launch_fill_all(d, dev_channel(d, channel), value, nullptr, d.stream);
        if (hipStreamSynchronize(d.stream) != hipSuccess) throw std::runtime_error("fill failed");
    });
    return CMG_OK;
}

cmg_status cmg_grid_copy_channel(cmg_grid* g, int src, int dst)
{
    touch_channel(g, dst);
    return op_copy(g, src, dst);
}

// ---------------------------------------------------------- device-resident exchange

cmg_status cmg_grid_set_channel_device(cmg_grid* g, int channel, const float* d_src)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels || !d_src) return CMG_ERR_BAD_ARG;
    touch_channel(g, channel);
    CAPI_GUARD({
        ensure_device(g);
        auto& d = g->dev;
        float* dst = dev_channel(d, channel);
        if (hipMemcpyAsync(dst, d_src, size_t(d.N) * kCells * sizeof(float), hipMemcpyDeviceToDevice, d.stream) != hipSuccess ||
            hipStreamSynchronize(d.stream) != hipSuccess)
            throw std::runtime_error("device channel copy failed");
        g->cells.ch[channel].clear();   // a host assembly copy of this channel is stale now
    });
    return CMG_OK;
}

cmg_status cmg_grid_get_channel_device(cmg_grid* g, int channel, float* d_dst)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels || !d_dst) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_device(g);
        auto& d = g->dev;
        const float* src = dev_channel(d, channel);
        if (hipMemcpyAsync(d_dst, src, size_t(d.N) * kCells * sizeof(float), hipMemcpyDeviceToDevice, d.stream) != hipSuccess ||
            hipStreamSynchronize(d.stream) != hipSuccess)
            throw std::runtime_error("device channel copy failed");
    });
    return CMG_OK;
}

namespace {

// exactly one of the address array and the slab base, a usable stride for a slab
cmg_status check_tiles_arg(const void* addrs, const void* base, size_t stride)
{
    if ((addrs == nullptr) == (base == nullptr)) {
        set_error("give either the device array of tile addresses or the slab base, not both");
        return CMG_ERR_BAD_ARG;
    }
    if (base && (stride < 44880 || stride % 16 != 0)) {
        set_error("slab stride must be a multiple of 16 and at least sizeof(PoissonTile<float,15>) = 44880");
        return CMG_ERR_BAD_ARG;
    }
    return CMG_OK;
}

cmg_status report_bad_tiles(unsigned int bad)
{
    if (!bad) return CMG_OK;
    set_error(std::to_string(bad) + " tile address(es) null or not 16-byte aligned; those tiles were skipped");
    return CMG_ERR_BAD_ARG;
}

}  // namespace

cmg_status cmg_grid_gather_tile_channel(cmg_grid* g, const uint64_t* d_tile_addrs, const void* d_base, size_t stride,
                                        int tile_channel, int channel)
{
    if (auto st = need_final(g)) return st;
    if (auto st = check_tiles_arg(d_tile_addrs, d_base, stride)) return st;
    if (channel < 0 || channel >= kChannels || tile_channel < 0 || tile_channel >= 15) return CMG_ERR_BAD_ARG;
    touch_channel(g, channel);
    unsigned int bad = 0;
    CAPI_GUARD({
        ensure_device(g);
        auto& d = g->dev;
        bad = tile_channel_io(d, true, d_tile_addrs, const_cast<void*>(d_base), stride, tile_channel, dev_channel(d, channel));
        g->cells.ch[channel].clear();
    });
    return report_bad_tiles(bad);
}

cmg_status cmg_grid_scatter_tile_channel(cmg_grid* g, int channel, const uint64_t* d_tile_addrs, void* d_base, size_t stride,
                                         int tile_channel)
{
    if (auto st = need_final(g)) return st;
    if (auto st = check_tiles_arg(d_tile_addrs, d_base, stride)) return st;
    if (channel < 0 || channel >= kChannels || tile_channel < 0 || tile_channel >= 15) return CMG_ERR_BAD_ARG;
    unsigned int bad = 0;
    CAPI_GUARD({
        ensure_device(g);
        auto& d = g->dev;
        bad = tile_channel_io(d, false, d_tile_addrs, d_base, stride, tile_channel, dev_channel(d, channel));
    });
    return report_bad_tiles(bad);
}

cmg_status cmg_grid_gather_tile_types(cmg_grid* g, const uint64_t* d_tile_addrs, const void* d_base, size_t stride)
{
    if (auto st = need_final(g)) return st;
    if (auto st = check_tiles_arg(d_tile_addrs, d_base, stride)) return st;
    touch_types(g);
    unsigned int bad = 0;
    CAPI_GUARD({
        ensure_types(g);
        ensure_device(g);
        bad = tile_types_io(g->dev, true, d_tile_addrs, const_cast<void*>(d_base), stride);
        g->types_on_device = true;
        g->host_types_stale = true;   // copied back only when a host path reads the types
    });
    return report_bad_tiles(bad);
}

cmg_status cmg_grid_scatter_tile_types(cmg_grid* g, const uint64_t* d_tile_addrs, void* d_base, size_t stride)
{
    if (auto st = need_final(g)) return st;
    if (auto st = check_tiles_arg(d_tile_addrs, d_base, stride)) return st;
    unsigned int bad = 0;
    CAPI_GUARD({
        ensure_device(g);
        if (!g->types_on_device) upload_types(g);
        bad = tile_types_io(g->dev, false, d_tile_addrs, d_base, stride);
    });
    return report_bad_tiles(bad);
}

// ---------------------------------------------------------- region of interest

cmg_status cmg_mark_region_of_interest(cmg_grid* g, int subtree_level, uint8_t launch_types, int channel, int pred, float a,
                                       float b, int calc_locked)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels || pred < CMG_ROI_MAX_GT || pred > CMG_ROI_SPAN_GT) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_device(g);
        auto& d = g->dev;
        if (!g->types_on_device) upload_types(g);
        const size_t N = size_t(std::max(d.N, 1));
        auto zalloc = [&](auto*& p, size_t bytes) {
            if (p) return;
            if (hipMalloc(reinterpret_cast<void**>(&p), bytes) != hipSuccess || hipMemset(p, 0, bytes) != hipSuccess)
                throw std::runtime_error("hipMalloc roi flags failed");
        };
        zalloc(d.roi, N);
        zalloc(d.roi_locked, N);
        zalloc(d.roi_min, N * sizeof(float));
        zalloc(d.roi_max, N * sizeof(float));
        launch_roi_flags(d, dev_channel(d, channel), subtree_level, launch_types, pred, a, b, calc_locked != 0, d.roi,
                         d.roi_locked, d.roi_min, d.roi_max, d.stream);
        if (hipStreamSynchronize(d.stream) != hipSuccess) throw std::runtime_error("roi flags failed");
    });
    return CMG_OK;
}

cmg_status cmg_grid_get_roi(cmg_grid* g, uint8_t* interest, uint8_t* locked, float* tile_min, float* tile_max)
{
    if (auto st = need_final(g)) return st;
    auto& d = g->dev;
    if (!g->uploaded || !d.roi) {
        set_error("no region-of-interest marking on this grid yet");
        return CMG_ERR_STATE;
    }
    CAPI_GUARD({
        const size_t N = size_t(d.N);
        auto get = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes && hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
                throw std::runtime_error("hipMemcpy roi flags failed");
        };
        get(interest, d.roi, N);
        get(locked, d.roi_locked, N);
        get(tile_min, d.roi_min, N * sizeof(float));
        get(tile_max, d.roi_max, N * sizeof(float));
    });
    return CMG_OK;
}

cmg_status cmg_prepare_types_and_coeffs(cmg_grid* g, int c0, float R)
{
    if (auto st = need_final(g)) return st;
    if (c0 < 0 || c0 + 3 >= kChannels) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_types(g);
        assemble(*g->topo, g->cells, c0, R, -1, 0, nullptr);
        upload_types(g);
        for (int a = 0; a < 4; ++a) upload_channel(g, c0 + a);
        g->assembled_c0 = c0;
    });
    return CMG_OK;
}

cmg_status cmg_assemble_cutcell(cmg_grid* g, int sdf_kind, const float* sdf, int face_rule, int c0, float R)
{
    if (auto st = need_final(g)) return st;
    if (!sdf || c0 < 0 || c0 + 3 >= kChannels || (face_rule < -1 || face_rule > 1)) return CMG_ERR_BAD_ARG;
    if (sdf_kind != CMG_SDF_SPHERE && sdf_kind != CMG_SDF_TORUS && sdf_kind != CMG_SDF_STAR) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_types(g);
        set_solid_types(*g->topo, g->cells, sdf_kind, sdf);
        assemble(*g->topo, g->cells, c0, R, face_rule, sdf_kind, sdf);
        upload_types(g);
        for (int a = 0; a < 4; ++a) upload_channel(g, c0 + a);
        g->assembled_c0 = c0;
    });
    return CMG_OK;
}

cmg_status cmg_assemble_cutcell_device(cmg_grid* g, int sdf_kind, const float* sdf, int face_rule, int c0, float R,
                                      int wall_chk_level, int dirichlet_ytop)
{
    if (auto st = need_final(g)) return st;
    if (c0 < 0 || c0 + 3 >= kChannels || face_rule < -1 || face_rule > 1) return CMG_ERR_BAD_ARG;
    if (sdf && sdf_kind != CMG_SDF_SPHERE && sdf_kind != CMG_SDF_TORUS) return CMG_ERR_BAD_ARG;
    if (!sdf && face_rule >= 0) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_types(g);
        ensure_device(g);
        DevGrid& d = g->dev;
        if (!g->types_on_device) upload_types(g);
        if (wall_chk_level >= 0) launch_assemble_walls(d, wall_chk_level, dirichlet_ytop != 0, d.stream);
        if (sdf) launch_assemble_solid(d, sdf_kind, sdf, d.stream);
        launch_assemble_operator(d, c0, R, face_rule, sdf_kind, sdf, d.acount, d.stream);
        // the host copy of the types follows the device; the host coefficient data is stale
        if (hipMemcpy(g->cells.type.data(), d.ctype, g->cells.type.size(), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy cell types failed");
        for (int a = 0; a < 4; ++a) g->cells.ch[c0 + a].clear();
        g->types_on_device = true;
        g->assembled_c0 = c0;
        touch_types(g);
        touch_channels(g, c0, 4);
    });
    return CMG_OK;
}

cmg_status cmg_rhs_uniform(cmg_grid* g, int channel, uint64_t seed, int zero_mean)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_types(g);
        rhs_uniform(*g->topo, g->cells, channel, seed, zero_mean != 0);
        upload_channel(g, channel);
    });
    return CMG_OK;
}

cmg_status cmg_rhs_divergence(cmg_grid* g, int channel, int u0, uint64_t seed)
{
    if (auto st = need_final(g)) return st;
    if (channel < 0 || channel >= kChannels || u0 < 0 || u0 + 2 >= kChannels) return CMG_ERR_BAD_ARG;
    if (channel >= u0 && channel <= u0 + 2) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        ensure_types(g);
        rhs_divergence(*g->topo, g->cells, channel, u0, seed);
        upload_channel(g, channel);
        for (int a = 0; a < 3; ++a) upload_channel(g, u0 + a);
    });
    return CMG_OK;
}

void cmg_solve_opts_default(cmg_solve_opts* o)
{
    if (!o) return;
    *o = cmg_solve_opts{};
    // FluidEuler::project: solve(grid, false, 100, 1e-6, 2, 10, 1, pure_neumann)
    o->verbose = 0;
    o->max_iters = 100;
    o->relative_tolerance = 1e-6;
    o->level_iters = 2;
    o->coarsest_iters = 10;
    o->sync_stride = 1;
    o->is_pure_neumann = 0;
    o->cycle = CMG_CYCLE_FAS_W;
    o->smoother = CMG_SMOOTH_RBGS;
    o->omega = 1.5;
    o->coarsest_direct = 0;
    o->precision = CMG_PREC_FP32;
    o->outer_tolerance = 0;
    o->inner_tolerance = 0;
    o->warm_start = 0;
    o->use_graph = 1;
}

cmg_status cmg_solver_create(cmg_grid* g, int c0, float R_matrix, float R_restrict, float P, cmg_solver** out)
{
    return solver_create(g, c0, R_matrix, R_restrict, P, out);
}

void cmg_solver_destroy(cmg_solver* s) { solver_destroy(s); }

cmg_status cmg_geo_solver_create(cmg_grid* g, int family, float one_over_alpha, float prolong_coeff, cmg_solver** out)
{
    return solver_create_geo(g, family, one_over_alpha, prolong_coeff, out);
}

cmg_status cmg_gmg_damped_jacobi_solve(cmg_solver* s, const cmg_solve_opts* opts, cmg_solve_stats* stats)
{
    touch_work(s);
    return solver_damped_jacobi(s, opts, stats);
}

cmg_status cmg_solver_apply(cmg_solver* s, int x, int y)
{
    if (s) {
        touch_channel(solver_grid(s), x);
        touch_channel(solver_grid(s), y);
    }
    return solver_apply(s, x, y);
}

cmg_status cmg_solver_set_net(cmg_solver* s, const float* weights) { return solver_set_net(s, weights); }
cmg_status cmg_solver_net_infer(cmg_solver* s, int r, int diag, int y)
{
    if (s) touch_channel(solver_grid(s), y);
    return solver_net_infer(s, r, diag, y);
}
cmg_status cmg_trainer_create(int capacity_tiles, uint64_t seed, cmg_trainer** out)
{
    return trainer_create(capacity_tiles, seed, out);
}
void cmg_trainer_destroy(cmg_trainer* t) { trainer_destroy(t); }
cmg_status cmg_trainer_collect(cmg_trainer* t, cmg_grid* g, int r, int diag, int e, int* appended)
{
    return trainer_collect(t, g, r, diag, e, appended);
}
cmg_status cmg_trainer_add_tiles(cmg_trainer* t, const float* x, const float* e, int n) { return trainer_add_tiles(t, x, e, n); }
int cmg_trainer_size(const cmg_trainer* t) { return trainer_size(t); }
cmg_status cmg_trainer_get_weights(cmg_trainer* t, float* w) { return trainer_weights(t, w, nullptr); }
cmg_status cmg_trainer_set_weights(cmg_trainer* t, const float* w) { return trainer_weights(t, nullptr, w); }
cmg_status cmg_trainer_loss_grad(cmg_trainer* t, int first, int n, int bf16, double* loss, float* grad)
{
    return trainer_loss_grad(t, first, n, bf16, loss, grad);
}
cmg_status cmg_trainer_step(cmg_trainer* t, int batch, float lr, int bf16, double* loss)
{
    return trainer_step(t, batch, lr, bf16, loss);
}
cmg_status cmg_solver_learned_counts(cmg_solver* s, uint64_t* judged, uint64_t* kept)
{
    return solver_learned_counts(s, judged, kept);
}
cmg_status cmg_mfma_probe(int k, const uint16_t* A, const uint16_t* B, float* C) { return mfma_probe(k, A, B, C); }

cmg_status cmg_solve(cmg_solver* s, const cmg_solve_opts* opts, cmg_solve_stats* stats)
{
    touch_work(s);
    return solver_solve(s, opts, stats);
}

cmg_status cmg_solver_get_x64(cmg_solver* s, double* host) { return solver_get_x64(s, host); }
cmg_status cmg_solver_coded_levels(cmg_solver* s, uint32_t* mask, uint32_t* wide) { return solver_coded_levels(s, mask, wide); }

cmg_status cmg_op_full_neg_lap(cmg_grid* g, int x, int c0, int y)
{
    touch_channel(g, x);   // propagated ghosts and accumulated parents are left in x
    touch_channel(g, y);
    return op_full_neg_lap(g, x, c0, y);
}
cmg_status cmg_op_propagate_ghosts(cmg_grid* g, int ch)
{
    touch_channel(g, ch);
    return op_propagate_ghosts(g, ch);
}
cmg_status cmg_op_accumulate_leafs(cmg_grid* g, int ch)
{
    touch_channel(g, ch);
    return op_accumulate_leafs(g, ch);
}
cmg_status cmg_op_gs_color(cmg_grid* g, int level, int x, int c0, int rhs, int color, double omega)
{
    touch_channel(g, x);
    return op_gs_color(g, level, x, c0, rhs, color, omega);
}
cmg_status cmg_op_jacobi(cmg_grid* g, int level, int x_in, int x_out, int gch, int c0, int rhs, float omega)
{
    touch_channel(g, x_in);   // ghost values may be refreshed in place
    touch_channel(g, x_out);
    return op_jacobi(g, level, x_in, x_out, gch, c0, rhs, omega);
}
cmg_status cmg_op_residual_restrict(cmg_grid* g, int level, int x, int c0, int rhs, float Rr)
{
    touch_channel(g, x);
    touch_channel(g, rhs);
    return op_residual_restrict(g, level, x, c0, rhs, Rr);
}
cmg_status cmg_op_fas_save(cmg_grid* g, int level, int x, int c0, int rhs, int x0)
{
    touch_channel(g, rhs);
    touch_channel(g, x0);
    return op_fas_save(g, level, x, c0, rhs, x0);
}
cmg_status cmg_op_fas_prolong(cmg_grid* g, int level, int x, int x0, float P)
{
    touch_channel(g, x);
    return op_fas_prolong(g, level, x, x0, P);
}
cmg_status cmg_op_dot(cmg_grid* g, int a, int b, double* out) { return op_dot(g, a, b, out); }
cmg_status cmg_op_norm(cmg_grid* g, int order, int channel, int volume_weighted, int cell_types, double* out)
{
    return op_norm(g, order, channel, volume_weighted, cell_types, out);
}
cmg_status cmg_op_norm_noalloc(cmg_grid* g, int order, int channel, int volume_weighted, int cell_types, double* out)
{
    return op_norm_noalloc(g, order, channel, volume_weighted, cell_types, out);
}
cmg_status cmg_divergence(cmg_grid* g, int u0, int channel, int coeff_channel)
{
    touch_channels(g, u0, 3);   // faces accumulated to parents
    touch_channel(g, channel);
    return op_divergence(g, u0, channel, coeff_channel);
}
cmg_status cmg_add_gradient(cmg_grid* g, int x, int coeff_channel, int u0, int weighted)
{
    touch_channel(g, x);
    touch_channels(g, u0, 3);
    return op_add_gradient(g, x, coeff_channel, u0, weighted);
}
cmg_status cmg_mix_velocity(cmg_grid* g, int coeff_channel, int u_fluid, int u_solid, int u_mix)
{
    touch_channels(g, u_mix, 3);
    return op_mix_velocity(g, coeff_channel, u_fluid, u_solid, u_mix);
}
cmg_status cmg_op_cycle(cmg_solver* s, const cmg_solve_opts* opts, int x, int rhs, int x0)
{
    if (s) {
        touch_channel(solver_grid(s), x);
        touch_channel(solver_grid(s), rhs);
        touch_channel(solver_grid(s), x0);
    }
    return solver_cycle(s, opts, x, rhs, x0);
}

cmg_status cmg_mg_solve(cmg_solver* s, const cmg_solve_opts* opts, cmg_solve_stats* stats)
{
    touch_work(s);
    return solver_mg_solve(s, opts, stats);
}

cmg_status cmg_bench_kernel(cmg_solver* s, int kernel, int reps, double* avg_ms, int* tiles)
{
    touch_work(s);
    return bench_kernel(s, kernel, reps, avg_ms, tiles);
}

cmg_status cmg_set_kernel_forms(int persistent_mask, int stencil_threads)
{
    if (!set_kernel_forms(persistent_mask, stencil_threads)) {
        set_error("kernel forms: mask must be 0..127 and threads 0, 128, 256 or 512");
        return CMG_ERR_BAD_ARG;
    }
    return CMG_OK;
}

cmg_status cmg_get_kernel_forms(int* persistent_mask, int* stencil_threads)
{
    if (!persistent_mask || !stencil_threads) return CMG_ERR_BAD_ARG;
    get_kernel_forms(persistent_mask, stencil_threads);
    return CMG_OK;
}

cmg_status cmg_device_synchronize(void)
{
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("hipDeviceSynchronize failed");
        return CMG_ERR_HIP;
    }
    return CMG_OK;
}

const char* cmg_status_string(cmg_status st)
{
    switch (st) {
    case CMG_OK: return "ok";
    case CMG_ERR_BAD_ARG: return "bad argument";
    case CMG_ERR_STATE: return "call out of order";
    case CMG_ERR_HIP: return "HIP runtime error";
    case CMG_ERR_NONFINITE: return "non-finite value";
    case CMG_ERR_BREAKDOWN: return "CG breakdown";
    case CMG_ERR_NOT_CONVERGED: return "not converged";
    case CMG_ERR_HASH_FULL: return "hash table full";
    }
    return "unknown";
}

const char* cmg_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

// ---------------------------------------------------------------- HAGR blobs ----
// HADeviceGrid::dumpBinaryBlob / loadBinaryBlob (src/HAGrid.h:687-853) for
// PoissonTile<float, 15>: a 40-byte header {u32 magic 'HAGR', u32 version 1,
// u32 num_levels, i32 max_level, u32 tile_size 44880, u32 coord_size 12, f32 h0,
// (4 bytes padding), u64 tile_count}, num_levels u32 log2 hash sizes, then per
// tile (all-tiles order) u32 level, u8 type, 3 x i32 tile coordinate and the raw
// tile: float data[15][736] (cell v at [c][v]; entries 512..735 are node values,
// 0 here), 6 neighbour records of 32 bytes (coordinate, pointer, level, type; the
// reader ignores them), u8 cell types[512], two flags, i32 status, i32 serial.
namespace {
constexpr uint32_t kHagrMagic = 0x52474148u;
constexpr size_t kHagrHeader = 40, kTileBytes = 44880, kChanStride = 736, kNbrOff = 15 * 736 * 4, kTypeOff = kNbrOff + 6 * 32;
constexpr size_t kRecord = 4 + 1 + 12 + kTileBytes;
static_assert(kTypeOff + 512 + 16 == kTileBytes, "PoissonTile<float, 15> layout");

template <typename V>
void put(uint8_t*& p, V v)
{
    std::memcpy(p, &v, sizeof(V));
    p += sizeof(V);
}
template <typename V>
V get(const uint8_t* p)
{
    V v;
    std::memcpy(&v, p, sizeof(V));
    return v;
}
}  // namespace

extern "C" {

cmg_status cmg_grid_dump_blob(cmg_grid* g, uint8_t tile_types, int max_level, uint8_t* out, size_t* size)
{
    if (auto st = need_final(g)) return st;
    if (!size || !(tile_types & (CMG_LEAF | CMG_NONLEAF | CMG_GHOST))) return CMG_ERR_BAD_ARG;
    CAPI_GUARD({
        const Topology& t = *g->topo;
        const int ml = max_level < 0 ? t.max_level() : std::min(max_level, t.max_level());
        const int N = t.num_tiles();
        std::vector<int> sel;
        for (int i = 0; i < N; ++i)
            if ((t.t_type[i] & tile_types) && t.t_level[i] <= ml) sel.push_back(i);
        const size_t total = kHagrHeader + 4 * size_t(t.num_levels()) + sel.size() * kRecord;
        if (!out) {
            *size = total;
            return CMG_OK;
        }
        if (*size < total) throw std::invalid_argument("dump_blob: buffer smaller than the blob");
        // channel values: the device copy where it exists, else the host assembly data
        ensure_types(g);
        std::array<std::vector<float>, kChannels> ch;
        for (int c = 0; c < kChannels; ++c) {
            if (g->uploaded && g->dev.ch[c]) {
                ch[c].resize(size_t(N) * kCells);
                if (hipMemcpy(ch[c].data(), g->dev.ch[c], ch[c].size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
                    throw std::runtime_error("hipMemcpy dump channel failed");
            } else if (!g->cells.ch[c].empty()) {
                ch[c] = g->cells.ch[c];
            }
        }
        uint8_t* p = out;
        put<uint32_t>(p, kHagrMagic);
        put<uint32_t>(p, 1u);
        put<uint32_t>(p, uint32_t(t.num_levels()));
        put<int32_t>(p, ml);
        put<uint32_t>(p, uint32_t(kTileBytes));
        put<uint32_t>(p, 12u);
        put<float>(p, t.h0());
        put<uint32_t>(p, 0u);   // padding before the 8-byte count
        put<uint64_t>(p, uint64_t(sel.size()));
        for (int l = 0; l < t.num_levels(); ++l) put<uint32_t>(p, t.log2_hash(l));
        for (int i : sel) {
            put<uint32_t>(p, uint32_t(t.t_level[i]));
            put<uint8_t>(p, t.t_type[i]);
            for (int a = 0; a < 3; ++a) put<int32_t>(p, t.t_coord[i][a]);
            uint8_t* tile = p;
            std::memset(tile, 0, kTileBytes);
            for (int c = 0; c < kChannels; ++c)
                if (!ch[c].empty())
                    std::memcpy(tile + size_t(c) * kChanStride * 4, ch[c].data() + size_t(i) * kCells, kCells * sizeof(float));
            for (int f = 0; f < 6; ++f) {
                const int nb = t.nbr[size_t(i) * 6 + f];
                if (nb < 0) continue;
                uint8_t* r = tile + kNbrOff + size_t(f) * 32;
                for (int a = 0; a < 3; ++a) std::memcpy(r + 4 * a, &t.t_coord[nb][a], 4);
                const int32_t lv = t.t_level[nb];
                std::memcpy(r + 24, &lv, 4);
                r[28] = t.t_type[nb];
            }
            std::memcpy(tile + kTypeOff, g->cells.type.data() + size_t(i) * kCells, kCells);
            p += kTileBytes;
        }
        *size = total;
    });
    return CMG_OK;
}

cmg_status cmg_grid_load_blob(const uint8_t* data, size_t size, cmg_grid** out)
{
    if (!data || !out) return CMG_ERR_BAD_ARG;
    *out = nullptr;
    CAPI_GUARD({
        std::unique_ptr<cmg_grid> g;
        if (size < kHagrHeader) throw std::invalid_argument("load_blob: buffer too small");
        if (get<uint32_t>(data) != kHagrMagic) throw std::invalid_argument("load_blob: bad magic");
        if (get<uint32_t>(data + 4) != 1u) throw std::invalid_argument("load_blob: unsupported version");
        const uint32_t nl = get<uint32_t>(data + 8);
        if (get<uint32_t>(data + 16) != kTileBytes) throw std::invalid_argument("load_blob: tile size is not PoissonTile<float, 15>");
        if (get<uint32_t>(data + 20) != 12u) throw std::invalid_argument("load_blob: coordinate size mismatch");
        if (nl == 0 || nl > uint32_t(kMaxLevels)) throw std::invalid_argument("load_blob: num_levels invalid");
        const float h0 = get<float>(data + 24);
        const uint64_t count = get<uint64_t>(data + 32);
        if (!(h0 > 0)) throw std::invalid_argument("load_blob: h0 invalid");
        const size_t need = kHagrHeader + 4 * size_t(nl);
        if (size < need || (size - need) / kRecord < count) throw std::invalid_argument("load_blob: buffer truncated");
        std::vector<uint32_t> log2(nl);
        for (uint32_t l = 0; l < nl; ++l) log2[l] = get<uint32_t>(data + kHagrHeader + 4 * l);
        g.reset(new cmg_grid);
        g->topo = std::make_unique<Topology>(h0, int(nl), log2.data());
        const uint8_t* rec = data + need;
        for (uint64_t r = 0; r < count; ++r, rec += kRecord) {
            const uint32_t lv = get<uint32_t>(rec);
            const uint8_t ty = rec[4];
            if (lv >= nl || (ty != CMG_LEAF && ty != CMG_NONLEAF && ty != CMG_GHOST))
                throw std::invalid_argument("load_blob: bad tile record");
            g->topo->set_tile(int(lv), Coord{get<int32_t>(rec + 5), get<int32_t>(rec + 9), get<int32_t>(rec + 13)}, ty);
        }
        g->topo->compress();   // the blob's ghosts are tiles of their own: none is spawned
        g->topo->finalize();
        ensure_types(g.get());
        const int N = g->topo->num_tiles();
        rec = data + need;
        std::array<bool, kChannels> used{};
        std::array<std::vector<float>, kChannels> ch;
        for (int c = 0; c < kChannels; ++c) ch[c].assign(size_t(N) * kCells, 0.0f);
        for (uint64_t r = 0; r < count; ++r, rec += kRecord) {
            const int i = g->topo->tile_at(int(get<uint32_t>(rec)), Coord{get<int32_t>(rec + 5), get<int32_t>(rec + 9), get<int32_t>(rec + 13)});
            if (i < 0) throw std::invalid_argument("load_blob: a record's tile is missing after compression");
            const uint8_t* tile = rec + 17;
            std::memcpy(g->cells.type.data() + size_t(i) * kCells, tile + kTypeOff, kCells);
            for (int c = 0; c < kChannels; ++c) {
                float* dst = ch[c].data() + size_t(i) * kCells;
                std::memcpy(dst, tile + size_t(c) * kChanStride * 4, kCells * sizeof(float));
                for (int v = 0; v < kCells && !used[c]; ++v) used[c] = dst[v] != 0.0f;
            }
        }
        for (int c = 0; c < kChannels; ++c)
            if (used[c]) g->cells.ch[c] = std::move(ch[c]);
        *out = g.release();
    });
    return CMG_OK;
}

}  // extern "C"

