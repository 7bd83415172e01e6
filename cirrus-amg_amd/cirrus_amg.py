"""Thin ctypes tooling over libcirrus_amg.so (include/cirrus_amg.h).

The solver itself is C++17 + HIP; this module only marshals numpy arrays
across the C ABI.  Names follow the reference's classes:
``HADeviceGrid`` (src/HAGrid.h) and ``AMGSolver`` (src/AMGSolver.h).
There is no CPU fallback: every device call goes through the HIP library and
raises ``CmgError`` when it fails.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcirrus_amg.so")

LEAF, NONLEAF, GHOST = 1, 2, 4
ROI_MAX_GT, ROI_MIN_LT, ROI_RANGE_MEETS, ROI_SPAN_GT = 0, 1, 2, 3   # cmg_mark_region_of_interest predicates
DIRICHLET, INTERIOR, NEUMANN = 1, 2, 4
X_CHANNEL, B_CHANNEL, P_CHANNEL, AP_CHANNEL, Z_CHANNEL = 0, 1, 2, 3, 4   # src/PoissonTile.h:13-18

TARGET_UNIFORM, TARGET_CENTER_POINT, TARGET_SPHERE_EMPTY, TARGET_SPHERE_SOLID = 0, 1, 2, 3
TARGET_STAR_EMPTY, TARGET_STAR_SOLID, TARGET_SDF_BANDS, TARGET_CALLBACK = 4, 5, 6, 7
SDF_SPHERE, SDF_TORUS, SDF_STAR = 0, 1, 2
CYCLE_FAS_W, CYCLE_FAS_V, CYCLE_FAS_F, CYCLE_VCYCLE, CYCLE_MU = 0, 1, 2, 3, 4
SMOOTH_RBGS, SMOOTH_JACOBI = 0, 1
PREC_FP32, PREC_FP64_REFINE = 0, 1
SOLVER_AMG, SOLVER_GMG, SOLVER_CMG = 0, 1, 2

STATUS = {0: "ok", 1: "bad argument", 2: "call out of order", 3: "HIP runtime error",
          4: "non-finite value", 5: "CG breakdown", 6: "not converged", 7: "hash table full"}

LEVEL_TARGET_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int32))


class CmgError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = "%s: %s (%d)" % (where, STATUS.get(status, "?"), status)
        try:
            detail = lib().cmg_last_error().decode()
            if detail:
                msg += " -- " + detail
        except Exception:
            pass
        super().__init__(msg)


class Target(C.Structure):
    _fields_ = [("kind", C.c_int), ("min_level", C.c_int), ("max_level", C.c_int), ("sdf_kind", C.c_int),
                ("sdf", C.c_float * 8), ("band", C.c_float * 12), ("fn", LEVEL_TARGET_FN), ("user", C.c_void_p)]


class SolveOpts(C.Structure):
    _fields_ = [("verbose", C.c_int), ("max_iters", C.c_int), ("relative_tolerance", C.c_double),
                ("level_iters", C.c_int), ("coarsest_iters", C.c_int), ("sync_stride", C.c_int),
                ("is_pure_neumann", C.c_int), ("cycle", C.c_int), ("smoother", C.c_int), ("omega", C.c_double),
                ("coarsest_direct", C.c_int), ("precision", C.c_int), ("outer_tolerance", C.c_double),
                ("inner_tolerance", C.c_double), ("warm_start", C.c_int), ("use_graph", C.c_int), ("mu", C.c_int),
                ("flexible", C.c_int), ("learned", C.c_int), ("cycle_coef_fp16", C.c_int),
                ("no_true_residual", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("outer_iters", C.c_int), ("rel_res", C.c_double),
                ("true_rel_res", C.c_double), ("ms", C.c_double)]


EXPORTS = [
    "cmg_grid_create", "cmg_grid_destroy", "cmg_grid_set_tile", "cmg_grid_rebuild", "cmg_grid_iterative_refine",
    "cmg_grid_finalize", "cmg_grid_num_tiles", "cmg_grid_max_level", "cmg_grid_level_starts", "cmg_grid_get_tiles",
    "cmg_grid_get_slots", "cmg_grid_get_neighbors", "cmg_grid_get_parents", "cmg_grid_get_children",
    "cmg_grid_set_cell_types", "cmg_grid_get_cell_types", "cmg_grid_get_host_channel",
    "cmg_grid_set_leaf_types_preset", "cmg_grid_set_wall_types", "cmg_grid_set_channel", "cmg_grid_get_channel",
    "cmg_grid_fill_channel", "cmg_prepare_types_and_coeffs", "cmg_assemble_cutcell", "cmg_solve_opts_default",
    "cmg_solver_create", "cmg_solver_destroy", "cmg_solve", "cmg_solver_get_x64", "cmg_op_full_neg_lap",
    "cmg_op_propagate_ghosts", "cmg_op_accumulate_leafs", "cmg_op_gs_color", "cmg_op_jacobi",
    "cmg_op_residual_restrict", "cmg_op_fas_save", "cmg_op_fas_prolong", "cmg_op_dot", "cmg_op_cycle",
    "cmg_status_string", "cmg_last_error", "cmg_grid_copy_channel", "cmg_bench_kernel",
    "cmg_rhs_uniform", "cmg_rhs_divergence", "cmg_device_synchronize", "cmg_set_kernel_forms",
    "cmg_get_kernel_forms", "cmg_mg_solve", "cmg_geo_solver_create", "cmg_gmg_damped_jacobi_solve",
    "cmg_solver_apply", "cmg_op_norm", "cmg_solver_set_net", "cmg_solver_net_infer", "cmg_trainer_create",
    "cmg_trainer_destroy", "cmg_trainer_collect", "cmg_trainer_add_tiles", "cmg_trainer_size", "cmg_trainer_get_weights",
    "cmg_trainer_set_weights", "cmg_trainer_loss_grad", "cmg_trainer_step", "cmg_mfma_probe", "cmg_solver_learned_counts", "cmg_grid_dump_blob", "cmg_grid_load_blob", "cmg_assemble_cutcell_device", "cmg_divergence", "cmg_add_gradient", "cmg_mix_velocity",
    "cmg_grid_rebuild_device", "cmg_grid_iterative_refine_device", "cmg_grid_finalize_device",
    "cmg_grid_build_device", "cmg_op_norm_noalloc",
    "cmg_grid_set_channel_device", "cmg_grid_get_channel_device", "cmg_grid_gather_tile_channel",
    "cmg_grid_scatter_tile_channel", "cmg_grid_gather_tile_types", "cmg_grid_scatter_tile_types",
    "cmg_mark_region_of_interest", "cmg_grid_get_roi", "cmg_solver_coded_levels",
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcirrus_amg.so is not built: run `make -C cirrus-amg_amd` "
                               "or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, i, f, d, st = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_int
        sig = {
            "cmg_grid_create": (st, [f, i, vp, C.POINTER(vp)]), "cmg_grid_destroy": (None, [vp]),
            "cmg_grid_set_tile": (st, [vp, i, i, i, i, C.c_uint8]), "cmg_grid_rebuild": (st, [vp]),
            "cmg_grid_iterative_refine": (st, [vp, C.POINTER(Target)]), "cmg_grid_finalize": (st, [vp]),
            "cmg_grid_rebuild_device": (st, [vp]), "cmg_grid_finalize_device": (st, [vp]),
            "cmg_grid_iterative_refine_device": (st, [vp, C.POINTER(Target)]),
            "cmg_grid_build_device": (st, [vp, C.POINTER(Target)]),
            "cmg_grid_num_tiles": (i, [vp]), "cmg_grid_max_level": (i, [vp]),
            "cmg_grid_level_starts": (st, [vp, vp]), "cmg_grid_get_tiles": (st, [vp, vp, vp, vp]),
            "cmg_grid_get_slots": (st, [vp, vp]), "cmg_grid_get_neighbors": (st, [vp, vp]),
            "cmg_grid_get_parents": (st, [vp, vp]), "cmg_grid_get_children": (st, [vp, vp]),
            "cmg_grid_set_cell_types": (st, [vp, vp]), "cmg_grid_get_cell_types": (st, [vp, vp]),
            "cmg_grid_get_host_channel": (st, [vp, i, vp]),
            "cmg_grid_set_leaf_types_preset": (st, [vp, C.POINTER(Target)]),
            "cmg_grid_set_wall_types": (st, [vp, i, i]), "cmg_grid_set_channel": (st, [vp, i, vp]),
            "cmg_grid_get_channel": (st, [vp, i, vp]), "cmg_grid_fill_channel": (st, [vp, i, f]),
            "cmg_prepare_types_and_coeffs": (st, [vp, i, f]), "cmg_assemble_cutcell": (st, [vp, i, vp, i, i, f]),
            "cmg_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
            "cmg_solver_create": (st, [vp, i, f, f, f, C.POINTER(vp)]), "cmg_solver_destroy": (None, [vp]),
            "cmg_solve": (st, [vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
            "cmg_solver_get_x64": (st, [vp, vp]),
            "cmg_solver_coded_levels": (st, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
            "cmg_op_full_neg_lap": (st, [vp, i, i, i]), "cmg_op_propagate_ghosts": (st, [vp, i]),
            "cmg_op_accumulate_leafs": (st, [vp, i]), "cmg_op_gs_color": (st, [vp, i, i, i, i, i, d]),
            "cmg_op_jacobi": (st, [vp, i, i, i, i, i, i, f]), "cmg_op_residual_restrict": (st, [vp, i, i, i, i, f]),
            "cmg_op_fas_save": (st, [vp, i, i, i, i, i]), "cmg_op_fas_prolong": (st, [vp, i, i, i, f]),
            "cmg_op_dot": (st, [vp, i, i, C.POINTER(d)]),
            "cmg_op_cycle": (st, [vp, C.POINTER(SolveOpts), i, i, i]),
            "cmg_status_string": (C.c_char_p, [st]), "cmg_last_error": (C.c_char_p, []),
            "cmg_grid_copy_channel": (st, [vp, i, i]),
            "cmg_device_synchronize": (st, []),
            "cmg_set_kernel_forms": (st, [i, i]),
            "cmg_mg_solve": (st, [vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
            "cmg_geo_solver_create": (st, [vp, i, C.c_float, C.c_float, C.POINTER(vp)]),
            "cmg_gmg_damped_jacobi_solve": (st, [vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
            "cmg_solver_apply": (st, [vp, i, i]),
            "cmg_op_norm": (st, [vp, i, i, i, i, C.POINTER(d)]),
            "cmg_op_norm_noalloc": (st, [vp, i, i, i, i, C.POINTER(d)]),
            "cmg_grid_set_channel_device": (st, [vp, i, vp]), "cmg_grid_get_channel_device": (st, [vp, i, vp]),
            "cmg_grid_gather_tile_channel": (st, [vp, vp, vp, C.c_size_t, i, i]),
            "cmg_grid_scatter_tile_channel": (st, [vp, i, vp, vp, C.c_size_t, i]),
            "cmg_grid_gather_tile_types": (st, [vp, vp, vp, C.c_size_t]),
            "cmg_grid_scatter_tile_types": (st, [vp, vp, vp, C.c_size_t]),
            "cmg_mark_region_of_interest": (st, [vp, i, C.c_uint8, i, i, f, f, i]),
            "cmg_grid_get_roi": (st, [vp, vp, vp, vp, vp]),
            "cmg_solver_set_net": (st, [vp, vp]), "cmg_solver_net_infer": (st, [vp, i, i, i]),
            "cmg_trainer_create": (st, [i, C.c_uint64, C.POINTER(vp)]), "cmg_trainer_destroy": (None, [vp]),
            "cmg_trainer_collect": (st, [vp, vp, i, i, i, C.POINTER(i)]), "cmg_trainer_add_tiles": (st, [vp, vp, vp, i]),
            "cmg_trainer_size": (i, [vp]), "cmg_trainer_get_weights": (st, [vp, vp]),
            "cmg_trainer_set_weights": (st, [vp, vp]),
            "cmg_trainer_loss_grad": (st, [vp, i, i, i, C.POINTER(d), vp]),
            "cmg_trainer_step": (st, [vp, i, C.c_float, i, C.POINTER(d)]), "cmg_mfma_probe": (st, [i, vp, vp, vp]),
            "cmg_solver_learned_counts": (st, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
            "cmg_grid_dump_blob": (st, [vp, C.c_uint8, i, vp, C.POINTER(C.c_size_t)]),
            "cmg_grid_load_blob": (st, [vp, C.c_size_t, C.POINTER(vp)]),
            "cmg_assemble_cutcell_device": (st, [vp, i, vp, i, i, f, i, i]),
            "cmg_divergence": (st, [vp, i, i, i]), "cmg_add_gradient": (st, [vp, i, i, i, i]),
            "cmg_mix_velocity": (st, [vp, i, i, i, i]),
            "cmg_get_kernel_forms": (st, [C.POINTER(i), C.POINTER(i)]),
            "cmg_rhs_uniform": (st, [vp, i, C.c_uint64, i]), "cmg_rhs_divergence": (st, [vp, i, i, C.c_uint64]),
            "cmg_bench_kernel": (st, [vp, i, i, C.POINTER(d), C.POINTER(i)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(status, where):
    if status != 0:
        raise CmgError(status, where)


DEFAULT_LOG2_HASH = (16, 16, 16, 16, 16, 16, 20, 20, 16, 16)   # tests/TestGrids.cuh:149


def make_target(kind, min_level, max_level, sdf_kind=SDF_SPHERE, sdf=(), bands=None, fn=None):
    t = Target()
    t.kind, t.min_level, t.max_level, t.sdf_kind = kind, min_level, max_level, sdf_kind
    for k, v in enumerate(sdf):
        t.sdf[k] = v
    for lvl, b in (bands or {}).items():
        t.band[lvl] = b
    if fn is not None:
        t._cb = LEVEL_TARGET_FN(lambda user, level, c: int(fn(level, (c[0], c[1], c[2]))))
        t.fn = t._cb
    return t


class HADeviceGrid:
    """Octree of 8^3 tiles; construction mirrors HADeviceGrid (src/HAGrid.h:356-1326)."""

    def __init__(self, h0=0.125, log2_hash=DEFAULT_LOG2_HASH):
        self.h = C.c_void_p()
        self.num_levels = len(log2_hash)
        self.h0 = h0
        arr = (C.c_uint32 * len(log2_hash))(*log2_hash)
        _check(lib().cmg_grid_create(h0, len(log2_hash), arr, C.byref(self.h)), "cmg_grid_create")

    def close(self):
        if self.h:
            lib().cmg_grid_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # construction
    def setTileHost(self, level, coord, tile_type):
        _check(lib().cmg_grid_set_tile(self.h, level, *coord, tile_type), "cmg_grid_set_tile")

    # device=True runs the step's flag passes, hash insertions and table builds as kernels; the
    # resulting grid is the host construction's, slot for slot
    def rebuild(self, device=False):
        if device:
            _check(lib().cmg_grid_rebuild_device(self.h), "cmg_grid_rebuild_device")
        else:
            _check(lib().cmg_grid_rebuild(self.h), "cmg_grid_rebuild")

    def iterativeRefine(self, target, device=False):
        if device:
            _check(lib().cmg_grid_iterative_refine_device(self.h, C.byref(target)), "cmg_grid_iterative_refine_device")
        else:
            _check(lib().cmg_grid_iterative_refine(self.h, C.byref(target)), "cmg_grid_iterative_refine")

    def finalize(self, device=False):
        if device:
            _check(lib().cmg_grid_finalize_device(self.h), "cmg_grid_finalize_device")
        else:
            _check(lib().cmg_grid_finalize(self.h), "cmg_grid_finalize")

    def build_device(self, target):
        """rebuild + iterativeRefine + finalize on the device over one copy of the tables"""
        _check(lib().cmg_grid_build_device(self.h, C.byref(target)), "cmg_grid_build_device")

    @classmethod
    def test_grid(cls, target, h0=0.125, log2_hash=DEFAULT_LOG2_HASH, device=False):
        """CreateTestGrid (tests/TestGrids.cuh:144-169): one LEAF root, rebuild, refine, LEAF cell types."""
        g = cls(h0, log2_hash)
        g.setTileHost(0, (0, 0, 0), LEAF)
        if device:
            g.build_device(target)
        else:
            g.rebuild()
            g.iterativeRefine(target)
            g.finalize()
        g.set_leaf_types_preset(target)
        return g

    # topology views
    def dump_blob(self, tile_types=7, max_level=-1):
        """HADeviceGrid::dumpBinaryBlob (src/HAGrid.h:687-758): the grid as 'HAGR' bytes"""
        n = C.c_size_t(0)
        _check(lib().cmg_grid_dump_blob(self.h, tile_types, max_level, None, C.byref(n)), "dump_blob")
        buf = (C.c_uint8 * n.value)()
        _check(lib().cmg_grid_dump_blob(self.h, tile_types, max_level, buf, C.byref(n)), "dump_blob")
        return bytes(buf)

    @classmethod
    def load_blob(cls, blob):
        """HADeviceGrid::loadBinaryBlob (src/HAGrid.h:760-853)"""
        g = cls.__new__(cls)
        g.h = C.c_void_p()
        g.num_levels = int(np.frombuffer(blob, np.uint32, 1, 8)[0]) if len(blob) >= 40 else 0
        g.h0 = float(np.frombuffer(blob, np.float32, 1, 24)[0]) if len(blob) >= 40 else 0.0
        data = (C.c_uint8 * len(blob)).from_buffer_copy(blob)
        _check(lib().cmg_grid_load_blob(data, len(blob), C.byref(g.h)), "load_blob")
        return g

    @property
    def N(self):
        return lib().cmg_grid_num_tiles(self.h)

    @property
    def max_level(self):
        return lib().cmg_grid_max_level(self.h)

    def level_starts(self):
        s = np.zeros(self.num_levels + 1, np.int32)
        _check(lib().cmg_grid_level_starts(self.h, _ptr(s)), "level_starts")
        return s

    def tiles(self):
        n = self.N
        lv = np.zeros(n, np.int32); co = np.zeros((n, 3), np.int32); ty = np.zeros(n, np.uint8)
        _check(lib().cmg_grid_get_tiles(self.h, _ptr(lv), _ptr(co), _ptr(ty)), "get_tiles")
        return lv, co, ty

    def slots(self):
        s = np.zeros(self.N, np.int32); _check(lib().cmg_grid_get_slots(self.h, _ptr(s)), "get_slots"); return s

    def neighbors(self):
        a = np.zeros((self.N, 6), np.int32); _check(lib().cmg_grid_get_neighbors(self.h, _ptr(a)), "nbr"); return a

    def parents(self):
        a = np.zeros(self.N, np.int32); _check(lib().cmg_grid_get_parents(self.h, _ptr(a)), "parents"); return a

    def children(self):
        a = np.zeros((self.N, 8), np.int32); _check(lib().cmg_grid_get_children(self.h, _ptr(a)), "children"); return a

    # cell data
    def set_leaf_types_preset(self, target):
        _check(lib().cmg_grid_set_leaf_types_preset(self.h, C.byref(target)), "set_leaf_types_preset")

    def set_wall_types(self, chk_level, dirichlet_ytop):
        _check(lib().cmg_grid_set_wall_types(self.h, chk_level, int(dirichlet_ytop)), "set_wall_types")

    def cell_types(self):
        a = np.zeros((self.N, 512), np.uint8)
        _check(lib().cmg_grid_get_cell_types(self.h, _ptr(a)), "get_cell_types")
        return a

    def set_cell_types(self, types):
        a = np.ascontiguousarray(types, np.uint8)
        assert a.size == self.N * 512
        _check(lib().cmg_grid_set_cell_types(self.h, _ptr(a)), "set_cell_types")

    def host_channel(self, c):
        a = np.zeros((self.N, 512), np.float32)
        _check(lib().cmg_grid_get_host_channel(self.h, c, _ptr(a)), "get_host_channel")
        return a

    def set_channel(self, c, values):
        a = np.ascontiguousarray(values, np.float32)
        assert a.size == self.N * 512
        _check(lib().cmg_grid_set_channel(self.h, c, _ptr(a)), "set_channel")

    def channel(self, c):
        a = np.zeros((self.N, 512), np.float32)
        _check(lib().cmg_grid_get_channel(self.h, c, _ptr(a)), "get_channel")
        return a

    def copy_channel(self, src, dst):
        _check(lib().cmg_grid_copy_channel(self.h, src, dst), "copy_channel")

    def fill_channel(self, c, value):
        _check(lib().cmg_grid_fill_channel(self.h, c, value), "fill_channel")

    # inputs
    def rhs_uniform(self, c, seed, zero_mean=True):
        _check(lib().cmg_rhs_uniform(self.h, c, seed, int(zero_mean)), "rhs_uniform")

    def rhs_divergence(self, c, u0, seed):
        _check(lib().cmg_rhs_divergence(self.h, c, u0, seed), "rhs_divergence")

    # assembly
    def prepare_types_and_coeffs(self, coeff_channel, R_matrix):
        _check(lib().cmg_prepare_types_and_coeffs(self.h, coeff_channel, R_matrix), "prepare_types_and_coeffs")

    def assemble_cutcell(self, sdf_kind, sdf, face_rule, coeff_channel, R_matrix):
        arr = (C.c_float * 8)(*sdf)
        _check(lib().cmg_assemble_cutcell(self.h, sdf_kind, arr, face_rule, coeff_channel, R_matrix),
               "assemble_cutcell")

    # kernels
    def assemble_cutcell_device(self, sdf_kind, sdf, face_rule, coeff_channel, R_matrix, wall_chk_level=-1,
                                dirichlet_ytop=False):
        """the host assembly (optional wall typing, solid cells, operator) run on the device"""
        arr = (C.c_float * len(sdf))(*sdf) if sdf is not None else None
        _check(lib().cmg_assemble_cutcell_device(self.h, sdf_kind, arr, face_rule, coeff_channel, R_matrix,
                                                 wall_chk_level, int(dirichlet_ytop)), "cmg_assemble_cutcell_device")

    def full_neg_lap(self, x, c0, y):
        _check(lib().cmg_op_full_neg_lap(self.h, x, c0, y), "full_neg_lap")

    def propagate_ghosts(self, c):
        _check(lib().cmg_op_propagate_ghosts(self.h, c), "propagate_ghosts")

    def accumulate_leafs(self, c):
        _check(lib().cmg_op_accumulate_leafs(self.h, c), "accumulate_leafs")

    def gs_color(self, level, x, c0, rhs, color, omega):
        _check(lib().cmg_op_gs_color(self.h, level, x, c0, rhs, color, omega), "gs_color")

    def jacobi(self, level, x_in, x_out, ghost_ch, c0, rhs, omega):
        _check(lib().cmg_op_jacobi(self.h, level, x_in, x_out, ghost_ch, c0, rhs, omega), "jacobi")

    def residual_restrict(self, level, x, c0, rhs, R_restrict):
        _check(lib().cmg_op_residual_restrict(self.h, level, x, c0, rhs, R_restrict), "residual_restrict")

    def fas_save(self, level, x, c0, rhs, x0):
        _check(lib().cmg_op_fas_save(self.h, level, x, c0, rhs, x0), "fas_save")

    def fas_prolong(self, level, x, x0, P):
        _check(lib().cmg_op_fas_prolong(self.h, level, x, x0, P), "fas_prolong")

    def divergence(self, u0, channel, coeff_channel=-1):
        _check(lib().cmg_divergence(self.h, u0, channel, coeff_channel), "cmg_divergence")

    def add_gradient(self, x, coeff_channel, u0, weighted=False):
        _check(lib().cmg_add_gradient(self.h, x, coeff_channel, u0, int(weighted)), "cmg_add_gradient")

    def mix_velocity(self, coeff_channel, u_fluid, u_solid, u_mix):
        _check(lib().cmg_mix_velocity(self.h, coeff_channel, u_fluid, u_solid, u_mix), "cmg_mix_velocity")

    def norm(self, order, channel, volume_weighted=False, cell_types=2):
        """NormSync (src/PoissonGrid.cu:322-375) over LEAF tiles."""
        out = C.c_double()
        _check(lib().cmg_op_norm(self.h, order, channel, int(volume_weighted), cell_types, C.byref(out)), "cmg_op_norm")
        return out.value

    def norm_noalloc(self, order, channel, volume_weighted=False, cell_types=2):
        """norm() with no allocation: the grid and the channel must already be on the device."""
        out = C.c_double()
        _check(lib().cmg_op_norm_noalloc(self.h, order, channel, int(volume_weighted), cell_types, C.byref(out)),
               "cmg_op_norm_noalloc")
        return out.value

    def dot(self, a, b):
        out = C.c_double()
        _check(lib().cmg_op_dot(self.h, a, b, C.byref(out)), "dot")
        return out.value

    # device-resident exchange: device addresses as Python ints
    def set_channel_device(self, c, d_src):
        _check(lib().cmg_grid_set_channel_device(self.h, c, C.c_void_p(d_src)), "cmg_grid_set_channel_device")

    def get_channel_device(self, c, d_dst):
        _check(lib().cmg_grid_get_channel_device(self.h, c, C.c_void_p(d_dst)), "cmg_grid_get_channel_device")

    def gather_tile_channel(self, tile_channel, channel, d_addrs=None, d_base=None, stride=0):
        """channel <- tile_channel of PoissonTile<float,15> device images (src/PoissonTile.h:9-175)"""
        _check(lib().cmg_grid_gather_tile_channel(self.h, C.c_void_p(d_addrs), C.c_void_p(d_base), stride, tile_channel,
                                                  channel), "cmg_grid_gather_tile_channel")

    def scatter_tile_channel(self, channel, tile_channel, d_addrs=None, d_base=None, stride=0):
        _check(lib().cmg_grid_scatter_tile_channel(self.h, channel, C.c_void_p(d_addrs), C.c_void_p(d_base), stride,
                                                   tile_channel), "cmg_grid_scatter_tile_channel")

    def gather_tile_types(self, d_addrs=None, d_base=None, stride=0):
        _check(lib().cmg_grid_gather_tile_types(self.h, C.c_void_p(d_addrs), C.c_void_p(d_base), stride),
               "cmg_grid_gather_tile_types")

    def scatter_tile_types(self, d_addrs=None, d_base=None, stride=0):
        _check(lib().cmg_grid_scatter_tile_types(self.h, C.c_void_p(d_addrs), C.c_void_p(d_base), stride),
               "cmg_grid_scatter_tile_types")

    def mark_region_of_interest(self, channel, pred=0, a=0.0, b=0.0, subtree_level=-1, launch_types=LEAF,
                                calc_locked=False):
        """MarkRegionOfInterestWithChannelMinAndMax128Kernel (src/PoissonGridKernels.cuh:11-54)"""
        _check(lib().cmg_mark_region_of_interest(self.h, subtree_level, launch_types, channel, pred, a, b,
                                                 int(calc_locked)), "cmg_mark_region_of_interest")

    def roi(self):
        """(interest, locked, tile min, tile max) of the last marking"""
        n = self.N
        fi, fl = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        mn, mx = np.zeros(n, np.float32), np.zeros(n, np.float32)
        _check(lib().cmg_grid_get_roi(self.h, _ptr(fi), _ptr(fl), _ptr(mn), _ptr(mx)), "cmg_grid_get_roi")
        return fi, fl, mn, mx


def device_synchronize():
    _check(lib().cmg_device_synchronize(), "cmg_device_synchronize")


def set_kernel_forms(persistent_mask, stencil_threads):
    """Stencil kernel forms (bit 0 operator, bit 1 Jacobi, bit 2 restrict persistent, bit 3 stored
    diagonals, bit 4 unfused boundary work, bit 5 no pulled prolongation, bit 6 no coefficient
    codes; threads per tile, 0 for the library's defaults)."""
    _check(lib().cmg_set_kernel_forms(persistent_mask, stencil_threads), "cmg_set_kernel_forms")


def get_kernel_forms():
    m, t = C.c_int(), C.c_int()
    _check(lib().cmg_get_kernel_forms(C.byref(m), C.byref(t)), "cmg_get_kernel_forms")
    return m.value, t.value


def default_opts(**kw):
    o = SolveOpts()
    lib().cmg_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class AMGSolver:
    """AMGSolver(coeff_channel, R_matrix, R_restrict, P) (src/AMGSolver.h:19-74)."""

    def __init__(self, grid, coeff_channel, R_matrix_coeff=1.0, R_restrict_coeff=1.0, P_coeff=1.0):
        self.grid = grid
        self.h = C.c_void_p()
        _check(lib().cmg_solver_create(grid.h, coeff_channel, R_matrix_coeff, R_restrict_coeff, P_coeff,
                                       C.byref(self.h)), "cmg_solver_create")
        self.last_stats = None

    def close(self):
        if self.h:
            lib().cmg_solver_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, verbose=False, max_iters=100, relative_tolerance=1e-6, level_iters=2, coarsest_iters=10,
              sync_stride=1, is_pure_neumann=False, opts=None, raise_on_nonconvergence=False):
        """AMGSolver::solve (src/AMGSolver.cu:1399-1549): returns (iters, relative residual)."""
        o = opts or default_opts()
        o.verbose, o.max_iters, o.relative_tolerance = int(verbose), max_iters, relative_tolerance
        o.level_iters, o.coarsest_iters, o.sync_stride = level_iters, coarsest_iters, sync_stride
        o.is_pure_neumann = int(is_pure_neumann)
        stats = SolveStats()
        status = lib().cmg_solve(self.h, C.byref(o), C.byref(stats))
        self.last_stats = stats
        if status not in (0, 6) or (status == 6 and raise_on_nonconvergence):
            raise CmgError(status, "cmg_solve")
        return stats.iters, stats.rel_res

    def mg_solve(self, opts, max_iters=100, relative_tolerance=1e-6):
        """AMGSolver::FASMuCycleSolve (src/AMGSolver.cu:1311-1356): (loop index at exit, |r|/|b|)."""
        o = opts
        o.max_iters, o.relative_tolerance = max_iters, relative_tolerance
        stats = SolveStats()
        status = lib().cmg_mg_solve(self.h, C.byref(o), C.byref(stats))
        self.last_stats = stats
        if status not in (0, 6):
            raise CmgError(status, "cmg_mg_solve")
        return stats.iters, stats.rel_res

    def x64(self):
        a = np.zeros((self.grid.N, 512), np.float64)
        _check(lib().cmg_solver_get_x64(self.h, _ptr(a)), "get_x64")
        return a

    def coded_levels(self):
        """bit mask of the levels whose face coefficients the stencil kernels read as codes"""
        return self.code_levels()[0]

    def code_levels(self):
        """(coded levels, the subset with ten-bit codes) as bit masks"""
        m, w = C.c_uint32(), C.c_uint32()
        _check(lib().cmg_solver_coded_levels(self.h, C.byref(m), C.byref(w)), "cmg_solver_coded_levels")
        return m.value, w.value

    def apply(self, x, y):
        """y = A x with this solver's operator on LEAF tiles."""
        _check(lib().cmg_solver_apply(self.h, x, y), "cmg_solver_apply")

    def set_net(self, weights):
        """weights of the learned tile smoother (7776 floats)"""
        w = np.ascontiguousarray(weights, np.float32)
        assert w.size == NET_WEIGHTS
        _check(lib().cmg_solver_set_net(self.h, _ptr(w)), "cmg_solver_set_net")

    def net_infer(self, r, diag, y):
        _check(lib().cmg_solver_net_infer(self.h, r, diag, y), "cmg_solver_net_infer")

    def learned_counts(self):
        """(tiles judged, tiles that kept the net's correction) over the last learned solve"""
        j, k = C.c_uint64(), C.c_uint64()
        _check(lib().cmg_solver_learned_counts(self.h, C.byref(j), C.byref(k)), "cmg_solver_learned_counts")
        return j.value, k.value

    def bench_kernel(self, kernel, reps):
        ms, tiles = C.c_double(), C.c_int()
        _check(lib().cmg_bench_kernel(self.h, kernel, reps, C.byref(ms), C.byref(tiles)), "bench_kernel")
        return ms.value, tiles.value

    def cycle(self, opts, x, rhs, x0):
        _check(lib().cmg_op_cycle(self.h, C.byref(opts), x, rhs, x0), "cycle")


class _GeoSolver(AMGSolver):
    FAMILY = None

    def __init__(self, grid, one_over_alpha=1.0, prolong_coeff=1.0):
        self.grid = grid
        self.h = C.c_void_p()
        _check(lib().cmg_geo_solver_create(grid.h, self.FAMILY, one_over_alpha, prolong_coeff, C.byref(self.h)),
               "cmg_geo_solver_create")
        self.last_stats = None


class GMGSolver(_GeoSolver):
    """GMGSolver(one_over_alpha, prolong_coeff) (src/GMGSolver.h:14-50)."""
    FAMILY = SOLVER_GMG

    def damped_jacobi_solve(self, max_iters, relative_tolerance, omega):
        """GMGSolver::dampedJacobiSolve (src/GMGSolver.cu:390-470): (loop index at exit, |r|/|b|)."""
        o = default_opts(max_iters=max_iters, relative_tolerance=relative_tolerance, omega=omega)
        stats = SolveStats()
        status = lib().cmg_gmg_damped_jacobi_solve(self.h, C.byref(o), C.byref(stats))
        self.last_stats = stats
        if status not in (0, 6):
            raise CmgError(status, "cmg_gmg_damped_jacobi_solve")
        return stats.iters, stats.rel_res


class CMGSolver(_GeoSolver):
    """CMGSolver(one_over_alpha, prolong_coeff) (src/CMGSolver.h:12-47)."""
    FAMILY = SOLVER_CMG


NET_WEIGHTS, NET_VOLUME = 7776, 1000


class Trainer:
    """Training state of the learned tile smoother (device-resident tiles, Adam)."""

    def __init__(self, capacity, seed=3):
        self.h = C.c_void_p()
        _check(lib().cmg_trainer_create(capacity, seed, C.byref(self.h)), "cmg_trainer_create")

    def close(self):
        if self.h:
            lib().cmg_trainer_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self):
        return lib().cmg_trainer_size(self.h)

    def collect(self, grid, r, diag, e):
        n = C.c_int()
        _check(lib().cmg_trainer_collect(self.h, grid.h, r, diag, e, C.byref(n)), "cmg_trainer_collect")
        return n.value

    def add_tiles(self, x, e):
        x = np.ascontiguousarray(x, np.float32)
        e = np.ascontiguousarray(e, np.float32)
        _check(lib().cmg_trainer_add_tiles(self.h, _ptr(x), _ptr(e), x.shape[0]), "cmg_trainer_add_tiles")

    def weights(self):
        w = np.zeros(NET_WEIGHTS, np.float32)
        _check(lib().cmg_trainer_get_weights(self.h, _ptr(w)), "cmg_trainer_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _check(lib().cmg_trainer_set_weights(self.h, _ptr(w)), "cmg_trainer_set_weights")

    def loss_grad(self, first, n, bf16=False):
        loss = C.c_double()
        g = np.zeros(NET_WEIGHTS, np.float32)
        _check(lib().cmg_trainer_loss_grad(self.h, first, n, int(bf16), C.byref(loss), _ptr(g)), "cmg_trainer_loss_grad")
        return loss.value, g

    def step(self, batch, lr=1e-3, bf16=True):
        loss = C.c_double()
        _check(lib().cmg_trainer_step(self.h, batch, lr, int(bf16), C.byref(loss)), "cmg_trainer_step")
        return loss.value


def mfma_probe(A_bits, B_bits):
    """C = A B on one MFMA instruction: A is 16 x K, B is K x 16 bfloat16 bit
    patterns, K = 16 or 32."""
    A = np.ascontiguousarray(A_bits, np.uint16)
    B = np.ascontiguousarray(B_bits, np.uint16)
    K = A.shape[1]
    if A.shape != (16, K) or B.shape != (K, 16) or K not in (16, 32):
        raise ValueError("mfma_probe: A must be 16 x K and B K x 16 with K = 16 or 32")
    Cm = np.zeros((16, 16), np.float32)
    _check(lib().cmg_mfma_probe(K, _ptr(A), _ptr(B), _ptr(Cm)), "cmg_mfma_probe")
    return Cm
