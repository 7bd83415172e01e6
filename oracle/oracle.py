"""ctypes view of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The shipped library never loads
it.  See oracle.h for the reference functions each entry point restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

LEAF, NONLEAF, GHOST = 1, 2, 4
DIRICHLET, INTERIOR, NEUMANN = 1, 2, 4
TARGET_UNIFORM, TARGET_CENTER_POINT, TARGET_SPHERE_EMPTY, TARGET_SPHERE_SOLID = 0, 1, 2, 3
TARGET_STAR_EMPTY, TARGET_STAR_SOLID, TARGET_SDF_BANDS = 4, 5, 6
SDF_SPHERE, SDF_TORUS, SDF_STAR = 0, 1, 2
MAX_LEVELS = 12


class Target(C.Structure):
    _fields_ = [("kind", C.c_int), ("min_level", C.c_int), ("max_level", C.c_int),
                ("sdf_kind", C.c_int), ("sdf", C.c_float * 8), ("band", C.c_float * MAX_LEVELS)]


class Cycle(C.Structure):
    _fields_ = [("mu", C.c_int), ("smoother", C.c_int), ("omega", C.c_double),
                ("level_iters", C.c_int), ("coarsest_iters", C.c_int), ("coarsest_direct", C.c_int),
                ("R_restrict", C.c_float), ("P", C.c_float), ("tmp_ch", C.c_int), ("kind", C.c_int)]


class SolveOut(C.Structure):
    _fields_ = [("iters", C.c_int), ("rel_res", C.c_double)]


class CSR(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int64), ("rowptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double)), ("cell", C.POINTER(C.c_int32))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, i, f, d, u64 = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_uint64
        sig = {
            "og_create": (vp, [f, i, vp]), "og_free": (None, [vp]),
            "og_hash": (C.c_uint32, [C.c_int32, C.c_int32, C.c_int32, C.c_uint32]),
            "og_set_tile": (i, [vp, i, i, i, i, C.c_uint8]),
            "og_compress": (None, [vp]), "og_spawn_ghosts": (None, [vp]),
            "og_refine_step": (i, [vp, vp, vp]), "og_iterative_refine": (None, [vp, vp]),
            "og_create_test_grid": (vp, [vp, vp, i, f]), "og_finalize": (None, [vp]),
            "og_num_tiles": (i, [vp]), "og_max_level": (i, [vp]), "og_level_range": (None, [vp, vp]),
            "og_get_tiles": (None, [vp, vp, vp, vp]), "og_get_slots": (None, [vp, vp]),
            "og_get_neighbors": (None, [vp, vp]), "og_get_parents": (None, [vp, vp]),
            "og_get_children": (None, [vp, vp]),
            "og_channel": (C.POINTER(C.c_float), [vp, i]), "og_cell_types": (C.POINTER(C.c_uint8), [vp]),
            "og_h": (f, [vp, i]),
            "og_set_leaf_types_preset": (None, [vp, vp]), "og_set_wall_types": (None, [vp, i, i]),
            "og_set_solid_types": (None, [vp, i, vp]),
            "og_coarsen_types_and_coeffs": (None, [vp, i, f]),
            "og_cutcell_coeffs": (None, [vp, i, vp, i, i, f]),
            "og_prepare_system": (None, [vp, i, f]),
            "og_face_fluid_ratio": (f, [f, f, f, f]), "og_sdf": (f, [i, vp, f, f, f]),
            "og_face_supersampled": (f, [i, vp, vp, i, f]),
            "og_roi_flags": (None, [vp, i, i, i, i, f, f, vp, vp, vp]),
            "og_propagate_ghosts": (None, [vp, i]), "og_accumulate_leafs": (None, [vp, i]),
            "og_neg_lap_leafs": (None, [vp, i, i, i]), "og_full_neg_lap": (None, [vp, i, i, i]),
            "og_gs_color": (None, [vp, i, i, i, i, i, d]),
            "og_gauss_seidel": (None, [vp, i, i, i, i, i, i, d]),
            "og_jacobi": (None, [vp, i, i, i, i, i, i, f]),
            "og_residual_restrict": (None, [vp, i, i, i, i, f]),
            "og_fas_save": (None, [vp, i, i, i, i, i]), "og_fas_prolong": (None, [vp, i, i, i, f]),
            "og_dot": (d, [vp, i, i]), "og_recenter": (None, [vp, i]),
            "og_fas_cycle": (None, [vp, vp, i, i, i, i]),
            "og_pcg": (SolveOut, [vp, vp, i, i, d, i]),
            "og_pcg_ex": (SolveOut, [vp, vp, i, i, d, i, i]),
            "og_fas_mu_solve": (SolveOut, [vp, vp, i, i, d]),
            "og_norm": (d, [vp, i, i, i, i]),
            "og_net_forward": (None, [vp, vp, vp]), "og_net_loss_grad": (d, [vp, i, vp, vp, vp]),
            "og_adam_step": (None, [vp, vp, vp, vp, i, i, d, d, d, d]),
            "og_divergence": (None, [vp, i, i, i]), "og_add_gradient": (None, [vp, i, i, i, i]),
            "og_mix_velocity": (None, [vp, i, i, i, i]),
            "og_geo_neg_lap_tiles": (None, [vp, i, i, i, i, i, i]), "og_geo_full_neg_lap": (None, [vp, i, i, i, i]),
            "og_geo_gs_color": (None, [vp, i, i, i, i, i]), "og_geo_residual_restrict": (None, [vp, i, i, i, i, f]),
            "og_geo_prolong": (None, [vp, i, i, i, f]), "og_geo_vcycle": (None, [vp, i, i, i, i, i, i, f, f]),
            "og_geo_pcg": (SolveOut, [vp, i, i, i, f, f, i, d, i]),
            "og_gmg_damped_jacobi": (SolveOut, [vp, i, d, f]),
            "og_restrict_source": (None, [vp, i, i, f]), "og_prolong_add": (None, [vp, i, i, f]),
            "og_deduct_flux": (None, [vp, i, i, i, i, i]), "og_set_zero_tiles": (None, [vp, i, i, i]),
            "og_vcycle": (None, [vp, vp, i, i, i, i]), "og_mu_cycle": (None, [vp, vp, i, i, i, i]),
            "og_fas_fcycle": (None, [vp, vp, i, i, i, i]), "og_precondition": (None, [vp, vp, i, i, i, i]),
            "og_assemble_csr": (C.POINTER(CSR), [vp, i]), "og_csr_free": (None, [C.POINTER(CSR)]),
            "og_csr_matvec": (None, [C.POINTER(CSR), vp, vp]),
            "og_bicgstab": (i, [C.POINTER(CSR), vp, vp, d, i, vp, i]),
            "og_cg": (i, [C.POINTER(CSR), vp, vp, d, i, vp, i]),
            "og_csr_asym": (d, [C.POINTER(CSR)]),
            "og_uniform_pm1": (f, [vp]), "og_normal": (f, [vp]), "og_splitmix64": (u64, [vp]),
            "og_rhs_uniform": (None, [vp, i, u64, i]), "og_rhs_divergence": (None, [vp, i, i, u64]),
            "og_num_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


DEFAULT_LOG2_HASH = (16, 16, 16, 16, 16, 16, 20, 20, 16, 16)   # tests/TestGrids.cuh:149


def make_target(kind, min_level, max_level, sdf_kind=SDF_SPHERE, sdf=(), bands=None):
    t = Target()
    t.kind, t.min_level, t.max_level, t.sdf_kind = kind, min_level, max_level, sdf_kind
    for i, v in enumerate(sdf):
        t.sdf[i] = v
    if bands:
        for lvl, b in bands.items():
            t.band[lvl] = b
    return t


CYCLE_FAS_MU, CYCLE_FAS_F, CYCLE_VCYCLE, CYCLE_MU = 0, 1, 2, 3
GEO_GMG, GEO_CMG = 0, 1
OG_INTERIOR_T = 2


def make_cycle(mu=2, smoother=0, omega=1.5, level_iters=2, coarsest_iters=10, coarsest_direct=0,
               R_restrict=1.0, P=1.0, tmp_ch=5, kind=CYCLE_FAS_MU):
    return Cycle(mu, smoother, omega, level_iters, coarsest_iters, coarsest_direct, R_restrict, P, tmp_ch, kind)


class Grid:
    """Oracle octree grid (restates HADeviceGrid construction)."""

    def __init__(self, target=None, log2_hash=DEFAULT_LOG2_HASH, h0=0.125, handle=None):
        L = lib()
        self._lh = (C.c_uint32 * len(log2_hash))(*log2_hash)
        self.num_levels = len(log2_hash)
        if handle is not None:
            self.h = handle
        else:
            self.target = target
            self.h = L.og_create_test_grid(C.byref(target), self._lh, len(log2_hash), h0)

    def close(self):
        if self.h:
            lib().og_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def N(self):
        return lib().og_num_tiles(self.h)

    @property
    def max_level(self):
        return lib().og_max_level(self.h)

    def level_range(self):
        s = np.zeros(self.num_levels + 1, np.int32)
        lib().og_level_range(self.h, _ptr(s))
        return s

    def tiles(self):
        n = self.N
        lv = np.zeros(n, np.int32); co = np.zeros((n, 3), np.int32); ty = np.zeros(n, np.uint8)
        lib().og_get_tiles(self.h, _ptr(lv), _ptr(co), _ptr(ty))
        return lv, co, ty

    def slots(self):
        s = np.zeros(self.N, np.int32); lib().og_get_slots(self.h, _ptr(s)); return s

    def neighbors(self):
        a = np.zeros((self.N, 6), np.int32); lib().og_get_neighbors(self.h, _ptr(a)); return a

    def parents(self):
        a = np.zeros(self.N, np.int32); lib().og_get_parents(self.h, _ptr(a)); return a

    def children(self):
        a = np.zeros((self.N, 8), np.int32); lib().og_get_children(self.h, _ptr(a)); return a

    def channel(self, c):
        p = lib().og_channel(self.h, c)
        return np.ctypeslib.as_array(p, shape=(self.N, 512))

    def cell_types(self):
        p = lib().og_cell_types(self.h)
        return np.ctypeslib.as_array(p, shape=(self.N, 512))

    def hval(self, level):
        return lib().og_h(self.h, level)

    # --- cell types / assembly
    def set_leaf_types_preset(self, target=None):
        lib().og_set_leaf_types_preset(self.h, C.byref(target or self.target))

    def set_wall_types(self, chk_level, dirichlet_ytop):
        lib().og_set_wall_types(self.h, chk_level, int(dirichlet_ytop))

    def coarsen_types_and_coeffs(self, c0=11, R=0.5):
        lib().og_coarsen_types_and_coeffs(self.h, c0, R)

    def cutcell_coeffs(self, sdf_kind, sdf, face_rule, c0=11, R=0.5):
        arr = (C.c_float * 8)(*sdf)
        lib().og_cutcell_coeffs(self.h, sdf_kind, arr, face_rule, c0, R)

    # --- ops
    def full_neg_lap(self, x, c0, y):
        lib().og_full_neg_lap(self.h, x, c0, y)

    def roi_flags(self, ch, pred=0, a=0.0, b=0.0, subtree_level=-1, launch_types=1):
        """MarkRegionOfInterestWithChannelMinAndMax128Kernel (src/PoissonGridKernels.cuh:11-54)"""
        n = self.N
        fi = np.zeros(n, np.uint8)
        mn, mx = np.zeros(n, np.float32), np.zeros(n, np.float32)
        lib().og_roi_flags(self.h, ch, subtree_level, launch_types, pred, a, b, _ptr(fi), _ptr(mn), _ptr(mx))
        return fi, mn, mx

    def dot(self, a, b):
        return lib().og_dot(self.h, a, b)

    def divergence(self, u0, ch, c0=-1):
        lib().og_divergence(self.h, u0, ch, c0)

    def add_gradient(self, x, c0, u0, weighted=False):
        lib().og_add_gradient(self.h, x, c0, u0, int(weighted))

    def mix_velocity(self, c0, uf0, us0, um0):
        lib().og_mix_velocity(self.h, c0, uf0, us0, um0)

    def norm(self, order, ch, volume_weighted=False, cell_types=OG_INTERIOR_T):
        return lib().og_norm(self.h, order, ch, int(volume_weighted), cell_types)

    def pcg(self, cycle, c0=11, max_iters=100, tol=1e-6, pure_neumann=False, flexible=False):
        out = lib().og_pcg_ex(self.h, C.byref(cycle), c0, max_iters, tol, int(pure_neumann), int(flexible))
        return out.iters, out.rel_res

    def fas_cycle(self, cycle, x, rhs, x0, c0=11):
        lib().og_fas_cycle(self.h, C.byref(cycle), x, rhs, x0, c0)

    def geo_pcg(self, mode, level_iters=2, coarsest_iters=10, one_over_alpha=1.0, P=1.0, max_iters=100, tol=1e-6,
                pure_neumann=False):
        out = lib().og_geo_pcg(self.h, mode, level_iters, coarsest_iters, one_over_alpha, P, max_iters, tol,
                               int(pure_neumann))
        return out.iters, out.rel_res

    def gmg_damped_jacobi(self, max_iters, tol, omega):
        out = lib().og_gmg_damped_jacobi(self.h, max_iters, tol, omega)
        return out.iters, out.rel_res

    def precondition(self, cycle, z, r, s, c0=11):
        lib().og_precondition(self.h, C.byref(cycle), z, r, s, c0)

    def fas_mu_solve(self, cycle, c0=11, max_iters=100, tol=1e-6):
        out = lib().og_fas_mu_solve(self.h, C.byref(cycle), c0, max_iters, tol)
        return out.iters, out.rel_res

    def rhs_uniform(self, ch, seed, zero_mean=True):
        lib().og_rhs_uniform(self.h, ch, seed, int(zero_mean))

    def rhs_divergence(self, ch, u0, seed):
        lib().og_rhs_divergence(self.h, ch, u0, seed)

    def assemble_csr(self, c0=11):
        return CSRMatrix(lib().og_assemble_csr(self.h, c0))


class CSRMatrix:
    def __init__(self, p):
        self.p = p
        m = p.contents
        self.n, self.nnz = m.n, m.nnz
        self.cell = np.ctypeslib.as_array(m.cell, shape=(self.n,)).copy() if self.n else np.zeros(0, np.int32)

    def to_scipy(self):
        import scipy.sparse as sp
        m = self.p.contents
        rp = np.ctypeslib.as_array(m.rowptr, shape=(self.n + 1,)).copy()
        col = np.ctypeslib.as_array(m.col, shape=(self.nnz,)).copy()
        val = np.ctypeslib.as_array(m.val, shape=(self.nnz,)).copy()
        return sp.csr_matrix((val, col, rp), shape=(self.n, self.n))

    def matvec(self, x):
        x = np.ascontiguousarray(x, np.float64); y = np.zeros(self.n)
        lib().og_csr_matvec(self.p, _ptr(x), _ptr(y)); return y

    def bicgstab(self, b, tol=1e-12, max_iters=20000, nthreads=0):
        b = np.ascontiguousarray(b, np.float64); x = np.zeros(self.n); rr = C.c_double()
        it = lib().og_bicgstab(self.p, _ptr(b), _ptr(x), tol, max_iters, C.byref(rr), nthreads)
        return x, it, rr.value

    def cg(self, b, tol=1e-6, max_iters=20000, nthreads=0, x0=None):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        rr = C.c_double()
        it = lib().og_cg(self.p, _ptr(b), _ptr(x), tol, max_iters, C.byref(rr), nthreads)
        return x, it, rr.value

    def asym(self):
        return lib().og_csr_asym(self.p)

    def close(self):
        if self.p:
            lib().og_csr_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NET_W = 7776


def _dptr(a):
    return a.ctypes.data_as(C.c_void_p)


def net_forward(w, x):
    """FP64 forward pass of the learned tile smoother on one 10^3 tile."""
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(1000)
    y = np.zeros(1000, np.float64)
    lib().og_net_forward(_dptr(w), _dptr(x), _dptr(y))
    return y


def net_loss_grad(w, x, e, with_grad=True):
    """mean loss over tiles x, e (n x 1000 float32) and its FP64 gradient"""
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    e = np.ascontiguousarray(e, np.float32)
    g = np.zeros(NET_W, np.float64) if with_grad else None
    loss = lib().og_net_loss_grad(_dptr(w), x.shape[0], _dptr(x), _dptr(e), _dptr(g) if with_grad else None)
    return loss, g


def adam_step(w, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    lib().og_adam_step(_dptr(w), _dptr(m), _dptr(v), _dptr(np.ascontiguousarray(g, np.float64)), w.size, t, lr, b1, b2, eps)


def face_supersampled(sdf_kind, sdf, gi, axis, h):
    """4x4 supersampled open fraction of the negative face of global cell gi along axis."""
    prm = (C.c_float * max(1, len(sdf)))(*sdf)
    g = (C.c_int * 3)(*gi)
    return lib().og_face_supersampled(sdf_kind, prm, g, axis, h)
