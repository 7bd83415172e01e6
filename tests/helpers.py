"""Shared builders: the same octree / operator / inputs on the oracle and on the library."""
import numpy as np

import cirrus_amg as P
import oracle as O

C0 = 11          # coefficient channels 11..14, as FluidEuler (cirrus_cutcell/FluidParams.h)
R_MATRIX = 0.5   # AMGSolver(c0, 0.5, 1, 1) in FluidEuler::project

SPHERE_C2 = (0.5, 0.5, 0.5, 0.2)
TORUS_C3 = (0.5, 0.5, 0.5, 0.15, 0.05)

# name -> (target kind, min level, max level, sdf kind, sdf, bands, assembly)
CASES = {
    "uniform3": (O.TARGET_UNIFORM, 3, 3, 0, (), None, "plain"),
    "uniform2": (O.TARGET_UNIFORM, 2, 2, 0, (), None, "plain"),
    "sphere_empty24": (O.TARGET_SPHERE_EMPTY, 2, 4, 0, (), None, "plain"),
    "star_empty24": (O.TARGET_STAR_EMPTY, 2, 4, 0, (), None, "plain"),
    "sphere_solid24": (O.TARGET_SPHERE_SOLID, 2, 4, O.SDF_SPHERE, (0.5, 0.5, 0.5, 0.25), None, "corner"),
    "center_point14": (O.TARGET_CENTER_POINT, 1, 4, 0, (), None, "plain"),
    "c2_small": (O.TARGET_SDF_BANDS, 2, 4, O.SDF_SPHERE, SPHERE_C2, {4: 0.05, 3: 0.15}, "solid"),
    "c3_small": (O.TARGET_SDF_BANDS, 2, 5, O.SDF_TORUS, TORUS_C3, {5: 0.02, 4: 0.04, 3: 0.08}, "supersampled"),
    # the C2 octree itself (8,217 tiles; its finest level is large enough for sibling blocks)
    "c2_full": (O.TARGET_SDF_BANDS, 3, 5, O.SDF_SPHERE, SPHERE_C2, {5: 0.05, 4: 0.15}, "solid"),
}


def build_pair(name, with_device=True):
    """Return (oracle grid, library grid) with identical topology, cell types and operator."""
    kind, lo, hi, sk, sdf, bands, asm = CASES[name]
    to = O.make_target(kind, lo, hi, sk, sdf, bands)
    go = O.Grid(to)
    go.set_leaf_types_preset()
    tp = P.make_target(kind, lo, hi, sk, sdf, bands)
    gp = P.HADeviceGrid.test_grid(tp)
    if asm == "plain":
        go.coarsen_types_and_coeffs(C0, R_MATRIX)
        gp.prepare_types_and_coeffs(C0, R_MATRIX)
    elif asm == "solid":
        go.cutcell_coeffs(sk, sdf, -1, C0, R_MATRIX)
        gp.assemble_cutcell(sk, sdf, -1, C0, R_MATRIX)
    elif asm == "corner":
        go.set_wall_types(lo, True); gp.set_wall_types(lo, True)
        go.cutcell_coeffs(sk, sdf, 0, C0, R_MATRIX)
        gp.assemble_cutcell(sk, sdf, 0, C0, R_MATRIX)
    elif asm == "supersampled":
        go.set_wall_types(lo, True); gp.set_wall_types(lo, True)
        go.cutcell_coeffs(sk, sdf, 1, C0, R_MATRIX)
        gp.assemble_cutcell(sk, sdf, 1, C0, R_MATRIX)
    return go, gp


def random_fill(n_tiles, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, size=(n_tiles, 512)) * scale).astype(np.float32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)
