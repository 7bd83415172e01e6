"""A solver whose grid's operator changes between two solves must solve the new operator.

The solver keeps data derived from the operator across solves (the level-0 FP64
inverse, the FP16 coefficient copies of the cycle, the derived-diagonal check)
and rebuilds it when the coefficient channels or the cell types are written.

* Scaling every coefficient by 2 scales the solution by 1/2 exactly: every
  operation of the solve (FP16 copies, Jacobi, transfers, the Gauss-Jordan
  inverse of level 0, FP64 dots) commutes with a power-of-two scale, so the
  second solve on the same solver must return x1 / 2 bit for bit with the same
  iteration count.  A stale level-0 inverse or stale FP16 copies would not.
* A different cut-cell body (types and coefficients change, host or device
  assembly): the same solver and a freshly made one return the same bits.
"""
import numpy as np
import pytest

import cirrus_amg as P
import helpers as H

pytestmark = pytest.mark.gpu


def _opts(fp16):
    return P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                          cycle_coef_fp16=fp16, no_true_residual=1)


def _solve(solver, gp, b, fp16):
    gp.set_channel(1, b)
    iters, rel = solver.solve(opts=_opts(fp16), max_iters=100, relative_tolerance=1e-6)
    assert rel < 1e-6
    return iters, gp.channel(0).copy()


@pytest.mark.parametrize("fp16", [0, 1])
def test_scaled_operator_gives_half_the_solution(fp16):
    print("\nscale operator, fp16 cycle", fp16, flush=True)
    _, gp = H.build_pair("c2_small")
    gp.rhs_uniform(1, 5, True)
    b = gp.channel(1).copy()
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it1, x1 = _solve(solver, gp, b, fp16)
    for a in range(4):
        gp.set_channel(H.C0 + a, gp.channel(H.C0 + a) * np.float32(2.0))
    it2, x2 = _solve(solver, gp, b, fp16)
    assert it2 == it1
    assert np.array_equal(H.bits(x2), H.bits(x1 * np.float32(0.5)))


@pytest.mark.parametrize("fp16", [0, 1])
@pytest.mark.parametrize("device_assembly", [False, True])
def test_new_body_same_solver_matches_fresh_solver(fp16, device_assembly):
    print("\nnew body, fp16 cycle", fp16, "device assembly", device_assembly, flush=True)
    kind, lo, hi, sk, sdf, bands, _ = H.CASES["c2_small"]
    _, gp = H.build_pair("c2_small")
    gp.rhs_uniform(1, 6, True)
    b = gp.channel(1).copy()
    old = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    _, x_first = _solve(old, gp, b, fp16)
    types_before = gp.cell_types().copy()
    # a larger sphere: more NEUMANN cells, other face coefficients on every level
    gp.set_leaf_types_preset(P.make_target(kind, lo, hi, sk, sdf, bands))
    bigger = (0.5, 0.5, 0.5, 0.24)
    if device_assembly:
        gp.assemble_cutcell_device(sk, bigger, -1, H.C0, H.R_MATRIX)
    else:
        gp.assemble_cutcell(sk, bigger, -1, H.C0, H.R_MATRIX)
    assert not np.array_equal(gp.cell_types(), types_before)
    b = np.where(gp.cell_types() == 2, b, np.float32(0))   # right-hand side on the new INTERIOR cells only
    it_old, x_old = _solve(old, gp, b, fp16)
    fresh = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it_new, x_new = _solve(fresh, gp, b, fp16)
    assert it_old == it_new
    assert np.array_equal(H.bits(x_old), H.bits(x_new))
    assert not np.array_equal(H.bits(x_old), H.bits(x_first))
