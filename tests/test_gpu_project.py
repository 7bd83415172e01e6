"""Projection wrappers on the GPU against the oracle: divergence, gradient, velocity mix."""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

CASES = ["sphere_empty24", "c2_small", "c3_small", "center_point14"]


def _load(go, gp, ch, values):
    go.channel(ch)[:] = values
    gp.set_channel(ch, values)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("with_coeff", [False, True])
def test_divergence_bitexact(case, with_coeff):
    print("\ndivergence", case, with_coeff, flush=True)
    go, gp = H.build_pair(case)
    for a in range(3):
        _load(go, gp, 5 + a, H.random_fill(go.N, 160 + a))
    c0 = H.C0 if with_coeff else -1
    go.divergence(5, 2, c0)
    gp.divergence(5, 2, c0)
    for ch in (2, 5, 6, 7):   # b, and the ghost / NONLEAF faces left in u
        assert np.array_equal(H.bits(go.channel(ch)), H.bits(gp.channel(ch))), (case, ch)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("weighted", [False, True])
def test_add_gradient_bitexact(case, weighted):
    print("\ngradient", case, weighted, flush=True)
    go, gp = H.build_pair(case)
    _load(go, gp, 0, H.random_fill(go.N, 170))
    for a in range(3):
        _load(go, gp, 5 + a, H.random_fill(go.N, 171 + a))
    go.add_gradient(0, H.C0, 5, weighted)
    gp.add_gradient(0, H.C0, 5, weighted)
    for ch in (5, 6, 7, 0):
        assert np.array_equal(H.bits(go.channel(ch)), H.bits(gp.channel(ch))), (case, ch)


@pytest.mark.parametrize("case", ["c2_small", "c3_small"])
def test_mix_velocity_bitexact(case):
    go, gp = H.build_pair(case)
    for a in range(3):
        _load(go, gp, 5 + a, H.random_fill(go.N, 180 + a))
        _load(go, gp, 8 + a, H.random_fill(go.N, 183 + a))
    go.mix_velocity(H.C0, 5, 8, 5)
    gp.mix_velocity(H.C0, 5, 8, 5)
    lv, co, ty = go.tiles()
    for a in range(3):
        assert np.array_equal(H.bits(go.channel(5 + a)[ty == 1]), H.bits(gp.channel(5 + a)[ty == 1])), (case, a)


def test_projection_on_device_removes_divergence():
    """FluidEuler::project's sequence on the device: b = div u, solve, u += grad p;
    the divergence left on cells away from the box faces is at rounding level."""
    import cirrus_amg as P
    go, gp = H.build_pair("c2_small")
    gp.rhs_divergence(1, 5, 7)
    before = gp.norm(-1, 1)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    opts = P.default_opts(precision=P.PREC_FP64_REFINE, outer_tolerance=1e-10, inner_tolerance=1e-4)
    solver.solve(max_iters=200, relative_tolerance=1e-6, opts=opts)
    gp.add_gradient(0, H.C0, 5, False)
    gp.divergence(5, 2, -1)
    import test_oracle as TO
    inner = TO._inner_cells(go)
    after = np.abs(gp.channel(2)[inner]).max()
    assert after <= 1e-5 * before, (before, after)
