"""Exact coefficient codes (kernel forms bit 6 clear, the default).

Per level, the stencil kernels may read a cell's face coefficients as indices into a table
of that level's distinct FP32 values (with the cell type in the same 32-bit word): one-byte
codes for at most 255 distinct values, ten-bit codes for at most 1023.  The table holds the
values' own bits, so every result must equal, bit for bit, the one computed from the FP32
channels (bit 6 set); the levels coded, and those with ten-bit codes, must be exactly the
ones the distinct bit patterns of their face coefficients call for.
"""
import numpy as np
import pytest

import cirrus_amg as P
import helpers as H

pytestmark = pytest.mark.gpu

CODES_OFF = 64


def _expected_codes(go):
    """(coded, ten-bit) level masks: bit l of the first set when level l's three face-coefficient
    channels hold at most 1023 distinct bit patterns (and, above 255, its cell types are 0,
    DIRICHLET, INTERIOR or NEUMANN), of the second when there are more than 255"""
    lv = go.tiles()[0]
    types = go.cell_types()
    coded = wide = 0
    for level in range(int(lv.max()) + 1):
        rows = lv == level
        keys = np.unique(np.concatenate([go.channel(H.C0 + a)[rows].view(np.uint32).ravel() for a in range(3)]))
        if len(keys) <= 255:
            coded |= 1 << level
        elif len(keys) <= 1023 and np.isin(types[rows], (0, 1, 2, 4)).all():
            coded |= 1 << level
            wide |= 1 << level
    return coded, wide


def _expected_levels(go):
    return _expected_codes(go)[0]


def _rhs(go, gp, seed):
    go.rhs_uniform(1, seed, True)
    gp.set_channel(1, go.channel(1))
    return go.channel(1).copy()


SOLVES = {
    # FluidEuler::project settings (cirrus_cutcell/FluidEuler.cu:702-706): FAS W, RB-GS 1.5
    "fluideuler_w": dict(),
    # V(2,2) damped Jacobi, direct level 0: the fused and pulled sweep forms
    "v_jacobi_direct": dict(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1),
    # the same with the cycle on FP16 copies: only the PCG operator reads the codes
    "v_jacobi_fp16": dict(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                          cycle_coef_fp16=1),
}


@pytest.mark.parametrize("case", ["uniform3", "sphere_solid24", "c3_small", "c2_full"])
@pytest.mark.parametrize("kind", sorted(SOLVES))
def test_coded_solve_matches_fp32_channels(case, kind):
    """A whole PCG solve reading the codes against the same solve on the FP32 channels, with
    one thread count for both (the per-tile partial sums group by threads)."""
    print("\ncoded", case, kind, flush=True)
    go, gp = H.build_pair(case)
    b = _rhs(go, gp, 41)
    opts = P.default_opts(**SOLVES[kind])
    saved = P.get_kernel_forms()
    out = []
    try:
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        for mask in (0, CODES_OFF):
            P.set_kernel_forms(mask, 128)
            gp.set_channel(1, b)
            it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, opts=opts)
            out.append((it, rel, gp.channel(0).copy(), solver.code_levels()))
    finally:
        P.set_kernel_forms(*saved)
    (it0, rel0, x0, m0), (it1, rel1, x1, m1) = out
    # sphere_solid24's finest level (674 distinct values) takes ten-bit codes; the others one-byte codes
    assert m0 == _expected_codes(go) and m0[0], (case, m0)
    assert bool(m0[1]) == (case == "sphere_solid24"), (case, m0)
    assert m1 == (0, 0)
    assert rel0 < 1e-6 and np.isfinite(x0).all()
    assert (it0, rel0) == (it1, rel1)
    assert np.array_equal(H.bits(x0), H.bits(x1)), (case, kind)


@pytest.mark.parametrize("case", ["sphere_empty24", "c3_small"])
def test_mixed_coded_and_fp32_levels(case):
    """Random face coefficients on the finest level leave it uncoded while the coarser levels
    keep their codes: the cycle (per-level forms) and the operator (per-tile forms in one
    launch) still give the FP32 channels' bits."""
    go, gp = H.build_pair(case)
    L = go.max_level
    lv = go.tiles()[0]
    rng = np.random.default_rng(7)
    for a in range(3):
        c = go.channel(H.C0 + a).copy()
        fine = lv == L
        c[fine] = c[fine] * rng.uniform(0.5, 1.0, size=c[fine].shape).astype(np.float32)
        gp.set_channel(H.C0 + a, c)
        go.channel(H.C0 + a)[:] = c
    b = _rhs(go, gp, 43)
    x = H.random_fill(go.N, 44)
    opts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_iters=4)
    saved = P.get_kernel_forms()
    out = []
    try:
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        for mask in (0, CODES_OFF):
            P.set_kernel_forms(mask, 128)
            gp.set_channel(1, b)
            solver.cycle(opts, 4, 1, 3)
            z = gp.channel(4).copy()
            gp.set_channel(0, x)
            solver.apply(0, 3)
            out.append((z, gp.channel(3).copy(), solver.coded_levels()))
    finally:
        P.set_kernel_forms(*saved)
    (z0, y0, m0), (z1, y1, m1) = out
    expected = _expected_levels(go)
    assert not expected & (1 << L) and expected, hex(expected)
    assert m0 == expected and m1 == 0
    leaf = go.tiles()[2] == 1
    assert np.isfinite(z0).all()
    assert np.array_equal(H.bits(z0), H.bits(z1)), case
    assert np.array_equal(H.bits(y0[leaf]), H.bits(y1[leaf])), case


def test_codes_follow_coefficient_writes():
    """Writing a coefficient channel drops the codes (coded_levels 0); the next solve builds
    them from the new values and matches the FP32 channels of the new operator."""
    go, gp = H.build_pair("c3_small")
    b = _rhs(go, gp, 45)
    opts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1)
    saved = P.get_kernel_forms()
    try:
        P.set_kernel_forms(0, 128)
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        solver.solve(max_iters=100, relative_tolerance=1e-6, opts=opts)
        assert solver.coded_levels() == _expected_levels(go)
        # the operator times 2 (exact): new values on every level
        for a in range(4):
            gp.set_channel(H.C0 + a, go.channel(H.C0 + a) * np.float32(2.0))
        assert solver.coded_levels() == 0
        gp.set_channel(1, b)
        it0, rel0 = solver.solve(max_iters=100, relative_tolerance=1e-6, opts=opts)
        x0 = gp.channel(0).copy()
        assert solver.coded_levels() == _expected_levels(go)
        P.set_kernel_forms(CODES_OFF, 128)
        gp.set_channel(1, b)
        it1, rel1 = solver.solve(max_iters=100, relative_tolerance=1e-6, opts=opts)
        x1 = gp.channel(0).copy()
    finally:
        P.set_kernel_forms(*saved)
    assert (it0, rel0) == (it1, rel1)
    assert np.array_equal(H.bits(x0), H.bits(x1))


@pytest.mark.parametrize("case", ["c3_small", "sphere_empty24"])
def test_ten_bit_codes_on_the_finest_level(case):
    """The finest level's face coefficients redrawn from 700 values (ten-bit codes there, one-byte
    codes below): the cycle's sweeps, restriction and FAS steps on the ten-bit words, and the
    operator (FP32 channels on that level), give the FP32 channels' bits."""
    go, gp = H.build_pair(case)
    L = go.max_level
    lv = go.tiles()[0]
    fine = lv == L
    rng = np.random.default_rng(11)
    for a in range(3):
        c = go.channel(H.C0 + a).copy()
        scale = np.float32(np.abs(c[fine]).max())
        pool = (scale * np.linspace(0.5, 1.0, 700)).astype(np.float32)
        f = c[fine]
        pick = pool[rng.integers(0, 700, size=f.shape)]
        c[fine] = np.where(f != 0, np.where(f < 0, -pick, pick), f).astype(np.float32)
        gp.set_channel(H.C0 + a, c)
        go.channel(H.C0 + a)[:] = c
    coded, wide = _expected_codes(go)
    assert wide == 1 << L and coded & (1 << L), (hex(coded), hex(wide))
    b = _rhs(go, gp, 47)
    x = H.random_fill(go.N, 48)
    saved = P.get_kernel_forms()
    runs = {
        "v_jacobi": P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_iters=4),
        "w_rbgs": P.default_opts(),
    }
    out = []
    try:
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        for mask in (0, CODES_OFF):
            P.set_kernel_forms(mask, 128)
            res = []
            for name in sorted(runs):
                gp.set_channel(1, b)
                solver.cycle(runs[name], 4, 1, 3)
                res.append(gp.channel(4).copy())
            gp.set_channel(0, x)
            solver.apply(0, 3)
            res.append(gp.channel(3).copy())
            out.append((res, solver.code_levels()))
    finally:
        P.set_kernel_forms(*saved)
    (r0, m0), (r1, m1) = out
    assert m0 == (coded, wide) and m1 == (0, 0)
    leaf = go.tiles()[2] == 1
    for k, (a0, a1) in enumerate(zip(r0, r1)):
        if k == len(r0) - 1:
            a0, a1 = a0[leaf], a1[leaf]
        assert np.isfinite(a0).all()
        assert np.array_equal(H.bits(a0), H.bits(a1)), (case, k)
