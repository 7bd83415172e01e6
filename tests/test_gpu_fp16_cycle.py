"""The preconditioning cycle on FP16 coefficient copies (cycle_coef_fp16).

The device converts each coefficient channel to FP16 bits of c / h (h = the
tile's spacing, a power of two times h0, so the scale is exact) and its cycle
kernels read those.  The oracle's cycle run on coefficients rounded the same
way, fp16(c / h) * h, must give the same bits; PCG with that cycle keeps the
FP32 operator, so it reaches the same tolerance and the same solution (to the
FP64 solution's 1e-5) as the FP32 cycle."""
import numpy as np
import pytest

import cirrus_amg as P
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu


def _round_coefficients(go, h0=0.125):
    """c -> float32(float16(c / h)) * h on the oracle's four coefficient channels"""
    lv = go.tiles()[0]
    h = (np.float32(h0) / (2.0 ** lv).astype(np.float32))[:, None].astype(np.float32)
    for a in range(4):
        c = go.channel(H.C0 + a)
        c[:] = (c / h).astype(np.float16).astype(np.float32) * h


def test_rounding_is_not_the_identity():
    """The rounding changes the operator (so the bit-exact test below can fail)."""
    go, _ = H.build_pair("c3_small", with_device=False)
    before = [go.channel(H.C0 + a).copy() for a in range(4)]
    _round_coefficients(go)
    changed = sum(int((go.channel(H.C0 + a) != before[a]).sum()) for a in range(4))
    assert changed > 1000


@pytest.mark.parametrize("case", ["uniform3", "sphere_empty24", "c3_small"])
@pytest.mark.parametrize("kind", ["v_jacobi_rbgs_coarsest", "w_rbgs"])
def test_fp16_cycle_matches_oracle_on_rounded_coefficients(case, kind):
    print("\nfp16 cycle", case, kind, flush=True)
    go, gp = H.build_pair(case)
    go.rhs_uniform(1, 5, True)
    gp.set_channel(1, go.channel(1))
    _round_coefficients(go)
    if kind == "v_jacobi_rbgs_coarsest":
        popts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_iters=8,
                               cycle_coef_fp16=1)
        ocyc = O.make_cycle(mu=1, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_iters=8)
    else:
        popts = P.default_opts(cycle_coef_fp16=1)
        ocyc = O.make_cycle(mu=2, smoother=0, omega=1.5, level_iters=2, coarsest_iters=10)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    go.fas_cycle(ocyc, 4, 1, 3, H.C0)
    solver.cycle(popts, 4, 1, 3)
    zo, zp = go.channel(4), gp.channel(4)
    assert np.isfinite(zp).all()
    assert np.array_equal(H.bits(zo), H.bits(zp)), (case, kind, np.abs(zo - zp).max())


@pytest.mark.parametrize("case", ["c2_small", "c3_small"])
def test_pcg_with_fp16_cycle_reaches_fp32_solution(case):
    go, gp = H.build_pair(case)
    go.rhs_uniform(1, 14, True)
    b = go.channel(1).copy()
    A = go.assemble_csr(H.C0)
    x64, _, rr = A.bicgstab(b.ravel()[A.cell].astype(np.float64), 1e-13, 20000)
    assert rr < 1e-11
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    res = {}
    for q in (0, 1):
        gp.set_channel(1, b)
        o = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                           cycle_coef_fp16=q)
        it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=o)
        x = gp.channel(0).ravel()[A.cell].astype(np.float64)
        res[q] = (it, rel, float(solver.last_stats.true_rel_res), np.linalg.norm(x - x64) / np.linalg.norm(x64))
    print("\n", case, res, flush=True)
    assert res[1][1] < 1e-6 and res[1][2] < 2e-6
    assert res[1][0] <= res[0][0] + 1
    assert res[1][3] <= 1e-5


def test_fp16_cycle_only_as_a_preconditioner():
    """A stand-alone multigrid solve would converge to the FP16 operator: refused."""
    _, gp = H.build_pair("uniform3")
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    with pytest.raises(P.CmgError) as e:
        solver.mg_solve(P.default_opts(cycle_coef_fp16=1))
    assert e.value.status == 1
    with pytest.raises(P.CmgError):
        solver.solve(opts=P.default_opts(cycle_coef_fp16=2))


def test_fp16_copies_out_of_range_are_refused():
    """A coefficient whose c / h does not fit FP16 makes the solve fail with
    CMG_ERR_BAD_ARG instead of running the cycle on infinities."""
    go, gp = H.build_pair("uniform3")
    go.rhs_uniform(1, 2, True)
    gp.set_channel(1, go.channel(1))
    c = gp.channel(H.C0).copy()
    c[0, 0] = 1e6          # c / h = 6.4e7 on a level-3 tile, beyond FP16's 65504
    gp.set_channel(H.C0, c)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    with pytest.raises(P.CmgError) as e:
        solver.solve(opts=P.default_opts(cycle_coef_fp16=1))
    assert e.value.status == 1
    assert "FP16" in str(e.value)


def test_fp16_copies_of_tiny_diagonals_are_refused():
    """Coefficients scaled by 1e-6 (a dt / rho of that size): the diagonals' c / h
    (about 6e-6) fall below FP16's smallest normal, which would make the cycle divide
    by 0 or by a few bits; the solve is refused up front.  The same operator solves
    with the FP32 cycle."""
    go, gp = H.build_pair("uniform3")
    go.rhs_uniform(1, 2, True)
    b = go.channel(1).copy()
    gp.set_channel(1, b)
    for a in range(4):
        gp.set_channel(H.C0 + a, gp.channel(H.C0 + a) * np.float32(1e-6))
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    with pytest.raises(P.CmgError) as e:
        solver.solve(opts=P.default_opts(cycle_coef_fp16=1))
    assert e.value.status == 1
    assert "smallest normal" in str(e.value)
    gp.set_channel(1, b * np.float32(1e-6))
    it, rel = solver.solve(relative_tolerance=1e-6, opts=P.default_opts(cycle_coef_fp16=0))
    assert rel < 1e-6
