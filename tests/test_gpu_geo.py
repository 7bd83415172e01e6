"""GMGSolver and CMGSolver on the GPU against the oracle (src/GMGSolver.cu, src/CMGSolver.cu)."""
import numpy as np
import pytest

import cirrus_amg as P
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

FAMILIES = [(P.GMGSolver, O.GEO_GMG), (P.CMGSolver, O.GEO_CMG)]
IDS = ["gmg", "cmg"]
CASES = ["uniform3", "sphere_empty24", "star_empty24", "c3_small"]


def _load(go, gp, ch, values):
    go.channel(ch)[:] = values
    gp.set_channel(ch, values)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("family", FAMILIES, ids=IDS)
def test_geo_full_operator_bitexact(case, family):
    cls, mode = family
    print("\ngeo apply", case, mode, flush=True)
    go, gp = H.build_pair(case)
    solver = cls(gp)
    _load(go, gp, 0, H.random_fill(go.N, 41))
    O.lib().og_geo_full_neg_lap(go.h, mode, 0, 3, 0)
    solver.apply(0, 3)
    leaf = go.tiles()[2] == 1
    assert np.array_equal(H.bits(go.channel(3)[leaf]), H.bits(gp.channel(3)[leaf])), (case, mode)
    # the propagated ghost (and CMG's accumulated parent) values stay in x, as in the reference
    assert np.array_equal(H.bits(go.channel(0)), H.bits(gp.channel(0))), (case, mode)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("family", FAMILIES, ids=IDS)
def test_geo_vcycle_bitexact(case, family):
    cls, mode = family
    print("\ngeo vcycle", case, mode, flush=True)
    go, gp = H.build_pair(case)
    solver = cls(gp)
    for ch, seed in ((1, 51), (3, 52), (4, 53)):
        _load(go, gp, ch, H.random_fill(go.N, seed))
    O.lib().og_geo_vcycle(go.h, mode, 4, 1, 3, 2, 10, np.float32(1.0), np.float32(1.0))
    solver.cycle(P.default_opts(level_iters=2, coarsest_iters=10), 4, 1, 3)
    for ch in (4, 3):
        assert np.array_equal(H.bits(go.channel(ch)), H.bits(gp.channel(ch))), (case, mode, ch)


@pytest.mark.parametrize("case", ["sphere_empty24", "c3_small"])
@pytest.mark.parametrize("family", FAMILIES, ids=IDS)
def test_geo_pcg_matches_oracle(case, family):
    cls, mode = family
    print("\ngeo pcg", case, mode, flush=True)
    go, gp = H.build_pair(case)
    go.rhs_uniform(1, 61, True)
    gp.set_channel(1, go.channel(1))
    solver = cls(gp)
    it, rel = solver.solve(max_iters=200, relative_tolerance=1e-6, level_iters=2, coarsest_iters=10)
    oit, orel = go.geo_pcg(mode, max_iters=200, tol=1e-6)
    print(it, rel, oit, orel, flush=True)
    assert abs(it - oit) <= 1 and rel < 1e-6 and orel < 1e-6
    leaf = go.tiles()[2] == 1
    xo, xp = go.channel(0)[leaf].astype(np.float64), gp.channel(0)[leaf].astype(np.float64)
    assert np.linalg.norm(xo - xp) <= 1e-5 * np.linalg.norm(xo)


def test_gmg_damped_jacobi_matches_oracle():
    go, gp = H.build_pair("uniform2")
    go.rhs_uniform(1, 71, True)
    gp.set_channel(1, go.channel(1))
    solver = P.GMGSolver(gp)
    it, rel = solver.damped_jacobi_solve(3000, 1e-2, 2.0 / 3.0)
    oit, orel = go.gmg_damped_jacobi(3000, 1e-2, 2.0 / 3.0)
    assert it == oit and rel < 1e-2
    assert abs(rel - orel) <= 1e-6 * orel
    assert np.array_equal(H.bits(go.channel(0)), H.bits(gp.channel(0)))


def test_geo_solver_rejects_amg_only_options():
    _, gp = H.build_pair("uniform2")
    solver = P.CMGSolver(gp)
    with pytest.raises(P.CmgError):
        solver.solve(opts=P.default_opts(precision=P.PREC_FP64_REFINE, outer_tolerance=1e-10))
    with pytest.raises(P.CmgError):
        solver.mg_solve(P.default_opts())
