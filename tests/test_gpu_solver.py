"""Multigrid cycles and PCG solves on the GPU against the oracle and the FP64 CSR oracle."""
import numpy as np
import pytest

import cirrus_amg as P
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu


def _rhs(go, gp, seed):
    go.rhs_uniform(1, seed, True)
    gp.set_channel(1, go.channel(1))
    return go.channel(1).copy()


def _opts(kind):
    if kind == "ref_w":      # FluidEuler::project settings (cirrus_cutcell/FluidEuler.cu:702-706)
        return P.default_opts(), O.make_cycle(mu=2, smoother=0, omega=1.5, level_iters=2, coarsest_iters=10)
    if kind == "v_jacobi_gs":  # V(2,2) damped Jacobi, RB-GS coarsest (deterministic end to end)
        return (P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_iters=8),
                O.make_cycle(mu=1, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_iters=8))
    if kind == "w_jacobi_gs":  # W(2,2) damped Jacobi: the prolongation rides on the second coarse visit only
        return (P.default_opts(smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_iters=8),
                O.make_cycle(mu=2, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_iters=8))
    if kind == "v_jacobi_direct":
        return (P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1),
                O.make_cycle(mu=1, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_direct=1))
    raise KeyError(kind)


@pytest.mark.parametrize("case", ["uniform3", "sphere_empty24", "c3_small", "c2_full"])
@pytest.mark.parametrize("kind,mask", [("ref_w", 0), ("ref_w", 16), ("v_jacobi_gs", 0), ("v_jacobi_gs", 16),
                                       ("w_jacobi_gs", 0), ("w_jacobi_gs", 16)])
def test_cycle_bitexact(case, kind, mask):
    """mask 16 keeps the FAS save and the prolongation in launches of their own and the
    coarsest iterations one launch per pass; by default a coarse level's first pre-sweep
    does the save, its last post-sweep the prolongation to the finer level, and a
    single-tile level 0 iterates in one workgroup.  Both give the oracle's bits."""
    print("\ncycle", case, kind, mask, flush=True)
    go, gp = H.build_pair(case)
    _rhs(go, gp, 5)
    popts, ocyc = _opts(kind)
    saved = P.get_kernel_forms()
    try:
        P.set_kernel_forms(mask, saved[1])
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        go.fas_cycle(ocyc, 4, 1, 3, H.C0)
        solver.cycle(popts, 4, 1, 3)
    finally:
        P.set_kernel_forms(*saved)
    zo, zp = go.channel(4), gp.channel(4)
    assert np.isfinite(zp).all()
    assert np.array_equal(H.bits(zo), H.bits(zp)), (case, kind, np.abs(zo - zp).max())


def test_cycle_direct_coarsest_close():
    go, gp = H.build_pair("uniform3")
    _rhs(go, gp, 6)
    popts, ocyc = _opts("v_jacobi_direct")
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    go.fas_cycle(ocyc, 4, 1, 3, H.C0)
    solver.cycle(popts, 4, 1, 3)
    zo, zp = go.channel(4), gp.channel(4)
    assert np.linalg.norm(zo - zp) <= 1e-5 * np.linalg.norm(zo)


def _fp64_solution(go):
    A = go.assemble_csr(H.C0)
    b = go.channel(1).ravel()[A.cell].astype(np.float64)
    x, it, rr = A.bicgstab(b, 1e-13, 20000)
    assert rr < 1e-11
    return A, b, x


@pytest.mark.parametrize("case", ["uniform3", "sphere_empty24", "c2_small", "c3_small"])
@pytest.mark.parametrize("kind", ["ref_w", "v_jacobi_direct"])
def test_pcg_matches_fp64_oracle(case, kind):
    print("\npcg", case, kind, flush=True)
    go, gp = H.build_pair(case)
    _rhs(go, gp, 7)
    A, b, x64 = _fp64_solution(go)
    popts, ocyc = _opts(kind)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, level_iters=2,
                           coarsest_iters=popts.coarsest_iters, opts=popts)
    st = solver.last_stats
    x = gp.channel(0).ravel()[A.cell].astype(np.float64)
    err = np.linalg.norm(x - x64) / np.linalg.norm(x64)
    # the oracle's own FP32 PCG for the iteration count
    oit, orel = go.pcg(ocyc, H.C0, 100, 1e-6)
    print(case, kind, "gpu iters", it, rel, "true", st.true_rel_res, "oracle iters", oit, "err", err, "ms", st.ms)
    assert rel < 1e-6 and st.true_rel_res < 5e-6
    assert abs(it - oit) <= 1
    assert err < 1e-5


@pytest.mark.parametrize("case", ["uniform3", "c3_small"])
def test_fp64_refinement(case):
    print("\nrefine", case, flush=True)
    go, gp = H.build_pair(case)
    _rhs(go, gp, 8)
    A, b, x64 = _fp64_solution(go)
    popts, _ = _opts("v_jacobi_direct")
    popts.precision = P.PREC_FP64_REFINE
    popts.outer_tolerance = 1e-10
    popts.inner_tolerance = 1e-4
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it, rel = solver.solve(max_iters=200, relative_tolerance=1e-10, opts=popts)
    st = solver.last_stats
    x = solver.x64().ravel()[A.cell]
    err = np.linalg.norm(x - x64) / np.linalg.norm(x64)
    res = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    print(case, "outer", st.outer_iters, "inner", it, "rel", rel, "csr res", res, "err", err)
    assert rel < 1e-10 and res < 1e-9
    assert err < 1e-8


def test_uniform_dst_closed_form():
    """Config-1 operator is h * (7-point Dirichlet Laplacian): DST-I gives the exact solution."""
    from scipy.fft import dstn, idstn
    go, gp = H.build_pair("uniform3")
    b = _rhs(go, gp, 0)
    lv, co, ty = go.tiles()
    n, h = 64, go.hval(3)
    leaf = np.where(ty == 1)[0]
    v = np.arange(512)
    gi = (co[leaf, 0, None] * 8 + (v >> 6)).ravel()
    gj = (co[leaf, 1, None] * 8 + ((v >> 3) & 7)).ravel()
    gk = (co[leaf, 2, None] * 8 + (v & 7)).ravel()
    bb = np.zeros((n, n, n)); bb[gi, gj, gk] = b[leaf].ravel()
    k = np.arange(1, n + 1); lam1 = 2 - 2 * np.cos(np.pi * k / (n + 1))
    lam = h * (lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :])
    xe = idstn(dstn(bb, type=1) / lam, type=1)
    popts, _ = _opts("v_jacobi_direct")
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it, rel = solver.solve(relative_tolerance=1e-6, opts=popts)
    x = gp.channel(0)[leaf].ravel()
    err = np.linalg.norm(x - xe[gi, gj, gk]) / np.linalg.norm(xe)
    print("dst iters", it, "err", err)
    assert err < 1e-5


def test_warm_start_and_zero_rhs():
    go, gp = H.build_pair("sphere_empty24")
    b = _rhs(go, gp, 9)
    popts, _ = _opts("v_jacobi_direct")
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it0, _ = solver.solve(relative_tolerance=1e-6, opts=popts)
    x0 = gp.channel(0)
    # a slightly different right-hand side from the previous solution
    gp.set_channel(1, b * np.float32(1.001))
    popts.warm_start = 1
    it1, rel1 = solver.solve(relative_tolerance=1e-6, opts=popts)
    assert it1 < it0 and rel1 < 1e-6
    # zero right-hand side: zero iterations, x = 0 (AMGSolver.cu:1415-1420)
    gp.set_channel(1, np.zeros_like(b))
    popts.warm_start = 0
    it2, rel2 = solver.solve(relative_tolerance=1e-6, opts=popts)
    assert it2 == 0 and rel2 == 0 and not gp.channel(0).any()
    del x0


def test_nonfinite_rhs_stops_the_solve():
    """A non-finite b stops the solve on the device (the |b|^2 reduction sets the stop flag)
    and the solve reports it (AMGSolver.cu:1410-1425 tests |b| first); the iterations launched
    on the published stopping test all see the flag.  The next solve with a finite b converges."""
    go, gp = H.build_pair("sphere_empty24")
    b = _rhs(go, gp, 13)
    popts, _ = _opts("v_jacobi_direct")
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    leaf = np.where(go.tiles()[2] == 1)[0]
    interior = np.where(go.cell_types()[leaf[0]] == 2)[0]
    bad = b.copy()
    bad[leaf[0], interior[0]] = np.nan
    gp.set_channel(1, bad)
    with pytest.raises(P.CmgError) as e:
        solver.solve(max_iters=50, relative_tolerance=1e-6, opts=popts)
    assert e.value.status == 4, e.value
    gp.set_channel(1, b)
    it, rel = solver.solve(max_iters=50, relative_tolerance=1e-6, opts=popts)
    assert 0 < it < 50 and rel < 1e-6


def test_pure_neumann():
    """All-Neumann box (tests/SolverTests.cu:789-805): singular operator, recentred residual."""
    go, gp = H.build_pair("uniform2")
    types = go.cell_types().copy()
    lv, co, ty = go.tiles()
    for t in np.where(ty == 1)[0]:
        for v in range(512):
            g = co[t] * 8 + np.array([v >> 6, (v >> 3) & 7, v & 7])
            if (g == 0).any() or (g == 31).any():
                types[t, v] = 4
    go2 = O.Grid(O.make_target(O.TARGET_UNIFORM, 2, 2))
    go2.cell_types()[:] = types
    go2.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    gp.set_cell_types(types)
    gp.prepare_types_and_coeffs(H.C0, H.R_MATRIX)
    go2.rhs_uniform(1, 3, True)
    gp.set_channel(1, go2.channel(1))
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it, rel = solver.solve(relative_tolerance=1e-6, is_pure_neumann=True)
    oit, orel = go2.pcg(O.make_cycle(), H.C0, 100, 1e-6, True)
    print("pure neumann gpu", it, rel, "oracle", oit, orel)
    assert rel < 1e-6 and abs(it - oit) <= 1


@pytest.mark.parametrize("forms", [(0, 256), (7, 512), (2, 128)], ids=["onetile_nt256", "persistent_nt512", "mixed_nt128"])
def test_pcg_same_across_kernel_forms(forms):
    """The kernel forms change only the grouping of the FP64 partial sums: the
    same solve converges in the same number of iterations to the same x."""
    print("\nforms", forms, flush=True)
    go, gp = H.build_pair("c3_small")
    b = _rhs(go, gp, 9)
    opts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1)
    saved = P.get_kernel_forms()
    try:
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        it0, rel0 = solver.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=opts)
        x0 = gp.channel(0).astype(np.float64)
        P.set_kernel_forms(*forms)
        gp.set_channel(1, b)
        it1, rel1 = solver.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=opts)
        x1 = gp.channel(0).astype(np.float64)
    finally:
        P.set_kernel_forms(*saved)
    assert rel0 < 1e-6 and rel1 < 1e-6
    assert abs(it1 - it0) <= 1, (it0, it1)
    assert np.linalg.norm(x1 - x0) <= 1e-5 * np.linalg.norm(x0)


def test_sync_stride_semantics():
    """AMGSolver::solve's sync_stride (src/AMGSolver.cu:1410-1549): the stopping
    test runs after iterations stride, 2*stride, ...; stride <= 0 runs exactly
    max_iters iterations and returns (max_iters, -1)."""
    go, gp = H.build_pair("sphere_empty24")
    b = _rhs(go, gp, 12)
    opts = P.default_opts()
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it1, rel1 = solver.solve(max_iters=100, relative_tolerance=1e-6, sync_stride=1, opts=opts)
    assert rel1 < 1e-6
    for stride in (2, 3, 5):
        gp.set_channel(1, b)
        it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, sync_stride=stride, opts=opts)
        assert it % stride == 0 and it1 <= it < it1 + stride and rel < 1e-6, (stride, it, it1)
    gp.set_channel(1, b)
    it, rel = solver.solve(max_iters=4, relative_tolerance=1e-6, sync_stride=-1, opts=opts)
    assert (it, rel) == (4, -1.0)
    x_nosync = gp.channel(0).copy()
    gp.set_channel(1, b)
    it, rel = solver.solve(max_iters=4, relative_tolerance=0.0, sync_stride=1, opts=opts)
    assert it == 4
    assert np.array_equal(H.bits(x_nosync), H.bits(gp.channel(0)))


def test_cpp_wrapper_solves():
    """The C++ caller (include/cirrus_amg.hpp) runs FluidEuler::project's solve
    call, the V-cycle form and the unsynchronised form on the 64^3 grid."""
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(here, "cpp", "_build", "test_amg_wrapper")
    assert os.path.exists(exe), "build tests/cpp first (make -C tests/cpp)"
    print("\ncpp wrapper gpu", flush=True)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "gpu part OK" in r.stdout


@pytest.mark.parametrize("case", ["uniform3", "sphere_empty24", "c3_small"])
@pytest.mark.parametrize("kind,smoother", [(P.CYCLE_FAS_F, 0), (P.CYCLE_VCYCLE, 0), (P.CYCLE_MU, 0),
                                           (P.CYCLE_FAS_F, 1), (P.CYCLE_MU, 1)])
def test_cycle_kinds_bitexact(case, kind, smoother):
    """FASFCycle, VCycle and muCycle (src/AMGSolver.cu:1059-1174,1358-1387) on the
    device give the oracle's bits."""
    print("\ncycle kind", case, kind, smoother, flush=True)
    go, gp = H.build_pair(case)
    rng = np.random.default_rng(31)
    for ch in (1, 3, 4, 5):
        v = rng.uniform(-1, 1, (go.N, 512)).astype(np.float32)
        go.channel(ch)[:] = v
        gp.set_channel(ch, v)
    okind = {P.CYCLE_FAS_F: O.CYCLE_FAS_F, P.CYCLE_VCYCLE: O.CYCLE_VCYCLE, P.CYCLE_MU: O.CYCLE_MU}[kind]
    omega = 1.5 if smoother == 0 else 2.0 / 3.0
    popts = P.default_opts(cycle=kind, smoother=smoother, omega=omega, coarsest_iters=8)
    ocyc = O.make_cycle(mu=2, smoother=smoother, omega=omega, level_iters=2, coarsest_iters=8, kind=okind)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    go.precondition(ocyc, 4, 1, 3, H.C0)
    solver.cycle(popts, 4, 1, 3)
    for ch in (4, 3, 1):
        a, b = go.channel(ch), gp.channel(ch)
        assert np.array_equal(H.bits(a), H.bits(b)), (case, kind, smoother, ch, np.abs(a - b).max())


@pytest.mark.parametrize("case,kind", [("sphere_empty24", P.CYCLE_FAS_F), ("sphere_empty24", P.CYCLE_VCYCLE),
                                       ("sphere_empty24", P.CYCLE_MU), ("c3_small", P.CYCLE_FAS_F),
                                       ("c3_small", P.CYCLE_VCYCLE), ("c3_small", P.CYCLE_MU)])
def test_pcg_each_cycle_kind_matches_oracle(case, kind):
    """PCG preconditioned by each cycle kind follows the oracle's PCG (same
    iteration count +-1, same residual reached); where it converges the
    solution is within 1e-5 of the FP64 solution."""
    go, gp = H.build_pair(case)
    _rhs(go, gp, 14)
    A, _, x64 = _fp64_solution(go)
    okind = {P.CYCLE_FAS_F: O.CYCLE_FAS_F, P.CYCLE_VCYCLE: O.CYCLE_VCYCLE, P.CYCLE_MU: O.CYCLE_MU}[kind]
    go2, _ = H.build_pair(case, with_device=False)
    go2.channel(1)[:] = go.channel(1)
    oit, orel = go2.pcg(O.make_cycle(kind=okind), H.C0, 100, 1e-6)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, opts=P.default_opts(cycle=kind))
    print("\npcg kind", case, kind, it, rel, oit, orel, flush=True)
    assert abs(it - oit) <= 1, (it, oit)
    if orel < 1e-6:
        assert rel < 1e-6
        x = gp.channel(0).ravel()[A.cell]
        assert np.linalg.norm(x - x64) <= 1e-5 * np.linalg.norm(x64)
    else:
        assert abs(np.log10(rel) - np.log10(orel)) < 0.5, (rel, orel)


@pytest.mark.parametrize("case", ["uniform3", "c3_small"])
def test_mg_solve_matches_oracle(case):
    """FASMuCycleSolve (src/AMGSolver.cu:1311-1356): the same cycles, the same
    reference return value, the same solution bits."""
    print("\nmg solve", case, flush=True)
    go, gp = H.build_pair(case)
    _rhs(go, gp, 21)
    opts = P.default_opts()
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it, rel = solver.mg_solve(opts, max_iters=100, relative_tolerance=1e-6)
    oit, orel = go.fas_mu_solve(O.make_cycle(mu=2), H.C0, 100, 1e-6)
    assert rel < 1e-6 and it == oit, (it, oit, rel, orel)
    assert abs(rel - orel) <= 1e-6 * orel
    leaf = go.tiles()[2] == 1
    assert np.array_equal(H.bits(go.channel(0)[leaf]), H.bits(gp.channel(0)[leaf]))


@pytest.mark.parametrize("case", ["sphere_empty24", "c3_small"])
def test_flexible_cg_matches_oracle(case):
    go, gp = H.build_pair(case)
    _rhs(go, gp, 17)
    A, _, x64 = _fp64_solution(go)
    go2, _ = H.build_pair(case, with_device=False)
    go2.channel(1)[:] = go.channel(1)
    cyc = O.make_cycle(mu=1, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_direct=1)
    oit, orel = go2.pcg(cyc, H.C0, 100, 1e-6, flexible=True)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    opts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1, flexible=1)
    it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=opts)
    assert rel < 1e-6 and abs(it - oit) <= 1, (it, oit)
    x = gp.channel(0).ravel()[A.cell]
    assert np.linalg.norm(x - x64) <= 1e-5 * np.linalg.norm(x64)


@pytest.mark.parametrize("case", ["sphere_empty24", "c2_small", "c3_small"])
def test_derived_diagonals_change_nothing(case):
    """Solves that derive LEAF diagonals from the face coefficients give the bits
    of solves that load the stored diagonals."""
    go, gp = H.build_pair(case)
    b = _rhs(go, gp, 23)
    opts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1)
    saved = P.get_kernel_forms()
    out = []
    try:
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        for mask in (0, 8):
            P.set_kernel_forms(mask, 256)
            gp.set_channel(1, b)
            it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=opts)
            out.append((it, rel, gp.channel(0).copy()))
    finally:
        P.set_kernel_forms(*saved)
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(H.bits(out[0][2]), H.bits(out[1][2]))


@pytest.mark.parametrize("cycle,flexible", [(P.CYCLE_FAS_V, 0), (P.CYCLE_FAS_W, 0), (P.CYCLE_FAS_V, 1)],
                         ids=["fas_v", "fas_w", "fas_v_flexible"])
@pytest.mark.parametrize("case", ["sphere_empty24", "c3_small"])
def test_fused_boundary_kernels(case, cycle, flexible):
    """Fusing the preconditioner's first finest-level Jacobi sweep into the PCG
    x/r update gives the bits of the separate launch; summing r.z in the last
    finest-level sweep (non-flexible PCG) regroups only FP64 partial sums."""
    go, gp = H.build_pair(case)
    b = _rhs(go, gp, 29)
    opts = P.default_opts(cycle=cycle, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1, flexible=flexible)
    saved = P.get_kernel_forms()
    out = []
    try:
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        for mask in (0, 16):
            P.set_kernel_forms(mask, 256)
            gp.set_channel(1, b)
            it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=opts)
            out.append((it, rel, gp.channel(0).copy()))
    finally:
        P.set_kernel_forms(*saved)
    assert out[0][1] < 1e-6
    assert out[0][0] == out[1][0]
    if flexible:
        assert out[0][1] == out[1][1]
        assert np.array_equal(H.bits(out[0][2]), H.bits(out[1][2]))
    else:
        assert abs(out[0][1] - out[1][1]) <= 1e-3 * out[1][1]
        x0, x1 = out[0][2].astype(np.float64), out[1][2].astype(np.float64)
        assert np.linalg.norm(x0 - x1) <= 1e-5 * np.linalg.norm(x1)


def test_solver_refuses_a_refinalized_grid():
    """A solver's buffers are sized for the grid it was made on; after the grid is
    finalized again every solver call is refused (CMG_ERR_STATE), a new solver works."""
    go, gp = H.build_pair("uniform3")
    _rhs(go, gp, 3)
    s = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    s.solve()
    gp.finalize()
    gp.set_leaf_types_preset(P.make_target(P.TARGET_UNIFORM, 3, 3))
    gp.prepare_types_and_coeffs(H.C0, H.R_MATRIX)
    gp.set_channel(1, go.channel(1))
    with pytest.raises(P.CmgError) as e:
        s.solve()
    assert e.value.status == 2
    with pytest.raises(P.CmgError):
        s.cycle(P.default_opts(), 4, 1, 3)
    it, rel = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0).solve()
    assert rel < 1e-6


def test_true_residual_can_be_skipped():
    """no_true_residual = 1 leaves the solve itself unchanged (the same bits of x, the same
    iterations and recursive residual) and reports true_rel_res = -1; the default computes it."""
    go, gp = H.build_pair("c2_small")
    go.rhs_uniform(1, 21, True)
    b = go.channel(1).copy()
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    out = []
    for skip in (0, 1, 0):
        print("\nno_true_residual", skip, flush=True)
        gp.set_channel(1, b)
        o = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                           cycle_coef_fp16=1, no_true_residual=skip)
        it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=o)
        out.append((it, rel, solver.last_stats.true_rel_res, gp.channel(0).copy()))
    assert out[0][0] == out[1][0] == out[2][0] and out[0][1] == out[1][1] == out[2][1]
    assert np.array_equal(H.bits(out[0][3]), H.bits(out[1][3]))
    assert np.array_equal(H.bits(out[0][3]), H.bits(out[2][3]))
    assert 0 < out[0][2] < 2e-6 and out[1][2] == -1 and out[2][2] == out[0][2]


@pytest.mark.parametrize("case", ["sphere_empty24", "c2_small", "c3_small"])
@pytest.mark.parametrize("q16", [0, 1])
def test_pulled_prolongation_changes_no_bit(case, q16):
    """By default a PCG solve's finest post-sweeps compute the FAS prolongation into the
    finest level on the fly (forms bit 5 set: fas_prolong_v writes it first).  The LEAF
    values the sweeps see are the same bits either way, so the whole solve is: the same
    iterations, the same residual, the same solution bits."""
    print("\npulled prolongation", case, q16, flush=True)
    go, gp = H.build_pair(case)
    b = _rhs(go, gp, 27)
    opts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                          cycle_coef_fp16=q16)
    saved = P.get_kernel_forms()
    out = []
    try:
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        for mask in (0, 32):
            P.set_kernel_forms(mask, saved[1])
            gp.set_channel(1, b)
            it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=opts)
            out.append((it, rel, gp.channel(0).copy()))
    finally:
        P.set_kernel_forms(*saved)
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1], (out[0][:2], out[1][:2])
    assert np.array_equal(H.bits(out[0][2]), H.bits(out[1][2]))


@pytest.mark.parametrize("case", ["uniform3", "c3_small"])
@pytest.mark.parametrize("cycle,iters", [(P.CYCLE_FAS_V, 1), (P.CYCLE_FAS_V, 3), (P.CYCLE_FAS_W, 2), (P.CYCLE_FAS_W, 3)])
def test_fused_cycle_forms_match_unfused(case, cycle, iters):
    """Every fused form of the FAS cycle with damped Jacobi (the FAS save in a coarse level's
    first sweep, the prolongation carried by the coarser level's last sweep, levels 1-0 in one
    workgroup around the direct solve) against the unfused launches (forms bit 4), bit for bit,
    for odd sweep counts (the ping-pong ends in the scratch buffer) and for W-cycles (two coarse
    visits, the direct solve run twice between the bottom workgroup's phases)."""
    go, gp = H.build_pair(case)
    _rhs(go, gp, 31)
    popts = P.default_opts(cycle=cycle, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1)
    popts.level_iters = iters
    saved = P.get_kernel_forms()
    out = []
    try:
        solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
        for mask in (0, 16):
            P.set_kernel_forms(mask, saved[1])
            gp.set_channel(1, go.channel(1))
            solver.cycle(popts, 4, 1, 3)
            out.append(gp.channel(4).copy())
    finally:
        P.set_kernel_forms(*saved)
    assert np.isfinite(out[0]).all()
    assert np.array_equal(H.bits(out[0]), H.bits(out[1])), (case, cycle, iters)
