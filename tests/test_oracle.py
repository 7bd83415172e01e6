"""CPU checks that pin the oracle: closed forms, invariants and hand-computed cases.

The reference ships no golden vectors or known-answer files for this path, so
the oracle is pinned by what the mathematics fixes: the DST-I solution of the
uniform Dirichlet operator, the closed-form coarse operators (-H, 6H), the
SplitMix64 test vector, FaceFluidRatio cases computed by hand, and the
topology invariants of a 2:1 graded octree with ghost tiles.
"""
import os

import numpy as np
import pytest

import helpers as H
import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_hash_matches_independent_formula():
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(-2**20, 2**20, 3))
        log2 = int(rng.integers(4, 21))
        ref = ((x * 1) ^ (y * 2654435761) ^ (z * 805459861)) & 0xFFFFFFFF & ((1 << log2) - 1)
        assert O.lib().og_hash(x, y, z, log2) == ref


def test_splitmix64_test_vector():
    import ctypes as C
    s = C.c_uint64(0)
    outs = [O.lib().og_splitmix64(C.byref(s)) for _ in range(3)]
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_counts_and_coarse_closed_form():
    g = O.Grid(O.make_target(O.TARGET_UNIFORM, 3, 3))
    lv, co, ty = g.tiles()
    assert [(int(((lv == l) & (ty == 1)).sum()), int(((lv == l) & (ty == 2)).sum()), int(((lv == l) & (ty == 4)).sum()))
            for l in range(4)] == [(0, 1, 0), (0, 8, 0), (0, 64, 0), (512, 0, 0)]
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    for l in range(4):
        h = g.hval(l)
        m = lv == l
        for a in range(3):
            assert np.all(g.channel(H.C0 + a)[m] == -h)
        assert np.all(g.channel(H.C0 + 3)[m] == 6 * h)   # (-h, 6h) -> (-H, 6H) with R = 1/2


def _dst_solution(n, h, bb):
    from scipy.fft import dstn, idstn
    k = np.arange(1, n + 1)
    lam1 = 2 - 2 * np.cos(np.pi * k / (n + 1))
    lam = h * (lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :])
    return idstn(dstn(bb, type=1) / lam, type=1)


@pytest.mark.parametrize("level", [1, 2])
def test_csr_matches_dst_golden(level):
    """Uniform Dirichlet operator = h * 7-point Laplacian; golden DST-I vectors in tests/golden."""
    gold = np.load(os.path.join(GOLDEN, "dst_uniform_l%d.npz" % level))
    g = O.Grid(O.make_target(O.TARGET_UNIFORM, level, level))
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    g.rhs_uniform(1, 0, True)
    A = g.assemble_csr(H.C0)
    lv, co, ty = g.tiles()
    t, v = A.cell // 512, A.cell % 512
    gi = co[t, 0] * 8 + (v >> 6); gj = co[t, 1] * 8 + ((v >> 3) & 7); gk = co[t, 2] * 8 + (v & 7)
    n = 8 << level
    bb = np.zeros((n, n, n)); bb[gi, gj, gk] = g.channel(1).ravel()[A.cell]
    assert np.array_equal(bb.astype(np.float32), gold["b"])
    x = gold["x"][gi, gj, gk]
    r = A.matvec(x) - bb[gi, gj, gk]
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(bb)
    assert A.asym() == 0.0


def test_graded_operator_is_nonsymmetric_at_tjunctions():
    g = O.Grid(O.make_target(O.TARGET_SPHERE_EMPTY, 2, 4))
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    A = g.assemble_csr(H.C0)
    # fine->coarse weight -h/2, coarse->fine weight -H/8 = -h/4: relative asymmetry h/4 / (6h) ... > 0
    assert A.asym() > 1e-3


def test_face_fluid_ratio_cases():
    r = O.lib().og_face_fluid_ratio
    assert r(1, 1, 1, 1) == 1.0
    assert r(-1, -1, -1, -1) == 0.0
    # straight cut through the middle of the face: corners 0,1 inside, 2,3 outside -> 1/2
    assert r(-1, -1, 1, 1) == pytest.approx(0.5)
    # one corner outside with both edges cut at 3/4 -> 0.5 * 0.75 * 0.75
    assert r(-1, -1, -1, 3) == pytest.approx(0.5 * 0.75 * 0.75)
    # one corner inside with both edges cut at 1/2 -> 1 - 0.5 * 0.5 * 0.5
    assert r(-1, 1, 1, 1) == pytest.approx(1 - 0.125)
    # tiny open fraction is floored at 0.1, zero stays zero
    assert r(-1, -1, -1, 0.01) == pytest.approx(0.1)


@pytest.mark.parametrize("name", ["sphere_empty24", "star_empty24", "c2_small", "c3_small", "center_point14"])
def test_topology_invariants(name):
    kind, lo, hi, sk, sdf, bands, asm = H.CASES[name]
    g = O.Grid(O.make_target(kind, lo, hi, sk, sdf, bands))
    lv, co, ty = g.tiles()
    nbr, par, ch = g.neighbors(), g.parents(), g.children()
    N = g.N
    for t in range(N):
        # neighbour symmetry
        for a in range(3):
            u = nbr[t, a]
            if u >= 0:
                assert nbr[u, a + 3] == t and lv[u] == lv[t]
                assert np.array_equal(co[u], co[t] - np.eye(3, dtype=np.int32)[a])
        if lv[t] > 0:
            assert par[t] >= 0 and np.array_equal(co[par[t]], co[t] >> 1)
        if ty[t] == O.NONLEAF:
            assert (ch[t] >= 0).all() and all(ty[c] & (O.LEAF | O.NONLEAF) for c in ch[t])
        if ty[t] == O.GHOST:
            # spawned only next to an active tile, under a LEAF parent
            assert ty[par[t]] == O.LEAF
            assert any(nbr[t, q] >= 0 and ty[nbr[t, q]] & (O.LEAF | O.NONLEAF) for q in range(6))
        if ty[t] == O.LEAF:
            # 2:1 grading: a same-level NONLEAF face neighbour has LEAF children on the shared face
            for q in range(6):
                u = nbr[t, q]
                if u >= 0 and ty[u] == O.NONLEAF:
                    a, side = q % 3, (0 if q >= 3 else 1)
                    for o in range(8):
                        if ((o >> a) & 1) == side:
                            assert ty[ch[u, o]] == O.LEAF
    # every empty face neighbour of an active tile whose would-be parent is a LEAF is a ghost
    for t in range(N):
        if ty[t] & (O.LEAF | O.NONLEAF) and lv[t] > 0:
            for q in range(6):
                if nbr[t, q] < 0:
                    a = q % 3
                    c = co[t].copy(); c[a] += -1 if q < 3 else 1
                    pc = c >> 1
                    pm = (lv == lv[t] - 1) & (co == pc).all(axis=1)
                    assert not (pm.any() and ty[np.where(pm)[0][0]] == O.LEAF)


def test_oracle_pcg_reference_cycle():
    g = O.Grid(O.make_target(O.TARGET_SPHERE_EMPTY, 2, 4))
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    g.rhs_uniform(1, 4, True)
    A = g.assemble_csr(H.C0)
    b = g.channel(1).ravel()[A.cell].astype(np.float64)
    x64, it, rr = A.bicgstab(b, 1e-13)
    iters, rel = g.pcg(O.make_cycle(), H.C0, 100, 1e-6)
    x = g.channel(0).ravel()[A.cell]
    assert rel < 1e-6 and iters <= 10
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-5


def test_oracle_fas_is_correction_scheme_on_uniform():
    """On a grid without coarse leaves a linear FAS V-cycle equals the correction scheme:
    the result is linear in b (V(b1 + b2) = V(b1) + V(b2) up to rounding)."""
    g = O.Grid(O.make_target(O.TARGET_UNIFORM, 2, 2))
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    cyc = O.make_cycle(mu=1, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_direct=1)
    outs = []
    rng = np.random.default_rng(1)
    b1 = (rng.uniform(-1, 1, (g.N, 512)) * (g.cell_types() == 2)).astype(np.float32)
    b2 = (rng.uniform(-1, 1, (g.N, 512)) * (g.cell_types() == 2)).astype(np.float32)
    for b in (b1, b2, b1 + b2):
        g.channel(1)[:] = b
        g.fas_cycle(cyc, 4, 1, 3, H.C0)
        outs.append(g.channel(4).astype(np.float64).copy())
    leaf = g.tiles()[2] == 1
    d = outs[2][leaf] - outs[0][leaf] - outs[1][leaf]
    assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(outs[2][leaf])


@pytest.mark.parametrize("name", ["c3_small", "c2_small", "sphere_solid24"])
def test_no_closed_interior_cells(name):
    """After assembly every LEAF INTERIOR cell has a positive diagonal and every
    connected fluid region touches a Dirichlet boundary (the system is regular)."""
    from scipy.sparse.csgraph import connected_components
    go, _ = H.build_pair(name)
    lv, co, ty = go.tiles()
    T, D = go.cell_types(), go.channel(H.C0 + 3)
    assert (D[(ty == 1)[:, None] & (T == 2)] > 0).all()
    A = go.assemble_csr(H.C0).to_scipy()
    n, lab = connected_components(A, directed=False)
    rowsum = np.asarray(A.sum(axis=1)).ravel()
    touches = np.zeros(n, bool)
    np.logical_or.at(touches, lab, rowsum > 1e-7 * A.diagonal())
    assert touches.all()


# ---- non-FAS transfers, VCycle / muCycle / FASFCycle, FASMuCycleSolve ---------

def _graded(name="sphere_empty24"):
    go, _ = H.build_pair(name, with_device=False)
    return go


def _centres(go, level_filter=None):
    """global cell index (i, j, k) of every cell: (N, 512, 3)"""
    lv, co, ty = go.tiles()
    v = np.arange(512)
    loc = np.stack([v >> 6, (v >> 3) & 7, v & 7], axis=1)
    return co[:, None, :] * 8 + loc[None, :, :]


def test_restrict_source_sums_interior_children():
    """parent = R * (sum of the INTERIOR children): a field linear in the global
    index restricts to 8 R times its mean over the 2^3 block, i.e. the linear
    function at the parent's centre (catches swapped axes or octants)."""
    go = _graded("c2_small")
    lv, co, ty = go.tiles()
    gi = _centres(go).astype(np.float64)
    T = go.cell_types()
    lin = lambda g: 3.0 * g[..., 0] - 2.0 * g[..., 1] + 0.5 * g[..., 2] + 1.0
    level = go.max_level
    go.channel(1)[:] = lin(gi).astype(np.float32)
    go.channel(1)[T != 2] = 0   # only INTERIOR cells contribute; keep the test exact
    before = go.channel(1).copy()
    R = 0.5
    O.lib().og_restrict_source(go.h, level, 1, np.float32(R))
    after = go.channel(1)
    par = go.parents()
    kids = np.where((lv == level) & ((ty == 1) | (ty == 2)))[0]
    checked = 0
    for t in kids[:40]:
        p = par[t]
        for q in range(0, 512, 37):
            c = gi[t, q].astype(int)
            blk = [(c[0] & ~1) + a for a in (0, 1)], [(c[1] & ~1) + b for b in (0, 1)], [(c[2] & ~1) + e for e in (0, 1)]
            loc = [np.where((gi[t] == (x, y, z)).all(axis=1))[0][0] for x in blk[0] for y in blk[1] for z in blk[2]]
            if not (T[t, loc] == 2).all():
                continue
            pc = c >> 1
            pv = np.where((gi[p] == pc).all(axis=1))[0][0]
            want = R * 8 * lin(np.array([c[0] & ~1, c[1] & ~1, c[2] & ~1]) + 0.5)
            assert abs(after[p, pv] - want) <= 1e-5 * max(1.0, abs(want)), (t, q)
            checked += 1
    assert checked > 50
    # tiles of other levels are untouched
    other = lv != level - 1
    assert np.array_equal(after[other], before[other])


def test_prolong_add_injects_parent_values():
    """x += P * parent x: with fine x = 0 every INTERIOR fine cell gets P times
    the value of the parent cell containing it; GHOST cells get the parent value."""
    go = _graded("sphere_empty24")
    lv, co, ty = go.tiles()
    gi = _centres(go)
    T = go.cell_types()
    level = go.max_level
    f = lambda g: (7.0 * g[..., 0] + 3.0 * g[..., 1] - g[..., 2]).astype(np.float32)
    x = np.where(lv[:, None] == level - 1, f(gi), 0).astype(np.float32)
    go.channel(0)[:] = x
    O.lib().og_prolong_add(go.h, level, 0, np.float32(2.0))
    out = go.channel(0)
    sel = lv == level
    par_val = f(gi >> 1)
    is_ghost = (ty == 4)[:, None]
    want = np.where(T == 2, np.where(is_ghost, par_val, 2.0 * par_val), 0)
    has_parent = go.parents() >= 0
    m = sel & has_parent
    assert np.array_equal(out[m], want[m].astype(np.float32))


def test_deduct_flux_is_the_operator_on_nonleaf_values():
    """DeductFlux with f = 0 equals minus the full operator applied to an x that
    is nonzero only on NONLEAF tiles of the level (the outer-NONLEAF terms)."""
    go = _graded("c2_small")
    lv, co, ty = go.tiles()
    level = go.max_level - 1
    rng = np.random.default_rng(5)
    # x on the LEAF children of NONLEAF tiles at `level`; accumulate gives the NONLEAF means
    par = go.parents()
    x = np.zeros((go.N, 512), np.float32)
    kids = (lv == level + 1) & (ty == 1) & (ty[par] == 2)
    x[kids] = rng.uniform(-1, 1, (int(kids.sum()), 512)).astype(np.float32)
    go.channel(0)[:] = x
    go.full_neg_lap(0, H.C0, 3)   # propagate + accumulate + operator on LEAF tiles
    y = go.channel(3).copy()
    xs = go.channel(0).copy()      # now holds the NONLEAF means
    go.channel(0)[:] = xs
    go.channel(5)[:] = 0
    O.lib().og_deduct_flux(go.h, level, 0, H.C0, 5, 6)
    rhs = go.channel(6)
    leafs = (lv == level) & (ty == 1)
    assert np.array_equal(rhs[leafs], -y[leafs])
    nz = np.count_nonzero(rhs[leafs])
    assert nz > 100


def test_deduct_flux_without_nonleaf_values_is_identity():
    go = _graded("c2_small")
    lv, co, ty = go.tiles()
    level = go.max_level - 1
    f = np.random.default_rng(6).uniform(-1, 1, (go.N, 512)).astype(np.float32)
    go.channel(0)[:] = np.where((ty == 2)[:, None], 0, np.random.default_rng(7).uniform(-1, 1, (go.N, 512))).astype(np.float32)
    go.channel(5)[:] = f
    O.lib().og_deduct_flux(go.h, level, 0, H.C0, 5, 6)
    T = go.cell_types()
    leafs = (lv == level) & (ty == 1)
    want = np.where(T == 2, f, 0)
    assert np.array_equal(go.channel(6)[leafs], want[leafs])


def test_mu_cycle_equals_fas_on_uniform():
    """Without coarse leaves the non-FAS muCycle (coarse x zeroed, no flux terms)
    is the correction scheme, which a linear FAS cycle equals up to rounding."""
    g = O.Grid(O.make_target(O.TARGET_UNIFORM, 3, 3))
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    g.rhs_uniform(1, 3, True)
    r = g.channel(1).copy()
    outs = []
    for kind in (O.CYCLE_FAS_MU, O.CYCLE_MU):
        g.channel(1)[:] = r
        cyc = O.make_cycle(mu=2, smoother=0, omega=1.5, level_iters=2, coarsest_iters=10, kind=kind)
        g.precondition(cyc, 4, 1, 3, H.C0)
        outs.append(g.channel(4).astype(np.float64).copy())
    leaf = g.tiles()[2] == 1
    d = outs[1][leaf] - outs[0][leaf]
    assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(outs[0][leaf])


@pytest.mark.parametrize("kind", [O.CYCLE_FAS_F, O.CYCLE_VCYCLE, O.CYCLE_MU])
def test_pcg_with_each_cycle_kind(kind):
    g = O.Grid(O.make_target(O.TARGET_SPHERE_EMPTY, 2, 4))
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    g.rhs_uniform(1, 4, True)
    A = g.assemble_csr(H.C0)
    b = g.channel(1).ravel()[A.cell].astype(np.float64)
    x64, it, rr = A.bicgstab(b, 1e-13)
    iters, rel = g.pcg(O.make_cycle(kind=kind), H.C0, 200, 1e-6)
    x = g.channel(0).ravel()[A.cell]
    assert rel < 1e-6, (kind, iters, rel)
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-5


def test_fas_mu_solve_converges_to_the_solution():
    g = O.Grid(O.make_target(O.TARGET_SPHERE_EMPTY, 2, 4))
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    g.rhs_uniform(1, 8, True)
    A = g.assemble_csr(H.C0)
    b = g.channel(1).ravel()[A.cell].astype(np.float64)
    x64, it, rr = A.bicgstab(b, 1e-13)
    iters, rel = g.fas_mu_solve(O.make_cycle(mu=2), H.C0, 100, 1e-6)
    assert rel < 1e-6 and 0 < iters < 100
    x = g.channel(0).ravel()[A.cell]
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-5
    # the reference returns the loop index at exit: one cycle fewer than it ran
    g.channel(1)[:] = 0
    g.channel(1).ravel()[A.cell] = b.astype(np.float32)
    # with max_iters = iters the loop ends without breaking: index = iters, not converged
    it2, rel2 = g.fas_mu_solve(O.make_cycle(mu=2), H.C0, iters, 1e-6)
    assert it2 == iters and rel2 >= 1e-6


# ---- geometric solvers (GMGSolver, CMGSolver) ------------------------------------

@pytest.mark.parametrize("mode", [O.GEO_GMG, O.GEO_CMG])
def test_geo_operator_is_the_7_point_laplacian_on_uniform(mode):
    """On a uniform grid both geometric operators reduce to h (6 x0 - sum x_nb)
    with x = 0 outside the box: the AMG operator of the same grid."""
    g = O.Grid(O.make_target(O.TARGET_UNIFORM, 2, 2))
    g.set_leaf_types_preset()
    g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
    x = np.random.default_rng(2).uniform(-1, 1, (g.N, 512)).astype(np.float32)
    g.channel(0)[:] = x
    g.full_neg_lap(0, H.C0, 3)
    amg = g.channel(3).astype(np.float64).copy()
    g.channel(0)[:] = x
    O.lib().og_geo_full_neg_lap(g.h, mode, 0, 4, 0)
    geo = g.channel(4).astype(np.float64)
    leaf = g.tiles()[2] == 1
    assert np.abs(geo[leaf] - amg[leaf]).max() <= 1e-6 * np.abs(amg[leaf]).max()


@pytest.mark.parametrize("mode", [O.GEO_GMG, O.GEO_CMG])
def test_geo_operator_annihilates_constants_away_from_dirichlet(mode):
    """A constant field (ghosts take their parents' constant value) has zero
    Laplacian at every INTERIOR cell whose face neighbours all exist and are
    not DIRICHLET."""
    go = _graded("sphere_empty24")
    lv, co, ty = go.tiles()
    go.channel(0)[:] = 3.0
    O.lib().og_geo_full_neg_lap(go.h, mode, 0, 3, 0)
    y = go.channel(3)
    T = go.cell_types()
    nbr = go.neighbors()
    v = np.arange(512)
    loc = np.stack([v >> 6, (v >> 3) & 7, v & 7], axis=1)
    checked = 0
    for t in np.where(ty == 1)[0]:
        ok = np.ones(512, bool)
        for a in range(3):
            for side, q in ((-1, a), (1, a + 3)):
                edge = loc[:, a] == (0 if side < 0 else 7)
                if nbr[t, q] < 0:
                    ok &= ~edge
                    continue
                nl = loc.copy()
                nl[:, a] = np.where(edge, 7 if side < 0 else 0, loc[:, a] + side)
                nv = (nl[:, 0] << 6) | (nl[:, 1] << 3) | nl[:, 2]
                nt = np.where(edge, nbr[t, q], t)
                ok &= T[nt, nv] != 1
        m = ok & (T[t] == 2)
        assert np.abs(y[t, m]).max(initial=0) <= 1e-5, t
        checked += int(m.sum())
    assert checked > 10000


@pytest.mark.parametrize("mode", [O.GEO_GMG, O.GEO_CMG])
def test_geo_pcg_solves_its_own_operator(mode):
    """PCG with the family's V-cycle converges to the solution of the family's
    operator, assembled here column by column from the oracle's own apply."""
    go = _graded("sphere_empty24")
    lv, co, ty = go.tiles()
    T = go.cell_types()
    cells = np.argwhere((ty == 1)[:, None] & (T == 2))
    rng = np.random.default_rng(9)
    b = np.zeros((go.N, 512), np.float32)
    b[cells[:, 0], cells[:, 1]] = rng.uniform(-1, 1, len(cells)).astype(np.float32)
    go.channel(1)[:] = b
    it, rel = go.geo_pcg(mode, max_iters=200, tol=1e-6)
    assert rel < 1e-6, (it, rel)
    x = go.channel(0)[cells[:, 0], cells[:, 1]].astype(np.float64)
    # residual of the FP32 solution with the family's operator, recomputed in FP64 from probes
    go.channel(0)[:] = 0
    go.channel(0)[cells[:, 0], cells[:, 1]] = x.astype(np.float32)
    O.lib().og_geo_full_neg_lap(go.h, mode, 0, 3, 0)
    ax = go.channel(3)[cells[:, 0], cells[:, 1]].astype(np.float64)
    bb = b[cells[:, 0], cells[:, 1]].astype(np.float64)
    assert np.linalg.norm(bb - ax) <= 2e-6 * np.linalg.norm(bb)


def test_gmg_damped_jacobi_converges():
    g = O.Grid(O.make_target(O.TARGET_UNIFORM, 1, 1))
    g.set_leaf_types_preset()
    g.rhs_uniform(1, 5, True)
    it, rel = g.gmg_damped_jacobi(5000, 1e-2, 2.0 / 3.0)
    assert rel < 1e-2 and 10 < it < 5000
    # the loop index at exit is returned; one iteration fewer does not converge
    g.rhs_uniform(1, 5, True)
    it2, rel2 = g.gmg_damped_jacobi(it, 1e-2, 2.0 / 3.0)
    assert it2 == it and rel2 >= 1e-2


def test_norm_closed_forms():
    """NormSync: a constant c on the selected cells has every norm |c|; a single
    nonzero value v gives Linf |v|, L1 |v| w / sum w and L2 |v| sqrt(w / sum w)."""
    go = _graded("sphere_empty24")
    lv, co, ty = go.tiles()
    T = go.cell_types()
    sel = (ty == 1)[:, None] & (T == 2)
    go.channel(2)[:] = np.where(sel, -2.5, 7.0).astype(np.float32)
    for order in (-1, 1, 2, 3):
        for vw in ((False, True) if order != -1 else (False,)):
            assert abs(go.norm(order, 2, vw) - 2.5) <= 1e-12 * 2.5, (order, vw)
    go.channel(2)[:] = 0
    t, v = np.argwhere(sel)[len(np.argwhere(sel)) // 2]
    go.channel(2)[t, v] = -4.0
    h = 0.125 / 2 ** lv
    w = (h ** 3)[:, None] * np.ones((1, 512))
    W = w[sel].sum()
    assert go.norm(-1, 2) == 4.0
    assert abs(go.norm(1, 2, True) - 4.0 * w[t, v] / W) <= 1e-12
    assert abs(go.norm(2, 2, True) - 4.0 * np.sqrt(w[t, v] / W)) <= 1e-12
    assert abs(go.norm(2, 2, False) - 4.0 / np.sqrt(sel.sum())) <= 1e-12


def test_flexible_cg_equals_pcg_with_a_fixed_preconditioner():
    """With a fixed (linear) preconditioner flexible CG's Polak-Ribiere beta equals
    Fletcher-Reeves in exact arithmetic: same iteration count (+-1), same answer."""
    out = []
    for flexible in (False, True):
        g = O.Grid(O.make_target(O.TARGET_SPHERE_EMPTY, 2, 4))
        g.set_leaf_types_preset()
        g.coarsen_types_and_coeffs(H.C0, H.R_MATRIX)
        g.rhs_uniform(1, 4, True)
        cyc = O.make_cycle(mu=1, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_direct=1)
        it, rel = g.pcg(cyc, H.C0, 100, 1e-8, flexible=flexible)
        out.append((it, rel, g.channel(0).astype(np.float64).copy()))
    (i0, r0, x0), (i1, r1, x1) = out
    assert r0 < 1e-8 and r1 < 1e-8 and abs(i0 - i1) <= 1
    assert np.linalg.norm(x1 - x0) <= 1e-6 * np.linalg.norm(x0)


# ---- projection wrappers ----------------------------------------------------------

def _inner_cells(go):
    """LEAF INTERIOR cells none of whose faces lies on the domain boundary"""
    lv, co, ty = go.tiles()
    T = go.cell_types()
    nbr = go.neighbors()
    v = np.arange(512)
    loc = np.stack([v >> 6, (v >> 3) & 7, v & 7], 1)
    inner = np.zeros((go.N, 512), bool)
    for t in np.where(ty == 1)[0]:
        ok = T[t] == 2
        for a in range(3):
            ok = ok & ~((loc[:, a] == 7) & (nbr[t, a + 3] < 0)) & ~((loc[:, a] == 0) & (nbr[t, a] < 0))
        inner[t] = ok
    return inner


@pytest.mark.parametrize("case", ["uniform3", "sphere_empty24", "c2_small"])
@pytest.mark.parametrize("weighted", [False, True])
def test_projection_removes_divergence(case, weighted):
    """div(u + grad p) = div u - A p: after solving A p = div u the divergence
    vanishes to rounding on every cell away from the box faces (whose outer
    faces carry no velocity), on graded and cut-cell grids alike."""
    go = _graded(case)
    go.rhs_divergence(1, 5, 7)
    b = go.channel(1).copy()
    A = go.assemble_csr(H.C0)
    x, it, rr = A.bicgstab(b.ravel()[A.cell].astype(np.float64), 1e-12, 20000)
    go.channel(0)[:] = 0
    go.channel(0).ravel()[A.cell] = x.astype(np.float32)
    go.add_gradient(0, H.C0, 5, weighted)
    go.divergence(5, 2, -1)
    inner = _inner_cells(go)
    before, after = np.abs(b[inner]).max(), np.abs(go.channel(2)[inner]).max()
    assert after <= 1e-5 * before, (case, before, after)


def test_divergence_with_coefficients_of_a_gradient_is_minus_the_operator():
    """On the uniform grid div_c(grad_w x) = -A x on cells away from the box."""
    go = _graded("uniform3")
    x = np.random.default_rng(3).uniform(-1, 1, (go.N, 512)).astype(np.float32)
    go.channel(0)[:] = x
    go.full_neg_lap(0, H.C0, 3)
    ax = go.channel(3).astype(np.float64).copy()
    for a in range(3):
        go.channel(5 + a)[:] = 0
    go.channel(0)[:] = x
    go.add_gradient(0, H.C0, 5, True)
    go.divergence(5, 2, H.C0)
    inner = _inner_cells(go)
    d = go.channel(2).astype(np.float64)
    # sum (u0 c0 h - u1 c1 h) with u = grad x, c = -h  ->  -h (6 x - sum x_nb) ... = -A x
    assert np.abs(d[inner] + ax[inner]).max() <= 1e-5 * np.abs(ax[inner]).max()


def test_mix_velocity_keeps_a_common_velocity():
    """r u + (1 - r) u = u for every fluid ratio r (the two branches of the mix)."""
    go = _graded("c2_small")
    lv, co, ty = go.tiles()
    for a in range(3):
        go.channel(5 + a)[:] = 1.5
        go.channel(8 + a)[:] = 1.5
    go.mix_velocity(H.C0, 5, 8, 2)
    for a in range(3):
        m = go.channel(2 + a)[ty == 1]
        assert np.abs(m - 1.5).max() <= 1e-6


# ---- learned tile smoother oracle ---------------------------------------------------

def test_net_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    w = rng.normal(0, 0.2, O.NET_W)
    x = rng.normal(0, 1, (2, 1000)).astype(np.float32)
    e = rng.normal(0, 1, (2, 1000)).astype(np.float32)
    _, g = O.net_loss_grad(w, x, e)
    for i in (3, 431, 432 + 17, 432 + 6911, 7344, 7775):   # each layer, first and last weights
        eps = 1e-6 * max(1.0, abs(w[i]))
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (O.net_loss_grad(wp, x, e, False)[0] - O.net_loss_grad(wm, x, e, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_net_is_positively_homogeneous_and_local_loss_is_a_laplacian():
    """No biases and ReLU: net(c x) = c net(x) for c > 0.  A perfect prediction
    (y = e on the tile) has zero loss; y - e constant on the volume leaves only
    the boundary layer of the interior Laplacian."""
    rng = np.random.default_rng(12)
    w = rng.normal(0, 0.3, O.NET_W)
    x = rng.normal(0, 1, 1000)
    y1, y2 = O.net_forward(w, x), O.net_forward(w, 2.5 * x)
    assert np.allclose(y2, 2.5 * y1, rtol=1e-12, atol=1e-12)
    e = y1.astype(np.float32)[None, :]
    loss, _ = O.net_loss_grad(w, x[None, :].astype(np.float32), e, False)
    assert loss < 1e-6 * np.sum(y1 ** 2)


def test_adam_first_step_is_a_signed_learning_rate():
    """At t = 1 the bias-corrected moments are g and g^2: w -= lr g / (|g| + eps)."""
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([3.0, -0.25, 0.0])
    m, v = np.zeros(3), np.zeros(3)
    O.adam_step(w, m, v, g, 1, lr=1e-3)
    assert np.allclose(w, [1.0 - 1e-3 * 3 / (3 + 1e-8), -2.0 + 1e-3 * 0.25 / (0.25 + 1e-8), 0.5], rtol=0, atol=1e-15)
