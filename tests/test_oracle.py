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
