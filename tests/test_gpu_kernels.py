"""Kernel-by-kernel parity of the HIP library against the oracle (bit-exact).

Each test feeds identical inputs (random values on every tile, including the
GHOST and NONLEAF tiles a kernel must ignore or read) to the oracle function
and to the device kernel, and compares the outputs bit for bit.
"""
import numpy as np
import pytest

import helpers as H

pytestmark = pytest.mark.gpu

KERNEL_CASES = ["uniform3", "sphere_empty24", "sphere_solid24", "c2_small", "c3_small", "center_point14", "c2_full"]


@pytest.fixture(scope="module", params=KERNEL_CASES)
def pair(request):
    print("\ncase", request.param, flush=True)
    go, gp = H.build_pair(request.param)
    # the device takes the operator the host assembled; make the channels identical anyway
    for a in range(4):
        gp.set_channel(H.C0 + a, go.channel(H.C0 + a))
    return request.param, go, gp


def _load(go, gp, ch, values):
    go.channel(ch)[:] = values
    gp.set_channel(ch, values)


def _leaf_rows(go):
    return go.tiles()[2] == 1


def test_apply_bitexact(pair):
    name, go, gp = pair
    x = H.random_fill(go.N, 1)
    _load(go, gp, 0, x)
    go.full_neg_lap(0, H.C0, 3)
    gp.full_neg_lap(0, H.C0, 3)
    leaf = _leaf_rows(go)
    yo = go.channel(3)[leaf]
    yp = gp.channel(3)[leaf]
    assert np.isfinite(yp).all()
    assert np.array_equal(H.bits(yo), H.bits(yp)), name


@pytest.mark.parametrize("color", [0, 1])
def test_gs_color_bitexact(pair, color):
    name, go, gp = pair
    lib = __import__("oracle").lib()
    for level in range(go.max_level + 1):
        x = H.random_fill(go.N, 10 + level)
        b = H.random_fill(go.N, 20 + level)
        _load(go, gp, 0, x)
        _load(go, gp, 1, b)
        lib.og_gs_color(go.h, level, 0, H.C0, 1, color, 1.5)
        gp.gs_color(level, 0, H.C0, 1, color, 1.5)
        assert np.array_equal(H.bits(go.channel(0)), H.bits(gp.channel(0))), (name, level)


def test_jacobi_bitexact(pair):
    name, go, gp = pair
    lib = __import__("oracle").lib()
    for level in range(go.max_level + 1):
        _load(go, gp, 0, H.random_fill(go.N, 30 + level))
        _load(go, gp, 1, H.random_fill(go.N, 40 + level))
        _load(go, gp, 5, H.random_fill(go.N, 50 + level))
        _load(go, gp, 6, H.random_fill(go.N, 60 + level))
        lib.og_jacobi(go.h, level, 5, 6, 0, H.C0, 1, np.float32(2.0 / 3.0))
        gp.jacobi(level, 5, 6, 0, H.C0, 1, np.float32(2.0 / 3.0))
        assert np.array_equal(H.bits(go.channel(6)), H.bits(gp.channel(6))), (name, level)


def test_restrict_save_prolong_bitexact(pair):
    name, go, gp = pair
    lib = __import__("oracle").lib()
    for level in range(1, go.max_level + 1):
        for ch, seed in ((0, 70), (1, 71), (3, 72)):
            _load(go, gp, ch, H.random_fill(go.N, seed + level))
        lib.og_residual_restrict(go.h, level, 0, H.C0, 1, 1.0)
        gp.residual_restrict(level, 0, H.C0, 1, 1.0)
        assert np.array_equal(H.bits(go.channel(1)), H.bits(gp.channel(1))), ("restrict b", name, level)
        assert np.array_equal(H.bits(go.channel(0)), H.bits(gp.channel(0))), ("restrict x", name, level)
        lib.og_fas_save(go.h, level - 1, 0, H.C0, 1, 3)
        gp.fas_save(level - 1, 0, H.C0, 1, 3)
        assert np.array_equal(H.bits(go.channel(1)), H.bits(gp.channel(1))), ("save b", name, level)
        assert np.array_equal(H.bits(go.channel(3)), H.bits(gp.channel(3))), ("save x0", name, level)
        lib.og_fas_prolong(go.h, level, 0, 3, 1.0)
        gp.fas_prolong(level, 0, 3, 1.0)
        assert np.array_equal(H.bits(go.channel(0)), H.bits(gp.channel(0))), ("prolong", name, level)


def test_propagate_accumulate_bitexact(pair):
    name, go, gp = pair
    lib = __import__("oracle").lib()
    _load(go, gp, 2, H.random_fill(go.N, 80))
    lib.og_propagate_ghosts(go.h, 2)
    gp.propagate_ghosts(2)
    lib.og_accumulate_leafs(go.h, 2)
    gp.accumulate_leafs(2)
    assert np.array_equal(H.bits(go.channel(2)), H.bits(gp.channel(2))), name


def test_dot_fp64(pair):
    name, go, gp = pair
    a, b = H.random_fill(go.N, 90), H.random_fill(go.N, 91)
    _load(go, gp, 2, a)
    _load(go, gp, 3, b)
    leaf = _leaf_rows(go)
    ref = float(np.sum(a[leaf].astype(np.float64) * b[leaf].astype(np.float64)))
    got = gp.dot(2, 3)
    assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)) * 100, (got, ref)


FORMS = [(mask, threads) for mask in (0, 7) for threads in (128, 256, 512)]


@pytest.fixture
def restore_forms():
    P = H.P
    saved = P.get_kernel_forms()
    yield P
    P.set_kernel_forms(*saved)


@pytest.mark.parametrize("forms", FORMS, ids=["mask%d_nt%d" % f for f in FORMS])
def test_kernel_forms_bitexact(pair, forms, restore_forms):
    """Every form of the stencil kernels (persistent or one tile per workgroup,
    128/256/512 threads per tile) gives the oracle's bits."""
    name, go, gp = pair
    P = restore_forms
    P.set_kernel_forms(*forms)
    assert P.get_kernel_forms() == forms
    lib = __import__("oracle").lib()
    _load(go, gp, 0, H.random_fill(go.N, 101))
    go.full_neg_lap(0, H.C0, 3)
    gp.full_neg_lap(0, H.C0, 3)
    leaf = _leaf_rows(go)
    assert np.array_equal(H.bits(go.channel(3)[leaf]), H.bits(gp.channel(3)[leaf])), ("apply", name, forms)
    for level in range(go.max_level + 1):
        _load(go, gp, 0, H.random_fill(go.N, 100 + level))   # ghost values (the apply above changed the oracle's)
        _load(go, gp, 1, H.random_fill(go.N, 110 + level))
        _load(go, gp, 5, H.random_fill(go.N, 120 + level))
        _load(go, gp, 6, H.random_fill(go.N, 130 + level))
        lib.og_jacobi(go.h, level, 5, 6, 0, H.C0, 1, np.float32(2.0 / 3.0))
        gp.jacobi(level, 5, 6, 0, H.C0, 1, np.float32(2.0 / 3.0))
        assert np.array_equal(H.bits(go.channel(6)), H.bits(gp.channel(6))), ("jacobi", name, level, forms)
    for level in range(1, go.max_level + 1):
        for ch, seed in ((0, 140), (1, 141)):
            _load(go, gp, ch, H.random_fill(go.N, seed + level))
        lib.og_residual_restrict(go.h, level, 0, H.C0, 1, 1.0)
        gp.residual_restrict(level, 0, H.C0, 1, 1.0)
        assert np.array_equal(H.bits(go.channel(1)), H.bits(gp.channel(1))), ("restrict b", name, level, forms)
        assert np.array_equal(H.bits(go.channel(0)), H.bits(gp.channel(0))), ("restrict x", name, level, forms)



@pytest.mark.parametrize("order,vw,types", [(-1, False, 2), (1, True, 2), (2, True, 2), (2, False, 7), (3, False, 2)])
def test_norm_matches_oracle(pair, order, vw, types):
    name, go, gp = pair
    _load(go, gp, 2, H.random_fill(go.N, 150))
    ref = go.norm(order, 2, vw, types)
    got = gp.norm(order, 2, vw, types)
    assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (name, order, vw, got, ref)
