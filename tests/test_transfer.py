"""Carrying a cell field from one octree to another (the warm start of the
moving-torus solve sequence): host-only checks on two sphere octrees."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cirrus-amg_amd"))

import cirrus_amg as P  # noqa: E402
import c5_sequence as S  # noqa: E402


def _grid(cx):
    t = P.make_target(P.TARGET_SDF_BANDS, 3, 5, P.SDF_SPHERE, (cx, 0.5, 0.5, 0.2), {5: 0.05, 4: 0.15})
    g = P.HADeviceGrid.test_grid(t)
    g.assemble_cutcell(P.SDF_SPHERE, (cx, 0.5, 0.5, 0.2), -1, 11, 0.5)
    return g


def _centres(g):
    lv, co, _ = g.tiles()
    h = 0.125 / 2.0 ** lv
    c = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(512, 3)
    return (co[:, None, :] * 8 + c[None] + 0.5) * h[:, None, None]


def _mask(g):
    _, _, ty = g.tiles()
    return (g.cell_types().reshape(g.N, 512) == 2) & ((ty & 1) > 0)[:, None]


@pytest.fixture(scope="module")
def grids():
    return _grid(0.5), _grid(0.55)


def _linear(p):
    return 1.0 + 2.0 * p[..., 0] - 3.0 * p[..., 1] + 0.5 * p[..., 2]


def test_same_octree_is_identity(grids):
    a, _ = grids
    x = np.random.default_rng(0).normal(size=(a.N, 512)).astype(np.float32)
    out = S.transfer_leaf_values(a, x, a)
    m = _mask(a)
    assert np.array_equal(out[m], x[m])
    assert not out[~m].any()


def test_linear_field(grids):
    """Means of children and copies keep a linear field exact (cell centres of a
    parent are the means of its children's centres); injection from a coarser
    tile is off by at most |grad| h_coarse / 2 per axis."""
    a, b = grids
    xa = _linear(_centres(a)).astype(np.float32)
    out = S.transfer_leaf_values(a, xa, b)
    m = _mask(b)
    ref = _linear(_centres(b))
    lv_b, co_b, _ = b.tiles()
    keys = {(int(l), *map(int, c)) for l, c, t in zip(*a.tiles()) if t & 3}
    same = np.array([(int(l), *map(int, c)) in keys for l, c in zip(lv_b, co_b)])
    err = np.abs(out - ref)
    assert err[m & same[:, None]].max() <= 1e-5
    hmax = 0.125 / 2.0 ** 3
    assert err[m].max() <= (2.0 + 3.0 + 0.5) * hmax / 2 + 1e-5
