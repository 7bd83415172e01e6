"""Pins of oracle functions that the reference's own tests cannot pin.

* hash-slot insertion order: an octree with forced collisions whose slots were
  worked out by hand from the reference's refine and ghost steps (hand_cases.py);
* the 4x4 supersampled face rule (the cut-cell rule of the 1024^3 torus workload):
  faces cut by nearly planar surfaces at known sub-face positions, and the solid
  area of whole planes of faces against closed-form cross-sections of a sphere
  and a torus;
* og_normal: moments, range and a Kolmogorov-Smirnov test against N(0, 1);
* og_net_forward: a single-tap net against a shift worked out by hand, and the
  three-layer net against torch's conv3d in FP64.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from hand_cases import HAND_BANDS, HAND_LOG2_HASH, HAND_SPHERE, check_hand_octree


def test_hand_octree_insertion_order():
    g = O.Grid(O.make_target(O.TARGET_SDF_BANDS, 1, 2, O.SDF_SPHERE, HAND_SPHERE, HAND_BANDS), HAND_LOG2_HASH)
    check_hand_octree(g)


# ---------------------------------------------------------------- supersampled faces
H64 = 1.0 / 64
GI = (10, 20, 30)      # a cell well inside the unit cube at spacing 1/64
BIG = 10.0             # sphere radius: its surface is planar to 3e-6 across a face


def _plane_solid_below(normal, point):
    """a sphere whose solid inside is (approximately) the half space n.(p - point) < 0"""
    n = np.asarray(normal, np.float64)
    n /= np.linalg.norm(n)
    c = np.asarray(point, np.float64) - BIG * n
    return (float(c[0]), float(c[1]), float(c[2]), BIG)


def _ss(prm, axis, gi=GI, h=H64):
    return O.face_supersampled(O.SDF_SPHERE, prm, gi, axis, h)


def test_supersampled_face_hand_cases():
    x, y, z = (v * H64 for v in GI)   # the z-face of cell GI spans [x, x+h] x [y, y+h] at height z
    # samples sit at sub-face centres 0.125, 0.375, 0.625, 0.875 (in cells) along each face axis
    # solid x < x + h/2: two sample columns of four are solid
    assert _ss(_plane_solid_below((1, 0, 0), (x + 0.5 * H64, y, z)), 2) == 0.5
    # solid x < x + h/4: one column (offset 0.125) solid
    assert _ss(_plane_solid_below((1, 0, 0), (x + 0.25 * H64, y, z)), 2) == 0.75
    # solid y < y + 3h/4: three rows solid
    assert _ss(_plane_solid_below((0, 1, 0), (x, y + 0.75 * H64, z)), 2) == 0.25
    # solid where (x' + y') < 1.625 h: only the sample (0.875, 0.875) is open -> 1/16, floored to 0.1
    f = _ss(_plane_solid_below((1, 1, 0), (x + 0.8125 * H64, y + 0.8125 * H64, z)), 2)
    assert f == np.float32(0.1)
    # solid where (x' + y') < 1.375 h: open samples (0.875, 0.875), (0.625, 0.875), (0.875, 0.625)
    # (sums 1.75, 1.5, 1.5; the next ones sum to 1.25) -> 3/16, above the floor
    assert _ss(_plane_solid_below((1, 1, 0), (x + 0.6875 * H64, y + 0.6875 * H64, z)), 2) == 3.0 / 16.0
    # the face lies at z = GI[2] h (the cell's lower face), not at the cell centre:
    # solid z < z + h/4 covers it, solid z < z - h/4 leaves it open
    assert _ss(_plane_solid_below((0, 0, 1), (x, y, z + 0.25 * H64)), 2) == 0.0
    assert _ss(_plane_solid_below((0, 0, 1), (x, y, z - 0.25 * H64)), 2) == 1.0
    # an x-face (axis 0) samples y and z: solid z < z + 3h/4 -> one of four rows open
    assert _ss(_plane_solid_below((0, 0, 1), (x, y, z + 0.75 * H64)), 0) == 0.25
    # open means SDF >= 0 (outside the body): the mirrored half space gives the complement
    assert _ss(_plane_solid_below((-1, 0, 0), (x + 0.25 * H64, y, z)), 2) == 0.25


def _plane_solid_area(kind, prm, axis, plane_index, h, lo, hi):
    """sum of (1 - open fraction) h^2 over the faces of one grid plane inside [lo, hi)^2"""
    n0, n1 = int(lo / h), int(math.ceil(hi / h))
    tot = 0.0
    for a in range(n0, n1):
        for b in range(n0, n1):
            gi = [0, 0, 0]
            gi[axis], gi[(axis + 1) % 3], gi[(axis + 2) % 3] = plane_index, a, b
            tot += (1.0 - O.face_supersampled(kind, prm, gi, axis, h)) * h * h
    return tot


# 4x4 midpoint sampling: a face's error is bounded by the sub-faces the boundary crosses
# and averages out along the curve (observed: 3e-5 to 1.5e-4 relative at h = 1/1024).
# A face plane half a cell off (the cell centre instead of the face) changes the
# off-centre torus section by 1 %, a sign flip by far more.
AREA_RTOL = 1e-3


@pytest.mark.parametrize("case", ["torus_mid", "torus_off", "sphere_off"])
def test_supersampled_planes_match_closed_form_sections(case):
    h = 1.0 / 1024
    R, r = 0.15, 0.05
    if case == "torus_mid":       # annulus between radii R - r and R + r
        got = _plane_solid_area(O.SDF_TORUS, (0.5, 0.5, 0.5, R, r), 2, 512, h, 0.28, 0.72)
        want = 4 * math.pi * R * r
    elif case == "torus_off":     # z = 0.5 + 32 h: annulus of half width sqrt(r^2 - dz^2)
        dz = 32 * h
        got = _plane_solid_area(O.SDF_TORUS, (0.5, 0.5, 0.5, R, r), 2, 512 + 32, h, 0.28, 0.72)
        want = 4 * math.pi * R * math.sqrt(r * r - dz * dz)
    else:                         # x = 0.5 + 64 h through a sphere of radius 0.2: a disc
        dx = 64 * h
        got = _plane_solid_area(O.SDF_SPHERE, (0.5, 0.5, 0.5, 0.2), 0, 512 + 64, h, 0.28, 0.72)
        want = math.pi * (0.2 ** 2 - dx * dx)
    assert abs(got / want - 1) < AREA_RTOL, (case, got, want)


# ---------------------------------------------------------------- og_normal
def test_normal_draws_are_standard_normal():
    from scipy import stats
    n = 1_000_000
    s = C.c_uint64(12345)
    f = O.lib().og_normal
    z = np.fromiter((f(C.byref(s)) for _ in range(n)), np.float64, n)
    se = 1 / math.sqrt(n)
    assert abs(z.mean()) < 5 * se
    assert abs(z.var() - 1) < 5 * math.sqrt(2) * se
    assert abs(stats.skew(z)) < 5 * math.sqrt(6) * se
    assert abs(stats.kurtosis(z)) < 5 * math.sqrt(24) * se
    # u1 in (0, 1] on 24 bits: |z| <= sqrt(-2 ln 2^-24)
    assert np.abs(z).max() <= math.sqrt(-2 * math.log(2.0 ** -24)) + 1e-6
    assert stats.kstest(z, "norm").pvalue > 1e-3


# ---------------------------------------------------------------- og_net_forward
def _vol(a):
    return np.asarray(a, np.float64).reshape(10, 10, 10)


def test_net_single_tap_is_a_shift():
    # W1[0] = tap (di, dj, dk) = (+1, 0, -1), W2[0][0] and W3[0] = centre tap: y = relu(x[i+1, j, k-1])
    w = np.zeros(O.NET_W)
    w[(1 + 1) * 9 + (0 + 1) * 3 + (-1 + 1)] = 1.0
    w[432 + 13] = 1.0
    w[7344 + 13] = 1.0
    x = np.random.default_rng(1).standard_normal(1000)
    y = _vol(O.net_forward(w, x))
    xv = _vol(x)
    want = np.zeros((10, 10, 10))
    want[:9, :, 1:] = np.maximum(xv[1:, :, :9], 0)
    assert np.array_equal(y, want)


def test_net_forward_matches_torch_conv3d():
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(7)
    for trial in range(3):
        w = rng.standard_normal(O.NET_W) * 0.3
        x = rng.standard_normal(1000)
        y = _vol(O.net_forward(w, x))
        t = torch.from_numpy(np.ascontiguousarray(x)).double().view(1, 1, 10, 10, 10)
        w1 = torch.from_numpy(w[:432].copy()).view(16, 1, 3, 3, 3)
        w2 = torch.from_numpy(w[432:7344].copy()).view(16, 16, 3, 3, 3)
        w3 = torch.from_numpy(w[7344:].copy()).view(1, 16, 3, 3, 3)
        a1 = F.relu(F.conv3d(t, w1, padding=1))
        a2 = F.relu(F.conv3d(a1, w2, padding=1))
        yt = F.conv3d(a2, w3, padding=1).view(10, 10, 10).numpy()
        assert np.allclose(y, yt, rtol=1e-12, atol=1e-12 * np.abs(yt).max()), trial


# ---------------------------------------------------------------- og_roi_flags
def test_roi_flags_are_per_tile_min_max_predicates():
    """og_roi_flags against numpy's per-tile min / max and the tile table (subtree rule included)"""
    import helpers as H
    go, _ = H.build_pair("sphere_empty24", with_device=False)
    n = go.N
    x = H.random_fill(n, 21)
    x[np.arange(n) % 3 == 0] = -np.abs(x[np.arange(n) % 3 == 0])      # some tiles entirely negative
    go.channel(5)[:] = x
    lv, _, ty = go.tiles()
    mn, mx = x.min(axis=1), x.max(axis=1)
    sub = int(lv.max()) - 1
    stype = np.where((lv == sub) & (ty == 2), 1, ty)                   # NONLEAF at the subtree level acts as LEAF
    for pred, a, b, want in [(0, 0.0, 0.0, mx > 0), (1, -0.9, 0.0, mn < -0.9), (2, 0.95, 2.0, (mn <= 2.0) & (mx >= 0.95)),
                             (3, 1.9, 0.0, (mx - mn) > 1.9)]:
        for level, types in [(-1, 1), (sub, 1), (-1, 6)]:
            sel = ((np.where(lv == level, stype, ty) if level >= 0 else ty) & types) != 0
            fi, tmn, tmx = go.roi_flags(5, pred, a, b, level, types)
            assert np.array_equal(fi.astype(bool), want & sel), (pred, level, types)
            assert np.array_equal(tmn[sel], mn[sel]) and np.array_equal(tmx[sel], mx[sel])
            assert 0 < sel.sum() < n
