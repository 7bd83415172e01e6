"""Device-resident exchange and region-of-interest flags on the GPU.

* cmg_mark_region_of_interest against the oracle (flags bit for bit, tile min / max
  equal) for every predicate, a subtree level and several tile-type selections;
* reference tile images (PoissonTile<float,15>, 44,880 bytes, in device memory
  allocated here through the HIP runtime the library uses): gather / scatter of channels and cell types through a
  device array of tile addresses and through a slab base + stride, leaving every
  other byte of the images alone;
* a FluidEuler::project-shaped pass that never touches host memory (gather the
  velocities and types from the tile images, divergence, solve, gradient, scatter)
  against the same pass through host channel copies, bit for bit;
* a HAGR blob loaded back and solved on the device like the grid it came from.
"""
import numpy as np
import pytest

import cirrus_amg as P
import helpers as H
from devmem import DeviceBuffer

pytestmark = pytest.mark.gpu

TILE = 44880
CH_FLOATS = 736
TYPE_OFF = 15 * 736 * 4 + 6 * 32


@pytest.mark.parametrize("case", ["sphere_empty24", "c3_small"])
def test_roi_flags_match_oracle(case):
    print("\nroi", case, flush=True)
    go, gp = H.build_pair(case)
    x = H.random_fill(go.N, 31)
    x[np.arange(go.N) % 3 == 0] = -np.abs(x[np.arange(go.N) % 3 == 0])   # some tiles entirely negative
    go.channel(5)[:] = x
    gp.set_channel(5, x)
    sub = go.max_level - 1
    for pred, a, b in [(P.ROI_MAX_GT, 0.0, 0.0), (P.ROI_MIN_LT, -0.9, 0.0), (P.ROI_RANGE_MEETS, 0.95, 2.0),
                       (P.ROI_SPAN_GT, 1.9, 0.0)]:
        for level, types in [(-1, P.LEAF), (sub, P.LEAF), (-1, P.NONLEAF | P.GHOST)]:
            fo, mno, mxo = go.roi_flags(5, pred, a, b, level, types)
            gp.mark_region_of_interest(5, pred, a, b, level, types, calc_locked=True)
            fi, fl, mn, mx = gp.roi()
            assert np.array_equal(fi, fo), (pred, level, types)
            assert np.array_equal(fl, fo)
            sel = fo.astype(bool) | (mno != 0) | (mxo != 0)
            assert np.array_equal(mn[sel], mno[sel]) and np.array_equal(mx[sel], mxo[sel])
    # calc_locked = 0 leaves the locked flags of the previous marking
    gp.mark_region_of_interest(5, P.ROI_MAX_GT, 10.0, 0.0, -1, P.LEAF, calc_locked=False)
    fi2, fl2, _, _ = gp.roi()
    assert not fi2.any() and np.array_equal(fl2, fl)


def _images(n, seed):
    """n tile images with random bytes in device memory, and a device array of their addresses;
    returns (images buffer, address buffer, host copy of the images)"""
    host = np.random.default_rng(seed).integers(0, 256, (n, TILE), dtype=np.uint8)
    img = DeviceBuffer(n * TILE)
    img.upload(host)
    addrs = DeviceBuffer(n * 8)
    addrs.upload(np.arange(n, dtype=np.uint64) * np.uint64(TILE) + np.uint64(img.addr))
    return img, addrs, host


def _floats(host):
    return host.view(np.float32).reshape(host.shape[0], TILE // 4)


def test_tile_images_gather_and_scatter():
    _, gp = H.build_pair("c2_small")
    n = gp.N
    img, addrs, before = _images(n, 41)
    # gather tile channel 7 (array of addresses) and the types (slab form)
    gp.gather_tile_channel(7, 9, d_addrs=addrs.addr)
    assert np.array_equal(H.bits(gp.channel(9)), H.bits(_floats(before)[:, 7 * CH_FLOATS:7 * CH_FLOATS + 512]))
    types_before = gp.cell_types().copy()
    gp.gather_tile_types(d_base=img.addr, stride=TILE)
    assert np.array_equal(gp.cell_types().reshape(n, 512), before[:, TYPE_OFF:TYPE_OFF + 512])
    gp.set_cell_types(types_before)
    # scatter a known channel into tile channel 3 (slab form) and the types (address form)
    vals = H.random_fill(n, 42)
    gp.set_channel(10, vals)
    gp.scatter_tile_channel(10, 3, d_base=img.addr, stride=TILE)
    gp.scatter_tile_types(d_addrs=addrs.addr)
    after = img.download().reshape(n, TILE)
    want = before.copy()
    _floats(want)[:, 3 * CH_FLOATS:3 * CH_FLOATS + 512] = vals
    want[:, TYPE_OFF:TYPE_OFF + 512] = types_before.reshape(n, 512)
    assert np.array_equal(after, want)          # nothing else in the images moved
    # device channel copies
    buf = DeviceBuffer(n * 512 * 4)
    gp.get_channel_device(10, buf.addr)
    assert np.array_equal(H.bits(buf.download(np.float32).reshape(n, 512)), H.bits(vals))
    buf.upload(vals * np.float32(2.0))
    gp.set_channel_device(11, buf.addr)
    assert np.array_equal(H.bits(gp.channel(11)), H.bits(vals * np.float32(2.0)))


def test_tile_image_errors():
    _, gp = H.build_pair("uniform2")
    n = gp.N
    img, addrs, _ = _images(n, 43)
    with pytest.raises(P.CmgError) as e:     # both forms at once
        gp.gather_tile_channel(0, 5, d_addrs=addrs.addr, d_base=img.addr, stride=TILE)
    assert e.value.status == 1
    with pytest.raises(P.CmgError):          # stride shorter than a tile
        gp.gather_tile_channel(0, 5, d_base=img.addr, stride=TILE - 16)
    a = np.arange(n, dtype=np.uint64) * np.uint64(TILE) + np.uint64(img.addr)
    a[3] = 0
    a[5] += np.uint64(4)                     # not 16-byte aligned
    bad = DeviceBuffer(n * 8)
    bad.upload(a)
    with pytest.raises(P.CmgError) as e:
        gp.gather_tile_channel(0, 5, d_addrs=bad.addr)
    assert e.value.status == 1 and "2 tile address" in str(e.value)


def _project(gp, fp16=1):
    o = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                       cycle_coef_fp16=fp16, no_true_residual=1)
    gp.divergence(5, 1, H.C0)
    s = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    iters, rel = s.solve(opts=o, max_iters=100, relative_tolerance=1e-6)
    assert rel < 1e-6
    gp.add_gradient(0, H.C0, 5, 1)
    return iters


def test_device_resident_projection_matches_host_path():
    """FluidEuler::project (cirrus_cutcell/FluidEuler.cu:688-752) with velocities living in
    reference tile images: no host copy of any per-cell array"""
    _, gd = H.build_pair("c2_small")
    _, gh = H.build_pair("c2_small")
    n = gd.N
    u = [H.random_fill(n, 50 + a) for a in range(3)]
    host = np.random.default_rng(44).integers(0, 256, (n, TILE), dtype=np.uint8)
    for a in range(3):
        _floats(host)[:, (6 + a) * CH_FLOATS:(6 + a) * CH_FLOATS + 512] = u[a]
    host[:, TYPE_OFF:TYPE_OFF + 512] = gd.cell_types().reshape(n, 512)
    img = DeviceBuffer(n * TILE)
    img.upload(host)
    addrs = DeviceBuffer(n * 8)
    addrs.upload(np.arange(n, dtype=np.uint64) * np.uint64(TILE) + np.uint64(img.addr))
    # device path: tiles -> solve -> tiles
    gd.gather_tile_types(d_addrs=addrs.addr)
    for a in range(3):
        gd.gather_tile_channel(6 + a, 5 + a, d_addrs=addrs.addr)
    it_d = _project(gd)
    for a in range(3):
        gd.scatter_tile_channel(5 + a, 6 + a, d_addrs=addrs.addr)
    # host path on an identical grid
    for a in range(3):
        gh.set_channel(5 + a, u[a])
    it_h = _project(gh)
    assert it_d == it_h
    out = _floats(img.download().reshape(n, TILE))
    for a in range(3):
        assert np.array_equal(H.bits(out[:, (6 + a) * CH_FLOATS:(6 + a) * CH_FLOATS + 512]), H.bits(gh.channel(5 + a)))
    assert np.array_equal(H.bits(gd.channel(0)), H.bits(gh.channel(0)))


def test_hagr_blob_solves_like_its_grid():
    """a grid dumped as HAGR (src/HAGrid.h:687-758), loaded back (:760-853) and solved on the device"""
    go, gp = H.build_pair("c2_small")
    gp.rhs_divergence(1, 5, 7)
    go.rhs_divergence(1, 5, 7)
    blob = gp.dump_blob()
    g2 = P.HADeviceGrid.load_blob(blob)
    o = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1)
    res = []
    for g in (gp, g2):
        s = P.AMGSolver(g, H.C0, H.R_MATRIX, 1.0, 1.0)
        it, rel = s.solve(opts=o, max_iters=100, relative_tolerance=1e-6)
        assert rel < 1e-6
        res.append((it, g.channel(0).copy()))
    assert res[0][0] == res[1][0]
    assert np.array_equal(H.bits(res[0][1]), H.bits(res[1][1]))
    A = go.assemble_csr(H.C0)
    b = go.channel(1).ravel()[A.cell].astype(np.float64)
    x64, _, _ = A.bicgstab(b, 1e-13, 20000)
    x = res[1][1].ravel()[A.cell].astype(np.float64)
    assert np.linalg.norm(x - x64) <= 1e-5 * np.linalg.norm(x64)
