"""Octree construction on the device against the host construction and the
oracle, slot for slot: hash slots, all-tiles order, tile types, level ranges,
neighbour / parent / child tables, for every level-target kind (device-evaluated
and host-evaluated), for the benchmark octrees, for mixed host/device step
sequences, and for a hash table that overflows."""
import time

import numpy as np
import pytest

import cirrus_amg as P
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

TORUS_BANDS_C3 = {7: 0.01, 6: 0.02, 5: 0.04, 4: 0.08}


def _topology(g):
    return [*g.tiles(), g.slots(), g.neighbors(), g.parents(), g.children(), g.level_starts(),
            np.array([g.max_level])]


def _assert_same(a, b, what):
    ta, tb = _topology(a), _topology(b)
    for k, (x, y) in enumerate(zip(ta, tb)):
        assert x.shape == y.shape and np.array_equal(x, y), (what, k)


def _build(target, device, log2_hash=P.DEFAULT_LOG2_HASH):
    g = P.HADeviceGrid(0.125, log2_hash)
    g.setTileHost(0, (0, 0, 0), P.LEAF)
    g.rebuild(device)
    g.iterativeRefine(target, device)
    g.finalize(device)
    return g


def _case_target(name):
    kind, lo, hi, sk, sdf, bands, _ = H.CASES[name]
    return P.make_target(kind, lo, hi, sk, sdf, bands)


@pytest.mark.parametrize("case", sorted(H.CASES))
def test_device_build_matches_host(case):
    print("\ncase", case, flush=True)
    _assert_same(_build(_case_target(case), False), _build(_case_target(case), True), case)


@pytest.mark.parametrize("case", ["sphere_empty24", "star_empty24", "c3_small"])
def test_device_build_matches_oracle(case):
    print("\ncase", case, flush=True)
    kind, lo, hi, sk, sdf, bands, _ = H.CASES[case]
    go = O.Grid(O.make_target(kind, lo, hi, sk, sdf, bands))
    gd = _build(_case_target(case), True)
    for a, b in zip(go.tiles(), gd.tiles()):
        assert np.array_equal(a, b)
    assert np.array_equal(go.slots(), gd.slots())
    assert np.array_equal(go.neighbors(), gd.neighbors())
    assert np.array_equal(go.parents(), gd.parents())
    assert np.array_equal(go.children(), gd.children())
    assert np.array_equal(go.level_range(), gd.level_starts())


@pytest.mark.parametrize("kind", [P.TARGET_STAR_SOLID, P.TARGET_SPHERE_SOLID, P.TARGET_CENTER_POINT])
def test_device_build_other_presets(kind):
    print("\nkind", kind, flush=True)
    t = lambda: P.make_target(kind, 1, 5)
    _assert_same(_build(t(), False), _build(t(), True), kind)


def test_device_build_callback_target():
    """A host callback target (evaluated on the host, one level at a time) builds the same octree."""
    seen = []

    def target(level, coord):
        seen.append(level)
        return 4 if (coord[0] + 2 * coord[1] + 3 * coord[2]) % 5 == 0 else 2

    ref = _build(P.make_target(P.TARGET_CALLBACK, 2, 4, fn=target), False)
    n_host = len(seen)
    seen.clear()
    got = _build(P.make_target(P.TARGET_CALLBACK, 2, 4, fn=target), True)
    _assert_same(ref, got, "callback")
    assert len(seen) == n_host   # the callback sees the same LEAF tiles


@pytest.mark.parametrize("name,target", [
    ("c2_256_sphere_bands", lambda: P.make_target(P.TARGET_SDF_BANDS, 2, 5, P.SDF_SPHERE, H.SPHERE_C2,
                                                  {5: 0.05, 4: 0.15})),
    ("c3_1024_torus_bands", lambda: P.make_target(P.TARGET_SDF_BANDS, 3, 7, P.SDF_TORUS, H.TORUS_C3, TORUS_BANDS_C3)),
])
def test_device_build_benchmark_octrees(name, target):
    print("\ncase", name, flush=True)

    def one_call():
        g = P.HADeviceGrid()
        g.setTileHost(0, (0, 0, 0), P.LEAF)
        g.build_device(target())
        return g

    def best_of(fn, reps=3):
        best, out = None, None
        for _ in range(reps):
            if out is not None:
                out.close()
            t0 = time.perf_counter()
            out = fn()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return out, best

    _build(target(), True).close()   # first call pays module load and allocator warm-up
    gh, th = best_of(lambda: _build(target(), False))
    gd, td = best_of(lambda: _build(target(), True))
    g1, t1 = best_of(one_call)
    print(f"{name}: {gh.N} tiles, best of 3: host construction {th * 1e3:.1f} ms, device construction "
          f"{td * 1e3:.1f} ms (three calls), {t1 * 1e3:.1f} ms (cmg_grid_build_device)")
    _assert_same(gh, gd, name)
    _assert_same(gh, g1, name + " one call")


def test_mixed_host_and_device_steps():
    """Device steps leave the host tables exact: host and device steps interleave freely."""
    t = lambda: _case_target("c2_small")
    ref = _build(t(), False)
    for refine_dev, finalize_dev in ((True, False), (False, True)):
        g = P.HADeviceGrid()
        g.setTileHost(0, (0, 0, 0), P.LEAF)
        g.rebuild(not refine_dev)
        g.iterativeRefine(t(), refine_dev)
        g.finalize(finalize_dev)
        _assert_same(ref, g, (refine_dev, finalize_dev))
    # host edits after a device refinement: an extra LEAF tile next to the octree, rebuilt on either side
    out = []
    for dev in (False, True):
        g = P.HADeviceGrid()
        g.setTileHost(0, (0, 0, 0), P.LEAF)
        g.rebuild(dev)
        g.iterativeRefine(t(), dev)
        g.setTileHost(1, (3, 3, 3), P.LEAF)
        g.rebuild(dev)
        g.finalize(dev)
        out.append(g)
    _assert_same(out[0], out[1], "edit+rebuild")


@pytest.mark.parametrize("log2_hash", [(1, 4, 7, 9, 12, 13, 16, 16, 16, 16), (1, 4, 8, 9, 12, 13, 16, 16, 16, 16)])
def test_device_build_dense_tables(log2_hash):
    """Tables near full (long probe clusters, many displacements in each insertion round):
    the device slots still equal the sequential insertion's."""
    t = lambda: P.make_target(P.TARGET_SDF_BANDS, 2, 5, P.SDF_SPHERE, H.SPHERE_C2, {5: 0.05, 4: 0.15})
    gh = _build(t(), False, log2_hash)
    lv, _, _ = gh.tiles()
    load = [int((lv == l).sum()) / 2 ** log2_hash[l] for l in range(6)]
    print("\nload factors", [round(x, 2) for x in load], flush=True)
    assert max(load) > 0.9
    _assert_same(gh, _build(t(), True, log2_hash), log2_hash)


def test_device_build_hash_full():
    """A level table too small for the refinement fails with CMG_ERR_HASH_FULL on both sides."""
    small = (4, 4, 5, 16, 16, 16, 16, 16, 16, 16)   # level 2 needs 64 slots
    for dev in (False, True):
        with pytest.raises(P.CmgError) as e:
            _build(P.make_target(P.TARGET_UNIFORM, 2, 2), dev, small)
        assert e.value.status == 7, dev


def test_device_built_grid_solves():
    """Assembly and a solve on a device-built octree reproduce the host-built grid's solve bit for bit."""
    kind, lo, hi, sk, sdf, bands, _ = H.CASES["c3_small"]
    res = []
    for dev in (False, True):
        g = P.HADeviceGrid.test_grid(P.make_target(kind, lo, hi, sk, sdf, bands), device=dev)
        g.set_wall_types(lo, True)
        g.assemble_cutcell(sk, sdf, 1, H.C0, H.R_MATRIX)
        g.set_channel(1, H.random_fill(g.N, 3))
        s = P.AMGSolver(g, H.C0, H.R_MATRIX, 1.0, 1.0)
        o = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1)
        it, rel = s.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=o)
        res.append((it, rel, g.channel(0).copy()))
        s.close()
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1]
    assert np.array_equal(H.bits(res[0][2]), H.bits(res[1][2]))


@pytest.mark.parametrize("one_call", [False, True])
def test_device_build_hand_octree(one_call):
    """device construction gives the slots worked out by hand from the reference's steps (hand_cases.py)"""
    from hand_cases import HAND_BANDS, HAND_LOG2_HASH, HAND_SPHERE, check_hand_octree
    print("\nhand octree, one call" if one_call else "\nhand octree, three calls", flush=True)
    t = P.make_target(P.TARGET_SDF_BANDS, 1, 2, P.SDF_SPHERE, HAND_SPHERE, HAND_BANDS)
    if one_call:
        g = P.HADeviceGrid.test_grid(t, log2_hash=HAND_LOG2_HASH, device=True)
    else:
        g = _build(t, True, HAND_LOG2_HASH)
    check_hand_octree(g)
