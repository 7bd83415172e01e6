"""Library host side (C++17 topology, assembly, inputs) against the oracle, bit for bit.

Runs without a GPU: the library uploads to the device lazily, so construction
and assembly are compared on the host copies.
"""
import numpy as np
import pytest

import cirrus_amg as P
import helpers as H


@pytest.fixture(scope="module", params=sorted(H.CASES))
def pair(request):
    return request.param, H.build_pair(request.param)


def test_topology_bitexact(pair):
    name, (go, gp) = pair
    for a, b in zip(go.tiles(), gp.tiles()):
        assert np.array_equal(a, b), name
    assert np.array_equal(go.slots(), gp.slots())
    assert np.array_equal(go.neighbors(), gp.neighbors())
    assert np.array_equal(go.parents(), gp.parents())
    assert np.array_equal(go.children(), gp.children())
    assert np.array_equal(go.level_range(), gp.level_starts())


def test_assembly_bitexact(pair):
    name, (go, gp) = pair
    assert np.array_equal(go.cell_types(), gp.cell_types()), name
    for a in range(4):
        assert np.array_equal(H.bits(go.channel(H.C0 + a)), H.bits(gp.host_channel(H.C0 + a))), (name, a)


def test_inputs_bitexact(pair):
    name, (go, gp) = pair
    go.rhs_uniform(1, 11, True)
    gp.rhs_uniform(1, 11, True)
    assert np.array_equal(H.bits(go.channel(1)), H.bits(gp.host_channel(1)))
    go.rhs_divergence(1, 5, 12)
    gp.rhs_divergence(1, 5, 12)
    assert np.array_equal(H.bits(go.channel(1)), H.bits(gp.host_channel(1)))
    for a in range(3):
        assert np.array_equal(H.bits(go.channel(5 + a)), H.bits(gp.host_channel(5 + a)))


def test_callback_target_matches_preset():
    """A level target given as a host callback builds the same octree as the preset."""
    kind, lo, hi = P.TARGET_SPHERE_EMPTY, 2, 4
    ref = P.HADeviceGrid.test_grid(P.make_target(kind, lo, hi))

    def target(level, coord):
        h = 0.125 / (1 << level)
        mn = np.array(coord, np.float32) * np.float32(8.0) * np.float32(h)
        mx = np.array([c + 1 for c in coord], np.float32) * np.float32(8.0) * np.float32(h)
        inside = 0
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    p = np.array([mx[0] if di else mn[0], mx[1] if dj else mn[1], mx[2] if dk else mn[2]], np.float32)
                    d = p - np.float32(0.5)
                    if np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])) - np.float32(0.25) < 0:
                        inside += 1
        return lo if inside in (0, 8) else hi

    g = P.HADeviceGrid()
    g.setTileHost(0, (0, 0, 0), P.LEAF)
    g.rebuild()
    g.iterativeRefine(P.make_target(P.TARGET_CALLBACK, lo, hi, fn=target))
    g.finalize()
    for a, b in zip(ref.tiles(), g.tiles()):
        assert np.array_equal(a, b)


def test_hand_octree_insertion_order_host():
    """the library's host construction gives the hand-derived slots (hand_cases.py)"""
    from hand_cases import HAND_BANDS, HAND_LOG2_HASH, HAND_SPHERE, check_hand_octree
    t = P.make_target(P.TARGET_SDF_BANDS, 1, 2, P.SDF_SPHERE, HAND_SPHERE, HAND_BANDS)
    check_hand_octree(P.HADeviceGrid.test_grid(t, log2_hash=HAND_LOG2_HASH))
