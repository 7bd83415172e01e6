"""cmg_op_norm_noalloc (no allocation) against cmg_op_norm, through the C ABI.

Two grids in a row, the second smaller than the first, so the second call runs on
a device scratch array that still holds the first call's partials.  The two entry
points sum the same per-tile partials in the same order, so the results must be
equal bit for bit.
"""
import numpy as np
import pytest

import cirrus_amg as P
from helpers import build_pair, random_fill

CASES = [(-1, False, 2), (1, False, 2), (2, False, 2), (2, True, 2), (3, True, 2), (1, True, 7), (2, False, 1)]


def _compare(gp, channel, seed):
    gp.set_channel(channel, random_fill(gp.N, seed, 3.0))
    for order, vw, types in CASES:
        a = gp.norm_noalloc(order, channel, vw, types)      # before cmg_op_norm on this grid
        b = gp.norm(order, channel, vw, types)
        c = gp.norm_noalloc(order, channel, vw, types)
        assert np.float64(a).tobytes() == np.float64(b).tobytes(), (order, vw, types, a, b)
        assert np.float64(c).tobytes() == np.float64(b).tobytes(), (order, vw, types, c, b)
        if types & 2:
            assert b > 0      # INTERIOR cells exist on every grid here


@pytest.mark.gpu
def test_noalloc_norm_matches_norm_two_calls_shrinking():
    _, big = build_pair("c2_small")
    _, small = build_pair("uniform2")
    assert small.N < big.N
    _compare(big, 5, 11)
    _compare(small, 6, 12)
    # back to the large grid: its partials were overwritten by the small one's
    _compare(big, 5, 13)


@pytest.mark.gpu
def test_noalloc_norm_refuses_what_would_allocate():
    _, gp = build_pair("uniform2", with_device=False)
    fresh = P.HADeviceGrid.test_grid(P.make_target(P.TARGET_UNIFORM, 2, 2))
    with pytest.raises(P.CmgError) as e:
        fresh.norm_noalloc(2, 0)        # not on the device yet
    assert "not on the device" in str(e.value)
    gp.set_channel(5, random_fill(gp.N, 3))
    with pytest.raises(P.CmgError) as e:
        gp.norm_noalloc(2, 9)           # channel 9 never allocated
    assert "not allocated" in str(e.value)
    with pytest.raises(P.CmgError):
        gp.norm_noalloc(0, 5)           # order 0 is not a norm
