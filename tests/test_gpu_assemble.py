"""The cut-cell operator assembled on the device against the host assembly, bit
for bit: cell types (walls, solid cells, ghosts, NONLEAF types, closed cells) and
the four coefficient channels on every tile."""
import time

import numpy as np
import pytest

import cirrus_amg as P
import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["uniform3", "sphere_empty24", "star_empty24", "center_point14", "c2_small",
                                  "sphere_solid24", "c3_small"])
def test_device_assembly_matches_host(case):
    print("\ncase", case, flush=True)
    kind, lo, hi, sk, sdf, bands, asm = H.CASES[case]
    _, gh = H.build_pair(case)
    gd = P.HADeviceGrid.test_grid(P.make_target(kind, lo, hi, sk, sdf, bands))
    walls = lo if asm in ("corner", "supersampled") else -1
    rule = {"plain": -1, "solid": -1, "corner": 0, "supersampled": 1}[asm]
    gd.assemble_cutcell_device(sk, sdf if asm != "plain" else None, rule, H.C0, H.R_MATRIX, walls, True)
    assert np.array_equal(gh.cell_types(), gd.cell_types())
    for a in range(4):
        assert np.array_equal(H.bits(gh.channel(H.C0 + a)), H.bits(gd.channel(H.C0 + a))), a


def test_device_assembly_then_solve():
    """A solve on the device-assembled operator matches the host-assembled one."""
    kind, lo, hi, sk, sdf, bands, asm = H.CASES["c3_small"]
    _, gh = H.build_pair("c3_small")
    gd = P.HADeviceGrid.test_grid(P.make_target(kind, lo, hi, sk, sdf, bands))
    t0 = time.perf_counter()
    gd.assemble_cutcell_device(sk, sdf, 1, H.C0, H.R_MATRIX, lo, True)
    print("device assembly s", time.perf_counter() - t0)
    b = H.random_fill(gh.N, 5)
    out = []
    for g in (gh, gd):
        g.set_channel(1, b)
        s = P.AMGSolver(g, H.C0, H.R_MATRIX, 1.0, 1.0)
        o = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1)
        it, rel = s.solve(max_iters=100, relative_tolerance=1e-6, coarsest_iters=0, opts=o)
        out.append((it, rel, g.channel(0).copy()))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(H.bits(out[0][2]), H.bits(out[1][2]))
