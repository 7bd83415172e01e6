"""The C ABI library loads, exports every declared symbol, and reports errors by status."""
import ctypes as C
import os
import re
import subprocess

import pytest

import cirrus_amg as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cirrus_amg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cmg_[a-z0-9_]+)\s*\(", text)) - {"cmg_level_target_fn"})


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(P.LIB_PATH)
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(P.EXPORTS)


def test_host_errors_are_status_codes():
    lib = P.lib()
    h = C.c_void_p()
    arr = (C.c_uint32 * 3)(8, 8, 8)
    assert lib.cmg_grid_create(-1.0, 3, arr, C.byref(h)) == 1          # BAD_ARG
    assert lib.cmg_grid_create(0.125, 3, arr, C.byref(h)) == 0
    assert lib.cmg_grid_set_tile(h, 5, 0, 0, 0, 1) == 1                # level out of range
    assert lib.cmg_grid_set_tile(h, 0, 0, 0, 0, 3) == 1                # not a tile type
    assert lib.cmg_grid_get_slots(h, None) == 2                        # not finalized
    assert lib.cmg_grid_set_tile(h, 0, 0, 0, 0, 1) == 0
    assert lib.cmg_grid_rebuild(h) == 0
    assert lib.cmg_grid_finalize(h) == 0
    assert lib.cmg_grid_num_tiles(h) == 1
    assert lib.cmg_prepare_types_and_coeffs(h, 13, 0.5) == 1           # c0+3 out of range
    lib.cmg_grid_destroy(h)


def test_hash_full_is_reported():
    g = P.HADeviceGrid(0.125, (4, 2, 2))     # 4 slots at levels 1 and 2
    g.setTileHost(0, (0, 0, 0), P.LEAF)
    g.rebuild()
    with pytest.raises(P.CmgError) as e:
        g.iterativeRefine(P.make_target(P.TARGET_UNIFORM, 2, 2))
    assert e.value.status == 7


def test_kernel_forms_validation():
    saved = P.get_kernel_forms()
    try:
        for bad in ((128, 256), (-1, 256), (7, 64), (0, 1024)):
            with pytest.raises(P.CmgError) as e:
                P.set_kernel_forms(*bad)
            assert e.value.status == P.lib().cmg_set_kernel_forms(*bad)
        P.set_kernel_forms(5, 128)
        assert P.get_kernel_forms() == (5, 128)
    finally:
        P.set_kernel_forms(*saved)


def _wrapper_binary():
    here = os.path.dirname(os.path.abspath(__file__))
    subprocess.run(["make", "-s", "-C", os.path.join(here, "cpp")], check=True)
    return os.path.join(here, "cpp", "_build", "test_amg_wrapper")


def test_cpp_wrapper_cpu_part():
    """A C++ caller written against the reference's class API (HADeviceGrid,
    AMGSolver) builds and runs its host part through include/cirrus_amg.hpp."""
    r = subprocess.run([_wrapper_binary(), "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cpu part OK" in r.stdout
