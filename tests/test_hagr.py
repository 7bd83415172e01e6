"""HAGR blobs (HADeviceGrid::dumpBinaryBlob / loadBinaryBlob, src/HAGrid.h:687-853):
the byte layout fixed by the reference's header and PoissonTile<float, 15>, round
trips, filtering and malformed input.  Host only."""
import struct

import numpy as np
import pytest

import cirrus_amg as P
import helpers as H

RECORD = 4 + 1 + 12 + 44880


@pytest.fixture(scope="module")
def grid():
    _, gp = H.build_pair("c2_small", with_device=False)
    return gp


def test_header_and_sizes(grid):
    blob = grid.dump_blob()
    magic, version, nl, ml, tile_size, coord_size, h0, pad, count = struct.unpack_from("<IIIiIIfIQ", blob, 0)
    assert magic == 0x52474148 and blob[:4] == b"HAGR" and version == 1
    assert (tile_size, coord_size) == (44880, 12)
    assert nl == len(P.DEFAULT_LOG2_HASH) and ml == grid.max_level and h0 == np.float32(0.125)
    assert list(struct.unpack_from("<%dI" % nl, blob, 40)) == list(P.DEFAULT_LOG2_HASH)
    assert count == grid.N and len(blob) == 40 + 4 * nl + count * RECORD


def test_records_follow_the_all_tiles_order(grid):
    blob = grid.dump_blob()
    lv, co, ty = grid.tiles()
    types = grid.cell_types().reshape(grid.N, 512)
    base = 40 + 4 * len(P.DEFAULT_LOG2_HASH)
    c0 = grid.host_channel(H.C0)
    for t in (0, grid.N // 2, grid.N - 1):
        off = base + t * RECORD
        level, typ = struct.unpack_from("<IB", blob, off)
        coord = struct.unpack_from("<3i", blob, off + 5)
        assert (level, typ, coord) == (lv[t], ty[t], tuple(co[t]))
        tile = np.frombuffer(blob, np.uint8, 44880, off + 17)
        data = tile[:15 * 736 * 4].view(np.float32).reshape(15, 736)
        assert np.array_equal(data[H.C0, :512], c0.reshape(grid.N, 512)[t])
        assert not data[:, 512:].any()
        assert np.array_equal(tile[15 * 736 * 4 + 192:15 * 736 * 4 + 192 + 512], types[t])


def test_round_trip(grid):
    blob = grid.dump_blob()
    g2 = P.HADeviceGrid.load_blob(blob)
    for a, b in zip(grid.tiles(), g2.tiles()):
        assert np.array_equal(a, b)
    assert np.array_equal(grid.neighbors(), g2.neighbors())
    assert np.array_equal(grid.cell_types(), g2.cell_types())
    for c in range(H.C0, H.C0 + 4):
        assert np.array_equal(grid.host_channel(c), g2.host_channel(c))
    assert g2.dump_blob() == blob


def test_filtered_dump(grid):
    lv, co, ty = grid.tiles()
    leaf = grid.dump_blob(tile_types=P.LEAF)
    count = struct.unpack_from("<Q", leaf, 32)[0]
    assert count == int((ty == P.LEAF).sum())
    lo = grid.dump_blob(max_level=1)
    assert struct.unpack_from("<i", lo, 12)[0] == 1
    assert struct.unpack_from("<Q", lo, 32)[0] == int((lv <= 1).sum())


def test_malformed_blobs(grid):
    blob = bytearray(grid.dump_blob())
    for bad in (b"XAGR" + blob[4:], blob[:4] + struct.pack("<I", 2) + blob[8:],
                blob[:16] + struct.pack("<I", 44881) + blob[20:], blob[:-1], blob[:20]):
        with pytest.raises(P.CmgError) as e:
            P.HADeviceGrid.load_blob(bytes(bad))
        assert e.value.status == 1
