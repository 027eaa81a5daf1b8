"""Binary navmesh fixture ("GMAPFNAV" v1) through the C ABI, no GPU needed:
round trips, the byte layout read back by an independent numpy parser,
the committed bow-tie fixture, and the decoder's rejection of malformed
input with the FillMapManually error contract (src/GMAPF.cu:199-217)."""
import os
import struct

import numpy as np
import pytest

import gmapf_amd as G
from gmapf_amd import scenarios as S
from tests import helpers as H

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bowtie.nav")


def fresh():
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=4))
    return g


def built(obstacles, w, h):
    g = fresh()
    g.fill_map_start(w, h)
    for o in obstacles:
        g.fill_map_obstacle(o)
    g.fill_map_end()
    return g


def fnv1a64(b: bytes) -> int:
    h = 1469598103934665603
    for x in b:
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def parse(b: bytes):
    """The layout, read field by field (independent of the C++ decoder)."""
    assert b[:8] == b"GMAPFNAV"
    version, hdr, V, E = struct.unpack_from("<4I", b, 8)
    bbox = np.frombuffer(b, np.float32, 4, 24)
    assert struct.unpack_from("<6I", b, 40) == (0,) * 6
    o = hdr
    tri = np.frombuffer(b, np.float32, 6 * V, o).reshape(V, 6)
    o += 24 * V
    cen = np.frombuffer(b, np.float32, 2 * V, o).reshape(V, 2)
    o += 8 * V
    row = np.frombuffer(b, np.uint32, V + 1, o)
    o += 4 * (V + 1)
    col = np.frombuffer(b, np.uint32, E, o)
    o += 4 * E
    (ck,) = struct.unpack_from("<Q", b, o)
    assert o + 8 == len(b)
    return dict(version=version, hdr=hdr, V=V, E=E, bbox=bbox, tri=tri, cen=cen, row=row, col=col, ck=ck, end=o)


def maze_map():
    return built(S.maze_obstacles(3, cells=8, cell=64.0, corridor=16.0), 512, 512)


@pytest.mark.parametrize("which", ["bowtie", "maze", "manual"])
def test_round_trip_gives_identical_map(which):
    if which == "bowtie":
        g = built(S.bowtie_obstacles(), 100, 100)
    elif which == "maze":
        g = maze_map()
    else:
        tri, nbr = H.grid_map(5, 12, 9, cell=3.0, p=0.3)
        g = fresh()
        g.fill_map_manually(tri, nbr)
    data = g.nav_bytes()
    h = fresh()
    h.fill_map_from_nav(data)
    for a, b in zip(g.map_arrays(), h.map_arrays()):
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                                     b.view(np.uint32) if b.dtype == np.float32 else b)
    assert h.nav_bytes() == data
    # point location on the loaded map is the same exact function
    tri, _, cen = g.map_arrays()
    lo, hi = tri.reshape(-1, 2).min(0), tri.reshape(-1, 2).max(0)
    r = np.random.default_rng(1)
    pts = np.concatenate([cen, (lo + (hi - lo) * r.random((500, 2))).astype(np.float32)])
    assert np.array_equal(g.locate(pts), h.locate(pts))


def test_layout_matches_the_specification():
    g = maze_map()
    tri, nbr, cen = g.map_arrays()
    b = g.nav_bytes()
    p = parse(b)
    V = len(tri)
    assert (p["version"], p["hdr"], p["V"]) == (1, 64, V)
    assert np.array_equal(p["tri"].view(np.uint32), tri.view(np.uint32))
    assert np.array_equal(p["cen"].view(np.uint32), cen.view(np.uint32))
    pts = tri.reshape(-1, 2)
    assert np.array_equal(p["bbox"], np.array([pts[:, 0].min(), pts[:, 1].min(), pts[:, 0].max(), pts[:, 1].max()],
                                              np.float32))
    deg = (nbr >= 0).sum(1)
    assert p["row"][0] == 0 and np.array_equal(np.diff(p["row"].astype(np.int64)), deg)
    assert p["E"] == deg.sum() == p["row"][-1]
    assert np.array_equal(p["col"].astype(np.int64), nbr[nbr >= 0].astype(np.int64))  # row-major, slot order
    assert p["ck"] == fnv1a64(b[:p["end"]])


def test_file_round_trip(tmp_path):
    g = built(S.bowtie_obstacles(), 100, 100)
    path = str(tmp_path / "m.nav")
    g.save_nav(path)
    assert open(path, "rb").read() == g.nav_bytes()
    h = fresh()
    h.load_nav(path)
    assert h.nav_bytes() == g.nav_bytes()


def test_committed_bowtie_fixture_is_what_fill_map_end_builds():
    data = open(GOLDEN, "rb").read()
    g = built(S.bowtie_obstacles(), 100, 100)
    assert g.nav_bytes() == data
    p = parse(data)
    # README map: 100 x 100 with the self-crossing wall split at (45, 50);
    # walkable area = 100^2 minus the two wall triangles, each with a 20-unit
    # side on x = 30 or x = 60 and its apex at (45, 50): 2 x (20 x 15 / 2)
    tri = p["tri"].astype(np.float64)
    a = 0.5 * np.abs((tri[:, 2] - tri[:, 0]) * (tri[:, 5] - tri[:, 1]) - (tri[:, 4] - tri[:, 0]) * (tri[:, 3] - tri[:, 1]))
    assert abs(a.sum() - (100 * 100 - 2 * 20 * 15 / 2)) < 1e-3
    assert np.array_equal(p["bbox"], np.array([0, 0, 100, 100], np.float32))
    h = fresh()
    h.load_nav(GOLDEN)
    assert h.locate(np.array([[45.0, 45.0], [45.0, 50.0], [10.0, 90.0]], np.float32)).tolist()[2] >= 0
    assert h.locate(np.array([[35.0, 50.0]], np.float32))[0] < 0  # inside the wall


def _reseal(b: bytearray):
    end = len(b) - 8
    struct.pack_into("<Q", b, end, fnv1a64(bytes(b[:end])))
    return bytes(b)


@pytest.mark.parametrize("case", ["short", "magic", "version", "size", "checksum", "rows", "id", "degree"])
def test_malformed_input_is_rejected(case):
    good = bytearray(maze_map().nav_bytes())
    p = parse(bytes(good))
    V, E = p["V"], p["E"]
    row_off = 64 + 32 * V
    col_off = row_off + 4 * (V + 1)
    b = bytearray(good)
    want = {"short": "shorter", "magic": "magic", "version": "version", "size": "size",
            "checksum": "checksum", "rows": "row", "id": "out of range", "degree": "row"}[case]
    if case == "short":
        b = b[:40]
    elif case == "magic":
        b[0:8] = b"GMAPFNAW"
    elif case == "version":
        struct.pack_into("<I", b, 8, 2)
    elif case == "size":
        b = b[:-12] + b[-8:]
    elif case == "checksum":
        b[100] ^= 1
    elif case == "rows":
        struct.pack_into("<I", b, row_off, 1)
        b = bytearray(_reseal(b))
    elif case == "id":
        struct.pack_into("<I", b, col_off, V)
        b = bytearray(_reseal(b))
    elif case == "degree":
        r0, r1 = struct.unpack_from("<2I", b, row_off)
        struct.pack_into("<I", b, row_off + 4, r0 + 4 if r0 + 4 <= E else r1)
        b = bytearray(_reseal(b))
    h = fresh()
    with pytest.raises(G.GmapfError) as e:
        h.fill_map_from_nav(bytes(b))
    assert e.value.code == -1 and want in e.value.msg, e.value.msg
    # nothing was loaded: the map can still be filled
    h.fill_map_from_nav(bytes(good))


def test_call_order_follows_fill_map_manually(tmp_path):
    data = built(S.bowtie_obstacles(), 100, 100).nav_bytes()
    g = G.GMAPF()
    with pytest.raises(G.GmapfError) as e:
        g.fill_map_from_nav(data)  # not initialised
    assert e.value.code == -2
    g.init(G.Config(agents_max_count=2))
    with pytest.raises(G.GmapfError) as e:
        g.nav_bytes()  # nothing filled yet
    assert e.value.code == -2
    with pytest.raises(G.GmapfError) as e:
        g.load_nav(str(tmp_path / "missing.nav"))
    assert e.value.code == -1
    g.fill_map_from_nav(data)
    with pytest.raises(G.GmapfError) as e:
        g.fill_map_from_nav(data)  # map already filled (GMAPF.cu:203-206)
    assert e.value.code == -2
    g.add_agent(10, 10)
