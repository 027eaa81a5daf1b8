"""C-ABI checks that need no GPU: every declared symbol is exported, host-side
API calls follow the reference's ordering/error contract
(src/GMAPF.cu:58-253), and host point location equals the reference scan."""
import ctypes
import os
import re

import numpy as np
import pytest

import gmapf_amd as G
from gmapf_amd import scenarios as S
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        names += re.findall(r"^\s*(?:int|const char\*)\s+(gmapf_\w+)\s*\(", txt, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(G.library_path())
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_config_default_matches_reference_readme_shape():
    lib = G.load_library()
    c = G._Config()
    assert lib.gmapf_config_default(ctypes.byref(c)) == 0
    assert c.path_finder_threads_per_agents >= 2 and c.path_finder_queue_capacity >= 1
    assert c.path_finder_heuristic_k == 1.0


def test_call_order_errors():
    g = G.GMAPF()
    with pytest.raises(G.GmapfError) as e:
        g.add_agent(1, 1)  # GMAPF.cu:221-222 "not initialized"
    assert e.value.code == -2
    with pytest.raises(G.GmapfError) as e:
        g.async_step(0.1)
    assert e.value.code == -2
    g.init(G.Config(agents_max_count=2))
    with pytest.raises(G.GmapfError) as e:
        g.add_agent(1, 1)  # no map yet
    assert e.value.code == -2
    with pytest.raises(G.GmapfError):
        g.profiler_dump()  # GMAPF.cu:328-329 "profiler disabled"
    tri, nbr = H.grid_map(1, 6, 6, cell=2.0, p=0.0)
    g.fill_map_manually(tri, nbr)
    with pytest.raises(G.GmapfError) as e:
        g.fill_map_manually(tri, nbr)  # GMAPF.cu:205-206 "map is already filled"
    assert e.value.code == -2
    with pytest.raises(G.GmapfError) as e:
        g.add_agent(100, 100)  # GMAPF.cu:227-228 "initial pos ... is invalid"
    assert e.value.code == -3
    a = g.add_agent(1, 1)
    b = g.add_agent(3, 3)
    assert (a, b) == (0, 1)  # sequential ids (CuList.h:33-40)
    with pytest.raises(G.GmapfError):
        g.add_agent(5, 5)  # AgentsMaxCount reached
    with pytest.raises(G.GmapfError) as e:
        g.set_agent_targ_pos(0, -5, 1)  # GMAPF.cu:244-245 "target pos ... is invalid"
    assert e.value.code == -3
    with pytest.raises(G.GmapfError) as e:
        g.set_agent_targ_pos(7, 1, 1)
    assert e.value.code == -1
    g.set_agent_targ_pos(0, 9, 9)
    assert g.get_agent_state(0) == G.SEARCH  # GMAPF.cu:248
    assert g.get_agent_state(1) == G.IDLE
    assert g.get_agent_pos(1) == (3.0, 3.0)
    st = g.set_agent_targets([0, 1], [[9, 9], [-1, -1]], raise_on_error=False)
    assert list(st) == [0, -3]
    g.close()


def test_searches_per_wave_bounded_by_lanes():
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=1, path_finder_threads_per_agents=4, searches_per_wave=17))
    tri, nbr = H.grid_map(1, 6, 6, cell=2.0, p=0.0)
    g.fill_map_manually(tri, nbr)
    with pytest.raises(G.GmapfError) as e:
        g.manual_start()  # 4 lanes per search: at most 16 searches per 64-lane workgroup
    assert e.value.code == -1
    g.close()


@pytest.mark.parametrize("k", [-1.0, -1e-30, float("nan"), float("inf")])
def test_heuristic_weight_must_be_finite_and_not_negative(k):
    """HeuristicK < 0 or not finite is refused at ManualStart (the
    reference takes any float; a negative weight ranks nodes farther from
    the other end first, and the search's queue keys assume F >= 1)."""
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=1, path_finder_heuristic_k=k))
    tri, nbr = H.grid_map(1, 6, 6, cell=2.0, p=0.0)
    g.fill_map_manually(tri, nbr)
    with pytest.raises(G.GmapfError) as e:
        g.manual_start()
    assert e.value.code == -1
    g.close()


def test_negative_search_timing_rejected():
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=1, search_timing=-1))
    tri, nbr = H.grid_map(1, 6, 6, cell=2.0, p=0.0)
    g.fill_map_manually(tri, nbr)
    with pytest.raises(G.GmapfError) as e:
        g.manual_start()
    assert e.value.code == -1
    g.close()


def test_bad_neighbour_index_rejected():
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=1))
    tri, nbr = H.grid_map(1, 3, 3, cell=1.0, p=0.0)
    nbr = nbr.copy()
    nbr[0, 0] = len(tri) + 5
    with pytest.raises(G.GmapfError) as e:
        g.fill_map_manually(tri, nbr)
    assert e.value.code == -1


def test_host_locate_equals_reference_scan(oracle):
    tri, nbr = H.grid_map(4, 30, 22, cell=3.0, p=0.3)
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=1))
    g.fill_map_manually(tri, nbr)
    r = S.stream(3, 3)
    pts = np.array([[r.uniform() * 100 - 5, r.uniform() * 76 - 5] for _ in range(20000)], np.float32)
    gx, gy = np.meshgrid(np.arange(-1.5, 92, 1.5, dtype=np.float32), np.arange(-1.5, 68, 0.75, dtype=np.float32))
    pts = np.concatenate([pts, np.stack([gx.ravel(), gy.ravel()], 1),
                          np.array([[np.nan, 1.0], [1.0, np.inf]], np.float32)]).astype(np.float32)
    brute = oracle.OracleMap(tri, nbr).locate(pts, grid=False)
    assert np.array_equal(g.locate(pts), brute)
    # NaN passes every sign test (CuNodesMap.h:260-270): triangle 0
    assert brute[-2] == 0


def test_fill_map_end_bowtie_navmesh_invariants():
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=1))
    g.fill_map_start(100, 100)
    g.fill_map_obstacle(S.bowtie_obstacles()[0])
    g.fill_map_end()
    tri, nbr, cen = g.map_arrays()
    area = np.abs((tri[:, 2] - tri[:, 0]) * (tri[:, 5] - tri[:, 1]) - (tri[:, 4] - tri[:, 0]) * (tri[:, 3] - tri[:, 1])) / 2
    # square minus the two lobes of the self-crossing wall: 10000 - 2 * (20 * 15 / 2)
    assert abs(area.sum() - 9700.0) < 1e-3
    assert np.all(area > 0)
    for i, row in enumerate(nbr):
        seen_invalid = False
        for n in row:
            if n < 0:
                seen_invalid = True
                continue
            assert not seen_invalid  # compacted, -1 tail (GMAPF.cu:185-190)
            assert i in nbr[n]  # symmetric adjacency
    # the crossing point (45, 50) of the two wall edges is a vertex
    v = tri.reshape(-1, 2)
    assert ((v[:, 0] == 45) & (v[:, 1] == 50)).any()
    np.testing.assert_array_equal(cen, S.centroids(tri))
