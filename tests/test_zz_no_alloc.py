"""The allocation-free bulk point location (gmapf_locate_device_noalloc)
against the allocating one (gmapf_locate_device) through the C ABI: equal
results on two calls in a row, the second smaller, and sizes beyond the
staging buffer rejected before any device work."""
import numpy as np
import pytest

import gmapf_amd as G
from gmapf_amd import scenarios as S
from tests import helpers as H

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 20


def _engine(seed):
    tri, nbr = H.grid_map(seed, 40, 30, cell=3.0, p=0.3)
    g = H.make_gpu(tri, nbr, G.Config(agents_max_count=1, path_finder_parallel_agents=1),
                   [S.centroids(tri)[0]])
    g.manual_start()
    return g


def _points(seed, n):
    r = np.random.RandomState(seed)
    pts = (r.rand(n, 2) * np.float32([126.0, 96.0]) - 3.0).astype(np.float32)
    gx, gy = np.meshgrid(np.arange(0, 121, 1.5, dtype=np.float32), np.arange(0, 91, 1.5, dtype=np.float32))
    lattice = np.stack([gx.ravel(), gy.ravel()], 1)  # vertices and edge points: boundary-inclusive cases
    return np.concatenate([lattice, pts]).astype(np.float32)[:n]


def test_noalloc_locate_equals_allocating_locate_twice():
    g = _engine(7)
    print("noalloc locate: large call", flush=True)
    big = _points(1, 400000)
    a1 = g.locate(big, device=True)
    b1 = g.locate(big, noalloc=True)
    assert np.array_equal(a1, b1)
    assert (b1 >= 0).any() and (b1 < 0).any()
    print("noalloc locate: smaller call", flush=True)
    small = _points(2, 3000)
    a2 = g.locate(small, device=True)
    b2 = g.locate(small, noalloc=True)
    assert np.array_equal(a2, b2)
    assert np.array_equal(b2, g.locate(small))  # host grid, same exact answer
    g.close()


def test_noalloc_locate_shared_by_handles_and_bounded():
    g1, g2 = _engine(3), _engine(4)
    pts = _points(5, 20000)
    print("noalloc locate: two handles", flush=True)
    for g in (g1, g2, g1):
        assert np.array_equal(g.locate(pts, noalloc=True), g.locate(pts, device=True))
    print("noalloc locate: the largest call", flush=True)
    full = _points(6, MAX_POINTS)
    assert len(full) == MAX_POINTS
    assert np.array_equal(g2.locate(full, noalloc=True), g2.locate(full, device=True))
    with pytest.raises(G.GmapfError) as e:
        g1.locate(np.zeros((MAX_POINTS + 1, 2), np.float32), noalloc=True)
    assert e.value.code == -1 and "at most 1048576" in e.value.msg
    assert g1.locate(np.zeros((0, 2), np.float32), noalloc=True).shape == (0,)
    g1.close()
    g2.close()
