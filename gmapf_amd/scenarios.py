"""Seeded workload generators for the per-frame step.

All randomness comes from SplitMix64 streams keyed by (seed, purpose), so a
workload is a pure function of its seed on every machine.

* ``grid_navmesh``      -- unit-square grids split into two triangles per free
                           cell (a FillMapManually map with compacted
                           neighbour lists);
* ``bowtie_map`` etc.   -- obstacle sets for FillMapStart/FillMapObstacle/
                           FillMapEnd (map sizes, obstacle shapes and agent
                           placements of the five benchmark workloads).
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014), 64-bit state, one output per call."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform_f32(self, lo: float, hi: float) -> float:
        return float(np.float32(lo + (hi - lo) * self.uniform()))

    def randint(self, n: int) -> int:
        return int(self.uniform() * n)


def stream(seed: int, purpose: int) -> SplitMix64:
    return SplitMix64((seed * 0x100000001B3 + purpose * 0x9E3779B97F4A7C15) & MASK64)


# ---------------------------------------------------------------- grid maps
def grid_navmesh(blocked: np.ndarray, cell: float = 1.0, x0: float = 0.0, y0: float = 0.0):
    """Triangulate the free cells of a boolean grid (blocked[y, x]).

    Each free cell (x, y) gives triangle A = (x0,y0),(x1,y0),(x1,y1) and
    B = (x0,y0),(x1,y1),(x0,y1).  A borders B, the cell below (its B) and the
    cell to the right (its B); B borders A, the cell above (its A) and the
    cell to the left (its A).  Neighbour lists are compacted with a -1 tail.
    Returns (tri (V,6) float32, nbr (V,3) int32).
    """
    blocked = np.asarray(blocked, dtype=bool)
    ny, nx = blocked.shape
    idx = -np.ones((ny, nx), dtype=np.int64)
    free = np.argwhere(~blocked)  # (y, x) in row-major order
    idx[free[:, 0], free[:, 1]] = np.arange(len(free))
    V = 2 * len(free)
    tri = np.zeros((V, 6), dtype=np.float32)
    nbr = -np.ones((V, 3), dtype=np.int32)
    for k, (y, x) in enumerate(free):
        xa, ya = np.float32(x0 + x * cell), np.float32(y0 + y * cell)
        xb, yb = np.float32(x0 + (x + 1) * cell), np.float32(y0 + (y + 1) * cell)
        a, b = 2 * k, 2 * k + 1
        tri[a] = (xa, ya, xb, ya, xb, yb)
        tri[b] = (xa, ya, xb, yb, xa, yb)
        na = [b]
        if y > 0 and idx[y - 1, x] >= 0:
            na.append(2 * idx[y - 1, x] + 1)
        if x + 1 < nx and idx[y, x + 1] >= 0:
            na.append(2 * idx[y, x + 1] + 1)
        nb = [a]
        if y + 1 < ny and idx[y + 1, x] >= 0:
            nb.append(2 * idx[y + 1, x])
        if x > 0 and idx[y, x - 1] >= 0:
            nb.append(2 * idx[y, x - 1])
        nbr[a, : len(na)] = na
        nbr[b, : len(nb)] = nb
    return tri, nbr


def random_blocked(seed: int, nx: int, ny: int, p: float) -> np.ndarray:
    r = stream(seed, 1)
    b = np.zeros((ny, nx), dtype=bool)
    for y in range(ny):
        for x in range(nx):
            b[y, x] = r.uniform() < p
    return b


def centroids(tri: np.ndarray) -> np.ndarray:
    """((P1 + P2) + P3) / 3 in FP32, the reference's PCenter (GMAPF.cu:215)."""
    t = tri.astype(np.float32)
    cx = ((t[:, 0] + t[:, 2]) + t[:, 4]) / np.float32(3.0)
    cy = ((t[:, 1] + t[:, 3]) + t[:, 5]) / np.float32(3.0)
    return np.stack([cx, cy], axis=1).astype(np.float32)


def components(nbr: np.ndarray) -> np.ndarray:
    """Connected-component label per node (undirected, over compacted lists)."""
    V = len(nbr)
    lab = -np.ones(V, dtype=np.int64)
    c = 0
    for s in range(V):
        if lab[s] >= 0:
            continue
        lab[s] = c
        stack = [s]
        while stack:
            u = stack.pop()
            for v in nbr[u]:
                if v < 0:
                    break
                if lab[v] < 0:
                    lab[v] = c
                    stack.append(v)
        c += 1
    return lab


def uniform_free_points(seed: int, purpose: int, n: int, w: float, h: float, locate, batch=4096):
    """n points uniform in [0,w)x[0,h) kept only where ``locate`` finds a node."""
    r = stream(seed, purpose)
    out = []
    while len(out) < n:
        m = max(batch, 2 * (n - len(out)))
        pts = np.array([[r.uniform() * w, r.uniform() * h] for _ in range(m)], dtype=np.float32)
        ok = locate(pts) >= 0
        for p in pts[ok]:
            out.append(p)
            if len(out) == n:
                break
    return np.array(out, dtype=np.float32).reshape(-1, 2)
