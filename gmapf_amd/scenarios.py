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


# ------------------------------------------------ counter-based vector draws
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def draws(seed: int, purpose: int, counters) -> np.ndarray:
    """SplitMix64 outputs number ``counters`` (uint64 array) of the stream
    (seed, purpose), as doubles in [0, 1)."""
    key = np.uint64((seed * 0x100000001B3 + purpose * 0x9E3779B97F4A7C15) & MASK64)
    c = np.asarray(counters, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = key + (c + np.uint64(1)) * _GOLD
    return (_mix(z) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


# --------------------------------------------------------------- workloads
class Scenario:
    """A map (FillMapStart + obstacles), agents and run parameters."""

    def __init__(self, name, width, height, obstacles, starts, targets, *, radius, speed, T, Q, K=1.0,
                 frames, dt=1.0 / 60, seed=1, P=None, retarget=None, notes=""):
        self.name = name
        self.width, self.height = float(width), float(height)
        self.obstacles = obstacles
        self.starts = np.asarray(starts, dtype=np.float32).reshape(-1, 2)
        self.targets = np.asarray(targets, dtype=np.float32).reshape(-1, 2)
        self.radius, self.speed = radius, speed
        self.T, self.Q, self.K = T, Q, K
        self.frames, self.dt = frames, dt
        self.seed = seed
        self.P = P or len(self.starts)
        self.retarget = retarget
        self.notes = notes

    @property
    def n(self):
        return len(self.starts)

    def build(self, g):
        """FillMapStart / FillMapObstacle / FillMapEnd on a GMAPF-like object."""
        g.fill_map_start(self.width, self.height)
        for o in self.obstacles:
            g.fill_map_obstacle(o)
        g.fill_map_end()


def _free_points(seed, purpose, n, x0, y0, w, h, locate, start_counter=0):
    out = np.zeros((0, 2), np.float32)
    c = start_counter
    while len(out) < n:
        m = max(4096, 3 * (n - len(out)))
        k = np.arange(c, c + m, dtype=np.uint64)
        c += m
        px = (x0 + w * draws(seed, purpose, 2 * k)).astype(np.float32)
        py = (y0 + h * draws(seed, purpose, 2 * k + 1)).astype(np.float32)
        pts = np.stack([px, py], 1)
        out = np.concatenate([out, pts[locate(pts) >= 0]])
    return out[:n]


def bowtie_obstacles():
    # the reference README's example wall (README.md:66-72): a self-crossing
    # polygon, split at (45, 50) by the triangulator
    return [np.array([(30, 40), (30, 60), (60, 40), (60, 60)], np.float32)]


def rect_obstacles(seed, W, H, n, smin, smax, gap=2.0):
    """n axis-aligned rectangles, sides U[smin, smax], at least `gap` apart and
    from the map border (2 units keeps every vertex clear of the triangulator's
    1-unit snap tolerance)."""
    r = stream(seed, 21)
    rects = []
    # accepted rects bucketed by the cell of their lower-left corner: a rect
    # that can touch a candidate has its corner within smax + gap of the
    # candidate's corner, so the 3x3 cells around it hold every such rect
    cs = smax + gap
    buckets = {}
    tries = 0
    while len(rects) < n and tries < 200000:
        tries += 1
        w = smin + (smax - smin) * r.uniform()
        h = smin + (smax - smin) * r.uniform()
        x = gap + (W - w - 2 * gap) * r.uniform()
        y = gap + (H - h - 2 * gap) * r.uniform()
        x, y, w, h = (round(v * 16) / 16 for v in (x, y, w, h))  # 1/16-unit grid
        cx, cy = int(x // cs), int(y // cs)
        near = (b for i in (cx - 1, cx, cx + 1) for j in (cy - 1, cy, cy + 1) for b in buckets.get((i, j), ()))
        if any(x < b[2] + gap and b[0] < x + w + gap and y < b[3] + gap and b[1] < y + h + gap for b in near):
            continue
        rects.append((x, y, x + w, y + h))
        buckets.setdefault((cx, cy), []).append(rects[-1])
    return [np.array([(a, b), (a, d), (c, d), (c, b)], np.float32) for a, b, c, d in rects]


def maze_obstacles(seed, cells=64, cell=64.0, corridor=16.0):
    """Spanning-tree maze: a cells x cells grid, random spanning tree
    (Kruskal over SplitMix64 edge weights), corridors `corridor` wide between
    the centres of tree-adjacent cells.  Walls are axis-aligned blocks on the
    lines {0, a, a+corridor, cell} of every cell (a = (cell-corridor)/2), so
    adjacent blocks share whole edges."""
    a = (cell - corridor) / 2
    b = a + corridor
    r = stream(seed, 31)
    edges = []
    for y in range(cells):
        for x in range(cells):
            if x + 1 < cells:
                edges.append((r.next_u64(), y * cells + x, y * cells + x + 1))
            if y + 1 < cells:
                edges.append((r.next_u64(), y * cells + x, (y + 1) * cells + x))
    edges.sort()
    parent = list(range(cells * cells))

    def find(u):
        while parent[u] != u:
            parent[u] = parent[parent[u]]
            u = parent[u]
        return u

    open_e = set()
    for _, u, v in edges:
        ru, rv = find(u), find(v)
        if ru != rv:
            parent[ru] = rv
            open_e.add((u, v))
    blocks = []
    for y in range(cells):
        for x in range(cells):
            ox, oy = x * cell, y * cell
            c = y * cells + x
            for (bx0, by0, bx1, by1) in ((0, 0, a, a), (b, 0, cell, a), (0, b, a, cell), (b, b, cell, cell)):
                blocks.append((ox + bx0, oy + by0, ox + bx1, oy + by1))
            if not (x + 1 < cells and (c, c + 1) in open_e):
                blocks.append((ox + b, oy + a, ox + cell, oy + b))      # right arm
            if not (x > 0 and (c - 1, c) in open_e):
                blocks.append((ox, oy + a, ox + a, oy + b))             # left arm
            if not (y + 1 < cells and (c, c + cells) in open_e):
                blocks.append((ox + a, oy + b, ox + b, oy + cell))      # top arm
            if not (y > 0 and (c - cells, c) in open_e):
                blocks.append((ox + a, oy, ox + b, oy + a))             # bottom arm
    return [np.array([(p, q), (p, s), (o, s), (o, q)], np.float32) for p, q, o, s in blocks]


def convex_obstacles(seed, W, H, n, rmin, rmax, vmin=6, vmax=12, gap=4.0, margin=300.0):
    """n non-overlapping convex polygons: vmin..vmax vertices on a circle of
    radius U[rmin, rmax] at sorted jittered angles (vertices on a 1/8 grid),
    kept out of the two side bands of width `margin`."""
    r = stream(seed, 41)
    circ, obs = [], []
    tries = 0
    while len(obs) < n and tries < 100000:
        tries += 1
        rad = rmin + (rmax - rmin) * r.uniform()
        cx = margin + rad + (W - 2 * margin - 2 * rad) * r.uniform()
        cy = rad + gap + (H - 2 * rad - 2 * gap) * r.uniform()
        if any(math.hypot(cx - px, cy - py) < rad + pr + gap for px, py, pr in circ):
            continue
        k = vmin + r.randint(vmax - vmin + 1)
        step = 2 * math.pi / k
        ang = [i * step + (r.uniform() - 0.5) * 0.5 * step for i in range(k)]
        pts = [(round((cx + rad * math.cos(t)) * 8) / 8, round((cy + rad * math.sin(t)) * 8) / 8) for t in ang]
        circ.append((cx, cy, rad))
        obs.append(np.array(pts, np.float32))
    return obs


def log_uniform_targets(seed, purpose, starts, dmin, dmax, W, H, locate, counters, ids=None):
    """Per agent: distance exp(U[ln dmin, ln dmax]) in a uniform direction,
    redrawn until the point is on the navmesh.  Agent ids[i] (default i)
    draws from its own counter counters[ids[i]], advanced in place, so an
    agent's k-th target does not depend on which other agents re-path."""
    starts = np.asarray(starts, np.float32)
    n = len(starts)
    ids = np.arange(n) if ids is None else np.asarray(ids)
    out = np.zeros((n, 2), np.float32)
    todo = np.arange(n)
    while len(todo):
        aid = ids[todo]
        base = (np.asarray(counters[aid], dtype=np.uint64) * np.uint64(2)
                + np.asarray(aid, dtype=np.uint64) * np.uint64(1 << 40))
        u1 = draws(seed, purpose, base)
        u2 = draws(seed, purpose, base + np.uint64(1))
        d = np.exp(math.log(dmin) + (math.log(dmax) - math.log(dmin)) * u1)
        th = 2 * math.pi * u2
        p = np.stack([starts[todo, 0] + d * np.cos(th), starts[todo, 1] + d * np.sin(th)], 1).astype(np.float32)
        counters[aid] += 1
        ok = (p[:, 0] >= 0) & (p[:, 0] <= W) & (p[:, 1] >= 0) & (p[:, 1] <= H)
        ok[ok] = locate(p[ok]) >= 0
        out[todo[ok]] = p[ok]
        todo = todo[~ok]
    return out


def config1(locate_factory, seed=1):
    """100x100, bow-tie wall, 64 agents r=2 v=1, T=4 Q=16 K=1, 100 frames."""
    obs = bowtie_obstacles()
    loc = locate_factory(100, 100, obs)
    st = _free_points(seed, 101, 64, 0, 0, 100, 100, loc)
    tg = _free_points(seed, 102, 64, 0, 0, 100, 100, loc)
    return Scenario("bowtie-100", 100, 100, obs, st, tg, radius=2.0, speed=1.0, T=4, Q=16, K=1.0, frames=100,
                    seed=seed)


def config2(locate_factory, seed=1, n_agents=2048, n_rects=400, size=1000.0):
    """1000x1000, 400 rects sides U[5,40], 2048 agents requesting at frame 0, T=8 Q=32, 300 frames."""
    obs = rect_obstacles(seed, size, size, n_rects, 5.0, 40.0)
    loc = locate_factory(size, size, obs)
    st = _free_points(seed, 201, n_agents, 0, 0, size, size, loc)
    tg = _free_points(seed, 202, n_agents, 0, 0, size, size, loc)
    return Scenario(f"rects-{int(size)}", size, size, obs, st, tg, radius=2.0, speed=1.0, T=8, Q=32,
                    frames=300, seed=seed)


def config3(locate_factory, seed=1, n_agents=16384, cells=64, T=8, Q=32):
    """4096x4096 spanning-tree maze (64-unit cells, 16-wide corridors),
    16,384 agents, start-target distance log-uniform over 10-4000."""
    W = cells * 64.0
    obs = maze_obstacles(seed, cells)
    loc = locate_factory(W, W, obs)
    st = _free_points(seed, 301, n_agents, 0, 0, W, W, loc)
    counters = np.zeros(n_agents, dtype=np.int64)
    tg = log_uniform_targets(seed, 302, st, 10.0, min(4000.0, W), W, W, loc, counters)
    sc = Scenario(f"maze-{int(W)}", W, W, obs, st, tg, radius=2.0, speed=1.0, T=T, Q=Q, frames=120, seed=seed)
    sc.counters = counters
    sc.locate = loc
    return sc


def config4(locate_factory, seed=1, n_agents=65536, size=2048.0, band=256.0):
    """2048x2048, 50 convex obstacles (6-12 vertices, radius 20-80), 65,536
    agents r=2 v=1 in two opposing streams between the side bands; hash cell 4."""
    obs = convex_obstacles(seed, size, size, 50, 20.0, 80.0, margin=band + 8)
    loc = locate_factory(size, size, obs)
    h = n_agents // 2
    left = _free_points(seed, 401, h, 0, 0, band, size, loc)
    right = _free_points(seed, 402, n_agents - h, size - band, 0, band, size, loc)
    tl = _free_points(seed, 403, h, size - band, 0, band, size, loc)
    tr = _free_points(seed, 404, n_agents - h, 0, 0, band, size, loc)
    st = np.concatenate([left, right])
    tg = np.concatenate([tl, tr])
    return Scenario(f"crowd-{int(size)}", size, size, obs, st, tg, radius=2.0, speed=1.0, T=8, Q=32, frames=600,
                    seed=seed)
