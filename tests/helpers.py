"""Shared helpers: drive the product (C ABI via gmapf_amd) and the oracle on
identical seeded inputs and compare them."""
from __future__ import annotations

import numpy as np

import gmapf_amd as G
from gmapf_amd import scenarios as S


def grid_map(seed, nx, ny, cell=4.0, p=0.25):
    blocked = S.random_blocked(seed, nx, ny, p)
    return S.grid_navmesh(blocked, cell)


def connected_pairs(seed, nbr, n, same_component=True):
    """n (start, target) node pairs, distinct, mostly in one component."""
    lab = S.components(nbr)
    r = S.stream(seed, 7)
    V = len(nbr)
    out = []
    while len(out) < n:
        a, b = r.randint(V), r.randint(V)
        if a == b:
            continue
        if same_component and lab[a] != lab[b] and r.uniform() < 0.9:
            continue
        out.append((a, b))
    return out


def make_gpu(tri, nbr, cfg: G.Config, agents_xy, targets_xy=None):
    g = G.GMAPF()
    g.init(cfg)
    g.fill_map_manually(tri, nbr)
    for x, y in agents_xy:
        g.add_agent(x, y)
    if targets_xy is not None:
        for i, (x, y) in enumerate(targets_xy):
            if x is not None:
                g.set_agent_targ_pos(i, x, y)
    return g


def default_hash_cell(radius):
    """The library's avoidance hash cell when the config gives none: just above
    cd = 2.5 r, in the library's FP32 arithmetic (the avoidance result does not
    depend on the cell; the oracle uses the same one to share the grid)."""
    return float(np.float32(np.float32(radius) * np.float32(2.5)) * np.float32(1.0 + 1.0 / 64.0))


def make_oracle(oracle_mod, tri, nbr, cfg: G.Config, agents_xy, targets_xy=None, omap=None):
    omap = omap or oracle_mod.OracleMap(tri, nbr)
    budget = cfg.round_budget if cfg.step_mode == G.MODE_BATCHED else 0
    w = oracle_mod.OracleWorld(
        omap, agents_max=cfg.agents_max_count, radius=cfg.agent_radius, speed=cfg.agent_speed,
        parallel=cfg.path_finder_parallel_agents, threads=cfg.path_finder_threads_per_agents,
        queue_cap=cfg.path_finder_queue_capacity, heur_k=cfg.path_finder_heuristic_k,
        round_budget=budget, hash_cell=cfg.hash_cell if cfg.hash_cell > 0 else default_hash_cell(cfg.agent_radius),
        use_grid_locate=1)
    for x, y in agents_xy:
        assert w.add_agent(x, y) >= 0
    if targets_xy is not None:
        for i, (x, y) in enumerate(targets_xy):
            if x is not None:
                assert w.set_target(i, x, y) == 0
    return w


def compare_worlds(g: G.GMAPF, w, frame, check_corridors=True):
    gxy, gst, gnd = g.agents(with_nodes=True)
    oxy, ost, ond = w.agents()
    bad_st = np.nonzero(gst != ost)[0]
    assert len(bad_st) == 0, f"frame {frame}: state mismatch at agents {bad_st[:10]} gpu={gst[bad_st[:10]]} oracle={ost[bad_st[:10]]}"
    # bitwise position equality (stated tolerance 1e-4 map units; the
    # implementation is bit-exact so compare the bits, NaN included)
    gb, ob = gxy.view(np.uint32), oxy.view(np.uint32)
    same = (gb == ob) | (np.isnan(gxy) & np.isnan(oxy))  # NaN payloads may differ
    bad = np.nonzero(~same.all(axis=1))[0]
    if len(bad):
        d = np.abs(gxy[bad] - oxy[bad]).max()
        raise AssertionError(f"frame {frame}: {len(bad)} positions differ (max abs {d}), first {bad[:5]} "
                             f"gpu={gxy[bad[:3]]} oracle={oxy[bad[:3]]}")
    assert (gnd == ond).all(), f"frame {frame}: CurrNodeIdx mismatch"
    if check_corridors:
        for i in np.nonzero(gst == G.MOVE)[0]:
            a, b = g.corridor(int(i)), w.corridor(int(i))
            assert (a is None) == (b is None), f"frame {frame} agent {i}: corridor presence"
            if a is not None:
                assert np.array_equal(a, b), f"frame {frame} agent {i}: corridor differs"
