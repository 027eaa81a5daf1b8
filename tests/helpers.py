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


def compare_hash(g: G.GMAPF, w, frame):
    """The spatial hash of the last frame that moved agents: cell keys and the
    stable counting sort (ascending key, ascending agent index inside a cell)
    equal the oracle's bit for bit."""
    okeys, oorder = w.last_sort()
    if len(okeys) == 0 or not len(oorder):
        return False  # no crowd phase yet
    gkeys, gorder = g.hash_state()
    assert np.array_equal(gkeys, okeys), f"frame {frame}: hash keys differ"
    assert np.array_equal(gorder, oorder), f"frame {frame}: hash order differs"
    return True


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


# ------------------------------------------------------------------ request admission (F1) closed forms
# PathFinder.h:85-116: Search agents are taken in ascending index; one whose
# node equals its target node turns Idle (before the capacity test); the
# first P requesting agents are admitted, the rest stay Search.  The maps are
# disjoint path graphs (one per agent), so a search from one end of a path of
# n edges to the other with T = 2 meets after R = ceil(n / 2) rounds, and in
# the batched form its corridor walk takes W = ceil(max(1, floor(n / 2)) / 2)
# more rounds.

def path_components(lengths):
    """Disjoint path graphs, component c with lengths[c] edges on its own
    row; returns tri, nbr and node(c, k)."""
    base = np.concatenate([[0], np.cumsum([n + 1 for n in lengths])]).astype(int)
    V = int(base[-1])
    tri = np.zeros((V, 6), np.float32)
    nbr = -np.ones((V, 3), np.int32)
    for c, n in enumerate(lengths):
        for k in range(n + 1):
            v = base[c] + k
            y = 10.0 * c
            tri[v] = (3 * k, y, 3 * k + 3, y, 3 * k, y + 3)
            l = [base[c] + x for x in (k - 1, k + 1) if 0 <= x <= n]
            nbr[v, :len(l)] = l
    return tri, nbr, (lambda c, k: int(base[c] + k))


def path_rounds(n, budget):
    """Rounds a search over n edges occupies its slot: the search, plus the
    corridor walk in the batched form (budget > 0)."""
    R = -(-n // 2)
    return R + (-(-max(1, n // 2) // 2) if budget > 0 else 0)


class Adapter:
    """The same few calls on the HIP engine (gmapf_amd.GMAPF) and the oracle
    world (pyoracle.OracleWorld)."""

    def __init__(self, w):
        self.w = w
        self.gpu = isinstance(w, G.GMAPF)

    def step(self):
        self.w.step(1.0 / 60)

    def set_target(self, i, xy):
        if self.gpu:
            self.w.set_agent_targ_pos(i, float(xy[0]), float(xy[1]))
        else:
            assert self.w.set_target(i, float(xy[0]), float(xy[1])) == 0

    def states(self):
        return self.w.agents()[1]

    def stats(self):
        return self.w.stats()

    def corridor(self, i):
        return self.w.corridor(i)


def admission_world(oracle_mod, gpu, lengths, P, budget, starts, targets):
    """Agents at the given (component, k) starts with (component, k) targets
    on path_components(lengths), speed 0 (they never leave their start node),
    T = 2, Q = 4."""
    tri, nbr, node = path_components(lengths)
    cen = S.centroids(tri)
    sxy = [cen[node(*s)] for s in starts]
    txy = [cen[node(*t)] for t in targets]
    mode = G.MODE_BATCHED if budget > 0 else G.MODE_PLAIN
    cfg = G.Config(agents_max_count=len(starts), agent_radius=0.1, agent_speed=0.0,
                   path_finder_parallel_agents=P, path_finder_threads_per_agents=2, path_finder_queue_capacity=4,
                   step_mode=mode, round_budget=max(1, budget))
    if gpu:
        w = make_gpu(tri, nbr, cfg, sxy, txy)
    else:
        w = make_oracle(oracle_mod, tri, nbr, cfg, sxy, txy)
    return Adapter(w), cen, node


def check_deferral_in_ascending_index(oracle_mod, gpu):
    """Plain form, P = 3, ten agents, agent 4's target is its own node: frame
    0 admits agents 0, 1, 2 and makes agent 4 Idle (the same-node test comes
    before the capacity test); every later frame admits the next three
    requesting agents in index order (each search ends in its frame)."""
    N, L = 10, 6
    starts = [(c, 0) for c in range(N)]
    targets = [(c, L) for c in range(N)]
    targets[4] = (4, 0)
    w, _, _ = admission_world(oracle_mod, gpu, [L] * N, 3, 0, starts, targets)
    requesting = [i for i in range(N) if i != 4]
    for f in range(4):
        w.step()
        st = w.states()
        done = requesting[:3 * (f + 1)]
        exp = [G.MOVE if i in done else (G.IDLE if i == 4 else G.SEARCH) for i in range(N)]
        assert list(st) == exp, (f, list(st), exp)
        assert w.stats()["admitted"] == len(done), (f, w.stats())
    assert w.stats()["found"] == 9


def check_slots_free_in_index_order(oracle_mod, gpu, budget):
    """Batched form, budget B, P = 2, four agents on paths of n edges: agents
    0 and 1 hold both slots for D = ceil(path_rounds / B) frames and move
    after frame D - 1; agents 2 and 3 are admitted in frame D and move after
    frame 2D - 1."""
    n = 13
    w, _, _ = admission_world(oracle_mod, gpu, [n] * 4, 2, budget, [(c, 0) for c in range(4)],
                              [(c, n) for c in range(4)])
    D = -(-path_rounds(n, budget) // budget)
    first = {}
    for f in range(2 * D + 1):
        w.step()
        st = w.states()
        for i in range(4):
            if st[i] == G.MOVE and i not in first:
                first[i] = f
        adm = w.stats()["admitted"]
        assert adm == (2 if f < D else 4), (f, adm)
    assert first == {0: D - 1, 1: D - 1, 2: 2 * D - 1, 3: 2 * D - 1}, (first, D)


def check_retarget_cancels_search_in_flight(oracle_mod, gpu):
    """Batched form, budget 1, P = 1: agent 0 searches 0 -> 40 on its path,
    agent 1 waits for the slot.  Before frame 5 agent 0 is re-targeted to
    node 30: its search is dropped, its slot freed, and agent 0 (the lower
    index) takes the slot again with the new target, starting from scratch:
    it moves after frame 5 + path_rounds(30) - 1 with corridor 1..30, agent 1
    is admitted the frame after, and the dropped search's rounds count."""
    w, cen, node = admission_world(oracle_mod, gpu, [40, 8], 1, 1, [(0, 0), (1, 0)], [(0, 40), (1, 8)])
    move0 = 5 + path_rounds(30, 1) - 1
    move1 = move0 + 1 + path_rounds(8, 1) - 1
    for f in range(move1 + 2):
        if f == 5:
            w.set_target(0, cen[node(0, 30)])
        w.step()
        st = w.states()
        exp0 = G.MOVE if f >= move0 else G.SEARCH
        exp1 = G.MOVE if f >= move1 else G.SEARCH
        assert (st[0], st[1]) == (exp0, exp1), (f, st, move0, move1)
        assert w.stats()["admitted"] == (1 if f < 5 else (2 if f <= move0 else 3)), (f, w.stats())
    assert list(w.corridor(0)) == [node(0, k) for k in range(1, 31)]
    s = w.stats()
    assert s["found"] == 2 and s["rounds"] == 5 + -(-30 // 2) + -(-8 // 2), s
