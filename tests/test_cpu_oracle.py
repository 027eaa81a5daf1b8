"""Pins of the CPU oracle (oracle/oracle.cpp) that do not reuse its code:
closed forms, invariants, brute force on tiny inputs and the bounded-queue
model.  The reference ships no tests or golden vectors, so these are what
fix the oracle's behaviour (see the oracle header)."""
import math

import numpy as np
import pytest

from gmapf_amd import scenarios as S
from tests import helpers as H


# --------------------------------------------------------------- queue model
def test_queue_keeps_q_smallest_by_f_then_node(oracle):
    """CuQueue.h:33-70 with (F, node) ties: the queue must behave like the
    brute-force model 'keep the Q smallest items, pop the minimum'."""
    r = np.random.RandomState(7)
    for Q in (1, 2, 3, 5, 16):
        for trial in range(40):
            n = 60
            ops = (r.rand(n) < 0.35).astype(np.int32)
            F = r.randint(0, 6, n).astype(np.float32)  # many equal F values
            node = r.permutation(1000)[:n].astype(np.int32)
            out = oracle.queue_script(Q, ops, F, node)
            model = []
            for i in range(n):
                if ops[i] == 0:
                    model.append((F[i], node[i]))
                    model.sort()
                    del model[Q:]
                else:
                    exp = model.pop(0)[1] if model else -1
                    assert out[i] == exp, (Q, trial, i)


# ------------------------------------------------------- tiny-graph searches
def brute_search(nbr, cen, S_, Tg, T, Q, K):
    """Independent serial statement of one search (Appendix-A rules):
    per round, lanes 0..T-1 pop their minimum (F, node); neighbours in list
    order; first claim wins; first meet in (lane, slot) order; frontier item
    i goes to lane i mod T; queues keep the Q smallest."""
    def d2(a, b):
        x = np.float32(cen[a][0]) - np.float32(cen[b][0])
        y = np.float32(cen[a][1]) - np.float32(cen[b][1])
        return np.float32(x * x) + np.float32(y * y)

    K = np.float32(K)
    F0 = np.float32(d2(S_, Tg) * K)
    queues = [[] for _ in range(T)]
    queues[0].append((F0, S_))
    queues[1].append((F0, Tg))
    info = {S_: (None, 0, 1), Tg: (None, 0, 0)}  # node -> (parent, G, dir)
    found = None
    rounds = 0
    while found is None:
        frontier = []
        for lane in range(T):
            if not queues[lane]:
                continue
            queues[lane].sort()
            _, c = queues[lane].pop(0)
            _, G, dr = info[c]
            for n in nbr[c]:
                if n < 0:
                    break
                h = d2(n, Tg) if dr else d2(n, S_)
                F = np.float32(np.float32(G + 1) + np.float32(h * K))
                if n not in info:
                    info[n] = (c, G + 1, dr)
                    frontier.append((F, n))
                elif info[n][2] != dr and found is None:
                    found = (c, n) if dr else (n, c)
        for i, it in enumerate(frontier):
            qq = queues[i % T]
            qq.append(it)
            qq.sort()
            del qq[Q:]
        rounds += 1
        if found is None and not any(queues):
            return False, [], rounds
    mS, mT = found
    left = []
    n = mS
    while n != S_:
        left.append(n)
        n = info[n][0]
    right = []
    n = mT
    while n != Tg:
        right.append(n)
        n = info[n][0]
    return True, left[::-1] + right + [Tg], rounds


def tiny_graph(seed, V, extra):
    """Random connected-ish graph with <= 3 neighbours per node (as a navmesh)."""
    r = np.random.RandomState(seed)
    adj = [[] for _ in range(V)]
    order = r.permutation(V)
    for k in range(1, V):  # random tree
        a, b = order[k], order[r.randint(0, k)]
        if len(adj[a]) < 3 and len(adj[b]) < 3:
            adj[a].append(b)
            adj[b].append(a)
    for _ in range(extra):
        a, b = r.randint(0, V, 2)
        if a != b and b not in adj[a] and len(adj[a]) < 3 and len(adj[b]) < 3:
            adj[a].append(b)
            adj[b].append(a)
    nbr = -np.ones((V, 3), np.int32)
    for i, l in enumerate(adj):
        nbr[i, :len(l)] = l
    cen = (r.randint(0, 64, (V, 2)) / 4.0).astype(np.float32)
    # triangles whose centroid is exactly cen (for the oracle's map input)
    tri = np.zeros((V, 6), np.float32)
    tri[:, 0], tri[:, 1] = cen[:, 0] - 1, cen[:, 1] - 1
    tri[:, 2], tri[:, 3] = cen[:, 0] + 2, cen[:, 1] - 1
    tri[:, 4], tri[:, 5] = cen[:, 0] - 1, cen[:, 1] + 2
    return tri, nbr


@pytest.mark.parametrize("seed", range(12))
def test_search_equals_brute_force_on_tiny_graphs(oracle, seed):
    V = 10 + 3 * seed
    tri, nbr = tiny_graph(seed, V, extra=seed)
    om = oracle.OracleMap(tri, nbr)
    cen = om.centroids()
    r = np.random.RandomState(100 + seed)
    for T, Q, K in ((2, 1, 1.0), (2, 2, 0.0), (4, 3, 1.0), (8, 16, 0.5), (3, 2, 2.0)):
        for _ in range(12):
            a, b = r.randint(0, V, 2)
            if a == b:
                continue
            f1, c1, st = om.search(int(a), int(b), T, Q, K)
            f2, c2, rounds = brute_search(nbr, cen, int(a), int(b), T, Q, K)
            assert f1 == f2 and list(c1) == c2, (seed, T, Q, K, a, b)
            assert st["rounds"] == rounds


def path_graph(n):
    V = n + 1
    nbr = -np.ones((V, 3), np.int32)
    for k in range(V):
        l = [x for x in (k - 1, k + 1) if 0 <= x < V]
        nbr[k, :len(l)] = l
    tri = np.zeros((V, 6), np.float32)
    for k in range(V):
        tri[k] = (3 * k, 0, 3 * k + 3, 0, 3 * k, 3)
    return tri, nbr


@pytest.mark.parametrize("n", range(1, 12))
def test_path_graph_closed_form(oracle, n):
    """On the path 0..n with T=2 each side advances one node per round, so
    the two fronts meet after ceil(n/2) rounds and the corridor is 1..n."""
    tri, nbr = path_graph(n)
    om = oracle.OracleMap(tri, nbr)
    for K in (0.0, 1.0):
        f, c, st = om.search(0, n, 2, 4, K)
        assert f and list(c) == list(range(1, n + 1))
        assert st["rounds"] == math.ceil(n / 2)
        assert st["length"] == len(c)


@pytest.mark.parametrize("B", [0, 1, 2, 3, 5, 8, 1000])
def test_batched_corridor_walk_closed_form(oracle, B):
    """Batched form (budget B > 0): after the meet the corridor is walked at
    two parent steps per round on each side, inside the budget, carried over
    frames, and the agent moves when both sides are done.  On the path 0..n
    with T=2 the meet takes R = ceil(n/2) rounds at mS = floor(n/2),
    mT = mS + 1, so G(mS) = floor(n/2) and G(mT) = ceil(n/2) - 1; the walk
    takes W = ceil(max(1, G(mS)) / 2) rounds and the agent first shows Move
    after frame ceil((R + W) / B) - 1.  The plain form (B = 0) moves it after
    frame 0 whatever n."""
    for n in range(1, 41):
        tri, nbr = path_graph(n)
        om = oracle.OracleMap(tri, nbr)
        f, c, st = om.search(0, n, 2, 4, 1.0)
        R = math.ceil(n / 2)
        assert f and st["rounds"] == R and st["mS"] == n // 2 and st["mT"] == n // 2 + 1
        W = math.ceil(max(1, n // 2) / 2)
        expect = 0 if B == 0 else math.ceil((R + W) / B) - 1
        w = oracle.OracleWorld(om, agents_max=1, radius=0.1, speed=1.0, parallel=1, threads=2, queue_cap=4,
                               round_budget=B)
        cen = om.centroids()
        assert w.add_agent(*cen[0]) == 0
        assert w.set_target(0, *cen[n]) == 0
        first = None
        for frame in range(expect + 2):
            w.step(1.0 / 60)
            _, state, _ = w.agents()
            if state[0] == 2:  # MOVE
                first = frame
                break
            assert state[0] == 1, (n, B, frame)  # SEARCH until the corridor is complete
        assert first == expect, (n, B, first, expect)
        assert list(w.corridor(0)) == list(range(1, n + 1))
        s = w.stats()
        assert s["found"] == 1 and s["rounds"] == R and s["corridor_nodes"] == n


def test_tree_corridor_is_the_unique_path(oracle):
    """In a tree every found corridor is the unique tree path; with Q >= V
    nothing is dropped, so every search must succeed."""
    tri, nbr = tiny_graph(5, 200, extra=0)
    lab = S.components(nbr)
    om = oracle.OracleMap(tri, nbr)
    adj = {i: [n for n in row if n >= 0] for i, row in enumerate(nbr)}

    def tree_path(a, b):
        prev = {a: None}
        st = [a]
        while st:
            u = st.pop()
            for v in adj[u]:
                if v not in prev:
                    prev[v] = u
                    st.append(v)
        out = []
        while b != a:
            out.append(b)
            b = prev[b]
        return out[::-1]

    r = np.random.RandomState(3)
    checked = 0
    for _ in range(300):
        a, b = r.randint(0, 200, 2)
        if a == b:
            continue
        for T in (2, 4, 8):
            f, c, _ = om.search(int(a), int(b), T, 256, 1.0)
            if lab[a] != lab[b]:
                assert not f
            else:
                assert f and list(c) == tree_path(int(a), int(b))
                checked += 1
    assert checked > 100


def test_corridor_invariants_on_grid_maps(oracle):
    tri, nbr = H.grid_map(8, 30, 30, cell=2.0, p=0.25)
    om = oracle.OracleMap(tri, nbr)
    for a, b in H.connected_pairs(2, nbr, 200):
        for T, Q in ((2, 4), (8, 16)):
            f, c, st = om.search(a, b, T, Q, 1.0)
            if not f:
                continue
            path = [a] + list(c)
            assert c[-1] == b and a not in c and len(set(c)) == len(c)
            for u, v in zip(path, path[1:]):
                assert v in nbr[u]
            assert st["length"] == len(c)


# ------------------------------------------------------------- point location
def test_oracle_grid_locate_equals_brute(oracle):
    tri, nbr = H.grid_map(12, 25, 25, cell=2.5, p=0.2)
    om = oracle.OracleMap(tri, nbr)
    r = np.random.RandomState(1)
    pts = np.concatenate([r.rand(20000, 2) * 70 - 4,
                          np.stack(np.meshgrid(np.arange(0, 63, 1.25), np.arange(0, 63, 1.25)), -1).reshape(-1, 2)])
    pts = pts.astype(np.float32)
    assert np.array_equal(om.locate(pts, grid=True), om.locate(pts, grid=False))


# ------------------------------------------------------------ crowd closed forms
def _strip_world(oracle, n_agents_xy, targets, radius=0.5, speed=1.0, hash_cell=0.0):
    tri, nbr = S.grid_navmesh(np.zeros((1, 40), bool), 4.0)  # 160 x 4 strip
    om = oracle.OracleMap(tri, nbr)
    w = oracle.OracleWorld(om, agents_max=len(n_agents_xy), radius=radius, speed=speed,
                           parallel=len(n_agents_xy), threads=2, queue_cap=64, hash_cell=hash_cell)
    for x, y in n_agents_xy:
        assert w.add_agent(x, y) >= 0
    for i, t in enumerate(targets):
        if t is not None:
            assert w.set_target(i, *t) == 0
    return w, om


def test_move_closed_form(oracle):
    """AgentsMover.h:57-70: toward the step point by speed*dt, snapping when
    within one step.  One agent, target in the adjacent triangle: frame 1
    finds the path, frame 2 sets the step point (centroid of the target
    triangle), frames 3.. move along the straight segment."""
    w, om = _strip_world(oracle, [(0.5, 3.5)], [(3.5, 0.5)], speed=6.0)
    cen = om.centroids()
    dt = 1.0 / 60
    w.step(dt)
    w.step(dt)
    si, sn, sxy = w.steps()
    assert si[0] == 0 and np.array_equal(sxy[0], cen[sn[0]])
    p0 = np.array([0.5, 3.5])
    goal = cen[sn[0]].astype(np.float64)
    d = np.linalg.norm(goal - p0)
    checked = 0
    for k in range(1, 200):
        w.step(dt)
        xy, st, _ = w.agents()
        si, sn2, sxy = w.steps()
        if si[0] == 0:  # still aiming at the centroid of the target triangle
            expect = p0 + (goal - p0) / d * min(k * 6.0 * dt, d)
            assert np.abs(xy[0] - expect).max() < 1e-4
            checked += 1
        else:  # inside the last corridor node: aims at the target position itself
            assert si[0] == 1 and np.array_equal(sxy[0], np.float32([3.5, 0.5])) or st[0] == 0
        if st[0] == 0:  # arrived within FLT_EPSILON + r (AgentsMover.h:162)
            assert np.linalg.norm(xy[0] - np.float32([3.5, 0.5])) <= 0.5 + 1e-6
            break
    assert checked > 5 and st[0] == 0


def test_two_agent_avoidance_closed_form(oracle):
    """AgentsMover.h:96-126: two agents closer than cd = 2.5 r push each other
    by (cd - d) along their separation, rotated by 10 degrees; a third
    moving agent far away makes the moving set non-empty."""
    r = 0.5
    cd = 2.5 * r
    a, b = (40.0, 2.0), (40.8, 2.0)  # d = 0.8 < cd
    w, om = _strip_world(oracle, [a, b, (1.0, 1.0)], [None, None, (150.0, 2.0)], radius=r, speed=0.0)
    cs, sn = oracle.rotation_constants()
    w.step(1.0 / 60)  # frame 1: agent 2 searches (moving set still empty)
    xy, _, _ = w.agents()
    assert np.array_equal(xy[0], np.float32(a)) and np.array_equal(xy[1], np.float32(b))
    w.step(1.0 / 60)  # frame 2: agent 2 is moving -> avoidance for all agents
    xy, _, _ = w.agents()
    d = 0.8
    push = cd - d
    for i, sign in ((0, -1.0), (1, 1.0)):
        vx, vy = sign * push, 0.0
        rx, ry = vx * cs - vy * sn, vx * sn + vy * cs
        exp = np.array([a, b][i]) + np.array([rx, ry])
        assert np.abs(xy[i] - exp).max() < 1e-5, (i, xy[i], exp)


def test_hash_avoidance_equals_brute_force_bitwise(oracle):
    """The oracle's spatial-hash avoidance must equal its O(N^2) form bit for
    bit (same neighbour sets, terms summed in ascending agent index)."""
    tri, nbr = H.grid_map(31, 40, 40, cell=2.0, p=0.1)
    om = oracle.OracleMap(tri, nbr)
    cen = om.centroids()
    lab = S.components(nbr)
    nodes = np.nonzero(lab == np.bincount(lab).argmax())[0]
    r = np.random.RandomState(4)
    pick = r.permutation(nodes)[:600]
    starts = cen[pick] + (r.rand(600, 2).astype(np.float32) - 0.5) * 0.2
    targs = cen[r.permutation(nodes)[:600]]
    worlds = []
    for hc in (0.0, 1.0, 2.5):
        w = oracle.OracleWorld(om, agents_max=600, radius=0.5, speed=2.0, parallel=600, threads=4, queue_cap=16,
                               hash_cell=hc)
        for p in starts:
            w.add_agent(*p)
        for i, t in enumerate(targs):
            w.set_target(i, *t)
        for _ in range(40):
            w.step(1.0 / 60)
        worlds.append(w.agents()[0])
    for other in worlds[1:]:
        same = (worlds[0].view(np.uint32) == other.view(np.uint32)) | (np.isnan(worlds[0]) & np.isnan(other))
        assert same.all()


def test_counting_sort_is_stable(oracle):
    tri, nbr = H.grid_map(2, 20, 20, cell=2.0, p=0.0)
    om = oracle.OracleMap(tri, nbr)
    w = oracle.OracleWorld(om, agents_max=500, radius=0.5, speed=1.0, parallel=500, threads=2, queue_cap=8,
                           hash_cell=1.0)
    r = np.random.RandomState(9)
    for p in r.rand(500, 2) * 39:
        w.add_agent(*p)
    w.set_target(0, 20.0, 20.0)
    w.step(1.0 / 60)
    w.step(1.0 / 60)
    keys, order = w.last_sort()
    valid = np.nonzero(keys >= 0)[0]
    assert np.array_equal(order, valid[np.argsort(keys[valid], kind="stable")])


# ------------------------------------------------------- request admission (F1)
def test_admission_defers_in_ascending_index_and_same_node_goes_idle(oracle):
    H.check_deferral_in_ascending_index(oracle, gpu=False)


@pytest.mark.parametrize("budget", [1, 2, 5])
def test_admission_fills_freed_slots_in_index_order(oracle, budget):
    H.check_slots_free_in_index_order(oracle, False, budget)


def test_retarget_cancels_search_in_flight(oracle):
    H.check_retarget_cancels_search_in_flight(oracle, gpu=False)
