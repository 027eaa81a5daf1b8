"""GPU parity: the HIP step (through the C ABI) against the CPU oracle on
identical seeded inputs.  Integer outputs (states, corridors, node ids) must
be equal; FP32 positions are compared bit for bit (the stated tolerance of
1e-4 map units is looser than what is checked)."""
import numpy as np
import pytest

import gmapf_amd as G
from gmapf_amd import scenarios as S
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("T,Q,spw,wide", [(2, 16, 0, False), (4, 16, 0, False), (8, 32, 0, False), (16, 8, 0, False),
                                          (32, 4, 0, False), (64, 2, 0, False), (4, 1, 0, False),
                                          (4, 16, 8, False), (2, 16, 7, False), (8, 16, 1, False),
                                          # the other compile-time queue kernels and the
                                          # deeper heaps of the autotuner's grid
                                          (2, 32, 0, False), (4, 32, 0, False), (4, 64, 0, False),
                                          (16, 128, 0, False), (64, 256, 0, False),
                                          # keys without G and side (the form maps of more
                                          # than 32,768 nodes take) in the compile-time
                                          # queue kernels
                                          (2, 16, 0, True), (4, 16, 0, True), (8, 16, 0, True), (2, 32, 0, True),
                                          (4, 32, 0, True), (8, 32, 0, True)])
def test_single_frame_search_parity(oracle, T, Q, spw, wide):
    """One frame, every agent searches once; corridors and flags vs
    oracle.search.  spw > 0: fewer searches per workgroup than 64 / T.
    wide: queue keys without G and side, which then come from the popped
    node's visited entry."""
    tri, nbr = H.grid_map(11, 48, 40, cell=4.0, p=0.3)
    cen = S.centroids(tri)
    pairs = H.connected_pairs(5, nbr, 300)
    cfg = G.Config(agents_max_count=len(pairs), path_finder_parallel_agents=len(pairs),
                   path_finder_threads_per_agents=T, path_finder_queue_capacity=Q,
                   agent_radius=0.25, agent_speed=1.0, searches_per_wave=spw, search_wide_keys=wide)
    g = H.make_gpu(tri, nbr, cfg, [cen[a] for a, _ in pairs], [cen[b] for _, b in pairs])
    print(f"search parity T={T} Q={Q} spw={spw} wide={wide} V={len(tri)}", flush=True)
    g.step(1.0 / 60)
    om = oracle.OracleMap(tri, nbr)
    _, st = g.agents()
    nfound = 0
    for i, (a, b) in enumerate(pairs):
        found, cor, _ = om.search(a, b, T, Q, 1.0)
        assert st[i] == (G.MOVE if found else G.IDLE), f"agent {i} ({a}->{b}) state"
        if found:
            nfound += 1
            assert np.array_equal(g.corridor(i), cor), f"agent {i} ({a}->{b}) corridor"
    s = g.stats()
    assert s["found"] == nfound and s["not_found"] == len(pairs) - nfound
    assert nfound > 0


@pytest.mark.parametrize("K", [0.0, 0.25, 3.0])
def test_search_parity_at_other_heuristic_weights(oracle, K):
    """HeuristicK 0 (the seeds' F is +0, every other F is G + 1), below and
    above 1: flags and corridors equal the oracle's."""
    tri, nbr = H.grid_map(13, 40, 40, cell=4.0, p=0.3)
    cen = S.centroids(tri)
    pairs = H.connected_pairs(7, nbr, 200)
    cfg = G.Config(agents_max_count=len(pairs), path_finder_parallel_agents=len(pairs),
                   path_finder_threads_per_agents=4, path_finder_queue_capacity=16, path_finder_heuristic_k=K,
                   agent_radius=0.25, agent_speed=1.0)
    g = H.make_gpu(tri, nbr, cfg, [cen[a] for a, _ in pairs], [cen[b] for _, b in pairs])
    print(f"search parity K={K} V={len(tri)}", flush=True)
    g.step(1.0 / 60)
    om = oracle.OracleMap(tri, nbr)
    _, st = g.agents()
    nfound = 0
    for i, (a, b) in enumerate(pairs):
        found, cor, _ = om.search(a, b, 4, 16, K)
        assert st[i] == (G.MOVE if found else G.IDLE), f"agent {i} ({a}->{b}) state"
        if found:
            nfound += 1
            assert np.array_equal(g.corridor(i), cor), f"agent {i} ({a}->{b}) corridor"
    assert nfound > 0
    g.close()


def _one_way_map(hub):
    """A grid navmesh whose neighbour lists are no longer a mesh's (the
    FillMapManually input may be any lists): the back link of about one edge
    in seven dropped, and with `hub` one node appended to one list in twenty
    of those with room, so that it has far more than four in-neighbours."""
    tri, nbr = H.grid_map(11, 48, 40, cell=4.0, p=0.3)
    nbr = np.array(nbr, dtype=np.int64)
    r = S.stream(31, 3)
    V = len(nbr)
    for i in range(V):
        for k in range(3):
            j = nbr[i, k]
            if j < 0:
                break
            if j > i and r.uniform() < 0.15:  # drop j -> i, keep the list compacted
                row = [x for x in nbr[j] if x >= 0 and x != i]
                nbr[j] = row + [-1] * (3 - len(row))
    indeg = np.bincount(nbr[nbr >= 0], minlength=V)
    if hub:
        h = int(np.argmax(indeg))
        for i in range(V):
            row = [x for x in nbr[i] if x >= 0]
            if len(row) < 3 and i != h and h not in row and r.uniform() < 0.05:
                nbr[i] = row + [h] + [-1] * (2 - len(row))
        indeg = np.bincount(nbr[nbr >= 0], minlength=V)
        assert indeg.max() > 4
    else:
        assert indeg.max() <= 4
    return tri, nbr


@pytest.mark.parametrize("hub,mode,budget", [(False, G.MODE_PLAIN, 0), (True, G.MODE_PLAIN, 0),
                                             (False, G.MODE_BATCHED, 3), (True, G.MODE_BATCHED, 3)])
def test_world_parity_on_one_way_neighbour_lists(oracle, hub, mode, budget):
    """Agents search and move over neighbour lists with one-way links (the
    search follows a node's own list; a corridor walks parents, which need
    not list their children), and with `hub` over a node listed by far more
    than three nodes.  Every frame equals the oracle's (states, positions,
    nodes; corridors every fifth frame); the batched form walks corridors
    over several frames (parked walks)."""
    tri, nbr = _one_way_map(hub)
    cen = S.centroids(tri)
    r = S.stream(6, 3)
    V = len(tri)
    perm = np.array(sorted(range(V), key=lambda _: r.next_u64()))
    n = 200
    starts = [cen[perm[i]] for i in range(n)]  # distinct start nodes
    targs = [cen[r.randint(V)] for _ in range(n)]
    cfg = G.Config(agents_max_count=n, path_finder_parallel_agents=n, path_finder_threads_per_agents=4,
                   path_finder_queue_capacity=16, agent_radius=0.5, agent_speed=2.0, step_mode=mode,
                   round_budget=max(1, budget))
    print(f"one-way lists hub={hub} mode={mode} V={V}", flush=True)
    g = H.make_gpu(tri, nbr, cfg, starts, targs)
    w = H.make_oracle(oracle, tri, nbr, cfg, starts, targs)
    for f in range(60):
        g.step(1.0 / 60)
        w.step(1.0 / 60)
        H.compare_worlds(g, w, f, check_corridors=(f % 5 == 0))
    gs, os_ = g.stats(), w.stats()
    for k in ("found", "not_found", "expansions", "discoveries", "corridor_nodes", "rounds", "admitted"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    assert gs["found"] > 50


def _world(seed, nx, ny, n_agents, T, Q, mode=G.MODE_PLAIN, budget=0, p=0.2, P=None, spw=0, wide=False, wg=0):
    tri, nbr = H.grid_map(seed, nx, ny, cell=4.0, p=p)
    cen = S.centroids(tri)
    lab = S.components(nbr)
    big = np.bincount(lab).argmax()
    nodes = np.nonzero(lab == big)[0]
    r = S.stream(seed, 3)
    perm = np.array(sorted(range(len(nodes)), key=lambda _: r.next_u64()))
    starts = [cen[nodes[perm[i]]] for i in range(n_agents)]  # distinct start nodes
    targs = [cen[nodes[r.randint(len(nodes))]] for _ in range(n_agents)]
    cfg = G.Config(agents_max_count=n_agents, path_finder_parallel_agents=P or n_agents,
                   path_finder_threads_per_agents=T, path_finder_queue_capacity=Q,
                   agent_radius=0.5, agent_speed=2.0, step_mode=mode, round_budget=budget or 1,
                   searches_per_wave=spw, search_wide_keys=wide, search_workgroups=wg)
    return tri, nbr, cfg, starts, targs


@pytest.mark.parametrize("mode,budget,P,spw,T,Q,wide,wg", [(G.MODE_PLAIN, 0, None, 0, 8, 16, False, 0),
                                                        (G.MODE_PLAIN, 0, 37, 0, 8, 16, False, 0),
                                                        (G.MODE_BATCHED, 3, None, 0, 8, 16, False, 0),
                                                        (G.MODE_BATCHED, 1, 50, 0, 8, 16, False, 0),
                                                        (G.MODE_BATCHED, 3, None, 3, 8, 16, False, 0),
                                                        # queues parked and restored past 16 items
                                                        (G.MODE_BATCHED, 2, None, 0, 4, 64, False, 0),
                                                        (G.MODE_BATCHED, 3, 60, 0, 2, 32, False, 0),
                                                        # keys without G and side, parked and restored
                                                        (G.MODE_BATCHED, 3, None, 0, 4, 16, True, 0),
                                                        (G.MODE_BATCHED, 2, 60, 0, 2, 32, True, 0),
                                                        (G.MODE_PLAIN, 0, 37, 0, 4, 16, True, 0),
                                                        # a slot per agent, 24 and 32 segments for 160
                                                        # agents: the segments take the slots by ticket
                                                        (G.MODE_BATCHED, 3, None, 0, 8, 16, False, 3),
                                                        (G.MODE_BATCHED, 2, None, 0, 4, 16, False, 2)])
def test_multi_frame_world_parity(oracle, mode, budget, P, spw, T, Q, wide, wg):
    """Agents search, move, avoid and progress for many frames; every frame
    must equal the oracle bit for bit (positions, states, nodes, corridors)."""
    tri, nbr, cfg, starts, targs = _world(21, 40, 40, 160, T, Q, mode, budget, P=P, spw=spw, wide=wide, wg=wg)
    print(f"world parity mode={mode} budget={budget} P={P} spw={spw} T={T} Q={Q} wide={wide} wg={wg}", flush=True)
    g = H.make_gpu(tri, nbr, cfg, starts, targs)
    w = H.make_oracle(oracle, tri, nbr, cfg, starts, targs)
    dt = 1.0 / 60
    hashed = 0
    for f in range(120):
        g.step(dt)
        w.step(dt)
        H.compare_worlds(g, w, f, check_corridors=(f % 10 == 0))
        hashed += H.compare_hash(g, w, f)
        if f == 60:  # re-target a few agents mid-run, including searching ones
            for i in range(0, 160, 7):
                if i % 2 == 0:  # two changes in one frame: the last one wins
                    x0, y0 = targs[(i * 5) % 160]
                    g.set_agent_targ_pos(i, x0, y0)
                    assert w.set_target(i, x0, y0) == 0
                x, y = targs[(i * 13) % 160]
                g.set_agent_targ_pos(i, x, y)
                assert w.set_target(i, x, y) == 0
    gs, os_ = g.stats(), w.stats()
    for k in ("found", "not_found", "expansions", "discoveries", "corridor_nodes", "rounds", "admitted"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    assert hashed > 100


@pytest.mark.parametrize("mode,budget", [(G.MODE_PLAIN, 0), (G.MODE_BATCHED, 2)])
def test_target_updates_made_during_a_step_apply_at_the_next(oracle, mode, budget):
    """SetAgentTargPos while a step runs takes effect at the next step
    (the reference loses it, README.md:91): the same changes made during
    GPU step f and after oracle frame f give the same world, bit for bit,
    including two changes of one agent (the last one wins) and changes on
    consecutive frames."""
    tri, nbr, cfg, starts, targs = _world(27, 36, 36, 128, 4, 16, mode, budget)
    print(f"late target updates mode={mode} budget={budget}", flush=True)
    g = H.make_gpu(tri, nbr, cfg, starts, targs)
    w = H.make_oracle(oracle, tri, nbr, cfg, starts, targs)
    dt = 1.0 / 60
    for f in range(80):
        changes = []
        if f in (20, 21, 45):
            for i in range(f % 3, 128, 5):
                if i % 2 == 0:
                    changes.append((i, *targs[(i * 7 + f) % 128]))
                changes.append((i, *targs[(i * 11 + f) % 128]))
        g.async_step(dt)
        for i, x, y in changes:  # while the GPU step runs
            g.set_agent_targ_pos(i, x, y)
        g.wait_step_end()
        w.step(dt)
        for i, x, y in changes:  # after the oracle frame
            assert w.set_target(i, x, y) == 0
        H.compare_worlds(g, w, f, check_corridors=(f % 10 == 0))
    gs, os_ = g.stats(), w.stats()
    assert gs["found"] > 0
    for k in ("found", "not_found", "expansions", "discoveries", "corridor_nodes", "rounds", "admitted"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    g.close()


@pytest.mark.parametrize("mode,budget", [(G.MODE_PLAIN, 0), (G.MODE_BATCHED, 2)])
def test_crowd_only_frames_and_back(oracle, mode, budget):
    """Once no agent searches and no target changes, a frame runs without
    admission and search (the moving set, hash, avoidance, progress and the
    readback only); a target change brings the full frame back.  Every frame
    equals the oracle bit for bit, through both transitions and through
    changes of the time step (every 25th frame is twice as long)."""
    tri, nbr, cfg, starts, targs = _world(41, 36, 36, 120, 4, 16, mode, budget)
    print(f"crowd-only frames mode={mode} budget={budget}", flush=True)
    g = H.make_gpu(tri, nbr, cfg, starts, targs)
    w = H.make_oracle(oracle, tri, nbr, cfg, starts, targs)
    seen_crowd_only = 0
    for f in range(160):
        dt = 1.0 / 30 if f % 25 == 0 else 1.0 / 60
        before = g.frame_kinds()["crowd_only"]
        g.step(dt)
        w.step(dt)
        H.compare_worlds(g, w, f, check_corridors=(f % 10 == 0))
        if g.frame_kinds()["crowd_only"] > before:
            seen_crowd_only += 1
        if f == 100:  # searches again: the next frame is a full one
            assert seen_crowd_only > 0
            k0 = g.frame_kinds()["crowd_only"]
            for i in range(0, 120, 9):
                x, y = targs[(i * 7) % 120]
                g.set_agent_targ_pos(i, x, y)
                assert w.set_target(i, x, y) == 0
        if f == 101:
            assert g.frame_kinds()["crowd_only"] == k0  # frame 101 admitted and searched
    gs, os_ = g.stats(), w.stats()
    for k in ("found", "not_found", "expansions", "discoveries", "corridor_nodes", "rounds", "admitted"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    assert gs["admitted"] > 120
    g.close()


def test_determinism_and_batched_equals_plain_corridors(oracle):
    """Same inputs twice -> identical bits; batched form gives the same
    corridors as the plain form for the same requests."""
    tri, nbr, cfg, starts, targs = _world(33, 32, 32, 96, 4, 16)
    runs = []
    for mode, frames in ((G.MODE_PLAIN, 40), (G.MODE_PLAIN, 40), (G.MODE_BATCHED, 400)):
        cfg.step_mode = mode
        cfg.round_budget = 2
        g = H.make_gpu(tri, nbr, cfg, starts, targs)
        for _ in range(frames):
            g.step(1.0 / 60)
        xy, st = g.agents()
        assert not (st == G.SEARCH).any()  # every frame-0 request has ended
        cors = [g.corridor(i) for i in range(len(starts))]
        runs.append((xy.copy(), st.copy(), cors))
        g.close()
    assert np.array_equal(runs[0][0].view(np.uint32), runs[1][0].view(np.uint32))
    assert np.array_equal(runs[0][1], runs[1][1])
    for a, b in zip(runs[0][2], runs[2][2]):
        assert (a is None) == (b is None)
        if a is not None:
            assert np.array_equal(a, b)


def test_search_timing_marks_only_timed_frames_and_changes_no_result(oracle):
    """search_timing k: frames 0, k, 2k, ... report the search launch's HIP
    event time, the others -1; the event nodes in the frame graph leave every
    result bit-identical to a run without them."""
    tri, nbr, cfg, starts, targs = _world(36, 32, 32, 96, 4, 16)
    cfg.step_mode = G.MODE_BATCHED
    cfg.round_budget = 2
    runs = []
    for k in (0, 3):
        cfg.search_timing = k
        g = H.make_gpu(tri, nbr, cfg, starts, targs)
        timed = []
        for f in range(40):
            g.step(1.0 / 60)
            timed.append(g.last_frame_ms()[2])
        xy, st = g.agents()
        runs.append((xy.copy(), st.copy(), [g.corridor(i) for i in range(len(starts))], g.stats()))
        g.close()
        if k == 0:
            assert all(t == -1.0 for t in timed)
        else:
            assert all((t > 0.0) == (f % k == 0) for f, t in enumerate(timed)), timed
            assert all(t == -1.0 for f, t in enumerate(timed) if f % k)
    assert np.array_equal(runs[0][0].view(np.uint32), runs[1][0].view(np.uint32))
    assert np.array_equal(runs[0][1], runs[1][1])
    for a, b in zip(runs[0][2], runs[1][2]):
        assert (a is None) == (b is None) and (a is None or np.array_equal(a, b))
    assert runs[0][3] == runs[1][3]


def test_hash_groups_agents_by_cell(oracle):
    """The counting sort lists every hashed agent once, grouped by cell in
    ascending key order and by ascending agent index inside a cell."""
    tri, nbr, cfg, starts, targs = _world(44, 30, 30, 400, 4, 16, p=0.1)
    g = H.make_gpu(tri, nbr, cfg, starts, targs)
    for _ in range(5):
        g.step(1.0 / 60)
    keys, order = g.hash_state()
    valid = np.nonzero(keys >= 0)[0]
    assert len(order) == len(valid) and np.array_equal(np.sort(order), valid)
    assert np.array_equal(order, valid[np.argsort(keys[valid], kind="stable")])
    assert len(np.unique(keys[valid])) > 10


def test_device_locate_matches_host_and_oracle(oracle):
    tri, nbr = H.grid_map(55, 30, 20, cell=3.0, p=0.3)
    cfg = G.Config(agents_max_count=1, path_finder_parallel_agents=1)
    g = H.make_gpu(tri, nbr, cfg, [S.centroids(tri)[0]])
    g.manual_start()
    r = S.stream(9, 1)
    pts = np.array([[r.uniform() * 96 - 3, r.uniform() * 66 - 3] for _ in range(20000)], np.float32)
    # grid vertices and edge midpoints: boundary-inclusive cases
    gx, gy = np.meshgrid(np.arange(0, 91, 1.5, dtype=np.float32), np.arange(0, 61, 1.5, dtype=np.float32))
    pts = np.concatenate([pts, np.stack([gx.ravel(), gy.ravel()], 1)]).astype(np.float32)
    dev = g.locate(pts, device=True)
    host = g.locate(pts)
    brute = oracle.OracleMap(tri, nbr).locate(pts, grid=False)
    assert np.array_equal(host, brute)
    assert np.array_equal(dev, brute)


@pytest.mark.parametrize("mode,budget", [(G.MODE_PLAIN, 0), (G.MODE_BATCHED, 2)])
def test_one_slot_many_searches_mark_wrap(oracle, mode, budget):
    """400 consecutive searches through one slot: the 7-bit visited-mark
    generation wraps three times (the slot's marks are cleared on each wrap);
    every corridor still equals the oracle's search from the agent's node."""
    tri, nbr = H.grid_map(17, 20, 16, cell=4.0, p=0.2)
    cen = S.centroids(tri)
    lab = S.components(nbr)
    big = np.bincount(lab).argmax()
    nodes = np.nonzero(lab == big)[0]
    r = S.stream(23, 5)
    cfg = G.Config(agents_max_count=3, path_finder_parallel_agents=1, path_finder_threads_per_agents=4,
                   path_finder_queue_capacity=8, agent_radius=0.25, agent_speed=1.0, step_mode=mode,
                   round_budget=max(1, budget))
    g = H.make_gpu(tri, nbr, cfg, [cen[nodes[k]] for k in range(3)])
    om = oracle.OracleMap(tri, nbr)
    print(f"mark wrap mode={mode} V={len(tri)}", flush=True)
    searches = found = 0
    while searches < 400:
        a = searches % 3
        _, _, cur = g.agents(with_nodes=True)
        tg = int(nodes[r.randint(len(nodes))])
        if tg == cur[a]:
            continue
        g.set_agent_targ_pos(a, *cen[tg])
        for _ in range(200):
            g.step(1.0 / 60)
            if g.get_agent_state(a) != G.SEARCH:
                break
        st = g.get_agent_state(a)
        assert st != G.SEARCH
        ok, cor, _ = om.search(int(cur[a]), tg, 4, 8, 1.0)
        assert (st == G.MOVE) == ok, f"search {searches}"
        if ok:
            found += 1
            assert np.array_equal(g.corridor(a), cor), f"search {searches}"
        searches += 1
    assert found > 300


@pytest.mark.parametrize("spread", [0.5, 20.0])
def test_avoidance_dense_cluster_equals_oracle(oracle, spread):
    """Agents packed within `spread` units of each other: at 0.5 each agent
    has more neighbours within cd than the per-thread neighbour list holds,
    so the avoidance pass takes them by rescans; at 20 the list serves.
    Positions, states and nodes must equal the oracle bit for bit."""
    tri, nbr = H.grid_map(3, 40, 40, cell=4.0, p=0.0)
    r = np.random.RandomState(5)
    n = 48
    starts = (np.array([80.0, 80.0]) + r.uniform(-spread, spread, (n, 2))).astype(np.float32)
    targs = (np.array([20.0, 140.0]) + r.uniform(-3, 3, (n, 2))).astype(np.float32)
    cfg = G.Config(agents_max_count=n, path_finder_parallel_agents=n, path_finder_threads_per_agents=4,
                   path_finder_queue_capacity=16, agent_radius=2.0, agent_speed=1.0)
    print(f"dense cluster spread={spread} agents={n}", flush=True)
    g = H.make_gpu(tri, nbr, cfg, starts, targs)
    w = H.make_oracle(oracle, tri, nbr, cfg, starts, targs)
    hashed = 0
    for f in range(40):
        g.step(1.0 / 60)
        w.step(1.0 / 60)
        H.compare_worlds(g, w, f, check_corridors=(f % 10 == 0))
        hashed += H.compare_hash(g, w, f)
    assert hashed > 30
    g.close()


# ------------------------------------------------------- request admission (F1)
def test_admission_defers_in_ascending_index_and_same_node_goes_idle(oracle):
    """The closed forms the oracle is pinned by, through the C ABI (plain
    form, P < N: the ranked admission path)."""
    print("admission closed form: deferral", flush=True)
    H.check_deferral_in_ascending_index(oracle, gpu=True)


@pytest.mark.parametrize("budget", [1, 2, 5])
def test_admission_fills_freed_slots_in_index_order(oracle, budget):
    print(f"admission closed form: freed slots budget={budget}", flush=True)
    H.check_slots_free_in_index_order(oracle, True, budget)


def test_retarget_cancels_search_in_flight(oracle):
    print("admission closed form: re-target cancels", flush=True)
    H.check_retarget_cancels_search_in_flight(oracle, gpu=True)


# ------------------------------------------------------- boundary
def _hip_current_device():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    d = ctypes.c_int(-1)
    assert hip.hipGetDevice(ctypes.byref(d)) == 0
    return d.value


def test_entry_points_leave_the_callers_device_current():
    """Every entry point that uses the device restores the caller's current
    device on return (the reference pushes and pops its context around each
    step, GMAPF.cu:82,127,266,308).  A one-GPU box can only show that the
    device is unchanged across start, steps, readbacks and destroy."""
    tri, nbr, cfg, starts, targs = _world(39, 20, 20, 32, 4, 16)
    before = _hip_current_device()
    g = H.make_gpu(tri, nbr, cfg, starts, targs)
    g.manual_start()
    assert _hip_current_device() == before
    for f in range(6):
        g.async_step(1.0 / 60)
        assert _hip_current_device() == before
        g.wait_step_end()
        assert _hip_current_device() == before
        g.agents(with_nodes=True)
        g.corridor(0)
        g.hash_state()
        assert _hip_current_device() == before
    g.locate(np.float32([starts[0]]), device=True)
    g.close()
    assert _hip_current_device() == before
