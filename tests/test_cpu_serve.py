"""Serving loop (include/gmapf_serve.h) without a GPU: argument checks, the
footprint formula and the admission plan (slots, one request per agent,
memory budget with FIFO stop, requests adopted from before start)."""
import numpy as np
import pytest

import gmapf_amd as G
from gmapf_amd.serving import Server
from tests import helpers as H


def _engine(P=4, T=8, Q=32, n_agents=16, targets=False, seed=3):
    tri, nbr = H.grid_map(seed, 12, 12, cell=4.0, p=0.0)
    cen = tri.reshape(-1, 3, 2).mean(axis=1)
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=64, path_finder_parallel_agents=P, path_finder_threads_per_agents=T,
                    path_finder_queue_capacity=Q, step_mode=G.MODE_BATCHED, round_budget=8))
    g.fill_map_manually(tri, nbr)
    for i in range(n_agents):
        g.add_agent(*cen[i])
    if targets:
        g.set_agent_targ_pos(0, *cen[-1])
        g.set_agent_targ_pos(1, *cen[-2])
    return g, cen, len(tri)


def test_create_rejects_bad_maps():
    g, _, _ = _engine()
    with pytest.raises(G.GmapfError):
        Server([])
    with pytest.raises(G.GmapfError):
        Server([g, g])
    empty = G.GMAPF()
    empty.init(G.Config())
    with pytest.raises(G.GmapfError):
        Server([empty])


def test_submit_checks_and_tickets():
    g, cen, _ = _engine()
    s = Server([g])
    assert s.submit(0, 0, *cen[5]) == 0
    assert s.submit(0, 1, *cen[6]) == 1
    with pytest.raises(G.GmapfError) as e:
        s.submit(1, 0, *cen[5])
    assert e.value.code == G_E_ARG
    with pytest.raises(G.GmapfError) as e:
        s.submit(0, 99, *cen[5])
    assert e.value.code == G_E_ARG
    with pytest.raises(G.GmapfError) as e:
        s.submit(0, 0, -100.0, -100.0)
    assert e.value.code == G_E_POS
    t, st = s.submit_many(0, [2, 3, 4], np.stack([cen[1], [-5.0, -5.0], cen[2]]))
    assert list(st) == [0, G_E_POS, 0] and list(t) == [2, -1, 3]
    assert s.stats()["pending"] == 4 and s.stats()["submitted"] == 4


G_E_ARG, G_E_POS, G_E_FULL = -1, -3, -7


def test_footprint_formula():
    g, _, V = _engine(T=16, Q=64)
    s = Server([g])
    assert s.footprint(0) == 16 * 64 * 8 + 16 * 4 + (V + 63) // 64 * 64 + V * 8


def test_plan_respects_slots_and_serialises_agents():
    a, cen, _ = _engine(P=2)
    b, cenb, _ = _engine(P=3, seed=4)
    s = Server([a, b])
    # map 0: agents 0,1,0(again),2 ; map 1: agents 5,6
    s.submit(0, 0, *cen[10])   # t0
    s.submit(0, 1, *cen[11])   # t1
    s.submit(0, 0, *cen[12])   # t2: agent 0 already taken in this scan
    s.submit(0, 2, *cen[13])   # t3: map 0 has only 2 slots
    s.submit(1, 5, *cenb[10])  # t4
    s.submit(1, 6, *cenb[11])  # t5
    assert list(s.plan()) == [0, 1, 4, 5]
    # planning has no side effects
    assert list(s.plan()) == [0, 1, 4, 5]
    assert s.stats()["pending"] == 6 and s.stats()["admitted"] == 0


def test_plan_memory_budget_stops_in_fifo_order():
    big, cen, _ = _engine(P=8)
    small, cens, _ = _engine(P=8, T=2, Q=1, seed=5)
    probe = Server([big, small])
    fb, fs = probe.footprint(0), probe.footprint(1)
    probe.close()
    s = Server([big, small], queue_memory_budget=2 * fb + fs)
    s.submit(1, 0, *cens[3])   # t0 small
    s.submit(0, 0, *cen[3])    # t1 big
    s.submit(0, 1, *cen[4])    # t2 big
    s.submit(0, 2, *cen[5])    # t3 big: over budget -> stop
    s.submit(1, 1, *cens[4])   # t4 small: behind the blocked request, not admitted
    assert list(s.plan()) == [0, 1, 2]
    with pytest.raises(G.GmapfError) as e:  # one request larger than the whole budget
        Server([big], queue_memory_budget=fb - 1).submit(0, 3, *cen[6])
    assert e.value.code == G_E_ARG


def test_pending_cap():
    g, cen, _ = _engine()
    s = Server([g], max_pending=2)
    s.submit(0, 0, *cen[1])
    s.submit(0, 1, *cen[2])
    with pytest.raises(G.GmapfError) as e:
        s.submit(0, 2, *cen[3])
    assert e.value.code == G_E_FULL
    assert s.stats()["rejected"] == 1


def test_agents_searching_before_start_hold_slots():
    g, cen, _ = _engine(P=3, targets=True)  # agents 0 and 1 are in Search
    s = Server([g])
    s.submit(0, 0, *cen[20])  # t0: agent 0 busy (adopted)
    s.submit(0, 4, *cen[21])  # t1: the one free slot
    s.submit(0, 5, *cen[22])  # t2: no slot left
    assert list(s.plan()) == [1]
    st = s.stats()
    assert st["in_flight"] == 2 and st["live_bytes"] == 2 * s.footprint(0)


def _grid_engine(n_cells, P, n_agents, seed=3):
    tri, nbr = H.grid_map(seed, n_cells, n_cells, cell=4.0, p=0.0)
    cen = tri.reshape(-1, 3, 2).mean(axis=1)
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=n_agents, path_finder_parallel_agents=P, path_finder_threads_per_agents=8,
                    path_finder_queue_capacity=32, step_mode=G.MODE_BATCHED, round_budget=8))
    g.fill_map_manually(tri, nbr)
    for i in range(n_agents):
        g.add_agent(*cen[i])
    return g, cen


def _flood(policy, budget_quanta=3, n_small=40, n_large=3):
    from gmapf_amd.serving import ADMIT_FAIR  # noqa: F401
    small, cs = _grid_engine(12, P=64, n_agents=64)
    large, cl = _grid_engine(40, P=8, n_agents=8)
    probe = Server([small, large])
    fs, fl = probe.footprint(0), probe.footprint(1)
    probe.close()
    s = Server([small, large], queue_memory_budget=budget_quanta * fl if budget_quanta else 0,
               admission_policy=policy)
    small_t = [s.submit(0, a, *cs[-1 - a]) for a in range(n_small)]  # the burst arrives first
    large_t = [s.submit(1, a, *cl[-1 - a]) for a in range(n_large)]
    return s, fs, fl, small_t, large_t, (small, large)


def test_fair_policy_admits_other_maps_during_a_burst():
    from gmapf_amd.serving import ADMIT_FAIR, ADMIT_FIFO
    s, fs, fl, small_t, large_t, keep = _flood(ADMIT_FIFO)
    fifo = set(s.plan().tolist())
    assert fifo and fifo <= set(small_t)  # the budget is spent on the burst, map 1 waits
    s, fs, fl, small_t, large_t, keep = _flood(ADMIT_FAIR)
    fair = s.plan().tolist()
    assert fair == s.plan().tolist()  # a plan changes no state
    assert set(fair) & set(large_t) == {large_t[0]}  # one quantum: one large request this frame
    # deficit round robin, one request per turn, from map 0: map 0, map 1,
    # then map 0 until its quantum is spent; the budget left goes to the
    # burst in arrival order
    q = max(fs, fl)
    n0 = q // fs
    assert fair[:2] == [small_t[0], large_t[0]] and fair[2:n0 + 1] == small_t[1:n0]
    extra = (3 * fl - (n0 * fs + fl)) // fs
    assert fair[n0 + 1:] == small_t[n0:n0 + extra]
    assert sum(fs if t in small_t else fl for t in fair) <= 3 * fl


def test_fair_policy_without_budget_admits_what_fifo_admits():
    from gmapf_amd.serving import ADMIT_FAIR, ADMIT_FIFO
    s, _, _, small_t, large_t, keep = _flood(ADMIT_FIFO, budget_quanta=0)
    fifo = set(s.plan().tolist())
    s, _, _, small_t, large_t, keep = _flood(ADMIT_FAIR, budget_quanta=0)
    fair = set(s.plan().tolist())
    assert fair == fifo == set(small_t) | set(large_t)


def test_fair_policy_keeps_agents_serialised_and_slots_bounded():
    from gmapf_amd.serving import ADMIT_FAIR
    g, cen = _grid_engine(12, P=3, n_agents=8)
    s = Server([g], admission_policy=ADMIT_FAIR)
    t = [s.submit(0, 0, *cen[20]), s.submit(0, 0, *cen[21]), s.submit(0, 1, *cen[22]),
         s.submit(0, 2, *cen[23]), s.submit(0, 3, *cen[24])]
    assert s.plan().tolist() == [t[0], t[2], t[3]]  # agent 0 once; 3 slots


def test_unknown_admission_policy_is_rejected():
    g, _ = _grid_engine(12, P=3, n_agents=2)
    with pytest.raises(G.GmapfError):
        Server([g], admission_policy=5)
