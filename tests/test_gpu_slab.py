"""Corridor slab: with a slab too small for every corridor at once, found
searches wait in their slots (status 4) and compaction releases the
corridors of agents that stopped moving.  Every agent still receives the
corridor the unconstrained engine gives it and the oracle's search
(PathFinder.h:276-306 layout), only later."""
import numpy as np
import pytest

import gmapf_amd as G
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _run(tri, nbr, starts, targets, mode, slab, frames=800, P=0, wg=0):
    cfg = G.Config(agents_max_count=len(starts), agent_radius=0.5, agent_speed=240.0,
                   path_finder_parallel_agents=P or len(starts), path_finder_threads_per_agents=4,
                   path_finder_queue_capacity=16, step_mode=mode, round_budget=4, corridor_slab=slab,
                   search_workgroups=wg)
    g = H.make_gpu(tri, nbr, cfg, starts, targets)
    first = {}  # agent -> corridor when it first moved
    seen_move_frame = {}
    for f in range(frames):
        g.step(1.0 / 60)
        st = g.agent_states()
        for i in np.nonzero(st == G.MOVE)[0]:
            i = int(i)
            if i not in first:
                first[i] = g.corridor(i)
                seen_move_frame[i] = f
        if not (st == G.SEARCH).any():
            break
    # after compactions an agent that stopped moving has no corridor; one that
    # still moves keeps the one it was given (never another buffer's words)
    for i in range(len(starts)):
        c = g.corridor(i)
        assert c is None or (i in first and np.array_equal(c, first[i])), i
    info = g.slab_info()
    stats = g.stats()
    g.close()
    return first, seen_move_frame, info, stats


@pytest.mark.parametrize("mode,P,wg", [(G.MODE_PLAIN, 0, 0), (G.MODE_BATCHED, 0, 0), (G.MODE_BATCHED, 60, 2),
                                      (G.MODE_BATCHED, 0, 2)])
def test_small_slab_defers_and_compacts_without_changing_corridors(oracle, mode, P, wg):
    """P, wg: fewer slots than agents and a persistent grid of wg workgroups
    (32 segments for 60 slots), so the workgroups take searches by atomic
    ticket and a segment whose deferred corridor still does not fit takes
    the next ticket.  P = 0 with wg: a slot per agent and 32 segments for
    160 slots, taken by ticket too."""
    tri, nbr = H.grid_map(11, 24, 24, cell=4.0, p=0.15)
    pairs = H.connected_pairs(12, nbr, 160)
    cen = tri.reshape(-1, 3, 2).mean(axis=1).astype(np.float32)
    starts = np.array([cen[a] for a, _ in pairs], np.float32)
    targets = np.array([cen[b] for _, b in pairs], np.float32)
    V = len(tri)
    print(f"slab test mode={mode} P={P} wg={wg} V={V} agents={len(pairs)}", flush=True)
    ref, ref_f, ref_info, ref_stats = _run(tri, nbr, starts, targets, mode, 0, P=P, wg=wg)
    assert ref_info["capacity"] >= len(pairs) * V or ref_info["capacity"] >= len(pairs) * 1024
    assert ref_info["deferred"] == 0 and ref_info["compactions"] == 0
    small, small_f, info, stats = _run(tri, nbr, starts, targets, mode, 2 * V, P=P, wg=wg)
    assert info["capacity"] == 2 * V
    assert info["deferred"] > 0, info
    assert info["compactions"] > 0, info
    assert info["in_use"] <= info["capacity"]
    assert small.keys() == ref.keys() and len(ref) > 0.5 * len(pairs)
    assert stats["found"] == ref_stats["found"] == len(ref)
    om = oracle.OracleMap(tri, nbr)
    for i, c in ref.items():
        assert c is not None and np.array_equal(small[i], c), i
        a, b = pairs[i]
        found, oc, _ = om.search(a, b, 4, 16)
        assert found and np.array_equal(oc, c), i
    # the small slab delays some agents, never makes one earlier
    assert all(small_f[i] >= ref_f[i] for i in ref) and any(small_f[i] > ref_f[i] for i in ref)
