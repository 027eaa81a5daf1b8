"""Frame-level parity at the benchmark configurations, on navmeshes built by
FillMapStart/FillMapObstacle/FillMapEnd: every frame the HIP step must equal
the oracle bit for bit (positions, states, CurrNodeIdx; corridors sampled).
Tolerance stated by the task: 1e-4 map units for positions; the check below
is exact equality, which is stricter."""
import numpy as np
import pytest

import gmapf_amd as G
from gmapf_amd import scenarios as S
from tests import helpers as H

pytestmark = pytest.mark.gpu


def locator():
    keep = []

    def make(W, Hh, obs):
        g = G.GMAPF()
        g.init(G.Config(agents_max_count=1))
        g.fill_map_start(W, Hh)
        for o in obs:
            g.fill_map_obstacle(o)
        g.fill_map_end()
        keep.append(g)
        return lambda p: g.locate(p)

    return make, keep


def run_pair(oracle, sc, mode=G.MODE_PLAIN, budget=1, frames=None, check_every=1, hash_cell=0.0):
    frames = frames or sc.frames
    cfg = G.Config(agents_max_count=sc.n, agent_radius=sc.radius, agent_speed=sc.speed,
                   path_finder_parallel_agents=sc.P, path_finder_threads_per_agents=sc.T,
                   path_finder_queue_capacity=sc.Q, path_finder_heuristic_k=sc.K, step_mode=mode,
                   round_budget=budget, hash_cell=hash_cell)
    g = G.GMAPF()
    g.init(cfg)
    sc.build(g)
    for p in sc.starts:
        g.add_agent(*p)
    g.set_agent_targets(np.arange(sc.n, dtype=np.int32), sc.targets)
    tri, nbr, _ = g.map_arrays()
    w = H.make_oracle(oracle, tri, nbr, cfg, sc.starts, sc.targets)
    for f in range(frames):
        g.step(sc.dt)
        w.step(sc.dt)
        if f % check_every == 0 or f == frames - 1:
            H.compare_worlds(g, w, f, check_corridors=(f % (10 * check_every) == 0))
    gs, os_ = g.stats(), w.stats()
    for k in ("found", "not_found", "expansions", "discoveries", "corridor_nodes", "rounds", "admitted"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    xy, st = g.agents()
    g.close()
    return gs, xy, st


@pytest.mark.parametrize("mode,budget", [(G.MODE_PLAIN, 1), (G.MODE_BATCHED, 2)])
def test_config1_bowtie(oracle, mode, budget):
    make, keep = locator()
    sc = S.config1(make)
    print(f"config1 mode={mode}", flush=True)
    gs, xy, st = run_pair(oracle, sc, mode, budget)
    assert gs["found"] > 0


def test_config2_rects_full(oracle):
    make, keep = locator()
    sc = S.config2(make)
    print(f"config2 {sc.n} agents", flush=True)
    gs, xy, st = run_pair(oracle, sc, frames=300, check_every=10)
    assert gs["found"] > 0.9 * sc.n


def test_config4_crowd_scaled(oracle):
    """Two opposing streams; 8192 agents (the full 65,536 run is the
    determinism test below and the crowd benchmark)."""
    make, keep = locator()
    sc = S.config4(make, n_agents=8192)
    print(f"config4 scaled {sc.n} agents", flush=True)
    gs, xy, st = run_pair(oracle, sc, frames=150, check_every=15, hash_cell=4.0)
    assert gs["found"] > 0.9 * sc.n


def test_config4_full_deterministic():
    """65,536 agents, 120 frames, twice: bitwise identical, no NaN."""
    make, keep = locator()
    sc = S.config4(make)
    outs = []
    for rep in range(2):
        cfg = G.Config(agents_max_count=sc.n, agent_radius=sc.radius, agent_speed=sc.speed,
                       path_finder_parallel_agents=sc.P, path_finder_threads_per_agents=sc.T,
                       path_finder_queue_capacity=sc.Q, hash_cell=4.0)
        g = G.GMAPF()
        g.init(cfg)
        sc.build(g)
        for p in sc.starts:
            g.add_agent(*p)
        g.set_agent_targets(np.arange(sc.n, dtype=np.int32), sc.targets)
        print(f"config4 full rep {rep}", flush=True)
        for f in range(120):
            g.step(sc.dt)
        xy, st = g.agents()
        outs.append((xy.copy(), st.copy(), g.stats()))
        g.close()
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1], outs[1][1])
    assert not np.isnan(outs[0][0]).any()
    assert outs[0][2] == outs[1][2]
