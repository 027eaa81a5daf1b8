"""An engine whose map comes from a "GMAPFNAV" fixture steps exactly like the
engine that triangulated the map (FillMapStart/FillMapObstacle/FillMapEnd,
src/GMAPF.cu:133-217) and like the oracle: same positions, states and
corridors every frame."""
import numpy as np
import pytest

import gmapf_amd as G
from gmapf_amd import scenarios as S
from tests import helpers as H
from tests.test_gpu_configs import locator

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,budget", [(G.MODE_PLAIN, 1), (G.MODE_BATCHED, 3)])
def test_loaded_map_steps_like_the_triangulated_one(oracle, mode, budget):
    make, keep = locator()
    sc = S.config2(make, n_agents=512, n_rects=120, size=500.0)
    cfg = G.Config(agents_max_count=sc.n, agent_radius=sc.radius, agent_speed=sc.speed,
                   path_finder_parallel_agents=sc.P, path_finder_threads_per_agents=sc.T,
                   path_finder_queue_capacity=sc.Q, path_finder_heuristic_k=sc.K, step_mode=mode,
                   round_budget=budget)
    a = G.GMAPF()
    a.init(cfg)
    sc.build(a)
    data = a.nav_bytes()
    b = G.GMAPF()
    b.init(cfg)
    b.fill_map_from_nav(data)
    ids = np.arange(sc.n, dtype=np.int32)
    for g in (a, b):
        for p in sc.starts:
            g.add_agent(*p)
        g.set_agent_targets(ids, sc.targets)
    tri, nbr, _ = b.map_arrays()
    w = H.make_oracle(oracle, tri, nbr, cfg, sc.starts, sc.targets)
    print(f"nav parity mode={mode} V={len(tri)} agents={sc.n}", flush=True)
    for f in range(60):
        a.step(sc.dt)
        b.step(sc.dt)
        w.step(sc.dt)
        xa, sa = a.agents()
        xb, sb = b.agents()
        assert np.array_equal(sa, sb) and np.array_equal(xa.view(np.uint32), xb.view(np.uint32)), f
        H.compare_worlds(b, w, f, check_corridors=(f % 10 == 0))
        if f % 10 == 0:
            for i in range(0, sc.n, 7):
                ca, cb = a.corridor(i), b.corridor(i)
                assert (ca is None) == (cb is None) and (ca is None or np.array_equal(ca, cb)), (f, i)
    assert b.stats()["found"] == a.stats()["found"] > 0
    a.close()
    b.close()
