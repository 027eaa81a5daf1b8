"""Serving loop on the GPU: every request's result in the mix (several maps,
continuous Poisson arrivals, searches spanning frames and overlapping with
new admissions under a memory budget) equals its result alone on the device
and the oracle's search, and the server's counters stay consistent."""
import numpy as np
import pytest

import gmapf_amd as G
from gmapf_amd.runner import Locators
from gmapf_amd.serving import ADMIT_FAIR, ADMIT_FIFO, Server, build_map, serving_maps

pytestmark = pytest.mark.gpu

T, Q = 8, 32


def _maps(locs):
    # two small rect maps at the serving workload's obstacle density
    return serving_maps(locs, seed=11, sizes=(160.0, 320.0), agents=(200, 400), slots=(24, 48),
                        target_pool=4096)


def _drive(srv, smaps, engines, frames, lam, seed, corridors):
    rng = np.random.Generator(np.random.PCG64(seed))
    done_all = []
    for f in range(frames):
        k = int(rng.poisson(lam)) if lam else 0
        if k:
            mi = rng.integers(0, len(smaps), size=k).astype(np.int32)
            ag = np.array([rng.integers(0, smaps[m].n_agents) for m in mi], dtype=np.int32)
            xy = np.stack([smaps[m].targets[rng.integers(0, len(smaps[m].targets))] for m in mi])
            _, st = srv.submit_many(mi, ag, xy)
            assert (st == 0).all()
            for m, a, p in zip(mi, ag, xy):
                corridors.setdefault("targets", {}).setdefault((int(m), int(a)), []).append(p)
        if f % 50 == 0:
            print(f"serve frame {f} submitted {k}", flush=True)
        srv.step(1.0 / 60)
        done = srv.poll()
        for d in done:
            if d["found"]:
                c = engines[d["map"]].corridor(int(d["agent"]))
                corridors[int(d["ticket"])] = None if c is None else c.copy()
        done_all.append(done)
        s = srv.stats()
        assert s["submitted"] == s["admitted"] + s["pending"] + s["rejected"]
        assert s["completed"] == s["found"] + s["no_path"]
        assert s["admitted"] == s["completed"] + s["in_flight"]
        if not lam and s["pending"] == 0 and s["in_flight"] == 0:
            break
    return np.concatenate(done_all)


def _check_alone(oracle, engines, done, pick, log):
    """Each picked request alone -- one at a time on a fresh engine of the same
    map, agent placed on the centroid of the request's start node -- gives the
    found flag and corridor it got in the mix, and both equal the oracle's
    search.  Returns the number checked."""
    by_map = {}
    for i in pick:
        by_map.setdefault(int(done[i]["map"]), []).append(done[i])
    # ticket -> target position, in submission order per (map, agent)
    tickets_by_key = {}
    for d in sorted(done, key=lambda d: d["ticket"]):
        tickets_by_key.setdefault((int(d["map"]), int(d["agent"])), []).append(int(d["ticket"]))
    target_of = {}
    for key, ts in tickets_by_key.items():
        for t, p in zip(ts, log["targets"][key]):
            target_of[t] = p
    checked = 0
    for m, reqs in by_map.items():
        tri, nbr, cen = engines[m].map_arrays()
        om = oracle.OracleMap(tri, nbr)
        starts = np.array([cen[int(d["start_node"])] for d in reqs], dtype=np.float32)
        alone = G.GMAPF()
        alone.init(G.Config(agents_max_count=len(reqs), agent_radius=2.0, agent_speed=1.0,
                            path_finder_parallel_agents=1, path_finder_threads_per_agents=T,
                            path_finder_queue_capacity=Q, step_mode=G.MODE_PLAIN))
        alone.fill_map_manually(tri, nbr)
        for x, y in starts:
            alone.add_agent(x, y)
        _, _, nodes = alone.agents(with_nodes=True)
        assert list(nodes) == [int(d["start_node"]) for d in reqs]
        alone.manual_start()
        print(f"alone map {m}: {len(reqs)} tickets", flush=True)
        for k, d in enumerate(reqs):
            tx, ty = target_of[int(d["ticket"])]
            alone.set_agent_targ_pos(k, tx, ty)
            for _ in range(10):
                alone.step(1.0 / 60)
                if alone.get_agent_state(k) != G.SEARCH:
                    break
            assert alone.get_agent_state(k) != G.SEARCH
            mix = log.get(int(d["ticket"]))
            got = alone.corridor(k) if alone.get_agent_state(k) == G.MOVE else None
            assert (got is not None) == bool(d["found"])
            tg = int(om.locate(np.array([[tx, ty]], np.float32))[0])
            S_node = int(d["start_node"])
            if S_node != tg:
                of, oc, _ = om.search(S_node, tg, T, Q)
                assert of == bool(d["found"])
                if of:
                    np.testing.assert_array_equal(oc, got)
            if d["found"]:
                np.testing.assert_array_equal(mix, got)
            checked += 1
        alone.close()
    return checked


@pytest.mark.parametrize("policy", [ADMIT_FIFO, ADMIT_FAIR])
def test_results_alone_equal_results_in_the_mix(oracle, policy):
    locs = Locators()
    smaps = _maps(locs)
    engines = [build_map(sm, T, Q, round_budget=3) for sm in smaps]
    fb = [(e.launch_info()["V"] + 63) // 64 * 64 + e.launch_info()["V"] * 8 + T * Q * 8 + T * 4 for e in engines]
    budget = 12 * fb[1] + 12 * fb[0]  # fewer than the slots: the budget binds
    srv = Server(engines, queue_memory_budget=budget, record_start_nodes=True, admission_policy=policy)
    assert [srv.footprint(m) for m in range(2)] == fb
    log = {}
    # requests keep arriving while earlier ones are searching (round budget 3)
    done = _drive(srv, smaps, engines, 40, 12.0, 5, log)
    # drain: no more arrivals
    more = _drive(srv, smaps, engines, 5000, 0, 6, log)
    done = np.concatenate([done, more])
    st = srv.stats()
    assert st["pending"] == 0 and st["in_flight"] == 0, st
    assert st["completed"] == st["submitted"] == len(done)
    assert st["peak_live_bytes"] <= budget
    assert (done["admit_frame"] > done["submit_frame"]).any(), "the budget never held a request back"
    assert (done["done_frame"] > done["admit_frame"]).any(), "no search spanned frames"
    assert done["found"].sum() > 0.5 * len(done)

    rng = np.random.Generator(np.random.PCG64(9))
    pick = rng.choice(len(done), size=min(40, len(done)), replace=False)
    checked = _check_alone(oracle, engines, done, pick, log)
    assert checked == len(pick)
    srv.close()
    for e in engines:
        e.close()


def test_three_maps_overlap_and_budget_bounds_concurrency():
    locs = Locators()
    smaps = serving_maps(locs, seed=2, sizes=(128.0, 256.0, 512.0), agents=(128, 512, 1024),
                         slots=(64, 128, 128), target_pool=4096)
    engines = [build_map(sm, 8, 16, round_budget=32) for sm in smaps]
    probe = Server(engines)
    fp = [probe.footprint(m) for m in range(3)]
    probe.close()
    budget = 100 * fp[2]
    srv = Server(engines, queue_memory_budget=budget)
    log = {}
    _drive(srv, smaps, engines, 60, 20.0, 3, log)
    _drive(srv, smaps, engines, 5000, 0, 4, log)
    st = srv.stats()
    assert st["pending"] == 0 and st["in_flight"] == 0
    assert st["completed"] == st["submitted"] > 900
    assert 0 < st["peak_live_bytes"] <= budget
    ms = srv.last_frame_ms()
    assert ms[0] > 0 and ms[3] >= 0
    srv.close()
    for e in engines:
        e.close()


def _burst_run(policy, smaps, n_small, n_large, budget):
    engines = [build_map(sm, 8, 16, round_budget=16) for sm in smaps]
    srv = Server(engines, queue_memory_budget=budget, admission_policy=policy)
    # every agent gets one request: its start node is its initial position,
    # so each request's search is the same under both policies
    ms = np.array([0] * n_small + [1] * n_large, np.int32)
    ag = np.array(list(range(n_small)) + list(range(n_large)), np.int32)
    xy = np.stack([smaps[0].targets[i] for i in range(n_small)] + [smaps[1].targets[i] for i in range(n_large)])
    t, st = srv.submit_many(ms, ag, xy)  # the burst on map 0 arrives first
    assert (st == 0).all()
    done = []
    for f in range(3000):
        srv.step(1.0 / 60)
        done.append(srv.poll())
        s = srv.stats()
        if s["pending"] == 0 and s["in_flight"] == 0:
            break
    done = np.concatenate(done)
    assert srv.stats()["peak_live_bytes"] <= budget
    corr = {int(d["ticket"]): engines[int(d["map"])].corridor(int(d["agent"])) for d in done if d["found"]}
    srv.close()
    for e in engines:
        e.close()
    return done, corr


def test_fair_policy_bounds_the_wait_of_a_map_behind_a_burst():
    locs = Locators()
    smaps = serving_maps(locs, seed=4, sizes=(160.0, 640.0), agents=(600, 32), slots=(600, 8),
                         target_pool=4096)
    probe = Server([build_map(sm, 8, 16, round_budget=16) for sm in smaps])
    fl = probe.footprint(1)
    probe.close()
    budget = 3 * fl
    print("burst fifo", flush=True)
    fifo, cf = _burst_run(ADMIT_FIFO, smaps, 600, 8, budget)
    print("burst fair", flush=True)
    fair, ca = _burst_run(ADMIT_FAIR, smaps, 600, 8, budget)
    assert len(fifo) == len(fair) == 608
    # same requests, same results: the policy changes only when a search runs
    for a, b in ((fifo, fair),):
        ka = {int(d["ticket"]): int(d["found"]) for d in a}
        kb = {int(d["ticket"]): int(d["found"]) for d in b}
        assert ka == kb
    assert cf.keys() == ca.keys() and all(np.array_equal(cf[k], ca[k]) for k in cf)
    wait = lambda d: d["admit_frame"] - d["submit_frame"]  # noqa: E731
    lf, la = fifo[fifo["map"] == 1], fair[fair["map"] == 1]
    # FIFO: map 1's requests wait behind the whole burst; fair: map 1 is served
    # in the first admission and every later one it can afford
    assert la["admit_frame"].min() == 0 < lf["admit_frame"].min()
    assert wait(la).max() < wait(lf).max()
    print(f"map-1 admission wait, frames: fifo max {wait(lf).max()} mean {wait(lf).mean():.1f}; "
          f"fair max {wait(la).max()} mean {wait(la).mean():.1f}", flush=True)


def test_serving_workload_full_size_alone_equals_mix_equals_oracle(oracle):
    """The serving workload at its size: rect maps 256^2, 1024^2 and 4096^2
    (1,024 / 16,384 / 32,768 agents, 1,024 / 4,096 / 4,096 slots) resident
    together, Poisson arrivals at 50,000 requests/s (about 833 a frame) for
    60 frames under a 2 GiB search-memory budget, then drained.  Counters
    stay consistent every frame, and 100 sampled tickets per map give the
    same result alone as in the mix, equal to the oracle's search."""
    locs = Locators()
    smaps = serving_maps(locs)
    assert [sm.size for sm in smaps] == [256.0, 1024.0, 4096.0]
    engines = [build_map(sm, T, Q, round_budget=128) for sm in smaps]
    srv = Server(engines, queue_memory_budget=2 << 30, record_start_nodes=True)
    log = {}
    lam = 50000.0 / 60.0
    print("serving full size: arrivals", flush=True)
    done = _drive(srv, smaps, engines, 60, lam, 17, log)
    print("serving full size: drain", flush=True)
    done = np.concatenate([done, _drive(srv, smaps, engines, 2000, 0, 18, log)])
    st = srv.stats()
    assert st["pending"] == 0 and st["in_flight"] == 0, st
    assert st["completed"] == st["submitted"] == len(done)
    assert st["submitted"] > 40000
    assert st["peak_live_bytes"] <= 2 << 30
    rng = np.random.Generator(np.random.PCG64(21))
    pick = []
    for m in range(3):
        idx = np.nonzero(done["map"] == m)[0]
        assert len(idx) > 1000
        pick += list(rng.choice(idx, size=100, replace=False))
    assert _check_alone(oracle, engines, done, pick, log) == 300
    srv.close()
    for e in engines:
        e.close()
