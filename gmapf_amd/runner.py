"""Frame-loop driver shared by bench.py and the autotuner: builds a
scenario's map and agents through the public API, runs frames with
continuous re-path (an agent whose search ended gets its next log-uniform
target before the next frame) and reports solves/s and frame-time
percentiles."""
from __future__ import annotations

import numpy as np

from . import GMAPF, IDLE, MODE_BATCHED, MODE_PLAIN, MOVE, SEARCH, Config
from . import scenarios as S


class Locators:
    """Keeps host-only GMAPF objects alive that serve exact point location
    for scenario generation (FillMap* + the host grid, no GPU)."""

    def __init__(self):
        self.keep = []

    def __call__(self, W, H, obs):
        g = GMAPF()
        g.init(Config(agents_max_count=1))
        g.fill_map_start(W, H)
        for o in obs:
            g.fill_map_obstacle(o)
        g.fill_map_end()
        self.keep.append(g)
        return lambda p: g.locate(p)


def make_world(sc, mode, T, Q, budget, profile=False, hash_cell=0.0, spw=0, timing=0, wg=0):
    cfg = Config(agents_max_count=sc.n, agent_radius=sc.radius, agent_speed=sc.speed,
                 path_finder_parallel_agents=sc.P, path_finder_threads_per_agents=T,
                 path_finder_queue_capacity=Q, path_finder_heuristic_k=sc.K, step_mode=mode,
                 round_budget=max(1, budget), is_profiler=profile, hash_cell=hash_cell,
                 searches_per_wave=spw, search_timing=timing, search_workgroups=wg)
    g = GMAPF()
    g.init(cfg)
    sc.build(g)
    for x, y in sc.starts:
        g.add_agent(x, y)
    g.set_agent_targets(np.arange(sc.n, dtype=np.int32), sc.targets)
    g.manual_start()
    return g


def retarget(g, sc, counters, hold=0, since=None, frame=0):
    """Next log-uniform targets for the agents whose search ended.  hold > 0:
    a found agent first moves for `hold` frames (since[i] is the frame it was
    first seen moving, -1 when not moving); an agent that arrives earlier, or
    whose search failed, re-paths at once."""
    xy, state = g.agents()
    if hold <= 0:
        done = np.nonzero(state != SEARCH)[0]
    else:
        moving = state == MOVE
        since[moving & (since < 0)] = frame
        since[~moving] = -1
        done = np.nonzero((state == IDLE) | (moving & (frame - since >= hold)))[0]
        since[done] = -1
    if len(done):
        tg = S.log_uniform_targets(sc.seed, 302, xy[done], 10.0, min(4000.0, sc.width), sc.width, sc.height,
                                   sc.locate, counters, ids=done)
        g.set_agent_targets(done.astype(np.int32), tg)


def run_frames(sc, mode, T, Q, budget, warmup, steps, profile=False, log=None, spw=0, timing=0, settle=0, hold=0,
               wg=0):
    """Continuous re-path frames; returns timing and counter statistics.
    `settle` + `warmup` frames run untimed before the `steps` measured ones
    (the mix of short and long searches in flight settles over the first
    hundred frames).  `hold` > 0: found agents move that many frames before
    re-pathing (the frame then also moves, hashes and steers agents).
    `spw`: searches per search workgroup (0: 64 / T); `wg`: cap on the
    persistent search grid (0: automatic).  `timing` k > 0: HIP events around
    the search launch of every k-th frame; the search time and bytes are
    averaged over the timed frames among the measured ones."""
    g = make_world(sc, mode, T, Q, budget, profile, spw=spw, timing=timing, wg=wg)
    counters = np.array(sc.counters, copy=True)
    since = -np.ones(sc.n, dtype=np.int64)
    frame_ms, found_d, exp_d, disc_d, corr_d, search_ms, search_b, host_ms = [], [], [], [], [], [], [], []
    moving_n = []
    prev = g.stats()
    warmup = settle + warmup
    for f in range(warmup + steps):
        if f >= warmup:
            moving_n.append(int((g.agent_states() == MOVE).sum()))
        g.async_step(sc.dt)
        g.wait_step_end()
        ms = g.last_frame_ms()
        st = g.stats()
        if f >= warmup:
            frame_ms.append(ms[0])
            host_ms.append(ms[5:8])
            found_d.append(st["found"] - prev["found"])
            exp_d.append(st["expansions"] - prev["expansions"])
            disc_d.append(st["discoveries"] - prev["discoveries"])
            corr_d.append(st["corridor_nodes"] - prev["corridor_nodes"])
            if ms[2] >= 0:
                # algorithmic bytes of the launch: 48 per expansion, 16 per
                # discovery, 8 per corridor node
                search_ms.append(ms[2])
                search_b.append(48.0 * exp_d[-1] + 16.0 * disc_d[-1] + 8.0 * corr_d[-1])
        prev = st
        retarget(g, sc, counters, hold, since, f)
        if log and (f == warmup - 1 or (f >= warmup and (f - warmup) % 30 == 29)):
            log(f"mode={mode} frame {f}: {ms[0]:.2f} ms, found so far {st['found']}")
    final = g.stats()
    info = g.launch_info()
    g.close()
    fm = np.array(frame_ms)
    tot_s = fm.sum() / 1e3
    return dict(
        solves_per_s=float(np.sum(found_d) / tot_s) if tot_s > 0 else 0.0,
        p50_ms=float(np.percentile(fm, 50)), p99_ms=float(np.percentile(fm, 99)),
        mean_ms=float(fm.mean()), found=int(np.sum(found_d)), expansions=int(np.sum(exp_d)),
        expansions_per_s=float(np.sum(exp_d) / tot_s) if tot_s > 0 else 0.0,
        # the search kernel's HIP-event time per timed frame, and the
        # algorithmic bytes of the same launches
        search_ms=float(np.mean(search_ms)) if search_ms else None,
        search_bytes=float(np.mean(search_b)) if search_b else None,
        search_frames=len(search_ms),
        host_enqueue_wait_readback_ms=[round(float(x), 4) for x in np.mean(np.array(host_ms), axis=0)],
        moving_agents_mean=float(np.mean(moving_n)) if moving_n else 0.0,
        not_found=final["not_found"], stats=final, launch=info)


__all__ = ["Locators", "make_world", "retarget", "run_frames", "MODE_PLAIN", "MODE_BATCHED"]
