"""Benchmark of the per-frame step on one MI355X.

Workload (headline): the 4096x4096 spanning-tree maze (64-unit cells,
16-wide corridors), 16,384 agents, start-target distances log-uniform over
10-4000 units, FP32.  Every agent re-paths as soon as its previous search
ended (continuous re-path load).  The same request stream runs through the
plain step (every admitted search runs to completion inside its frame) and
the batched step (persistent work-queue search with a per-frame round budget,
frame replayed from a captured hipGraph), both with the same (T, Q): the
autotuned pair from gmapf_amd/tuned.json when present.

Metric: path solves/s (found searches / summed frame wall time, frame =
AsyncStep entry -> WaitStepEnd return) of the batched step; p50/p99 frame ms,
the plain step's numbers and their ratio are reported beside it, with the
search kernel's HBM roofline and a single-core CPU oracle baseline.  Every
run first settles for --settle untimed frames (the mix of short and long
searches in flight shifts for about a hundred frames), then --warmup, then
the --steps timed frames.  A second pair of runs holds every found agent on
its corridor for --hold frames before it re-paths, so the timed frames also
move, hash and steer agents (the full frame of the reference's step).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --workload {serving, crowd, bowtie, rects}
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gmapf_amd as G  # noqa: E402
from gmapf_amd import scenarios as S  # noqa: E402
from gmapf_amd.autotune import TUNE_FILE  # noqa: E402
from gmapf_amd.runner import Locators, make_world, retarget, run_frames  # noqa: E402

HBM_PEAK_GBS = 8000.0
# frames between the timed search launches of the timed region
SEARCH_TIMING_EVERY = 8


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cpu():
    """Model, clock and thread count of the host the CPU baseline runs on."""
    model, mhz = None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model is None:
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("cpu MHz") and mhz is None:
                    mhz = round(float(line.split(":", 1)[1]), 1)
    except (OSError, ValueError):
        pass
    return dict(cpu_model=model, cpu_mhz=mhz, host_threads=os.cpu_count())


def tuned_entry(name, n):
    try:
        with open(TUNE_FILE) as f:
            return json.load(f)[f"{name}/{n}"]
    except (OSError, KeyError, ValueError):
        return None


def tuned_point(name, n):
    """(T, Q, budget, searches per workgroup) of the tuned point, or None."""
    e = tuned_entry(name, n)
    return (int(e["T"]), int(e["Q"]), int(e["budget"]), int(e.get("spw", 0))) if e else None


def measured_traffic(T, Q, budget, n):
    """HBM bytes per k_search launch from the committed rocprofv3 counter
    summary, when it was taken with the same (T, Q, budget, agents)."""
    path = os.path.join(ROOT, "profiles", "k_search_maze16k_pmc.md")
    try:
        txt = open(path).read()
    except OSError:
        return None
    if f"profile_config: T={T} Q={Q} budget={budget} agents={n}" not in txt:
        return None
    for line in txt.splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def roofline_search(bat, T, Q, budget, n):
    """Search-kernel HBM roofline over the timed batched frames: algorithmic
    bytes 48*E + 16*D + 8*L per launch (E expansions, D discoveries, L
    corridor nodes; device counters, frame by frame) over the kernel's
    HIP-event time on its own stream (events recorded around the launch in
    the captured frame graph of every k-th frame of the timed region),
    averaged over those launches."""
    ms = bat["search_ms"] or 0.0
    b = bat["search_bytes"] or 0.0
    gbs = b / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    return dict(bound="hbm", achieved=round(float(gbs), 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(float(gbs) / HBM_PEAK_GBS, 5), traffic=measured_traffic(T, Q, budget, n),
                bytes_per_launch=b, ms_per_launch=ms)


def cpu_baseline(sc, loc, T, Q, seconds=12.0):
    """Single-core oracle (the CPU restatement) on a bounded sample of the
    same request stream: every agent's start node with its 1st, 2nd, ...
    log-uniform target, searched until ~`seconds` of CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    g = loc.keep[-1]
    tri, nbr, _ = g.map_arrays()
    om = pyoracle.OracleMap(tri, nbr)
    S_nodes = g.locate(sc.starts)
    counters = np.array(sc.counters, copy=True)
    targets = sc.targets
    found = n = rep = 0
    t0 = time.process_time()
    while time.process_time() - t0 < seconds:
        T_nodes = g.locate(targets)
        for i in range(sc.n):
            if time.process_time() - t0 >= seconds:
                break
            if S_nodes[i] != T_nodes[i]:
                ok, _, _ = om.search(int(S_nodes[i]), int(T_nodes[i]), T, Q, sc.K)
                found += int(ok)
                n += 1
        rep += 1
        targets = S.log_uniform_targets(sc.seed, 302, sc.starts, 10.0, min(4000.0, sc.width), sc.width,
                                        sc.height, sc.locate, counters)
    el = time.process_time() - t0
    return dict(value=round(found / el, 2), unit="solves/s", cores=1, kind="port",
                sample=f"{n} searches ({rep} passes over the {sc.n} agents' start nodes with log-uniform targets) "
                       f"of the maze workload, single-threaded oracle, {el:.1f} s CPU", **host_cpu())


def serving_main(args):
    """Serving workload: rect maps 256^2 / 1024^2 / 4096^2 (1,024 / 16,384 /
    32,768 agents), Poisson re-path arrivals at 1k, 10k and 50k requests/s,
    admission by a 2 GiB search-memory budget; one JSON line."""
    from gmapf_amd.serving import run_serving, serving_maps

    loc = Locators()
    smaps = serving_maps(loc)
    T, Q, B = args.T or 8, args.Q or 32, args.budget or 128
    steps = args.steps if args.steps_given else 300
    runs = []
    for lam in ([args.lam] if args.lam else (1000.0, 10000.0, 50000.0)):
        r = run_serving(smaps, lam, steps, warmup=args.warmup, T=T, Q=Q, round_budget=B, log=log)
        log("serving", {k: v for k, v in r.items() if k != "stats"})
        runs.append(r)
    top = runs[-1]
    # roofline of the search launches: a separate run at the top rate with HIP
    # events around every map's search launch in every 8th frame (their graph
    # nodes lengthen a frame, so the reported runs carry none); the largest
    # map's launch is the dominant kernel
    tr = run_serving(smaps, top["lam"], steps, warmup=args.warmup, T=T, Q=Q, round_budget=B,
                     timing=SEARCH_TIMING_EVERY)
    per_map = [dict(gbs=round(t["bytes"] / (t["ms"] / 1e3) / 1e9, 2) if t["ms"] > 0 else None,
                    ms_per_launch=round(t["ms"] / t["launches"], 4) if t["launches"] else None,
                    bytes_per_launch=round(t["bytes"] / t["launches"], 1) if t["launches"] else None)
               for t in tr["search_timed"]]
    big = tr["search_timed"][-1]
    gbs = big["bytes"] / (big["ms"] / 1e3) / 1e9 if big["ms"] > 0 else 0.0
    roof = dict(bound="hbm", achieved=round(gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(gbs / HBM_PEAK_GBS, 5), traffic=None)
    log("serving search roofline per map", per_map)
    cpu = None if args.skip_cpu else serving_cpu_baseline(smaps, loc, T, Q)
    out = {
        "metric": "serving loop: path solves/s with p50/p99 frame ms against a 16.6 ms frame budget",
        "value": round(top["solves_per_s"], 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(top["mean_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int32",
        "data": "synthetic (seeded rect maps, uniform free-space starts/targets, PCG64 Poisson arrivals)",
        "config": {
            "workload": "serving: 3 device-resident rect maps 256^2/1024^2/4096^2 (400 rects per 10^6 unit^2, "
                        "sides 5-40) with 1,024/16,384/32,768 agents, Poisson re-path arrivals, "
                        "2 GiB search-memory admission budget; value at 50k requests/s",
            "navmesh_nodes": top["nodes"],
            "T": T, "Q": Q, "round_budget": B,
            "footprint_bytes_per_request": top["footprint_bytes"],
            "runs": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()
                      if k not in ("stats", "nodes", "footprint_bytes", "search_timed")} | {
                         "admitted": r["stats"]["admitted"], "pending_at_end": r["stats"]["pending"],
                         "peak_live_bytes": r["stats"]["peak_live_bytes"]} for r in runs],
            "search_launch_per_map": per_map,
            "parallelism": "single GPU, maps on separate streams",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def serving_cpu_baseline(smaps, loc, T, Q, seconds=12.0):
    """Single-core oracle on a bounded sample of the serving request mix: a
    uniformly chosen map, one of its agents' start nodes and a target node from
    the map's pool (the same draw as the serving runs), searched until
    ~`seconds` of CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    oms, starts, pools = [], [], []
    for k, sm in enumerate(smaps):
        g = loc.keep[k]
        tri, nbr, _ = g.map_arrays()
        oms.append(pyoracle.OracleMap(tri, nbr))
        starts.append(g.locate(sm.starts))
        pools.append(g.locate(sm.targets))
    rng = np.random.Generator(np.random.PCG64(11))
    found = n = 0
    t0 = time.process_time()
    while time.process_time() - t0 < seconds:
        m = int(rng.integers(0, len(smaps)))
        s_node = int(starts[m][int(rng.integers(0, len(starts[m])))])
        t_node = int(pools[m][int(rng.integers(0, len(pools[m])))])
        if s_node == t_node:
            continue
        ok, _, _ = oms[m].search(s_node, t_node, T, Q, 1.0)
        found += int(ok)
        n += 1
    el = time.process_time() - t0
    return dict(value=round(found / el, 2), unit="solves/s", cores=1, kind="port",
                sample=f"{n} searches of the serving request mix (uniform map, start node of a uniform agent, "
                       f"target node from the map's pool), single-threaded oracle, {el:.1f} s CPU", **host_cpu())


def crowd_bytes(g, cell, n_moving):
    """Algorithmic HBM bytes of one crowd phase (move, hash, avoidance,
    progress) from the hash the frame built: key N*(8+4); histogram and scan
    N*4 + cells*8; scatter N*20; avoidance N*(8 + s*4 + 8) + 8 per candidate
    read from the s stencil cells; move/progress 24 per moving agent.  The
    accept test's triangle reads are not counted."""
    keys, _ = g.hash_state()
    tri, _, _ = g.map_arrays()
    xy = tri.reshape(-1, 2)
    inv = np.float32(1.0) / np.float32(cell)
    hnx = int((np.float32(xy[:, 0].max()) - np.float32(xy[:, 0].min())) * inv) + 1
    hny = int((np.float32(xy[:, 1].max()) - np.float32(xy[:, 1].min())) * inv) + 1
    k = int(np.floor(np.float32(2.5 * 2.0) / np.float32(cell))) + 1
    cnt = np.bincount(keys[keys >= 0], minlength=hnx * hny).reshape(hny, hnx).astype(np.int64)
    pad = np.pad(cnt, k)
    cs = pad.cumsum(0).cumsum(1)
    cs = np.pad(cs, ((1, 0), (1, 0)))
    w = 2 * k + 1
    box = cs[w:, w:] - cs[:-w, w:] - cs[w:, :-w] + cs[:-w, :-w]
    cand = int(box.reshape(-1)[keys[keys >= 0]].sum())
    N = len(keys)
    s = w * w
    b = N * 12 + N * 4 + hnx * hny * 8 + N * 20 + N * (16 + 4 * s) + 8 * cand + 24 * n_moving
    return float(b), cand


def default_hash_cell(radius):
    """The library's default avoidance hash cell (hash_cell <= 0): just above
    cd = 2.5 r, so neighbours within cd lie in a 3 x 3 cell stencil."""
    return float(np.float32(np.float32(radius) * np.float32(2.5)) * np.float32(1.0 + 1.0 / 64.0))


def crowd_frames(sc, cell, warmup, steps, budget):
    """Frame times (ms) of the crowd workload with the given hash cell, the
    mean host split of a timed frame (enqueue, wait for the device, readback)
    and the frame kinds (all frames, frames without admission and search)."""
    g = make_world(sc, G.MODE_BATCHED, sc.T, sc.Q, budget, hash_cell=cell)
    fm, split = [], []
    for f in range(warmup + steps):
        g.async_step(sc.dt)
        g.wait_step_end()
        lm = g.last_frame_ms()
        if f >= warmup:
            fm.append(lm[0])
            split.append(lm[5:8])
        if f % 100 == 0:
            log(f"crowd cell {cell:.4g} frame {f}: {lm[0]:.3f} ms")
    st_final = g.stats()
    st_final["host_split_ms"] = [round(float(v), 4) for v in np.mean(split, axis=0)]
    st_final["frame_kinds"] = g.frame_kinds()
    g.close()
    return np.array(fm), st_final


def crowd_main(args):
    """Crowd workload: 2048^2 with 50 convex obstacles, 65,536 agents in two
    opposing streams, hash cell 4 (5 x 5 stencil) as the workload specifies;
    the library's default cell (just above cd, 3 x 3 stencil) is measured
    beside it.  One JSON line (agent-steps/s)."""
    loc = Locators()
    sc = S.config4(loc)
    steps = args.steps if args.steps_given else 600
    T, Q, B = sc.T, sc.Q, args.budget or 256
    cell = args.hash_cell or 4.0
    stencil = int(np.floor(np.float32(2.5 * sc.radius) / np.float32(cell))) + 1
    fm, st_final = crowd_frames(sc, cell, args.warmup, steps, B)
    dcell = default_hash_cell(sc.radius)
    fd, _ = crowd_frames(sc, dcell, args.warmup, steps, B)
    # roofline of the crowd phase: profiled frames (per-phase HIP events)
    gp = make_world(sc, G.MODE_BATCHED, T, Q, B, profile=True, hash_cell=cell)
    tot_b = tot_ms = 0.0
    cands = []
    for f in range(40):
        _, st_before = gp.agents()
        gp.async_step(sc.dt)
        gp.wait_step_end()
        if f >= 10:
            b, cand = crowd_bytes(gp, cell, int((st_before == G.MOVE).sum()))
            tot_b += b
            tot_ms += gp.last_frame_ms()[4]
            cands.append(cand)
    gp.close()
    gbs = tot_b / (tot_ms / 1e3) / 1e9
    roof = dict(bound="hbm", achieved=round(gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(gbs / HBM_PEAK_GBS, 5), traffic=None)
    cpu = None if args.skip_cpu else crowd_cpu_baseline(sc, loc, cell)
    out = {
        "metric": "crowd step: agent-steps/s at 65,536 agents (move + hash + avoidance + progress)",
        "value": round(sc.n * len(fm) / (fm.sum() / 1e3), 1),
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(fm.mean()), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int32",
        "data": "synthetic (seeded SplitMix64 convex obstacles and band agents)",
        "config": {
            "workload": f"{sc.name}: 50 convex obstacles (6-12 vertices, radius 20-80), {sc.n} agents r=2 v=1 in "
                        f"two opposing streams, hash cell {cell:.4g} ({2 * stencil + 1}x{2 * stencil + 1} stencil), "
                        "dt 1/60",
            "p50_ms": round(float(np.percentile(fm, 50)), 4), "p99_ms": round(float(np.percentile(fm, 99)), 4),
            "crowd_phase_ms": round(tot_ms / 30, 4), "crowd_bytes_per_frame": round(tot_b / 30, 1),
            "neighbour_candidates_per_frame": int(np.mean(cands)),
            "searches_found": st_final["found"], "T": T, "Q": Q, "round_budget": B,
            "host_split_ms": dict(zip(("enqueue", "device_wait", "readback"), st_final["host_split_ms"])),
            "frames_without_search": st_final["frame_kinds"]["crowd_only"],
            "default_cell": {"hash_cell": round(dcell, 4), "stencil": "3x3",
                             "agent_steps_per_s": round(sc.n * len(fd) / (fd.sum() / 1e3), 1),
                             "p50_ms": round(float(np.percentile(fd, 50)), 4),
                             "p99_ms": round(float(np.percentile(fd, 99)), 4)},
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def crowd_cpu_baseline(sc, loc, cell, seconds=12.0):
    """Single-core oracle world (exact hash avoidance) on the same scenario:
    frame 0 (the searches) untimed, then frames until ~`seconds` of CPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    g = loc.keep[-1]
    tri, nbr, _ = g.map_arrays()
    om = pyoracle.OracleMap(tri, nbr)
    w = pyoracle.OracleWorld(om, agents_max=sc.n, radius=sc.radius, speed=sc.speed, parallel=sc.P,
                             threads=sc.T, queue_cap=sc.Q, heur_k=sc.K, round_budget=0, hash_cell=cell,
                             use_grid_locate=1)
    for x, y in sc.starts:
        w.add_agent(x, y)
    for i, (x, y) in enumerate(sc.targets):
        w.set_target(i, x, y)
    w.step(sc.dt)
    t0 = time.process_time()
    n = 0
    while time.process_time() - t0 < seconds:
        w.step(sc.dt)
        n += 1
    el = time.process_time() - t0
    return dict(value=round(sc.n * n / el, 1), unit="agent-steps/s", cores=1, kind="port",
                sample=f"{n} frames of the {sc.n}-agent crowd after the search frame, single-threaded oracle "
                       f"with the exact hash accelerator, {el:.1f} s CPU", **host_cpu())


def _oracle_world(sc, g, cell):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    tri, nbr, _ = g.map_arrays()
    om = pyoracle.OracleMap(tri, nbr)
    w = pyoracle.OracleWorld(om, agents_max=sc.n, radius=sc.radius, speed=sc.speed, parallel=sc.P,
                             threads=sc.T, queue_cap=sc.Q, heur_k=sc.K, round_budget=0, hash_cell=cell,
                             use_grid_locate=1)
    for x, y in sc.starts:
        w.add_agent(x, y)
    for i, (x, y) in enumerate(sc.targets):
        w.set_target(i, x, y)
    return w


def _side_line(metric, value, unit, steps, warmup, ms, config, roofline, cpu, hib=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms, "higher_is_better": hib, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32+int32", "data": "synthetic (seeded SplitMix64 maps and agents)", "config": config,
            "roofline": roofline, "cpu_baseline": cpu}


def bowtie_main(args):
    """Bow-tie map (100^2, one self-intersecting wall polygon), 64 agents,
    T=4 Q=16 K=1, 100 frames at dt 1/60: the plain GPU step and the
    single-core CPU oracle frame by frame, with the frames whose positions,
    states and nodes are bitwise equal counted.  The GPU frame here is a few
    launches' latency (64 searches), so the oracle is the faster one."""
    loc = Locators()
    sc = S.config1(loc)
    frames = args.steps if args.steps_given else sc.frames
    g = make_world(sc, G.MODE_PLAIN, sc.T, sc.Q, 0)
    w = _oracle_world(sc, g, default_hash_cell(sc.radius))
    gpu_ms, cpu_ms, equal = [], [], 0
    for f in range(frames):
        g.async_step(sc.dt)
        g.wait_step_end()
        gpu_ms.append(g.last_frame_ms()[0])
        t0 = time.perf_counter()
        w.step(sc.dt)
        cpu_ms.append((time.perf_counter() - t0) * 1e3)
        gxy, gst, gnd = g.agents(with_nodes=True)
        oxy, ost, ond = w.agents()
        equal += int(np.array_equal(gst, ost) and np.array_equal(gnd, ond)
                     and np.array_equal(gxy.view(np.uint32), oxy.view(np.uint32)))
    st = g.stats()
    g.close()
    gm, cm = np.array(gpu_ms), np.array(cpu_ms)
    log(f"bowtie: gpu {gm.mean():.4f} ms/frame, oracle {cm.mean():.4f} ms/frame, {equal}/{frames} frames equal")
    cfg = {"workload": f"{sc.name}: 100x100 map, bow-tie wall (30,40),(30,60),(60,40),(60,60), {sc.n} agents r=2 v=1, "
                       f"T={sc.T} Q={sc.Q} K=1, plain step, dt 1/60",
           "frames": frames, "bitwise_equal_frames": equal, "found": st["found"], "not_found": st["not_found"],
           "gpu_p50_ms": round(float(np.percentile(gm, 50)), 4), "gpu_p99_ms": round(float(np.percentile(gm, 99)), 4),
           "cpu_oracle_p50_ms": round(float(np.percentile(cm, 50)), 4),
           "cpu_oracle_p99_ms": round(float(np.percentile(cm, 99)), 4),
           "roofline_note": "64 searches on a 100-node map: the frame is launch latency, not memory traffic",
           "parallelism": "single GPU"}
    cpu = dict(value=round(float(cm.mean()), 4), unit="ms/frame", cores=1, kind="port",
               sample=f"the same {frames} frames, single-threaded oracle", **host_cpu())
    print(json.dumps(_side_line("bow-tie map, 64 agents: plain GPU step frame time beside the single-core CPU oracle",
                                round(float(gm.mean()), 4), "ms/frame", frames, 0, round(float(gm.mean()), 4), cfg,
                                None, cpu, hib=False)))


def rects_main(args):
    """Rect map (1000^2, 400 non-overlapping rects with sides 5-40), 2048
    agents all requesting in frame 0, plain step at T=8 Q=32, 300 frames:
    solves/s over the run, frame p50/p99, and the search launch of frame 0
    against its HBM roofline; the single-core oracle on the same run."""
    loc = Locators()
    sc = S.config2(loc)
    frames = args.steps if args.steps_given else sc.frames
    g = make_world(sc, G.MODE_PLAIN, sc.T, sc.Q, 0, timing=frames + 1)  # times frame 0's search only
    fm = []
    st0 = search_ms = None
    for f in range(frames):
        g.async_step(sc.dt)
        g.wait_step_end()
        ms = g.last_frame_ms()
        fm.append(ms[0])
        if f == 0:
            st0, search_ms = g.stats(), ms[2]
    st = g.stats()
    V = g.launch_info()["V"]
    g.close()
    fm = np.array(fm)
    b0 = 48.0 * st0["expansions"] + 16.0 * st0["discoveries"] + 8.0 * st0["corridor_nodes"]
    gbs = b0 / (search_ms / 1e3) / 1e9
    roof = dict(bound="hbm", achieved=round(gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 5),
                traffic=None)
    cpu = None
    if not args.skip_cpu:
        w = _oracle_world(sc, loc.keep[-1], default_hash_cell(sc.radius))
        t0 = time.process_time()
        n = 0
        while n < frames and time.process_time() - t0 < 15.0:
            w.step(sc.dt)
            n += 1
        el = time.process_time() - t0
        found = w.stats()["found"]
        cpu = dict(value=round(found / el, 2), unit="solves/s", cores=1, kind="port",
                   sample=f"{n} of the {frames} frames ({found} searches found in frame 0, then the crowd), "
                          f"single-threaded oracle, {el:.3f} s CPU", **host_cpu())
    solves = st["found"] / (fm.sum() / 1e3)
    cfg = {"workload": f"{sc.name}: 1000x1000 map, 400 seeded non-overlapping rects (sides 5-40), {sc.n} agents r=2 v=1 "
                       f"all requesting in frame 0, T={sc.T} Q={sc.Q}, plain step, dt 1/60",
           "navmesh_nodes": V, "frames": frames, "found": st["found"], "not_found": st["not_found"],
           "p50_ms": round(float(np.percentile(fm, 50)), 4), "p99_ms": round(float(np.percentile(fm, 99)), 4),
           "frame0_ms": round(float(fm[0]), 4), "frame0_solves_per_s": round(st0["found"] / (fm[0] / 1e3), 1),
           "frame0_search_kernel_ms": round(float(search_ms), 4), "parallelism": "single GPU"}
    log("rects", cfg)
    print(json.dumps(_side_line("rect map, 2048 agents: plain-step path solves/s over 300 frames, with p50/p99 frame ms",
                                round(solves, 1), "solves/s", frames, 0, round(float(fm.mean()), 4), cfg, roof, cpu)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=("maze", "serving", "crowd", "bowtie", "rects"), default="maze")
    ap.add_argument("--settle", type=int, default=100, help="untimed frames before the warm-up ones")
    ap.add_argument("--hold", type=int, default=30, help="moving variant: frames a found agent moves first")
    ap.add_argument("--skip-moving", action="store_true")
    ap.add_argument("--agents", type=int, default=16384)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--T", type=int, default=0)
    ap.add_argument("--Q", type=int, default=0)
    ap.add_argument("--budget", type=int, default=0)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--lam", type=float, default=0.0, help="serving: one arrival rate (requests/s) only")
    ap.add_argument("--hash-cell", type=float, default=0.0, help="crowd: avoidance hash cell (0: library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the step runs on one GPU; sharding over several GPUs is "
                                   "designed but not built, run with --gpus 1"}))
        sys.exit(2)
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 120
    if args.workload == "serving":
        return serving_main(args)
    if args.workload == "crowd":
        return crowd_main(args)
    if args.workload == "bowtie":
        return bowtie_main(args)
    if args.workload == "rects":
        return rects_main(args)
    t_all = time.time()
    loc = Locators()
    sc = S.config3(loc, n_agents=args.agents, cells=args.cells)
    tp = tuned_point(sc.name, sc.n)
    T, Q, B, spw = tp if tp else (8, 32, 192, 0)
    T, Q, B = args.T or T, args.Q or Q, args.budget or B
    if args.T or args.Q:
        spw = 0
    log(f"workload {sc.name}: {sc.n} agents, T={T} Q={Q} budget={B} spw={spw} (autotuned={tp is not None})")
    plain = None
    settle = max(0, args.settle)
    if not args.skip_plain:
        plain = run_frames(sc, G.MODE_PLAIN, T, Q, 0, args.warmup, args.steps, log=log, settle=settle)
        log("plain", {k: v for k, v in plain.items() if k not in ("stats", "launch")})
    # HIP events around the search launch of every k-th frame (their two
    # graph nodes add ~20 us to a frame, so the other frames carry none)
    tk = max(1, min(SEARCH_TIMING_EVERY, args.steps))
    bat = run_frames(sc, G.MODE_BATCHED, T, Q, B, args.warmup, args.steps, log=log, timing=tk, settle=settle,
                     spw=spw)
    log("batched", {k: v for k, v in bat.items() if k not in ("stats", "launch")})
    # the complete frame: found agents move `hold` frames before re-pathing,
    # so the timed frames also move, hash, steer and advance agents
    moving = None
    if not args.skip_moving and args.hold > 0:
        # agents re-path in waves every `hold` frames, so the window covers at
        # least four of them whatever --steps is (a shorter one lands on a
        # phase of the cycle, and the plain step's long search frames come
        # once per cycle)
        mk = max(args.steps, 4 * args.hold)
        mb = run_frames(sc, G.MODE_BATCHED, T, Q, B, args.warmup, mk, settle=settle, hold=args.hold, spw=spw)
        mp = None if args.skip_plain else run_frames(sc, G.MODE_PLAIN, T, Q, 0, args.warmup, mk,
                                                      settle=settle, hold=args.hold)
        moving = {"hold_frames": args.hold, "frames": mk, "moving_agents_mean": round(mb["moving_agents_mean"], 1),
                  "batched_solves_per_s": round(mb["solves_per_s"], 1), "batched_mean_ms": round(mb["mean_ms"], 4),
                  "plain_mean_ms": round(mp["mean_ms"], 4) if mp else None,
                  "plain_p50_ms": round(mp["p50_ms"], 4) if mp else None,
                  "batched_p50_ms": round(mb["p50_ms"], 4), "batched_p99_ms": round(mb["p99_ms"], 4),
                  "batched_expansions_per_s": round(mb["expansions_per_s"], 1),
                  "plain_solves_per_s": round(mp["solves_per_s"], 1) if mp else None,
                  "plain_p99_ms": round(mp["p99_ms"], 4) if mp else None,
                  "batched_over_plain": round(mb["solves_per_s"] / mp["solves_per_s"], 3) if mp else None,
                  "batched_over_plain_expansions": (round(mb["expansions_per_s"] / mp["expansions_per_s"], 3)
                                                    if mp else None)}
        log("moving", moving)
    # the autotuner's fastest point, when it is not the reported one
    fastest = None
    te = tuned_entry(sc.name, sc.n) if tp and not (args.T or args.Q or args.budget) else None
    fe = te.get("fastest") if te else None
    if fe and (fe["T"], fe["Q"], fe["budget"], fe.get("spw", 0)) != (T, Q, B, spw):
        fb = run_frames(sc, G.MODE_BATCHED, fe["T"], fe["Q"], fe["budget"], args.warmup, args.steps, timing=tk,
                        settle=settle, spw=fe.get("spw", 0))
        fp = None if args.skip_plain else run_frames(sc, G.MODE_PLAIN, fe["T"], fe["Q"], 0, args.warmup, args.steps,
                                                     settle=settle)
        fastest = dict(T=fe["T"], Q=fe["Q"], round_budget=fe["budget"], searches_per_wave=fe.get("spw", 0),
                       solves_per_s=round(fb["solves_per_s"], 1),
                       p99_ms=round(fb["p99_ms"], 4),
                       plain_solves_per_s=round(fp["solves_per_s"], 1) if fp else None,
                       batched_over_plain=round(fb["solves_per_s"] / fp["solves_per_s"], 3) if fp else None)
        log("fastest", fastest)
    roof = roofline_search(bat, T, Q, B, sc.n)
    log("roofline", roof)
    cpu = None if args.skip_cpu else cpu_baseline(sc, loc, T, Q)
    out = {
        "metric": "path solves/s @16k agents (batched step), with p99 frame ms and plain-step ratio",
        "value": round(bat["solves_per_s"], 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bat["mean_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int32",
        "data": "synthetic (seeded SplitMix64 maze and agents)",
        "config": {
            "workload": f"{sc.name} spanning-tree maze (64-unit cells, 16-wide corridors), {sc.n} agents, "
                        "log-uniform 10-4000 start-target distance, continuous re-path",
            "navmesh_nodes": bat["launch"]["V"],
            "T": T, "Q": Q, "round_budget": B, "searches_per_wave": spw or 64 // T, "autotuned": tp is not None,
            "batched_p50_ms": round(bat["p50_ms"], 4), "batched_p99_ms": round(bat["p99_ms"], 4),
            "batched_expansions_per_s": round(bat["expansions_per_s"], 1),
            "plain_solves_per_s": round(plain["solves_per_s"], 1) if plain else None,
            "plain_p50_ms": round(plain["p50_ms"], 4) if plain else None,
            "plain_p99_ms": round(plain["p99_ms"], 4) if plain else None,
            "batched_over_plain": (round(bat["solves_per_s"] / plain["solves_per_s"], 3)
                                   if plain and plain["solves_per_s"] else None),
            "plain_expansions_per_s": round(plain["expansions_per_s"], 1) if plain else None,
            "batched_over_plain_expansions": (round(bat["expansions_per_s"] / plain["expansions_per_s"], 3)
                                              if plain and plain["expansions_per_s"] else None),
            "settle_frames": settle,
            "moving_variant": moving,
            "search_kernel_ms_per_launch": round(roof["ms_per_launch"], 4),
            "search_kernel_timed_launches": bat["search_frames"],
            "autotune_choice": ("fastest point with batched >= 5x plain at p99 <= 4 ms"
                                if te and te.get("target") else "fastest point at p99 <= 4 ms") if te else None,
            "fastest": fastest,
            "parallelism": "single GPU",
        },
        "roofline": {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "cpu_baseline": cpu,
    }
    log(f"total bench wall {time.time() - t_all:.1f}s")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
