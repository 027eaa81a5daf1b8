"""Serving loop over several device-resident maps: Python mirror of
``include/gmapf_serve.h`` plus the serving workload (three rect maps of
256^2, 1024^2 and 4096^2 with Poisson re-path arrivals).

The server itself is native (``gmapf_amd/csrc/server.cpp``); this module
binds it with ctypes and drives it."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import GMAPF, MODE_BATCHED, Config, GmapfError, load_library
from . import scenarios as S

__all__ = ["Server", "ServingMap", "serving_maps", "build_map", "run_serving", "DONE_DTYPE", "ADMIT_FIFO",
           "ADMIT_FAIR"]


class _ServerConfig(C.Structure):
    _fields_ = [("queue_memory_budget", C.c_longlong), ("max_pending", C.c_int),
                ("record_start_nodes", C.c_int), ("admission_policy", C.c_int)]


ADMIT_FIFO, ADMIT_FAIR = 0, 1


class _Done(C.Structure):
    _fields_ = [("ticket", C.c_longlong), ("map", C.c_int), ("agent", C.c_int), ("found", C.c_int),
                ("start_node", C.c_int), ("submit_frame", C.c_longlong), ("admit_frame", C.c_longlong),
                ("done_frame", C.c_longlong), ("latency_ms", C.c_double)]


DONE_DTYPE = np.dtype([("ticket", np.int64), ("map", np.int32), ("agent", np.int32), ("found", np.int32),
                       ("start_node", np.int32), ("submit_frame", np.int64), ("admit_frame", np.int64),
                       ("done_frame", np.int64), ("latency_ms", np.float64)])
assert DONE_DTYPE.itemsize == C.sizeof(_Done)

_STATS = ["frames", "submitted", "admitted", "completed", "found", "no_path", "pending", "in_flight",
          "live_bytes", "peak_live_bytes", "rejected", "held_by_budget", "maps"]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Server:
    """gmapf_server over GMAPF engines (one per map); the server owns their
    frame loop (see include/gmapf_serve.h)."""

    def __init__(self, maps, queue_memory_budget=0, max_pending=0, record_start_nodes=False,
                 admission_policy=ADMIT_FIFO):
        self._lib = load_library()
        self.maps = list(maps)
        hs = (C.c_void_p * len(self.maps))(*[m._h for m in self.maps])
        cfg = _ServerConfig(int(queue_memory_budget), int(max_pending), int(bool(record_start_nodes)),
                            int(admission_policy))
        s = C.c_void_p()
        rc = self._lib.gmapf_server_create(C.byref(s), hs, len(self.maps), C.byref(cfg))
        if rc != 0:
            raise GmapfError(rc, "server create failed (map not filled or bad arguments)")
        self._s = s

    def _check(self, rc):
        if rc != 0:
            raise GmapfError(rc, self._lib.gmapf_server_last_error(self._s).decode(errors="replace"))

    def close(self):
        if getattr(self, "_s", None):
            self._lib.gmapf_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, map_idx, agent, x, y) -> int:
        t = C.c_longlong()
        self._check(self._lib.gmapf_server_submit(self._s, int(map_idx), int(agent), float(x), float(y),
                                                  C.byref(t)))
        return t.value

    def submit_many(self, map_idx, agents, xy):
        n = len(agents)
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(map_idx, np.int32), (n,)))
        a = np.ascontiguousarray(agents, dtype=np.int32)
        xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float32).reshape(-1, 2))
        tickets = np.full(n, -1, dtype=np.int64)
        st = np.zeros(n, dtype=np.int32)
        self._check(self._lib.gmapf_server_submit_many(self._s, n, _ptr(m), _ptr(a), _ptr(xy), _ptr(tickets),
                                                       _ptr(st)))
        return tickets, st

    def async_step(self, dt):
        self._check(self._lib.gmapf_server_async_step(self._s, float(dt)))

    def wait_step_end(self):
        self._check(self._lib.gmapf_server_wait_step_end(self._s))

    def step(self, dt):
        self.async_step(dt)
        self.wait_step_end()

    def poll(self, cap=1 << 16) -> np.ndarray:
        parts = []
        while True:
            buf = np.zeros(cap, dtype=DONE_DTYPE)
            n = C.c_int()
            self._check(self._lib.gmapf_server_poll(self._s, _ptr(buf), cap, C.byref(n)))
            parts.append(buf[:n.value])
            if n.value < cap:
                break
        return np.concatenate(parts)

    def plan(self, cap=1 << 20) -> np.ndarray:
        buf = np.zeros(cap, dtype=np.int64)
        n = C.c_int()
        self._check(self._lib.gmapf_server_plan(self._s, _ptr(buf), cap, C.byref(n)))
        return buf[:min(n.value, cap)].copy()

    def stats(self) -> dict:
        out = np.zeros(16, dtype=np.int64)
        self._check(self._lib.gmapf_server_get_stats(self._s, _ptr(out)))
        return dict(zip(_STATS, (int(x) for x in out)))

    def last_frame_ms(self) -> np.ndarray:
        out = np.zeros(8, dtype=np.float64)
        self._check(self._lib.gmapf_server_last_frame_ms(self._s, _ptr(out)))
        return out

    def footprint(self, map_idx) -> int:
        b = C.c_longlong()
        self._check(self._lib.gmapf_server_footprint(self._s, int(map_idx), C.byref(b)))
        return b.value


# --------------------------------------------------------------- workload
@dataclass
class ServingMap:
    size: float
    n_rects: int
    n_agents: int
    P: int
    obstacles: list = None
    starts: np.ndarray = None
    targets: np.ndarray = None  # pool of uniform free-space targets


def serving_maps(locate_factory, seed=1, sizes=(256.0, 1024.0, 4096.0), agents=(1024, 16384, 32768),
                 slots=(1024, 4096, 4096), target_pool=1 << 16):
    """Rect maps at the obstacle density of the 1000^2 / 400-rect map
    (400 per 10^6 unit^2, sides U[5,40]); starts and a pool of targets
    uniform over free space."""
    out = []
    for k, (W, n, P) in enumerate(zip(sizes, agents, slots)):
        n_rects = int(round(400 * W * W / 1e6))
        obs = S.rect_obstacles(seed + 100 * k, W, W, n_rects, 5.0, 40.0)
        loc = locate_factory(W, W, obs)
        st = S._free_points(seed + 100 * k, 501, n, 0, 0, W, W, loc)
        tg = S._free_points(seed + 100 * k, 502, target_pool, 0, 0, W, W, loc)
        out.append(ServingMap(W, n_rects, n, P, obs, st, tg))
    return out


def build_map(sm: ServingMap, T=8, Q=32, round_budget=256, mode=MODE_BATCHED, agents_max=32768, K=1.0,
              timing=0):
    g = GMAPF()
    g.init(Config(agents_max_count=max(agents_max, sm.n_agents), agent_radius=2.0, agent_speed=1.0,
                  path_finder_parallel_agents=sm.P, path_finder_threads_per_agents=T,
                  path_finder_queue_capacity=Q, path_finder_heuristic_k=K, step_mode=mode,
                  round_budget=round_budget, search_timing=timing))
    g.fill_map_start(sm.size, sm.size)
    for o in sm.obstacles:
        g.fill_map_obstacle(o)
    g.fill_map_end()
    for x, y in sm.starts:
        g.add_agent(x, y)
    g.manual_start()
    return g


def run_serving(smaps, lam, frames, warmup=10, dt=1.0 / 60, seed=7, T=8, Q=32, round_budget=256,
                queue_memory_budget=2 << 30, log=None, timing=0):
    """Poisson(lam * dt) re-path arrivals per frame, each for a uniformly
    chosen map, a uniformly chosen agent of it and a target from the map's
    free-space pool; returns frame-latency percentiles, solves/s and request
    latencies over the measured frames.  timing = k > 0: HIP events around
    every map's search launch in every k-th frame, and per map the
    algorithmic bytes (48 per expansion, 16 per discovery, 8 per corridor
    node) and durations of those launches ("search_timed")."""
    engines = [build_map(sm, T, Q, round_budget, timing=timing) for sm in smaps]
    timed = [[0.0, 0.0, 0] for _ in engines]  # bytes, ms, launches
    eprev = [e.stats() for e in engines] if timing else None
    srv = Server(engines, queue_memory_budget=queue_memory_budget)
    rng = np.random.Generator(np.random.PCG64(seed))
    n_ag = np.array([sm.n_agents for sm in smaps], dtype=np.int64)
    pool = min(len(sm.targets) for sm in smaps)
    pools = np.stack([sm.targets[:pool] for sm in smaps])
    frame_ms, found_n, lat, parts = [], 0, [], []
    prev = srv.stats()
    for f in range(warmup + frames):
        k = int(rng.poisson(lam * dt))
        if k:
            mi = rng.integers(0, len(smaps), size=k)
            ag = (rng.random(k) * n_ag[mi]).astype(np.int32)
            xy = pools[mi, rng.integers(0, pool, size=k)]
            srv.submit_many(mi.astype(np.int32), ag, xy)
        srv.step(dt)
        st = srv.stats()
        done = srv.poll()
        if timing:
            for m, e in enumerate(engines):
                es, ems = e.stats(), e.last_frame_ms()
                if f >= warmup and ems[2] >= 0:
                    timed[m][0] += 48.0 * (es["expansions"] - eprev[m]["expansions"]) + 16.0 * (
                        es["discoveries"] - eprev[m]["discoveries"]) + 8.0 * (
                        es["corridor_nodes"] - eprev[m]["corridor_nodes"])
                    timed[m][1] += float(ems[2])
                    timed[m][2] += 1
                eprev[m] = es
        if f >= warmup:
            ms = srv.last_frame_ms()
            frame_ms.append(ms[0])
            parts.append(ms[1:5])
            found_n += st["found"] - prev["found"]
            lat.append(done["latency_ms"])
        prev = st
        if log and f >= warmup and (f - warmup) % 60 == 59:
            log(f"lam={lam} frame {f}: {frame_ms[-1]:.2f} ms pending {st['pending']} in flight {st['in_flight']}")
    fm = np.array(frame_ms)
    lat = np.concatenate(lat) if lat else np.zeros(0)
    final = srv.stats()
    info = [e.launch_info() for e in engines]
    foot = [srv.footprint(m) for m in range(len(engines))]
    srv.close()
    for e in engines:
        e.close()
    tot = fm.sum() / 1e3
    return dict(
        lam=lam, frames=frames, p50_ms=float(np.percentile(fm, 50)), p99_ms=float(np.percentile(fm, 99)),
        max_ms=float(fm.max()), mean_ms=float(fm.mean()), solves_per_s=float(found_n / tot) if tot else 0.0,
        frame_budget_ms=16.6, frames_over_budget=int((fm > 16.6).sum()),
        request_latency_p50_ms=float(np.percentile(lat, 50)) if len(lat) else None,
        request_latency_p99_ms=float(np.percentile(lat, 99)) if len(lat) else None,
        host_admit_launch_wait_retire_ms=[round(float(x), 4) for x in np.mean(np.array(parts), axis=0)],
        stats=final, footprint_bytes=foot, nodes=[i["V"] for i in info],
        search_timed=[dict(bytes=b, ms=t, launches=n) for b, t, n in timed] if timing else None)
