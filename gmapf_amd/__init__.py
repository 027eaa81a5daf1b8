"""gmapf_amd -- Python mirror of the gmapf::GMAPF API over the C ABI.

The product is the C ABI in ``include/gmapf_c.h`` (implemented by
``libgmapf_amd.so``: host C++ runtime + gfx950 HIP kernels).  This module
binds it with ctypes so tests and the benchmark drive exactly what a C or C++
caller drives.  Method names follow the reference class
(include/gmapf/GMAPF.h:30-58 of the reference): ``init``, ``fill_map_start``,
``fill_map_obstacle``, ``fill_map_end``, ``fill_map_manually``, ``add_agent``,
``set_agent_targ_pos`` (alias ``set_agent_target_pos``), ``manual_start``,
``async_step``, ``wait_step_end``, ``get_agent_pos``, ``get_agent_state``,
``profiler_dump``.  Misuse raises :class:`GmapfError` where the reference
throws ``std::exception``.

There is no CPU fallback: if the shared library is missing, loading fails.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

__all__ = [
    "GMAPF", "Config", "GmapfError", "IDLE", "SEARCH", "MOVE", "MODE_PLAIN", "MODE_BATCHED",
    "load_library", "library_path",
]

IDLE, SEARCH, MOVE = 0, 1, 2
MODE_PLAIN, MODE_BATCHED = 0, 1

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    return os.path.join(_HERE, "libgmapf_amd.so")


class GmapfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gmapf error {code}: {msg}")
        self.code = code
        self.msg = msg


class _Config(C.Structure):
    _fields_ = [
        ("agents_max_count", C.c_int),
        ("agent_radius", C.c_float),
        ("agent_speed", C.c_float),
        ("path_finder_parallel_agents", C.c_int),
        ("path_finder_threads_per_agents", C.c_int),
        ("path_finder_queue_capacity", C.c_int),
        ("path_finder_heuristic_k", C.c_float),
        ("is_debug", C.c_int),
        ("is_profiler", C.c_int),
        ("step_mode", C.c_int),
        ("round_budget", C.c_int),
        ("hash_cell", C.c_float),
        ("rot_cs", C.c_float),
        ("rot_sn", C.c_float),
        ("device", C.c_int),
        ("corridor_slab", C.c_longlong),
        ("searches_per_wave", C.c_int),
        ("search_timing", C.c_int),
        ("search_workgroups", C.c_int),
        ("search_wide_keys", C.c_int),
    ]


class _Node(C.Structure):
    _fields_ = [
        ("p1", C.c_float * 2),
        ("p2", C.c_float * 2),
        ("p3", C.c_float * 2),
        ("pcenter", C.c_float * 2),
        ("neibs_idx", C.c_int * 3),
    ]


@dataclass
class Config:
    """Reference gmapf::Config (Config.h:5-17) plus build knobs."""
    agents_max_count: int = 2048
    agent_radius: float = 2.0
    agent_speed: float = 1.0
    path_finder_parallel_agents: int = 2048
    path_finder_threads_per_agents: int = 4
    path_finder_queue_capacity: int = 16
    path_finder_heuristic_k: float = 1.0
    is_debug: bool = False
    is_profiler: bool = False
    step_mode: int = MODE_PLAIN
    round_budget: int = 256
    hash_cell: float = 0.0
    rot_cs: float = 0.0
    rot_sn: float = 0.0
    device: int = 0
    corridor_slab: int = 0
    searches_per_wave: int = 0
    search_timing: int = 0  # HIP events around the search launch of every k-th frame; 0: none
    search_workgroups: int = 0  # batched form: cap on the persistent search grid; 0: automatic
    search_wide_keys: bool = False  # keys without G and side also on maps of <= 32,768 nodes

    def _to_c(self) -> _Config:
        c = _Config()
        for f in fields(self):
            v = getattr(self, f.name)
            setattr(c, f.name, int(v) if isinstance(v, bool) else v)
        return c


_LIB = None

_SIGS = {
    "gmapf_config_default": [C.c_void_p],
    "gmapf_create": [C.POINTER(C.c_void_p)],
    "gmapf_destroy": [C.c_void_p],
    "gmapf_init": [C.c_void_p, C.POINTER(_Config)],
    "gmapf_manual_start": [C.c_void_p],
    "gmapf_fill_map_start": [C.c_void_p, C.c_float, C.c_float],
    "gmapf_fill_map_obstacle": [C.c_void_p, C.c_void_p, C.c_int],
    "gmapf_fill_map_end": [C.c_void_p],
    "gmapf_fill_map_manually": [C.c_void_p, C.c_void_p, C.c_int],
    "gmapf_nav_encode": [C.c_void_p, C.c_void_p, C.c_longlong, C.POINTER(C.c_longlong)],
    "gmapf_fill_map_from_nav": [C.c_void_p, C.c_void_p, C.c_longlong],
    "gmapf_save_nav": [C.c_void_p, C.c_char_p],
    "gmapf_load_nav": [C.c_void_p, C.c_char_p],
    "gmapf_add_agent": [C.c_void_p, C.c_float, C.c_float, C.POINTER(C.c_int)],
    "gmapf_set_agent_targ_pos": [C.c_void_p, C.c_int, C.c_float, C.c_float],
    "gmapf_async_step": [C.c_void_p, C.c_float],
    "gmapf_wait_step_end": [C.c_void_p],
    "gmapf_get_agent_pos": [C.c_void_p, C.c_int, C.c_void_p],
    "gmapf_get_agent_state": [C.c_void_p, C.c_int, C.POINTER(C.c_int)],
    "gmapf_agent_positions": [C.c_void_p, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_int)],
    "gmapf_profiler_dump": [C.c_void_p],
    "gmapf_agent_count": [C.c_void_p, C.POINTER(C.c_int)],
    "gmapf_get_agents_soa": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "gmapf_set_agent_targ_pos_many": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
    "gmapf_get_agent_corridor": [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)],
    "gmapf_get_agent_steps": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "gmapf_map_size": [C.c_void_p, C.POINTER(C.c_int)],
    "gmapf_get_map": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "gmapf_locate": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "gmapf_locate_device": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "gmapf_locate_device_noalloc": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "gmapf_get_stats": [C.c_void_p, C.c_void_p],
    "gmapf_last_frame_ms": [C.c_void_p, C.c_void_p],
    "gmapf_get_hash": [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)],
    "gmapf_get_launch_info": [C.c_void_p, C.c_void_p],
    "gmapf_get_frame_kinds": [C.c_void_p, C.c_void_p],
    "gmapf_get_slab_info": [C.c_void_p, C.c_void_p],
    "gmapf_get_config": [C.c_void_p, C.POINTER(_Config)],
    "gmapf_agent_states": [C.c_void_p, C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int)],
    # serving loop (include/gmapf_serve.h)
    "gmapf_server_create": [C.POINTER(C.c_void_p), C.c_void_p, C.c_int, C.c_void_p],
    "gmapf_server_destroy": [C.c_void_p],
    "gmapf_server_submit": [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float, C.POINTER(C.c_longlong)],
    "gmapf_server_submit_many": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p],
    "gmapf_server_async_step": [C.c_void_p, C.c_float],
    "gmapf_server_wait_step_end": [C.c_void_p],
    "gmapf_server_poll": [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)],
    "gmapf_server_plan": [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)],
    "gmapf_server_get_stats": [C.c_void_p, C.c_void_p],
    "gmapf_server_last_frame_ms": [C.c_void_p, C.c_void_p],
    "gmapf_server_footprint": [C.c_void_p, C.c_int, C.POINTER(C.c_longlong)],
}


def load_library(path: str | None = None):
    """Load libgmapf_amd.so (raises OSError when it is not built)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or library_path()
    if not os.path.exists(p):
        raise OSError(f"{p} is not built; run __graft_entry__.build() or make -C gmapf_amd/csrc")
    lib = C.CDLL(p)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.gmapf_last_error.argtypes = [C.c_void_p]
    lib.gmapf_last_error.restype = C.c_char_p
    lib.gmapf_server_last_error.argtypes = [C.c_void_p]
    lib.gmapf_server_last_error.restype = C.c_char_p
    if path is None:
        _LIB = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class GMAPF:
    """One gmapf::GMAPF object (GMAPF.h:30-98) behind the C ABI."""

    def __init__(self):
        self._lib = load_library()
        h = C.c_void_p()
        self._check(self._lib.gmapf_create(C.byref(h)), h)
        self._h = h

    # -- plumbing
    def _check(self, rc: int, h=None):
        if rc != 0:
            hh = h if h is not None else self._h
            msg = self._lib.gmapf_last_error(hh).decode(errors="replace") if hh else ""
            raise GmapfError(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gmapf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- lifecycle
    def init(self, config: Config):
        c = config._to_c()
        self._check(self._lib.gmapf_init(self._h, C.byref(c)))
        self.config = config

    def manual_start(self):
        self._check(self._lib.gmapf_manual_start(self._h))

    # -- map
    def fill_map_start(self, width: float, height: float | None = None):
        if height is None:  # FillMapStart(V2Float) form
            width, height = width
        self._check(self._lib.gmapf_fill_map_start(self._h, float(width), float(height)))

    def fill_map_obstacle(self, vertices):
        v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32).reshape(-1, 2))
        self._check(self._lib.gmapf_fill_map_obstacle(self._h, _ptr(v), len(v)))

    def fill_map_end(self):
        self._check(self._lib.gmapf_fill_map_end(self._h))

    def fill_map_manually(self, tri, nbr):
        tri = np.asarray(tri, dtype=np.float32).reshape(-1, 6)
        nbr = np.asarray(nbr, dtype=np.int32).reshape(-1, 3)
        n = len(tri)
        nodes = (_Node * n)()
        buf = np.frombuffer(nodes, dtype=np.uint8).view(
            np.dtype([("p", np.float32, 8), ("n", np.int32, 3)]))
        buf["p"][:, :6] = tri
        buf["n"][:] = nbr
        self._check(self._lib.gmapf_fill_map_manually(self._h, nodes, n))

    # -- map fixture ("GMAPFNAV" byte string, no reference counterpart)
    def nav_bytes(self) -> bytes:
        n = C.c_longlong()
        self._check(self._lib.gmapf_nav_encode(self._h, None, 0, C.byref(n)))
        buf = (C.c_ubyte * n.value)()
        self._check(self._lib.gmapf_nav_encode(self._h, buf, n.value, C.byref(n)))
        return bytes(buf)

    def fill_map_from_nav(self, data: bytes):
        data = bytes(data)
        self._check(self._lib.gmapf_fill_map_from_nav(self._h, data, len(data)))

    def save_nav(self, path: str):
        self._check(self._lib.gmapf_save_nav(self._h, os.fsencode(path)))

    def load_nav(self, path: str):
        self._check(self._lib.gmapf_load_nav(self._h, os.fsencode(path)))

    # -- agents
    def add_agent(self, x: float, y: float) -> int:
        i = C.c_int()
        self._check(self._lib.gmapf_add_agent(self._h, float(x), float(y), C.byref(i)))
        return i.value

    def set_agent_targ_pos(self, agent_id: int, x: float, y: float):
        self._check(self._lib.gmapf_set_agent_targ_pos(self._h, int(agent_id), float(x), float(y)))

    set_agent_target_pos = set_agent_targ_pos

    def set_agent_targets(self, ids, xy, raise_on_error=True):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float32).reshape(-1, 2))
        st = np.zeros(len(ids), dtype=np.int32)
        rc = self._lib.gmapf_set_agent_targ_pos_many(self._h, len(ids), _ptr(ids), _ptr(xy), _ptr(st))
        self._check(rc)
        if raise_on_error and (st != 0).any():
            k = int(np.nonzero(st)[0][0])
            raise GmapfError(int(st[k]), f"target {k} rejected")
        return st

    # -- frame
    def async_step(self, dt: float):
        self._check(self._lib.gmapf_async_step(self._h, float(dt)))

    def wait_step_end(self):
        self._check(self._lib.gmapf_wait_step_end(self._h))

    def step(self, dt: float):
        self.async_step(dt)
        self.wait_step_end()

    # -- readback
    def get_agent_pos(self, agent_id: int):
        xy = np.zeros(2, dtype=np.float32)
        self._check(self._lib.gmapf_get_agent_pos(self._h, int(agent_id), _ptr(xy)))
        return float(xy[0]), float(xy[1])

    def get_agent_state(self, agent_id: int) -> int:
        s = C.c_int()
        self._check(self._lib.gmapf_get_agent_state(self._h, int(agent_id), C.byref(s)))
        return s.value

    def agent_count(self) -> int:
        n = C.c_int()
        self._check(self._lib.gmapf_agent_count(self._h, C.byref(n)))
        return n.value

    def get_config(self) -> Config:
        c = _Config()
        self._check(self._lib.gmapf_get_config(self._h, C.byref(c)))
        kw = {}
        for f in fields(Config):
            v = getattr(c, f.name)
            kw[f.name] = bool(v) if f.type in (bool, "bool") else v
        return Config(**kw)

    def agent_states(self) -> np.ndarray:
        """Copy of the host state mirror (refreshed by every wait_step_end)."""
        p = C.POINTER(C.c_int)()
        n = C.c_int()
        self._check(self._lib.gmapf_agent_states(self._h, C.byref(p), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def agents(self, with_nodes=False):
        n = self.agent_count()
        xy = np.zeros((n, 2), dtype=np.float32)
        st = np.zeros(n, dtype=np.int32)
        nd = np.zeros(n, dtype=np.int32) if with_nodes else None
        self._check(self._lib.gmapf_get_agents_soa(
            self._h, _ptr(xy), _ptr(st), _ptr(nd) if with_nodes else None))
        return (xy, st, nd) if with_nodes else (xy, st)

    def corridor(self, agent_id: int):
        n = C.c_int()
        self._check(self._lib.gmapf_get_agent_corridor(self._h, int(agent_id), None, 0, C.byref(n)))
        if n.value < 0:
            return None
        buf = np.zeros(n.value, dtype=np.int32)
        self._check(self._lib.gmapf_get_agent_corridor(self._h, int(agent_id), _ptr(buf), n.value,
                                                       C.byref(n)))
        return buf

    def steps(self):
        n = self.agent_count()
        si = np.zeros(n, dtype=np.int32)
        sn = np.zeros(n, dtype=np.int32)
        sxy = np.zeros((n, 2), dtype=np.float32)
        self._check(self._lib.gmapf_get_agent_steps(self._h, _ptr(si), _ptr(sn), _ptr(sxy)))
        return si, sn, sxy

    def map_arrays(self):
        v = C.c_int()
        self._check(self._lib.gmapf_map_size(self._h, C.byref(v)))
        V = v.value
        tri = np.zeros((V, 6), dtype=np.float32)
        nbr = np.zeros((V, 3), dtype=np.int32)
        cen = np.zeros((V, 2), dtype=np.float32)
        self._check(self._lib.gmapf_get_map(self._h, _ptr(tri), _ptr(nbr), _ptr(cen)))
        return tri, nbr, cen

    def locate(self, xy, device=False, noalloc=False):
        """Min containing triangle per point, or -1: on the host, or on the
        GPU (device=True); noalloc=True stages through the library's fixed
        per-GPU buffer instead of allocating (at most 2^20 points)."""
        xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float32).reshape(-1, 2))
        out = np.zeros(len(xy), dtype=np.int32)
        if noalloc:
            fn = self._lib.gmapf_locate_device_noalloc
        else:
            fn = self._lib.gmapf_locate_device if device else self._lib.gmapf_locate
        self._check(fn(self._h, len(xy), _ptr(xy), _ptr(out)))
        return out

    def stats(self):
        out = np.zeros(8, dtype=np.int64)
        self._check(self._lib.gmapf_get_stats(self._h, _ptr(out)))
        keys = ["found", "not_found", "expansions", "discoveries", "corridor_nodes", "rounds",
                "admitted", "frames"]
        return dict(zip(keys, (int(x) for x in out)))

    def last_frame_ms(self):
        out = np.zeros(8, dtype=np.float64)
        self._check(self._lib.gmapf_last_frame_ms(self._h, _ptr(out)))
        return out

    def hash_state(self):
        n = self.agent_count()
        keys = np.zeros(max(n, 1), dtype=np.int32)
        order = np.zeros(max(n, 1), dtype=np.int32)
        tot = C.c_int()
        self._check(self._lib.gmapf_get_hash(self._h, _ptr(keys), _ptr(order), C.byref(tot)))
        return keys[:n], order[:tot.value]

    def slab_info(self):
        out = (C.c_longlong * 4)()
        self._check(self._lib.gmapf_get_slab_info(self._h, out))
        return dict(capacity=out[0], in_use=out[1], compactions=out[2], deferred=out[3])

    def frame_kinds(self):
        """Frames run so far, and those among them that had nothing to admit
        or search (the crowd phase and the readback only)."""
        out = np.zeros(2, dtype=np.int64)
        self._check(self._lib.gmapf_get_frame_kinds(self._h, _ptr(out)))
        return dict(frames=int(out[0]), crowd_only=int(out[1]))

    def launch_info(self):
        out = np.zeros(8, dtype=np.int32)
        self._check(self._lib.gmapf_get_launch_info(self._h, _ptr(out)))
        return dict(zip(["T", "Q", "P", "budget", "grid", "mode", "V", "persistent"],
                        (int(x) for x in out)))

    def profiler_dump(self):
        self._check(self._lib.gmapf_profiler_dump(self._h))
