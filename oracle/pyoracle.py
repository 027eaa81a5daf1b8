"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the checker; the product (gmapf_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcConfig(C.Structure):
    _fields_ = [
        ("agents_max", C.c_int),
        ("radius", C.c_float),
        ("speed", C.c_float),
        ("parallel", C.c_int),
        ("threads", C.c_int),
        ("queue_cap", C.c_int),
        ("heur_k", C.c_float),
        ("round_budget", C.c_int),
        ("rot_cs", C.c_float),
        ("rot_sn", C.c_float),
        ("hash_cell", C.c_float),
        ("use_grid_locate", C.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        p = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(p):
            build()
        L = C.CDLL(p)
        vp, ip, fp = C.c_void_p, C.c_int, C.c_float
        sig = {
            "orc_map_create": ([ip, vp, vp], vp),
            "orc_map_destroy": ([vp], None),
            "orc_map_count": ([vp], ip),
            "orc_map_centroids": ([vp, vp], None),
            "orc_locate_brute": ([vp, fp, fp], ip),
            "orc_locate_grid": ([vp, fp, fp], ip),
            "orc_locate_many": ([vp, ip, vp, vp, ip], None),
            "orc_queue_script": ([ip, ip, vp, vp, vp, vp], None),
            "orc_search": ([vp, ip, ip, ip, ip, fp, vp, ip, C.POINTER(C.c_int), vp], ip),
            "orc_world_create": ([vp, C.POINTER(OrcConfig)], vp),
            "orc_world_destroy": ([vp], None),
            "orc_world_add_agent": ([vp, fp, fp], ip),
            "orc_world_set_target": ([vp, ip, fp, fp], ip),
            "orc_world_step": ([vp, fp], None),
            "orc_world_count": ([vp], ip),
            "orc_world_get": ([vp, vp, vp, vp], None),
            "orc_world_corridor": ([vp, ip, vp, ip], ip),
            "orc_world_get_steps": ([vp, vp, vp, vp], None),
            "orc_world_stats": ([vp, vp], None),
            "orc_world_last_sort": ([vp, vp, vp], ip),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleMap:
    def __init__(self, tri, nbr):
        self.tri = np.ascontiguousarray(np.asarray(tri, dtype=np.float32).reshape(-1, 6))
        self.nbr = np.ascontiguousarray(np.asarray(nbr, dtype=np.int32).reshape(-1, 3))
        self.V = len(self.tri)
        self.h = lib().orc_map_create(self.V, _p(self.tri), _p(self.nbr))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_map_destroy(self.h)
            self.h = None

    def centroids(self):
        c = np.zeros((self.V, 2), dtype=np.float32)
        lib().orc_map_centroids(self.h, _p(c))
        return c

    def locate(self, xy, grid=True):
        xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float32).reshape(-1, 2))
        out = np.zeros(len(xy), dtype=np.int32)
        lib().orc_locate_many(self.h, len(xy), _p(xy), _p(out), 1 if grid else 0)
        return out

    def search(self, S, Tg, T, Q, K=1.0):
        cap = self.V + 1
        buf = np.zeros(cap, dtype=np.int32)
        n = C.c_int()
        st = np.zeros(6, dtype=np.int64)
        found = lib().orc_search(self.h, int(S), int(Tg), int(T), int(Q), float(K), _p(buf), cap,
                                 C.byref(n), _p(st))
        return bool(found), buf[: n.value].copy(), dict(
            rounds=int(st[0]), expansions=int(st[1]), discoveries=int(st[2]), length=int(st[3]),
            mS=int(st[4]), mT=int(st[5]))


def queue_script(Q, ops, F, node):
    ops = np.ascontiguousarray(ops, dtype=np.int32)
    F = np.ascontiguousarray(F, dtype=np.float32)
    node = np.ascontiguousarray(node, dtype=np.int32)
    out = np.zeros(len(ops), dtype=np.int32)
    lib().orc_queue_script(int(Q), len(ops), _p(ops), _p(F), _p(node), _p(out))
    return out


def rotation_constants():
    """FP32 cos/sin of the 10-degree turn: angle = 10 * pi_f / 180 in FP32,
    cos/sin in double, rounded once (same recipe as the library's default)."""
    ang = np.float32(np.float32(10.0) * np.float32(3.14159265358979)) / np.float32(180.0)
    return float(np.float32(np.cos(np.float64(ang)))), float(np.float32(np.sin(np.float64(ang))))


class OracleWorld:
    def __init__(self, omap: OracleMap, *, agents_max, radius, speed, parallel, threads, queue_cap,
                 heur_k=1.0, round_budget=0, hash_cell=0.0, use_grid_locate=1, rot=None):
        cs, sn = rot if rot is not None else rotation_constants()
        self.cfg = OrcConfig(agents_max, radius, speed, parallel, threads, queue_cap, heur_k,
                             round_budget, cs, sn, hash_cell, use_grid_locate)
        self.map = omap
        self.h = lib().orc_world_create(omap.h, C.byref(self.cfg))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_world_destroy(self.h)
            self.h = None

    def add_agent(self, x, y):
        return lib().orc_world_add_agent(self.h, float(x), float(y))

    def set_target(self, i, x, y):
        return lib().orc_world_set_target(self.h, int(i), float(x), float(y))

    def step(self, dt):
        lib().orc_world_step(self.h, float(dt))

    def count(self):
        return lib().orc_world_count(self.h)

    def agents(self):
        n = self.count()
        xy = np.zeros((n, 2), dtype=np.float32)
        st = np.zeros(n, dtype=np.int32)
        nd = np.zeros(n, dtype=np.int32)
        lib().orc_world_get(self.h, _p(xy), _p(st), _p(nd))
        return xy, st, nd

    def corridor(self, i):
        cap = self.map.V + 1
        buf = np.zeros(cap, dtype=np.int32)
        n = lib().orc_world_corridor(self.h, int(i), _p(buf), cap)
        return None if n < 0 else buf[:n].copy()

    def steps(self):
        n = self.count()
        si = np.zeros(n, dtype=np.int32)
        sn = np.zeros(n, dtype=np.int32)
        sxy = np.zeros((n, 2), dtype=np.float32)
        lib().orc_world_get_steps(self.h, _p(si), _p(sn), _p(sxy))
        return si, sn, sxy

    def stats(self):
        out = np.zeros(8, dtype=np.int64)
        lib().orc_world_stats(self.h, _p(out))
        keys = ["found", "not_found", "expansions", "discoveries", "corridor_nodes", "rounds",
                "admitted", "frames"]
        return dict(zip(keys, (int(x) for x in out)))

    def last_sort(self):
        n = self.count()
        order = np.zeros(max(n, 1), dtype=np.int32)
        keys = np.zeros(max(n, 1), dtype=np.int32)
        m = lib().orc_world_last_sort(self.h, _p(order), _p(keys))
        return keys[:n], order[:m]
