"""Writes tests/golden/cdt_ref.npz: triangulations produced by the reference's
vendored CDT library (built from the reference sources by oracle/ref.mk into
oracle/_ref/ref_cdt) with the reference's FillMapEnd recipe
(src/GMAPF.cu:161-196), for maps whose constrained Delaunay triangulation
is unique (vertices in general position).  Run in the container that has
the reference checkout:

    make -f oracle/ref.mk -C oracle && python tests/golden/make_cdt_golden.py
"""
import math
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from gmapf_amd import scenarios as S  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_cdt")


def general_position_case(seed):
    """Rectangle + random convex polygons with irregular float vertices."""
    W = float(np.float32(100 + seed * 37.3))
    H = float(np.float32(80 + seed * 11.7))
    r = S.stream(seed, 11)
    obs, circ = [], []
    tries = 0
    while len(obs) < 5 + 3 * seed and tries < 10000:
        tries += 1
        cx, cy = r.uniform() * W, r.uniform() * H
        rad = 3 + 9 * r.uniform()
        if cx - rad < 2 or cy - rad < 2 or cx + rad > W - 2 or cy + rad > H - 2:
            continue
        if any(math.hypot(cx - a, cy - b) < rad + rb + 2 for a, b, rb in circ):
            continue
        k = 3 + r.randint(6)
        ang = sorted(r.uniform() * 2 * math.pi for _ in range(k))
        obs.append(np.array([(cx + rad * math.cos(a), cy + rad * math.sin(a)) for a in ang], np.float32))
        circ.append((cx, cy, rad))
    return W, H, obs


def run_ref(W, H, obs):
    v = [(0.0, 0.0), (0.0, H), (W, H), (W, 0.0)]
    e = [(0, 1), (1, 2), (2, 3), (3, 0)]
    for o in obs:
        s = len(v)
        for i, p in enumerate(o):
            v.append((float(p[0]), float(p[1])))
            if i > 0:
                e.append((len(v) - 2, len(v) - 1))
        e.append((len(v) - 1, s))
    txt = f"{len(v)}\n" + "".join(f"{np.float32(x).item()!r} {np.float32(y).item()!r}\n" for x, y in v)
    txt += f"{len(e)}\n" + "".join(f"{a} {b}\n" for a, b in e)
    out = subprocess.run([REF], input=txt, capture_output=True, text=True, check=True).stdout.split("\n")
    n = int(out[0])
    rows = np.array([list(map(float, l.split())) for l in out[1:1 + n]])
    return rows[:, :6].astype(np.float32)


def main():
    data = {}
    for seed in range(8):
        W, H, obs = general_position_case(seed)
        tri = run_ref(W, H, obs)
        data[f"c{seed}_wh"] = np.array([W, H], np.float32)
        data[f"c{seed}_nobs"] = np.array([len(o) for o in obs], np.int32)
        data[f"c{seed}_obs"] = np.concatenate(obs).astype(np.float32)
        data[f"c{seed}_tri"] = tri
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "cdt_ref.npz"), **data)
    print("wrote", len(data) // 4, "cases")


if __name__ == "__main__":
    main()
