"""Autotuner for the batched step: threads per search T, queue capacity Q and
the per-frame round budget, chosen by measurement on the target workload.

For every (T, Q) whose LDS footprint fits, continuous re-path runs of the
batched form at each round budget of BUDGETS measure solves/s (found
searches only: a small Q that fails fast does not count) and the p99 frame
time over the benchmark's own window (5 warm-up + 120 frames: the mix of
short and long searches in flight shifts towards long ones over the first
hundred frames, so shorter probes favour small budgets).  The fastest (T, Q, budget) with p99 within the frame target wins after
a longer confirmation run of the best few.  Results are cached per
(workload, agents) in gmapf_amd/tuned.json.

    python -m gmapf_amd.autotune [--agents 16384] [--target-ms 4.0]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import MODE_BATCHED, GmapfError
from . import scenarios as S
from .runner import Locators, run_frames

TUNE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
T_SET = (2, 4, 8, 16, 32, 64)
Q_SET = (16, 32, 64, 128, 256)
LDS_BYTES = 160 * 1024


BUDGETS = (48, 64, 96, 128, 192, 256)


def lds_fits(Q):
    return 64 * Q * 8 + 4096 <= LDS_BYTES


def tune(sc, target_ms=4.0, budgets=BUDGETS, warmup=5, steps=120, confirm_steps=120, top=3, log=print):
    table = []
    for T in T_SET:
        for Q in Q_SET:
            if not lds_fits(Q):
                continue
            for B in budgets:
                t0 = time.time()
                try:
                    r = run_frames(sc, MODE_BATCHED, T, Q, B, warmup, steps)
                except GmapfError as e:  # does not fit this device
                    log(f"[tune] T={T} Q={Q}: {e}")
                    break
                row = dict(T=T, Q=Q, budget=B, solves_per_s=r["solves_per_s"], p50_ms=r["p50_ms"],
                           p99_ms=r["p99_ms"], found=r["found"], secs=round(time.time() - t0, 1))
                table.append(row)
                log(f"[tune] T={T:2d} Q={Q:3d} B={B:3d}: {r['solves_per_s']:.4g} solves/s "
                    f"p50 {r['p50_ms']:.2f} p99 {r['p99_ms']:.2f} ms")
                if r["p50_ms"] > target_ms:  # larger budgets only make frames longer
                    break
    ok = [r for r in table if r["p99_ms"] <= target_ms] or sorted(table, key=lambda r: r["p99_ms"])[:top]
    ranked = sorted(ok, key=lambda r: -r["solves_per_s"])[:top]
    best = None
    for r in ranked:
        c = run_frames(sc, MODE_BATCHED, r["T"], r["Q"], r["budget"], warmup, confirm_steps)
        r["solves_confirm"] = c["solves_per_s"]
        r["p99_confirm"] = c["p99_ms"]
        log(f"[tune] confirm T={r['T']} Q={r['Q']} B={r['budget']}: {c['solves_per_s']:.4g} solves/s "
            f"p99 {c['p99_ms']:.2f} ms")
        if c["p99_ms"] <= target_ms and (best is None or c["solves_per_s"] > best["solves_confirm"]):
            best = r
    if best is None:
        best = min(ranked, key=lambda r: r["p99_confirm"])
    return best, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--agents", type=int, default=16384)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--target-ms", type=float, default=4.0)
    ap.add_argument("--out", default=TUNE_FILE)
    args = ap.parse_args()
    loc = Locators()
    sc = S.config3(loc, n_agents=args.agents, cells=args.cells)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    best, table = tune(sc, args.target_ms, log=log)
    try:
        with open(args.out) as f:
            d = json.load(f)
    except (OSError, ValueError):
        d = {}
    d[f"{sc.name}/{sc.n}"] = dict(T=best["T"], Q=best["Q"], budget=best["budget"],
                                 solves_per_s=best.get("solves_confirm"), p99_ms=best.get("p99_confirm"),
                                 target_ms=args.target_ms, table=table)
    with open(args.out, "w") as f:
        json.dump(d, f, indent=1)
    print(json.dumps({k: best[k] for k in ("T", "Q", "budget")}))


if __name__ == "__main__":
    main()
