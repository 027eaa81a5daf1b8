"""bench.py's command line on the CPU: more than one GPU is refused with a
JSON error line before any GPU work, and the search-kernel roofline follows
from the timed launches' bytes and durations."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in out["error"]


def test_roofline_from_timed_launches():
    sys.path.insert(0, ROOT)
    import bench

    bat = {"search_ms": 0.5, "search_bytes": 2.0e8}
    r = bench.roofline_search(bat, 4, 16, 128, 16384)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] == pytest.approx(400.0)  # 2e8 B / 0.5 ms
    assert r["frac"] == pytest.approx(0.05)
    none = bench.roofline_search({"search_ms": None, "search_bytes": None}, 4, 16, 128, 16384)
    assert none["achieved"] == 0.0 and none["frac"] == 0.0
