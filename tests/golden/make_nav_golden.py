"""Writes tests/golden/bowtie.nav: the "GMAPFNAV" fixture of the 100x100 map
with the reference README's bow-tie wall (README.md:66-72), as the library's
FillMapStart / FillMapObstacle / FillMapEnd produce it.  The test that reads
it pins the byte format and the triangulation of that map.

    python tests/golden/make_nav_golden.py
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import gmapf_amd as G  # noqa: E402
from gmapf_amd import scenarios as S  # noqa: E402


def bowtie_map_bytes() -> bytes:
    g = G.GMAPF()
    g.init(G.Config(agents_max_count=1))
    g.fill_map_start(100, 100)
    for o in S.bowtie_obstacles():
        g.fill_map_obstacle(o)
    g.fill_map_end()
    b = g.nav_bytes()
    g.close()
    return b


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bowtie.nav")
    data = bowtie_map_bytes()
    with open(out, "wb") as f:
        f.write(data)
    print(f"wrote {out}: {len(data)} bytes")
