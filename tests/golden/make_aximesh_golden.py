"""Writes tests/golden/aximesh_golden.npz: the reference's own mesher
(aximesh/source/aximesh.cpp, compiled in place by `make -C oracle ref` into
oracle/_ref/libref_aximesh.so) run on the inputs below.

Run from the repository root after `make -C oracle ref`.  Per case the file
holds the input (points, segments, parameters) and the reference's output
(points, triangles, neighbour links, renumbered segments, or the error text):
data, not source.  The end-to-end cases are the ones the reference's harness
drives through generateMesh (aximesh/tools/aximesh_cli.cpp:540-625); the
others are discretised domains the product meshes (the nozzle outline, a
channel with an obstacle either way round, a boundary with interior points,
a jittered outline, non-default parameters) and an input on which the
reference stops with an error."""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "axisim-software_amd"))

import aximesh_ref  # noqa: E402
import axs  # noqa: E402  (outline helpers only: plain numpy, no library call)

OUT = ROOT / "tests" / "golden" / "aximesh_golden.npz"

S3 = 0.86602540378443865


def circle(cz, cr, radius, n, group, clockwise=False):
    th = np.linspace(0.0, 2.0 * np.pi, n + 1)[:-1]
    p = [(cz + radius * np.cos(t), cr + radius * np.sin(t)) for t in th]
    edges = [(p[i][0], p[i][1], p[(i + 1) % n][0], p[(i + 1) % n][1], group) for i in range(n)]
    if clockwise:
        edges = [(e[2], e[3], e[0], e[1], e[4]) for e in reversed(edges)]
    return edges


BOX = [(0.0, 0.0, 4.0, 0.0, 3), (4.0, 0.0, 4.0, 1.0, 1), (4.0, 1.0, 0.0, 1.0, 2), (0.0, 1.0, 0.0, 0.0, 0)]


def cases() -> dict:
    c = {
        # aximesh_cli.cpp:540-625, generateMesh end to end
        "triangle": ([(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)], []),
        "square": ([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)], []),
        "rhombus-tall": ([(-1.0, 0.0), (0.0, -2.0), (1.0, 0.0), (0.0, 2.0)], []),
        "rhombus-wide": ([(-1.0, 0.0), (0.0, -0.5), (1.0, 0.0), (0.0, 0.5)], []),
        "collinear-hull": ([(0.0, 0.0), (1.0, 0.0), (2.0, 0.0), (2.0, 1.0), (0.0, 1.0)], []),
        "hexagon-fan": ([(1.0, 0.0), (0.5, S3), (-0.5, S3), (-1.0, 0.0), (-0.5, -S3), (0.5, -S3), (0.0, 0.0)], []),
        "boundary-only": ([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)], [(0, 1), (1, 2), (2, 3), (3, 0)]),
        "square-diagonal": ([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)],
                            [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]),
        "rhombus-forced": ([(-1.0, 0.0), (0.0, -0.5), (1.0, 0.0), (0.0, 0.5)],
                           [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]),
        "blocked-edge": ([(0.0, 0.0), (4.0, 0.0), (2.0, 0.2), (2.0, -2.0)], [(0, 1)]),
        "zigzag": ([(0.0, 0.0), (6.0, 0.0), (1.0, 0.8), (2.0, -0.7), (3.0, 0.9), (4.0, -0.6), (5.0, 0.7)],
                   [(0, 1)]),
        "notched-domain": ([(0.0, 0.0), (2.0, 0.0), (2.0, 2.0), (1.0, 1.0), (0.0, 2.0)],
                           [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]),
    }
    p, s, _ = axs.outline_boundary([axs.nozzle_outline()], 0.002)
    c["nozzle-2mm"] = (p, s)
    p, s, _ = axs.outline_boundary([BOX, circle(1.0, 0.5, 0.2, 24, 4)], 0.1)
    c["obstacle"] = (p, s)
    p, s, _ = axs.outline_boundary([BOX, circle(1.0, 0.5, 0.2, 24, 4, clockwise=True)], 0.1)
    c["obstacle-cw"] = (p, s)
    rng = np.random.default_rng(20261204)
    p, s, _ = axs.outline_boundary([BOX], 0.05)
    c["box-cloud"] = (np.vstack([p, 0.05 + rng.random((100, 2)) * [3.9, 0.9]]), s)
    p, s, _ = axs.outline_boundary([axs.nozzle_outline()], 0.001)
    c["nozzle-jittered"] = (p + rng.normal(0.0, 1e-5, p.shape), s)
    p, s, _ = axs.outline_boundary([axs.nozzle_outline()], 6.25e-5)
    c["nozzle-stalls"] = (p, s)
    return c


PARAMS = {"nozzle-2mm-params": ("nozzle-2mm", (1.4, 0.1, 1.0))}


def main() -> None:
    if not aximesh_ref.available():
        raise SystemExit("build the reference mesher first: make -C oracle ref")
    out = {}
    all_cases = cases()
    runs = [(name, name, (1.4, 0.2, 1.4)) for name in all_cases] + [(k, v[0], v[1]) for k, v in PARAMS.items()]
    for key, name, prm in runs:
        pts, segs = all_cases[name]
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, 2)
        segs = np.asarray(segs, dtype=np.int32).reshape(-1, 2)
        r = aximesh_ref.generate(pts, segs, *prm)
        out[f"{key}.in_points"] = pts
        out[f"{key}.in_segments"] = segs
        out[f"{key}.params"] = np.array(prm)
        out[f"{key}.error"] = np.array(r["error"])
        for k in ("points", "tri_v", "tri_adj", "segments"):
            out[f"{key}.{k}"] = r[k]
        print(f"{key:20s} {len(pts):6d} in, {len(r['points']):6d} points, {len(r['tri_v']):6d} triangles"
              f"{'  error: ' + r['error'] if r['error'] else ''}")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
