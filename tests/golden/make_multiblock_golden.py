"""Writes tests/golden/multiblock_*.npz: the reference's own toPackedMesh
(structs/multiblock.h, compiled in place by `make -C oracle ref` into
oracle/_ref/ref_multiblock_golden) run on the block layouts below.

Run from the repository root after `make -C oracle ref`.  The fixtures are the
packer's output arrays only (data, not source)."""
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
REF_BIN = ROOT / "oracle" / "_ref" / "ref_multiblock_golden"

KEYS = ["cell_center_z", "cell_center_r", "cell_volume", "cell_area2d", "cell_face_start", "cell_face_ids",
        "face_owner", "face_neighbor", "face_normal_z", "face_normal_r", "face_center_z", "face_center_r",
        "face_area", "face_group"]
INT_KEYS = {"cell_face_start", "cell_face_ids", "face_owner", "face_neighbor", "face_group"}

# (z0, r0, axial length, axial cells, axial coef, radial length, radial cells, radial coef, gW, gE, gS, gN)
LAYOUTS = {
    "pipe_32x8": [(0.0, 0.0, 0.2, 32, 1.0, 0.01, 8, 1.0, 0, 1, 3, 2)],
    "expansion_small": [
        (0.0, 0.0, 0.04, 12, 1.0, 0.002, 8, 1 / 1.05, 0, -1, 3, 2),
        (0.04, 0.0, 0.26, 40, 1.0, 0.002, 8, 1 / 1.05, -1, 1, 3, -1),
        (0.04, 0.002, 0.26, 40, 1.0, 0.004, 12, 1 / 1.05, 2, 1, -1, 2),
    ],
}


def run_reference(blocks) -> dict:
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "mesh.bin"
        args = [str(REF_BIN), str(out), str(len(blocks))]
        for b in blocks:
            args += [repr(float(x)) if isinstance(x, float) else str(x) for x in b]
        subprocess.run(args, check=True)
        raw = out.read_bytes()
    arrays, pos = {}, 0
    for k in KEYS:
        n = int(np.frombuffer(raw, dtype=np.int64, count=1, offset=pos)[0])
        pos += 8
        dt = np.int32 if k in INT_KEYS else np.float64
        arrays[k] = np.frombuffer(raw, dtype=dt, count=n, offset=pos).copy()
        pos += n * np.dtype(dt).itemsize
    return arrays


def main():
    if not REF_BIN.exists():
        sys.exit("build the reference checker first: make -C oracle ref")
    for name, blocks in LAYOUTS.items():
        a = run_reference(blocks)
        np.savez_compressed(ROOT / "tests" / "golden" / f"multiblock_{name}.npz", **a)
        print(name, a["cell_center_z"].size, "cells", a["face_owner"].size, "faces")


if __name__ == "__main__":
    main()
