"""Train the learned initial guess of the pressure correction on large p'
systems and save its weights (build extension; the reference has no learned
part).

Training pairs come from the solver itself.  For each training case (seeded
nozzles of 200k to 1M cells that differ from the benchmark nozzle in seeds and
geometry, and the 238,717-cell sudden expansion) SIMPLE runs from a cold start
and the p' system (A, b) is captured after outer iterations 1, 2, 4, ..., 4096.
Each system is restricted to the level the solver runs the network on (the
first multigrid level with at most 16384 cells): the Galerkin operator
A_k = P^T A P, b_k = P^T b, and the target is that level's converged
correction A_k^{-1} b_k (a direct solve).  The network is trained on those
pairs on the GPU by its own forward and backward kernels.

Writes, beside this file:
  data/guess_weights.npy   float32 parameters (Net(layers=4))
  data/guess_training.json the cases, iterations, pair count, loss curve

Usage:  python axisim-software_amd/train_guess.py [--epochs N] [--quick]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

import axs  # noqa: E402
import learned_guess as lg  # noqa: E402

WEIGHTS = HERE / "data" / "guess_weights.npy"
RECORD = HERE / "data" / "guess_training.json"
ITERATIONS = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]


def training_cases(quick: bool = False):
    """(name, case) pairs: seeded nozzles of 200k, 500k and 1M cells with their
    own throat radius and cone angle (the benchmark nozzle uses seeds 3001/3002
    and the nominal geometry), and the sudden expansion at its full size."""
    out = []
    sizes = [(2.0e5, 6101, 0.0019, 18.0, 450.0), (5.0e5, 6201, 0.0021, 22.0, 550.0), (1.0e6, 6301, 0.0020, 20.0, 500.0)]
    if quick:
        sizes = [(2.0e4, 6001, 0.0019, 18.0, 450.0)]
    for cells, seed, r_throat, cone, re in sizes:
        mesh = axs.nozzle_mesh(axs.nozzle_spacing_for(cells), seed_points=seed, seed_jitter=seed + 1,
                               r_throat=r_throat, cone_deg=cone)
        out.append((f"nozzle_{mesh.n_cells}_seed{seed}", axs.nozzle_case(mesh=mesh, re_throat=re, mg_max_levels=6)))
    if not quick:
        out.append(("expansion_238717", axs.expansion_case()))
    return out


def build_pairs(cases, iterations, log=print):
    pairs, meta = [], []
    for name, case in cases:
        t = time.time()
        mesh = case.mesh
        fs, fn = mesh.arrays["cell_face_start"], lg.slot_neighbours(mesh)
        mg = axs.Multigrid(fs, fn, max_levels=case.config.mg_max_levels)
        hier, k = lg.Hierarchy(mg), lg.guess_level(mg)
        systems = lg.solver_systems(case, list(iterations))
        log(f"{name}: {len(systems)} p' systems captured ({time.time() - t:.1f} s)")
        for it, (AC, AF, b) in zip(iterations, systems):
            pairs.append(lg.coarse_pair(hier, fs, fn, AC, AF, b, k))
            log(f"{name}: pair after iteration {it}")
        meta.append({"case": name, "cells": mesh.n_cells, "level": k, "level_cells": mg.level_info(k)[0],
                     "iterations": list(iterations)})
        log(f"{name}: {len(systems)} pairs on level {k} ({mg.level_info(k)[0]} cells), {time.time() - t:.1f} s")
        mg.close()
    return pairs, meta


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--quick", action="store_true", help="one 20k-cell nozzle, few iterations (a smoke run)")
    ap.add_argument("--out", type=Path, default=WEIGHTS)
    args = ap.parse_args()
    log = lambda m: print(f"[train_guess] {m}", flush=True)  # noqa: E731
    iterations = ITERATIONS[:5] if args.quick else ITERATIONS
    pairs, meta = build_pairs(training_cases(args.quick), iterations, log)
    net = axs.Net(layers=4, seed=args.seed)
    for q in pairs:
        net.add_sample(*q)
    t = time.time()
    losses = net.train(args.epochs, batch=8, lr=args.lr, seed=args.seed)
    log(f"{len(pairs)} pairs, {args.epochs} epochs in {time.time() - t:.1f} s: loss {losses[0]:.5f} -> "
        f"{losses[-1]:.5f}")
    args.out.parent.mkdir(parents=True, exist_ok=True)
    np.save(args.out, net.get_params())
    record = {"pairs": len(pairs), "epochs": args.epochs, "lr": args.lr, "seed": args.seed, "layers": 4,
              "cases": meta, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
              "loss_every_25": [float(x) for x in losses[::25]],
              "target": "A_k^{-1} b_k on the first multigrid level with at most 16384 cells (direct solve)"}
    args.out.with_name(RECORD.name if args.out == WEIGHTS else args.out.stem + ".json").write_text(
        json.dumps(record, indent=1))
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
