"""Benchmark: SIMPLE outer iterations/s on the ~1M-cell unstructured nozzle.

One "step" is one SIMPLE outer iteration (assembly, 2 x 20 momentum Jacobi
sweeps, Rhie-Chow, two non-orthogonal pressure-correction passes with 3
agglomeration-multigrid V-cycles each, corrections, residual norms), replayed
as one hipGraph.  Inputs are resident in HBM when the timed region starts.

Prints one JSON line (see the repository's benchmark contract):
  value       outer iterations/s on one MI355X (FP64 reference arithmetic)
  roofline    the dominant kernel (the fused u+v momentum Jacobi sweeps, several
              sweeps per launch) measured with HIP events on the solver's
              stream, against the 8 TB/s HBM peak; the one-sweep kernel beside it
  mg_vcycle   the p' multigrid solve (3 V-cycles) against the same peak, warm and cold caches
  to_tolerance the nozzle, the 128 x 32 pipe and the 238,717-cell expansion
              from a cold start to the 1e-6 residual, FP64 and mixed: iterations,
              wall time, and the mixed fields' error against the FP64 fields
  batched_sweep 64 cases of 20k-500k cells in one solver, BDF2 pulsatile, each
              case stopping on its own residuals; aggregate cell-updates/s and
              the cases' latency percentiles
  cpu_baseline the FP64 oracle built on this host (-O3 -march=native), a bounded
              sample (20 iterations) of the same workload on every thread this
              process may use (`cores`), the one-thread rate beside it, and the
              pipe run to 1e-6 on the host (`c1`)
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "axisim-software_amd"))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK_GBS = 8000.0


def _cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or the platform's string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=float, default=1.0e6)
    ap.add_argument("--mg-levels", type=int, default=6,
                    help="multigrid levels of the nozzle case (its configuration: 4:1 coarsening, 6 levels)")
    ap.add_argument("--cpu-iters", type=int, default=20, help="oracle sample size (outer iterations)")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the runs to the 1e-6 residual (nozzle, pipe and expansion, FP64 and mixed)")
    ap.add_argument("--converge-cap", type=int, default=60000, help="iteration cap of the runs to 1e-6")
    ap.add_argument("--direct-every", type=int, default=16,
                    help="mixed path with the direct coarsest solve: outer iterations per inverse formation")
    ap.add_argument("--no-guess", action="store_true", help="skip the learned-guess run to 1e-6")
    ap.add_argument("--expansion-cap", type=int, default=20000,
                    help="iteration cap of the 238,717-cell expansion runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default workgroup sizes")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched-sweep leg")
    ap.add_argument("--c4-cases", type=int, default=64)
    ap.add_argument("--c4-steps", type=int, default=200)
    ap.add_argument("--c4-iters", type=int, default=5, help="outer iterations per time step at most")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only launch the measured kernels and a 512 MiB copy, for rocprofv3 --pmc passes")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the solver does not shard "
              f"across GPUs)", file=sys.stderr)
        return 2
    if args.traffic_probe:
        return traffic_probe()

    import numpy as np

    import axs

    t0 = time.time()
    case = axs.nozzle_case(cells=args.cells, max_iter=args.warmup, mg_max_levels=args.mg_levels)
    m = case.mesh
    log(f"mesh: {m.n_cells} cells, {m.n_faces} faces ({time.time() - t0:.1f} s)")

    s = axs.make_solver(case)
    t1 = time.time()
    s.run()  # cold start + warm-up iterations (also builds hierarchy and graph)
    st = s.stats()
    log(f"warm-up: {st.iterations} iterations, setup {st.setup_ms:.0f} ms, mg levels {st.mg_levels} "
        f"({time.time() - t1:.1f} s)")
    tune = None
    if not args.no_autotune:
        d_ms, t_ms = s.autotune(20)
        tune = {"default_ms": round(d_ms, 4), "tuned_ms": round(t_ms, 4)}
        case.config.max_iter = 2  # a continued run adds max_iter iterations
        s.configure(case.config)
        s.run(continue_run=True)  # rebuilds the iteration graph with the chosen sizes
        log(f"autotune: tuned kernels {d_ms:.3f} -> {t_ms:.3f} ms per pass")

    pre_iters = s.stats().iterations
    ms = s.run_timed(args.steps)
    it_s = args.steps / (ms / 1000.0)
    log(f"timed: {args.steps} iterations in {ms:.1f} ms -> {it_s:.1f} it/s")

    # dominant kernel: the u+v momentum Jacobi sweeps, `fused` sweeps per launch
    # (20 sweeps per momentum solve); the one-sweep kernel beside it
    st = s.stats()
    fused, fused_cells = st.sweeps_per_launch, st.sweep_block_cells
    # template arguments as rocprofv3 names the kernels: slots per row (3 on
    # triangle meshes), and cells per block and sweeps per launch when fused
    regs = 3 if int(np.diff(m.arrays["cell_face_start"]).max()) <= 3 else 4
    k1_name = f"kJacobiUV<{regs}>"
    k1_ms, k1_bytes = s.bench_kernel(axs.KERNEL_JACOBI_UV, 200)
    k1_gbs = k1_bytes / (k1_ms * 1e-3) / 1e9
    log(f"{k1_name} (one sweep): {k1_ms * 1000:.1f} us/launch, {k1_gbs:.0f} GB/s")
    if fused >= 2:
        k_ms, k_bytes = s.bench_kernel(axs.KERNEL_JACOBI_UV_FUSED, 200)
        kname = f"kJacobiUVBlock<{fused_cells},{fused},{regs}>"
    else:
        k_ms, k_bytes, kname = k1_ms, k1_bytes, k1_name
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    log(f"{kname}: {k_ms * 1000:.1f} us/launch ({fused} sweeps), {achieved:.0f} GB/s algorithmic")
    # the same launches with the caches flushed before each one
    kc_ms, _ = s.bench_kernel(axs.KERNEL_JACOBI_UV_FUSED if fused >= 2 else axs.KERNEL_JACOBI_UV, 50, cold=True)
    k1c_ms, _ = s.bench_kernel(axs.KERNEL_JACOBI_UV, 50, cold=True)
    cold_gbs, k1_cold_gbs = k_bytes / (kc_ms * 1e-3) / 1e9, k1_bytes / (k1c_ms * 1e-3) / 1e9
    log(f"cold caches: {kname} {kc_ms * 1000:.1f} us ({cold_gbs:.0f} GB/s), {k1_name} {k1c_ms * 1000:.1f} us "
        f"({k1_cold_gbs:.0f} GB/s)")

    # p' system of the last iteration -> stand-alone multigrid solve timing
    fs = m.arrays["cell_face_start"]
    mg = axs.Multigrid(fs, _slot_neighbours(m), max_levels=args.mg_levels)
    AC, AF, b = s.field("pp_ac"), s.field("pp_af"), s.field("pp_b")
    mg_ms, mg_bytes, _ = mg.bench(AC, AF, b, cycles=case.config.mg_max_iter, repeats=20)
    mg_gbs = mg_bytes / (mg_ms * 1e-3) / 1e9
    log(f"mg solve: {mg_ms:.3f} ms, {mg_gbs:.0f} GB/s (tail from level {mg.tail_info()[0]})")
    mgc_ms, _, _ = mg.bench(AC, AF, b, cycles=case.config.mg_max_iter, repeats=20, cold=True)
    mgc_gbs = mg_bytes / (mgc_ms * 1e-3) / 1e9
    log(f"mg solve, cold caches: {mgc_ms:.3f} ms, {mgc_gbs:.0f} GB/s")
    mg32c_ms, mg32c_bytes, _ = mg.bench(AC, AF, b, cycles=case.config.mg_max_iter, repeats=20,
                                        precision=axs.PREC_MIXED)
    mg32c_gbs = mg32c_bytes / (mg32c_ms * 1e-3) / 1e9
    log(f"mg solve (FP32 cycles, FP64 defect per cycle): {mg32c_ms:.3f} ms, {mg32c_gbs:.0f} GB/s")
    mg.set_mixed(axs.REFINE_SOLVE)
    mg32_ms, mg32_bytes, _ = mg.bench(AC, AF, b, cycles=case.config.mg_max_iter, repeats=20,
                                      precision=axs.PREC_MIXED)
    mg.set_mixed(axs.REFINE_CYCLE)
    mg32_gbs = mg32_bytes / (mg32_ms * 1e-3) / 1e9
    log(f"mg solve (FP32 cycles, FP64 defect per solve): {mg32_ms:.3f} ms, {mg32_gbs:.0f} GB/s")

    # the same run on the mixed-precision path: FP32 momentum sweeps and V-cycles
    # inside FP64 defect correction; fields compared with the FP64 run after the
    # same number of iterations from the same cold start
    mixed = None
    if not args.no_mixed:
        def mixed_leg(refine, coarse_direct=0, precision=axs.PREC_MIXED):
            mcase = axs.nozzle_case(mesh=m, max_iter=pre_iters, precision=precision,
                                    mg_max_levels=args.mg_levels, mixed_refine=refine, coarse_direct=coarse_direct)
            sm = axs.make_solver(mcase)
            sm.run()
            mms = sm.run_timed(args.steps)
            err = {}
            for name in ("u", "v", "p"):
                a, r = sm.field(name), s.field(name)
                err[name] = float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-300))
            return args.steps / (mms / 1000.0), mms, err, sm.stats().graph_nodes

        # the p' solve refined once (one FP64 defect, FP32 cycles with the
        # level-0 work fused across them): the default
        mit_s, mms, err, nodes = mixed_leg(axs.REFINE_SOLVE)
        log(f"mixed: {mit_s:.1f} it/s ({mit_s / it_s:.2f}x), rel L2 vs FP64 after "
            f"{pre_iters + args.steps} iterations {err}")
        # the FP64 defect before every V-cycle
        cit_s, cms, cerr, cnodes = mixed_leg(axs.REFINE_CYCLE)
        log(f"mixed, refined every cycle: {cit_s:.1f} it/s ({cit_s / it_s:.2f}x)")
        # with the FP64 direct solve of the coarsest level, its inverse formed
        # on every outer iteration and on every --direct-every-th
        dit_s, dms, derr, dnodes = mixed_leg(axs.REFINE_SOLVE, coarse_direct=1)
        log(f"mixed, direct coarsest solve: {dit_s:.1f} it/s ({dit_s / it_s:.2f}x)")
        kit_s, kms, kerr, knodes = mixed_leg(axs.REFINE_SOLVE, coarse_direct=args.direct_every)
        log(f"mixed, direct coarsest solve formed every {args.direct_every} iterations: {kit_s:.1f} it/s "
            f"({kit_s / it_s:.2f}x)")
        # the FP64 path with the same direct coarsest solve
        fit_s, fms, _, fnodes = mixed_leg(axs.REFINE_SOLVE, coarse_direct=args.direct_every, precision=axs.PREC_FP64)
        log(f"fp64, direct coarsest solve formed every {args.direct_every} iterations: {fit_s:.1f} it/s "
            f"({fit_s / it_s:.2f}x)")
        mixed = {"value": round(mit_s, 3), "unit": "outer_iter/s", "ms_per_step": round(mms / args.steps, 5),
                 "speedup_vs_fp64": round(mit_s / it_s, 3),
                 "dtype": "f32 sweeps/V-cycles, f64 defects, assembly, fluxes and residuals",
                 "refine": "one f64 defect per p' solve",
                 "graph_nodes_per_iteration": {"fp64": s.stats().graph_nodes, "mixed": nodes},
                 "rel_l2_vs_fp64": err, "iterations_compared": pre_iters + args.steps,
                 "mg_solve_ms": round(mg32_ms, 4), "mg_solve_gbs": round(mg32_gbs, 1),
                 "refined_every_cycle": {"value": round(cit_s, 3), "speedup_vs_fp64": round(cit_s / it_s, 3),
                                         "ms_per_step": round(cms / args.steps, 5), "graph_nodes_per_iteration": cnodes,
                                         "rel_l2_vs_fp64": cerr, "mg_solve_ms": round(mg32c_ms, 4),
                                         "mg_solve_gbs": round(mg32c_gbs, 1)},
                 "direct_coarsest": {"value": round(dit_s, 3), "speedup_vs_fp64": round(dit_s / it_s, 3),
                                     "ms_per_step": round(dms / args.steps, 5), "graph_nodes_per_iteration": dnodes,
                                     "note": "inverse formed every outer iteration; per iteration, its gain is in "
                                             "the outer iterations to the tolerance (to_tolerance.nozzle_1M)"},
                 "direct_coarsest_interval": {"formed_every": args.direct_every, "value": round(kit_s, 3),
                                              "speedup_vs_fp64": round(kit_s / it_s, 3),
                                              "ms_per_step": round(kms / args.steps, 5),
                                              "graph_nodes_per_iteration": knodes},
                 "fp64_direct_coarsest_interval": {"formed_every": args.direct_every, "value": round(fit_s, 3),
                                                   "speedup_vs_fp64": round(fit_s / it_s, 3),
                                                   "ms_per_step": round(fms / args.steps, 5),
                                                   "graph_nodes_per_iteration": fnodes}}

    # runs to the 1e-6 residual from a cold start: the nozzle (the metric's
    # workload), the pipe and the expansion, each FP64 and mixed, with the
    # per-field error of the mixed run against the FP64 run
    to_tol = None
    nozzle_fp64 = {}  # the nozzle's converged FP64 fields
    if not args.no_converge:
        to_tol = {
            "tolerance": 1e-6,
            "direct_formed_every": args.direct_every,
            "nozzle_1M": converge_pair(lambda **kw: axs.nozzle_case(mesh=m, mg_max_levels=args.mg_levels, **kw),
                                       args.converge_cap, args.direct_every, keep=nozzle_fp64, one_cycle=True),
            "pipe_128x32": converge_pair(lambda **kw: axs.pipe_case(**kw), args.converge_cap, args.direct_every),
            "expansion_238k": converge_pair(lambda **kw: axs.expansion_case(**kw), args.expansion_cap,
                                            args.direct_every),
        }

    # the learned initial guess (weights trained by train_guess.py on other
    # nozzles and the expansion): the nozzle to 1e-6 on the full mixed path
    # (FP32 cycles, FP64 refinement and direct coarsest solve) with the guess
    # before each outer iteration's first p' solve
    guess = None
    weights = ROOT / "axisim-software_amd" / "data" / "guess_weights.npy"
    if not args.no_guess and not args.no_converge and weights.exists():
        net = axs.Net(layers=4)
        net.set_params(np.load(weights))
        gcase = axs.nozzle_case(mesh=m, mg_max_levels=args.mg_levels, max_iter=args.converge_cap,
                                precision=axs.PREC_MIXED, coarse_direct=args.direct_every)
        sg = axs.make_solver(gcase)
        sg.set_learned_guess(net, -1)
        t = time.time()
        sg.run()
        wall = time.time() - t
        gst = sg.stats()
        nz = to_tol["nozzle_1M"]
        plain, plain32, direct = nz["fp64"], nz["mixed"], nz["mixed_direct"]
        err = {}
        for f in ("u", "v", "p"):
            ref = nozzle_fp64[f]
            err[f] = float(np.linalg.norm(sg.field(f) - ref) / max(np.linalg.norm(ref), 1e-300))
        guess = {"workload": "the 1M nozzle from a cold start to 1e-6, mixed precision (FP32 cycles, FP64 "
                             f"refinement and direct coarsest solve, its inverse formed every {args.direct_every} "
                             "outer iterations) with the learned guess",
                 "weights": "axisim-software_amd/data/guess_weights.npy (train_guess.py)",
                 "iterations": gst.iterations, "converged": bool(gst.converged), "wall_s": round(wall, 3),
                 "solve_s": round(gst.solve_ms / 1e3, 3),
                 "iterations_saved_vs_mixed_direct": direct["iterations"] - gst.iterations,
                 "wall_vs_fp64": round(gst.solve_ms / 1e3 / plain["solve_s"], 4),
                 "speedup_to_tolerance_vs_fp64": round(plain["solve_s"] / (gst.solve_ms / 1e3), 3),
                 "wall_vs_mixed_direct": round(gst.solve_ms / 1e3 / direct["solve_s"], 4),
                 "plain_fp64": {"iterations": plain["iterations"], "solve_s": plain["solve_s"]},
                 "mixed": {"iterations": plain32["iterations"], "solve_s": plain32["solve_s"]},
                 "mixed_direct": {"iterations": direct["iterations"], "solve_s": direct["solve_s"]},
                 "rel_l2_vs_fp64": err,
                 "note": "speedups at the same residual tolerance; the direct-coarsest fields are further from "
                         "the converged ones there (to_tolerance.nozzle_1M.mixed_direct_tight is the comparison "
                         "at matched accuracy)"}
        log(f"learned guess: {gst.iterations} iterations to 1e-6 in {wall:.2f} s (mixed + direct "
            f"{direct['iterations']}, mixed {plain32['iterations']}, fp64 {plain['iterations']})")
        del sg
        guess["per_solve"] = guess_per_solve(net, m, args.mg_levels)
        del net

    # batched sweep: many cases of uneven size in one solver, every launch
    # covers all; each case stops on its own residuals in every time step
    c4 = None
    if not args.no_c4:
        tc = time.time()
        cases = axs.sweep_cases(args.c4_cases)
        cells = sum(c.mesh.n_cells for c in cases)
        batch = axs.batch_case(cases, transient=1, time_scheme=axs.TIME_BDF2, dt=1e-3, t_end=2e-3,
                               max_iter=args.c4_iters, use_non_orth=1, structured_mesh=0,
                               gradient_scheme=axs.GRAD_LSQ, convection_scheme=axs.CONV_UPWIND,
                               save_key_frame_iter=args.c4_steps)
        log(f"c4: {len(cases)} cases, {cells} cells ({time.time() - tc:.1f} s to generate)")
        sb = axs.make_solver(batch, per_case=True)
        sb.run()  # two time steps: hierarchy, graph, warm caches
        c4tune = None if args.no_autotune else sb.autotune(10)
        batch.config.t_end = args.c4_steps * 1e-3
        sb.configure(batch.config)
        t2 = time.time()
        sb.run()
        wall = time.time() - t2
        sst = sb.stats()
        cups = cells * sst.iterations / (sst.solve_ms / 1e3)
        # a case's latency: the run's iterations up to the one where it stopped
        # (in the last time step) times the mean iteration time of the run
        per_it = sst.solve_ms / 1e3 / max(1, sst.iterations)
        cst = [sb.case_stats(i) for i in range(len(cases))]
        lat = np.array([k.stop_iteration * per_it for k in cst])
        active = np.array([k.iterations for k in cst])
        log(f"c4: {args.c4_steps} BDF2 steps, {sst.iterations} outer iterations in {wall:.2f} s -> "
            f"{cups:.3e} cell-updates/s; case latency p50 {np.percentile(lat, 50):.3f} s p95 "
            f"{np.percentile(lat, 95):.3f} s")
        templates = [c.info["template"] for c in cases]
        c4 = {"cases": len(cases), "cells": cells, "templates": {t: templates.count(t) for t in set(templates)},
              "time_steps": args.c4_steps, "dt": 1e-3, "max_iter_per_step": args.c4_iters,
              "per_case_convergence": True,
              "frames_captured": sb.frame_count(), "save_key_frame_iter": args.c4_steps,
              "outer_iterations": sst.iterations, "solve_s": round(sst.solve_ms / 1e3, 3),
              "cell_updates_per_s": round(cups, 1),
              "case_latency_s": {"p50": round(float(np.percentile(lat, 50)), 4),
                                 "p95": round(float(np.percentile(lat, 95)), 4),
                                 "max": round(float(lat.max()), 4),
                                 "definition": "iterations of the run until the case stopped in its last time "
                                               "step x the run's mean iteration time"},
              "case_iterations": {"p50": float(np.percentile(active, 50)), "p95": float(np.percentile(active, 95)),
                                  "min": int(active.min()), "max": int(active.max())},
              "cases_converged_last_step": int(sum(k.converged for k in cst)),
              "note": ("every case ran the per-step iteration cap in every step (its residuals, scaled by the "
                       "step's first iterations, stay above the tolerance), so each case's latency is the run's"
                       if int(active.min()) == sst.iterations else
                       "cases stop at different iterations of their time steps"),
              "mg_levels": sst.mg_levels,
              "work_split": {"sweeps_per_launch": sst.sweeps_per_launch, "sweep_block_cells": sst.sweep_block_cells,
                             "autotuned": c4tune is not None},
              "autotune_ms": None if c4tune is None else [round(c4tune[0], 4), round(c4tune[1], 4)]}
        del sb

    cpu = None
    if not args.no_cpu_baseline:
        import oracle_binding
        from oracle_binding import Oracle, set_threads

        flags = native_oracle(oracle_binding)
        cfg = case.config
        cfg.max_iter = args.cpu_iters

        def oracle_rate(threads):
            used = set_threads(threads)
            o = Oracle(m, case.bcs, cfg)
            tc = time.time()
            o.run()
            return o.iterations, time.time() - tc, used

        # the threads this process may run on (the host's share, not its total)
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        it1, dt1, _ = oracle_rate(1)
        itn, dtn, used = oracle_rate(avail)
        set_threads(1)
        what = (f"outer iterations of the same {m.n_cells}-cell nozzle case from a cold start, FP64 oracle "
                f"(oracle/oracle.cpp built for this host; OpenMP over its per-cell and per-face loops, the "
                f"multigrid transfers per coarse cell and the residual sums in fixed chunks)")
        # the pipe (128 x 32) to its 1e-6 residual on the host, one thread and all
        c1 = {}
        for th in (1, avail):
            used1 = set_threads(th)
            pc = axs.pipe_case()
            o = Oracle(pc.mesh, pc.bcs, pc.config)
            tc = time.time()
            o.run()
            d = time.time() - tc
            c1[f"threads_{used1}"] = {"iterations": o.iterations, "converged": bool(o.converged),
                                      "wall_s": round(d, 3), "value": round(o.iterations / d, 2)}
        set_threads(1)
        cpu = {"value": round(itn / dtn, 4), "unit": "outer_iter/s", "cores": used, "kind": "port",
               "sample": f"{itn} {what}, {used} threads, {dtn:.1f} s", "iterations": itn,
               "compiler_flags": flags,
               "one_thread": {"value": round(it1 / dt1, 4), "cores": 1, "sample": f"{it1} of the same, {dt1:.1f} s"},
               "c1": {"workload": "pipe 128x32 (4096 cells) from a cold start to the 1e-6 residual", **c1},
               "host_cpu": _cpu_model(), "host_threads": os.cpu_count()}
        log(f"cpu baseline: {cpu['value']} it/s on {used} threads, {cpu['one_thread']['value']} it/s on one; "
            f"pipe to 1e-6: {c1}")

    out = {
        "metric": "SIMPLE outer iter/s at 1M cells (unstructured nozzle, non-orthogonal correction, agglomeration MG)",
        "value": round(it_s, 3),
        "unit": "outer_iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Poisson-disk + 5% jitter + Delaunay nozzle mesh (seeds 3001/3002), cold start",
        "config": {
            "workload": "nozzle_1M_triangles_simple_fp64",
            "cells": m.n_cells, "faces": m.n_faces, "face_refs": int(fs[-1]),
            "mg_levels": st.mg_levels, "convection": "upwind", "gradient": "lsq", "non_orth": True,
            "momentum_sweeps": case.config.linear_max_iter, "mg_cycles": case.config.mg_max_iter,
            "cell_updates_per_s": round(it_s * m.n_cells, 1),
            "bytes_per_iteration": st.bytes_per_iteration,
            "graph_nodes_per_iteration": st.graph_nodes,
            "iteration_hbm_frac": round(st.bytes_per_iteration * it_s / 1e9 / HBM_PEAK_GBS, 4),
        },
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "frac_cold": round(cold_gbs / HBM_PEAK_GBS, 4), "us_per_launch_cold": round(kc_ms * 1000, 3),
                     "traffic": measured_traffic(kname),
                     "bytes_per_launch": k_bytes, "us_per_launch": round(k_ms * 1000, 3),
                     "sweeps_per_launch": fused, "us_per_sweep": round(k_ms * 1000 / max(1, fused), 3),
                     "one_sweep_kernel": {"kernel": k1_name, "achieved": round(k1_gbs, 1),
                                          "frac": round(k1_gbs / HBM_PEAK_GBS, 4),
                                          "frac_cold": round(k1_cold_gbs / HBM_PEAK_GBS, 4),
                                          "traffic": measured_traffic(k1_name), "bytes_per_launch": k1_bytes,
                                          "us_per_launch": round(k1_ms * 1000, 3)}},
        "mg_vcycle": {"bound": "hbm", "achieved": round(mg_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(mg_gbs / HBM_PEAK_GBS, 4), "ms_per_solve": round(mg_ms, 4),
                      "frac_cold": round(mgc_gbs / HBM_PEAK_GBS, 4), "ms_per_solve_cold": round(mgc_ms, 4),
                      "bytes_per_solve": mg_bytes, "cycles": case.config.mg_max_iter},
        "autotune": tune,
        "mixed_precision": mixed,
        "to_tolerance": to_tol,
        "learned_guess": guess,
        "batched_sweep": c4,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def guess_per_solve(net, m, mg_levels: int, iterations=(20, 200, 2000), checkpoints=(3, 30, 300, 3000)) -> dict:
    """The learned guess as the start of single p' solves: p' systems of the 1M
    nozzle captured after the given outer iterations, each solved by FP64
    V-cycles from zero and from the network's start (its guess on the guess
    level, prolonged); the relative residual (to that of x = 0) at the start and
    after 3 (the cycles SIMPLE runs), 30, 300 and 3000 cycles."""
    import numpy as np

    import axs
    import learned_guess as lg

    t0 = time.time()
    case = axs.nozzle_case(mesh=m, mg_max_levels=mg_levels)
    systems = lg.solver_systems(case, list(iterations))
    fs, fn = m.arrays["cell_face_start"], _slot_neighbours(m)
    mg = axs.Multigrid(fs, fn, max_levels=mg_levels)
    hier, k = lg.Hierarchy(mg), lg.guess_level(mg)
    k_cells = mg.level_info(k)[0]
    rows = []
    for it, (AC, AF, b) in zip(iterations, systems):
        A = lg.csr_matrix(fs, fn, AC, AF)
        starts = {"zero": np.zeros_like(b), "guess": lg.guess_start(net, hier, fs, fn, AC, AF, b, k)}
        row = {"outer_iteration": it}
        for name, x in starts.items():
            hist, done = {"0": float(f"{lg.relative_residual(A, b, x):.4g}")}, 0
            for c in checkpoints:
                x = mg.solve(AC, AF, b, x, cycles=c - done)  # continues from x: the same cycles
                done = c
                hist[str(c)] = float(f"{lg.relative_residual(A, b, x):.4g}")
            row[name] = hist
        rows.append(row)
        log(f"guess per solve, system after iteration {it}: residual after 3 / 3000 cycles: zero "
            f"{row['zero']['3']:.3g} / {row['zero']['3000']:.3g}, guess {row['guess']['3']:.3g} / "
            f"{row['guess']['3000']:.3g}")
    mg.close()
    return {"what": "relative residual of single p' solves from zero and from the network's start, by V-cycles",
            "guess_level": k, "guess_level_cells": k_cells, "systems": rows, "seconds": round(time.time() - t0, 1)}


def converge_pair(make, cap: int, every: int, keep: dict | None = None, tight_tol: float = 1e-9,
                  one_cycle: bool = False) -> dict:
    """One case from a cold start to its residual tolerance (at most `cap`
    iterations), FP64, mixed, and both with the FP64 direct solve of the
    coarsest level (its inverse formed on every `every`-th outer iteration),
    those two also to `tight_tol` when they converged: iterations, wall time
    and rate of each, and each run's fields' relative L2 error against the
    plain FP64 fields (kept in `keep` when given).  one_cycle: also the mixed
    path with the direct solve and one V-cycle per p' solve (mg_max_iter = 1)
    to both tolerances."""
    import numpy as np

    import axs

    out, fields = {}, {}
    runs = (("fp64", dict(precision=axs.PREC_FP64)), ("mixed", dict(precision=axs.PREC_MIXED)),
            ("mixed_direct", dict(precision=axs.PREC_MIXED, coarse_direct=every)),
            ("fp64_direct", dict(precision=axs.PREC_FP64, coarse_direct=every)))
    for name, kw in runs:
        case = make(max_iter=cap, **kw)
        s = axs.make_solver(case)
        t = time.time()
        s.run()
        wall = time.time() - t
        st = s.stats()
        out[name] = {"iterations": st.iterations, "converged": bool(st.converged), "cap_hit": st.iterations >= cap,
                     "wall_s": round(wall, 3), "solve_s": round(st.solve_ms / 1e3, 3),
                     "it_per_s": round(st.iterations / (st.solve_ms / 1e3), 2),
                     "last_residual": [float(f"{x:.4g}") for x in st.last_residual[:3]]}
        fields[name] = {f: s.field(f) for f in ("u", "v", "p")}
        log(f"{case.name} {name}: {st.iterations} iterations, converged {bool(st.converged)}, {wall:.2f} s")
        del s
    if keep is not None:
        keep.update(fields["fp64"])
    out["cells"] = case.mesh.n_cells
    # the residual test stops the direct-coarsest runs with their momentum
    # residuals well above FP64's (the continuity residual decides both), so
    # their fields sit further from the converged ones at the same tolerance:
    # the same path once more to a 1000x tighter tolerance, whose fields are
    # at least as close to the converged ones as FP64's at 1e-6 (on the
    # nozzle, measured against a run to 1e-10: FP64 at 1e-6 u 1.1e-5,
    # v 6.8e-5, p 4.4e-5; mixed with the direct solve at 1e-9 u 3.8e-6,
    # v 1.1e-4, p 3.6e-6; at 1e-6 u 2.0e-3, v 4.4e-2, p 1.2e-3)
    legs = [("mixed_direct", dict(precision=axs.PREC_MIXED, coarse_direct=every)),
            ("fp64_direct", dict(precision=axs.PREC_FP64, coarse_direct=every))]
    if one_cycle:
        kw = dict(precision=axs.PREC_MIXED, coarse_direct=every, mg_max_iter=1)
        case = make(max_iter=cap, **kw)
        s = axs.make_solver(case)
        s.run()
        st = s.stats()
        out["mixed_direct_1cycle"] = {"iterations": st.iterations, "converged": bool(st.converged),
                                      "solve_s": round(st.solve_ms / 1e3, 3),
                                      "it_per_s": round(st.iterations / (st.solve_ms / 1e3), 2)}
        fields["mixed_direct_1cycle"] = {f: s.field(f) for f in ("u", "v", "p")}
        log(f"{case.name} mixed_direct, one V-cycle per p' solve: {st.iterations} iterations, "
            f"converged {bool(st.converged)}, {st.solve_ms / 1e3:.2f} s")
        del s
        legs.append(("mixed_direct_1cycle", kw))
    for base, kw in legs:
        if not out[base]["converged"]:
            continue
        case = make(max_iter=cap, tol=tight_tol, **kw)
        s = axs.make_solver(case)
        t = time.time()
        s.run()
        wall = time.time() - t
        st = s.stats()
        out[f"{base}_tight"] = {
            "tolerance": tight_tol, "iterations": st.iterations, "converged": bool(st.converged),
            "cap_hit": st.iterations >= cap, "wall_s": round(wall, 3), "solve_s": round(st.solve_ms / 1e3, 3),
            "it_per_s": round(st.iterations / (st.solve_ms / 1e3), 2),
            "last_residual": [float(f"{x:.4g}") for x in st.last_residual[:3]]}
        fields[f"{base}_tight"] = {f: s.field(f) for f in ("u", "v", "p")}
        log(f"{case.name} {base} to {tight_tol:g}: {st.iterations} iterations, "
            f"converged {bool(st.converged)}, {wall:.2f} s")
        del s
    for name in ("mixed", "mixed_direct", "mixed_direct_tight", "fp64_direct", "fp64_direct_tight",
                 "mixed_direct_1cycle", "mixed_direct_1cycle_tight"):
        if name not in fields:
            continue
        out[f"{name}_rel_l2_vs_fp64"] = {
            f: float(np.linalg.norm(fields[name][f] - fields["fp64"][f]) / max(np.linalg.norm(fields["fp64"][f]), 1e-300))
            for f in ("u", "v", "p")}
        out[f"{name}_speedup"] = round(out["fp64"]["solve_s"] / out[name]["solve_s"], 3)
    return out


def native_oracle(binding) -> str:
    """Build oracle/oracle.cpp for this host's CPU (-O3 -march=native) into a
    temporary directory and load it; the flags used (the prebuilt library's
    when the compiler is missing or fails)."""
    import shutil
    import subprocess
    import tempfile

    sys.path.insert(0, str(ROOT / "axisim-software_amd"))
    from build_lib import ORACLE_FLAGS

    flags = ORACLE_FLAGS + ["-march=native"]
    if shutil.which("g++") is None:
        return " ".join(ORACLE_FLAGS) + " (prebuilt)"
    out = Path(tempfile.mkdtemp(prefix="axs_oracle_")) / "liboracle_native.so"
    r = subprocess.run(["g++", *flags, "-fPIC", "-shared", f"-I{ROOT / 'include'}", str(ROOT / "oracle" / "oracle.cpp"),
                        "-o", str(out)], capture_output=True, text=True)
    if r.returncode != 0:
        log(f"native oracle build failed, using the prebuilt one: {r.stderr[-300:]}")
        return " ".join(ORACLE_FLAGS) + " (prebuilt)"
    binding.use_library(out)
    return " ".join(flags)


TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_nozzle1M_pmc.txt"


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size), or None."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            parts = line.split()
            if len(parts) >= 2 and parts[0] == f"traffic_bytes_per_launch[{kernel}]":
                return float(parts[1])
    except OSError:
        pass
    return None


def traffic_probe() -> int:
    """The kernels the benchmark reports, launched on the 1M-cell nozzle from a
    few iterations in, then a plain copy of 512 MiB that calibrates the counters."""
    import torch

    import axs

    case = axs.nozzle_case(max_iter=5, mg_max_levels=6)
    s = axs.make_solver(case)
    s.run()
    for k in (axs.KERNEL_JACOBI_UV, axs.KERNEL_JACOBI_UV_FUSED, axs.KERNEL_MOMENTUM, axs.KERNEL_FACE_FLUX,
              axs.KERNEL_PP_COEFF):
        ms, by = s.bench_kernel(k, 10)
        log(f"kernel {k}: {ms * 1000:.1f} us, {by:.0f} algorithmic bytes")
    # the mixed path's kernels: 5 outer iterations on the same mesh
    sm = axs.make_solver(axs.nozzle_case(mesh=case.mesh, max_iter=5, mg_max_levels=6, precision=axs.PREC_MIXED))
    sm.run()
    log("mixed: 5 outer iterations")
    x = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    x.fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copy: {x.numel() * 8} bytes read and written per launch")
    return 0


def _slot_neighbours(m) -> "np.ndarray":
    import numpy as np

    a = m.arrays
    ids = a["cell_face_ids"]
    cells = np.repeat(np.arange(m.n_cells, dtype=np.int32), np.diff(a["cell_face_start"]))
    own = a["face_owner"][ids]
    nb = a["face_neighbor"][ids]
    return np.where(own == cells, nb, own).astype(np.int32)


if __name__ == "__main__":
    sys.exit(main())
