"""GPU parity at the benchmark configurations' full sizes, against the FP64
oracle run on every host thread this process may use (its results do not
depend on the thread count):

  C3  the 1M-cell unstructured nozzle, non-orthogonal correction, agglomeration
      multigrid with 6 levels, 3 outer iterations: every state array <= 1e-12
  C2  the 238,717-cell graded sudden expansion, SOU + LSQ, 10 outer iterations
  C4  the 64-case pulsatile BDF2 sweep (about 10M cells) as one batch, without
      multigrid, 2 time steps: sampled cases' slices bit-exact against the
      same cases run alone, and the smallest case against the oracle
  C3 variants at full size: the colour Gauss-Seidel momentum solver with the
      multi-colour Gauss-Seidel multigrid smoother; energy and Michaelis-Menten
      species transport; BDF2 with a pulsatile inlet (observed: bit for bit)
"""
import os

import numpy as np
import pytest

import axs
from oracle_binding import Oracle, set_threads
from test_gpu_parity import STATE_FIELDS, compare

pytestmark = pytest.mark.gpu


@pytest.fixture
def oracle_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    set_threads(min(16, n))
    yield
    set_threads(1)


def _both(case):
    print(f"{case.name}: {case.mesh.n_cells} cells, {case.config.max_iter} iterations", flush=True)
    s = axs.make_solver(case)
    s.run()
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    assert s.stats().iterations == o.iterations == case.config.max_iter
    return s, o


def test_c3_nozzle_full_size(require_gpu, oracle_threads):
    case = axs.nozzle_case(mg_max_levels=6, max_iter=3)
    assert case.mesh.n_cells > 990_000
    s, o = _both(case)
    assert s.stats().mg_levels == 6
    worst = compare(s, o, STATE_FIELDS, 1e-12)
    print("C3 worst rel L2", max(worst.values()))


def test_c2_expansion_full_size(require_gpu, oracle_threads):
    case = axs.expansion_case(max_iter=10)
    assert case.mesh.n_cells == 238_717
    s, o = _both(case)
    worst = compare(s, o, STATE_FIELDS, 1e-12)
    print("C2 worst rel L2", max(worst.values()))


def _single(case, cfg):
    c = axs.Case(case.name, case.mesh, case.bcs, axs.Config.from_buffer_copy(cfg), dict(case.info))
    return c


def test_c4_sweep_batch_sampled_cases(require_gpu, oracle_threads):
    cases = axs.sweep_cases()
    cfg = dict(transient=1, time_scheme=axs.TIME_BDF2, dt=1e-3, t_end=2e-3, max_iter=3, use_multigrid=0,
               use_non_orth=1, structured_mesh=0, gradient_scheme=axs.GRAD_LSQ,
               convection_scheme=axs.CONV_UPWIND, save_key_frame_iter=100)
    batch = axs.batch_case(cases, **cfg)
    for i in range(3):
        batch.config.residual[i].tol = 0.0  # every step runs max_iter iterations
    print(f"C4 batch: {len(cases)} cases, {batch.mesh.n_cells} cells", flush=True)
    sb = axs.make_solver(batch)
    sb.run()
    assert sb.stats().iterations == 6
    sizes = [c.mesh.n_cells for c in cases]
    rng = np.random.default_rng(4004)
    samples = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), *map(int, rng.choice(len(cases), 2, replace=False))})
    fields = {k: sb.field(k) for k in ("u", "v", "p", "pp")}
    for i in samples:
        first, n = batch.info["ranges"][i]
        single = _single(cases[i], batch.config)
        print(f"case {i} ({cases[i].info['template']}, {n} cells) alone", flush=True)
        s = axs.make_solver(single)
        s.run()
        for k, a in fields.items():
            np.testing.assert_array_equal(a[first:first + n], s.field(k), err_msg=f"case {i} {k}")
        if i == int(np.argmin(sizes)):
            o = Oracle(single.mesh, single.bcs, single.config)
            o.run()
            compare(s, o, ["u", "v", "p", "pp", "mdot", "u_old", "u_old2"], 1e-12)


def test_c3_nozzle_gauss_seidel_paths(require_gpu, oracle_threads):
    """The 1M nozzle with the colour Gauss-Seidel momentum solver and the
    multi-colour Gauss-Seidel multigrid smoother, 2 iterations."""
    case = axs.nozzle_case(mg_max_levels=6, max_iter=2, linear_solver=axs.LINEAR_GS_RB,
                           mg_smoother=axs.MG_SMOOTH_RBGS)
    s, o = _both(case)
    worst = compare(s, o, STATE_FIELDS, 1e-12)
    print("C3 Gauss-Seidel worst rel L2", max(worst.values()))


def test_c3_nozzle_scalar_transport(require_gpu, oracle_threads):
    """The 1M nozzle with energy and Michaelis-Menten species transport
    (integer-exponent kinetics: no pow on either side), 2 iterations."""
    case = axs.nozzle_case(mg_max_levels=6, max_iter=2)
    sc = axs.scalar_pipe_case(nz=4, nr=2)
    case.bcs = sc.bcs  # the same four groups: inlet, outlet, wall, axis
    case.bcs[axs.VAR_U].value[axs.G_INLET] = case.info["u_in"]
    c = case.config
    c.solve_energy = c.solve_concentration = 1
    c.residual[axs.RES_TEMPERATURE].enabled = c.residual[axs.RES_CONCENTRATION].enabled = 1
    s, o = _both(case)
    worst = compare(s, o, STATE_FIELDS + ["t", "c", "wall_cw", "wall_ocr"], 1e-12)
    print("C3 scalars worst rel L2", max(worst.values()))


def test_c3_nozzle_transient_pulsatile(require_gpu, oracle_threads):
    """The 1M nozzle, BDF2 with a pulsatile inlet, 2 steps of 2 iterations."""
    case = axs.nozzle_case(mg_max_levels=6, max_iter=2, transient=1, time_scheme=axs.TIME_BDF2, dt=1e-3,
                           t_end=2e-3)
    u = case.bcs[axs.VAR_U]
    u.type[axs.G_INLET] = axs.BC_PULSATILE
    u.amplitude = [0.0] * len(u.type)
    u.frequency = [0.0] * len(u.type)
    u.amplitude[axs.G_INLET], u.frequency[axs.G_INLET] = 0.3, 50.0
    for i in range(3):
        case.config.residual[i].tol = 0.0
    print(f"{case.name}: {case.mesh.n_cells} cells, 2 BDF2 steps", flush=True)
    s = axs.make_solver(case)
    s.run()
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    assert s.stats().iterations == o.iterations == 4
    worst = compare(s, o, ["u", "v", "p", "pp", "mdot", "u_old", "u_old2", "v_old", "v_old2"], 1e-12)
    print("C3 transient worst rel L2", max(worst.values()))
