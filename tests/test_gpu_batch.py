"""Batched runs (build extension): many cases packed into one union mesh, one
solver and one launch per kernel.  Every per-cell and per-face step treats each
case as it would alone, so without the multigrid (whose coarsest levels are
shared) a batch reproduces the individual runs bit for bit."""
import numpy as np
import pytest

import axs

pytestmark = pytest.mark.gpu

COMMON = dict(gradient_scheme=axs.GRAD_LSQ, convection_scheme=axs.CONV_UPWIND, use_non_orth=1, structured_mesh=0)


def three_cases(**cfg):
    kw = dict(COMMON, **cfg)
    return [
        axs.pipe_case(nz=16, nr=8, **kw),
        axs.pipe_case(nz=24, nr=6, L=0.1, R=0.005, u_in=0.01, **kw),
        axs.nozzle_case(mesh=axs.nozzle_mesh(axs.nozzle_spacing_for(3000)), re_throat=100, **kw),
    ]


def run(case):
    s = axs.make_solver(case)
    s.run()
    return s


def test_batch_without_multigrid_is_bit_exact(require_gpu):
    cfg = dict(use_multigrid=0, max_iter=12, tol=0.0)
    cases = three_cases(**cfg)
    batch = axs.batch_case(cases)
    sb = run(batch)
    assert sb.stats().iterations == 12
    for c, (first, n) in zip(cases, batch.info["ranges"]):
        s = run(c)
        for name in ("u", "v", "p", "pp"):
            np.testing.assert_array_equal(sb.field(name)[first:first + n], s.field(name), err_msg=f"{c.name} {name}")


def test_batch_transient_pulsatile_runs_every_case(require_gpu):
    """BDF2 with a pulsatile inlet per case (different amplitudes and
    frequencies): without the multigrid, bit-exact against each case alone."""
    cfg = dict(use_multigrid=0, transient=1, time_scheme=axs.TIME_BDF2, dt=1e-3, t_end=5e-3, max_iter=4,
               tol=1e-12)
    cases = three_cases(**cfg)
    for i, c in enumerate(cases):
        u = c.bcs[axs.VAR_U]
        u.type[axs.G_INLET] = axs.BC_PULSATILE
        u.amplitude = [0.0] * len(u.type)
        u.frequency = [0.0] * len(u.type)
        u.amplitude[axs.G_INLET] = 0.3
        u.frequency[axs.G_INLET] = 1.0 + 50.0 * i
    batch = axs.batch_case(cases)
    sb = run(batch)
    for c, (first, n) in zip(cases, batch.info["ranges"]):
        s = run(c)
        for name in ("u", "v", "p"):
            np.testing.assert_array_equal(sb.field(name)[first:first + n], s.field(name), err_msg=f"{c.name} {name}")


def test_batch_with_multigrid_reaches_the_same_steady_state(require_gpu):
    cases = three_cases(max_iter=20000, tol=1e-10)
    batch = axs.batch_case(cases)
    sb = run(batch)
    assert sb.stats().converged
    for c, (first, n) in zip(cases, batch.info["ranges"]):
        s = run(c)
        assert s.stats().converged
        for name in ("u", "p"):
            a, r = sb.field(name)[first:first + n], s.field(name)
            err = np.linalg.norm(a - r) / np.linalg.norm(r)
            print(c.name, name, "rel err", err, "iterations", sb.stats().iterations, s.stats().iterations)
            assert err < 1e-5


def _alone(case, cfg):
    c = axs.Case(case.name, case.mesh, case.bcs, axs.Config.from_buffer_copy(cfg), dict(case.info))
    s = axs.make_solver(c)
    s.run()
    return s


def test_per_case_convergence_matches_each_case_alone(require_gpu):
    """Each case of a batch stops on its own residuals: without multigrid it
    runs the iterations it runs alone and ends with the same fields, bit for bit."""
    kw = dict(COMMON, use_multigrid=0, max_iter=6000, tol=1e-4)
    cases = [axs.pipe_case(nz=12, nr=4, **kw), axs.pipe_case(nz=16, nr=3, L=0.1, R=0.005, u_in=0.01, **kw),
             axs.pipe_case(nz=10, nr=6, L=0.05, R=0.004, u_in=0.003, **kw)]
    batch = axs.batch_case(cases)
    sb = axs.make_solver(batch, per_case=True)
    sb.run()
    assert sb.stats().converged
    iters = []
    for i, (c, (first, n)) in enumerate(zip(cases, batch.info["ranges"])):
        s = _alone(c, batch.config)
        st = sb.case_stats(i)
        iters.append((st.iterations, s.stats().iterations))
        assert st.converged and st.iterations == s.stats().iterations, (i, st.iterations, s.stats().iterations)
        assert st.stop_iteration == st.iterations  # a steady run from iteration 0
        np.testing.assert_allclose(np.array(st.last_residual), s.history(0)[-1:].tolist() + [
            s.history(r)[-1] for r in range(1, axs.RES_COUNT)], rtol=0, atol=0)
        for name in ("u", "v", "p", "pp"):
            np.testing.assert_array_equal(sb.field(name)[first:first + n], s.field(name), err_msg=f"case {i} {name}")
        f0, f1 = batch.info["face_starts"][i], batch.info["face_starts"][i + 1]
        np.testing.assert_array_equal(sb.field("mdot")[f0:f1], s.field("mdot"))
    print("per-case iterations (batch, alone):", iters)
    assert len({a for a, _ in iters}) > 1  # the cases do stop at different iterations
    assert sb.stats().iterations == max(a for a, _ in iters)


def test_per_case_convergence_with_multigrid(require_gpu):
    """With multigrid the union's hierarchy coarsens further than a small case's
    own (it stops at 64 cells over all cases), so a case's iteration path
    differs from its run alone; each still stops on its own residuals at the
    same solution."""
    cases = three_cases(max_iter=4000, tol=1e-5)
    batch = axs.batch_case(cases)
    sb = axs.make_solver(batch, per_case=True)
    sb.run()
    for i, (c, (first, n)) in enumerate(zip(cases, batch.info["ranges"])):
        s = _alone(c, batch.config)
        st = sb.case_stats(i)
        assert st.converged and s.stats().converged
        print(f"case {i}: {st.iterations} iterations in the batch, {s.stats().iterations} alone")
        assert max(st.last_residual[:3]) <= 1e-5
        # both stopped at residuals below 1e-5: the fields agree to that order,
        # measured on the velocity scale (the radial velocity is ~1e-3 of it)
        uscale = np.linalg.norm(s.field("u"))
        for name, scale in (("u", uscale), ("v", uscale), ("p", np.linalg.norm(s.field("p")))):
            a, b = sb.field(name)[first:first + n], s.field(name)
            assert np.linalg.norm(a - b) <= 1e-5 * scale, (i, name)


def test_per_case_transient_steps(require_gpu):
    cfg = dict(use_multigrid=0, transient=1, time_scheme=axs.TIME_BDF2, dt=0.0625, t_end=0.25, max_iter=40,
               tol=1e-4)
    cases = three_cases(**cfg)
    for c in cases:
        u = c.bcs[axs.VAR_U]
        u.type[axs.G_INLET] = axs.BC_PULSATILE
        u.amplitude = [0.0] * len(u.type)
        u.frequency = [0.0] * len(u.type)
        u.amplitude[axs.G_INLET], u.frequency[axs.G_INLET] = 0.3, 2.0
    batch = axs.batch_case(cases)
    sb = axs.make_solver(batch, per_case=True)
    sb.run()
    for i, (c, (first, n)) in enumerate(zip(cases, batch.info["ranges"])):
        s = _alone(c, batch.config)
        assert sb.case_stats(i).iterations == s.stats().iterations, i
        # it stopped within the last time step
        assert sb.stats().iterations - 40 < sb.case_stats(i).stop_iteration <= sb.stats().iterations
        for name in ("u", "v", "p", "u_old", "u_old2"):
            np.testing.assert_array_equal(sb.field(name)[first:first + n], s.field(name), err_msg=f"case {i} {name}")
