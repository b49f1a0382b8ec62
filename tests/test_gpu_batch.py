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
