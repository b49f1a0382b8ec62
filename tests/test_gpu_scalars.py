"""GPU parity of the scalar transport (energy, concentration, Michaelis-Menten
and Hill reactive walls) against the FP64 oracle on the same inputs.

The kinetics with integer exponents (Michaelis-Menten, Hill n = 1) evaluate no
pow() and stay bit-exact; a Hill law with n = 2 and substrate inhibition calls
pow() on both sides (the device library's and the C library's), whose last bits
may differ, so those runs are held to 1e-11.
"""
import numpy as np
import pytest

import axs
from oracle_binding import Oracle
from test_gpu_parity import STATE_FIELDS, compare

pytestmark = pytest.mark.gpu

SCALAR_FIELDS = ["t", "c", "t_ac", "t_b", "t_invac", "t_af", "c_ac", "c_b", "c_invac", "c_af", "wall_cw", "wall_ocr"]


def run_both(case):
    s = axs.make_solver(case)
    s.run()
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    assert s.stats().iterations == o.iterations
    return s, o


def test_scalar_pipe_michaelis_menten(require_gpu):
    case = axs.scalar_pipe_case(nz=32, nr=8, max_iter=12)
    s, o = run_both(case)
    compare(s, o, STATE_FIELDS + SCALAR_FIELDS, 1e-13)
    for r in range(axs.RES_COUNT):
        np.testing.assert_allclose(s.history(r), o.history(r), rtol=1e-10, atol=0)
    np.testing.assert_allclose(s.wall_diagnostic(), o.wall_diagnostic(), rtol=1e-12)


@pytest.mark.parametrize("scheme", [axs.CONV_CENTRAL, axs.CONV_SOU])
def test_scalar_nozzle_hill_inhibition(require_gpu, scheme):
    mesh = axs.nozzle_mesh(spacing=4e-4)
    case = axs.nozzle_case(mesh=mesh, max_iter=8, convection_scheme=scheme)
    sc = axs.scalar_pipe_case(nz=4, nr=2, wall_kinetics=axs.BC_HILL, hill_n=2.0, inhib_v2=0.5, vmax=3e-6)
    case.bcs = sc.bcs  # same four groups: inlet, outlet, wall, axis
    case.bcs[axs.VAR_U].value[axs.G_INLET] = case.info["u_in"]
    c = case.config
    c.solve_energy = c.solve_concentration = 1
    c.residual[axs.RES_TEMPERATURE].enabled = c.residual[axs.RES_CONCENTRATION].enabled = 1
    s, o = run_both(case)
    compare(s, o, STATE_FIELDS + SCALAR_FIELDS + ["grad_tz", "grad_tr", "grad_cz", "grad_cr"], 1e-11)


def test_scalar_transient_bdf2_frames(require_gpu):
    case = axs.scalar_pipe_case(nz=24, nr=6, transient=1, time_scheme=axs.TIME_BDF2, dt=0.0625, t_end=0.25,
                                max_iter=3, save_key_frame_iter=1)
    s, o = run_both(case)
    compare(s, o, ["u", "v", "p", "t", "c", "t_old", "t_old2", "c_old", "c_old2"], 1e-13)
    fs, fo = s.frames(), o.frames()
    assert [t for t, _ in fs] == [t for t, _ in fo] and len(fs) == 5
    for (_, a), (_, b) in zip(fs, fo):
        for k in ("u", "t", "c"):
            np.testing.assert_allclose(a[k], b[k], rtol=1e-13, atol=0, err_msg=k)


def test_scalar_pipe_converged_with_colour_gauss_seidel(require_gpu):
    case = axs.scalar_pipe_case(nz=32, nr=8, linear_solver=axs.LINEAR_GS_RB)
    for i in range(axs.RES_COUNT):
        case.config.residual[i].tol = 1e-6
    s, o = run_both(case)
    assert s.stats().converged and o.converged
    compare(s, o, ["u", "v", "p", "t", "c", "grad_tz", "grad_cr"], 1e-8)
    # the concentration decreases downstream at the wall, the temperature tends to the wall value
    a = case.mesh.arrays
    c = s.field("c")
    near = np.argsort(-a["cell_center_r"])[:32]
    assert c[near].min() < 0.2


def test_scalar_cases_batched_per_case(require_gpu):
    """Scalar cases of different sizes and wall kinetics in one batch, each on
    its own residuals (five of them): without multigrid each stops at the
    iteration it stops at alone, with the same fields bit for bit."""
    kw = dict(use_multigrid=0, max_iter=6000, tol=1e-4)
    cases = [axs.scalar_pipe_case(nz=12, nr=4, **kw),
             axs.scalar_pipe_case(nz=16, nr=3, L=0.1, R=0.005, u_in=0.01, vmax=2e-6, **kw),
             axs.scalar_pipe_case(nz=10, nr=6, wall_kinetics=axs.BC_HILL, t_in=320.0, c_in=0.3, **kw)]
    batch = axs.batch_case(cases)
    sb = axs.make_solver(batch, per_case=True)
    sb.run()
    iters = []
    for i, (c, (first, n)) in enumerate(zip(cases, batch.info["ranges"])):
        alone = axs.Case(c.name, c.mesh, c.bcs, axs.Config.from_buffer_copy(batch.config), dict(c.info))
        s = axs.make_solver(alone)
        s.run()
        st = sb.case_stats(i)
        iters.append((st.iterations, s.stats().iterations))
        assert st.iterations == s.stats().iterations and st.converged == s.stats().converged, (i, iters[-1])
        for name in ("u", "v", "p", "t", "c"):
            np.testing.assert_array_equal(sb.field(name)[first:first + n], s.field(name), err_msg=f"case {i} {name}")
        f0, f1 = batch.info["face_starts"][i], batch.info["face_starts"][i + 1]
        np.testing.assert_array_equal(sb.field("wall_cw")[f0:f1], s.field("wall_cw"))
    print("per-case iterations (batch, alone):", iters)
