"""GPU parity: the HIP FP64 path (libaxs.so, through the C ABI) against the CPU
oracle on identical meshes and inputs.

Tolerances: the FP64 path keeps the reference's operation order per cell and
both sides are built without FMA contraction, so per-iteration arrays agree
bit for bit except where a sum is reordered (the multigrid uses deterministic
gathers in the oracle's order, so those agree too).  The asserted bound is
1e-12 relative L2 per array after a few iterations, and the north-star bound
of 1e-8 relative L2 per field at convergence with the outer-iteration count
within +-1.
"""
import numpy as np
import pytest

import axs
from oracle_binding import Oracle

pytestmark = pytest.mark.gpu

STATE_FIELDS = ["u", "v", "p", "pp", "mdot", "du", "dv", "grad_pz", "grad_pr", "u_ac", "u_b", "u_invac", "u_af",
                "v_ac", "v_b", "v_invac", "v_af", "pp_ac", "pp_b", "pp_invac", "pp_af"]


def rel_l2(a, b):
    na = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / na if na > 0 else d


def run_both(case, iters=None):
    if iters is not None:
        case.config.max_iter = iters
    s = axs.make_solver(case)
    s.run()
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    return s, o


def compare(s, o, names, tol):
    worst = {}
    for name in names:
        a, b = s.field(name), o.field(name)
        assert a.shape == b.shape, name
        assert np.all(np.isfinite(a)), name
        e = rel_l2(a, b)
        worst[name] = e
    bad = {k: v for k, v in worst.items() if v > tol}
    assert not bad, f"fields beyond {tol}: {bad}"
    return worst


def small_pipe(**kw):
    return axs.pipe_case(nz=32, nr=8, **kw)


def test_pipe_one_iteration_all_arrays(require_gpu):
    s, o = run_both(small_pipe(), iters=1)
    compare(s, o, STATE_FIELDS, 1e-13)
    assert s.stats().iterations == o.iterations == 1


def test_pipe_c1_twenty_iterations(require_gpu):
    s, o = run_both(axs.pipe_case(), iters=20)
    compare(s, o, STATE_FIELDS, 1e-12)
    for r in range(3):
        np.testing.assert_allclose(s.history(r), o.history(r), rtol=1e-10, atol=0)


def test_pipe_c1_converged(require_gpu):
    case = axs.pipe_case()
    s, o = run_both(case)
    st = s.stats()
    assert o.converged and st.converged
    assert abs(st.iterations - o.iterations) <= 1, (st.iterations, o.iterations)
    compare(s, o, ["u", "v", "p", "mdot", "grad_uz", "grad_ur", "grad_vz", "grad_vr", "grad_pz", "grad_pr"], 1e-8)
    # Hagen-Poiseuille in the developed region
    m = case.mesh
    u = s.field("u")
    z, r = m.arrays["cell_center_z"], m.arrays["cell_center_r"]
    sel = z > 0.15
    ua = 0.01 * (1 - (r[sel] / 0.01) ** 2)
    assert np.sqrt(np.mean((u[sel] - ua) ** 2)) / 0.01 < 0.005


def test_pipe_c1_gauss_seidel_multigrid(require_gpu):
    """C1 with the multi-colour Gauss-Seidel multigrid smoother (build
    extension): per-iteration state against the oracle's restatement of the same
    smoother, then convergence within one iteration of it."""
    s, o = run_both(axs.pipe_case(mg_smoother=axs.MG_SMOOTH_RBGS), iters=20)
    compare(s, o, STATE_FIELDS, 1e-12)
    case = axs.pipe_case(mg_smoother=axs.MG_SMOOTH_RBGS)
    s, o = run_both(case)
    st = s.stats()
    print("iterations GS-MG", st.iterations, "oracle", o.iterations)
    assert o.converged and st.converged
    assert abs(st.iterations - o.iterations) <= 1, (st.iterations, o.iterations)
    compare(s, o, ["u", "v", "p", "mdot"], 1e-8)


@pytest.mark.parametrize("scheme", [axs.CONV_CENTRAL, axs.CONV_SOU])
@pytest.mark.parametrize("grad", [axs.GRAD_GREEN_GAUSS, axs.GRAD_LSQ])
def test_schemes_small_pipe(require_gpu, scheme, grad):
    case = small_pipe(convection_scheme=scheme, gradient_scheme=grad)
    s, o = run_both(case, iters=8)
    compare(s, o, STATE_FIELDS, 1e-12)


def test_expansion_small(require_gpu):
    case = axs.expansion_case(nz=(12, 40), nr=(8, 12))
    s, o = run_both(case, iters=10)
    compare(s, o, STATE_FIELDS, 1e-12)


def test_nozzle_small_nonortho(require_gpu):
    mesh = axs.nozzle_mesh(spacing=4e-4)
    case = axs.nozzle_case(mesh=mesh)
    s, o = run_both(case, iters=6)
    compare(s, o, STATE_FIELDS, 1e-12)


def test_gauss_seidel_momentum(require_gpu):
    case = small_pipe(linear_solver=axs.LINEAR_GS_RB)
    s, o = run_both(case, iters=5)
    compare(s, o, STATE_FIELDS, 1e-12)
    assert s.stats().n_colors == 2


def test_no_multigrid_pressure(require_gpu):
    case = small_pipe(use_multigrid=0)
    s, o = run_both(case, iters=5)
    compare(s, o, STATE_FIELDS, 1e-12)


def test_transient_bdf2_pulsatile(require_gpu):
    case = small_pipe(transient=1, time_scheme=axs.TIME_BDF2, dt=1e-2, t_end=0.05, max_iter=4)
    case.bcs = axs.default_bcs(axs.GROUP_TYPES, {(axs.G_INLET, axs.VAR_U): (axs.BC_PULSATILE, 0.005, 0.3, 1.0)})
    s, o = run_both(case)
    assert s.stats().iterations == o.iterations
    compare(s, o, ["u", "v", "p", "mdot"], 1e-12)


def test_fully_developed_inlet_c1(require_gpu):
    """C1 with a fully-developed (parabolic) inlet profile v (1 - (r/R)^2)."""
    case = axs.pipe_case()
    u = case.bcs[axs.VAR_U]
    u.type[axs.G_INLET] = axs.BC_FULLY_DEVELOPED
    u.value[axs.G_INLET] = 0.01
    u.length = [0.0] * len(u.type)
    u.length[axs.G_INLET] = 0.01
    s, o = run_both(case, iters=30)
    compare(s, o, STATE_FIELDS, 1e-12)


def test_derived_outputs_c1(require_gpu):
    """createSolutions' host-derived fields: continuity, |V|, cell Reynolds number."""
    s, o = run_both(axs.pipe_case(), iters=25)
    compare(s, o, ["continuity", "vmag", "cell_re", "grad_uz", "grad_ur", "grad_vz", "grad_vr"], 1e-12)


def _pulsatile_pipe(t_end, dt=1e-2):
    case = small_pipe(transient=1, time_scheme=axs.TIME_BDF2, dt=dt, t_end=t_end, max_iter=3, save_key_frame_iter=1)
    case.bcs = axs.default_bcs(axs.GROUP_TYPES, {(axs.G_INLET, axs.VAR_U): (axs.BC_PULSATILE, 0.005, 0.3, 5.0)})
    return case


def test_transient_frames_and_continued_pulsatile(require_gpu):
    """Frames every step (interval 1) from the start state; a continued run
    appends frames and keeps the inlet's phase (steps at 4 dt and 5 dt)."""
    dt = 0.0625
    case = _pulsatile_pipe(3 * dt, dt)
    s, o = run_both(case)
    case.config.t_end = 2 * dt
    s.configure(case.config)
    s.run(continue_run=True)
    o.configure(case.config)
    o.run(continue_run=True)
    fs, fo = s.frames(), o.frames()
    assert [t for t, _ in fs] == [t for t, _ in fo]
    assert [round(t / dt, 9) for t, _ in fs] == [0.0, 1.0, 2.0, 3.0, 4.0, 5.0]
    for (_, a), (_, b) in zip(fs, fo):
        for k in ("u", "v", "p"):
            assert rel_l2(a[k], b[k]) <= 1e-12, k
    compare(s, o, ["u", "v", "p", "mdot", "u_old", "v_old", "u_old2", "v_old2"], 1e-12)
    # a fresh run drops the frames and starts again at t = 0
    s.run()
    assert [round(t / dt, 9) for t, _ in s.frames()] == [0.0, 1.0, 2.0]


def test_bench_kernel_leaves_the_state_alone(require_gpu):
    """Timing kernels between runs does not change what a continued run computes."""
    case = small_pipe(max_iter=5)
    s1, s2 = axs.make_solver(case), axs.make_solver(case)
    s1.run()
    s2.run()
    before = {k: s1.field(k) for k in ("u", "v", "p", "mdot", "pp")}
    kernels = [axs.KERNEL_JACOBI_UV, axs.KERNEL_MOMENTUM, axs.KERNEL_MG_SMOOTH, axs.KERNEL_FACE_FLUX,
               axs.KERNEL_PP_COEFF]
    if s1.stats().sweeps_per_launch >= 2:
        kernels.append(axs.KERNEL_JACOBI_UV_FUSED)
    for k in kernels:
        s1.bench_kernel(k, 3)
        s1.bench_kernel(k, 2, cold=True)
    for k, a in before.items():
        np.testing.assert_array_equal(s1.field(k), a, err_msg=k)
    case.config.max_iter = 2
    for s in (s1, s2):
        s.configure(case.config)
        s.run(continue_run=True)
    for k in ("u", "v", "p", "mdot", "pp"):
        np.testing.assert_array_equal(s1.field(k), s2.field(k), err_msg=k)
