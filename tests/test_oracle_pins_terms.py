"""Pins of the oracle's face terms (oracle/oracle.cpp) by identities that hold
independently of how the oracle writes them: each test feeds one oracle kernel
crafted fields on a real mesh and compares with a closed form computed here
from the mesh geometry alone.

  central differencing (K3)        A phi - b = sum_f F_f phi(x_f) - phi_P sum_f F_f for a
                                   linear phi on a uniform grid (face value = midpoint value)
  second-order upwind (K3)         the same identity with phi_f = phi(x_f) inside the
                                   limiter bounds; on a crafted step the face value clips to
                                   the pair's max downstream of the upwind cell, min upstream
  non-orthogonal diffusion (K3)    b - A phi = sum_f Gamma A_f n_f . G for a linear phi with
                                   its exact gradient G on jittered triangles
  p' right-hand side cross term    b - A p' = sum_f rho D A_f n_f . G (K12 + K14, same field)
  mass-flux update cross term      mdot_new - mdot_old = -rho D A_f n_f . G (K16)
  Rhie-Chow (K13)                  linear p: mdot = rho A u(x_f).n; cubic p = z^3 on a uniform
                                   grid: the correction is the classic third-difference term
                                   rho A D h^2/2 n_z; a Dirichlet-p outlet face with p linear and
                                   zero at the face: mdot = rho A u_P
  BDF1 / BDF2 (K5)                 BDF1 exact for linear-in-time, BDF2 exact for quadratic,
                                   BDF1 not exact for quadratic
  pulsatile valueAtTime            U (1 + A sin(2 pi f t)) at t = 0, 1/12f, 1/4f, 3/4f
  fully-developed inlet            face value v (1 - (r_f/L)^2); its transport row equals a
                                   Dirichlet row with that value
  Neumann face value               phi_P + g d_PB with d_PB = |(x_f - x_P) . n|
  old-level rotation (BDF2)        after k steps u_old = the frame of step k-1, u_old2 of k-2
"""
import math

import numpy as np
import pytest

import axs
from oracle_binding import Oracle


def cells_faces(mesh):
    a = mesh.arrays
    fs, ids = a["cell_face_start"], a["cell_face_ids"]
    return fs, ids, a["face_owner"], a["face_neighbor"]


def interior_cells(mesh):
    fs, ids, own, nb = cells_faces(mesh)
    return np.array([np.all(nb[ids[fs[c]:fs[c + 1]]] >= 0) for c in range(mesh.n_cells)])


def row_apply(mesh, AC, AF, phi):
    """(A phi)_n = AC_n phi_n + sum_k AF_k phi_nb(k) in the cell's slot order."""
    fs, ids, own, nb = cells_faces(mesh)
    out = AC * phi
    for c in range(mesh.n_cells):
        for k in range(fs[c], fs[c + 1]):
            f = ids[k]
            o = nb[f] if own[f] == c else own[f]
            if nb[f] >= 0:
                out[c] += AF[k] * phi[o]
    return out


def outward(mesh, c, f):
    a = mesh.arrays
    s = 1.0 if a["face_owner"][f] == c else -1.0
    return s * a["face_normal_z"][f], s * a["face_normal_r"][f]


@pytest.fixture(scope="module")
def quad():
    case = axs.pipe_case(nz=12, nr=6)
    return case, Oracle(case.mesh, case.bcs, case.config)


@pytest.fixture(scope="module")
def tri():
    case = axs.nozzle_case(spacing=6e-4)
    return case, Oracle(case.mesh, case.bcs, case.config)


def _convective_identity(mesh, AC, AF, b, phi, mdot, face_value):
    """A phi - b = sum_f F_f phi_f - phi_P sum_f F_f on interior cells."""
    fs, ids, own, nb = cells_faces(mesh)
    lhs = row_apply(mesh, AC, AF, phi) - b
    inner = interior_cells(mesh)
    for c in np.where(inner)[0]:
        F = np.array([mdot[f] if own[f] == c else -mdot[f] for f in ids[fs[c]:fs[c + 1]]])
        exp = np.sum(F * face_value[ids[fs[c]:fs[c + 1]]]) - phi[c] * F.sum()
        assert abs(lhs[c] - exp) <= 1e-12 * (np.abs(F).sum() * np.abs(phi).max())


@pytest.mark.parametrize("scheme", [axs.CONV_CENTRAL, axs.CONV_SOU])
def test_higher_order_face_value_linear_field(quad, scheme):
    case, o = quad
    a = case.mesh.arrays
    rng = np.random.default_rng(11)
    phi = 2.0 + 30.0 * a["cell_center_z"] - 50.0 * a["cell_center_r"]
    gz = np.full(case.mesh.n_cells, 30.0)
    gr = np.full(case.mesh.n_cells, -50.0)
    mdot = rng.uniform(-1.0, 1.0, case.mesh.n_faces)
    AC, AF, b = o.transport(axs.VAR_U, phi, gz, gr, mdot, 0, 0.0, 1, scheme)
    phi_f = 2.0 + 30.0 * a["face_center_z"] - 50.0 * a["face_center_r"]
    _convective_identity(case.mesh, AC, AF, b, phi, mdot, phi_f)
    # and upwind alone is not that identity (the test can fail)
    AC0, AF0, b0 = o.transport(axs.VAR_U, phi, gz, gr, mdot, 0, 0.0, 1, axs.CONV_UPWIND)
    with pytest.raises(AssertionError):
        _convective_identity(case.mesh, AC0, AF0, b0, phi, mdot, phi_f)


def test_sou_limiter_clips_on_a_step(quad):
    case, o = quad
    a = case.mesh.arrays
    zc = a["cell_center_z"]
    z0 = np.median(zc)
    phi = np.where(zc >= z0, 1.0, 0.0)
    h = 0.2 / 12
    gz = np.full(case.mesh.n_cells, 100.0 / h)   # far steeper than the step: always overshoots
    gr = np.zeros(case.mesh.n_cells)
    rng = np.random.default_rng(5)
    mdot = rng.choice([-1.0, 1.0], case.mesh.n_faces) * rng.uniform(0.5, 1.0, case.mesh.n_faces)
    AC, AF, b = o.transport(axs.VAR_U, phi, gz, gr, mdot, 0, 0.0, 1, axs.CONV_SOU)
    own, nb = a["face_owner"], a["face_neighbor"]
    fz = a["face_center_z"]
    face = np.zeros(case.mesh.n_faces)
    for f in range(case.mesh.n_faces):
        if nb[f] < 0:
            continue
        up = own[f] if mdot[f] > 0 else nb[f]
        pair = (phi[own[f]], phi[nb[f]])
        # x_f - x_U . G: +50 downstream in z of the upwind cell, -50 upstream, 0 radially
        if abs(fz[f] - zc[up]) < 1e-12:
            face[f] = phi[up]
        else:
            face[f] = max(pair) if fz[f] > zc[up] else min(pair)
    _convective_identity(case.mesh, AC, AF, b, phi, mdot, face)


def test_non_orthogonal_diffusion_exact_for_linear_field(tri):
    case, o = tri
    m = case.mesh
    a = m.arrays
    G = (4.0, -9.0)
    phi = 1.0 + G[0] * a["cell_center_z"] + G[1] * a["cell_center_r"]
    gz, gr = np.full(m.n_cells, G[0]), np.full(m.n_cells, G[1])
    gamma = 1.7e-3
    fs, ids, own, nb = cells_faces(m)
    inner = interior_cells(m)

    def defect(non_ortho):
        AC, AF, b = o.transport(axs.VAR_U, phi, gz, gr, None, non_ortho, gamma, 0, axs.CONV_UPWIND)
        return b - row_apply(m, AC, AF, phi)

    exp = np.zeros(m.n_cells)
    for c in range(m.n_cells):
        for f in ids[fs[c]:fs[c + 1]]:
            nz, nr = outward(m, c, f)
            exp[c] += gamma * a["face_area"][f] * (nz * G[0] + nr * G[1])
    d1 = defect(1)
    scale = np.abs(exp[inner]).max()
    np.testing.assert_allclose(d1[inner], exp[inner], rtol=0, atol=1e-9 * scale)
    # without the cross term the jittered mesh leaves an error
    assert np.abs(defect(0)[inner] - exp[inner]).max() > 1e-3 * scale


def test_pp_rhs_and_mass_flux_cross_terms(tri):
    case, o = tri
    m = case.mesh
    a = m.arrays
    rho, D = 1000.0, 2.5e-6
    DU = np.full(m.n_cells, D)
    G = (-3.0, 7.0)
    pp = 0.5 + G[0] * a["cell_center_z"] + G[1] * a["cell_center_r"]
    gz, gr = np.full(m.n_cells, G[0]), np.full(m.n_cells, G[1])
    AC, AF = o.pp_coeff(rho, DU, DU)
    b = o.pp_rhs(rho, np.zeros(m.n_faces), DU, DU, gz, gr, 1)
    fs, ids, own, nb = cells_faces(m)
    inner = interior_cells(m)
    exp = np.zeros(m.n_cells)
    for c in range(m.n_cells):
        for f in ids[fs[c]:fs[c + 1]]:
            nz, nr = outward(m, c, f)
            exp[c] += rho * D * a["face_area"][f] * (nz * G[0] + nr * G[1])
    got = b - row_apply(m, AC, AF, pp)
    scale = np.abs(exp[inner]).max()
    np.testing.assert_allclose(got[inner], exp[inner], rtol=0, atol=1e-9 * scale)
    # updateMassFlux with the cross term: the corrected flux drops by rho D A n.G
    mdot0 = np.random.default_rng(2).uniform(-1, 1, m.n_faces)
    mdot1 = o.mass_flux_update(rho, mdot0, pp, DU, DU, gz, gr, 1)
    inner_f = (nb >= 0) & (a["face_area"] > 1e-30)
    dn = rho * D * a["face_area"] * (a["face_normal_z"] * G[0] + a["face_normal_r"] * G[1])
    np.testing.assert_allclose((mdot1 - mdot0)[inner_f], -dn[inner_f], rtol=0, atol=1e-9 * np.abs(dn).max())
    # and the orthogonal part alone does not satisfy it on this mesh
    mdot2 = o.mass_flux_update(rho, mdot0, pp, DU, DU, gz, gr, 0)
    assert np.abs((mdot2 - mdot0)[inner_f] + dn[inner_f]).max() > 1e-3 * np.abs(dn).max()


def test_rhie_chow_linear_pressure_and_outlet(quad):
    case, o = quad
    m = case.mesh
    a = m.arrays
    rho, D, U, L = 1000.0, 3e-6, 0.02, 0.2
    G = 4.0
    z, r = a["cell_center_z"], a["cell_center_r"]
    u = U + 0.5 * z - 2.0 * r
    v = 0.25 * z + 1.5 * r
    p = G * (z - L)  # zero at the outlet plane
    ones = np.ones(m.n_cells)
    md = o.face_flux(rho, u, v, p, G * ones, 0 * ones, D * ones, D * ones)
    nb = a["face_neighbor"]
    A = a["face_area"]
    fz, fr = a["face_center_z"], a["face_center_r"]
    uf = U + 0.5 * fz - 2.0 * fr
    vf = 0.25 * fz + 1.5 * fr
    inner = (nb >= 0) & (A > 1e-30)
    exp = rho * A * (uf * a["face_normal_z"] + vf * a["face_normal_r"])
    np.testing.assert_allclose(md[inner], exp[inner], rtol=1e-12, atol=1e-15)
    assert np.all(md[A <= 1e-30] == 0.0)
    # outlet (Dirichlet p = 0, u Neumann 0): face velocity u_P, compact = cell gradient
    outlet = (nb < 0) & (a["face_group"] == axs.G_OUTLET)
    own = a["face_owner"]
    np.testing.assert_allclose(md[outlet], (rho * A * u[own])[outlet], rtol=1e-12)


def test_rhie_chow_third_difference_for_cubic_pressure(quad):
    case, o = quad
    m = case.mesh
    a = m.arrays
    rho, D = 1000.0, 5e-6
    h = 0.2 / 12
    z = a["cell_center_z"]
    zero = np.zeros(m.n_cells)
    p = z ** 3
    md = o.face_flux(rho, zero, zero, p, 3.0 * z ** 2, zero, np.full(m.n_cells, D), np.full(m.n_cells, D))
    nb, A, nz = a["face_neighbor"], a["face_area"], a["face_normal_z"]
    axial = (nb >= 0) & (np.abs(np.abs(nz) - 1.0) < 1e-12)
    radial = (nb >= 0) & (np.abs(nz) < 1e-12) & (A > 1e-30)
    np.testing.assert_allclose(md[axial], (rho * A * D * h * h / 2.0 * nz)[axial], rtol=1e-6)
    np.testing.assert_allclose(md[radial], 0.0, atol=1e-7 * np.abs(md[axial]).max())


def test_bdf_exact_for_polynomials_in_time(quad):
    case, o = quad
    vol = case.mesh.arrays["cell_volume"]
    cap, dt, t = 1000.0, 1e-3, 0.25
    c0, c1, c2 = 0.3, -2.0, 40.0
    y = lambda s: c0 + c1 * s + c2 * s * s
    dy = lambda s: c1 + 2 * c2 * s
    n = case.mesh.n_cells
    old, old2 = np.full(n, y(t)), np.full(n, y(t - dt))
    # BDF2 on a quadratic: x = (b + cap V y'(t+dt)) / AC = y(t+dt) exactly
    AC, b = o.transient(old, old2, cap, dt)
    np.testing.assert_allclose(AC, 1.5 * cap * vol / dt, rtol=1e-15)
    np.testing.assert_allclose((b + cap * vol * dy(t + dt)) / AC, y(t + dt), rtol=1e-12)
    # BDF1 on a linear function: exact; on the quadratic: off by c2 dt^2
    AC1, b1 = o.transient(old, None, cap, dt)
    np.testing.assert_allclose(AC1, cap * vol / dt, rtol=1e-15)
    lin = lambda s: c0 + c1 * s
    AC1, b1 = o.transient(np.full(n, lin(t)), None, cap, dt)
    np.testing.assert_allclose((b1 + cap * vol * c1) / AC1, lin(t + dt), rtol=1e-12)
    AC1, b1 = o.transient(old, None, cap, dt)
    err = (b1 + cap * vol * dy(t + dt)) / AC1 - y(t + dt)
    np.testing.assert_allclose(err, c2 * dt * dt, rtol=1e-6)


def test_pulsatile_value_at_time():
    case = axs.pipe_case(nz=8, nr=4)
    U, A, f = 0.05, 0.3, 1.25
    u = case.bcs[axs.VAR_U]
    u.type[axs.G_INLET] = axs.BC_PULSATILE
    u.value[axs.G_INLET] = U
    u.amplitude = [0.0] * 4
    u.frequency = [0.0] * 4
    u.amplitude[axs.G_INLET], u.frequency[axs.G_INLET] = A, f
    o = Oracle(case.mesh, case.bcs, case.config)
    for t, exp in ((0.0, U), (1 / (12 * f), U * (1 + A / 2)), (1 / (4 * f), U * (1 + A)),
                   (3 / (4 * f), U * (1 - A)), (2.0 / f, U)):
        assert math.isclose(o.bc_value(axs.VAR_U, axs.G_INLET, t), exp, rel_tol=1e-12, abs_tol=1e-16)
    # a steady group keeps its value
    assert o.bc_value(axs.VAR_U, axs.G_WALL, 0.3) == 0.0


def test_fully_developed_inlet_face_values_and_row():
    R, v = 0.01, 0.02
    case = axs.pipe_case(nz=6, nr=5, R=R)
    u = case.bcs[axs.VAR_U]
    u.type[axs.G_INLET] = axs.BC_FULLY_DEVELOPED
    u.value[axs.G_INLET] = v
    u.length = [0.0] * 4
    u.length[axs.G_INLET] = R
    o = Oracle(case.mesh, case.bcs, case.config)
    a = case.mesh.arrays
    phi = np.random.default_rng(1).uniform(size=case.mesh.n_cells)
    fv = o.face_values(axs.VAR_U, phi)
    inlet = (a["face_neighbor"] < 0) & (a["face_group"] == axs.G_INLET)
    np.testing.assert_allclose(fv[inlet], v * (1 - (a["face_center_r"][inlet] / R) ** 2), rtol=1e-14)
    # one radial cell: the inlet row equals a Dirichlet row with the parabola's face value
    one = axs.pipe_case(nz=4, nr=1, R=R)
    for var in (axs.VAR_U,):
        bc = one.bcs[var]
        bc.length = [0.0] * 4
        bc.length[axs.G_INLET] = R
    fd = one.bcs[axs.VAR_U]
    fd.type[axs.G_INLET], fd.value[axs.G_INLET] = axs.BC_FULLY_DEVELOPED, v
    o_fd = Oracle(one.mesh, one.bcs, one.config)
    dirich = axs.pipe_case(nz=4, nr=1, R=R)
    dirich.bcs[axs.VAR_U].value[axs.G_INLET] = v * (1 - 0.5 ** 2)
    o_d = Oracle(dirich.mesh, dirich.bcs, dirich.config)
    n = one.mesh.n_cells
    phi = np.linspace(0.01, 0.02, n)
    mdot = -np.abs(np.random.default_rng(4).uniform(0.1, 1.0, one.mesh.n_faces))
    for conv in (0, 1):
        r1 = o_fd.transport(axs.VAR_U, phi, None, None, mdot, 0, 1e-3, conv, axs.CONV_UPWIND)
        r2 = o_d.transport(axs.VAR_U, phi, None, None, mdot, 0, 1e-3, conv, axs.CONV_UPWIND)
        for x, y in zip(r1, r2):
            np.testing.assert_allclose(x, y, rtol=1e-14, atol=0)


def test_neumann_face_value(quad):
    case = axs.pipe_case(nz=12, nr=6)
    g = 2.5
    case.bcs[axs.VAR_U].value[axs.G_OUTLET] = g  # u is Neumann at the outlet
    o = Oracle(case.mesh, case.bcs, case.config)
    a = case.mesh.arrays
    phi = np.random.default_rng(9).uniform(size=case.mesh.n_cells)
    fv = o.face_values(axs.VAR_U, phi)
    out = np.where((a["face_neighbor"] < 0) & (a["face_group"] == axs.G_OUTLET))[0]
    own = a["face_owner"][out]
    dpb = np.abs((a["face_center_z"][out] - a["cell_center_z"][own]) * a["face_normal_z"][out]
                 + (a["face_center_r"][out] - a["cell_center_r"][own]) * a["face_normal_r"][out])
    np.testing.assert_allclose(fv[out], phi[own] + g * dpb, rtol=1e-14)


def test_bdf2_old_level_rotation_and_frames():
    case = axs.pipe_case(nz=16, nr=6, transient=1, time_scheme=axs.TIME_BDF2, dt=1e-2, t_end=3e-2,
                         max_iter=4, save_key_frame_iter=1)
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    fr = o.frames()
    # the start state, then every step (interval 1)
    assert [round(t / 1e-2, 9) for t, _ in fr] == [0.0, 1.0, 2.0, 3.0]
    np.testing.assert_array_equal(fr[0][1]["u"], 0.0)
    np.testing.assert_array_equal(o.field("u"), fr[3][1]["u"])
    np.testing.assert_array_equal(o.field("u_old"), fr[2][1]["u"])
    np.testing.assert_array_equal(o.field("u_old2"), fr[1][1]["u"])
    np.testing.assert_array_equal(o.field("v_old2"), fr[1][1]["v"])


def test_frames_interval_last_step_and_cap():
    case = axs.pipe_case(nz=16, nr=6, transient=1, dt=1e-2, t_end=5e-2, max_iter=2, save_key_frame_iter=2)
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    # steps 0, 2, 4 by the interval (4 is also the last) after the start state
    assert [round(t / 1e-2, 9) for t, _ in o.frames()] == [0.0, 1.0, 3.0, 5.0]
    case.config.max_time_frames = 2
    o2 = Oracle(case.mesh, case.bcs, case.config)
    o2.run()
    assert [round(t / 1e-2, 9) for t, _ in o2.frames()] == [0.0, 1.0]


def test_continued_pulsatile_run_keeps_phase():
    dt, f, A, U = 0.0625, 1.0, 0.3, 0.01
    case = axs.pipe_case(nz=16, nr=6, transient=1, dt=dt, t_end=3 * dt, max_iter=3, save_key_frame_iter=1)
    u = case.bcs[axs.VAR_U]
    u.type[axs.G_INLET] = axs.BC_PULSATILE
    u.value[axs.G_INLET] = U
    u.amplitude = [0.0] * 4
    u.frequency = [0.0] * 4
    u.amplitude[axs.G_INLET], u.frequency[axs.G_INLET] = A, f
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    assert math.isclose(o.bc_current(axs.VAR_U, axs.G_INLET), U * (1 + A * math.sin(2 * math.pi * f * 3 * dt)),
                        rel_tol=1e-14)
    case.config.t_end = 2 * dt
    o.configure(case.config)
    o.run(continue_run=True)
    # the continued steps are at 4 dt and 5 dt (the last frame was at 3 dt)
    assert math.isclose(o.bc_current(axs.VAR_U, axs.G_INLET), U * (1 + A * math.sin(2 * math.pi * f * 5 * dt)),
                        rel_tol=1e-14)
    assert [round(t / dt, 9) for t, _ in o.frames()] == [0.0, 1.0, 2.0, 3.0, 4.0, 5.0]
    # a fresh (not continued) run starts again from t = 0
    o.run(continue_run=False)
    assert [round(t / dt, 9) for t, _ in o.frames()] == [0.0, 1.0, 2.0]
