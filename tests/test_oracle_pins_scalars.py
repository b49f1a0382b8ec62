"""Pins of the oracle's scalar transport (energy, concentration and the
Michaelis-Menten / Hill reactive walls) by closed forms it shares no code with:

  conduction           no flow, Dirichlet values at the inlet and outlet planes,
                       insulated wall and axis: the converged field is linear in z
                       at the cell centres, T = T_in + (T_out - T_in) z / L (exact
                       for the finite-volume scheme on a uniform grid), for the
                       energy equation (k / (rho cp)) and for species (D)
  Michaelis-Menten     the wall Newton solve of h (c_P - c_w) = Vmax c_w / (Km + c_w)
                       equals the positive root of the quadratic
                       h c_w^2 + (h Km - h c_P + Vmax) c_w - h c_P Km = 0
  Hill                 n = 1 reproduces Michaelis-Menten; n = 2 with substrate
                       inhibition leaves a flux residual below 1e-12 of h c_P
  wall consumption     area * J(c_w) per reactive face, and the diagnostic's sums,
                       recomputed here from the wall concentrations
"""
import math

import numpy as np
import pytest

import axs
from oracle_binding import Oracle, wall_concentration


def conduction_case(which):
    case = axs.scalar_pipe_case(nz=16, nr=4, u_in=0.0, energy=which == "energy",
                                concentration=which == "species", max_iter=1500)
    var = axs.VAR_T if which == "energy" else axs.VAR_C
    bc = case.bcs[var]
    bc.type[axs.G_INLET], bc.value[axs.G_INLET] = axs.BC_DIRICHLET, 1.0
    bc.type[axs.G_OUTLET], bc.value[axs.G_OUTLET] = axs.BC_DIRICHLET, 0.25
    bc.type[axs.G_WALL], bc.value[axs.G_WALL] = axs.BC_NEUMANN, 0.0
    c = case.config
    c.diffusivity = 2e-3   # fast species conduction: the closed form does not depend on it
    c.alpha_u = 1.0        # temperature relaxes with alpha_u; unrelaxed it converges in a few hundred iterations
    for i in range(axs.RES_COUNT):
        c.residual[i].tol = 0.0  # run all max_iter iterations
    return case, var


@pytest.mark.parametrize("which", ["energy", "species"])
def test_conduction_is_linear_in_z(which):
    case, var = conduction_case(which)
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    z = case.mesh.arrays["cell_center_z"]
    phi = o.field("t" if var == axs.VAR_T else "c")
    np.testing.assert_allclose(phi, 1.0 + (0.25 - 1.0) * z / 0.2, rtol=0, atol=1e-11)


@pytest.mark.parametrize("cp,h,vmax,km", [(0.2, 2e-5, 1e-6, 0.01), (1.0, 1e-3, 5e-3, 0.5), (3e-3, 1e-4, 1e-5, 1e-3)])
def test_michaelis_menten_wall_is_the_quadratic_root(cp, h, vmax, km):
    cw = wall_concentration(False, vmax, km, 1.0, 1.0, 0.0, 0.0, cp, h)
    bq = h * km - h * cp + vmax
    root = (-bq + math.sqrt(bq * bq + 4 * h * h * cp * km)) / (2 * h)
    assert math.isclose(cw, root, rel_tol=1e-10)
    assert 0.0 <= cw <= cp
    # a Hill law with n = 1 is the Michaelis-Menten law
    assert math.isclose(wall_concentration(True, vmax, km, 1.0, 1.0, 0.0, 0.0, cp, h), cw, rel_tol=1e-12)


def test_hill_with_inhibition_balances_the_wall_flux():
    cp, h, vmax, km, n, m, k2, v2 = 0.2, 2e-5, 3e-6, 0.05, 2.0, 2.0, 0.08, 0.6
    cw = wall_concentration(True, vmax, km, n, m, k2, v2, cp, h)
    J = vmax * cw ** n / (km ** n + cw ** n) * (1.0 - v2 * cw ** m / (k2 ** m + cw ** m))
    assert abs(h * (cp - cw) - J) <= 1e-12 * h * cp
    assert 0.0 < cw < cp
    # without inhibition the wall takes more, so its concentration is lower
    assert wall_concentration(True, vmax, km, n, m, k2, 0.0, cp, h) < cw


def test_wall_consumption_and_diagnostic():
    case = axs.scalar_pipe_case(nz=32, nr=8, max_iter=40)
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    a = case.mesh.arrays
    bc = case.bcs[axs.VAR_C]
    vmax, km = bc.vmax[axs.G_WALL], bc.km[axs.G_WALL]
    wall = np.where((a["face_neighbor"] < 0) & (a["face_group"] == axs.G_WALL))[0]
    cw, ocr = o.field("wall_cw"), o.field("wall_ocr")
    assert np.all(cw[wall] > 0) and np.all(cw[wall] < 0.2)
    np.testing.assert_allclose(ocr[wall], a["face_area"][wall] * vmax * cw[wall] / (km + cw[wall]), rtol=1e-14)
    d = o.wall_diagnostic()
    assert d[3] == wall.size
    np.testing.assert_allclose([d[0], d[4], d[8], d[9]], [ocr[wall].sum(), cw[wall].sum(), vmax * wall.size,
                                                         km * wall.size], rtol=1e-12)
    # inflow supply: -mdot * c_in over the inlet faces
    inlet = np.where((a["face_neighbor"] < 0) & (a["face_group"] == axs.G_INLET))[0]
    md = o.field("mdot")
    assert math.isclose(d[1], np.sum(-md[inlet] * 0.2), rel_tol=1e-12)
    # the wall consumes part of what comes in
    assert 0 < d[0] < d[1]


def test_reactive_wall_row_is_the_wall_flux():
    """Transport of a uniform concentration c0 (= the inlet value) with no
    convection: every interior, inlet, outlet and axis face balances, so each
    row's b - A c0 is what its reactive wall faces take, -sum area J(c_w), with
    c_w the Michaelis-Menten root for h = 1 / (d_PF / D + R_wall)."""
    rw = 2.0e3
    case = axs.scalar_pipe_case(nz=12, nr=5, wall_resistance=rw)
    o = Oracle(case.mesh, case.bcs, case.config)
    a = case.mesh.arrays
    D, c0 = case.config.diffusivity, 0.2
    bc = case.bcs[axs.VAR_C]
    vmax, km = bc.vmax[axs.G_WALL], bc.km[axs.G_WALL]
    phi = np.full(case.mesh.n_cells, c0)
    AC, AF, b = o.transport(axs.VAR_C, phi, None, None, None, 0, D, 0, axs.CONV_UPWIND)
    fs, ids = a["cell_face_start"], a["cell_face_ids"]
    rowsum = AC.copy()
    for c in range(case.mesh.n_cells):
        rowsum[c] += AF[fs[c]:fs[c + 1]].sum()
    got = b - rowsum * c0
    exp = np.zeros(case.mesh.n_cells)
    wall = np.where((a["face_neighbor"] < 0) & (a["face_group"] == axs.G_WALL))[0]
    for f in wall:
        c = a["face_owner"][f]
        dpf = abs((a["face_center_z"][f] - a["cell_center_z"][c]) * a["face_normal_z"][f]
                  + (a["face_center_r"][f] - a["cell_center_r"][c]) * a["face_normal_r"][f])
        h = 1.0 / (dpf / D + rw)
        bq = h * km - h * c0 + vmax
        cw = (-bq + math.sqrt(bq * bq + 4 * h * h * c0 * km)) / (2 * h)
        exp[c] -= a["face_area"][f] * vmax * cw / (km + cw)
    assert np.count_nonzero(exp) == wall.size
    np.testing.assert_allclose(got, exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())
