"""Output formats on the host (no GPU): the .axislv solver payload (format 5)
against the byte layout the reference writes field by field
(manager/source/file_manager.cpp:1482-1549; ConfigSolver 40 bytes,
ConfigSimple 24, FluidPropertyConfig 40, VariableUnits 15, SolverFieldOption 2,
then per residual type / norm / scaling / enabled / tolerance), its clamps, and
the multiblock cell outlines the NumPy export welds."""
import struct

import numpy as np
import pytest

import axs

LAYOUT = "<II" + "15B" + "2B" + "5B3xi4B2di4x" + "5d" + "2i2d" + "4Bd" * 5


def custom_config():
    c = axs.default_config(convection_scheme=axs.CONV_SOU, gradient_scheme=axs.GRAD_GREEN_GAUSS,
                           linear_solver=axs.LINEAR_GS_RB, time_scheme=axs.TIME_BDF2, linear_max_iter=13,
                           use_multigrid=0, add_convection=1, transient=1, use_non_orth=1, dt=2.5e-3, t_end=0.75,
                           save_key_frame_iter=7, rho=1050.0, mu=3.5e-3, cp=3600.0, k=0.5, diffusivity=2e-9,
                           max_iter=321, check_conv=3, solve_energy=1, solve_concentration=0)
    for i in range(axs.RES_COUNT):
        c.residual[i].enabled = i % 2
        c.residual[i].norm = i % 3
        c.residual[i].scaling = (i + 1) % 4
        c.residual[i].tol = 10.0 ** -(i + 2)
    return c


def test_solver_file_byte_layout(tmp_path):
    c = custom_config()
    x = axs.SolverFile()
    for i in range(15):
        x.var_units[i] = i + 1
    x.mom_tol, x.pp_tol = 2e-8, 3e-5
    for i in range(axs.RES_COUNT):
        x.residual_type[i] = i % 3
    path = tmp_path / "s.axislv"
    axs.save_solver_file(path, c, x)
    data = path.read_bytes()
    assert len(data) == struct.calcsize(LAYOUT) == 8 + 15 + 2 + 40 + 40 + 24 + 5 * 12
    res = []
    for i in range(axs.RES_COUNT):
        res += [i % 3, i % 3, (i + 1) % 4, i % 2, 10.0 ** -(i + 2)]
    expected = struct.pack(LAYOUT, 0x53585641, 5, *range(1, 16), 1, 0,
                           0, axs.CONV_SOU, axs.GRAD_GREEN_GAUSS, axs.LINEAR_GS_RB, axs.TIME_BDF2, 13, 0, 1, 1, 1,
                           2.5e-3, 0.75, 7,
                           1050.0, 3.5e-3, 3600.0, 0.5, 2e-9,
                           321, 3, 2e-8, 3e-5, *res)
    assert data[:4] == b"AVXS"  # 0x53585641 little-endian ("AXVS" read as a number)
    assert data == expected


def test_solver_file_round_trip(tmp_path):
    c = custom_config()
    path = tmp_path / "s.axislv"
    axs.save_solver_file(path, c)
    base = axs.default_config(mg_max_iter=9, alpha_u=0.5)
    got, extra = axs.load_solver_file(path, base)
    for name, _ in axs.Config._fields_:
        if name == "residual":
            for i in range(axs.RES_COUNT):
                for f, _ in axs.ResidualConfig._fields_:
                    assert getattr(got.residual[i], f) == getattr(c.residual[i], f), (i, f)
        elif name in ("mg_max_iter", "alpha_u"):
            assert getattr(got, name) == getattr(base, name)  # not in the file: kept
        else:
            assert getattr(got, name) == getattr(c, name), name
    assert extra.mom_tol == 1e-8 and extra.pp_tol == 1e-5 and extra.velocity_solver == 0


def test_solver_file_clamps(tmp_path):
    c = axs.default_config(linear_max_iter=0, save_key_frame_iter=-1, max_iter=0, check_conv=0,
                           convection_scheme=axs.CONV_QUICK, time_scheme=7, gradient_scheme=5, rho=-1.0)
    c.residual[0].norm = 9
    c.residual[1].scaling = axs.SCALE_CONTINUITY
    path = tmp_path / "s.axislv"
    axs.save_solver_file(path, c)
    got, _ = axs.load_solver_file(path)
    assert (got.linear_max_iter, got.save_key_frame_iter, got.max_iter, got.check_conv) == (20, 2, 50, 1)
    assert got.convection_scheme == axs.CONV_UPWIND and got.time_scheme == axs.TIME_BDF1
    assert got.gradient_scheme == axs.GRAD_LSQ
    assert (got.rho, got.mu, got.cp, got.k, got.diffusivity) == (998.0, 0.0010518, 4180.0, 0.6, 3.0277e-9)
    assert got.residual[0].norm == axs.NORM_LINF and got.residual[1].scaling == axs.SCALE_NONE


def test_solver_file_rejects_other_files(tmp_path):
    path = tmp_path / "s.axislv"
    axs.save_solver_file(path, axs.default_config())
    data = bytearray(path.read_bytes())
    bad = tmp_path / "v6.axislv"
    bad.write_bytes(bytes(data[:4]) + struct.pack("<I", 6) + bytes(data[8:]))
    with pytest.raises(axs.AxsError) as e:
        axs.load_solver_file(bad)
    assert e.value.args and "version" in str(e.value)
    short = tmp_path / "short.axislv"
    short.write_bytes(bytes(data[:50]))
    with pytest.raises(axs.AxsError):
        axs.load_solver_file(short)
    other = tmp_path / "other.bin"
    other.write_bytes(b"\0" * 200)
    with pytest.raises(axs.AxsError):
        axs.load_solver_file(other)


LAYOUT_V4 = "<II" + "15B" + "2B" + "iiBBB5xdd" + "iii" + "5d" + "iiddB7x" + "i" + "iiiBd" * 5


def test_solver_file_reads_version_4(tmp_path):
    """The version-4 payload (readSolverPayloadLegacy, file_manager.cpp:321-373):
    int-width enums, the velocity solver / convection scheme / keyframe interval
    as loose ints, the corrector as ConfigSimple's trailing byte (a pass count:
    2 means on), the gradient scheme after it; the multigrid flag, which v4 did
    not hold, comes back on whatever the caller had."""
    res = []
    for i in range(axs.RES_COUNT):
        res += [7 if i == 0 else i % 3, i % 3, (i + 1) % 4, i % 2, 10.0 ** -(i + 3)]
    data = struct.pack(LAYOUT_V4, 0x53585641, 4, *range(2, 17), 1, 1,
                       axs.LINEAR_GS_RB, 17, 1, 1, axs.TIME_BDF2, 2.5e-3, 0.5,
                       0, axs.CONV_SOU, 3,
                       1060.0, 3.5e-3, 3600.0, 0.52, 1e-9,
                       250, 2, 3e-8, 4e-5, 2,
                       axs.GRAD_GREEN_GAUSS, *res)
    assert len(data) == 8 + 15 + 2 + 32 + 12 + 40 + 32 + 4 + 5 * 21
    path = tmp_path / "v4.axislv"
    path.write_bytes(data)
    got, extra = axs.load_solver_file(path, axs.default_config(use_multigrid=0, alpha_u=0.5))
    assert list(extra.var_units) == list(range(2, 17))
    assert (got.solve_energy, got.solve_concentration) == (1, 1)
    assert (got.linear_solver, got.linear_max_iter, got.add_convection, got.transient) == (axs.LINEAR_GS_RB, 17, 1, 1)
    assert (got.time_scheme, got.dt, got.t_end) == (axs.TIME_BDF2, 2.5e-3, 0.5)
    assert (got.convection_scheme, got.save_key_frame_iter, got.gradient_scheme) == (axs.CONV_SOU, 3,
                                                                                     axs.GRAD_GREEN_GAUSS)
    assert (got.rho, got.mu, got.cp, got.k, got.diffusivity) == (1060.0, 3.5e-3, 3600.0, 0.52, 1e-9)
    assert (got.max_iter, got.check_conv, extra.mom_tol, extra.pp_tol) == (250, 2, 3e-8, 4e-5)
    assert got.use_non_orth == 1 and got.use_multigrid == 1 and got.alpha_u == 0.5
    assert extra.residual_type[0] == 1  # out of range -> RESIDUAL_RAW
    for i in range(axs.RES_COUNT):
        assert (got.residual[i].norm, got.residual[i].scaling, got.residual[i].enabled) == (i % 3, (i + 1) % 4, i % 2)
        assert got.residual[i].tol == 10.0 ** -(i + 3)
    # a version-4 file cut short, and version 3, are refused
    short = tmp_path / "short.axislv"
    short.write_bytes(data[:-10])
    with pytest.raises(axs.AxsError):
        axs.load_solver_file(short)
    v3 = tmp_path / "v3.axislv"
    v3.write_bytes(data[:4] + struct.pack("<I", 3) + data[8:])
    with pytest.raises(axs.AxsError):
        axs.load_solver_file(v3)


def test_multiblock_quads_match_cells():
    case = axs.expansion_case(nz=(6, 10), nr=(4, 5))
    m = case.mesh
    q = m.quads
    assert q.shape == (m.n_cells, 4, 2)
    z, r = q[:, :, 0], q[:, :, 1]
    area = 0.5 * np.abs(np.sum(z * np.roll(r, -1, axis=1) - np.roll(z, -1, axis=1) * r, axis=1))
    np.testing.assert_allclose(area, m.arrays["cell_area2d"], rtol=1e-12)
    # each quad's vertices surround its centre
    assert np.all(z.min(axis=1) < m.arrays["cell_center_z"]) and np.all(z.max(axis=1) > m.arrays["cell_center_z"])
    assert m.polygons().corners == 4
