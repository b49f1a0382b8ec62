"""NumPy export of a GPU solve (saveSolutionNpy, manager/source/file_manager.cpp:
2074-2309): the tables hold exactly the solver's fields in the results panel's
column order beside the cell geometry, the outlines index the mesh's own
corners (multiblock quads welded into shared points, triangles as meshed), each
transient frame is a table of its own and meta.json describes them."""
import json

import numpy as np
import pytest

import axs

pytestmark = pytest.mark.gpu

COLUMNS = {"Axial Velocity": "u", "Radial Velocity": "v", "Velocity Magnitude": "vmag", "Pressure": "p",
           "Temperature": "t", "dT/dz": "grad_tz", "dT/dr": "grad_tr", "Concentration": "c", "dC/dz": "grad_cz",
           "dC/dr": "grad_cr", "Continuity": "continuity", "Cell Reynolds Number": "cell_re", "dU/dz": "grad_uz",
           "dU/dr": "grad_ur", "dV/dz": "grad_vz", "dV/dr": "grad_vr"}


def load(path):
    raw = path.read_bytes()
    hl = raw[8] | (raw[9] << 8)
    assert raw[:8] == b"\x93NUMPY\x01\x00" and (10 + hl) % 64 == 0 and raw[10 + hl - 1:10 + hl] == b"\n"
    return np.load(path, allow_pickle=False)


def check_solution(s, case, d, scalars=()):
    meta = json.loads((d / "meta.json").read_text())
    names = ["Axial Velocity", "Radial Velocity", "Velocity Magnitude", "Pressure"]
    if "t" in scalars:
        names += ["Temperature", "dT/dz", "dT/dr"]
    if "c" in scalars:
        names += ["Concentration", "dC/dz", "dC/dr"]
    names += ["Continuity", "Cell Reynolds Number", "dU/dz", "dU/dr", "dV/dz", "dV/dr"]
    assert meta["columns"] == ["z", "r", "volume", "area2D"] + names
    assert meta["frameColumns"] == names and meta["cells"] == case.mesh.n_cells
    sol = load(d / "solution.npy")
    assert sol.dtype == np.float64 and sol.shape == (case.mesh.n_cells, 4 + len(names))
    a = case.mesh.arrays
    for k, key in enumerate(("cell_center_z", "cell_center_r", "cell_volume", "cell_area2d")):
        np.testing.assert_array_equal(sol[:, k], a[key])
    for k, name in enumerate(names):
        np.testing.assert_array_equal(sol[:, 4 + k], s.field(COLUMNS[name]), err_msg=name)
    assert meta["fluid"]["rho"] == case.config.rho and meta["fluid"]["D"] == case.config.diffusivity
    return meta


def test_pipe_export_with_welded_quads(require_gpu, tmp_path):
    case = axs.pipe_case(nz=32, nr=8, max_iter=30)
    s = axs.make_solver(case)
    s.run()
    s.save_solution_npy(tmp_path)
    meta = check_solution(s, case, tmp_path)
    assert meta["meshType"] == "Structured" and meta["multiBlock"] is True and meta["structured"] is False
    pts, cells = load(tmp_path / "points.npy"), load(tmp_path / "cells.npy")
    assert cells.dtype == np.int32 and cells.shape == (case.mesh.n_cells, 4)
    assert pts.shape == (33 * 9, 2) and meta["points"] == 33 * 9 and meta["cellCorners"] == 4
    np.testing.assert_array_equal(pts[cells], case.mesh.quads)
    assert meta["frames"] == [] and not list(tmp_path.glob("frame_*.npy"))


def test_expansion_export_welds_block_interfaces(require_gpu, tmp_path):
    case = axs.expansion_case(nz=(6, 10), nr=(4, 5), max_iter=5)
    s = axs.make_solver(case)
    s.run()
    s.save_solution_npy(tmp_path)
    pts, cells = load(tmp_path / "points.npy"), load(tmp_path / "cells.npy")
    # three blocks' nodes, each interface node once: (6+1)(4+1) + (10+1)(4+1) + (10+1)(5+1) - 5 - 11
    assert pts.shape[0] == 7 * 5 + 11 * 5 + 11 * 6 - 5 - 11
    # an interface node keeps the first block's coordinates (the blocks may differ in the last bit)
    np.testing.assert_allclose(pts[cells], case.mesh.quads, rtol=0, atol=1e-15)


def test_nozzle_export_triangles(require_gpu, tmp_path):
    case = axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=6e-4), max_iter=5)
    s = axs.make_solver(case)
    s.run()
    s.save_solution_npy(tmp_path)
    meta = check_solution(s, case, tmp_path)
    assert meta["meshType"] == "Unstructured" and meta["cellCorners"] == 3
    np.testing.assert_array_equal(load(tmp_path / "points.npy"), case.mesh.points)
    np.testing.assert_array_equal(load(tmp_path / "cells.npy"), case.mesh.tris)


def test_transient_scalar_export_frames(require_gpu, tmp_path):
    case = axs.scalar_pipe_case(nz=16, nr=4, transient=1, dt=0.0625, t_end=0.1875, max_iter=3, save_key_frame_iter=1)
    s = axs.make_solver(case)
    s.run()
    s.save_solution_npy(tmp_path)
    meta = check_solution(s, case, tmp_path, scalars=("t", "c"))
    frames = s.frames()
    assert [f["file"] for f in meta["frames"]] == [f"frame_{i:04d}.npy" for i in range(len(frames))]
    assert [f["time"] for f in meta["frames"]] == [t for t, _ in frames]
    names = meta["frameColumns"]
    for i, (_, fields) in enumerate(frames):
        tab = load(tmp_path / f"frame_{i:04d}.npy")
        assert tab.shape == (case.mesh.n_cells, len(names))
        for k, name in enumerate(names):
            want = fields.get(COLUMNS[name])
            if want is None:
                assert not tab[:, k].any(), name  # no per-step data: zeros
            else:
                np.testing.assert_array_equal(tab[:, k], want, err_msg=name)
