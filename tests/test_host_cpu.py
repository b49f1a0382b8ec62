"""CPU tests of the product's host side (no GPU): the C ABI exports, the
case generator against the reference's own packer, and the host builders
(colouring) against the oracle's restatement."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import axs
from oracle_binding import Oracle, oracle_coloring

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"


def declared_symbols():
    text = (ROOT / "include" / "axs.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 30
    lib = axs.lib()
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(axs.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (axs_\w+)", out))
    assert set(syms) <= exported


def test_library_has_gfx950_code_object():
    data = axs.LIB_PATH.read_bytes()
    assert b"gfx950" in data


def test_no_gpu_here_is_an_error_not_a_fallback():
    import ctypes

    h = ctypes.c_void_p()
    rc = axs.lib().axs_create(ctypes.byref(h), 0)
    if rc == 0:
        axs.lib().axs_destroy(h)
        pytest.skip("a GPU is visible")
    assert rc in (axs_err for axs_err in (1, 2))
    assert axs.lib().axs_last_error(None)


LAYOUTS = {
    "pipe_32x8": [dict(z0=0.0, r0=0.0, axial=(0.2, 32), radial=(0.01, 8), groups=(0, 1, 3, 2))],
    "expansion_small": [
        dict(z0=0.0, r0=0.0, axial=(0.04, 12), radial=(0.002, 8, 1 / 1.05), groups=(0, -1, 3, 2)),
        dict(z0=0.04, r0=0.0, axial=(0.26, 40), radial=(0.002, 8, 1 / 1.05), groups=(-1, 1, 3, -1)),
        dict(z0=0.04, r0=0.002, axial=(0.26, 40), radial=(0.004, 12, 1 / 1.05), groups=(2, 1, -1, 2)),
    ],
}


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_multiblock_matches_reference_packer_bit_exact(name):
    """Golden vectors written by the reference's toPackedMesh (multiblock.h)."""
    with np.load(GOLDEN / f"multiblock_{name}.npz", allow_pickle=False) as g:
        ref = {k: g[k] for k in g.files}
    m = axs.multiblock_mesh(LAYOUTS[name])
    for k, v in ref.items():
        got = m.arrays[k]
        assert got.shape == v.shape, k
        np.testing.assert_array_equal(got, v, err_msg=k)


def test_full_pipe_matches_reference_packer_when_available():
    binp = ROOT / "oracle" / "_ref" / "ref_multiblock_golden"
    if not binp.exists():
        pytest.skip("reference checker not built here")
    import sys

    sys.path.insert(0, str(GOLDEN))
    from make_multiblock_golden import run_reference

    ref = run_reference([(0.0, 0.0, 0.2, 128, 1.0, 0.01, 32, 1.0, 0, 1, 3, 2)])
    m = axs.pipe_case().mesh
    for k, v in ref.items():
        np.testing.assert_array_equal(m.arrays[k], v, err_msg=k)


@pytest.mark.parametrize("which", ["pipe", "nozzle"])
def test_coloring_matches_oracle(which):
    case = axs.pipe_case() if which == "pipe" else axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=4e-4))
    o = Oracle(case.mesh, case.bcs, case.config)
    fs, fn = o.csr()
    cs, order = axs.coloring(fs, fn)
    rcs, rorder = oracle_coloring(fs, fn)
    np.testing.assert_array_equal(cs, rcs)
    np.testing.assert_array_equal(order, rorder)
    if which == "pipe":
        assert cs.size - 1 == 2


def test_nozzle_mesh_deterministic():
    a = axs.nozzle_mesh(spacing=5e-4)
    b = axs.nozzle_mesh(spacing=5e-4)
    for k in a.arrays:
        np.testing.assert_array_equal(a.arrays[k], b.arrays[k])
    c = axs.nozzle_mesh(spacing=5e-4, seed_points=11)
    assert c.n_cells != a.n_cells or not np.array_equal(c.arrays["cell_center_z"], a.arrays["cell_center_z"])


def test_triangles_builder_face_numbering():
    """Two triangles sharing an edge: faces in (v0v1, v1v2, v2v0) first-visit
    order, owner = first triangle (createUnstructuredMesh, mesh.cpp:1917-1983)."""
    pts = np.array([[0.0, 1.0], [1.0, 1.0], [0.0, 2.0], [1.0, 2.0]])
    tris = np.array([[0, 1, 2], [1, 3, 2]])
    segs = [(0.0, 1.0, 1.0, 1.0, 0), (1.0, 1.0, 1.0, 2.0, 1), (1.0, 2.0, 0.0, 2.0, 2), (0.0, 2.0, 0.0, 1.0, 3)]
    m = axs.triangle_mesh(pts, tris, segs, tol=1e-3)
    a = m.arrays
    assert m.n_cells == 2 and m.n_faces == 5
    np.testing.assert_array_equal(a["face_owner"], [0, 0, 0, 1, 1])
    np.testing.assert_array_equal(a["face_neighbor"], [-1, 1, -1, -1, -1])
    np.testing.assert_array_equal(a["cell_face_ids"], [0, 1, 2, 3, 4, 1])
    np.testing.assert_array_equal(a["face_group"], [0, -1, 3, 1, 2])
    # the shared face's normal points from cell 0 towards cell 1
    assert a["face_normal_z"][1] > 0 and a["face_normal_r"][1] > 0
    np.testing.assert_allclose(a["cell_volume"], 2 * np.pi * a["cell_center_r"] * 0.5)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for several it says so and exits with
    status 2 before touching a device."""
    import sys

    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout.strip() == ""
