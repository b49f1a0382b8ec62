"""Momentum Jacobi sweeps fused per launch (build extension): S sweeps of u
and v in one workgroup per block of consecutive cells, the intermediate
iterates in LDS, the halo rings around each block recomputed by the block.
Each row is evaluated exactly as by the one-sweep kernel, so a run is
bit-identical to the unfused one for every block size and sweep count, also
when the sweep count does not divide the solve's sweeps -- for the FP64 sweeps
and for the mixed path's FP32 defect sweeps alike."""
import numpy as np
import pytest

import axs

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "p", "pp", "mdot")


def run(case, sweeps, cells=1024, continue_run=False, solver=None):
    s = solver or axs.make_solver(case)
    s.set_sweep_fusion(sweeps, cells)
    s.run(continue_run=continue_run)
    return s


@pytest.fixture(scope="module")
def nozzle_mesh():
    # about 6000 cells: 6 blocks of 1024 cells, 12 of 512
    return axs.nozzle_mesh(axs.nozzle_spacing_for(6000))


@pytest.mark.parametrize("linear_iters", [20, 5, 4])
@pytest.mark.parametrize("sweeps,cells", [(2, 512), (2, 1024), (3, 512), (3, 1024), (4, 512), (4, 1024), (5, 512), (5, 1024)])
def test_fused_sweeps_bit_exact(require_gpu, nozzle_mesh, sweeps, cells, linear_iters):
    case = axs.nozzle_case(mesh=nozzle_mesh, re_throat=100, max_iter=8, tol=0.0, linear_max_iter=linear_iters)
    ref = run(case, 1)
    assert ref.stats().sweeps_per_launch == 1
    s = run(case, sweeps, cells)
    assert s.stats().sweeps_per_launch == sweeps
    assert s.stats().iterations == ref.stats().iterations == 8
    for name in FIELDS:
        np.testing.assert_array_equal(s.field(name), ref.field(name), err_msg=name)


def test_fusion_change_between_continued_runs(require_gpu, nozzle_mesh):
    """Switching the sweep count between continued runs rebuilds the blocks and
    the graph; the continued solution is the unfused one."""
    case = axs.nozzle_case(mesh=nozzle_mesh, re_throat=100, max_iter=4, tol=0.0)
    ref = run(case, 1)
    run(case, 1, continue_run=True, solver=ref)
    s = run(case, 3, 512)
    run(case, 2, 1024, continue_run=True, solver=s)
    assert s.stats().iterations == ref.stats().iterations == 8
    for name in FIELDS:
        np.testing.assert_array_equal(s.field(name), ref.field(name), err_msg=name)


def test_fused_sweeps_bit_exact_full_size(require_gpu):
    """The 1M-cell nozzle: two outer iterations, fused (the default) against
    one sweep per launch."""
    case = axs.nozzle_case(cells=1.0e6, max_iter=2, tol=0.0)
    ref = run(case, 1)
    s = axs.make_solver(case)
    s.run()
    assert s.stats().sweeps_per_launch == 5
    for name in FIELDS:
        np.testing.assert_array_equal(s.field(name), ref.field(name), err_msg=name)


@pytest.mark.parametrize("sweeps", [3, 4, 5])
def test_fused_sweeps_quad_mesh_bit_exact(require_gpu, sweeps):
    """A structured quad pipe (4 slots per cell, Green-Gauss gradients) numbered
    in rows of 32 cells, so its halos fit: fused against one sweep per launch."""
    case = axs.pipe_case(nz=32, nr=640, max_iter=6, tol=0.0)
    ref = run(case, 1)
    s = run(case, sweeps, 0)
    assert s.stats().sweeps_per_launch == sweeps
    for name in FIELDS:
        np.testing.assert_array_equal(s.field(name), ref.field(name), err_msg=name)


@pytest.mark.parametrize("sweeps", [3, 4, 5])
def test_fused_sweeps_patch_blocks_bit_exact(require_gpu, sweeps):
    """Rows of 640 cells: consecutive-index blocks are strips whose halos do
    not fit, so the blocks are breadth-first patches; fused against one sweep
    per launch, bit for bit."""
    case = axs.pipe_case(nz=640, nr=40, max_iter=6, tol=0.0)
    ref = run(case, 1)
    s = run(case, sweeps, 0)
    assert s.stats().sweeps_per_launch == sweeps
    for name in FIELDS:
        np.testing.assert_array_equal(s.field(name), ref.field(name), err_msg=name)


@pytest.mark.parametrize("linear_iters", [20, 7, 4])
@pytest.mark.parametrize("sweeps,cells", [(2, 512), (3, 1024), (4, 1024), (5, 1024)])
def test_fused_fp32_sweeps_bit_exact(require_gpu, nozzle_mesh, sweeps, cells, linear_iters):
    """The mixed path's FP32 defect sweeps fused over the same blocks: S sweeps
    per launch against one per launch, bit for bit.  Fused, the first launch
    forms the FP64 defect in its first sweep and the last ends in u += delta;
    unfused, those are launches of their own (sweep counts that do not divide
    the solve's sweeps, and solves of fewer sweeps than one launch holds)."""
    case = axs.nozzle_case(mesh=nozzle_mesh, re_throat=100, max_iter=8, tol=0.0, precision=axs.PREC_MIXED,
                           linear_max_iter=linear_iters)
    ref = run(case, 1)
    s = run(case, sweeps, cells)
    assert s.stats().sweeps_per_launch == sweeps
    assert s.stats().iterations == ref.stats().iterations == 8
    for name in FIELDS:
        np.testing.assert_array_equal(s.field(name), ref.field(name), err_msg=name)


@pytest.mark.parametrize("sweeps", [3, 5])
def test_fused_fp32_sweeps_quad_and_patch_blocks_bit_exact(require_gpu, sweeps):
    """The FP32 defect sweeps with 4-slot rows: the quad pipe numbered in short
    rows (index blocks) and in long rows (breadth-first patches), fused against
    one sweep per launch."""
    for nz, nr in ((32, 640), (640, 40)):
        case = axs.pipe_case(nz=nz, nr=nr, max_iter=6, tol=0.0, precision=axs.PREC_MIXED)
        ref = run(case, 1)
        s = run(case, sweeps, 0)
        assert s.stats().sweeps_per_launch == sweeps
        for name in FIELDS:
            np.testing.assert_array_equal(s.field(name), ref.field(name), err_msg=f"{nz}x{nr} {name}")
