"""GPU multigrid: hierarchy maps bit-exact against the oracle's restatement of
buildLevels (multigrid.cu:79-325), and the device V-cycle against the
oracle's vCycle (multigrid.cu:607-662) on the same system."""
import numpy as np
import pytest

import axs
from oracle_binding import Oracle, OracleMG

pytestmark = pytest.mark.gpu


def pp_system(case, iters=3):
    case.config.max_iter = iters
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    fs, fn = o.csr()
    return fs, fn, o.field("pp_ac"), o.field("pp_af"), o.field("pp_b")


@pytest.mark.parametrize("which", ["pipe", "nozzle"])
def test_maps_bit_exact(require_gpu, which):
    case = axs.pipe_case() if which == "pipe" else axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=2.5e-4))
    o = Oracle(case.mesh, case.bcs, case.config)
    fs, fn = o.csr()
    g = axs.Multigrid(fs, fn)
    r = OracleMG(fs, fn)
    assert g.levels == r.levels
    for l in range(g.levels):
        a, b = g.maps(l), r.maps(l)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"level {l} {k}")


@pytest.mark.parametrize("cycles", [1, 3])
def test_vcycle_matches_oracle(require_gpu, cycles):
    case = axs.pipe_case()
    fs, fn, AC, AF, b = pp_system(case)
    x0 = np.zeros_like(b)
    g = axs.Multigrid(fs, fn)
    r = OracleMG(fs, fn)
    xg = g.solve(AC, AF, b, x0, cycles=cycles)
    xr = r.solve(AC, AF, b, x0, cycles=cycles)
    np.testing.assert_allclose(xg, xr, rtol=0, atol=1e-13 * np.abs(xr).max())


# (pre, post, coarse): the first sweep from zero folded into the restriction
# (pre >= 1), none to fold (pre 0, coarse 0), several left after it (pre 2)
SWEEPS = [(1, 1, 30), (0, 1, 30), (2, 2, 7), (1, 0, 0), (3, 1, 1)]


@pytest.mark.parametrize("sweeps", SWEEPS, ids=lambda s: "pre%d_post%d_coarse%d" % s)
@pytest.mark.parametrize("size", ["tail_at_1", "tail_below"])
def test_vcycle_sweep_counts_bit_exact(require_gpu, sweeps, size):
    pre, post, coarse = sweeps
    case = axs.pipe_case(nz=32, nr=8) if size == "tail_at_1" else axs.pipe_case()
    fs, fn, AC, AF, b = pp_system(case)
    g = axs.Multigrid(fs, fn)
    tail, lds = g.tail_info()
    print("levels", [g.level_info(l)[0] for l in range(g.levels)], "tail", tail, "lds", lds)
    assert (tail == 1) if size == "tail_at_1" else (tail > 1)
    x0 = np.linspace(-1.0, 1.0, b.size) * 1e-3  # a non-zero start for level 0
    xg = g.solve(AC, AF, b, x0, cycles=2, pre=pre, post=post, coarse=coarse)
    xr = OracleMG(fs, fn).solve(AC, AF, b, x0, cycles=2, pre=pre, post=post, coarse=coarse)
    np.testing.assert_array_equal(xg, xr)


def test_vcycles_match_textbook_on_nozzle(require_gpu):
    """Deep hierarchy (two-pass restriction on level 0, one-pass below it, the
    LDS-resident tail) against the textbook FP64 cycle and the oracle."""
    from mg_textbook import TextbookMG, spmat

    case = axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=1.1e-4))
    fs, fn, AC, AF, b = pp_system(case)
    g = axs.Multigrid(fs, fn)
    sizes = [g.level_info(l)[0] for l in range(g.levels)]
    print("levels", sizes, "tail", g.tail_info(), "fp32 tail", g.tail_info(axs.PREC_MIXED))
    assert g.levels >= 5
    assert sizes[0] > 65536 >= sizes[1]  # both restriction forms run
    assert 0 < g.tail_info()[0] < g.levels - 1  # the tail holds more than the coarsest level
    x = g.solve(AC, AF, b, np.zeros_like(b), cycles=3)
    tb = TextbookMG(spmat(fs, fn, AC, AF), [g.maps(l)["cell_to_coarse"] for l in range(g.levels - 1)])
    ref = tb.solve(b, np.zeros_like(b), 3)
    np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    # and against the oracle bit for bit
    xr = OracleMG(fs, fn).solve(AC, AF, b, np.zeros_like(b), cycles=3)
    np.testing.assert_array_equal(x, xr)


@pytest.mark.parametrize("which", ["pipe", "nozzle"])
@pytest.mark.parametrize("sweeps", [(1, 1, 30), (2, 0, 3)], ids=lambda s: "pre%d_post%d_coarse%d" % s)
def test_gauss_seidel_vcycle_bit_exact(require_gpu, which, sweeps):
    """Multi-colour Gauss-Seidel smoothing (per-colour launches on large levels,
    one workgroup on small ones) against the oracle, bit for bit."""
    pre, post, coarse = sweeps
    case = axs.pipe_case() if which == "pipe" else axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=1.1e-4))
    fs, fn, AC, AF, b = pp_system(case)
    g = axs.Multigrid(fs, fn)
    x0 = np.linspace(-1.0, 1.0, b.size) * 1e-3
    xg = g.solve(AC, AF, b, x0, cycles=2, pre=pre, post=post, coarse=coarse, smoother=axs.MG_SMOOTH_RBGS)
    xr = OracleMG(fs, fn).solve(AC, AF, b, x0, cycles=2, pre=pre, post=post, coarse=coarse, smoother=1)
    np.testing.assert_array_equal(xg, xr)
    # the FP32 cycles in FP64 refinement stay on the FP64 iterate
    x32 = g.solve(AC, AF, b, x0, cycles=2, pre=pre, post=post, coarse=coarse, smoother=axs.MG_SMOOTH_RBGS,
                  precision=axs.PREC_MIXED)
    assert np.linalg.norm(x32 - xr) / np.linalg.norm(xr) < 1e-5


@pytest.fixture(scope="module")
def nozzle_system():
    case = axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=1.1e-4))
    return pp_system(case)


@pytest.mark.parametrize("cycles", [1, 2, 4])
def test_fused_level0_cycles_bit_exact(require_gpu, nozzle_system, cycles):
    """Level 0 of a large mesh runs fused across cycles (sweep + residual, then
    prolongation + two sweeps + residual per later cycle): bit for bit the
    oracle's cycles, from a non-zero start and for any cycle count."""
    fs, fn, AC, AF, b = nozzle_system
    assert fs.size - 1 >= 16384
    g = axs.Multigrid(fs, fn)
    x0 = np.linspace(-1.0, 1.0, b.size) * 1e-3
    xg = g.solve(AC, AF, b, x0, cycles=cycles)
    xr = OracleMG(fs, fn).solve(AC, AF, b, x0, cycles=cycles)
    np.testing.assert_array_equal(xg, xr)
    # a sweep count the fused path does not take falls back to the one-pass kernels
    xg2 = g.solve(AC, AF, b, x0, cycles=cycles, pre=2, post=1)
    xr2 = OracleMG(fs, fn).solve(AC, AF, b, x0, cycles=cycles, pre=2, post=1)
    np.testing.assert_array_equal(xg2, xr2)


def test_fused_level0_quad_mesh_bit_exact(require_gpu):
    """The fused level-0 cycles on a quad mesh (4 slots per row) against the
    oracle, bit for bit."""
    case = axs.pipe_case(nz=32, nr=640)
    fs, fn, AC, AF, b = pp_system(case, iters=2)
    assert fs.size - 1 >= 16384
    x0 = np.linspace(-1.0, 1.0, b.size) * 1e-3
    xg = axs.Multigrid(fs, fn).solve(AC, AF, b, x0, cycles=3)
    xr = OracleMG(fs, fn).solve(AC, AF, b, x0, cycles=3)
    np.testing.assert_array_equal(xg, xr)


def test_fused_level0_patch_blocks_bit_exact(require_gpu):
    """Level 0 numbered in rows of 640 cells (breadth-first patch blocks) against
    the oracle, bit for bit."""
    case = axs.pipe_case(nz=640, nr=40)
    fs, fn, AC, AF, b = pp_system(case, iters=2)
    x0 = np.linspace(-1.0, 1.0, b.size) * 1e-3
    xg = axs.Multigrid(fs, fn).solve(AC, AF, b, x0, cycles=3)
    xr = OracleMG(fs, fn).solve(AC, AF, b, x0, cycles=3)
    np.testing.assert_array_equal(xg, xr)
