"""Mixed precision (build extension): FP32 momentum sweeps and FP32 V-cycles
inside FP64 defect correction, FP64 coarsest direct solve.  The plain FP64
path is the yardstick: the mixed path must reach the same converged state."""
import numpy as np
import pytest

import axs
from oracle_binding import Oracle, OracleMG

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def pp_system(case, iters=3):
    case.config.max_iter = iters
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    fs, fn = o.csr()
    return fs, fn, o.field("pp_ac"), o.field("pp_af"), o.field("pp_b")


def dense_residual(fs, fn, AC, AF, b, x):
    r = b - AC * x
    rows = np.repeat(np.arange(fs.size - 1), np.diff(fs))
    m = fn >= 0
    np.add.at(r, rows[m], -AF[m] * x[fn[m]])
    return r


def test_mixed_vcycle_is_fp64_vcycle_up_to_rounding(require_gpu):
    fs, fn, AC, AF, b = pp_system(axs.pipe_case())
    g = axs.Multigrid(fs, fn)
    x64 = g.solve(AC, AF, b, np.zeros_like(b), cycles=3)
    x32 = g.solve(AC, AF, b, np.zeros_like(b), cycles=3, precision=axs.PREC_MIXED)
    assert rel(x32, x64) < 1e-5


@pytest.mark.parametrize("direct,levels", [(False, 0), (True, 0), (True, 4), (True, 3)])
def test_mixed_cycle_matches_textbook(require_gpu, direct, levels):
    """FP32 V-cycles in FP64 refinement, with and without the FP64 coarsest
    solve, against the textbook FP64 cycle (scipy.sparse).  Note: with the
    exact coarsest solve this unsmoothed-aggregation cycle over-corrects and the
    L1 residual grows over the first cycles on this system -- the textbook cycle
    does the same; the reference's 30 coarse sweeps damp that correction.
    Four levels leave a coarsest level of a few hundred cells, three of over a
    thousand: their inverse is the blocked Gauss-Jordan one, applied one wave
    per row."""
    from mg_textbook import TextbookMG, spmat

    case = axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=2e-4))
    fs, fn, AC, AF, b = pp_system(case)
    g = axs.Multigrid(fs, fn, max_levels=levels)
    if levels:
        assert 256 < g.level_info(g.levels - 1)[0] <= 2048
    x = g.solve(AC, AF, b, np.zeros_like(b), cycles=3, precision=axs.PREC_MIXED, coarse_direct=int(direct))
    tb = TextbookMG(spmat(fs, fn, AC, AF), [g.maps(l)["cell_to_coarse"] for l in range(g.levels - 1)])
    ref = tb.solve(b, np.zeros_like(b), 3, direct=direct)
    assert rel(x, ref) < 1e-4


def test_mixed_refinement_tracks_fp64_cycles(require_gpu):
    """60 refinement cycles: the mixed iterate stays on the FP64 iterate
    (FP32 rounding scales with the shrinking defect)."""
    fs, fn, AC, AF, b = pp_system(axs.pipe_case(nz=32, nr=8))
    AC = AC * 1.02  # well-conditioned copy so that 60 cycles make progress
    g = axs.Multigrid(fs, fn)
    x64 = g.solve(AC, AF, b, np.zeros_like(b), cycles=60)
    x32 = g.solve(AC, AF, b, np.zeros_like(b), cycles=60, precision=axs.PREC_MIXED)
    r64 = np.abs(dense_residual(fs, fn, AC, AF, b, x64)).sum()
    r32 = np.abs(dense_residual(fs, fn, AC, AF, b, x32)).sum()
    assert r64 < 0.01 * np.abs(b).sum()
    assert rel(x32, x64) < 1e-6
    assert r32 < 1.01 * r64


@pytest.fixture(scope="module")
def nozzle_system():
    # 25k triangles: level 0 runs fused across cycles
    return pp_system(axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=1.1e-4)))


def solve_refined_once(g, AC, AF, b, x0, cycles, fused):
    g.set_mixed(axs.REFINE_SOLVE, fuse_level0=fused)
    try:
        return g.solve(AC, AF, b, x0, cycles=cycles, precision=axs.PREC_MIXED)
    finally:
        g.set_mixed(axs.REFINE_CYCLE, fuse_level0=True)


@pytest.mark.parametrize("cycles", [1, 2, 3, 4])
def test_refine_per_solve_fused_level0_bit_exact(require_gpu, nozzle_system, cycles):
    """One FP64 defect per solve, then every V-cycle in FP32 on the correction:
    level 0 fused across cycles (the FP32 copies through the FP64 path's
    level-0 blocks and fused level-1 restriction) against one-pass launches,
    bit for bit, from a non-zero start; and the answer is the FP64 cycles'
    up to FP32 rounding."""
    fs, fn, AC, AF, b = nozzle_system
    assert fs.size - 1 >= 16384
    g = axs.Multigrid(fs, fn)
    x0 = np.linspace(-1.0, 1.0, b.size) * 1e-3
    xf = solve_refined_once(g, AC, AF, b, x0, cycles, True)
    xu = solve_refined_once(g, AC, AF, b, x0, cycles, False)
    np.testing.assert_array_equal(xf, xu)
    x64 = g.solve(AC, AF, b, x0, cycles=cycles)
    assert rel(xf - x0, x64 - x0) < 1e-5


@pytest.mark.parametrize("nz,nr", [(32, 640), (640, 40)])
def test_refine_per_solve_fused_level0_quad_and_patch_blocks(require_gpu, nz, nr):
    """The same on quad meshes (4-slot rows): numbered in short rows (index
    blocks) and in long rows (breadth-first patches)."""
    fs, fn, AC, AF, b = pp_system(axs.pipe_case(nz=nz, nr=nr), iters=2)
    g = axs.Multigrid(fs, fn)
    x0 = np.zeros_like(b)
    xf = solve_refined_once(g, AC, AF, b, x0, 3, True)
    xu = solve_refined_once(g, AC, AF, b, x0, 3, False)
    np.testing.assert_array_equal(xf, xu)
    assert rel(xf, g.solve(AC, AF, b, x0, cycles=3)) < 1e-5


@pytest.mark.parametrize("refine", [axs.REFINE_CYCLE, axs.REFINE_SOLVE])
def test_mixed_simple_converges_to_fp64_solution(require_gpu, refine):
    case = axs.pipe_case()
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    assert o.converged
    mcase = axs.pipe_case(precision=axs.PREC_MIXED, mixed_refine=refine)
    s = axs.make_solver(mcase)
    s.run()
    st = s.stats()
    assert st.converged
    for name in ("u", "v", "p"):
        e = rel(s.field(name), o.field(name))
        print(name, "rel L2 vs FP64 oracle", e, "iterations", st.iterations, "vs", o.iterations)
        assert e < 1e-8, (name, e)
    assert abs(st.iterations - o.iterations) <= 1


@pytest.mark.parametrize("which,every", [("pipe", 1), ("nozzle", 1), ("pipe", 4), ("nozzle", 7)])
def test_mixed_direct_coarse_converges_to_fp64_solution(require_gpu, which, every):
    """With the exact FP64 coarsest solve the outer iterations take another
    path (fewer of them) to the same solution: both runs go to 1e-10 (the
    nozzle, whose small radial velocity needs it, to 1e-12), where the fields
    agree to 1e-8.  The nozzle at three levels has a coarsest level of a few
    hundred cells (the blocked inverse).  every > 1: the inverse formed on every
    `every`-th outer iteration and applied, one operator behind, in between."""
    if which == "pipe":
        make = lambda **kw: axs.pipe_case(tol=1e-10, **kw)
    else:
        mesh = axs.nozzle_mesh(axs.nozzle_spacing_for(6000))
        make = lambda **kw: axs.nozzle_case(mesh=mesh, re_throat=100, tol=1e-12, mg_max_levels=3, **kw)
    a = axs.make_solver(make())
    a.run()
    b = axs.make_solver(make(precision=axs.PREC_MIXED, coarse_direct=every))
    b.run()
    assert a.stats().converged and b.stats().converged
    print("iterations fp64", a.stats().iterations, "mixed + direct coarse", b.stats().iterations)
    for name in ("u", "v", "p"):
        e = rel(b.field(name), a.field(name))
        print(name, "rel L2 vs FP64", e)
        assert e < 1e-8, (name, e)


@pytest.mark.parametrize("refine", [axs.REFINE_CYCLE, axs.REFINE_SOLVE])
def test_mixed_expansion_converges_to_fp64_solution(require_gpu, refine):
    """The sudden expansion (SOU, least-squares gradients, graded multiblock
    grid) at 32,256 cells and throat Re 100, where the steady iteration
    converges to 1e-6 (at Re 200 and above it does not: the residuals stall,
    in FP64 as in the oracle): the mixed path ends on the FP64 path's fields."""
    kw = dict(nz=(48, 320), nr=(32, 64), re_throat=100.0, max_iter=20000)
    a = axs.make_solver(axs.expansion_case(**kw))
    a.run()
    b = axs.make_solver(axs.expansion_case(precision=axs.PREC_MIXED, mixed_refine=refine, **kw))
    b.run()
    assert a.stats().converged and b.stats().converged
    print("iterations fp64", a.stats().iterations, "mixed", b.stats().iterations)
    assert abs(a.stats().iterations - b.stats().iterations) <= 1
    for name in ("u", "v", "p"):
        e = rel(b.field(name), a.field(name))
        print(name, "rel L2 mixed vs FP64", e)
        assert e < 1e-8, (name, e)


@pytest.mark.parametrize("levels", [0, 3])
def test_direct_refresh_interval_first_iterations(require_gpu, levels):
    """coarse_direct = K forms the inverse on the first outer iteration of a
    run and on every K-th after it: one iteration is bit-identical to K = 1's,
    and the run to 1e-6 with the inverse one to four operators behind takes
    about as many outer iterations as K = 1 (measured on the 1M nozzle: 6903
    at K = 1, 6904-6915 at K = 2-64)."""
    mesh = axs.nozzle_mesh(axs.nozzle_spacing_for(6000))
    make = lambda **kw: axs.nozzle_case(mesh=mesh, re_throat=100, mg_max_levels=levels,
                                        precision=axs.PREC_MIXED, **kw)
    one = axs.make_solver(make(coarse_direct=1, max_iter=1))
    one.run()
    lag = axs.make_solver(make(coarse_direct=5, max_iter=1))
    lag.run()
    for name in ("u", "v", "p"):
        assert np.array_equal(one.field(name), lag.field(name)), name
    full = axs.make_solver(make(coarse_direct=1))
    full.run()
    every = axs.make_solver(make(coarse_direct=5))
    every.run()
    n1, n5 = full.stats().iterations, every.stats().iterations
    print("iterations to 1e-6: every iteration", n1, "every fifth", n5)
    assert full.stats().converged and every.stats().converged
    assert n5 <= 1.1 * n1 + 5


def test_direct_refresh_interval_rejects_negative(require_gpu):
    case = axs.pipe_case(nz=32, nr=8, precision=axs.PREC_MIXED, coarse_direct=-1)
    with pytest.raises(Exception):
        axs.make_solver(case).run()


@pytest.mark.parametrize("levels", [0, 4, 3])
def test_fp64_cycle_with_direct_coarse_matches_textbook(require_gpu, levels):
    """The FP64 V-cycles with the exact coarsest solve (the one-workgroup
    inverse inside the tail at full depth, the blocked inverse applied across
    the GPU at four levels) against the textbook FP64 cycle with a direct
    coarse solve (scipy.sparse): the same cycles up to summation order."""
    from mg_textbook import TextbookMG, spmat

    case = axs.nozzle_case(mesh=axs.nozzle_mesh(spacing=2e-4))
    fs, fn, AC, AF, b = pp_system(case)
    g = axs.Multigrid(fs, fn, max_levels=levels)
    if levels:
        assert 256 < g.level_info(g.levels - 1)[0] <= 2048
    x = g.solve(AC, AF, b, np.zeros_like(b), cycles=3, coarse_direct=1)
    tb = TextbookMG(spmat(fs, fn, AC, AF), [g.maps(l)["cell_to_coarse"] for l in range(g.levels - 1)])
    ref = tb.solve(b, np.zeros_like(b), 3, direct=True)
    e = rel(x, ref)
    print("rel L2 vs the textbook cycle", e)
    assert e < 1e-8


@pytest.mark.parametrize("which,every", [("pipe", 1), ("nozzle", 4)])
def test_fp64_direct_coarse_converges_to_fp64_solution(require_gpu, which, every):
    """The FP64 path with the exact coarsest solve (formed every `every`-th
    outer iteration) reaches the plain FP64 path's converged fields (both to
    1e-10, the nozzle to 1e-12) within 1e-8."""
    if which == "pipe":
        make = lambda **kw: axs.pipe_case(tol=1e-10, **kw)
    else:
        mesh = axs.nozzle_mesh(axs.nozzle_spacing_for(6000))
        make = lambda **kw: axs.nozzle_case(mesh=mesh, re_throat=100, tol=1e-12, mg_max_levels=3, **kw)
    a = axs.make_solver(make())
    a.run()
    b = axs.make_solver(make(coarse_direct=every))
    b.run()
    assert a.stats().converged and b.stats().converged
    print("iterations fp64", a.stats().iterations, "fp64 + direct coarse", b.stats().iterations)
    for name in ("u", "v", "p"):
        e = rel(b.field(name), a.field(name))
        print(name, "rel L2 vs FP64", e)
        assert e < 1e-8, (name, e)
