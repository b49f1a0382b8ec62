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


def test_vcycle_reduces_residual(require_gpu):
    case = axs.pipe_case()
    fs, fn, AC, AF, b = pp_system(case)
    g = axs.Multigrid(fs, fn)
    x = np.zeros_like(b)

    def resid(x):
        r = b - AC * x
        for c in range(fs.size - 1):
            for k in range(fs[c], fs[c + 1]):
                if fn[k] >= 0:
                    r[c] -= AF[k] * x[fn[k]]
        return np.abs(r).sum()

    r0 = resid(x)
    x = g.solve(AC, AF, b, x, cycles=10)
    assert resid(x) < 0.5 * r0
