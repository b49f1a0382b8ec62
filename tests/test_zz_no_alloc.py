"""axs_mg_solve_noalloc: the multigrid solve with its work arrays in the
library's static device buffer, called through the C ABI on device arrays,
against axs_mg_solve (host arrays, the hierarchy's own buffers) on the same
systems: two calls in a row, the second on a smaller hierarchy, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import axs
from oracle_binding import Oracle

pytestmark = pytest.mark.gpu


def pp_system(case, iters=3):
    case.config.max_iter = iters
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    fs, fn = o.csr()
    return fs, fn, o.field("pp_ac"), o.field("pp_af"), o.field("pp_b")


class Hip:
    """The HIP runtime libaxs.so runs on (loaded with it), for the test's own
    device arrays."""

    def __init__(self):
        axs.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]

    def upload(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p: C.c_void_p, n: int) -> np.ndarray:
        out = np.zeros(n)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self, p):
        assert self.L.hipFree(p) == 0


def test_noalloc_matches_allocating_solve(require_gpu):
    hip = Hip()
    big = axs.nozzle_case(mesh=axs.nozzle_mesh(axs.nozzle_spacing_for(60000)))
    small = axs.pipe_case(nz=32, nr=8)
    for name, case in (("nozzle 60k", big), ("pipe 32x8", small)):
        fs, fn, AC, AF, b = pp_system(case)
        mg = axs.Multigrid(fs, fn)
        print(f"{name}: {fs.size - 1} cells, {mg.levels} levels, tail {mg.tail_info()}", flush=True)
        x0 = np.random.default_rng(3).uniform(-1e-3, 1e-3, b.size)
        ref = mg.solve(AC, AF, b, x0, cycles=3)
        d = {k: hip.upload(v) for k, v in (("AC", AC), ("AF", AF), ("b", b), ("x", x0))}
        mg.solve_noalloc(d["AC"], d["AF"], d["b"], d["x"], cycles=3)
        hip.sync()
        got = hip.download(d["x"], b.size)
        np.testing.assert_array_equal(got, ref, err_msg=name)
        # the inputs are untouched
        np.testing.assert_array_equal(hip.download(d["b"], b.size), b)
        np.testing.assert_array_equal(hip.download(d["AF"], AF.size), AF)
        for p in d.values():
            hip.free(p)
        mg.close()


def test_noalloc_rejects_a_hierarchy_beyond_its_buffer(require_gpu):
    # a 4M-cell chain: level 0's second iterate and residual alone are 64 MiB
    n = 4 << 20
    fs = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    fn = np.stack([np.arange(n) - 1, np.arange(n) + 1], axis=1).astype(np.int32)
    fn[fn >= n] = -1
    mg = axs.Multigrid(fs, fn.ravel(), max_levels=2)
    rc = axs.lib().axs_mg_solve_noalloc(mg._h, C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), 3, 0.6,
                                        1, 1, 30, None)
    assert rc == 5  # AXS_ERR_RANGE, before anything is launched
    assert b"bytes" in axs.lib().axs_last_error(None)
    mg.close()
