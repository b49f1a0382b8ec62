"""CPU pins of the learned-guess network's restatement (oracle/net_reference.py):
its gradient against central finite differences of its own forward pass, the
transposed aggregation against a dense adjacency, and the BF16 rounding against
hand-checked values.  No GPU."""
import sys
from pathlib import Path

import numpy as np

import axs
from oracle_binding import Oracle

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import net_reference as nr  # noqa: E402


def pp_system(nz=8, nr_=4, iters=2):
    case = axs.pipe_case(nz=nz, nr=nr_)
    case.config.max_iter = iters
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    fs, fn = o.csr()
    return fs, fn, o.field("pp_ac"), o.field("pp_af"), o.field("pp_b")


def dense(fs, fn, AC, AF):
    n = fs.size - 1
    A = np.diag(AC.astype(np.float64))
    for i in range(n):
        for k in range(fs[i], fs[i + 1]):
            if fn[k] >= 0:
                A[i, fn[k]] += AF[k]
    return A


def test_bf16_rounding_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38],
                 np.float32)
    got = nr.bf16(x)
    # bf16 spacing at 1 is 2^-7: 1 + 2^-7 is exact; 1 + 3*2^-9 rounds up to it;
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    want = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -7, 1.0, 1.0 + 2 ** -6, -2.5, got[6]], np.float32)
    np.testing.assert_array_equal(got, want)
    assert abs(got[6] - 3.0e38) / 3.0e38 < 2 ** -8


def test_transposed_aggregation_is_adjacency_transpose():
    fs, fn, AC, AF, b = pp_system()
    g = nr.Graph(fs, fn, AC, AF)
    W = np.zeros((g.n, g.n))
    for i in range(g.n):
        for k in range(fs[i], fs[i + 1]):
            if fn[k] >= 0:
                W[i, fn[k]] += g.w64[k]
    X = np.random.default_rng(1).normal(size=(g.n, 5))
    np.testing.assert_allclose(g.aggregate(X, g.w64), W @ X, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(g.aggregate(X, g.weightsT(True)), W.T @ X, rtol=1e-13, atol=1e-14)


def test_gradient_matches_finite_differences():
    fs, fn, AC, AF, b = pp_system()
    g = nr.Graph(fs, fn, AC, AF)
    x_star = np.linalg.solve(dense(fs, fn, AC * 1.02, AF), b)
    layers = 3
    p = nr.init_params(layers, seed=3)
    loss, grad = nr.loss_grad(p, g, b, x_star, layers, exact=True)
    assert np.isfinite(loss) and loss > 0
    rng = np.random.default_rng(7)
    for _ in range(4):
        v = rng.normal(size=p.size)
        eps = 1e-6
        lp = nr.loss_grad(p + eps * v, g, b, x_star, layers, exact=True)[0]
        lm = nr.loss_grad(p - eps * v, g, b, x_star, layers, exact=True)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad @ v) <= 1e-5 * max(abs(fd), 1e-12), (fd, grad @ v)


def test_network_is_positively_homogeneous():
    """No biases and ReLU only: the guess scales with b (x(2b) = 2 x(b)),
    and the device-precision forward stays close to the exact one."""
    fs, fn, AC, AF, b = pp_system()
    g = nr.Graph(fs, fn, AC, AF)
    p = nr.init_params(4, seed=5)
    _, x1, _ = nr.forward(p, g, b, 4, exact=True)
    _, x2, _ = nr.forward(p, g, 2.0 * b, 4, exact=True)
    np.testing.assert_allclose(x2, 2.0 * x1, rtol=1e-12, atol=1e-300)
    _, xb, _ = nr.forward(p, g, b, 4, exact=False)
    assert np.linalg.norm(xb - x1) / np.linalg.norm(x1) < 3e-2
