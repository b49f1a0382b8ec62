"""Textbook aggregation multigrid with scipy.sparse, used as an independent
check of the device V-cycle variants (P = piecewise-constant injection from
the hierarchy maps, R = P^T, Galerkin A_c = P^T A P, damped Jacobi)."""
import numpy as np
import scipy.sparse as sp


def spmat(fs, fn, AC, AF):
    n = fs.size - 1
    rows = np.repeat(np.arange(n), np.diff(fs))
    m = fn >= 0
    return (sp.csr_matrix((AF[m], (rows[m], fn[m])), shape=(n, n)) + sp.diags(AC)).tocsr()


class TextbookMG:
    def __init__(self, A, maps):
        """maps: cell_to_coarse arrays of every level but the coarsest."""
        self.A = [A]
        self.P = []
        for c2c in maps:
            n, nc = c2c.size, int(c2c.max()) + 1
            P = sp.csr_matrix((np.ones(n), (np.arange(n), c2c)), shape=(n, nc))
            self.P.append(P)
            self.A.append((P.T @ self.A[-1] @ P).tocsr())
        self.inv = np.linalg.inv(self.A[-1].toarray())

    def vcycle(self, l, b, x, w=0.6, pre=1, post=1, coarse=30, direct=False):
        A = self.A[l]
        D = A.diagonal()
        if l == len(self.A) - 1:
            if direct:
                return self.inv @ b
            for _ in range(coarse):
                x = x + w * (b - A @ x) / D
            return x
        for _ in range(pre):
            x = x + w * (b - A @ x) / D
        rc = self.P[l].T @ (b - A @ x)
        xc = self.vcycle(l + 1, rc, np.zeros(rc.size), w, pre, post, coarse, direct)
        x = x + self.P[l] @ xc
        for _ in range(post):
            x = x + w * (b - A @ x) / D
        return x

    def solve(self, b, x, cycles, direct=False):
        for _ in range(cycles):
            x = self.vcycle(0, b, x, direct=direct)
        return x
