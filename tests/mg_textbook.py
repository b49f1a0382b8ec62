"""Textbook aggregation multigrid with scipy.sparse, used as an independent
check of the device V-cycle variants (P = piecewise-constant injection from
the hierarchy maps, R = P^T, Galerkin A_c = P^T A P, damped Jacobi or
multi-colour Gauss-Seidel in a given colour order)."""
import numpy as np
import scipy.sparse as sp


def spmat(fs, fn, AC, AF):
    n = fs.size - 1
    rows = np.repeat(np.arange(n), np.diff(fs))
    m = fn >= 0
    return (sp.csr_matrix((AF[m], (rows[m], fn[m])), shape=(n, n)) + sp.diags(AC)).tocsr()


class TextbookMG:
    def __init__(self, A, maps, colorings=None):
        """maps: cell_to_coarse arrays of every level but the coarsest;
        colorings: (color_start, cell_order) of every level, for Gauss-Seidel."""
        self.colorings = colorings
        self.A = [A]
        self.P = []
        for c2c in maps:
            n, nc = c2c.size, int(c2c.max()) + 1
            P = sp.csr_matrix((np.ones(n), (np.arange(n), c2c)), shape=(n, nc))
            self.P.append(P)
            self.A.append((P.T @ self.A[-1] @ P).tocsr())
        self.inv = np.linalg.inv(self.A[-1].toarray())

    def smooth(self, l, b, x, sweeps, w, gs):
        A = self.A[l]
        D = A.diagonal()
        if not gs:
            for _ in range(sweeps):
                x = x + w * (b - A @ x) / D
            return x
        x = x.copy()
        cs, order = self.colorings[l]
        for _ in range(sweeps):
            for c in range(cs.size - 1):
                for n in order[cs[c]:cs[c + 1]]:
                    if abs(D[n]) < 1e-30:
                        continue
                    row = A.getrow(n)
                    off = row.dot(x)[0] - D[n] * x[n]
                    x[n] = (b[n] - off) / D[n]
        return x

    def vcycle(self, l, b, x, w=0.6, pre=1, post=1, coarse=30, direct=False, gs=False):
        A = self.A[l]
        if l == len(self.A) - 1:
            if direct:
                return self.inv @ b
            return self.smooth(l, b, x, coarse, w, gs)
        x = self.smooth(l, b, x, pre, w, gs)
        rc = self.P[l].T @ (b - A @ x)
        xc = self.vcycle(l + 1, rc, np.zeros(rc.size), w, pre, post, coarse, direct, gs)
        x = x + self.P[l] @ xc
        return self.smooth(l, b, x, post, w, gs)

    def solve(self, b, x, cycles, direct=False, gs=False, **kw):
        for _ in range(cycles):
            x = self.vcycle(0, b, x, direct=direct, gs=gs, **kw)
        return x
