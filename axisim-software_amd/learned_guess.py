"""Training data and evaluation for the learned initial guess of the pressure
correction (build extension; the reference has no learned part).

The network runs on one coarse multigrid level k of the p' system A x = b.  A
training pair is that level's Galerkin system (A_k = P^T A P, b_k = P^T b) and
its solution A_k^{-1} b_k: prolonged back to the mesh, the exact coarse
solution is the best guess the level can express in the energy norm, so it is
what the network is taught to approximate.  Pairs come from the solver itself:
p' systems captured at chosen outer iterations of SIMPLE runs on seeded
geometries.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import axs

GUESS_CELLS = 16384  # the solver's default: the first level with at most this many cells


def slot_neighbours(mesh: axs.Mesh) -> np.ndarray:
    """Neighbour cell of every face slot (-1 on boundary faces), CSR order."""
    a = mesh.arrays
    ids = a["cell_face_ids"]
    cells = np.repeat(np.arange(mesh.n_cells, dtype=np.int32), np.diff(a["cell_face_start"]))
    own = a["face_owner"][ids]
    nb = a["face_neighbor"][ids]
    return np.where(own == cells, nb, own).astype(np.int32)


def csr_matrix(fs, fn, AC, AF):
    n = fs.size - 1
    rows = np.repeat(np.arange(n), np.diff(fs))
    m = fn >= 0
    return (sp.csr_matrix((AF[m], (rows[m], fn[m])), shape=(n, n)) + sp.diags(AC)).tocsr()


def guess_level(mg: axs.Multigrid, cells: int = GUESS_CELLS) -> int:
    for l in range(1, mg.levels):
        if mg.level_info(l)[0] <= cells:
            return l
    return -1


class Hierarchy:
    """Transfer operators of a multigrid hierarchy (piecewise-constant P)."""

    def __init__(self, mg: axs.Multigrid):
        self.maps = [mg.maps(l) for l in range(mg.levels)]
        self.P = []
        for l in range(mg.levels - 1):
            c2c = self.maps[l]["cell_to_coarse"]
            nc = mg.level_info(l + 1)[0]
            self.P.append(sp.csr_matrix((np.ones(c2c.size), (np.arange(c2c.size), c2c)), shape=(c2c.size, nc)))

    def restrict(self, v, level):
        for l in range(level):
            v = self.P[l].T @ v
        return v

    def prolong(self, v, level):
        for l in range(level - 1, -1, -1):
            v = self.P[l] @ v
        return v

    def galerkin(self, A, level):
        for l in range(level):
            A = (self.P[l].T @ A @ self.P[l]).tocsr()
        return A

    def level_system(self, A, level):
        """(face_start, face_neighbor, AC, AF) of level `level` in its CSR order."""
        Ak = self.galerkin(A, level)
        fs, fn = self.maps[level]["face_start"], self.maps[level]["face_neighbor"]
        rows = np.repeat(np.arange(fs.size - 1), np.diff(fs))
        AF = np.zeros(fn.size)
        m = fn >= 0
        AF[m] = np.asarray(Ak[rows[m], fn[m]]).ravel()
        return fs, fn, Ak.diagonal().copy(), AF


def coarse_pair(hier: Hierarchy, fs, fn, AC, AF, b, level):
    """The training pair of one p' system: level-k system and A_k^{-1} b_k."""
    A = csr_matrix(fs, fn, AC, AF)
    fsk, fnk, ACk, AFk = hier.level_system(A, level)
    bk = hier.restrict(b, level)
    xk = spla.spsolve(csr_matrix(fsk, fnk, ACk, AFk).tocsc(), bk)
    return fsk, fnk, ACk, AFk, bk, xk


def solver_systems(case, iterations, device: int = 0):
    """p' systems (AC, AF, b) of a SIMPLE run, captured after each of the given
    outer iterations (ascending)."""
    s = axs.make_solver(case, device)
    out = []
    cfg = case.config
    done = 0
    for it in iterations:
        cfg.max_iter = it
        s.configure(cfg)
        s.run(continue_run=done > 0)
        done = it
        out.append((s.field("pp_ac").copy(), s.field("pp_af").copy(), s.field("pp_b").copy()))
    s.close()
    return out


def relative_residual(A, b, x) -> float:
    return float(np.linalg.norm(b - A @ x) / max(np.linalg.norm(b), 1e-300))


def cycles_to_tolerance(mg: axs.Multigrid, A, AC, AF, b, x0, tol=1e-6, max_cycles=400, **kw):
    """V-cycles from x0 until the relative residual is at most tol (max_cycles
    if never), and the residual history."""
    x = x0.copy()
    hist = [relative_residual(A, b, x)]
    c = 0
    while hist[-1] > tol and c < max_cycles:
        x = mg.solve(AC, AF, b, x, cycles=1, **kw)
        c += 1
        hist.append(relative_residual(A, b, x))
    return c, np.array(hist)


def guess_start(net: axs.Net, hier: Hierarchy, fs, fn, AC, AF, b, level):
    """The network's start for A x = b: its (energy-scaled) guess on the
    Galerkin system of `level`, prolonged to the mesh."""
    A = csr_matrix(fs, fn, AC, AF)
    fsk, fnk, ACk, AFk = hier.level_system(A, level)
    xk = net.predict(fsk, fnk, ACk, AFk, hier.restrict(b, level))
    return hier.prolong(xk, level)
