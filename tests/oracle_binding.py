"""ctypes binding of oracle/liboracle.so -- the FP64 CPU restatement used as
the checker in the parity tests and as bench.py's CPU baseline."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import axs

ROOT = Path(__file__).resolve().parent.parent
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_L = None


def olib() -> C.CDLL:
    global _L
    if _L is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError("liboracle.so not built; run __graft_entry__.build()")
        L = C.CDLL(str(ORACLE_LIB))
        V = C.c_void_p
        sig = {
            "orc_mesh_quality": (C.c_int, [C.POINTER(axs.MeshView), _ip, _dp, _dp, _dp, _dp, _dp]),
            "orc_create": (V, [C.POINTER(axs.MeshView), C.POINTER(axs.Config)]),
            "orc_destroy": (None, [V]),
            "orc_set_bcs": (C.c_int, [V, C.c_int, C.POINTER(axs.BCView)]),
            "orc_run": (C.c_int, [V, C.c_int]),
            "orc_configure": (None, [V, C.POINTER(axs.Config)]),
            "orc_set_stop_at": (None, [V, C.c_int]),
            "orc_iterations": (C.c_int, [V]),
            "orc_converged": (C.c_int, [V]),
            "orc_get_residual_history": (C.c_int, [V, C.c_int, _dp, C.c_int]),
            "orc_get_field": (C.c_int, [V, C.c_int, _dp, C.c_size_t]),
            "orc_gradient": (C.c_int, [V, C.c_int, C.c_int, _dp, _dp, _dp]),
            "orc_build_csr": (C.c_int, [V, _ip, _ip]),
            "orc_build_coloring": (C.c_int, [C.c_int, _ip, _ip, _ip, _ip]),
            "orc_mg_create": (V, [C.c_int, _ip, _ip, C.c_int]),
            "orc_mg_destroy": (None, [V]),
            "orc_mg_levels": (C.c_int, [V]),
            "orc_mg_level_info": (C.c_int, [V, C.c_int, _ip, _ip]),
            "orc_mg_get_maps": (C.c_int, [V, C.c_int, _ip, _ip, _ip, _ip]),
            "orc_mg_solve": (C.c_int, [V, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                       C.c_int]),
            "orc_mg_get_operator": (C.c_int, [V, C.c_int, _dp, _dp]),
            "orc_residual_value": (None, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, _dp]),
            "orc_probe_face_values": (C.c_int, [V, C.c_int, _dp, _dp]),
            "orc_probe_transport": (C.c_int, [V, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                              C.c_int, C.c_double, _dp, _dp, _dp]),
            "orc_probe_transient": (C.c_int, [V, _dp, _dp, C.c_double, C.c_double, _dp, _dp]),
            "orc_probe_face_flux": (C.c_int, [V, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_probe_pp_coeff": (C.c_int, [V, C.c_double, _dp, _dp, _dp, _dp]),
            "orc_probe_pp_rhs": (C.c_int, [V, C.c_double, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp]),
            "orc_probe_mass_flux_update": (C.c_int, [V, C.c_double, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp]),
            "orc_probe_bc_value": (C.c_double, [V, C.c_int, C.c_int, C.c_double]),
            "orc_bc_current": (C.c_double, [V, C.c_int, C.c_int]),
            "orc_frame_count": (C.c_int, [V]),
            "orc_wall_diagnostic": (C.c_int, [V, _dp]),
            "orc_probe_wall_concentration": (C.c_double, [C.c_int] + [C.c_double] * 8),
            "orc_get_frame": (C.c_int, [V, C.c_int, C.c_int, _dp, _dp, C.c_size_t]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_get_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L.orc_set_threads(1)  # one thread unless a caller asks for more (set_threads)
        _L = L
    return _L


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dn(a):
    """Pointer to a float64 copy of a, or NULL for None."""
    return None if a is None else _d(_f64(a))


def _i(a):
    return a.ctypes.data_as(_ip)


def use_library(path) -> None:
    """Load the oracle from another build of oracle/oracle.cpp (bench.py's
    host-tuned CPU baseline); before the first call into the oracle."""
    global ORACLE_LIB
    if _L is not None:
        raise RuntimeError("the oracle library is already loaded")
    ORACLE_LIB = Path(path)


def set_threads(n: int) -> int:
    """Threads of the oracle's per-cell / per-face loops (results do not depend
    on it); returns the count in effect."""
    olib().orc_set_threads(int(n))
    return olib().orc_get_threads()


class Oracle:
    """CPU FP64 SIMPLE solver (oracle/oracle.cpp)."""

    def __init__(self, mesh: axs.Mesh, bcs: dict, cfg: axs.Config):
        self.mesh = mesh
        v = mesh.view()
        self._h = olib().orc_create(C.byref(v), C.byref(cfg))
        if not self._h:
            raise RuntimeError("orc_create failed")
        for var, bc in bcs.items():
            bv, keep = bc.view()
            if olib().orc_set_bcs(self._h, var, C.byref(bv)) != 0:
                raise RuntimeError("orc_set_bcs failed")

    def __del__(self):
        if getattr(self, "_h", None):
            olib().orc_destroy(self._h)
            self._h = None

    def run(self, continue_run: bool = False) -> int:
        return olib().orc_run(self._h, int(continue_run))

    def configure(self, cfg: axs.Config) -> None:
        olib().orc_configure(self._h, C.byref(cfg))

    def stop_at(self, iteration: int) -> None:
        olib().orc_set_stop_at(self._h, iteration)

    @property
    def iterations(self) -> int:
        return olib().orc_iterations(self._h)

    @property
    def converged(self) -> bool:
        return bool(olib().orc_converged(self._h))

    def history(self, res: int) -> np.ndarray:
        n = olib().orc_get_residual_history(self._h, res, None, 0)
        out = np.zeros(n)
        if n:
            olib().orc_get_residual_history(self._h, res, _d(out), n)
        return out

    def field(self, name: str) -> np.ndarray:
        fid = axs.FIELDS[name]
        m = self.mesh
        S = int(m.arrays["cell_face_start"][-1])
        if name in ("mdot", "inv_dpn", "wp", "dpb", "lsq_wz", "lsq_wr", "length2d", "wall_cw", "wall_ocr"):
            n = m.n_faces
        elif name.endswith("_af"):
            n = S
        else:
            n = m.n_cells
        out = np.zeros(n)
        rc = olib().orc_get_field(self._h, fid, _d(out), n)
        if rc != 0:
            raise RuntimeError(f"orc_get_field({name}) failed: {rc}")
        return out

    def gradient(self, var: int, scheme: int, phi: np.ndarray):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        gz = np.zeros_like(phi)
        gr = np.zeros_like(phi)
        olib().orc_gradient(self._h, var, scheme, _d(phi), _d(gz), _d(gr))
        return gz, gr

    # ---- single kernels on caller data (closed-form pins) ----
    def _n(self):
        return self.mesh.n_cells, self.mesh.n_faces, int(self.mesh.arrays["cell_face_start"][-1])

    def face_values(self, var: int, phi) -> np.ndarray:
        """phi at every face seen from its owner, boundary faces from var's BCs."""
        N, F, S = self._n()
        out = np.zeros(F)
        olib().orc_probe_face_values(self._h, var, _d(_f64(phi)), _d(out))
        return out

    def transport(self, var, phi, gz, gr, mdot, non_ortho, gamma, conv, scheme, flux_scale=1.0):
        """Transport coefficients (AC, AF, b) of one field from zero."""
        N, F, S = self._n()
        AC, AF, b = np.zeros(N), np.zeros(S), np.zeros(N)
        olib().orc_probe_transport(self._h, var, _dn(phi), _dn(gz), _dn(gr), _dn(mdot),
                                   int(non_ortho), gamma, int(conv), scheme, flux_scale, _d(AC), _d(AF), _d(b))
        return AC, AF, b

    def transient(self, old, old2, cap, dt):
        N, F, S = self._n()
        AC, b = np.zeros(N), np.zeros(N)
        olib().orc_probe_transient(self._h, _dn(old), _dn(old2), cap, dt, _d(AC), _d(b))
        return AC, b

    def face_flux(self, rho, u, v, p, gpz, gpr, DU, DV) -> np.ndarray:
        N, F, S = self._n()
        out = np.zeros(F)
        olib().orc_probe_face_flux(self._h, rho, *(_d(_f64(a)) for a in (u, v, p, gpz, gpr, DU, DV)), _d(out))
        return out

    def pp_coeff(self, rho, DU, DV):
        N, F, S = self._n()
        AC, AF = np.zeros(N), np.zeros(S)
        olib().orc_probe_pp_coeff(self._h, rho, _d(_f64(DU)), _d(_f64(DV)), _d(AC), _d(AF))
        return AC, AF

    def pp_rhs(self, rho, mdot, DU, DV, gpz, gpr, cross) -> np.ndarray:
        N, F, S = self._n()
        b = np.zeros(N)
        olib().orc_probe_pp_rhs(self._h, rho, *(_d(_f64(a)) for a in (mdot, DU, DV, gpz, gpr)), int(cross), _d(b))
        return b

    def mass_flux_update(self, rho, mdot, pp, DU, DV, gpz, gpr, non_ortho) -> np.ndarray:
        out = np.array(mdot, dtype=np.float64, copy=True)
        olib().orc_probe_mass_flux_update(self._h, rho, *(_d(_f64(a)) for a in (pp, DU, DV, gpz, gpr)),
                                          int(non_ortho), _d(out))
        return out

    def bc_value(self, var: int, group: int, t: float) -> float:
        return olib().orc_probe_bc_value(self._h, var, group, t)

    def bc_current(self, var: int, group: int) -> float:
        return olib().orc_bc_current(self._h, var, group)

    def frames(self) -> list:
        """[(time, {"u", "v", "p"[, "t"][, "c"]})] of the transient frames."""
        out = []
        N = self.mesh.n_cells
        for i in range(olib().orc_frame_count(self._h)):
            t = C.c_double()
            fields = {}
            for name, fid in (("u", 0), ("v", 1), ("p", 2), ("t", 32), ("c", 33)):
                a = np.zeros(N)
                if olib().orc_get_frame(self._h, i, fid, C.byref(t), _d(a), N) == 0:
                    fields[name] = a
            out.append((t.value, fields))
        return out

    def wall_diagnostic(self) -> np.ndarray:
        out = np.zeros(10)
        olib().orc_wall_diagnostic(self._h, _d(out))
        return out

    def csr(self):
        m = self.mesh
        fs = np.zeros(m.n_cells + 1, np.int32)
        fn = np.zeros(int(m.arrays["cell_face_start"][-1]), np.int32)
        olib().orc_build_csr(self._h, _i(fs), _i(fn))
        return fs, fn


def wall_concentration(hill, vmax, km, n, m, k2, v2, cp, h) -> float:
    """The oracle's wall Newton solve alone."""
    return olib().orc_probe_wall_concentration(int(hill), vmax, km, n, m, k2, v2, cp, h)


def oracle_coloring(fs: np.ndarray, fn: np.ndarray):
    n = fs.size - 1
    cs = np.zeros(n + 1, np.int32)
    order = np.zeros(n, np.int32)
    nc = olib().orc_build_coloring(n, _i(fs), _i(fn), _i(cs), _i(order))
    return cs[: nc + 1].copy(), order


class OracleMG:
    def __init__(self, fs: np.ndarray, fn: np.ndarray, max_levels: int = 0):
        self.fs = np.ascontiguousarray(fs, dtype=np.int32)
        self.fn = np.ascontiguousarray(fn, dtype=np.int32)
        self.n = self.fs.size - 1
        self._h = olib().orc_mg_create(self.n, _i(self.fs), _i(self.fn), max_levels)

    def __del__(self):
        if getattr(self, "_h", None):
            olib().orc_mg_destroy(self._h)
            self._h = None

    @property
    def levels(self) -> int:
        return olib().orc_mg_levels(self._h)

    def level_info(self, l: int):
        n, r = C.c_int(), C.c_int()
        olib().orc_mg_level_info(self._h, l, C.byref(n), C.byref(r))
        return n.value, r.value

    def maps(self, l: int) -> dict:
        n, r = self.level_info(l)
        coarsest = l == self.levels - 1
        fs = np.zeros(n + 1, np.int32)
        fn = np.zeros(r, np.int32)
        c2c = np.zeros(0 if coarsest else n, np.int32)
        s2s = np.zeros(0 if coarsest else r, np.int32)
        olib().orc_mg_get_maps(self._h, l, _i(fs), _i(fn), _i(c2c) if c2c.size else None,
                               _i(s2s) if s2s.size else None)
        return {"face_start": fs, "face_neighbor": fn, "cell_to_coarse": c2c, "slot_to_coarse_slot": s2s}

    def solve(self, AC, AF, b, x, cycles=3, weight=0.6, pre=1, post=1, coarse=30, smoother=0) -> np.ndarray:
        AC, AF, b = (np.ascontiguousarray(a, dtype=np.float64) for a in (AC, AF, b))
        x = np.array(x, dtype=np.float64, copy=True)
        olib().orc_mg_solve(self._h, _d(AC), _d(AF), _d(b), _d(x), cycles, weight, pre, post, coarse, smoother)
        return x

    def operator(self, l: int):
        n, r = self.level_info(l)
        AC = np.zeros(n)
        AF = np.zeros(r)
        olib().orc_mg_get_operator(self._h, l, _d(AC), _d(AF))
        return AC, AF


def residual_value(is_continuity, norm, scaling, res, scale, scale_iteration, scale_val=0.0):
    res = np.ascontiguousarray(res, dtype=np.float64)
    scale = np.ascontiguousarray(scale, dtype=np.float64)
    sv = C.c_double(scale_val)
    rv = C.c_double(0.0)
    olib().orc_residual_value(int(is_continuity), norm, scaling, _d(res), _d(scale), res.size, scale_iteration,
                              C.byref(sv), C.byref(rv))
    return rv.value, sv.value


def corner_rings(mesh):
    """(cornerStart, z, r) of the cells' corner rings: 4 per multiblock cell, the
    three points of each triangle; empty rings without either."""
    n = mesh.n_cells
    if mesh.quads is not None:
        return (np.arange(n + 1, dtype=np.int32) * 4, np.ascontiguousarray(mesh.quads[:, :, 0]).ravel(),
                np.ascontiguousarray(mesh.quads[:, :, 1]).ravel())
    if mesh.points is not None and mesh.tris is not None:
        t = np.asarray(mesh.tris).ravel()
        return (np.arange(n + 1, dtype=np.int32) * 3, np.ascontiguousarray(mesh.points[t, 0]),
                np.ascontiguousarray(mesh.points[t, 1]))
    return np.zeros(n + 1, np.int32), np.zeros(0), np.zeros(0)


def mesh_quality(mesh, rings=None):
    """The oracle's Quality::buildQuality: aspect ratio, non-orthogonality and
    skewness per cell; rings = (cornerStart, z, r) overrides the mesh's own."""
    cs, pz, pr = rings if rings is not None else corner_rings(mesh)
    cs = np.ascontiguousarray(cs, dtype=np.int32)
    pz, pr = _f64(pz), _f64(pr)
    n = mesh.n_cells
    a, o, k = np.zeros(n), np.zeros(n), np.zeros(n)
    v = mesh.view()
    rc = olib().orc_mesh_quality(C.byref(v), _i(cs), _d(pz), _d(pr), _d(a), _d(o), _d(k))
    assert rc == 0
    return {"aspect_ratio": a, "non_orthogonality": o, "skewness": k}
