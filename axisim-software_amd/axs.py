"""ctypes binding of libaxs.so (include/axs.h) plus the benchmark cases.

The solver runs only through the HIP library: importing this module fails
loudly when libaxs.so has not been built, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libaxs.so"

# ---- enums (include/axs.h) ----
BC_DIRICHLET, BC_NEUMANN, BC_FULLY_DEVELOPED, BC_MICHAELIS_MENTEN, BC_HILL, BC_NONE, BC_PULSATILE = range(7)
BOUNDARY_WALL, BOUNDARY_VELOCITY_INLET, BOUNDARY_PRESSURE_OUTLET, BOUNDARY_SYMMETRY, BOUNDARY_FAR_FIELD = range(5)
CONV_UPWIND, CONV_CENTRAL, CONV_SOU, CONV_QUICK = range(4)
GRAD_GREEN_GAUSS, GRAD_LSQ = range(2)
LINEAR_JACOBI, LINEAR_GS_RB = range(2)
TIME_BDF1, TIME_BDF2 = range(2)
NORM_L1, NORM_L2, NORM_LINF = range(3)
SCALE_NONE, SCALE_N, SCALE_SQRT_N, SCALE_DIAGONAL, SCALE_CONTINUITY = range(5)
PREC_FP64, PREC_MIXED = range(2)
REFINE_CYCLE, REFINE_SOLVE = range(2)
MG_SMOOTH_JACOBI, MG_SMOOTH_RBGS = range(2)
VAR_U, VAR_V, VAR_P, VAR_T, VAR_C = range(5)
RES_U, RES_V, RES_CONTINUITY, RES_TEMPERATURE, RES_CONCENTRATION = range(5)
RES_COUNT = 5
KERNEL_JACOBI_UV, KERNEL_MOMENTUM, KERNEL_MG_SMOOTH, KERNEL_FACE_FLUX, KERNEL_PP_COEFF, KERNEL_JACOBI_UV_FUSED = range(6)

FIELDS = {
    "u": 0, "v": 1, "p": 2, "pp": 3, "grad_uz": 4, "grad_ur": 5, "grad_vz": 6, "grad_vr": 7,
    "grad_pz": 8, "grad_pr": 9, "continuity": 10, "vmag": 11, "cell_re": 12, "mdot": 13,
    "du": 14, "dv": 15,
    "u_ac": 16, "u_b": 17, "u_invac": 18, "u_af": 19,
    "v_ac": 20, "v_b": 21, "v_invac": 22, "v_af": 23,
    "pp_ac": 24, "pp_b": 25, "pp_invac": 26, "pp_af": 27,
    "u_old": 28, "v_old": 29, "u_old2": 30, "v_old2": 31,
    "t": 32, "c": 33, "grad_tz": 34, "grad_tr": 35, "grad_cz": 36, "grad_cr": 37,
    "t_ac": 50, "t_b": 51, "t_invac": 52, "t_af": 53, "c_ac": 54, "c_b": 55, "c_invac": 56, "c_af": 57,
    "t_old": 58, "c_old": 59, "t_old2": 60, "c_old2": 61, "wall_cw": 62, "wall_ocr": 63,
    "inv_dpn": 40, "wp": 41, "dpb": 42, "inv_a2d": 43, "lsq_wz": 44, "lsq_wr": 45,
    "lsq_inv_zz": 46, "lsq_inv_zr": 47, "lsq_inv_rr": 48, "length2d": 49,
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)
_fp = C.POINTER(C.c_float)


class MeshView(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int), ("n_faces", C.c_int),
        ("cell_center_z", _dp), ("cell_center_r", _dp), ("cell_volume", _dp), ("cell_area2d", _dp),
        ("cell_face_start", _ip), ("cell_face_ids", _ip),
        ("face_owner", _ip), ("face_neighbor", _ip),
        ("face_normal_z", _dp), ("face_normal_r", _dp), ("face_center_z", _dp), ("face_center_r", _dp),
        ("face_area", _dp), ("face_group", _ip),
    ]


class BCView(C.Structure):
    _fields_ = [
        ("n_groups", C.c_int), ("type", _u8p), ("boundary_type", _u8p), ("value", _dp),
        ("length", _dp), ("amplitude", _dp), ("frequency", _dp),
        ("vmax", _dp), ("km", _dp), ("hill_n", _dp), ("inhib_m", _dp), ("inhib_k2", _dp), ("inhib_v2", _dp),
        ("wall_resistance", _dp),
    ]


class ResidualConfig(C.Structure):
    _fields_ = [("enabled", C.c_int), ("norm", C.c_int), ("scaling", C.c_int), ("tol", C.c_double)]


class Config(C.Structure):
    _fields_ = [
        ("convection_scheme", C.c_int), ("gradient_scheme", C.c_int), ("linear_solver", C.c_int),
        ("time_scheme", C.c_int), ("linear_max_iter", C.c_int), ("use_multigrid", C.c_int),
        ("add_convection", C.c_int), ("transient", C.c_int), ("use_non_orth", C.c_int),
        ("dt", C.c_double), ("t_end", C.c_double),
        ("max_iter", C.c_int), ("check_conv", C.c_int),
        ("mg_max_iter", C.c_int), ("mg_weight", C.c_double), ("mg_coarse_sweeps", C.c_int),
        ("mg_pre_sweeps", C.c_int), ("mg_post_sweeps", C.c_int),
        ("alpha_u", C.c_double), ("alpha_p", C.c_double), ("rho", C.c_double), ("mu", C.c_double),
        ("residual", ResidualConfig * RES_COUNT),
        ("structured_mesh", C.c_int),
        ("mg_max_levels", C.c_int), ("mg_smoother", C.c_int), ("precision", C.c_int),
        ("coarse_direct", C.c_int),
        ("save_key_frame_iter", C.c_int), ("max_time_frames", C.c_int),
        ("solve_energy", C.c_int), ("solve_concentration", C.c_int),
        ("cp", C.c_double), ("k", C.c_double), ("diffusivity", C.c_double),
        ("mixed_refine", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("converged", C.c_int), ("stopped", C.c_int),
        ("n_cells", C.c_int), ("n_faces", C.c_int), ("n_face_refs", C.c_int),
        ("n_colors", C.c_int), ("mg_levels", C.c_int), ("mg_level_cells", C.c_int * 32),
        ("setup_ms", C.c_double), ("solve_ms", C.c_double), ("last_residual", C.c_double * RES_COUNT),
        ("bytes_per_iteration", C.c_double), ("vcycle_bytes", C.c_double),
        ("sweeps_per_launch", C.c_int), ("sweep_block_cells", C.c_int), ("graph_nodes", C.c_int),
    ]


class CaseStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("last_residual", C.c_double * RES_COUNT),
                ("stop_iteration", C.c_int)]


class BlockSpec(C.Structure):
    _fields_ = [
        ("z0", C.c_double), ("r0", C.c_double),
        ("axial_length", C.c_double), ("axial_cells", C.c_int), ("axial_grading", C.c_int),
        ("axial_coef", C.c_double),
        ("radial_length", C.c_double), ("radial_cells", C.c_int), ("radial_grading", C.c_int),
        ("radial_coef", C.c_double),
        ("edge_group", C.c_int * 4),
    ]


class NozzleSpec(C.Structure):
    _fields_ = [
        ("r_in", C.c_double), ("l_in", C.c_double), ("cone_deg", C.c_double), ("r_throat", C.c_double),
        ("l_throat", C.c_double), ("diff_deg", C.c_double), ("r_out", C.c_double), ("l_out", C.c_double),
        ("spacing", C.c_double), ("jitter", C.c_double),
        ("seed_points", C.c_uint64), ("seed_jitter", C.c_uint64),
    ]


class CellPolygons(C.Structure):
    _fields_ = [
        ("corners", C.c_int), ("quad_z", _dp), ("quad_r", _dp), ("n_points", C.c_int), ("point_z", _dp),
        ("point_r", _dp), ("tris", _ip), ("structured", C.c_int), ("multiblock", C.c_int),
    ]


class CdtParams(C.Structure):
    _fields_ = [("B", C.c_double), ("beta", C.c_double), ("classify_tol", C.c_double),
                ("max_points", C.c_longlong)]


class SolverFile(C.Structure):
    """What the .axislv solver payload holds beyond Config."""
    _fields_ = [
        ("var_units", C.c_uint8 * 15), ("velocity_solver", C.c_int), ("mom_tol", C.c_double),
        ("pp_tol", C.c_double), ("residual_type", C.c_int * RES_COUNT),
    ]


PROGRESS_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_char_p), _dp)


def load_library(path: os.PathLike | None = None) -> C.CDLL:
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"libaxs.so not built ({p}); run __graft_entry__.build() first")
    lib = C.CDLL(str(p))
    S = C.c_void_p
    sig = {
        "axs_config_defaults": (None, [C.POINTER(Config)]),
        "axs_create": (C.c_int, [C.POINTER(S), C.c_int]),
        "axs_destroy": (None, [S]),
        "axs_last_error": (C.c_char_p, [S]),
        "axs_set_mesh": (C.c_int, [S, C.POINTER(MeshView)]),
        "axs_set_bcs": (C.c_int, [S, C.c_int, C.POINTER(BCView)]),
        "axs_configure": (C.c_int, [S, C.POINTER(Config)]),
        "axs_run": (C.c_int, [S, PROGRESS_CB, C.c_void_p, C.c_int]),
        "axs_run_timed": (C.c_int, [S, C.c_int, _dp]),
        "axs_bench_kernel": (C.c_int, [S, C.c_int, C.c_int, _dp, _dp]),
        "axs_bench_kernel_cold": (C.c_int, [S, C.c_int, C.c_int, _dp, _dp]),
        "axs_autotune": (C.c_int, [S, C.c_int, _dp, _dp]),
        "axs_set_sweep_fusion": (C.c_int, [S, C.c_int, C.c_int]),
        "axs_request_stop": (None, [S]),
        "axs_can_continue": (C.c_int, [S, C.c_char_p, C.c_size_t]),
        "axs_get_field": (C.c_int, [S, C.c_int, _dp, C.c_size_t]),
        "axs_field_length": (C.c_size_t, [S, C.c_int]),
        "axs_get_residual_history": (C.c_int, [S, C.c_int, _dp, C.c_int, _ip]),
        "axs_get_stats": (C.c_int, [S, C.POINTER(Stats)]),
        "axs_frame_count": (C.c_int, [S]),
        "axs_get_wall_diagnostic": (C.c_int, [S, _dp]),
        "axs_mesh_get_quads": (C.c_int, [S, _ip, _dp, _dp]),
        "axs_set_cases": (C.c_int, [S, C.c_int, _ip, _ip]),
        "axs_get_case_stats": (C.c_int, [S, C.c_int, C.POINTER(CaseStats)]),
        "axs_save_solution_npy": (C.c_int, [S, C.POINTER(CellPolygons), C.c_char_p]),
        "axs_mesh_quality": (C.c_int, [C.POINTER(MeshView), C.POINTER(CellPolygons), C.c_int, _dp, _dp, _dp]),
        "axs_save_solver_file": (C.c_int, [C.c_char_p, C.POINTER(Config), C.POINTER(SolverFile)]),
        "axs_load_solver_file": (C.c_int, [C.c_char_p, C.POINTER(Config), C.POINTER(SolverFile)]),
        "axs_get_frame": (C.c_int, [S, C.c_int, C.c_int, _dp, _dp, C.c_size_t]),
        "axs_mg_create": (C.c_int, [C.POINTER(S), C.c_int, _ip, _ip, C.c_int]),
        "axs_mg_destroy": (None, [S]),
        "axs_mg_levels": (C.c_int, [S]),
        "axs_mg_level_info": (C.c_int, [S, C.c_int, _ip, _ip]),
        "axs_mg_tail_info": (C.c_int, [S, C.c_int, _ip, _ip]),
        "axs_mg_set_mixed": (C.c_int, [S, C.c_int, C.c_int]),
        "axs_mg_get_maps": (C.c_int, [S, C.c_int, _ip, _ip, _ip, _ip]),
        "axs_mg_solve": (C.c_int, [S, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int]),
        "axs_mg_solve_noalloc": (C.c_int, [S, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                           C.c_int, C.c_int, C.c_int, C.c_void_p]),
        "axs_mg_bench": (C.c_int, [S, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]),
        "axs_mg_bench_cold": (C.c_int, [S, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, _dp, _dp]),
        "axs_build_coloring": (C.c_int, [C.c_int, _ip, _ip, _ip, _ip, _ip]),
        "axs_net_create": (C.c_int, [C.POINTER(S), C.c_int, C.c_uint64, C.c_int]),
        "axs_net_destroy": (None, [S]),
        "axs_net_num_params": (C.c_int, [S]),
        "axs_net_get_params": (C.c_int, [S, _fp]),
        "axs_net_set_params": (C.c_int, [S, _fp]),
        "axs_net_add_sample": (C.c_int, [S, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _ip]),
        "axs_net_clear_samples": (C.c_int, [S]),
        "axs_net_loss_grad": (C.c_int, [S, C.c_int, _dp, _fp]),
        "axs_net_train": (C.c_int, [S, C.c_int, C.c_int, C.c_double, C.c_uint64, _dp]),
        "axs_net_predict": (C.c_int, [S, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]),
        "axs_set_learned_guess": (C.c_int, [S, S, C.c_int]),
        "axs_mesh_multiblock": (C.c_int, [C.POINTER(S), C.c_int, C.POINTER(BlockSpec)]),
        "axs_mesh_triangles": (C.c_int, [C.POINTER(S), C.c_int, _dp, _dp, C.c_int, _ip, C.c_int, _dp, _dp, _dp,
                                         _dp, _ip, C.c_double]),
        "axs_mesh_nozzle": (C.c_int, [C.POINTER(S), C.POINTER(NozzleSpec)]),
        "axs_mesh_get_view": (C.c_int, [S, C.POINTER(MeshView)]),
        "axs_mesh_concat": (C.c_int, [C.POINTER(S), C.c_int, C.POINTER(MeshView), C.c_int]),
        "axs_mesh_get_triangles": (C.c_int, [S, _ip, _dp, _dp, _ip, _ip]),
        "axs_mesh_destroy": (None, [S]),
        "axs_cdt_params_defaults": (None, [C.POINTER(CdtParams)]),
        "axs_cdt_generate": (C.c_int, [C.POINTER(S), C.c_int, _dp, _dp, C.c_int, _ip, C.POINTER(CdtParams)]),
        "axs_cdt_insert_points": (C.c_int, [C.POINTER(S), C.c_int, _dp, _dp]),
        "axs_cdt_counts": (C.c_int, [S, _ip, _ip, _ip]),
        "axs_cdt_get": (C.c_int, [S, _dp, _dp, _ip, _ip, _ip]),
        "axs_cdt_destroy": (None, [S]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


class AxsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"axs status {status}: {msg}")
        self.status = status


def _check(status: int, handle=None) -> None:
    if status != 0:
        msg = lib().axs_last_error(handle)
        raise AxsError(status, msg.decode() if msg else "")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


# ---------------------------------------------------------------------------
# Mesh: plain numpy arrays (FVMeshHostPacked layout)
# ---------------------------------------------------------------------------
MESH_KEYS_D = ["cell_center_z", "cell_center_r", "cell_volume", "cell_area2d", "face_normal_z", "face_normal_r",
               "face_center_z", "face_center_r", "face_area"]
MESH_KEYS_I = ["cell_face_start", "cell_face_ids", "face_owner", "face_neighbor", "face_group"]


@dataclass
class Mesh:
    arrays: dict
    points: np.ndarray | None = None  # (n, 2) for triangle meshes
    tris: np.ndarray | None = None
    quads: np.ndarray | None = None   # (cells, 4, 2) corner (z, r) of multiblock cells

    def polygons(self) -> "CellPolygons | None":
        """The cell outlines for a solution export (None: unknown)."""
        p = CellPolygons()
        keep = {}
        if self.quads is not None:
            keep["qz"] = np.ascontiguousarray(self.quads[:, :, 0]).ravel()
            keep["qr"] = np.ascontiguousarray(self.quads[:, :, 1]).ravel()
            p.corners, p.structured, p.multiblock = 4, 1, 1
            p.quad_z, p.quad_r = _dptr(keep["qz"]), _dptr(keep["qr"])
        elif self.points is not None and self.tris is not None:
            keep["pz"] = np.ascontiguousarray(self.points[:, 0])
            keep["pr"] = np.ascontiguousarray(self.points[:, 1])
            keep["t"] = np.ascontiguousarray(self.tris, dtype=np.int32).ravel()
            p.corners, p.n_points = 3, self.points.shape[0]
            p.point_z, p.point_r, p.tris = _dptr(keep["pz"]), _dptr(keep["pr"]), _iptr(keep["t"])
        else:
            return None
        p._keep = keep
        return p

    @property
    def n_cells(self) -> int:
        return int(self.arrays["cell_center_z"].size)

    @property
    def n_faces(self) -> int:
        return int(self.arrays["face_owner"].size)

    def view(self) -> MeshView:
        a = self.arrays
        v = MeshView()
        v.n_cells = self.n_cells
        v.n_faces = self.n_faces
        for k in MESH_KEYS_D:
            setattr(v, k, _dptr(a[k]))
        for k in MESH_KEYS_I:
            setattr(v, k, _iptr(a[k]))
        return v

    def save(self, path: os.PathLike) -> None:
        extra = {}
        if self.points is not None:
            extra = {"points": self.points, "tris": self.tris}
        if self.quads is not None:
            extra["quads"] = self.quads
        np.savez_compressed(path, **self.arrays, **extra)

    @staticmethod
    def load(path: os.PathLike) -> "Mesh":
        with np.load(path, allow_pickle=False) as z:
            arrays = {k: np.ascontiguousarray(z[k]) for k in MESH_KEYS_D + MESH_KEYS_I}
            pts = z["points"] if "points" in z.files else None
            tris = z["tris"] if "tris" in z.files else None
            quads = z["quads"] if "quads" in z.files else None
        return Mesh(arrays, pts, tris, quads)


def mesh_quality(mesh: Mesh, device: int = 0) -> dict:
    """Quality::buildQuality (solver/source/quality.cpp:160-166) on the GPU:
    per-cell aspect ratio, non-orthogonality and skewness (the last from the
    mesh's quads or triangles; zeros without them)."""
    n = mesh.n_cells
    out = {k: np.zeros(n) for k in ("aspect_ratio", "non_orthogonality", "skewness")}
    v = mesh.view()
    poly = mesh.polygons()
    _check(lib().axs_mesh_quality(C.byref(v), C.byref(poly) if poly is not None else None, device,
                                  _dptr(out["aspect_ratio"]), _dptr(out["non_orthogonality"]),
                                  _dptr(out["skewness"])))
    return out


def _mesh_from_handle(h) -> Mesh:
    v = MeshView()
    _check(lib().axs_mesh_get_view(h, C.byref(v)))
    n, f = v.n_cells, v.n_faces
    s = v.cell_face_start[n]
    sizes = {"cell_center_z": n, "cell_center_r": n, "cell_volume": n, "cell_area2d": n, "face_normal_z": f,
             "face_normal_r": f, "face_center_z": f, "face_center_r": f, "face_area": f,
             "cell_face_start": n + 1, "cell_face_ids": s, "face_owner": f, "face_neighbor": f, "face_group": f}
    arrays = {}
    for k, cnt in sizes.items():
        ptr = getattr(v, k)
        dt = np.float64 if k in MESH_KEYS_D else np.int32
        arrays[k] = np.ctypeslib.as_array(ptr, shape=(cnt,)).astype(dt, copy=True)
    npts, ntri = C.c_int(), C.c_int()
    _check(lib().axs_mesh_get_triangles(h, C.byref(npts), None, None, C.byref(ntri), None))
    pts = tris = None
    if ntri.value > 0:
        pz = np.zeros(npts.value)
        pr = np.zeros(npts.value)
        tr = np.zeros(3 * ntri.value, dtype=np.int32)
        _check(lib().axs_mesh_get_triangles(h, None, _dptr(pz), _dptr(pr), None, _iptr(tr)))
        pts = np.stack([pz, pr], axis=1)
        tris = tr.reshape(-1, 3)
    nq = C.c_int()
    _check(lib().axs_mesh_get_quads(h, C.byref(nq), None, None))
    quads = None
    if nq.value > 0:
        qz = np.zeros(4 * nq.value)
        qr = np.zeros(4 * nq.value)
        _check(lib().axs_mesh_get_quads(h, None, _dptr(qz), _dptr(qr)))
        quads = np.stack([qz.reshape(-1, 4), qr.reshape(-1, 4)], axis=2)
    lib().axs_mesh_destroy(h)
    return Mesh(arrays, pts, tris, quads)


def multiblock_mesh(blocks: list[dict]) -> Mesh:
    """blocks: dicts with z0, r0, axial=(length, cells[, coef]), radial=(length, cells[, coef]),
    groups=(west, east, south, north)."""
    specs = (BlockSpec * len(blocks))()
    for i, b in enumerate(blocks):
        s = specs[i]
        s.z0, s.r0 = b["z0"], b["r0"]
        al, ac = b["axial"][0], b["axial"][1]
        acoef = b["axial"][2] if len(b["axial"]) > 2 else 1.0
        rl, rc = b["radial"][0], b["radial"][1]
        rcoef = b["radial"][2] if len(b["radial"]) > 2 else 1.0
        s.axial_length, s.axial_cells, s.axial_grading, s.axial_coef = al, ac, int(acoef != 1.0), acoef
        s.radial_length, s.radial_cells, s.radial_grading, s.radial_coef = rl, rc, int(rcoef != 1.0), rcoef
        for e in range(4):
            s.edge_group[e] = b["groups"][e]
    h = C.c_void_p()
    _check(lib().axs_mesh_multiblock(C.byref(h), len(blocks), specs))
    return _mesh_from_handle(h)


def triangle_mesh(points: np.ndarray, tris: np.ndarray, segments: list[tuple], tol: float) -> Mesh:
    """segments: (z0, r0, z1, r1, group) boundary segments for face classification."""
    pz = np.ascontiguousarray(points[:, 0], dtype=np.float64)
    pr = np.ascontiguousarray(points[:, 1], dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.int32).ravel()
    seg = np.array([s[:4] for s in segments], dtype=np.float64).reshape(-1, 4)
    sg = np.array([s[4] for s in segments], dtype=np.int32)
    cols = [np.ascontiguousarray(seg[:, i]) for i in range(4)]
    h = C.c_void_p()
    _check(lib().axs_mesh_triangles(C.byref(h), pz.size, _dptr(pz), _dptr(pr), t.size // 3, _iptr(t), len(segments),
                                    _dptr(cols[0]), _dptr(cols[1]), _dptr(cols[2]), _dptr(cols[3]), _iptr(sg), tol))
    return _mesh_from_handle(h)


def nozzle_mesh(spacing: float, jitter: float = 0.05, seed_points: int = 3001, seed_jitter: int = 3002,
                **geometry) -> Mesh:
    g = dict(NOZZLE_GEOMETRY)
    g.update(geometry)
    s = NozzleSpec()
    for k, v in g.items():
        setattr(s, k, v)
    s.spacing, s.jitter, s.seed_points, s.seed_jitter = spacing, jitter, seed_points, seed_jitter
    h = C.c_void_p()
    _check(lib().axs_mesh_nozzle(C.byref(h), C.byref(s)))
    return _mesh_from_handle(h)


# ---------------------------------------------------------------------------
# Unstructured mesher (AxiMesh::generateMesh, aximesh/source/aximesh.cpp)
# ---------------------------------------------------------------------------
@dataclass
class CdtMesh:
    """AxiMesh::Mesh (aximesh/include/aximesh/aximesh.h:80-86): points (n, 2)
    in bin order then refinement points; triangles tri_v (t, 3) wound
    counter-clockwise with tri_adj[t, i] the triangle across edge
    v[i] -> v[i+1] (-1 on the boundary); segments (s, 2) renumbered."""
    points: np.ndarray
    tri_v: np.ndarray
    tri_adj: np.ndarray
    segments: np.ndarray


def _cdt_result(h) -> CdtMesh:
    n, t, s = C.c_int(), C.c_int(), C.c_int()
    try:
        _check(lib().axs_cdt_counts(h, C.byref(n), C.byref(t), C.byref(s)))
        x, y = np.zeros(n.value), np.zeros(n.value)
        v = np.zeros(3 * t.value, dtype=np.int32)
        adj = np.zeros(3 * t.value, dtype=np.int32)
        seg = np.zeros(2 * s.value, dtype=np.int32)
        _check(lib().axs_cdt_get(h, _dptr(x), _dptr(y), _iptr(v), _iptr(adj), _iptr(seg)))
    finally:
        lib().axs_cdt_destroy(h)
    return CdtMesh(np.stack([x, y], axis=1), v.reshape(-1, 3), adj.reshape(-1, 3), seg.reshape(-1, 2))


def cdt_params(B: float | None = None, beta: float | None = None, classify_tol: float | None = None,
               max_points: int = 0) -> CdtParams:
    """AxiMesh::Params with the reference defaults for whatever is not given
    (max_points: the build's point limit, 0 for the default)."""
    p = CdtParams()
    lib().axs_cdt_params_defaults(C.byref(p))
    if B is not None:
        p.B = B
    if beta is not None:
        p.beta = beta
    if classify_tol is not None:
        p.classify_tol = classify_tol
    p.max_points = max_points
    return p


def cdt_generate(points: np.ndarray, segments: np.ndarray, params: CdtParams | None = None) -> CdtMesh:
    """AxiMesh::generateMesh(points, segments, len(points), params):
    constrained Delaunay triangulation, flood-fill removal of the outside,
    frontal refinement (aximesh.cpp:1105-1175)."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 2)
    x, y = np.ascontiguousarray(pts[:, 0]), np.ascontiguousarray(pts[:, 1])
    seg = np.ascontiguousarray(np.asarray(segments, dtype=np.int32).reshape(-1, 2)).ravel()
    h = C.c_void_p()
    _check(lib().axs_cdt_generate(C.byref(h), x.size, _dptr(x), _dptr(y), seg.size // 2, _iptr(seg),
                                  C.byref(params) if params is not None else None))
    return _cdt_result(h)


def cdt_insert_points(points: np.ndarray) -> CdtMesh:
    """AxiMesh::insertPoints(points, len(points)) (aximesh.cpp:466-489): the
    Delaunay triangulation of the points in the given order inside the super
    triangle, whose three vertices follow the input points."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 2)
    x, y = np.ascontiguousarray(pts[:, 0]), np.ascontiguousarray(pts[:, 1])
    h = C.c_void_p()
    _check(lib().axs_cdt_insert_points(C.byref(h), x.size, _dptr(x), _dptr(y)))
    return _cdt_result(h)


def nozzle_outline(**geometry) -> list[tuple]:
    """The nozzle's boundary loop, counter-clockwise, as (z0, r0, z1, r1, group)
    edges: axis (3), outlet (1), wall (2), inlet (0)."""
    g = dict(NOZZLE_GEOMETRY)
    g.update(geometry)
    tc, td = np.tan(np.radians(g["cone_deg"])), np.tan(np.radians(g["diff_deg"]))
    z = [0.0, g["l_in"]]
    r = [g["r_in"], g["r_in"]]
    z.append(z[-1] + abs(g["r_in"] - g["r_throat"]) / tc); r.append(g["r_throat"])
    z.append(z[-1] + g["l_throat"]); r.append(g["r_throat"])
    z.append(z[-1] + abs(g["r_out"] - g["r_throat"]) / td); r.append(g["r_out"])
    z.append(z[-1] + g["l_out"]); r.append(g["r_out"])
    L = z[-1]
    edges = [(0.0, 0.0, L, 0.0, 3), (L, 0.0, L, r[-1], 1)]
    edges += [(z[i], r[i], z[i - 1], r[i - 1], 2) for i in range(len(z) - 1, 0, -1)]
    edges.append((0.0, r[0], 0.0, 0.0, 0))
    return edges


def outline_boundary(loops: list[list[tuple]], spacing: float):
    """Boundary points and segments of closed loops of (z0, r0, z1, r1, group)
    edges, each edge cut into ceil(length / spacing) equal pieces (the input
    the GUI hands the mesher, Mesh::runAxiMeshTriangulation,
    solver/source/mesh.cpp:1216-1223).  Returns points (n, 2), segments
    (n, 2) and each segment's group."""
    pts, segs, groups = [], [], []
    for loop in loops:
        first = len(pts)
        for z0, r0, z1, r1, g in loop:
            n = max(1, int(np.ceil(np.hypot(z1 - z0, r1 - r0) / spacing - 1e-9)))
            for k in range(n):
                t = k / n
                pts.append((z0 + t * (z1 - z0), r0 + t * (r1 - r0)))
                segs.append((len(pts) - 1, len(pts)))
                groups.append(g)
        segs[-1] = (segs[-1][0], first)
    return np.array(pts), np.array(segs, dtype=np.int32), np.array(groups, dtype=np.int32)


def cdt_mesh(loops: list[list[tuple]], spacing: float, params: CdtParams | None = None) -> tuple[Mesh, CdtMesh]:
    """The unstructured front end end to end: loops discretised at `spacing`,
    meshed by the constrained Delaunay mesher, and built into the
    finite-volume mesh (createUnstructuredMesh), boundary faces taking the
    group of the loop edge they lie on."""
    pts, segs, _ = outline_boundary(loops, spacing)
    m = cdt_generate(pts, segs, params)
    edges = [e for loop in loops for e in loop]
    return triangle_mesh(m.points, m.tri_v, edges, tol=1e-6 * spacing), m


# ---------------------------------------------------------------------------
# Boundary conditions
# ---------------------------------------------------------------------------
@dataclass
class BC:
    """Per-group boundary condition of one variable."""
    type: list
    value: list
    length: list | None = None
    boundary_type: list | None = None
    amplitude: list | None = None
    frequency: list | None = None
    # Michaelis-Menten / Hill wall kinetics per group (None: defaults)
    vmax: list | None = None
    km: list | None = None
    hill_n: list | None = None
    inhib_m: list | None = None
    inhib_k2: list | None = None
    inhib_v2: list | None = None
    wall_resistance: list | None = None

    KINETICS = ("vmax", "km", "hill_n", "inhib_m", "inhib_k2", "inhib_v2", "wall_resistance")

    def view(self):
        n = len(self.type)
        keep = {
            "type": np.asarray(self.type, dtype=np.uint8),
            "value": np.asarray(self.value, dtype=np.float64),
            "length": np.asarray(self.length if self.length is not None else [0.0] * n, dtype=np.float64),
            "bt": np.asarray(self.boundary_type if self.boundary_type is not None else [0] * n, dtype=np.uint8),
            "amp": np.asarray(self.amplitude if self.amplitude is not None else [0.0] * n, dtype=np.float64),
            "freq": np.asarray(self.frequency if self.frequency is not None else [0.0] * n, dtype=np.float64),
        }
        v = BCView()
        v.n_groups = n
        v.type = keep["type"].ctypes.data_as(_u8p)
        v.boundary_type = keep["bt"].ctypes.data_as(_u8p)
        v.value = _dptr(keep["value"])
        v.length = _dptr(keep["length"])
        v.amplitude = _dptr(keep["amp"])
        v.frequency = _dptr(keep["freq"])
        for name in self.KINETICS:
            vals = getattr(self, name)
            if vals is not None:
                keep[name] = np.asarray(vals, dtype=np.float64)
                setattr(v, name, _dptr(keep[name]))
        return v, keep


# Default BC type of each (boundary type, variable) (getDefaultBCType,
# manager/source/boundary_func.cpp:117-197)
_DEFAULT_BC = {  # u, v, p, T, C
    BOUNDARY_VELOCITY_INLET: (BC_DIRICHLET, BC_DIRICHLET, BC_NEUMANN, BC_DIRICHLET, BC_DIRICHLET),
    BOUNDARY_PRESSURE_OUTLET: (BC_NEUMANN, BC_NEUMANN, BC_DIRICHLET, BC_NEUMANN, BC_NEUMANN),
    BOUNDARY_WALL: (BC_DIRICHLET, BC_DIRICHLET, BC_NEUMANN, BC_NEUMANN, BC_NEUMANN),
    BOUNDARY_SYMMETRY: (BC_NEUMANN, BC_DIRICHLET, BC_NEUMANN, BC_NEUMANN, BC_NEUMANN),
    BOUNDARY_FAR_FIELD: (BC_DIRICHLET, BC_DIRICHLET, BC_NEUMANN, BC_NEUMANN, BC_NEUMANN),
}


def default_bcs(group_types: list[int], overrides: dict | None = None, scalars: bool = False) -> dict:
    """BCs of u, v, p (and T, C when `scalars`) for groups of the given boundary
    types, default values 0.  overrides: {(group, var): (type, value)}."""
    out = {}
    for var in ((VAR_U, VAR_V, VAR_P, VAR_T, VAR_C) if scalars else (VAR_U, VAR_V, VAR_P)):
        types = [_DEFAULT_BC[t][var] for t in group_types]
        values = [0.0] * len(group_types)
        out[var] = BC(types, values, boundary_type=list(group_types))
    for (g, var), spec in (overrides or {}).items():
        bc = out[var]
        bc.type[g] = spec[0]
        bc.value[g] = spec[1]
        if len(spec) > 2:
            bc.amplitude = bc.amplitude or [0.0] * len(group_types)
            bc.frequency = bc.frequency or [0.0] * len(group_types)
            bc.amplitude[g], bc.frequency[g] = spec[2], spec[3]
    return out


def save_solver_file(path: os.PathLike, cfg: Config, extra: SolverFile | None = None) -> None:
    """Write the .axislv solver payload (format 5) of cfg."""
    _check(lib().axs_save_solver_file(os.fsencode(path), C.byref(cfg), C.byref(extra) if extra is not None else None))


def load_solver_file(path: os.PathLike, cfg: Config | None = None) -> tuple[Config, SolverFile]:
    """Read a .axislv solver payload over cfg (default: the defaults)."""
    out = Config()
    if cfg is None:
        lib().axs_config_defaults(C.byref(out))
    else:
        C.pointer(out)[0] = cfg
    extra = SolverFile()
    _check(lib().axs_load_solver_file(os.fsencode(path), C.byref(out), C.byref(extra)))
    return out, extra


def default_config(**kw) -> Config:
    c = Config()
    lib().axs_config_defaults(C.byref(c))
    for k, v in kw.items():
        if k == "residual_tol":
            for i, t in enumerate(v):
                c.residual[i].tol = t
        else:
            setattr(c, k, v)
    return c


# ---------------------------------------------------------------------------
# Solver
# ---------------------------------------------------------------------------
class Solver:
    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().axs_create(C.byref(self._h), device))
        self._keep = []
        self.mesh: Mesh | None = None

    def close(self) -> None:
        if self._h:
            lib().axs_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mesh(self, mesh: Mesh) -> None:
        self.mesh = mesh
        v = mesh.view()
        _check(lib().axs_set_mesh(self._h, C.byref(v)), self._h)

    def set_bcs(self, bcs: dict) -> None:
        for var, bc in bcs.items():
            v, keep = bc.view()
            _check(lib().axs_set_bcs(self._h, var, C.byref(v)), self._h)

    def configure(self, cfg: Config) -> None:
        self.cfg = cfg
        _check(lib().axs_configure(self._h, C.byref(cfg)), self._h)

    def run(self, progress=None, continue_run: bool = False) -> None:
        if progress is None:
            cb = PROGRESS_CB()
        else:
            def _cb(user, it, n, names, values):
                progress(it, [values[i] for i in range(n)])
            cb = PROGRESS_CB(_cb)
        self._cb = cb
        _check(lib().axs_run(self._h, cb, None, int(continue_run)), self._h)

    def run_timed(self, n: int) -> float:
        ms = C.c_double()
        _check(lib().axs_run_timed(self._h, n, C.byref(ms)), self._h)
        return ms.value

    def autotune(self, repeats: int = 20) -> tuple[float, float]:
        """Choose workgroup sizes by timing (after a run); returns the summed
        kernel ms at the default and at the chosen sizes."""
        a, b = C.c_double(), C.c_double()
        _check(lib().axs_autotune(self._h, repeats, C.byref(a), C.byref(b)), self._h)
        return a.value, b.value

    def set_sweep_fusion(self, sweeps: int, block_cells: int = 0) -> None:
        """Momentum Jacobi sweeps per launch (1 to 5) and cells per fused
        block (512 or 1024; 0: the larger one whose halos fit) for the next
        run; bit-identical results."""
        _check(lib().axs_set_sweep_fusion(self._h, sweeps, block_cells), self._h)

    def set_learned_guess(self, net: "Net | None", level: int = -1) -> None:
        """Learned initial guess for the first p' solve of each outer iteration
        (None: off).  level: multigrid level the net runs on (-1: automatic)."""
        self._net = net
        _check(lib().axs_set_learned_guess(self._h, net._h if net is not None else None, level), self._h)

    def bench_kernel(self, kernel: int, repeats: int = 50, cold: bool = False) -> tuple[float, float]:
        """(average launch ms, algorithmic bytes per launch) of one kernel;
        cold: each launch after 512 MiB of writes have flushed the caches."""
        ms, by = C.c_double(), C.c_double()
        fn = lib().axs_bench_kernel_cold if cold else lib().axs_bench_kernel
        _check(fn(self._h, kernel, repeats, C.byref(ms), C.byref(by)), self._h)
        return ms.value, by.value

    def request_stop(self) -> None:
        lib().axs_request_stop(self._h)

    def can_continue(self) -> tuple[bool, str]:
        buf = C.create_string_buffer(256)
        ok = lib().axs_can_continue(self._h, buf, 256)
        return bool(ok), buf.value.decode()

    def field(self, name: str) -> np.ndarray:
        fid = FIELDS[name]
        n = lib().axs_field_length(self._h, fid)
        out = np.zeros(n)
        _check(lib().axs_get_field(self._h, fid, _dptr(out), n), self._h)
        return out

    def history(self, res: int) -> np.ndarray:
        cnt = C.c_int()
        _check(lib().axs_get_residual_history(self._h, res, None, 0, C.byref(cnt)), self._h)
        out = np.zeros(cnt.value)
        if cnt.value:
            _check(lib().axs_get_residual_history(self._h, res, _dptr(out), cnt.value, C.byref(cnt)), self._h)
        return out

    def set_cases(self, cell_starts, face_starts) -> None:
        """Batched run: cases as contiguous cell / face ranges (starts incl. the end)."""
        cs = np.ascontiguousarray(cell_starts, dtype=np.int32)
        fs = np.ascontiguousarray(face_starts, dtype=np.int32)
        _check(lib().axs_set_cases(self._h, cs.size - 1, _iptr(cs), _iptr(fs)), self._h)

    def case_stats(self, c: int) -> CaseStats:
        out = CaseStats()
        _check(lib().axs_get_case_stats(self._h, c, C.byref(out)), self._h)
        return out

    def frame_count(self) -> int:
        return lib().axs_frame_count(self._h)

    def frames(self) -> list:
        """[(time, {"u", "v", "p"[, "t"][, "c"]})] of the transient frames."""
        out = []
        N = self.mesh.n_cells
        solved = [("u", 0), ("v", 1), ("p", 2)]
        if getattr(self, "cfg", None) is not None:
            solved += [("t", 32)] * bool(self.cfg.solve_energy) + [("c", 33)] * bool(self.cfg.solve_concentration)
        for i in range(lib().axs_frame_count(self._h)):
            t = C.c_double()
            fields = {}
            for name, fid in solved:
                a = np.zeros(N)
                _check(lib().axs_get_frame(self._h, i, fid, C.byref(t), _dptr(a), N), self._h)
                fields[name] = a
            out.append((t.value, fields))
        return out

    def wall_diagnostic(self) -> np.ndarray:
        out = np.zeros(10)
        _check(lib().axs_get_wall_diagnostic(self._h, _dptr(out)), self._h)
        return out

    def save_solution_npy(self, directory: os.PathLike, mesh: "Mesh | None" = None) -> None:
        """NumPy export of the last solve (solution.npy, points/cells.npy,
        frame_NNNN.npy, meta.json) into `directory`."""
        mesh = mesh if mesh is not None else self.mesh
        poly = mesh.polygons() if mesh is not None else None
        _check(lib().axs_save_solution_npy(self._h, C.byref(poly) if poly is not None else None,
                                           os.fsencode(directory)), self._h)

    def stats(self) -> Stats:
        s = Stats()
        _check(lib().axs_get_stats(self._h, C.byref(s)), self._h)
        return s


class Multigrid:
    """Stand-alone agglomeration multigrid on the device."""

    def __init__(self, face_start: np.ndarray, face_neighbor: np.ndarray, max_levels: int = 0):
        self.fs = np.ascontiguousarray(face_start, dtype=np.int32)
        self.fn = np.ascontiguousarray(face_neighbor, dtype=np.int32)
        self.n = self.fs.size - 1
        self._h = C.c_void_p()
        _check(lib().axs_mg_create(C.byref(self._h), self.n, _iptr(self.fs), _iptr(self.fn), max_levels))

    def close(self):
        if self._h:
            lib().axs_mg_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def levels(self) -> int:
        return lib().axs_mg_levels(self._h)

    def level_info(self, level: int) -> tuple[int, int]:
        n, r = C.c_int(), C.c_int()
        _check(lib().axs_mg_level_info(self._h, level, C.byref(n), C.byref(r)))
        return n.value, r.value

    def tail_info(self, precision: int = PREC_FP64) -> tuple[int, int]:
        """(first level of the single-workgroup LDS tail or -1, its LDS bytes)"""
        lv, by = C.c_int(), C.c_int()
        _check(lib().axs_mg_tail_info(self._h, precision, C.byref(lv), C.byref(by)))
        return lv.value, by.value

    def maps(self, level: int) -> dict:
        n, r = self.level_info(level)
        fs = np.zeros(n + 1, np.int32)
        fn = np.zeros(r, np.int32)
        coarsest = level == self.levels - 1
        c2c = np.zeros(0 if coarsest else n, np.int32)
        s2s = np.zeros(0 if coarsest else r, np.int32)
        _check(lib().axs_mg_get_maps(self._h, level, _iptr(fs), _iptr(fn),
                                     _iptr(c2c) if c2c.size else None, _iptr(s2s) if s2s.size else None))
        return {"face_start": fs, "face_neighbor": fn, "cell_to_coarse": c2c, "slot_to_coarse_slot": s2s}

    def set_mixed(self, refine: int = REFINE_CYCLE, fuse_level0: bool = True) -> None:
        """Mixed-precision refinement (REFINE_CYCLE / REFINE_SOLVE) of the
        solves below, and whether level 0 may run fused across cycles."""
        _check(lib().axs_mg_set_mixed(self._h, refine, int(fuse_level0)))

    def solve(self, AC, AF, b, x, cycles=3, weight=0.6, pre=1, post=1, coarse=30, precision=PREC_FP64,
              coarse_direct=0, smoother=MG_SMOOTH_JACOBI) -> np.ndarray:
        AC, AF, b = (np.ascontiguousarray(a, dtype=np.float64) for a in (AC, AF, b))
        x = np.array(x, dtype=np.float64, copy=True)
        _check(lib().axs_mg_solve(self._h, _dptr(AC), _dptr(AF), _dptr(b), _dptr(x), cycles, weight, pre, post,
                                  coarse, precision, coarse_direct, smoother))
        return x

    def solve_noalloc(self, AC: int, AF: int, b: int, x: int, cycles=3, weight=0.6, pre=1, post=1, coarse=30,
                      stream: int | None = None) -> None:
        """axs_mg_solve_noalloc on device arrays given by address (x in/out);
        enqueued on `stream` (None: the hierarchy's stream), not waited for."""
        _check(lib().axs_mg_solve_noalloc(self._h, AC, AF, b, x, cycles, weight, pre, post, coarse, stream))

    def bench(self, AC, AF, b, cycles=3, weight=0.6, pre=1, post=1, coarse=30, precision=PREC_FP64,
              coarse_direct=0, smoother=MG_SMOOTH_JACOBI, repeats=20, cold=False):
        """(ms per solve, algorithmic bytes per solve, x); cold: each replay after
        a 512 MiB cache flush (x is then not returned: None)."""
        AC, AF, b = (np.ascontiguousarray(a, dtype=np.float64) for a in (AC, AF, b))
        ms, by = C.c_double(), C.c_double()
        if cold:
            _check(lib().axs_mg_bench_cold(self._h, _dptr(AC), _dptr(AF), _dptr(b), cycles, weight, pre, post,
                                           coarse, precision, coarse_direct, smoother, repeats, C.byref(ms),
                                           C.byref(by)))
            return ms.value, by.value, None
        x = np.zeros(self.n)
        _check(lib().axs_mg_bench(self._h, _dptr(AC), _dptr(AF), _dptr(b), cycles, weight, pre, post, coarse,
                                  precision, coarse_direct, smoother, repeats, C.byref(ms), C.byref(by),
                                  _dptr(x)))
        return ms.value, by.value, x


class Net:
    """Message-passing network for the learned initial guess, trained on the GPU."""

    def __init__(self, layers: int = 4, seed: int = 0, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().axs_net_create(C.byref(self._h), layers, seed, device))
        self.layers = layers

    def close(self):
        if self._h:
            lib().axs_net_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return lib().axs_net_num_params(self._h)

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        _check(lib().axs_net_get_params(self._h, _fptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(lib().axs_net_set_params(self._h, _fptr(p)))

    def add_sample(self, fs, fn, AC, AF, b, x) -> int:
        fs, fn = (np.ascontiguousarray(a, dtype=np.int32) for a in (fs, fn))
        AC, AF, b, x = (np.ascontiguousarray(a, dtype=np.float64) for a in (AC, AF, b, x))
        idx = C.c_int()
        _check(lib().axs_net_add_sample(self._h, fs.size - 1, _iptr(fs), _iptr(fn), _dptr(AC), _dptr(AF),
                                        _dptr(b), _dptr(x), C.byref(idx)))
        return idx.value

    def clear_samples(self) -> None:
        _check(lib().axs_net_clear_samples(self._h))

    def loss_grad(self, sample: int):
        loss = C.c_double()
        g = np.zeros(self.num_params, np.float32)
        _check(lib().axs_net_loss_grad(self._h, sample, C.byref(loss), _fptr(g)))
        return loss.value, g

    def train(self, epochs: int, batch: int = 8, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(epochs)
        _check(lib().axs_net_train(self._h, epochs, batch, lr, seed, _dptr(losses)))
        return losses

    def predict(self, fs, fn, AC, AF, b) -> np.ndarray:
        fs, fn = (np.ascontiguousarray(a, dtype=np.int32) for a in (fs, fn))
        AC, AF, b = (np.ascontiguousarray(a, dtype=np.float64) for a in (AC, AF, b))
        x = np.zeros(fs.size - 1)
        _check(lib().axs_net_predict(self._h, fs.size - 1, _iptr(fs), _iptr(fn), _dptr(AC), _dptr(AF), _dptr(b),
                                     _dptr(x)))
        return x


def coloring(face_start: np.ndarray, face_neighbor: np.ndarray):
    fs = np.ascontiguousarray(face_start, dtype=np.int32)
    fn = np.ascontiguousarray(face_neighbor, dtype=np.int32)
    n = fs.size - 1
    cs = np.zeros(n + 1, np.int32)
    order = np.zeros(n, np.int32)
    nc = C.c_int()
    _check(lib().axs_build_coloring(n, _iptr(fs), _iptr(fn), _iptr(cs), _iptr(order), C.byref(nc)))
    return cs[: nc.value + 1].copy(), order


# ---------------------------------------------------------------------------
# Benchmark / parity cases (the configurations of the benchmark suite)
# ---------------------------------------------------------------------------
# group ids used by every case
G_INLET, G_OUTLET, G_WALL, G_AXIS = 0, 1, 2, 3
GROUP_TYPES = [BOUNDARY_VELOCITY_INLET, BOUNDARY_PRESSURE_OUTLET, BOUNDARY_WALL, BOUNDARY_SYMMETRY]

NOZZLE_GEOMETRY = dict(r_in=0.006, l_in=0.03, cone_deg=20.0, r_throat=0.002, l_throat=0.05, diff_deg=10.0,
                       r_out=0.006, l_out=0.05)


@dataclass
class Case:
    name: str
    mesh: Mesh
    bcs: dict
    config: Config
    info: dict = field(default_factory=dict)


def pipe_case(nz: int = 128, nr: int = 32, L: float = 0.2, R: float = 0.01, u_in: float = 0.005,
              rho: float = 1000.0, mu: float = 1e-3, tol: float = 1e-6, max_iter: int = 200000, **cfg) -> Case:
    """Steady laminar pipe flow on a structured r-z grid (Re = rho*u_in*2R/mu = 100
    for the defaults): first-order upwind, Green-Gauss, FP64, weighted-Jacobi MG."""
    mesh = multiblock_mesh([dict(z0=0.0, r0=0.0, axial=(L, nz), radial=(R, nr),
                                 groups=(G_INLET, G_OUTLET, G_AXIS, G_WALL))])
    bcs = default_bcs(GROUP_TYPES, {(G_INLET, VAR_U): (BC_DIRICHLET, u_in)})
    c = default_config(rho=rho, mu=mu, gradient_scheme=GRAD_GREEN_GAUSS, convection_scheme=CONV_UPWIND,
                       max_iter=max_iter, structured_mesh=1, residual_tol=(tol,) * RES_COUNT)
    for k, v in cfg.items():
        setattr(c, k, v)
    return Case("pipe", mesh, bcs, c, dict(L=L, R=R, u_in=u_in, rho=rho, mu=mu))


def scalar_pipe_case(nz: int = 64, nr: int = 16, energy: bool = True, concentration: bool = True,
                     wall_kinetics: int = BC_MICHAELIS_MENTEN, t_in: float = 310.0, t_wall: float = 300.0,
                     c_in: float = 0.2, vmax: float = 1e-6, km: float = 0.01, hill_n: float = 1.0,
                     inhib_v2: float = 0.0, inhib_k2: float = 0.05, inhib_m: float = 2.0,
                     wall_resistance: float = 0.0, **kw) -> Case:
    """The pipe with heat and species transport: inlet temperature t_in into a
    wall held at t_wall; inlet concentration c_in consumed at the wall by
    Michaelis-Menten (or Hill) kinetics with Vmax, Km (n, optional substrate
    inhibition V2, K2, m) through an extra wall resistance."""
    case = pipe_case(nz=nz, nr=nr, **kw)
    u_in = case.info["u_in"]
    bcs = default_bcs(GROUP_TYPES, {(G_INLET, VAR_U): (BC_DIRICHLET, u_in), (G_INLET, VAR_T): (BC_DIRICHLET, t_in),
                                    (G_WALL, VAR_T): (BC_DIRICHLET, t_wall), (G_INLET, VAR_C): (BC_DIRICHLET, c_in)},
                      scalars=True)
    cb = bcs[VAR_C]
    n = len(cb.type)
    cb.type[G_WALL] = wall_kinetics
    cb.vmax, cb.km, cb.hill_n = [0.0] * n, [0.0] * n, [1.0] * n
    cb.inhib_m, cb.inhib_k2, cb.inhib_v2, cb.wall_resistance = [1.0] * n, [0.0] * n, [0.0] * n, [0.0] * n
    cb.vmax[G_WALL], cb.km[G_WALL], cb.hill_n[G_WALL] = vmax, km, hill_n
    cb.inhib_m[G_WALL], cb.inhib_k2[G_WALL], cb.inhib_v2[G_WALL] = inhib_m, inhib_k2, inhib_v2
    cb.wall_resistance[G_WALL] = wall_resistance
    case.bcs = bcs
    c = case.config
    c.solve_energy, c.solve_concentration = int(energy), int(concentration)
    c.residual[RES_TEMPERATURE].enabled = int(energy)
    c.residual[RES_CONCENTRATION].enabled = int(concentration)
    case.name = "scalar_pipe"
    return case


def expansion_case(nz=(137, 887), nr=(85, 171), bias: float = 1.05, re_throat: float = 500.0, rho=1000.0,
                   mu=1e-3, tol=1e-6, max_iter=200000, **cfg) -> Case:
    """Structured three-block sudden expansion: r 0.002 -> 0.006 at z = 0.04,
    length 0.3; geometric radial grading toward each wall; SOU + LSQ."""
    r1, r2, z1, L = 0.002, 0.006, 0.04, 0.3
    u_in = re_throat * mu / (rho * 2 * r1)
    # geometric progression ratio `bias` per cell toward the wall of each band:
    # band 1 (0..r1) grows toward r1, band 2 (r1..r2) grows toward r2
    inv = 1.0 / bias
    blocks = [
        dict(z0=0.0, r0=0.0, axial=(z1, nz[0]), radial=(r1, nr[0], inv), groups=(G_INLET, -1, G_AXIS, G_WALL)),
        dict(z0=z1, r0=0.0, axial=(L - z1, nz[1]), radial=(r1, nr[0], inv), groups=(-1, G_OUTLET, G_AXIS, -1)),
        dict(z0=z1, r0=r1, axial=(L - z1, nz[1]), radial=(r2 - r1, nr[1], inv), groups=(G_WALL, G_OUTLET, -1, G_WALL)),
    ]
    mesh = multiblock_mesh(blocks)
    bcs = default_bcs(GROUP_TYPES, {(G_INLET, VAR_U): (BC_DIRICHLET, u_in)})
    c = default_config(rho=rho, mu=mu, gradient_scheme=GRAD_LSQ, convection_scheme=CONV_SOU, max_iter=max_iter,
                       structured_mesh=1, residual_tol=(tol,) * RES_COUNT)
    for k, v in cfg.items():
        setattr(c, k, v)
    return Case("expansion", mesh, bcs, c, dict(u_in=u_in, r1=r1, r2=r2))


def nozzle_spacing_for(cells: float) -> float:
    """Poisson-disk radius giving about `cells` triangles on the nozzle (the
    sampler places ~0.72 points per r^2 of area, two triangles per point)."""
    g = NOZZLE_GEOMETRY
    lc = (g["r_in"] - g["r_throat"]) / math.tan(math.radians(g["cone_deg"]))
    ld = (g["r_out"] - g["r_throat"]) / math.tan(math.radians(g["diff_deg"]))
    area = (g["r_in"] * g["l_in"] + 0.5 * (g["r_in"] + g["r_throat"]) * lc + g["r_throat"] * g["l_throat"]
            + 0.5 * (g["r_out"] + g["r_throat"]) * ld + g["r_out"] * g["l_out"])
    return math.sqrt(area * 2 * NOZZLE_DENSITY / cells)


NOZZLE_DENSITY = 0.6273


def nozzle_case(cells: float = 1.0e6, spacing: float | None = None, re_throat=500.0, rho=1000.0, mu=1e-3, tol=1e-6,
                max_iter=200000, mesh: Mesh | None = None, **cfg) -> Case:
    """Unstructured triangular nozzle (Poisson-disk + 5 % jitter + Delaunay),
    non-orthogonal corrector on, agglomeration MG."""
    if mesh is None:
        mesh = nozzle_mesh(spacing if spacing is not None else nozzle_spacing_for(cells))
    g = NOZZLE_GEOMETRY
    u_throat = re_throat * mu / (rho * 2 * g["r_throat"])
    u_in = u_throat * (g["r_throat"] / g["r_in"]) ** 2
    bcs = default_bcs(GROUP_TYPES, {(G_INLET, VAR_U): (BC_DIRICHLET, u_in)})
    c = default_config(rho=rho, mu=mu, use_non_orth=1, max_iter=max_iter, residual_tol=(tol,) * RES_COUNT)
    for k, v in cfg.items():
        setattr(c, k, v)
    return Case("nozzle", mesh, bcs, c, dict(u_in=u_in, u_throat=u_throat))


def sweep_cases(count: int = 64, n_min: float = 2.0e4, n_max: float = 5.0e5, amplitude: float = 0.3,
                frequency: float = 1.0, seed_template: int = 4001, seed_size: int = 4002, seed_params: int = 4003,
                rho: float = 1000.0, mu: float = 1e-3) -> list[Case]:
    """A parameter sweep of axisymmetric cases for a batched run: template
    pipe / expansion / nozzle uniformly (seed_template), cell count
    log-uniform in [n_min, n_max] (seed_size), radius uniform in
    [0.002, 0.01] m and Reynolds number uniform in [50, 1000] (seed_params), and
    a pulsatile inlet u = mean (1 + amplitude sin(2 pi frequency t)).  Pipe and
    nozzle geometries scale with the radius; the expansion keeps its own."""
    rt, rn, rp = (np.random.default_rng(s) for s in (seed_template, seed_size, seed_params))
    out = []
    for i in range(count):
        t = int(rt.integers(3))
        N = float(np.exp(rn.uniform(np.log(n_min), np.log(n_max))))
        R = float(rp.uniform(0.002, 0.01))
        Re = float(rp.uniform(50.0, 1000.0))
        if t == 0:
            nr = max(4, int(round(math.sqrt(N / 8.0))))
            nz = max(8, int(round(N / nr)))
            case = pipe_case(nz=nz, nr=nr, L=20.0 * R, R=R, u_in=Re * mu / (rho * 2.0 * R), rho=rho, mu=mu)
        elif t == 1:
            f = math.sqrt(N / (137 * 85 + 887 * 256))
            case = expansion_case(nz=(max(4, round(137 * f)), max(8, round(887 * f))),
                                  nr=(max(4, round(85 * f)), max(4, round(171 * f))), re_throat=Re, rho=rho, mu=mu)
        else:
            scale = R / NOZZLE_GEOMETRY["r_in"]
            geo = {k: (v * scale if k not in ("cone_deg", "diff_deg") else v) for k, v in NOZZLE_GEOMETRY.items()}
            mesh = nozzle_mesh(nozzle_spacing_for(N) * scale, seed_points=5000 + i, seed_jitter=6000 + i, **geo)
            # nozzle_case sets the throat velocity from the unscaled throat radius
            case = nozzle_case(mesh=mesh, re_throat=Re / scale, rho=rho, mu=mu)
        u = case.bcs[VAR_U]
        u.type[G_INLET] = BC_PULSATILE
        u.amplitude = [0.0] * len(u.type)
        u.frequency = [0.0] * len(u.type)
        u.amplitude[G_INLET], u.frequency[G_INLET] = amplitude, frequency
        case.info.update(template=("pipe", "expansion", "nozzle")[t], radius=R, re=Re)
        out.append(case)
    return out


def concat_meshes(meshes: list[Mesh], groups_per_mesh: int = 4) -> Mesh:
    """Disjoint union of meshes for a batched run (mesh c's boundary group g
    becomes g + c * groups_per_mesh)."""
    views = (MeshView * len(meshes))()
    for i, m in enumerate(meshes):
        views[i] = m.view()
    h = C.c_void_p()
    _check(lib().axs_mesh_concat(C.byref(h), len(meshes), views, groups_per_mesh))
    return _mesh_from_handle(h)  # copies the arrays and releases the handle


def batch_case(cases: list[Case], groups_per_mesh: int = 4, **cfg) -> Case:
    """Many cases as one: the union mesh, every case's BCs on its own groups, the
    configuration of the first case (overridden by cfg).  info["ranges"] holds
    each case's (first cell, cell count) in the union."""
    mesh = concat_meshes([c.mesh for c in cases], groups_per_mesh)
    bcs = {}
    defaults = (("type", BC_NONE), ("value", 0.0), ("length", 0.0), ("boundary_type", 0), ("amplitude", 0.0),
                ("frequency", 0.0), ("vmax", 0.0), ("km", 0.0), ("hill_n", 1.0), ("inhib_m", 1.0),
                ("inhib_k2", 0.0), ("inhib_v2", 0.0), ("wall_resistance", 0.0))
    for var in sorted(cases[0].bcs):
        cols = {k: [] for k, _ in defaults}
        for c in cases:
            bc = c.bcs[var]
            n = len(bc.type)
            if n > groups_per_mesh:
                raise ValueError("a case has more boundary groups than groups_per_mesh")
            pad = groups_per_mesh - n
            for k, default in defaults:
                vals = getattr(bc, k)
                cols[k] += (list(vals) if vals is not None else [default] * n) + [default] * pad
        bcs[var] = BC(**cols)
    c0 = cases[0].config
    config = Config()
    C.pointer(config)[0] = c0
    for k, v in cfg.items():
        setattr(config, k, v)
    ranges, off = [], 0
    cell_starts, face_starts = [0], [0]
    for c in cases:
        ranges.append((off, c.mesh.n_cells))
        off += c.mesh.n_cells
        cell_starts.append(cell_starts[-1] + c.mesh.n_cells)
        face_starts.append(face_starts[-1] + c.mesh.n_faces)
    return Case("batch", mesh, bcs, config, dict(ranges=ranges, names=[c.name for c in cases],
                                                 cell_starts=cell_starts, face_starts=face_starts))


def make_solver(case: Case, device: int = 0, per_case: bool = False) -> Solver:
    """A solver on the case; per_case (batched cases): each case converges on
    its own residuals (axs_set_cases)."""
    s = Solver(device)
    s.set_mesh(case.mesh)
    s.set_bcs(case.bcs)
    s.configure(case.config)
    if per_case:
        s.set_cases(case.info["cell_starts"], case.info["face_starts"])
    return s
