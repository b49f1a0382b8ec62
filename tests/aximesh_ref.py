"""ctypes access to the reference's own mesher, compiled in place from
aximesh/source/aximesh.cpp by `make -C oracle ref` into
oracle/_ref/libref_aximesh.so (test infrastructure: a checker for the
product's mesher, present only where the reference sources were)."""
import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
REF_LIB = ROOT / "oracle" / "_ref" / "libref_aximesh.so"

_dp, _ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
_LIB = None


def available() -> bool:
    return REF_LIB.exists()


def _lib():
    global _LIB
    if _LIB is None:
        lib = C.CDLL(str(REF_LIB))
        lib.ref_aximesh_generate.restype = C.c_void_p
        lib.ref_aximesh_generate.argtypes = [C.c_int, _dp, _dp, C.c_int, _ip, C.c_double, C.c_double, C.c_double]
        lib.ref_aximesh_insert.restype = C.c_void_p
        lib.ref_aximesh_insert.argtypes = [C.c_int, _dp, _dp]
        lib.ref_aximesh_error.restype = C.c_char_p
        lib.ref_aximesh_error.argtypes = [C.c_void_p]
        lib.ref_aximesh_counts.argtypes = [C.c_void_p, _ip, _ip, _ip]
        lib.ref_aximesh_copy.argtypes = [C.c_void_p, _dp, _dp, _ip, _ip, _ip]
        lib.ref_aximesh_free.argtypes = [C.c_void_p]
        _LIB = lib
    return _LIB


def _take(h) -> dict:
    lib = _lib()
    try:
        err = lib.ref_aximesh_error(h).decode()
        n, t, s = C.c_int(), C.c_int(), C.c_int()
        lib.ref_aximesh_counts(h, C.byref(n), C.byref(t), C.byref(s))
        x, y = np.zeros(n.value), np.zeros(n.value)
        v = np.zeros(3 * t.value + 1, np.int32)
        a = np.zeros(3 * t.value + 1, np.int32)
        sg = np.zeros(2 * s.value + 1, np.int32)
        lib.ref_aximesh_copy(h, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), v.ctypes.data_as(_ip),
                             a.ctypes.data_as(_ip), sg.ctypes.data_as(_ip))
    finally:
        lib.ref_aximesh_free(h)
    return {"error": err, "points": np.stack([x, y], axis=1), "tri_v": v[:3 * t.value].reshape(-1, 3),
            "tri_adj": a[:3 * t.value].reshape(-1, 3), "segments": sg[:2 * s.value].reshape(-1, 2)}


def generate(points, segments, B=1.4, beta=0.2, classify_tol=1.4) -> dict:
    """AxiMesh::generateMesh(points, segments, len(points), {B, beta, classify_tol})."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 2)
    x, y = np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1])
    s = np.ascontiguousarray(np.asarray(segments, dtype=np.int32).reshape(-1, 2)).ravel()
    if s.size == 0:
        s = np.zeros(1, np.int32)
        ns = 0
    else:
        ns = s.size // 2
    h = _lib().ref_aximesh_generate(x.size, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), ns,
                                    s.ctypes.data_as(_ip), B, beta, classify_tol)
    return _take(h)


def insert_points(points) -> dict:
    """AxiMesh::insertPoints(points, len(points))."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 2)
    x, y = np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1])
    return _take(_lib().ref_aximesh_insert(x.size, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp)))
