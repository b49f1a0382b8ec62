"""Pins of the oracle's mesh quality metrics (Quality::buildQuality,
solver/source/quality.cpp:160-166) by closed forms on crafted cells:

  square cells             aspect ratio 1, non-orthogonality 1, skewness 0
  dz x dr rectangles       aspect ratio dz/dr (the centroid lies dz/2 and dr/2 from its faces)
  right isosceles triangle aspect ratio sqrt(2) (the centroid lies 2A/(3L) from an edge of
                           length L), skewness (90 - 60)/120 = (60 - 45)/60 = 1/4,
                           non-orthogonality 2/sqrt(5) (centroid to a leg's midpoint)
  two triangles of a square the diagonal face: centroid to centroid along its normal
  equilateral triangle     aspect ratio 1, skewness 0, non-orthogonality 1
  jittered nozzle          aspect ratio >= 1, 0 < non-orthogonality <= 1, 0 <= skewness <= 1
"""
import math

import numpy as np
import pytest

import axs
from oracle_binding import mesh_quality


def one_triangle(p0, p1, p2):
    pts = np.array([p0, p1, p2], dtype=float)
    segs = [(*pts[0], *pts[1], 0), (*pts[1], *pts[2], 1), (*pts[2], *pts[0], 2)]
    return axs.triangle_mesh(pts, np.array([[0, 1, 2]]), segs, tol=1e-6)


def test_square_cells():
    m = axs.pipe_case(nz=6, nr=6, L=0.06, R=0.06).mesh
    q = mesh_quality(m)
    np.testing.assert_allclose(q["aspect_ratio"], 1.0, rtol=1e-12)
    np.testing.assert_allclose(q["non_orthogonality"], 1.0, rtol=1e-12)
    np.testing.assert_allclose(q["skewness"], 0.0, atol=1e-12)


@pytest.mark.parametrize("nz,nr", [(10, 4), (5, 20)])
def test_rectangles(nz, nr):
    L, R = 0.2, 0.01
    m = axs.pipe_case(nz=nz, nr=nr, L=L, R=R).mesh
    q = mesh_quality(m)
    dz, dr = L / nz, R / nr
    np.testing.assert_allclose(q["aspect_ratio"], max(dz, dr) / min(dz, dr), rtol=1e-10)
    np.testing.assert_allclose(q["non_orthogonality"], 1.0, rtol=1e-12)
    np.testing.assert_allclose(q["skewness"], 0.0, atol=1e-12)


def test_right_isosceles_triangle():
    q = mesh_quality(one_triangle((0.0, 1.0), (1.0, 1.0), (0.0, 2.0)))
    assert math.isclose(q["aspect_ratio"][0], math.sqrt(2.0), rel_tol=1e-12)
    assert math.isclose(q["skewness"][0], 0.25, rel_tol=1e-12)
    assert math.isclose(q["non_orthogonality"][0], 2.0 / math.sqrt(5.0), rel_tol=1e-12)


def test_square_split_in_two_triangles():
    pts = np.array([[0.0, 1.0], [1.0, 1.0], [0.0, 2.0], [1.0, 2.0]])
    segs = [(0.0, 1.0, 1.0, 1.0, 0), (1.0, 1.0, 1.0, 2.0, 1), (1.0, 2.0, 0.0, 2.0, 2), (0.0, 2.0, 0.0, 1.0, 3)]
    m = axs.triangle_mesh(pts, np.array([[0, 1, 2], [1, 3, 2]]), segs, tol=1e-3)
    q = mesh_quality(m)
    # both halves are right isosceles triangles; the diagonal joins two centroids
    # along its normal, so the interior face adds a 1 and changes nothing
    np.testing.assert_allclose(q["aspect_ratio"], math.sqrt(2.0), rtol=1e-12)
    np.testing.assert_allclose(q["skewness"], 0.25, rtol=1e-12)
    np.testing.assert_allclose(q["non_orthogonality"], 2.0 / math.sqrt(5.0), rtol=1e-12)


def test_equilateral_triangle():
    q = mesh_quality(one_triangle((0.0, 1.0), (1.0, 1.0), (0.5, 1.0 + math.sqrt(3.0) / 2.0)))
    assert math.isclose(q["aspect_ratio"][0], 1.0, rel_tol=1e-12)
    assert abs(q["skewness"][0]) < 1e-12
    assert math.isclose(q["non_orthogonality"][0], 1.0, rel_tol=1e-12)


def test_no_corners_no_skewness():
    m = one_triangle((0.0, 1.0), (1.0, 1.0), (0.0, 2.0))
    q = mesh_quality(m, rings=(np.zeros(2, np.int32), np.zeros(0), np.zeros(0)))
    assert q["skewness"][0] == 0.0 and q["aspect_ratio"][0] > 1.0


def test_jittered_nozzle_ranges():
    m = axs.nozzle_mesh(spacing=4e-4)
    q = mesh_quality(m)
    assert np.all(q["aspect_ratio"] >= 1.0)
    assert np.all((q["non_orthogonality"] > 0.0) & (q["non_orthogonality"] <= 1.0))
    assert np.all((q["skewness"] >= 0.0) & (q["skewness"] <= 1.0))
    # Delaunay triangles of a Poisson-disk set: no sliver
    assert q["skewness"].max() < 0.9 and np.median(q["skewness"]) < 0.4
