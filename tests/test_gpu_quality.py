"""Mesh quality metrics on the GPU (axs_mesh_quality) against the oracle's
Quality::buildQuality (solver/source/quality.cpp:160-166): aspect ratio and
non-orthogonality bit for bit (square roots and divisions are correctly rounded
on both sides), skewness within 1e-13 (acos of the device library and of the
host's may differ in the last bit)."""
import numpy as np
import pytest

import axs
from oracle_binding import mesh_quality as oracle_quality

pytestmark = pytest.mark.gpu


def check(mesh):
    g, o = axs.mesh_quality(mesh), oracle_quality(mesh)
    np.testing.assert_array_equal(g["aspect_ratio"], o["aspect_ratio"])
    np.testing.assert_array_equal(g["non_orthogonality"], o["non_orthogonality"])
    np.testing.assert_allclose(g["skewness"], o["skewness"], rtol=0, atol=1e-13)
    return g


def test_quality_nozzle_1m(require_gpu):
    g = check(axs.nozzle_mesh(axs.nozzle_spacing_for(1.0e6)))
    print("nozzle: aspect ratio max", g["aspect_ratio"].max(), "skewness max", g["skewness"].max(),
          "non-orthogonality min", g["non_orthogonality"].min())


def test_quality_graded_expansion(require_gpu):
    g = check(axs.expansion_case().mesh)
    assert g["skewness"].max() < 1e-12  # axis-aligned graded quads
    assert g["aspect_ratio"].max() > 1.0


def test_quality_without_polygons(require_gpu):
    m = axs.pipe_case(nz=8, nr=4).mesh
    bare = axs.Mesh(m.arrays)
    g = axs.mesh_quality(bare)
    assert not g["skewness"].any()
    np.testing.assert_array_equal(g["aspect_ratio"], oracle_quality(m)["aspect_ratio"])
