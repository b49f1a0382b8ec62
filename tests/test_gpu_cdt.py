"""The solver on meshes from the product's own unstructured front end: a
boundary outline, discretised and meshed by the constrained Delaunay mesher
with frontal refinement (axs_cdt_generate, the reference's
AxiMesh::generateMesh), then built into the finite-volume mesh
(createUnstructuredMesh).  The GPU path against the oracle on the same mesh,
to the bounds of the other parity tests (per-iteration arrays within 1e-12
relative L2, residual histories within 1e-10), and the mesh's quality metrics
against the oracle's."""
import numpy as np
import pytest

import axs
from oracle_binding import Oracle
from oracle_binding import mesh_quality as oracle_quality

pytestmark = pytest.mark.gpu

STATE_FIELDS = ["u", "v", "p", "pp", "mdot", "du", "dv", "u_ac", "u_b", "u_af", "pp_ac", "pp_b", "pp_af"]


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def run_both(case, iters):
    case.config.max_iter = iters
    s = axs.make_solver(case)
    s.run()
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    return s, o


def check_parity(s, o):
    bad = {}
    for name in STATE_FIELDS:
        a, b = s.field(name), o.field(name)
        assert a.shape == b.shape and np.all(np.isfinite(a)), name
        e = rel_l2(a, b)
        if e > 1e-12:
            bad[name] = e
    assert not bad, f"fields beyond 1e-12: {bad}"
    for r in range(3):
        np.testing.assert_allclose(s.history(r), o.history(r), rtol=1e-10, atol=0)


def channel_with_ring(length=0.04, radius=0.01, ring_z=0.012, ring_r=0.005, ring_radius=0.002, n=32):
    """An axisymmetric channel with a torus across the flow: the box loop
    (axis 3, outlet 1, wall 2, inlet 0) and the torus's circle, clockwise, as
    more wall (2)."""
    box = [(0.0, 0.0, length, 0.0, axs.G_AXIS), (length, 0.0, length, radius, axs.G_OUTLET),
           (length, radius, 0.0, radius, axs.G_WALL), (0.0, radius, 0.0, 0.0, axs.G_INLET)]
    th = np.linspace(0.0, 2.0 * np.pi, n + 1)[:-1]
    p = [(ring_z + ring_radius * np.cos(t), ring_r + ring_radius * np.sin(t)) for t in th]
    ring = [(p[(i + 1) % n][0], p[(i + 1) % n][1], p[i][0], p[i][1], axs.G_WALL) for i in reversed(range(n))]
    return [box, ring]


def test_cdt_nozzle_matches_oracle(require_gpu):
    mesh, m = axs.cdt_mesh([axs.nozzle_outline()], 2.5e-4)
    print("CDT nozzle:", mesh.n_cells, "cells,", len(m.points), "points", flush=True)
    assert mesh.n_cells > 20000
    s, o = run_both(axs.nozzle_case(mesh=mesh), 20)
    check_parity(s, o)


def test_cdt_channel_with_ring_matches_oracle(require_gpu):
    mesh, m = axs.cdt_mesh(channel_with_ring(), 3e-4)
    print("CDT channel with ring:", mesh.n_cells, "cells", flush=True)
    u_in = 0.05
    bcs = axs.default_bcs(axs.GROUP_TYPES, {(axs.G_INLET, axs.VAR_U): (axs.BC_DIRICHLET, u_in)})
    cfg = axs.default_config(rho=1000.0, mu=1e-3, use_non_orth=1, max_iter=20)
    s, o = run_both(axs.Case("ring", mesh, bcs, cfg, {}), 20)
    check_parity(s, o)


def test_cdt_mesh_quality_matches_oracle(require_gpu):
    mesh, _ = axs.cdt_mesh([axs.nozzle_outline()], 2.5e-4)
    g, o = axs.mesh_quality(mesh), oracle_quality(mesh)
    np.testing.assert_array_equal(g["aspect_ratio"], o["aspect_ratio"])
    np.testing.assert_array_equal(g["non_orthogonality"], o["non_orthogonality"])
    np.testing.assert_allclose(g["skewness"], o["skewness"], rtol=0, atol=1e-13)
    # frontal refinement makes near-equilateral triangles away from the boundary
    print("CDT nozzle skewness: median", np.median(g["skewness"]), "max", g["skewness"].max())
    assert np.median(g["skewness"]) < 0.2
