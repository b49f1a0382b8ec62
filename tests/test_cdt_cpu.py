"""The unstructured mesher (AxiMesh::generateMesh / AxiMesh::insertPoints,
aximesh/source/aximesh.cpp) through the C ABI, on the CPU (it is host code).

* insertPoints on the reference harness's cases (aximesh/tools/aximesh_cli.cpp:
  446-490), whose expected triangle sets its authors checked against a
  brute-force Delaunay triangulation, plus the harness's invariants (2n+1
  triangles, counter-clockwise, areas tiling the super triangle, reciprocal
  links, three hull edges, no vertex inside a circumcircle).
* generateMesh against the reference's own output (tests/golden/
  aximesh_golden.npz, written by the reference compiled in place), point for
  point and bit for bit, triangle for triangle with its rotation and links,
  and against the compiled reference directly where oracle/_ref holds it.
  The harness's generateMesh counts (2n-h-2) predate the flood fill and the
  refinement and the reference itself no longer meets them (with no segments
  the flood fill keeps nothing; an interior segment flips inside and outside),
  so the product is held to the reference's output, and to the invariants
  that hold for a domain bounded by closed loops.
* errors where the reference throws, and the rejections where its behaviour
  is undefined.
* the reference's own harness (aximesh/tools/aximesh_cli.cpp), compiled in
  place once on the reference's mesher and once on the product's drop-in
  header and library (oracle/Makefile `harness`): the two reports agree line
  for line.
* the front end end to end: outline -> mesher -> finite-volume mesh, with the
  volume of revolution and the boundary groups."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import aximesh_ref
import axs

GOLDEN = Path(__file__).resolve().parent / "golden" / "aximesh_golden.npz"
HARNESS_REF = aximesh_ref.ROOT / "oracle" / "_ref" / "aximesh_cli_ref"
HARNESS_AXS = aximesh_ref.ROOT / "oracle" / "_ref" / "aximesh_cli_axs"


def orient(a, b, p):
    return (b[..., 0] - a[..., 0]) * (p[..., 1] - a[..., 1]) - (b[..., 1] - a[..., 1]) * (p[..., 0] - a[..., 0])


def canonical(tris):
    """Each triangle rotated to start at its lowest vertex (winding kept), sorted."""
    out = []
    for t in np.asarray(tris):
        i = int(np.argmin(t))
        out.append((int(t[i]), int(t[(i + 1) % 3]), int(t[(i + 2) % 3])))
    return sorted(out)


def links_reciprocal(v, adj) -> int:
    """Links that are not reciprocal: adj[t, i] must carry edge v[t, i+1] -> v[t, i]
    and point back at t."""
    bad = 0
    for t in range(len(v)):
        for i in range(3):
            u = adj[t, i]
            if u < 0:
                continue
            if u == t or u >= len(v):
                bad += 1
                continue
            a, b = v[t, i], v[t, (i + 1) % 3]
            if not any(v[u, j] == b and v[u, (j + 1) % 3] == a and adj[u, j] == t for j in range(3)):
                bad += 1
    return bad


def inside_circumcircle(A, B, Cc, P) -> bool:
    """The reference harness's check: circumcircle relative to C, strictly
    inside with a relative slack of 1e-9 so cocircular points read outside."""
    ax, ay = A[0] - Cc[0], A[1] - Cc[1]
    bx, by = B[0] - Cc[0], B[1] - Cc[1]
    d = 2.0 * (ax * by - ay * bx)
    if abs(d) < 1e-12:
        return False
    a2, b2 = ax * ax + ay * ay, bx * bx + by * by
    ux, uy = (by * a2 - ay * b2) / d, (ax * b2 - bx * a2) / d
    ox, oy = Cc[0] + ux, Cc[1] + uy
    return (P[0] - ox) ** 2 + (P[1] - oy) ** 2 < (ux * ux + uy * uy) * (1.0 - 1e-9)


# aximesh_cli.cpp:446-490: points in insertion order, expected triangles
# (vertices 0..n-1 the points, n..n+2 the super triangle); empty: cocircular
INSERT_CASES = {
    "one point": ([(0.30, 0.20)], [(0, 1, 2), (0, 2, 3), (0, 3, 1)]),
    "no swap needed": ([(0.25, 0.25), (0.75, 0.25), (0.50, 0.75)],
                       [(0, 1, 2), (0, 2, 5), (0, 3, 4), (0, 4, 1), (0, 5, 3), (1, 4, 5), (1, 5, 2)]),
    "one swap": ([(1.0, 1.0), (-1.0, 0.5)], [(0, 1, 3), (0, 3, 4), (0, 4, 1), (1, 2, 3), (1, 4, 2)]),
    "cascade": ([(1.0, 0.0), (0.6, 0.9), (-0.4, 1.0), (-1.0, 0.1), (-0.5, -0.8), (0.5, -0.7), (0.05, 0.05)],
                [(0, 1, 6), (0, 5, 8), (0, 6, 5), (0, 8, 9), (0, 9, 1), (1, 2, 6), (1, 9, 2), (2, 3, 6),
                 (2, 9, 3), (3, 4, 6), (3, 7, 4), (3, 9, 7), (4, 5, 6), (4, 7, 8), (4, 8, 5)]),
    "cocircular": ([(0.2, 0.2), (0.8, 0.2), (0.8, 0.8), (0.2, 0.8)], []),
}


@pytest.mark.parametrize("name", list(INSERT_CASES))
def test_insert_points_reference_cases(name):
    pts, expected = INSERT_CASES[name]
    n = len(pts)
    m = axs.cdt_insert_points(np.array(pts))
    P, v, adj = m.points, m.tri_v, m.tri_adj
    assert P.shape == (n + 3, 2)
    assert np.array_equal(P[n:], [[-100.0, -100.0], [100.0, -100.0], [0.0, 100.0]])
    assert len(v) == 2 * n + 1
    assert v.min() >= 0 and v.max() < n + 3
    area2 = orient(P[v[:, 0]], P[v[:, 1]], P[v[:, 2]])
    assert (area2 > 0).all()
    assert abs(area2.sum() - 40000.0) < 1e-6 * 40000.0
    assert not (adj == np.arange(len(v))[:, None]).any()
    assert links_reciprocal(v, adj) == 0
    assert (adj < 0).sum() == 3
    for t in v:
        for p in range(n + 3):
            if p not in t:
                assert not inside_circumcircle(P[t[0]], P[t[1]], P[t[2]], P[p])
    if expected:
        assert canonical(v) == sorted(expected)


def golden_keys():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return sorted({k.split(".")[0] for k in z.files})


def golden_case(key):
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k.split(".", 1)[1]: z[k] for k in z.files if k.split(".")[0] == key}


def assert_same_mesh(m, ref):
    assert m.points.shape == ref["points"].shape
    assert np.array_equal(m.points.view(np.uint64), ref["points"].view(np.uint64)), "points differ in their bits"
    assert np.array_equal(m.tri_v, ref["tri_v"])
    assert np.array_equal(m.tri_adj, ref["tri_adj"])
    assert np.array_equal(m.segments, ref["segments"])


@pytest.mark.parametrize("key", golden_keys())
def test_generate_matches_reference_golden(key):
    g = golden_case(key)
    B, beta, tol = (float(x) for x in g["params"])
    prm = axs.cdt_params(B=B, beta=beta, classify_tol=tol)
    err = str(g["error"])
    if err:
        with pytest.raises(axs.AxsError) as ex:
            axs.cdt_generate(g["in_points"], g["in_segments"], prm)
        assert ex.value.status == 1 and err in str(ex.value)
        return
    assert_same_mesh(axs.cdt_generate(g["in_points"], g["in_segments"], prm), g)


def _box():
    return [(0.0, 0.0, 4.0, 0.0, 3), (4.0, 0.0, 4.0, 1.0, 1), (4.0, 1.0, 0.0, 1.0, 2), (0.0, 1.0, 0.0, 0.0, 0)]


def _circle(cz, cr, radius, n, group, clockwise=False):
    th = np.linspace(0.0, 2.0 * np.pi, n + 1)[:-1]
    p = [(cz + radius * np.cos(t), cr + radius * np.sin(t)) for t in th]
    edges = [(p[i][0], p[i][1], p[(i + 1) % n][0], p[(i + 1) % n][1], group) for i in range(n)]
    if clockwise:
        edges = [(e[2], e[3], e[0], e[1], e[4]) for e in reversed(edges)]
    return edges


def live_cases():
    rng = np.random.default_rng(11)
    c = {}
    for h in (5e-4, 2.5e-4, 1.25e-4):
        p, s, _ = axs.outline_boundary([axs.nozzle_outline()], h)
        c[f"nozzle h={h}"] = (p, s, ())
    for cw in (False, True):
        p, s, _ = axs.outline_boundary([_box(), _circle(1.0, 0.5, 0.2, 48, 4, cw)], 0.03)
        c[f"obstacle cw={cw}"] = (p, s, ())
    p, s, _ = axs.outline_boundary([_box(), _circle(1.0, 0.5, 0.2, 24, 4), _circle(2.6, 0.4, 0.15, 16, 5)], 0.05)
    c["two obstacles"] = (p, s, ())
    p, s, _ = axs.outline_boundary([_box()], 0.05)
    c["box + 2000 interior points"] = (np.vstack([p, 0.05 + rng.random((2000, 2)) * [3.9, 0.9]]), s, ())
    pts = rng.random((500, 2)) * [3.0, 1.0]
    seg = [(int(a), int(b)) for a, b in rng.integers(0, 500, (10, 2)) if a != b]
    c["cloud with crossing segments"] = (pts, np.array(seg), ())
    p, s, _ = axs.outline_boundary([axs.nozzle_outline()], 5e-4)
    c["nozzle jittered"] = (p + rng.normal(0.0, 1e-5, p.shape), s, ())
    p, s, _ = axs.outline_boundary([axs.nozzle_outline()], 1e-3)
    c["nozzle beta 0.05"] = (p, s, (1.4, 0.05, 1.4))
    c["nozzle coarse accept"] = (p, s, (1.4, 0.3, 2.5))
    return c


@pytest.mark.skipif(not aximesh_ref.available(), reason="reference mesher not built (make -C oracle ref)")
@pytest.mark.parametrize("name", list(live_cases()))
def test_generate_matches_reference_build(name):
    pts, segs, prm = live_cases()[name]
    ref = aximesh_ref.generate(pts, segs, *prm) if prm else aximesh_ref.generate(pts, segs)
    params = axs.cdt_params(*prm) if prm else None
    if ref["error"]:
        with pytest.raises(axs.AxsError, match=ref["error"]):
            axs.cdt_generate(pts, segs, params)
        return
    assert len(ref["tri_v"]) > 0
    assert_same_mesh(axs.cdt_generate(pts, segs, params), ref)


@pytest.mark.skipif(not aximesh_ref.available(), reason="reference mesher not built (make -C oracle ref)")
def test_insert_points_matches_reference_build():
    rng = np.random.default_rng(5)
    lattice = np.stack(np.meshgrid(np.arange(12) / 11, np.arange(12) / 11), axis=-1).reshape(-1, 2)
    dup = np.round(rng.random((500, 2)) * 8) / 8
    # generic points; a shuffled lattice (cocircular quadruples everywhere);
    # repeated points, on which the reference's point search stalls
    for pts in (rng.random((2000, 2)), lattice[rng.permutation(len(lattice))], dup):
        ref = aximesh_ref.insert_points(pts)
        if ref["error"]:
            with pytest.raises(axs.AxsError, match=ref["error"]):
                axs.cdt_insert_points(pts)
            continue
        m = axs.cdt_insert_points(pts)
        assert np.array_equal(m.tri_v, ref["tri_v"]) and np.array_equal(m.tri_adj, ref["tri_adj"])
        assert np.array_equal(m.points, ref["points"])


def loop_area_and_volume(loops):
    """Signed area (shoelace) and volume of revolution about r = 0
    (V = -pi/3 sum (z1 - z0)(r0^2 + r0 r1 + r1^2)) of closed edge loops."""
    area = vol = 0.0
    for loop in loops:
        for z0, r0, z1, r1, _ in loop:
            area += 0.5 * (z0 * r1 - z1 * r0)
            vol += -np.pi / 3.0 * (z1 - z0) * (r0 * r0 + r0 * r1 + r1 * r1)
    return area, vol


@pytest.mark.parametrize("domain", ["nozzle", "obstacle"])
def test_domain_mesh_invariants(domain):
    if domain == "nozzle":
        loops, h, holes = [axs.nozzle_outline()], 5e-4, 0
    else:
        loops, h, holes = [_box(), _circle(1.0, 0.5, 0.2, 48, 4, clockwise=True)], 0.03, 1
    pts, segs, _ = axs.outline_boundary(loops, h)
    m = axs.cdt_generate(pts, segs)
    P, v, adj = m.points, m.tri_v, m.tri_adj
    area2 = orient(P[v[:, 0]], P[v[:, 1]], P[v[:, 2]])
    assert (area2 > 0).all()
    expected_area, _ = loop_area_and_volume(loops)
    assert abs(0.5 * area2.sum() - expected_area) < 1e-12 * abs(expected_area) * len(v)
    assert links_reciprocal(v, adj) == 0
    # every segment is a boundary edge, and the boundary is exactly the segments
    edges = {(int(v[t, i]), int(v[t, (i + 1) % 3])) for t in range(len(v)) for i in range(3) if adj[t, i] < 0}
    assert len(edges) == len(m.segments)
    for a, b in m.segments:
        assert (a, b) in edges or (b, a) in edges
    # Euler characteristic of a region with `holes` holes
    used = np.unique(v)
    n_edges = (3 * len(v) + len(edges)) // 2
    assert len(used) - n_edges + len(v) == 1 - holes
    # refinement points lie inside and the input points keep their coordinates
    n_in = len(pts)
    assert np.allclose(np.sort(P[:n_in], axis=0), np.sort(pts, axis=0), rtol=0, atol=1e-15 * np.abs(pts).max())
    # interior edges are locally Delaunay (the apex across is not inside the circumcircle)
    for t in range(0, len(v), max(1, len(v) // 2000)):
        for i in range(3):
            u = adj[t, i]
            if u < 0:
                continue
            j = [k for k in range(3) if adj[u, k] == t][0]
            apex = P[v[u, (j + 2) % 3]]
            assert not inside_circumcircle(P[v[t, 0]], P[v[t, 1]], P[v[t, 2]], apex)


def test_generate_rejects_inputs_the_reference_leaves_undefined():
    sq = np.array([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)])
    loop = np.array([(0, 1), (1, 2), (2, 3), (3, 0)])
    with pytest.raises(axs.AxsError) as ex:
        axs.cdt_generate(np.zeros((0, 2)), np.zeros((0, 2)))
    assert ex.value.status == 1
    with pytest.raises(axs.AxsError, match="zero extent"):
        axs.cdt_generate(np.array([(0.0, 0.0), (1.0, 0.0), (2.0, 0.0)]), [(0, 1)])
    with pytest.raises(axs.AxsError) as ex:
        axs.cdt_generate(sq, [(0, 4)])
    assert ex.value.status == 5
    with pytest.raises(axs.AxsError, match="itself"):
        axs.cdt_generate(sq, [(1, 1)])
    bad = sq.copy()
    bad[2, 0] = np.nan
    with pytest.raises(axs.AxsError, match="non-finite"):
        axs.cdt_generate(bad, loop)
    # the point limit (not in the reference) stops a refinement that would pass it
    p, s, _ = axs.outline_boundary([axs.nozzle_outline()], 0.002)
    with pytest.raises(axs.AxsError, match="point limit"):
        axs.cdt_generate(p, s, axs.cdt_params(max_points=200))
    assert len(axs.cdt_generate(p, s, axs.cdt_params(max_points=278)).points) == 278


def test_front_end_to_finite_volume_mesh():
    """Outline -> mesher -> createUnstructuredMesh: the cells are the mesher's
    triangles in its order, their volumes add up to the volume of revolution
    of the outline (Pappus is exact per triangle), and each boundary face
    takes the group of the loop edge it lies on."""
    loops = [axs.nozzle_outline()]
    h = 5e-4
    mesh, m = axs.cdt_mesh(loops, h)
    a = mesh.arrays
    assert mesh.n_cells == len(m.tri_v)
    assert np.array_equal(mesh.tris, m.tri_v)
    _, vol = loop_area_and_volume(loops)
    assert abs(a["cell_volume"].sum() - vol) < 1e-12 * vol
    _, _, groups = axs.outline_boundary(loops, h)
    bnd = a["face_neighbor"] < 0
    assert (a["face_group"][bnd] >= 0).all() and (a["face_group"][~bnd] == -1).all()
    for g in range(4):
        assert (a["face_group"][bnd] == g).sum() == (groups == g).sum()
        length = sum(np.hypot(e[2] - e[0], e[3] - e[1]) for e in loops[0] if e[4] == g)
        faces = bnd & (a["face_group"] == g)
        # face area = 2 pi r_f |edge|; the axis faces (r = 0) have none
        if g != 3:
            assert abs((a["face_area"][faces] / (2 * np.pi * a["face_center_r"][faces])).sum() - length) \
                < 1e-12 * length


@pytest.mark.skipif(not (HARNESS_REF.exists() and HARNESS_AXS.exists()),
                    reason="mesher harness not built (make -C oracle harness)")
def test_reference_harness_reports_the_same_on_the_drop_in():
    """The harness prints every case's triangles and checks; the reference's
    mesher also prints a debugging dump of each recovered segment
    (aximesh.cpp:729-737), which the drop-in does not, so those lines are
    dropped from its report before the comparison."""
    ref = subprocess.run([str(HARNESS_REF)], capture_output=True, text=True, timeout=120)
    dropin = subprocess.run([str(HARNESS_AXS)], capture_output=True, text=True, timeout=120)
    scaffold = re.compile(r"^    constraint \d+-\d+: |^      [ 0-9]{2,} \(.*\) -- ")
    ref_lines = [line for line in ref.stdout.splitlines() if not scaffold.match(line)]
    assert dropin.stdout.splitlines() == ref_lines
    assert dropin.returncode == ref.returncode
    # the insertion suite passes in full on both
    assert "5/5 cases passed" in dropin.stdout
