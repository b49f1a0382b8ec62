// cdt.hpp -- the unstructured mesher: constrained Delaunay triangulation of a
// boundary point set, flood-fill removal of the outside, and frontal Delaunay
// refinement (Rebay) to a size field taken from the boundary segments.
//
// Same input, same output as AxiMesh::generateMesh / AxiMesh::insertPoints
// (aximesh/include/aximesh/aximesh.h:87-100, aximesh/source/aximesh.cpp):
// the point order after the bin sort, every refinement point to the bit, the
// triangle list with its vertex rotation and neighbour links, and the
// remapped segments.  The triangle order becomes the cell order of the
// finite-volume mesh built from it, so it is part of the result.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace axs::cdt {

struct Pt {
    double x = 0.0, y = 0.0;
};

// v counter-clockwise; adj[i] is the triangle across edge v[i] -> v[i+1]
// (-1: none)
struct Tri {
    int v[3];
    int adj[3];
};

struct Seg {
    int a, b;
};

// AxiMesh::Params, aximesh.h:56-61
struct Params {
    double skinny = 1.4;      // B (the unused Ruppert pass's quality bound)
    double grading = 0.2;     // beta, the size field's Lipschitz bound
    double accept = 1.4;      // classify_tol, accepted when sqrt(3) R <= accept * h
    // not in the reference: refinement stops with an error at this many points
    // instead of exhausting memory on a size field that never settles
    long long maxPoints = 1LL << 25;
};

struct Result {
    std::vector<Pt> points;   // input points in bin order, then refinement points
    std::vector<Tri> tris;
    std::vector<Seg> segs;    // the input segments in the new point numbering
};

// Thrown where the reference throws std::runtime_error (a point search that
// goes round more triangles than exist, a size field that keeps changing),
// and for inputs on which the reference's behaviour is undefined.
struct MeshError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// generateMesh: the whole pipeline on n points and the segments joining them
// (caller indices).
Result generate(const std::vector<Pt>& points, const std::vector<Seg>& segs, const Params& prm);

// insertPoints: the super triangle, then each point in the given order; the
// three super vertices are appended to `points`.  Returns 2n + 1 triangles
// (fewer when a point falls outside the super triangle).
std::vector<Tri> insertAll(std::vector<Pt>& points);

}  // namespace axs::cdt
