// compat/aximesh/aximesh.h -- C++ drop-in for the mesher's public interface
// (aximesh/include/aximesh/aximesh.h:7-100) on top of the libaxs.so C ABI.
//
// Mesh::runAxiMeshTriangulation (solver/source/mesh.cpp:1208-1251) and the
// mesher's harness (aximesh/tools/aximesh_cli.cpp) use Point, Segment,
// Triangle, Mesh, Params, generateMesh and insertPoints; with this directory
// on the include path in place of aximesh/include they compile unchanged and
// get the same meshes.  Failures throw std::runtime_error with the reference's
// text, as the reference does.  The mesher's internal types (front edges,
// normalisation, segment and front states) are not part of the interface and
// are not declared here.
#pragma once

#include <vector>

namespace AxiMesh {

struct Point {
    double x = 0.0;
    double y = 0.0;
};

// counter-clockwise; adj[i] is the triangle across edge v[i] -> v[i+1] (-1: none)
struct Triangle {
    int v[3];
    int adj[3];
};

// an edge that must appear in the mesh, between two existing points
struct Segment {
    int a;
    int b;
};

struct Params {
    double B = 1.4;
    double beta = 0.2;
    double classify_tol = 1.4;
};

struct Mesh {
    std::vector<Point> points;       // bin order, then the refinement points
    std::vector<Triangle> triangles;
    std::vector<Segment> segments;   // renumbered with the points
};

// size must be points.size() (how the GUI calls it)
Mesh generateMesh(const std::vector<Point>& points, const std::vector<Segment>& segments, int size,
                  const Params& params = {});

// appends the three super-triangle vertices to points
std::vector<Triangle> insertPoints(std::vector<Point>& points, int size);

}  // namespace AxiMesh
