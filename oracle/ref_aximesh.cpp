// ref_aximesh.cpp -- C entry points over the reference's own mesher
// (aximesh/source/aximesh.cpp with aximesh/include/aximesh/aximesh.h).
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// reference's aximesh.cpp where it lies (never copied); the library goes to
// oracle/_ref/.  Tests and the golden-vector script load it with ctypes to
// compare the product's mesher with the reference, case by case.
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "aximesh/aximesh.h"

namespace {
struct Out {
    std::vector<AxiMesh::Point> points;
    std::vector<AxiMesh::Triangle> tris;
    std::vector<AxiMesh::Segment> segs;
    std::string error;
};
}  // namespace

extern "C" {

// AxiMesh::generateMesh with the given Params; never returns NULL (a thrown
// error is kept as text, ref_aximesh_error).
void* ref_aximesh_generate(int n, const double* x, const double* y, int nSeg, const int* seg, double B, double beta,
                           double classifyTol) {
    Out* o = new Out();
    std::vector<AxiMesh::Point> pts(n);
    for (int i = 0; i < n; i++) pts[i] = {x[i], y[i]};
    std::vector<AxiMesh::Segment> segs(nSeg);
    for (int k = 0; k < nSeg; k++) segs[k] = {seg[2 * k], seg[2 * k + 1]};
    AxiMesh::Params prm;
    prm.B = B;
    prm.beta = beta;
    prm.classify_tol = classifyTol;
    try {
        AxiMesh::Mesh m = AxiMesh::generateMesh(pts, segs, n, prm);
        o->points = std::move(m.points);
        o->tris = std::move(m.triangles);
        o->segs = std::move(m.segments);
    } catch (const std::exception& e) {
        o->error = e.what();
    }
    return o;
}

// AxiMesh::insertPoints on the points as given.
void* ref_aximesh_insert(int n, const double* x, const double* y) {
    Out* o = new Out();
    o->points.resize(n);
    for (int i = 0; i < n; i++) o->points[i] = {x[i], y[i]};
    try {
        o->tris = AxiMesh::insertPoints(o->points, n);
    } catch (const std::exception& e) {
        o->error = e.what();
    }
    return o;
}

const char* ref_aximesh_error(void* h) { return static_cast<Out*>(h)->error.c_str(); }

void ref_aximesh_counts(void* h, int* nPoints, int* nTris, int* nSeg) {
    const Out* o = static_cast<Out*>(h);
    *nPoints = (int)o->points.size();
    *nTris = (int)o->tris.size();
    *nSeg = (int)o->segs.size();
}

void ref_aximesh_copy(void* h, double* x, double* y, int* triV, int* triAdj, int* seg) {
    const Out* o = static_cast<Out*>(h);
    for (size_t i = 0; i < o->points.size(); i++) {
        x[i] = o->points[i].x;
        y[i] = o->points[i].y;
    }
    for (size_t t = 0; t < o->tris.size(); t++)
        for (int k = 0; k < 3; k++) {
            triV[3 * t + k] = o->tris[t].v[k];
            triAdj[3 * t + k] = o->tris[t].adj[k];
        }
    for (size_t k = 0; k < o->segs.size(); k++) {
        seg[2 * k] = o->segs[k].a;
        seg[2 * k + 1] = o->segs[k].b;
    }
}

void ref_aximesh_free(void* h) { delete static_cast<Out*>(h); }
}
