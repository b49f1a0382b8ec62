// compat/aximesh.cpp -- AxiMesh::generateMesh / insertPoints on the C ABI
// (axs_cdt_generate, axs_cdt_insert_points).
#include "aximesh/aximesh.h"

#include <stdexcept>
#include <string>

#include "axs.h"

namespace AxiMesh {
namespace {

[[noreturn]] void fail() {
    const char* why = axs_last_error(nullptr);
    throw std::runtime_error(why && *why ? why : "mesher failed");
}

Mesh take(axs_cdt* h) {
    int nP = 0, nT = 0, nS = 0;
    axs_cdt_counts(h, &nP, &nT, &nS);
    std::vector<double> x(nP), y(nP);
    std::vector<int> v(3 * (size_t)nT), adj(3 * (size_t)nT), seg(2 * (size_t)nS);
    axs_cdt_get(h, x.data(), y.data(), v.data(), adj.data(), seg.data());
    axs_cdt_destroy(h);
    Mesh m;
    m.points.resize(nP);
    for (int i = 0; i < nP; i++) m.points[i] = {x[i], y[i]};
    m.triangles.resize(nT);
    for (int t = 0; t < nT; t++)
        for (int k = 0; k < 3; k++) {
            m.triangles[t].v[k] = v[3 * t + k];
            m.triangles[t].adj[k] = adj[3 * t + k];
        }
    m.segments.resize(nS);
    for (int k = 0; k < nS; k++) m.segments[k] = {seg[2 * k], seg[2 * k + 1]};
    return m;
}

}  // namespace

Mesh generateMesh(const std::vector<Point>& points, const std::vector<Segment>& segments, int size,
                  const Params& params) {
    if (size != (int)points.size()) throw std::invalid_argument("generateMesh: size must be points.size()");
    std::vector<double> x(points.size()), y(points.size());
    for (size_t i = 0; i < points.size(); i++) {
        x[i] = points[i].x;
        y[i] = points[i].y;
    }
    std::vector<int> seg(2 * segments.size());
    for (size_t k = 0; k < segments.size(); k++) {
        seg[2 * k] = segments[k].a;
        seg[2 * k + 1] = segments[k].b;
    }
    axs_cdt_params p;
    axs_cdt_params_defaults(&p);
    p.B = params.B;
    p.beta = params.beta;
    p.classify_tol = params.classify_tol;
    axs_cdt* h = nullptr;
    if (axs_cdt_generate(&h, size, x.data(), y.data(), (int)segments.size(), seg.data(), &p) != AXS_OK) fail();
    return take(h);
}

std::vector<Triangle> insertPoints(std::vector<Point>& points, int size) {
    if (size != (int)points.size()) throw std::invalid_argument("insertPoints: size must be points.size()");
    std::vector<double> x(points.size()), y(points.size());
    for (size_t i = 0; i < points.size(); i++) {
        x[i] = points[i].x;
        y[i] = points[i].y;
    }
    axs_cdt* h = nullptr;
    if (axs_cdt_insert_points(&h, size, x.data(), y.data()) != AXS_OK) fail();
    Mesh m = take(h);
    points = std::move(m.points);
    return std::move(m.triangles);
}

}  // namespace AxiMesh
