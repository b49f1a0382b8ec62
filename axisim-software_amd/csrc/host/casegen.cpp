// casegen.cpp -- input producers: structured multiblock grids, triangle meshes
// and a seeded Poisson-disk + Delaunay nozzle mesher.
//
// The solver consumes one face-based mesh (HostMesh).  The two FVMesh producers
// restate the reference's own:
//   structured: generateFaces / makeRectBlock / autoDetectInterfaces /
//               toPackedMesh            structs/multiblock.h:59-136, 220-444
//   triangles:  createUnstructuredMesh  solver/source/mesh.cpp:1848-2002
//               + groupAtPoint          manager/source/math_func.cpp:152-195
// Boundary groups on the structured path come straight from the per-edge tags
// of each block (the reference re-derives the same ids geometrically in
// classifyBoundaryFaces, mesh.cpp:1835-1846).
//
// The nozzle mesher has no reference counterpart (the reference meshes with its
// own CDT or Gmsh); it writes the 1M-cell unstructured fixture.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <set>
#include <unordered_map>
#include <vector>

#include "mesh_types.hpp"

namespace axs {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Node {
    double z, r;
};

// generateFaces, multiblock.h:59-80
std::vector<double> zoneFaces(double start, double length, int cells, int grading, double coef) {
    std::vector<double> d(cells, length / cells);
    if (grading == 1 && cells > 1 && coef > 0.0 && coef != 1.0) {
        const double d0 = length * (coef - 1.0) / (std::pow(coef, cells) - 1.0);
        for (int k = 0; k < cells; k++) d[k] = d0 * std::pow(coef, k);
    }
    std::vector<double> f;
    f.reserve(cells + 1);
    double x = start;
    f.push_back(x);
    for (int k = 0; k < cells; k++) {
        x += d[k];
        f.push_back(x);
    }
    return f;
}

enum Edge { West = 0, East = 1, South = 2, North = 3 };

struct Block {
    int nr = 0, nz = 0, offset = 0;
    std::vector<Node> nodes;  // (nr+1)*(nz+1), node (I,J) at I*(nz+1)+J
    int group[4] = {-1, -1, -1, -1};
    const Node& node(int I, int J) const { return nodes[(size_t)I * (nz + 1) + J]; }
    int cell(int i, int j) const { return offset + i * nz + j; }
};

int edgeCount(const Block& b, int e) { return (e == West || e == East) ? b.nr : b.nz; }

void edgeCorners(const Block& b, int e, Node& s, Node& t) {
    switch (e) {
    case West: s = b.node(0, 0); t = b.node(b.nr, 0); break;
    case East: s = b.node(0, b.nz); t = b.node(b.nr, b.nz); break;
    case South: s = b.node(0, 0); t = b.node(0, b.nz); break;
    default: s = b.node(b.nr, 0); t = b.node(b.nr, b.nz); break;
    }
}

// local cell + node pair of position k along edge e (multiblock.h:298-306)
void edgeFace(const Block& b, int e, int k, int& c, Node& n0, Node& n1) {
    switch (e) {
    case West: c = b.cell(k, 0); n0 = b.node(k, 0); n1 = b.node(k + 1, 0); break;
    case East: c = b.cell(k, b.nz - 1); n0 = b.node(k, b.nz); n1 = b.node(k + 1, b.nz); break;
    case South: c = b.cell(0, k); n0 = b.node(0, k); n1 = b.node(0, k + 1); break;
    default: c = b.cell(b.nr - 1, k); n0 = b.node(b.nr, k); n1 = b.node(b.nr, k + 1); break;
    }
}

// quadCentroidArea, multiblock.h:272-295
void quadCentroid(const Node p[4], double& cz, double& cr, double& area) {
    double A = 0.0, Cz = 0.0, Cr = 0.0;
    for (int k = 0; k < 4; k++) {
        const Node& a = p[k];
        const Node& b = p[(k + 1) & 3];
        const double cross = a.z * b.r - b.z * a.r;
        A += cross;
        Cz += (a.z + b.z) * cross;
        Cr += (a.r + b.r) * cross;
    }
    A *= 0.5;
    if (std::fabs(A) < 1e-30) {
        cz = 0.25 * (p[0].z + p[1].z + p[2].z + p[3].z);
        cr = 0.25 * (p[0].r + p[1].r + p[2].r + p[3].r);
        area = 0.0;
        return;
    }
    cz = Cz / (6.0 * A);
    cr = Cr / (6.0 * A);
    area = std::fabs(A);
}

}  // namespace

std::string buildMultiblock(HostMesh& out, int nBlocks, const axs_block_spec* spec) {
    if (nBlocks <= 0 || !spec) return "no blocks";
    std::vector<Block> blocks(nBlocks);
    int total = 0;
    for (int bi = 0; bi < nBlocks; bi++) {
        const axs_block_spec& s = spec[bi];
        if (s.axial_cells <= 0 || s.radial_cells <= 0) return "block with no cells";
        const std::vector<double> zf = zoneFaces(s.z0, s.axial_length, s.axial_cells, s.axial_grading, s.axial_coef);
        const std::vector<double> rf = zoneFaces(s.r0, s.radial_length, s.radial_cells, s.radial_grading, s.radial_coef);
        Block& b = blocks[bi];
        b.nz = (int)zf.size() - 1;
        b.nr = (int)rf.size() - 1;
        b.nodes.resize((size_t)(b.nr + 1) * (b.nz + 1));
        for (int I = 0; I <= b.nr; I++)
            for (int J = 0; J <= b.nz; J++) b.nodes[(size_t)I * (b.nz + 1) + J] = Node{zf[J], rf[I]};
        for (int e = 0; e < 4; e++) b.group[e] = s.edge_group[e];
        b.offset = total;  // assignGlobalNumbering, multiblock.h:210-214
        total += b.nr * b.nz;
    }
    // autoDetectInterfaces, multiblock.h:220-251
    struct Itf { int a, b, ea, eb; bool rev; };
    std::vector<Itf> itfs;
    const double tol2 = 1e-9 * 1e-9;
    auto coincide = [&](const Node& p, const Node& q) {
        const double dz = p.z - q.z, dr = p.r - q.r;
        return dz * dz + dr * dr <= tol2;
    };
    for (int a = 0; a < nBlocks; a++)
        for (int b = a + 1; b < nBlocks; b++)
            for (int ea = 0; ea < 4; ea++)
                for (int eb = 0; eb < 4; eb++) {
                    Node as, ae, bs, be;
                    edgeCorners(blocks[a], ea, as, ae);
                    edgeCorners(blocks[b], eb, bs, be);
                    const bool aligned = coincide(as, bs) && coincide(ae, be);
                    const bool reversed = coincide(as, be) && coincide(ae, bs);
                    if (!aligned && !reversed) continue;
                    if (edgeCount(blocks[a], ea) != edgeCount(blocks[b], eb))
                        return "non-conformal seam between blocks";
                    itfs.push_back(Itf{a, b, ea, eb, reversed});
                }

    // toPackedMesh, multiblock.h:308-444
    HostMesh& m = out;
    m = HostMesh{};
    m.nCells = total;
    m.cellCenterZ.assign(total, 0.0);
    m.cellCenterR.assign(total, 0.0);
    m.cellArea2D.assign(total, 0.0);
    m.cellVolume.assign(total, 0.0);
    m.quadZ.assign((size_t)total * 4, 0.0);
    m.quadR.assign((size_t)total * 4, 0.0);
    for (const Block& b : blocks)
        for (int i = 0; i < b.nr; i++)
            for (int j = 0; j < b.nz; j++) {
                const int g = b.cell(i, j);
                const Node q[4] = {b.node(i, j), b.node(i, j + 1), b.node(i + 1, j + 1), b.node(i + 1, j)};
                for (int k = 0; k < 4; k++) {
                    m.quadZ[(size_t)g * 4 + k] = q[k].z;
                    m.quadR[(size_t)g * 4 + k] = q[k].r;
                }
                double cz, cr, area;
                quadCentroid(q, cz, cr, area);
                m.cellCenterZ[g] = cz;
                m.cellCenterR[g] = cr;
                m.cellArea2D[g] = area;
                m.cellVolume[g] = 2.0 * kPi * std::fabs(cr) * area;
            }
    std::vector<std::vector<int>> cellFaces(total);
    auto emit = [&](int owner, int nb, const Node& a, const Node& b, int group) {
        const int fid = (int)m.faceOwner.size();
        const double dz = b.z - a.z, dr = b.r - a.r;
        const double L = std::sqrt(dz * dz + dr * dr);
        const double fz = 0.5 * (a.z + b.z), fr = 0.5 * (a.r + b.r);
        double nz = (L > 0.0) ? dr / L : 0.0;
        double nr = (L > 0.0) ? -dz / L : 0.0;
        double tz, tr;
        if (nb >= 0) {
            tz = m.cellCenterZ[nb] - m.cellCenterZ[owner];
            tr = m.cellCenterR[nb] - m.cellCenterR[owner];
        } else {
            tz = fz - m.cellCenterZ[owner];
            tr = fr - m.cellCenterR[owner];
        }
        if (nz * tz + nr * tr < 0.0) {
            nz = -nz;
            nr = -nr;
        }
        m.faceOwner.push_back(owner);
        m.faceNeighbor.push_back(nb);
        m.faceNormalZ.push_back(nz);
        m.faceNormalR.push_back(nr);
        m.faceCenterZ.push_back(fz);
        m.faceCenterR.push_back(fr);
        m.faceArea.push_back(2.0 * kPi * std::fabs(fr) * L);
        m.faceGroup.push_back(group);
        cellFaces[owner].push_back(fid);
        if (nb >= 0) cellFaces[nb].push_back(fid);
    };
    for (const Block& b : blocks)
        for (int i = 0; i < b.nr; i++)
            for (int j = 0; j < b.nz; j++) {
                const int ow = b.cell(i, j);
                if (j + 1 < b.nz) emit(ow, b.cell(i, j + 1), b.node(i, j + 1), b.node(i + 1, j + 1), -1);
                if (i + 1 < b.nr) emit(ow, b.cell(i + 1, j), b.node(i + 1, j), b.node(i + 1, j + 1), -1);
            }
    std::set<long long> claimed;
    for (const Itf& it : itfs) {
        claimed.insert((long long)it.a * 4 + it.ea);
        claimed.insert((long long)it.b * 4 + it.eb);
        const Block& A = blocks[it.a];
        const Block& B = blocks[it.b];
        const int count = edgeCount(A, it.ea);
        for (int k = 0; k < count; k++) {
            int ca;
            Node n0, n1;
            edgeFace(A, it.ea, k, ca, n0, n1);
            const int kb = it.rev ? (count - 1 - k) : k;
            int cb;
            Node u0, u1;
            edgeFace(B, it.eb, kb, cb, u0, u1);
            emit(ca, cb, n0, n1, -1);
        }
    }
    for (int bi = 0; bi < nBlocks; bi++) {
        const Block& b = blocks[bi];
        for (int e = 0; e < 4; e++) {
            if (claimed.count((long long)bi * 4 + e)) continue;
            const int count = edgeCount(b, e);
            for (int k = 0; k < count; k++) {
                int c;
                Node n0, n1;
                edgeFace(b, e, k, c, n0, n1);
                emit(c, -1, n0, n1, b.group[e]);
            }
        }
    }
    m.cellFaceStart.assign(total + 1, 0);
    for (int c = 0; c < total; c++) m.cellFaceStart[c + 1] = m.cellFaceStart[c] + (int)cellFaces[c].size();
    m.cellFaceIDs.reserve(m.cellFaceStart[total]);
    for (int c = 0; c < total; c++)
        for (int f : cellFaces[c]) m.cellFaceIDs.push_back(f);
    m.nFaces = (int)m.faceOwner.size();
    return "";
}

// createUnstructuredMesh, mesh.cpp:1848-2002, with groupAtPoint classification
std::string buildTriangles(HostMesh& out, const std::vector<double>& pz, const std::vector<double>& pr,
                           const std::vector<int>& tris, int nSeg, const double* sz0, const double* sr0,
                           const double* sz1, const double* sr1, const int* sgroup, double tol) {
    const int nP = (int)pz.size();
    const int nT = (int)tris.size() / 3;
    if (nT <= 0) return "no triangles";
    for (int i = 0; i < 3 * nT; i++)
        if (tris[i] < 0 || tris[i] >= nP) return "triangle vertex out of range";
    HostMesh& m = out;
    m = HostMesh{};
    m.nCells = nT;
    m.pointZ = pz;
    m.pointR = pr;
    m.tris = tris;
    m.cellCenterZ.resize(nT);
    m.cellCenterR.resize(nT);
    m.cellArea2D.resize(nT);
    m.cellVolume.resize(nT);
    for (int c = 0; c < nT; c++) {
        const int a = tris[3 * c], b = tris[3 * c + 1], d = tris[3 * c + 2];
        // triangleCentroid / triangleArea2D, mesh.cpp:1549-1564
        m.cellCenterZ[c] = (pz[a] + pz[b] + pz[d]) / 3.0;
        m.cellCenterR[c] = (pr[a] + pr[b] + pr[d]) / 3.0;
        const double dz1 = pz[b] - pz[a], dr1 = pr[b] - pr[a];
        const double dz2 = pz[d] - pz[a], dr2 = pr[d] - pr[a];
        m.cellArea2D[c] = 0.5 * std::abs(dz1 * dr2 - dr1 * dz2);
        m.cellVolume[c] = 2.0 * kPi * m.cellCenterR[c] * m.cellArea2D[c];
    }
    std::unordered_map<uint64_t, int> edgeToFace;
    edgeToFace.reserve((size_t)nT * 2);
    std::vector<std::vector<int>> cellFaces(nT);
    auto addEdge = [&](int c, int v0, int v1) {
        const uint64_t key = ((uint64_t)(uint32_t)std::min(v0, v1) << 32) | (uint32_t)std::max(v0, v1);
        auto it = edgeToFace.find(key);
        if (it == edgeToFace.end()) {
            const double fz = 0.5 * (pz[v0] + pz[v1]);
            const double fr = 0.5 * (pr[v0] + pr[v1]);
            const double dz = pz[v1] - pz[v0], dr = pr[v1] - pr[v0];
            const double len = std::sqrt(dz * dz + dr * dr);
            double nz = dr, nr = -dz;
            const double mag = std::sqrt(nz * nz + nr * nr);
            if (mag > 1e-30) {
                nz /= mag;
                nr /= mag;
            }
            const double tz = fz - m.cellCenterZ[c], tr = fr - m.cellCenterR[c];
            if (nz * tz + nr * tr < 0.0) {
                nz *= -1.0;
                nr *= -1.0;
            }
            const int fid = (int)m.faceOwner.size();
            m.faceOwner.push_back(c);
            m.faceNeighbor.push_back(-1);
            m.faceGroup.push_back(-1);
            m.faceCenterZ.push_back(fz);
            m.faceCenterR.push_back(fr);
            m.faceArea.push_back(2.0 * kPi * fr * len);
            m.faceNormalZ.push_back(nz);
            m.faceNormalR.push_back(nr);
            edgeToFace[key] = fid;
            cellFaces[c].push_back(fid);
        } else {
            const int fid = it->second;
            m.faceNeighbor[fid] = c;
            m.faceGroup[fid] = -1;
            cellFaces[c].push_back(fid);
        }
    };
    for (int c = 0; c < nT; c++) {
        const int a = tris[3 * c], b = tris[3 * c + 1], d = tris[3 * c + 2];
        addEdge(c, a, b);
        addEdge(c, b, d);
        addEdge(c, d, a);
    }
    m.nFaces = (int)m.faceOwner.size();
    // classifyBoundaryFaces, mesh.cpp:1835-1846
    for (int f = 0; f < m.nFaces; f++) {
        if (m.faceNeighbor[f] >= 0) continue;
        int best = -1;
        double bestDist = tol;
        for (int s = 0; s < nSeg; s++) {
            if (sgroup[s] < 0) continue;
            const double abZ = sz1[s] - sz0[s], abR = sr1[s] - sr0[s];
            const double apZ = m.faceCenterZ[f] - sz0[s], apR = m.faceCenterR[f] - sr0[s];
            const double ab2 = abZ * abZ + abR * abR;
            double qz = sz0[s], qr = sr0[s];
            if (ab2 > 1e-30) {
                double t = (apZ * abZ + apR * abR) / ab2;
                t = std::clamp(t, 0.0, 1.0);
                qz = sz0[s] + t * abZ;
                qr = sr0[s] + t * abR;
            }
            const double dz = qz - m.faceCenterZ[f], dr = qr - m.faceCenterR[f];
            const double dist = std::sqrt(dz * dz + dr * dr);
            if (dist < bestDist) {
                bestDist = dist;
                best = sgroup[s];
            }
        }
        m.faceGroup[f] = best;
    }
    m.cellFaceStart.assign(nT + 1, 0);
    for (int c = 0; c < nT; c++) m.cellFaceStart[c + 1] = m.cellFaceStart[c] + (int)cellFaces[c].size();
    m.cellFaceIDs.reserve(m.cellFaceStart[nT]);
    for (int c = 0; c < nT; c++)
        for (int f : cellFaces[c]) m.cellFaceIDs.push_back(f);
    return "";
}

// ---------------------------------------------------------------------------
// Nozzle mesher: Poisson-disk (Bridson) points, jitter, Bowyer-Watson Delaunay.
// ---------------------------------------------------------------------------
namespace {

struct Rng {  // splitmix64; identical sequence on every platform
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Profile {  // wall radius R(z) as a polyline from z = 0 to z = L
    std::vector<double> z, r;
    double L() const { return z.back(); }
    double R(double x) const {
        if (x <= z[0]) return r[0];
        for (size_t i = 1; i < z.size(); i++)
            if (x <= z[i]) {
                const double t = (x - z[i - 1]) / (z[i] - z[i - 1]);
                return r[i - 1] + t * (r[i] - r[i - 1]);
            }
        return r.back();
    }
};

struct Seg {
    double z0, r0, z1, r1;
    int group;
};

double segDist(double pz, double pr, const Seg& s) {
    const double abZ = s.z1 - s.z0, abR = s.r1 - s.r0;
    const double ab2 = abZ * abZ + abR * abR;
    double t = ab2 > 0 ? ((pz - s.z0) * abZ + (pr - s.r0) * abR) / ab2 : 0.0;
    t = std::clamp(t, 0.0, 1.0);
    const double dz = s.z0 + t * abZ - pz, dr = s.r0 + t * abR - pr;
    return std::sqrt(dz * dz + dr * dr);
}

uint64_t hilbert(uint32_t x, uint32_t y, int order) {  // xy -> Hilbert index
    uint64_t d = 0;
    for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1 : 0;
        const uint32_t ry = (y & s) ? 1 : 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

struct Delaunay {
    std::vector<double> X, Y;
    struct Tri {
        int v[3];
        int n[3];  // n[i] = triangle across the edge opposite v[i]
    };
    std::vector<Tri> T;
    std::vector<int> stamp;
    int stampId = 0, last = 0;

    long double orient(int a, int b, int c) const {
        const long double ax = X[a], ay = Y[a];
        return ((long double)X[b] - ax) * ((long double)Y[c] - ay) -
               ((long double)Y[b] - ay) * ((long double)X[c] - ax);
    }
    long double orientP(int a, int b, double px, double py) const {
        const long double ax = X[a], ay = Y[a];
        return ((long double)X[b] - ax) * ((long double)py - ay) - ((long double)Y[b] - ay) * ((long double)px - ax);
    }
    bool inCircle(const Tri& t, double px, double py) const {
        const long double adx = (long double)X[t.v[0]] - px, ady = (long double)Y[t.v[0]] - py;
        const long double bdx = (long double)X[t.v[1]] - px, bdy = (long double)Y[t.v[1]] - py;
        const long double cdx = (long double)X[t.v[2]] - px, cdy = (long double)Y[t.v[2]] - py;
        const long double det = (adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) +
                                (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy) +
                                (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady);
        return det > 0;
    }

    int locate(double px, double py) {
        int t = last;
        int rot = 0;
        for (long guard = 0; guard < 100000000L; guard++) {
            const Tri& tr = T[t];
            bool moved = false;
            for (int e = 0; e < 3; e++) {
                const int i = (rot + e) % 3;
                if (orientP(tr.v[(i + 1) % 3], tr.v[(i + 2) % 3], px, py) < 0) {
                    if (tr.n[i] < 0) break;
                    t = tr.n[i];
                    moved = true;
                    break;
                }
            }
            rot = (rot + 1) % 3;
            if (!moved) return t;
        }
        return -1;
    }

    bool insert(int p) {
        const double px = X[p], py = Y[p];
        const int t0 = locate(px, py);
        if (t0 < 0) return false;
        stampId++;
        std::vector<int> cavity{t0};
        stamp[t0] = stampId;
        for (size_t i = 0; i < cavity.size(); i++) {
            const Tri& c = T[cavity[i]];
            for (int e = 0; e < 3; e++) {
                const int nb = c.n[e];
                if (nb < 0 || stamp[nb] == stampId) continue;
                if (inCircle(T[nb], px, py)) {
                    stamp[nb] = stampId;
                    cavity.push_back(nb);
                }
            }
        }
        struct BEdge { int a, b, out; };
        std::vector<BEdge> edges;
        for (int ci : cavity) {
            const Tri& c = T[ci];
            for (int e = 0; e < 3; e++) {
                const int nb = c.n[e];
                if (nb >= 0 && stamp[nb] == stampId) continue;
                edges.push_back(BEdge{c.v[(e + 1) % 3], c.v[(e + 2) % 3], nb});
            }
        }
        // A star-shaped cavity of c triangles has c + 2 boundary edges: reuse
        // every cavity slot, append the rest.
        std::vector<int> idx(edges.size());
        for (size_t i = 0; i < edges.size(); i++) {
            if (i < cavity.size()) {
                idx[i] = cavity[i];
            } else {
                idx[i] = (int)T.size();
                T.push_back(Tri{});
                stamp.push_back(0);
            }
        }
        for (size_t i = edges.size(); i < cavity.size(); i++) T[cavity[i]].v[0] = -1;  // dead
        auto startAt = [&](int v) {
            for (size_t i = 0; i < edges.size(); i++)
                if (edges[i].a == v) return idx[i];
            return -1;
        };
        auto endAt = [&](int v) {
            for (size_t i = 0; i < edges.size(); i++)
                if (edges[i].b == v) return idx[i];
            return -1;
        };
        for (size_t i = 0; i < edges.size(); i++) {
            Tri& t = T[idx[i]];
            const BEdge& e = edges[i];
            t.v[0] = e.a;
            t.v[1] = e.b;
            t.v[2] = p;
            t.n[2] = e.out;
            t.n[0] = startAt(e.b);  // across edge (b, p)
            t.n[1] = endAt(e.a);    // across edge (p, a)
            if (e.out >= 0) {
                Tri& o = T[e.out];
                for (int k = 0; k < 3; k++) {
                    const int x = o.v[(k + 1) % 3], y = o.v[(k + 2) % 3];
                    if ((x == e.a && y == e.b) || (x == e.b && y == e.a)) o.n[k] = idx[i];
                }
            }
        }
        last = idx[0];
        return true;
    }
};

}  // namespace

std::string buildNozzle(HostMesh& out, const axs_nozzle_spec& sp) {
    const double h = sp.spacing;
    if (!(h > 0.0) || !(sp.r_in > 0) || !(sp.r_throat > 0) || !(sp.r_out > 0)) return "bad nozzle spec";
    const double tc = std::tan(sp.cone_deg * kPi / 180.0), td = std::tan(sp.diff_deg * kPi / 180.0);
    if (!(tc > 0) || !(td > 0)) return "bad nozzle angles";
    Profile P;
    double z = 0.0;
    P.z.push_back(z); P.r.push_back(sp.r_in);
    z += sp.l_in;       P.z.push_back(z); P.r.push_back(sp.r_in);
    z += std::fabs(sp.r_in - sp.r_throat) / tc; P.z.push_back(z); P.r.push_back(sp.r_throat);
    z += sp.l_throat;   P.z.push_back(z); P.r.push_back(sp.r_throat);
    z += std::fabs(sp.r_out - sp.r_throat) / td; P.z.push_back(z); P.r.push_back(sp.r_out);
    z += sp.l_out;      P.z.push_back(z); P.r.push_back(sp.r_out);
    const double L = P.L();
    double Rmax = 0.0;
    for (double r : P.r) Rmax = std::max(Rmax, r);

    // boundary loop: axis (group 3), outlet (1), wall (2, reversed profile), inlet (0)
    std::vector<Seg> segs;
    segs.push_back(Seg{0.0, 0.0, L, 0.0, 3});
    segs.push_back(Seg{L, 0.0, L, P.r.back(), 1});
    for (size_t i = P.z.size() - 1; i > 0; i--) segs.push_back(Seg{P.z[i], P.r[i], P.z[i - 1], P.r[i - 1], 2});
    segs.push_back(Seg{0.0, P.r[0], 0.0, 0.0, 0});

    std::vector<double> X, Y;
    for (const Seg& s : segs) {  // boundary points, each loop vertex once
        const double len = std::hypot(s.z1 - s.z0, s.r1 - s.r0);
        if (len <= 0.0) continue;
        const int n = std::max(1, (int)std::ceil(len / h - 1e-9));
        for (int k = 0; k < n; k++) {
            const double t = (double)k / n;
            X.push_back(s.z0 + t * (s.z1 - s.z0));
            Y.push_back(s.r0 + t * (s.r1 - s.r0));
        }
    }
    const int nBoundary = (int)X.size();

    auto inside = [&](double pz, double pr, double margin) {
        if (pz < 0.0 || pz > L || pr < 0.0 || pr > P.R(pz)) return false;
        for (const Seg& s : segs)
            if (segDist(pz, pr, s) < margin) return false;
        return true;
    };

    // Bridson Poisson-disk sampling on a background grid of cell h / sqrt(2)
    const double cell = h / std::sqrt(2.0);
    const int gx = (int)(L / cell) + 3, gy = (int)(Rmax / cell) + 3;
    const int SLOTS = 4;
    std::vector<int> grid((size_t)gx * gy * SLOTS, -1);
    auto gridAdd = [&](int id) {
        const int ix = std::clamp((int)(X[id] / cell), 0, gx - 1), iy = std::clamp((int)(Y[id] / cell), 0, gy - 1);
        int* slot = &grid[((size_t)iy * gx + ix) * SLOTS];
        for (int k = 0; k < SLOTS; k++)
            if (slot[k] < 0) {
                slot[k] = id;
                return true;
            }
        return false;
    };
    auto farEnough = [&](double pz, double pr) {
        const int ix = (int)(pz / cell), iy = (int)(pr / cell);
        const double h2 = h * h;
        for (int dy = -2; dy <= 2; dy++)
            for (int dx = -2; dx <= 2; dx++) {
                const int x = ix + dx, y = iy + dy;
                if (x < 0 || y < 0 || x >= gx || y >= gy) continue;
                const int* slot = &grid[((size_t)y * gx + x) * SLOTS];
                for (int k = 0; k < SLOTS; k++) {
                    const int q = slot[k];
                    if (q < 0) continue;
                    const double ddx = X[q] - pz, ddy = Y[q] - pr;
                    if (ddx * ddx + ddy * ddy < h2) return false;
                }
            }
        return true;
    };
    for (int i = 0; i < nBoundary; i++)
        if (!gridAdd(i)) return "boundary points too dense for the sampling grid";
    Rng rng(sp.seed_points);
    std::vector<int> active;
    {
        const double sz = 0.5 * sp.l_in, sr = 0.5 * sp.r_in;
        if (!inside(sz, sr, 0.5 * h)) return "seed point outside the domain";
        X.push_back(sz);
        Y.push_back(sr);
        gridAdd((int)X.size() - 1);
        active.push_back((int)X.size() - 1);
    }
    const int kTries = 30;
    while (!active.empty()) {
        const size_t pick = (size_t)(rng.u01() * active.size()) % active.size();
        const int a = active[pick];
        bool found = false;
        for (int t = 0; t < kTries; t++) {
            const double ang = 2.0 * kPi * rng.u01();
            const double rad = h * (1.0 + rng.u01());
            const double cz = X[a] + rad * std::cos(ang), cr = Y[a] + rad * std::sin(ang);
            if (!inside(cz, cr, 0.5 * h)) continue;
            if (!farEnough(cz, cr)) continue;
            X.push_back(cz);
            Y.push_back(cr);
            if (!gridAdd((int)X.size() - 1)) return "sampling grid overflow";
            active.push_back((int)X.size() - 1);
            found = true;
            break;
        }
        if (!found) {
            active[pick] = active.back();
            active.pop_back();
        }
    }
    // jitter interior nodes by up to jitter*h (uniform in a disk)
    Rng jr(sp.seed_jitter);
    for (size_t i = nBoundary; i < X.size(); i++) {
        double dx, dy;
        do {
            dx = 2.0 * jr.u01() - 1.0;
            dy = 2.0 * jr.u01() - 1.0;
        } while (dx * dx + dy * dy > 1.0);
        const double nz = X[i] + sp.jitter * h * dx, nr = Y[i] + sp.jitter * h * dy;
        if (inside(nz, nr, 0.25 * h)) {
            X[i] = nz;
            Y[i] = nr;
        }
    }
    const int nPts = (int)X.size();

    // Delaunay: super triangle vertices 0..2, then points in Hilbert order
    Delaunay D;
    const double cx = 0.5 * L, cy = 0.5 * Rmax, span = 20.0 * std::max(L, Rmax);
    D.X = {cx - 2 * span, cx + 2 * span, cx};
    D.Y = {cy - span, cy - span, cy + 2 * span};
    std::vector<int> order(nPts);
    {
        std::vector<uint64_t> key(nPts);
        const double sc = 65535.0 / std::max(L, Rmax);
        for (int i = 0; i < nPts; i++)
            key[i] = hilbert((uint32_t)(X[i] * sc), (uint32_t)(Y[i] * sc), 16);
        for (int i = 0; i < nPts; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    }
    for (int i = 0; i < nPts; i++) {
        D.X.push_back(X[order[i]]);
        D.Y.push_back(Y[order[i]]);
    }
    D.T.push_back(Delaunay::Tri{{0, 1, 2}, {-1, -1, -1}});
    D.stamp.push_back(0);
    for (int i = 0; i < nPts; i++)
        if (!D.insert(3 + i)) return "Delaunay point location failed";

    // keep triangles of real vertices whose centroid is inside the domain
    struct OutTri { int a, b, c; uint64_t key; };
    std::vector<OutTri> kept;
    const double sc = 65535.0 / std::max(L, Rmax);
    for (const auto& t : D.T) {
        if (t.v[0] < 0) continue;
        if (t.v[0] < 3 || t.v[1] < 3 || t.v[2] < 3) continue;
        const double gz = (D.X[t.v[0]] + D.X[t.v[1]] + D.X[t.v[2]]) / 3.0;
        const double gr = (D.Y[t.v[0]] + D.Y[t.v[1]] + D.Y[t.v[2]]) / 3.0;
        if (!inside(gz, gr, 0.0)) continue;
        kept.push_back(OutTri{t.v[0] - 3, t.v[1] - 3, t.v[2] - 3,
                              hilbert((uint32_t)(gz * sc), (uint32_t)(gr * sc), 16)});
    }
    // cell order = Hilbert order of the centroids, so neighbouring cells sit
    // close in memory (coalesced stencil gathers)
    std::stable_sort(kept.begin(), kept.end(), [](const OutTri& a, const OutTri& b) { return a.key < b.key; });
    // renumber nodes by first use
    std::vector<int> remap(nPts, -1);
    std::vector<double> pz, pr;
    std::vector<int> tris;
    tris.reserve(kept.size() * 3);
    for (const OutTri& t : kept)
        for (int v : {t.a, t.b, t.c}) {
            if (remap[v] < 0) {
                remap[v] = (int)pz.size();
                pz.push_back(D.X[3 + v]);
                pr.push_back(D.Y[3 + v]);
            }
            tris.push_back(remap[v]);
        }
    std::vector<double> sz0, sr0, sz1, sr1;
    std::vector<int> sg;
    for (const Seg& s : segs) {
        sz0.push_back(s.z0);
        sr0.push_back(s.r0);
        sz1.push_back(s.z1);
        sr1.push_back(s.r1);
        sg.push_back(s.group);
    }
    const double tol = 1e-4 * std::max(std::max(L, Rmax), 1.0);  // classifyBoundaryFaces, mesh.cpp:1836
    return buildTriangles(out, pz, pr, tris, (int)segs.size(), sz0.data(), sr0.data(), sz1.data(), sr1.data(),
                          sg.data(), tol);
}

}  // namespace axs

namespace axs {

std::string concatMeshes(HostMesh& out, int count, const axs_mesh_view* meshes, int groupsPerMesh) {
    if (count <= 0 || !meshes) return "no meshes to concatenate";
    if (groupsPerMesh <= 0) return "groups per mesh must be positive";
    out = HostMesh{};
    long cells = 0, faces = 0, slots = 0;
    for (int c = 0; c < count; c++) {
        const axs_mesh_view& v = meshes[c];
        if (v.n_cells <= 0 || v.n_faces <= 0 || !v.cell_face_start) return "empty mesh in the batch";
        cells += v.n_cells;
        faces += v.n_faces;
        slots += v.cell_face_start[v.n_cells];
    }
    if (cells > 0x7fffffffL || faces > 0x7fffffffL || slots > 0x7fffffffL) return "batch too large for 32-bit indices";
    out.nCells = (int)cells;
    out.nFaces = (int)faces;
    out.cellFaceStart.reserve(cells + 1);
    out.cellFaceStart.push_back(0);
    int cellOff = 0, faceOff = 0;
    for (int c = 0; c < count; c++) {
        const axs_mesh_view& v = meshes[c];
        auto app = [](std::vector<double>& dst, const double* src, int n) { dst.insert(dst.end(), src, src + n); };
        app(out.cellCenterZ, v.cell_center_z, v.n_cells);
        app(out.cellCenterR, v.cell_center_r, v.n_cells);
        app(out.cellVolume, v.cell_volume, v.n_cells);
        app(out.cellArea2D, v.cell_area2d, v.n_cells);
        const int slot0 = out.cellFaceStart.back();
        for (int i = 0; i < v.n_cells; i++) out.cellFaceStart.push_back(slot0 + v.cell_face_start[i + 1]);
        for (int k = 0; k < v.cell_face_start[v.n_cells]; k++) out.cellFaceIDs.push_back(v.cell_face_ids[k] + faceOff);
        for (int f = 0; f < v.n_faces; f++) {
            out.faceOwner.push_back(v.face_owner[f] + cellOff);
            out.faceNeighbor.push_back(v.face_neighbor[f] >= 0 ? v.face_neighbor[f] + cellOff : -1);
            const int g = v.face_group[f];
            if (g >= groupsPerMesh) return "boundary group beyond groups per mesh";
            out.faceGroup.push_back(g >= 0 ? g + c * groupsPerMesh : g);
        }
        app(out.faceNormalZ, v.face_normal_z, v.n_faces);
        app(out.faceNormalR, v.face_normal_r, v.n_faces);
        app(out.faceCenterZ, v.face_center_z, v.n_faces);
        app(out.faceCenterR, v.face_center_r, v.n_faces);
        app(out.faceArea, v.face_area, v.n_faces);
        cellOff += v.n_cells;
        faceOff += v.n_faces;
    }
    return {};
}

}  // namespace axs
