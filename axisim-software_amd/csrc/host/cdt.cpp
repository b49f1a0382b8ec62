// cdt.cpp -- constrained Delaunay triangulation + frontal refinement.
//
// Pipeline (AxiMesh::generateMesh, aximesh/source/aximesh.cpp:1105-1175):
//   1. scale the points into the unit box by the larger extent
//   2. order them along a serpentine of n^(1/4) x n^(1/4) bins
//   3. Sloan's incremental Delaunay: super triangle, walk to the enclosing
//      triangle, split, Lawson swaps with Cline and Renka's test
//   4. recover each segment: list the edges it crosses, swap them out,
//      restore Delaunay on the new edges
//   5. flood fill from the super triangle, flipping inside/outside at every
//      segment edge, and keep the inside
//   6. size field: mean length of the segments at each point, limited to a
//      grading of beta per unit length (repeated sweeps over the edges)
//   7. Rebay's frontal refinement: from the front edge of the waiting
//      triangle with the largest circumradius, insert the point on its
//      Voronoi segment that makes a triangle of the local size
//   8. scale back
//
// Every floating-point expression keeps the reference's operand order, and
// every tie is broken as the reference breaks it (std::sort of the bin
// numbers, std::priority_queue of the front, LIFO stacks, first-match
// scans), so the output is the same, index for index and bit for bit.
#include "cdt.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <queue>

namespace axs::cdt {
namespace {

inline double dist2(const Pt& a, const Pt& b) {
    const double dx = a.x - b.x, dy = a.y - b.y;
    return dx * dx + dy * dy;
}

// twice the signed area of (a, b, p): positive when p is left of a -> b
inline double orient(const Pt& a, const Pt& b, const Pt& p) {
    return (b.x - a.x) * (p.y - a.y) - (b.y - a.y) * (p.x - a.x);
}

// (a - c) . (b - c)
inline double dotAt(const Pt& a, const Pt& b, const Pt& c) {
    const double ax = a.x - c.x, ay = a.y - c.y, bx = b.x - c.x, by = b.y - c.y;
    return ax * bx + ay * by;
}

// squared circumradius from the sides: |ab|^2 |bc|^2 |ca|^2 / (4 (2A)^2)
inline double circumR2(const Pt& a, const Pt& b, const Pt& c) {
    const double q = orient(a, b, c);
    return (dist2(a, b) * dist2(b, c) * dist2(c, a)) / (4.0 * q * q);
}

// circumcentre by Cramer's rule, relative to a
inline Pt circumcentre(const Pt& a, const Pt& b, const Pt& c) {
    const double bx = b.x - a.x, by = b.y - a.y, cx = c.x - a.x, cy = c.y - a.y;
    const double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
    const double d = 2.0 * orient(a, b, c);
    return {a.x + ((cy * b2 - by * c2) / d), a.y + ((bx * c2 - cx * b2) / d)};
}

// Cline and Renka's swap test for the quadrilateral p0 p2 p1 p (p0 -> p1 the
// shared edge, p2 and p the apexes on either side): true when p lies strictly
// inside the circumcircle of (p0, p1, p2), i.e. the angles at p2 and at p sum
// to more than pi; cocircular points do not swap.
inline bool swaps(const Pt& p0, const Pt& p1, const Pt& p2, const Pt& p) {
    const double x13 = p0.x - p2.x, x23 = p1.x - p2.x, x1p = p0.x - p.x, x2p = p1.x - p.x;
    const double y13 = p0.y - p2.y, y23 = p1.y - p2.y, y1p = p0.y - p.y, y2p = p1.y - p.y;
    const double cosA = x13 * x23 + y13 * y23;
    const double cosB = x2p * x1p + y2p * y1p;
    if (cosA >= 0 && cosB >= 0) return false;
    if (cosA < 0 && cosB < 0) return true;
    const double sinAB = (x13 * y23 - x23 * y13) * cosB + (x2p * y1p - x1p * y2p) * cosA;
    return sinAB < 0;
}

inline int slotOf(const int (&a)[3], int x) {
    for (int i = 0; i < 3; i++)
        if (a[i] == x) return i;
    return -1;
}

enum class Recovery { Exists, Crossing, Degenerate, Unresolved };

enum class Front : uint8_t { Accepted, Active, Waiting };

struct FrontEdge {
    int t, e;
    double len, crad;
    int v0, v1;
};

// the front is served largest circumradius first
struct SmallerRadius {
    bool operator()(const FrontEdge& a, const FrontEdge& b) const { return a.crad < b.crad; }
};

class Mesher {
public:
    std::vector<Pt>& P;
    std::vector<Tri> T;
    std::vector<int> touched;   // triangles split or swapped by the last insert
    std::vector<int> vtx;       // a triangle holding each vertex (segment recovery)

    explicit Mesher(std::vector<Pt>& pts) : P(pts) {}

    // -------- incremental Delaunay (insertVertex / insertPoints, aximesh.cpp:339-489)

    // Super triangle over the unit box, then points 0..n-1 in order, each walk
    // starting from the last triangle made.
    void triangulate(int n) {
        T.reserve(2 * (size_t)n + 1);
        T.push_back({{n, n + 1, n + 2}, {-1, -1, -1}});
        P.push_back({-100.0, -100.0});
        P.push_back({100.0, -100.0});
        P.push_back({0.0, 100.0});
        int start = 0;
        for (int p = 0; p < n; p++) {
            touched.clear();
            if (!insert(p, start)) continue;
            start = (int)T.size() - 1;
        }
    }

    // Walk from triangle t to point p (leaving each triangle across its first
    // edge that p is strictly right of), split the triangle holding it into
    // three with p as every piece's third vertex, and swap until Delaunay.
    // false when the walk leaves the triangulation.
    bool insert(int p, int t) {
        const Pt q = P[p];
        for (int steps = 0;; steps++) {
            if (t == -1) return false;
            if (steps > (int)T.size()) throw MeshError("point search has stalled");
            const Tri& c = T[t];
            int exit = -1;
            for (int e = 0; e < 3; e++)
                if (orient(P[c.v[e]], P[c.v[(e + 1) % 3]], q) < 0) {
                    exit = e;
                    break;
                }
            if (exit < 0) break;
            t = c.adj[exit];
        }
        const int a = T[t].v[0], b = T[t].v[1], c = T[t].v[2];
        const int n0 = T[t].adj[0], n1 = T[t].adj[1], n2 = T[t].adj[2];
        const int t1 = (int)T.size(), t2 = t1 + 1;
        T[t] = {{a, b, p}, {n0, t1, t2}};
        T.push_back({{b, c, p}, {n1, t2, t}});
        T.push_back({{c, a, p}, {n2, t, t1}});
        if (n1 != -1) relink(n1, t, t1);
        if (n2 != -1) relink(n2, t, t2);
        touched.push_back(t1);
        touched.push_back(t2);
        touched.push_back(t);

        // the triangles across the three edges facing p, last pushed first
        stack_.clear();
        for (int nb : {n0, n1, n2})
            if (nb != -1) stack_.push_back(nb);
        while (!stack_.empty()) {
            const int o = stack_.back();
            stack_.pop_back();
            const Tri& to = T[o];
            if (to.v[2] == p) continue;
            int e = -1;
            for (int i = 0; i < 3; i++) {
                const int nb = to.adj[i];
                if (nb != -1 && T[nb].v[2] == p) {
                    e = i;
                    break;
                }
            }
            if (e < 0) continue;
            if (!swaps(P[to.v[e]], P[to.v[(e + 1) % 3]], P[to.v[(e + 2) % 3]], q)) continue;
            const int f1 = to.adj[(e + 1) % 3], f2 = to.adj[(e + 2) % 3];
            flip(o, e);
            touched.push_back(o);
            if (f1 != -1) stack_.push_back(f1);
            if (f2 != -1) stack_.push_back(f2);
        }
        return true;
    }

    // Swap the diagonal that t shares with its neighbour u across edge e:
    // t = (v0 v1 v2), u = (v1 v0 w) become u = (v1 v2 w) and t = (v2 v0 w)
    // (flipEdge, aximesh.cpp:254-296).
    void flip(int t, int e) {
        const int u = T[t].adj[e];
        if (u < 0) throw MeshError("edge swap on a boundary edge");
        const int eu = slotOf(T[u].adj, t);
        if (eu < 0) throw MeshError("triangle links are not reciprocal");
        const Tri A = T[t], B = T[u];
        const int w = B.v[(eu + 2) % 3];
        const int v0 = A.v[e], v1 = A.v[(e + 1) % 3], v2 = A.v[(e + 2) % 3];
        const int a1 = A.adj[(e + 1) % 3], a2 = A.adj[(e + 2) % 3];
        const int b1 = B.adj[(eu + 1) % 3], b2 = B.adj[(eu + 2) % 3];
        T[u] = {{v1, v2, w}, {a1, t, b2}};
        T[t] = {{v2, v0, w}, {a2, b1, u}};
        if (b1 != -1) relink(b1, u, t);
        if (a1 != -1) relink(a1, t, u);
    }

    // -------- segment recovery (aximesh.cpp:491-746)

    // Where segment a-b leaves a: rotate around a from vtx[a] (checkSegment).
    Recovery classify(const Seg& s, int& tOut, int& eOut) const {
        const int a = s.a, b = s.b;
        const int t0 = vtx[a];
        if (t0 < 0) return Recovery::Unresolved;
        int t = t0;
        do {
            const int e = slotOf(T[t].v, a);
            if (e < 0) return Recovery::Unresolved;
            const int va = T[t].v[(e + 1) % 3], vb = T[t].v[(e + 2) % 3];
            if (va == b) return Recovery::Exists;
            const double oa = orient(P[a], P[va], P[b]);
            const double ob = orient(P[a], P[vb], P[b]);
            if (oa > 0 && ob < 0) {
                tOut = t;
                eOut = (e + 1) % 3;
                return Recovery::Crossing;
            }
            if (oa == 0 && dotAt(P[b], P[va], P[a]) > 0) {
                tOut = t;
                eOut = e;
                return Recovery::Degenerate;
            }
            t = T[t].adj[e];
            if (t < 0) return Recovery::Unresolved;
        } while (t != t0);
        return Recovery::Unresolved;
    }

    // March from a to b, listing the edges the segment crosses in order
    // (handleCrossing); Degenerate when it runs into a vertex first.
    Recovery march(const Seg& s, int t, int e) {
        const int a = s.a, b = s.b;
        int in = e;
        cross_.clear();
        cross_.push_back({T[t].v[e], T[t].v[(e + 1) % 3]});
        for (size_t steps = 0;; steps++) {
            const int nt = T[t].adj[e];
            if (nt < 0 || steps > T.size()) throw MeshError("segment recovery walked off the triangulation");
            for (int k = 0; k < 3; k++)
                if (T[nt].adj[k] == t) {
                    in = k;
                    break;
                }
            const int c = T[nt].v[(in + 2) % 3];
            if (c == b) return Recovery::Crossing;
            const double oc = orient(P[a], P[c], P[b]);
            if (oc == 0) return Recovery::Degenerate;
            if (oc < 0) {
                t = nt;
                e = (in + 1) % 3;
                cross_.push_back({T[nt].v[(in + 1) % 3], c});
            } else if (oc > 0) {
                t = nt;
                e = (in + 2) % 3;
                cross_.push_back({c, T[nt].v[in]});
            } else {
                throw MeshError("segment recovery met a non-finite orientation");
            }
        }
    }

    // Swap the crossed edges out, each once it is the diagonal of a convex
    // quadrilateral, re-queueing the others (handleCrossEdge).  Swapped edges
    // that still cross go back on the queue; the rest, other than a-b itself,
    // are collected for the Delaunay repair.
    void swapOut(const Seg& s) {
        const int a = s.a, b = s.b;
        fresh_.clear();
        const size_t k = cross_.size();
        const size_t limit = std::max<size_t>(size_t(1) << 22, 64 * k * k);
        size_t head = 0;
        while (head != cross_.size()) {
            if (head > limit) throw MeshError("segment recovery did not finish");
            const Seg c = cross_[head++];
            int t, e;
            if (!findEdge(c.a, c.b, t, e)) throw MeshError("a crossed edge is missing from the triangulation");
            const int u = T[t].adj[e];
            if (u < 0) throw MeshError("a crossed edge lies on the boundary");
            const int eu = slotOf(T[u].adj, t);
            const int w = T[u].v[(eu + 2) % 3];
            const int v0 = T[t].v[e], v1 = T[t].v[(e + 1) % 3], v2 = T[t].v[(e + 2) % 3];
            const double d0 = orient(P[v2], P[w], P[v0]);
            const double d1 = orient(P[v2], P[w], P[v1]);
            if (!(d0 < 0 && d1 > 0)) {
                cross_.push_back(c);
                continue;
            }
            flip(t, e);
            stamp(t, u);
            const double da = orient(P[a], P[v2], P[b]);
            const double db = orient(P[a], P[w], P[b]);
            if ((da > 0 && db < 0) || (da < 0 && db > 0))
                cross_.push_back({v2, w});
            else if (!((v2 == a && w == b) || (v2 == b && w == a)))
                fresh_.push_back({v2, w});
        }
    }

    // Swap the new edges until none fails the Delaunay test, sweeping the
    // list until a sweep swaps nothing (restoreDelaunayFromSegments).
    void repair() {
        bool swapped = true;
        for (long sweeps = 0; swapped; sweeps++) {
            if (sweeps > (1L << 22)) throw MeshError("Delaunay repair did not finish");
            swapped = false;
            for (Seg& s : fresh_) {
                int t, e;
                if (!findEdge(s.a, s.b, t, e)) continue;
                const int u = T[t].adj[e];
                if (u < 0) throw MeshError("a repaired edge lies on the boundary");
                const int eu = slotOf(T[u].adj, t);
                const int w = T[u].v[(eu + 2) % 3];
                const int v2 = T[t].v[(e + 2) % 3];
                if (!swaps(P[T[t].v[e]], P[T[t].v[(e + 1) % 3]], P[v2], P[w])) continue;
                flip(t, e);
                stamp(t, u);
                s = {v2, w};
                swapped = true;
            }
        }
    }

    // The segment loop (constrain, aximesh.cpp:712-746): segments already
    // present and segments through a vertex are left as they are.
    void recover(const Seg& s) {
        int t = -1, e = -1;
        if (classify(s, t, e) != Recovery::Crossing) return;
        if (march(s, t, e) != Recovery::Crossing) return;
        swapOut(s);
        repair();
    }

    // The triangle holding directed edge a -> b, rotating around a from
    // vtx[a] (findEdge, aximesh.cpp:237-252).
    bool findEdge(int a, int b, int& tOut, int& eOut) const {
        const int t0 = vtx[a];
        if (t0 < 0) return false;
        int t = t0;
        do {
            const int e = slotOf(T[t].v, a);
            if (e < 0) return false;
            if (T[t].v[(e + 1) % 3] == b) {
                tOut = t;
                eOut = e;
                return true;
            }
            t = T[t].adj[e];
        } while (t != t0 && t != -1);
        return false;
    }

    // -------- inside / outside (buildConstrainedEdge, removeFloodFill, aximesh.cpp:653-869)

    void keepInside(const std::vector<Seg>& segs) {
        const int nT = (int)T.size();
        std::vector<uint8_t> wall(3 * (size_t)nT, 0);
        for (const Seg& s : segs) {
            int t, e;
            if (!findEdge(s.a, s.b, t, e)) continue;
            wall[3 * t + e] = 1;
            const int u = T[t].adj[e];
            if (u == -1) continue;
            const int eu = slotOf(T[u].adj, t);
            if (eu < 0) throw MeshError("triangle links are not reciprocal");
            wall[3 * u + eu] = 1;
        }
        // -1 unvisited, 0 outside, 1 inside; seeded with every triangle on a
        // super vertex, depth first
        const int n = (int)P.size() - 3;
        std::vector<int8_t> side(nT, -1);
        std::vector<int> todo;
        for (int t = 0; t < nT; t++)
            if (T[t].v[0] >= n || T[t].v[1] >= n || T[t].v[2] >= n) {
                todo.push_back(t);
                side[t] = 0;
            }
        while (!todo.empty()) {
            const int t = todo.back();
            todo.pop_back();
            for (int e = 0; e < 3; e++) {
                const int nb = T[t].adj[e];
                if (nb == -1) continue;
                const int8_t s = wall[3 * t + e] == 1 ? int8_t(1 - side[t]) : side[t];
                if (side[nb] != -1) continue;
                side[nb] = s;
                todo.push_back(nb);
            }
        }
        std::vector<int> renum(nT, -1);
        int kept = 0;
        for (int t = 0; t < nT; t++)
            if (side[t] == 1) renum[t] = kept++;
        for (int t = 0; t < nT; t++) {
            if (renum[t] == -1) continue;
            Tri k = T[t];
            for (int i = 0; i < 3; i++) k.adj[i] = k.adj[i] == -1 ? -1 : renum[k.adj[i]];
            T[renum[t]] = k;
        }
        T.resize(kept);
        P.resize(n);
    }

    // -------- frontal refinement (frontalInit, advancingFront, aximesh.cpp:871-1103)

    std::vector<double> h;      // target edge length at each point
    std::vector<Front> state;   // per triangle
    std::priority_queue<FrontEdge, std::vector<FrontEdge>, SmallerRadius> front;

    void sizeField(const std::vector<Seg>& segs, double beta) {
        std::vector<double> lenSum(P.size());
        std::vector<int> count(P.size());
        for (const Seg& s : segs) {
            const double l = std::sqrt(dist2(P[s.a], P[s.b]));
            lenSum[s.a] += l;
            lenSum[s.b] += l;
            count[s.a] += 1;
            count[s.b] += 1;
        }
        h.assign(P.size(), std::numeric_limits<double>::max());
        for (size_t i = 0; i < h.size(); i++)
            if (count[i] != 0) h[i] = lenSum[i] / count[i];
        bool changed;
        int sweeps = 0;
        do {
            changed = false;
            if (sweeps++ > 100) throw MeshError("point search has stalled");
            for (const Tri& t : T)
                for (int e = 0; e < 3; e++) {
                    const int v = t.v[e], w = t.v[(e + 1) % 3];
                    const double len = std::sqrt(dist2(P[v], P[w]));
                    if (h[v] > h[w] + beta * len) {
                        h[v] = h[w] + beta * len;
                        changed = true;
                    }
                    if (h[w] > h[v] + beta * len) {
                        h[w] = h[v] + beta * len;
                        changed = true;
                    }
                }
        } while (changed);
    }

    // accepted when its circumradius R satisfies 3 R^2 <= (tol * mean h)^2
    void classifyTri(int t, double tol) {
        const Tri& k = T[t];
        const double hT = (h[k.v[0]] + h[k.v[1]] + h[k.v[2]]) / 3.0;
        const double target = tol * hT;
        const Pt &a = P[k.v[0]], &b = P[k.v[1]], &c = P[k.v[2]];
        const double ab = dist2(a, b), bc = dist2(b, c), ca = dist2(c, a);
        const double q = orient(a, b, c);
        const double r2 = (ab * bc * ca) / (4.0 * q * q);
        state[t] = (3.0 * r2 <= target * target) ? Front::Accepted : Front::Waiting;
    }

    // a waiting triangle next to the boundary or to an accepted one is active
    void activate(int t) {
        if (state[t] == Front::Accepted) return;
        for (int e = 0; e < 3; e++) {
            const int nb = T[t].adj[e];
            if (nb == -1 || state[nb] == Front::Accepted) {
                state[t] = Front::Active;
                break;
            }
        }
    }

    void pushFront(int t) {
        const Tri& k = T[t];
        for (int e = 0; e < 3; e++) {
            if (state[t] != Front::Accepted && (k.adj[e] == -1 || state[k.adj[e]] == Front::Accepted))
                front.push({t, e, std::sqrt(dist2(P[k.v[e]], P[k.v[(e + 1) % 3]])),
                            std::sqrt(circumR2(P[k.v[0]], P[k.v[1]], P[k.v[2]])), k.v[e], k.v[(e + 1) % 3]});
        }
    }

    bool current(const FrontEdge& f) const {
        const Tri& k = T[f.t];
        if (k.v[f.e] != f.v0) return false;
        if (k.v[(f.e + 1) % 3] != f.v1) return false;
        if (state[f.t] == Front::Accepted) return false;
        const int nb = k.adj[f.e];
        if (nb != -1 && state[nb] != Front::Accepted) return false;
        return true;
    }

    void refine(const Params& prm) {
        // the triangulation's last insert left its triangles here, numbered
        // before the outside was removed
        touched.clear();
        state.assign(T.size(), Front::Waiting);
        for (int t = 0; t < (int)T.size(); t++) classifyTri(t, prm.accept);
        for (int t = 0; t < (int)T.size(); t++) activate(t);
        for (int t = 0; t < (int)T.size(); t++) pushFront(t);

        const double rhoConst = 1 / std::sqrt(3.0);
        std::vector<int> dirty, mark;
        int epoch = 0;
        while (!front.empty()) {
            const FrontEdge f = front.top();
            front.pop();
            if (!current(f)) continue;
            // Rebay 1993, sec. 6.2: front edge PQ with midpoint M and half
            // length p; the new point X lies on the perpendicular bisector,
            // towards the triangle's circumcentre C, at distance
            // d = rho + sqrt(rho^2 - p^2) (rho = max(h_M / sqrt 3, p)), but
            // not beyond C (distance q)
            const Pt Pp = P[f.v0], Q = P[f.v1];
            const Pt apex = P[T[f.t].v[(f.e + 2) % 3]];
            const double len = f.len;
            const double p = len / 2.0;
            const Pt M = {0.5 * (Pp.x + Q.x), 0.5 * (Pp.y + Q.y)};
            const double ex = -(Q.y - Pp.y) / len;
            const double ey = (Q.x - Pp.x) / len;
            const double hM = 0.5 * (h[f.v0] + h[f.v1]);
            const double rhoM = rhoConst * hM;
            const double rho = (rhoM < p) ? p : rhoM;
            double d = rho + std::sqrt(rho * rho - p * p);
            const Pt C = circumcentre(Pp, Q, apex);
            const double q = (C.x - M.x) * ex + (C.y - M.y) * ey;
            if (q <= 0.0) continue;
            d = (q < d) ? q : d;
            if ((long long)P.size() >= prm.maxPoints) throw MeshError("refinement exceeded the point limit");
            P.push_back({M.x + d * ex, M.y + d * ey});
            if (!insert((int)P.size() - 1, f.t)) {
                P.pop_back();
                touched.clear();
                continue;
            }
            h.push_back(hM);
            state.resize(T.size(), Front::Waiting);
            mark.resize(T.size(), -1);
            ++epoch;
            dirty.clear();
            for (int t : touched) {
                if (mark[t] != epoch) {
                    mark[t] = epoch;
                    dirty.push_back(t);
                }
                for (int e = 0; e < 3; e++) {
                    const int nb = T[t].adj[e];
                    if (nb == -1) continue;
                    if (mark[nb] != epoch) {
                        mark[nb] = epoch;
                        dirty.push_back(nb);
                    }
                }
            }
            for (int t : dirty) classifyTri(t, prm.accept);
            for (int t : dirty) activate(t);
            for (int t : dirty) pushFront(t);
            touched.clear();
        }
    }

private:
    std::vector<int> stack_;
    std::vector<Seg> cross_, fresh_;

    void relink(int t, int from, int to) {
        for (int i = 0; i < 3; i++)
            if (T[t].adj[i] == from) {
                T[t].adj[i] = to;
                return;
            }
    }

    // after a swap, point every vertex of the two triangles at one of them
    void stamp(int t, int u) {
        for (int i = 0; i < 3; i++) {
            vtx[T[t].v[i]] = t;
            vtx[T[u].v[i]] = u;
        }
    }
};

}  // namespace

std::vector<Tri> insertAll(std::vector<Pt>& points) {
    Mesher m(points);
    m.triangulate((int)points.size());
    return std::move(m.T);
}

Result generate(const std::vector<Pt>& in, const std::vector<Seg>& segsIn, const Params& prm) {
    const int n = (int)in.size();
    if (n < 1) throw std::invalid_argument("the mesher needs points");
    for (const Pt& p : in)
        if (!std::isfinite(p.x) || !std::isfinite(p.y)) throw std::invalid_argument("non-finite point coordinate");
    for (const Seg& s : segsIn) {
        if (s.a < 0 || s.a >= n || s.b < 0 || s.b >= n) throw std::out_of_range("segment end out of range");
        if (s.a == s.b) throw std::invalid_argument("segment joins a point to itself");
    }

    // 1. bounding box and unit-box scaling (buildNormVariables, normalize)
    const auto [xlo, xhi] =
        std::minmax_element(in.begin(), in.end(), [](const Pt& a, const Pt& b) { return a.x < b.x; });
    const auto [ylo, yhi] =
        std::minmax_element(in.begin(), in.end(), [](const Pt& a, const Pt& b) { return a.y < b.y; });
    const double xMin = xlo->x, yMin = ylo->y;
    const double dx = xhi->x - xlo->x, dy = yhi->y - ylo->y;
    // the bin index divides by each extent
    if (!(dx > 0.0) || !(dy > 0.0)) throw std::invalid_argument("the points span no area (zero extent in x or y)");
    const double dMax = std::max(dx, dy);

    // 2. serpentine bin order (buildSortedPointIndices)
    const int nb = (int)std::lround(std::pow((double)n, 0.25));
    std::vector<int> bin(n), order(n);
    for (int k = 0; k < n; k++) {
        const int i = int(0.99 * nb * (in[k].x - xMin) / dx);
        const int j = int(0.99 * nb * (in[k].y - yMin) / dy);
        bin[k] = i % 2 == 0 ? i * nb + j : (i + 1) * nb - j - 1;
        order[k] = k;
    }
    std::sort(order.begin(), order.end(), [&bin](int a, int b) { return bin[a] < bin[b]; });
    std::vector<int> renum(n);
    for (int k = 0; k < n; k++) renum[order[k]] = k;

    Result r;
    r.points.reserve((size_t)n + 3);
    for (int k : order) r.points.push_back({(in[k].x - xMin) / dMax, (in[k].y - yMin) / dMax});

    // 3. Delaunay triangulation of the points
    Mesher m(r.points);
    m.triangulate(n);
    m.vtx.assign((size_t)n + 3, -1);
    for (int t = 0; t < (int)m.T.size(); t++)
        for (int i = 0; i < 3; i++) m.vtx[m.T[t].v[i]] = t;

    // 4. segments, renumbered, recovered in order
    r.segs.reserve(segsIn.size());
    for (const Seg& s : segsIn) r.segs.push_back({renum[s.a], renum[s.b]});
    for (const Seg& s : r.segs) m.recover(s);

    // 5. keep the inside; the super vertices go
    m.keepInside(r.segs);

    // 6-7. size field and frontal refinement
    m.sizeField(r.segs, prm.grading);
    m.refine(prm);

    // 8. back to the caller's coordinates
    for (Pt& p : r.points) p = {p.x * dMax + xMin, p.y * dMax + yMin};
    r.tris = std::move(m.T);
    return r;
}

}  // namespace axs::cdt
