// topology.cpp -- set-up builders: geometry, neighbour CSR, colouring and the
// agglomeration hierarchy with its GPU gather tables.
#include "topology.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace axs {

namespace {
constexpr double kPi = 3.14159265358979323846;
}

// preComputeVariables, memory_manger.cpp:512-617.  Same expressions in the same
// order, so the device reads bit-identical geometry.
void computeGeometry(const HostMesh& m, Geometry& g) {
    const int N = m.nCells, F = m.nFaces;
    g.invDPN.assign(F, 0.0);
    g.wP.assign(F, 0.0);
    g.dPB.assign(F, 0.0);
    g.length2D.assign(F, 0.0);
    g.invA2D.resize(N);
    for (int c = 0; c < N; c++) g.invA2D[c] = 2.0 * kPi * m.cellCenterR[c] / m.cellVolume[c];
    std::vector<double> offZ(F), offR(F);
    g.lsqWZ.resize(F);
    g.lsqWR.resize(F);
    for (int f = 0; f < F; f++) {
        g.length2D[f] = m.faceArea[f] / (2 * kPi * m.faceCenterR[f]);
        const int P = m.faceOwner[f], Nb = m.faceNeighbor[f];
        const double pfz = m.faceCenterZ[f] - m.cellCenterZ[P];
        const double pfr = m.faceCenterR[f] - m.cellCenterR[P];
        const double nz = m.faceNormalZ[f], nr = m.faceNormalR[f];
        if (Nb < 0) {
            g.dPB[f] = std::abs(pfz * nz + pfr * nr);
        } else {
            const double nfz = m.faceCenterZ[f] - m.cellCenterZ[Nb];
            const double nfr = m.faceCenterR[f] - m.cellCenterR[Nb];
            const double pnz = m.cellCenterZ[Nb] - m.cellCenterZ[P];
            const double pnr = m.cellCenterR[Nb] - m.cellCenterR[P];
            const double dP = std::abs(pfz * nz + pfr * nr);
            const double dN = std::abs(nfz * nz + nfr * nr);
            const double dPN = std::sqrt(pnz * pnz + pnr * pnr);
            g.invDPN[f] = 1.0 / dPN;
            g.wP[f] = dN / (dP + dN);
        }
        // least-squares offsets: neighbour centre, or face centre on a boundary
        const double dz = (Nb >= 0 ? m.cellCenterZ[Nb] : m.faceCenterZ[f]) - m.cellCenterZ[P];
        const double dr = (Nb >= 0 ? m.cellCenterR[Nb] : m.faceCenterR[f]) - m.cellCenterR[P];
        const double w = 1.0 / (dz * dz + dr * dr);
        offZ[f] = dz;
        offR[f] = dr;
        g.lsqWZ[f] = w * dz;
        g.lsqWR[f] = w * dr;
    }
    g.lsqInvZZ.assign(N, 0.0);
    g.lsqInvZR.assign(N, 0.0);
    g.lsqInvRR.assign(N, 0.0);
    for (int c = 0; c < N; c++) {
        double zz = 0.0, zr = 0.0, rr = 0.0;
        for (int k = m.cellFaceStart[c]; k < m.cellFaceStart[c + 1]; k++) {
            const int f = m.cellFaceIDs[k];
            zz += g.lsqWZ[f] * offZ[f];
            zr += g.lsqWZ[f] * offR[f];
            rr += g.lsqWR[f] * offR[f];
        }
        const double det = zz * rr - zr * zr;
        if (std::abs(det) <= 1.0e-30) continue;
        g.lsqInvZZ[c] = rr / det;
        g.lsqInvZR[c] = -zr / det;
        g.lsqInvRR[c] = zz / det;
    }
}

void computeSlotGeometry(const HostMesh& m, const Geometry& g, SlotGeometry& s) {
    const size_t S = m.cellFaceIDs.size();
    for (auto* v : {&s.area, &s.invDPN, &s.nz, &s.nr, &s.w, &s.tz, &s.tr, &s.lsqWZ, &s.lsqWR, &s.len2D})
        v->assign(S, 0.0);
    s.flags.assign(S, 0);
    for (int c = 0; c < m.nCells; c++)
        for (int k = m.cellFaceStart[c]; k < m.cellFaceStart[c + 1]; k++) {
            const int f = m.cellFaceIDs[k];
            if (f < 0 || f >= m.nFaces) continue;
            const int ow = m.faceOwner[f], nb = m.faceNeighbor[f];
            const bool own = ow == c;
            // orientation as in nonOrthoScalarDiffusionFlux / phiGradientGreenGauss:
            // the stored normal is flipped where c is the face's neighbour
            const bool flip = nb == c;
            const double nz = flip ? -m.faceNormalZ[f] : m.faceNormalZ[f];
            const double nr = flip ? -m.faceNormalR[f] : m.faceNormalR[f];
            s.area[k] = m.faceArea[f];
            s.invDPN[k] = g.invDPN[f];
            s.nz[k] = nz;
            s.nr[k] = nr;
            s.w[k] = own ? g.wP[f] : 1.0 - g.wP[f];
            s.lsqWZ[k] = own ? g.lsqWZ[f] : -g.lsqWZ[f];
            s.lsqWR[k] = own ? g.lsqWR[f] : -g.lsqWR[f];
            s.len2D[k] = g.length2D[f];
            s.flags[k] = (uint8_t)((own ? SlotGeometry::kOwner : 0) | (m.faceCenterR[f] > 1.0e-30 ? 0 : SlotGeometry::kAxis));
            if (nb >= 0) {
                const int o = own ? nb : ow;
                const double area = m.faceArea[f];
                const double dz = m.cellCenterZ[o] - m.cellCenterZ[c];
                const double dr = m.cellCenterR[o] - m.cellCenterR[c];
                const double aOverNd = area * g.invDPN[f];
                s.tz[k] = area * nz - aOverNd * dz;
                s.tr[k] = area * nr - aOverNd * dr;
            }
        }
}

void cellNeighbourCSR(const HostMesh& m, std::vector<int>& start, std::vector<int>& nb) {
    start.assign(m.cellFaceStart.begin(), m.cellFaceStart.end());
    start[0] = 0;
    nb.resize(m.cellFaceIDs.size());
    for (int c = 0; c < m.nCells; c++)
        for (int k = m.cellFaceStart[c]; k < m.cellFaceStart[c + 1]; k++) {
            const int f = m.cellFaceIDs[k];
            int other = -1;
            if (f >= 0 && f < m.nFaces) {
                if (m.faceOwner[f] == c) other = m.faceNeighbor[f];
                else if (m.faceNeighbor[f] == c) other = m.faceOwner[f];
            }
            nb[k] = other;
        }
}

void buildSweepBlocks(int nCells, const std::vector<int>& start, const std::vector<int>& nb, int cells, int rings,
                      int rowSlots, SweepBlocks& out, bool patches) {
    if (cells <= 0 || rings < 1 || rowSlots < 1) throw std::invalid_argument("bad sweep block parameters");
    out = SweepBlocks{};
    out.cells = cells;
    out.rings = rings;
    out.rowSlots = rowSlots;
    out.patches = patches;
    out.nBlocks = (nCells + cells - 1) / cells;
    out.haloStart.assign(1, 0);
    out.ringEnd.reserve((size_t)out.nBlocks * rings);
    out.locOwn.assign(start[nCells], (int16_t)-1);
    for (int c = 0; c < nCells; c++) out.maxSlots = std::max(out.maxSlots, start[c + 1] - start[c]);
    // the blocks' cells: index ranges, or breadth-first patches of `cells`
    // cells each (every patch full but the last; a patch continues from the
    // next unassigned cell when its region runs out)
    std::vector<int> blockOf(nCells, -1), members;
    if (patches) {
        out.own.assign((size_t)out.nBlocks * cells, -1);
        std::vector<int> queue;
        int seed = 0, b = 0, filled = 0;
        size_t qh = 0;
        queue.reserve(nCells);
        while (filled < nCells) {
            if (qh == queue.size()) {
                while (blockOf[seed] >= 0) seed++;
                blockOf[seed] = b;
                queue.push_back(seed);
            }
            const int c = queue[qh++];
            out.own[(size_t)b * cells + (filled - b * cells)] = c;
            if (++filled % cells == 0 || filled == nCells) {
                // the patch is full: the rest of the queue goes back to the pool
                for (size_t i = qh; i < queue.size(); i++) {
                    blockOf[queue[i]] = -1;
                    seed = std::min(seed, queue[i]);
                }
                queue.clear();
                qh = 0;
                std::sort(out.own.begin() + (size_t)b * cells, out.own.begin() + (size_t)b * cells + (filled - b * cells));
                b++;
                continue;
            }
            for (int k = start[c]; k < start[c + 1]; k++) {
                const int o = nb[k];
                if (o >= 0 && blockOf[o] < 0) {
                    blockOf[o] = b;
                    queue.push_back(o);
                }
            }
        }
    } else {
        for (int c = 0; c < nCells; c++) blockOf[c] = c / cells;
    }
    std::vector<int> stamp(nCells, -1), where(nCells, -1), frontier, ring;
    for (int b = 0; b < out.nBlocks; b++) {
        const int a = b * cells, e = std::min(nCells, a + cells);
        const size_t h0 = out.halo.size();
        members.clear();
        if (patches) {
            for (int i = 0; i < cells; i++)
                if (out.own[(size_t)b * cells + i] >= 0) members.push_back(out.own[(size_t)b * cells + i]);
        } else {
            for (int c = a; c < e; c++) members.push_back(c);
        }
        frontier.clear();
        for (size_t i = 0; i < members.size(); i++) {
            const int c = members[i];
            stamp[c] = b;
            where[c] = (int)i;  // LDS slot of a block cell: its position in the block
            frontier.push_back(c);
        }
        for (int r = 1; r <= rings; r++) {
            ring.clear();
            for (int c : frontier)
                for (int k = start[c]; k < start[c + 1]; k++) {
                    const int o = nb[k];
                    if (o >= 0 && stamp[o] != b) {
                        stamp[o] = b;
                        ring.push_back(o);
                    }
                }
            std::sort(ring.begin(), ring.end());
            out.halo.insert(out.halo.end(), ring.begin(), ring.end());
            out.ringEnd.push_back((int)(out.halo.size() - h0));
            frontier.swap(ring);
        }
        const int nh = (int)(out.halo.size() - h0);
        out.maxHalo = std::max(out.maxHalo, nh);
        for (int i = 0; i < nh; i++) where[out.halo[h0 + i]] = cells + i;
        // block cells sit at their position in the block, halo cells after
        // the block: every neighbour a later sweep reads from LDS is listed
        auto local = [&](int o) -> int16_t { return o < 0 ? (int16_t)-1 : (int16_t)where[o]; };
        for (int c : members)
            for (int k = start[c]; k < start[c + 1]; k++) out.locOwn[k] = local(nb[k]);
        // rows of rings 1 .. rings - 1, which the sweeps after the first read from LDS
        const int inner = rings >= 2 ? out.ringEnd[(size_t)b * rings + rings - 2] : 0;
        for (int i = 0; i < nh; i++) {
            const int c = out.halo[h0 + i];
            for (int q = 0; q < rowSlots; q++) {
                const int k = start[c] + q;
                out.locHalo.push_back((i < inner && k < start[c + 1]) ? local(nb[k]) : (int16_t)-1);
            }
        }
        for (int i = 0; i < nh; i++) where[out.halo[h0 + i]] = -1;
        for (int c : members) where[c] = -1;
        out.haloStart.push_back((int)out.halo.size());
    }
}

int greedyColoring(int n, const int* start, const int* nb, std::vector<int>& colorStart,
                   std::vector<int>& cellOrder) {
    colorStart.clear();
    cellOrder.clear();
    if (n <= 0) return 0;
    std::vector<int> color(n, -1);
    std::vector<int> usedBy;  // usedBy[c] == cell  <=>  colour c is taken by a neighbour of cell
    int nColors = 0;
    for (int c = 0; c < n; c++) {
        for (int k = start[c]; k < start[c + 1]; k++) {
            const int o = nb[k];
            if (o < 0 || color[o] < 0) continue;
            if (color[o] >= (int)usedBy.size()) usedBy.resize(color[o] + 1, -1);
            usedBy[color[o]] = c;
        }
        int pick = 0;
        while (pick < (int)usedBy.size() && usedBy[pick] == c) pick++;
        color[c] = pick;
        nColors = std::max(nColors, pick + 1);
    }
    colorStart.assign(nColors + 1, 0);
    for (int c = 0; c < n; c++) colorStart[color[c] + 1]++;
    for (int k = 0; k < nColors; k++) colorStart[k + 1] += colorStart[k];
    cellOrder.resize(n);
    std::vector<int> cursor(colorStart.begin(), colorStart.end() - 1);
    for (int c = 0; c < n; c++) cellOrder[cursor[color[c]]++] = c;
    return nColors;
}

namespace {

// buildAgglomerationMap, multigrid.cu:79-194: breadth-first groups of up to 4
// seeded in cell order, stranded singletons merged into their smallest
// neighbouring group, then renumbered by first appearance.
int aggregate(const Level& f, std::vector<int>& owner) {
    const int kTarget = 4;
    owner.assign(f.n, -1);
    int groups = 0;
    int grown[kTarget];
    for (int seed = 0; seed < f.n; seed++) {
        if (owner[seed] >= 0) continue;
        const int g = groups++;
        owner[seed] = g;
        int count = 1;
        grown[0] = seed;
        for (int q = 0; q < count && count < kTarget; q++) {
            const int cell = grown[q];
            for (int k = f.start[cell]; k < f.start[cell + 1] && count < kTarget; k++) {
                const int o = f.nb[k];
                if (o < 0 || owner[o] >= 0) continue;
                owner[o] = g;
                grown[count++] = o;
            }
        }
    }
    std::vector<int> size(groups, 0);
    for (int c = 0; c < f.n; c++) size[owner[c]]++;
    for (int c = 0; c < f.n; c++) {
        const int g = owner[c];
        if (size[g] != 1) continue;
        int best = -1;
        for (int k = f.start[c]; k < f.start[c + 1]; k++) {
            const int o = f.nb[k];
            if (o < 0) continue;
            const int go = owner[o];
            if (go == g) continue;
            if (best < 0 || size[go] < size[best]) best = go;
        }
        if (best < 0) continue;
        owner[c] = best;
        size[best]++;
        size[g] = 0;
    }
    std::vector<int> renum(groups, -1);
    int next = 0;
    for (int c = 0; c < f.n; c++) {
        int& r = renum[owner[c]];
        if (r < 0) r = next++;
        owner[c] = r;
    }
    return next;
}

// coarsenGrid, multigrid.cu:196-287: coarse CSR sorted by (coarse cell, coarse
// neighbour) and the fine-slot -> coarse-slot map; plus the gather tables.
Level coarsen(Level& f) {
    Level c;
    c.n = aggregate(f, f.toCoarse);
    struct Ref { int row, col, slot; };
    std::vector<Ref> refs;
    refs.reserve(f.nb.size());
    for (int i = 0; i < f.n; i++) {
        const int ci = f.toCoarse[i];
        for (int k = f.start[i]; k < f.start[i + 1]; k++) {
            const int o = f.nb[k];
            if (o < 0) continue;
            const int co = f.toCoarse[o];
            if (co != ci) refs.push_back(Ref{ci, co, k});
        }
    }
    // slots are visited in ascending order, so a stable sort on (row, col)
    // keeps each coarse entry's fine slots ascending
    std::stable_sort(refs.begin(), refs.end(),
                     [](const Ref& a, const Ref& b) { return a.row != b.row ? a.row < b.row : a.col < b.col; });
    f.slotToCoarseSlot.assign(f.nb.size(), -1);
    c.start.assign(c.n + 1, 0);
    f.cslotStart.clear();
    f.cslotFine.clear();
    for (size_t i = 0; i < refs.size(); i++) {
        const bool fresh = i == 0 || refs[i].row != refs[i - 1].row || refs[i].col != refs[i - 1].col;
        if (fresh) {
            c.nb.push_back(refs[i].col);
            c.start[refs[i].row + 1]++;
            f.cslotStart.push_back((int)f.cslotFine.size());
        }
        f.slotToCoarseSlot[refs[i].slot] = (int)c.nb.size() - 1;
        f.cslotFine.push_back(refs[i].slot);
    }
    f.cslotStart.push_back((int)f.cslotFine.size());
    for (int i = 0; i < c.n; i++) c.start[i + 1] += c.start[i];
    // members of each coarse cell, ascending fine index
    f.memberStart.assign(c.n + 1, 0);
    for (int i = 0; i < f.n; i++) f.memberStart[f.toCoarse[i] + 1]++;
    for (int i = 0; i < c.n; i++) f.memberStart[i + 1] += f.memberStart[i];
    f.members.resize(f.n);
    std::vector<int> cur(f.memberStart.begin(), f.memberStart.end() - 1);
    for (int i = 0; i < f.n; i++) f.members[cur[f.toCoarse[i]]++] = i;
    return c;
}

void clearCoarseMaps(Level& f) {
    f.toCoarse.clear();
    f.slotToCoarseSlot.clear();
    f.memberStart.clear();
    f.members.clear();
    f.cslotStart.clear();
    f.cslotFine.clear();
}

}  // namespace

std::vector<Level> buildHierarchy(int n, const std::vector<int>& start, const std::vector<int>& nb,
                                  int maxLevels) {
    const int kMinCells = 64;          // multigrid.cu:55
    const double kMaxShrink = 0.8;     // multigrid.cu:60
    std::vector<Level> levels;
    Level fine;
    fine.n = n;
    fine.start = start;
    fine.nb = nb;
    while (fine.n > kMinCells && !(maxLevels > 0 && (int)levels.size() + 1 >= maxLevels)) {
        Level coarse = coarsen(fine);
        if (coarse.n >= (int)(fine.n * kMaxShrink)) {
            clearCoarseMaps(fine);
            break;
        }
        levels.push_back(std::move(fine));
        fine = std::move(coarse);
    }
    levels.push_back(std::move(fine));
    return levels;
}

}  // namespace axs
