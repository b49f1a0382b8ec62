// oracle.cpp -- CPU FP64 restatement of AxiSim's SIMPLE pressure-velocity loop.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
// solver in libaxs.so.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load liboracle.so, and only as the checker or as the
// timed CPU baseline -- never as the product path.
//
// Each function restates one reference kernel or host routine in plain serial
// C++, in the reference's operation order, and cites the file:line it follows
// (paths relative to the reference repository root).  Built with
// -ffp-contract=off so that every multiply-add rounds exactly as written.
// Loops whose iterations each write only their own cell, face or row slots
// carry an OpenMP `parallel for` (the CPU baseline's multi-thread sample); the
// results do not depend on the thread count.  Reductions and scatters stay
// serial.  orc_set_threads chooses the count (the binding sets 1 at load).
//
// Pinning (see tests/test_oracle_*.py): geometry closed forms, gradient
// exactness for linear fields, dense P^T A P, dense Jacobi / Gauss-Seidel
// iteration matrices, M-matrix and symmetry identities of the assembled
// systems, Hagen-Poiseuille on the pipe case, hand-computed agglomeration
// maps, residual norm closed forms and the index-4 continuity scale capture;
// each face term on its own through the orc_probe_* entry points (exactness
// for linear / cubic fields and polynomials in time, crafted limiter cases,
// closed-form boundary values).  What stays unpinned is listed as "parity
// unpinned" in the module docstring of tests/test_oracle_pins.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include <omp.h>

#include "../include/axs.h"

namespace {

constexpr double PI = 3.14159265358979323846;  // std::numbers::pi, math_func / multiblock.h:269

// ---------------------------------------------------------------------------
// Data model (structs/solver_struct.h:103-120, 265-321, 333-397, 431-454)
// ---------------------------------------------------------------------------
struct Mesh {
    int N = 0, F = 0;
    std::vector<double> cZ, cR, vol, area2D;
    std::vector<int> faceStart, faceIDs;
    std::vector<int> owner, neighbor, group;
    std::vector<double> nZ, nR, fZ, fR, area;
    // preComputeVariables (memory_manger.cpp:512-617)
    std::vector<double> invDPN, wP, dPB, invA2D, lsqWZ, lsqWR, lsqInvZZ, lsqInvZR, lsqInvRR, length2D;
};

struct BC {
    int nGroups = 0;
    std::vector<uint8_t> type;
    std::vector<double> value, length, amplitude, frequency, mean;
    std::vector<uint8_t> pulsatile;
    // Michaelis-Menten / Hill kinetics and wall-layer resistance per group, with
    // Km^n and K2^m evaluated once (createBoundaryFieldHost/Device,
    // memory_manger.cpp:51-189)
    std::vector<double> vmax, km, hillN, inhibM, k2, v2, rtot, kmN, k2M;
};

// Reactive-wall state per face (FVFaceDevice cw / ocrWall): the wall
// concentration of the last Newton solve and the face's consumption area*J.
struct Wall {
    std::vector<double> cw, ocr;
};

struct Coeff {
    int N = 0, nRefs = 0;
    std::vector<double> AF, AC, b, invAC;
    std::vector<int> faceStart, faceNeighbor;
};

// memory_manger.cpp:512-617
void preCompute(Mesh& m) {
    const int N = m.N, F = m.F;
    m.invDPN.assign(F, 0.0);
    m.wP.assign(F, 0.0);
    m.dPB.assign(F, 0.0);
    m.invA2D.assign(N, 0.0);
    m.length2D.assign(F, 0.0);
    for (int c = 0; c < N; c++) m.invA2D[c] = 2.0 * PI * m.cR[c] / m.vol[c];
    for (int f = 0; f < F; f++) {
        m.length2D[f] = m.area[f] / (2 * PI * m.fR[f]);
        const int nb = m.neighbor[f], ow = m.owner[f];
        const double dzPF = m.fZ[f] - m.cZ[ow];
        const double drPF = m.fR[f] - m.cR[ow];
        if (nb < 0) {
            m.dPB[f] = std::abs(dzPF * m.nZ[f] + drPF * m.nR[f]);
            continue;
        }
        const double dzNF = m.fZ[f] - m.cZ[nb];
        const double drNF = m.fR[f] - m.cR[nb];
        const double dzPN = m.cZ[nb] - m.cZ[ow];
        const double drPN = m.cR[nb] - m.cR[ow];
        const double dPF = std::abs(dzPF * m.nZ[f] + drPF * m.nR[f]);
        const double dNF = std::abs(dzNF * m.nZ[f] + drNF * m.nR[f]);
        const double dPN = std::sqrt(dzPN * dzPN + drPN * drPN);
        const double denom = dPF + dNF;
        m.invDPN[f] = 1.0 / dPN;
        m.wP[f] = dNF / denom;
    }
    m.lsqWZ.assign(F, 0.0);
    m.lsqWR.assign(F, 0.0);
    m.lsqInvZZ.assign(N, 0.0);
    m.lsqInvZR.assign(N, 0.0);
    m.lsqInvRR.assign(N, 0.0);
    std::vector<double> dZ(F, 0.0), dR(F, 0.0);
    for (int f = 0; f < F; f++) {
        const int ow = m.owner[f], nb = m.neighbor[f];
        const double dz = (nb >= 0 ? m.cZ[nb] : m.fZ[f]) - m.cZ[ow];
        const double dr = (nb >= 0 ? m.cR[nb] : m.fR[f]) - m.cR[ow];
        const double w = 1.0 / (dz * dz + dr * dr);
        dZ[f] = dz;
        dR[f] = dr;
        m.lsqWZ[f] = w * dz;
        m.lsqWR[f] = w * dr;
    }
    for (int c = 0; c < N; c++) {
        double Szz = 0.0, Szr = 0.0, Srr = 0.0;
        for (int k = m.faceStart[c]; k < m.faceStart[c + 1]; k++) {
            const int f = m.faceIDs[k];
            Szz += m.lsqWZ[f] * dZ[f];
            Szr += m.lsqWZ[f] * dR[f];
            Srr += m.lsqWR[f] * dR[f];
        }
        const double det = Szz * Srr - Szr * Szr;
        if (std::abs(det) <= 1.0e-30) continue;
        m.lsqInvZZ[c] = Srr / det;
        m.lsqInvZR[c] = -Szr / det;
        m.lsqInvRR[c] = Szz / det;
    }
}

// buildCellFaceCSR, memory_manger.cpp:287-323
void buildCSR(const Mesh& m, std::vector<int>& fs, std::vector<int>& fn) {
    fs.assign(m.N + 1, 0);
    int total = 0;
    for (int c = 0; c < m.N; c++) {
        fs[c] = total;
        total += m.faceStart[c + 1] - m.faceStart[c];
    }
    fs[m.N] = total;
    fn.clear();
    fn.reserve(total);
    for (int c = 0; c < m.N; c++) {
        for (int k = m.faceStart[c]; k < m.faceStart[c + 1]; k++) {
            const int f = m.faceIDs[k];
            int nb = -1;
            if (f >= 0 && f < m.F) {
                if (m.owner[f] == c) nb = m.neighbor[f];
                else if (m.neighbor[f] == c) nb = m.owner[f];
            }
            fn.push_back(nb);
        }
    }
}

// buildMeshColoring, memory_manger.cpp:334-409
int buildColoring(int N, const int* fs, const int* fn, std::vector<int>& colorStart,
                  std::vector<int>& cellOrder) {
    colorStart.clear();
    cellOrder.clear();
    if (N <= 0) return 0;
    std::vector<int> color(N, -1), stamp;
    int nColors = 0;
    for (int c = 0; c < N; c++) {
        for (int k = fs[c]; k < fs[c + 1]; k++) {
            const int nb = fn[k];
            if (nb < 0 || color[nb] < 0) continue;
            if (color[nb] >= (int)stamp.size()) stamp.resize(color[nb] + 1, -1);
            stamp[color[nb]] = c;
        }
        int pick = 0;
        while (pick < (int)stamp.size() && stamp[pick] == c) pick++;
        color[c] = pick;
        if (pick >= nColors) nColors = pick + 1;
    }
    colorStart.assign(nColors + 1, 0);
    for (int c = 0; c < N; c++) colorStart[color[c] + 1]++;
    for (int k = 0; k < nColors; k++) colorStart[k + 1] += colorStart[k];
    cellOrder.assign(N, 0);
    std::vector<int> fill = colorStart;
    for (int c = 0; c < N; c++) cellOrder[fill[color[c]]++] = c;
    return nColors;
}

// ---------------------------------------------------------------------------
// Agglomeration hierarchy (multigrid.cu:41-325)
// ---------------------------------------------------------------------------
struct Grid {
    int nCells = 0;
    std::vector<int> faceStart, faceNeighbor, cellToCoarse, fineSlotToCoarseSlot;
    // the fine cells of each coarse cell in ascending order (CSR over the
    // coarse cells): restriction and the coarse operator gather per coarse
    // cell in that order, which is the order of a serial scatter over fine cells
    std::vector<int> memberStart, members;
    bool isCoarsest() const { return cellToCoarse.empty(); }
    int nRefs() const { return (int)faceNeighbor.size(); }
};

// buildAgglomerationMap, multigrid.cu:79-194
int buildAgglomerationMap(const Grid& fine, std::vector<int>& c2c) {
    const int target = 4;
    c2c.assign(fine.nCells, -1);
    int nCoarse = 0;
    std::vector<int> members;
    for (int n = 0; n < fine.nCells; n++) {
        if (c2c[n] >= 0) continue;
        const int c = nCoarse++;
        c2c[n] = c;
        members.clear();
        members.push_back(n);
        for (int mi = 0; mi < (int)members.size() && (int)members.size() < target; mi++) {
            const int cell = members[mi];
            for (int k = fine.faceStart[cell];
                 k < fine.faceStart[cell + 1] && (int)members.size() < target; k++) {
                const int nb = fine.faceNeighbor[k];
                if (nb < 0) continue;
                if (c2c[nb] >= 0) continue;
                c2c[nb] = c;
                members.push_back(nb);
            }
        }
    }
    std::vector<int> size(nCoarse, 0);
    for (int n = 0; n < fine.nCells; n++) size[c2c[n]]++;
    for (int n = 0; n < fine.nCells; n++) {
        const int c = c2c[n];
        if (size[c] != 1) continue;
        int best = -1, bestSize = 0;
        for (int k = fine.faceStart[n]; k < fine.faceStart[n + 1]; k++) {
            const int nb = fine.faceNeighbor[k];
            if (nb < 0) continue;
            const int cnb = c2c[nb];
            if (cnb == c) continue;
            if (best < 0 || size[cnb] < bestSize) {
                best = cnb;
                bestSize = size[cnb];
            }
        }
        if (best < 0) continue;
        c2c[n] = best;
        size[best]++;
        size[c] = 0;
    }
    std::vector<int> remap(nCoarse, -1);
    int nCompact = 0;
    for (int n = 0; n < fine.nCells; n++) {
        const int c = c2c[n];
        if (remap[c] < 0) remap[c] = nCompact++;
        c2c[n] = remap[c];
    }
    return nCompact;
}

// coarsenGrid, multigrid.cu:196-287
Grid coarsenGrid(Grid& fine) {
    Grid coarse;
    const int nCoarse = buildAgglomerationMap(fine, fine.cellToCoarse);
    coarse.nCells = nCoarse;
    struct Ref { int cc, cn, slot; };
    std::vector<Ref> refs;
    for (int n = 0; n < fine.nCells; n++) {
        const int cn = fine.cellToCoarse[n];
        for (int k = fine.faceStart[n]; k < fine.faceStart[n + 1]; k++) {
            const int nb = fine.faceNeighbor[k];
            if (nb < 0) continue;
            const int cnb = fine.cellToCoarse[nb];
            if (cnb != cn) refs.push_back(Ref{cn, cnb, k});
        }
    }
    std::stable_sort(refs.begin(), refs.end(), [](const Ref& a, const Ref& b) {
        if (a.cc != b.cc) return a.cc < b.cc;
        return a.cn < b.cn;
    });
    fine.fineSlotToCoarseSlot.assign(fine.nRefs(), -1);
    coarse.faceStart.assign(nCoarse + 1, 0);
    for (size_t i = 0; i < refs.size(); i++) {
        const bool newEntry = i == 0 || refs[i].cc != refs[i - 1].cc || refs[i].cn != refs[i - 1].cn;
        if (newEntry) {
            coarse.faceNeighbor.push_back(refs[i].cn);
            coarse.faceStart[refs[i].cc + 1]++;
        }
        fine.fineSlotToCoarseSlot[refs[i].slot] = (int)coarse.faceNeighbor.size() - 1;
    }
    for (int c = 0; c < nCoarse; c++) coarse.faceStart[c + 1] += coarse.faceStart[c];
    return coarse;
}

// buildLevels, multigrid.cu:298-325 (+ max_levels cap, a build extension:
// 0 keeps the reference rule)
std::vector<Grid> buildLevels(Grid fine, int maxLevels) {
    std::vector<Grid> levels;
    while (true) {
        if (fine.nCells <= 64) break;
        if (maxLevels > 0 && (int)levels.size() + 1 >= maxLevels) break;
        Grid coarse = coarsenGrid(fine);
        if (coarse.nCells >= (int)(fine.nCells * 0.8)) {
            fine.cellToCoarse.clear();
            fine.fineSlotToCoarseSlot.clear();
            break;
        }
        levels.push_back(std::move(fine));
        fine = std::move(coarse);
    }
    levels.push_back(std::move(fine));
    return levels;
}

struct MGLevel {
    Grid grid;
    Coeff coeff;
    std::vector<double> x, res;
    std::vector<int> colorStart, cellOrder;  // greedy colouring (Gauss-Seidel smoother)
};

struct MGConfig {
    int maxIter = 3;
    double weight = 0.6;
    int linearSweep = 30;
    int pre = 1, post = 1;
    int smoother = AXS_MG_SMOOTH_JACOBI;
};

// jacobiRow, multigrid.cu:447-484
void jacobiRow(const Coeff& c, const std::vector<double>& xOld, std::vector<double>& xNew, int n,
               double weight) {
    const double AC = c.AC[n];
    if (std::fabs(AC) < 1.0e-30) {
        xNew[n] = xOld[n];
        return;
    }
    double Ax = AC * xOld[n];
    for (int k = c.faceStart[n]; k < c.faceStart[n + 1]; k++) {
        const int nb = c.faceNeighbor[k];
        if (nb != -1) Ax += c.AF[k] * xOld[nb];
    }
    xNew[n] = xOld[n] + weight * (c.b[n] - Ax) / AC;
}

// smoothen / smoothenRegular / jacobiSingleBlock, multigrid.cu:486-541, 664-725.
// The single-block and the multi-block smoother compute the same sweeps.
void smoothen(MGLevel& L, double weight, int sweeps) {
    const int N = L.grid.nCells;
    if (N <= 0 || sweeps <= 0) return;
    std::vector<double> xNew(N);
    for (int s = 0; s < sweeps; s++) {
#pragma omp parallel for schedule(static)
        for (int n = 0; n < N; n++) jacobiRow(L.coeff, L.x, xNew, n, weight);
        L.x.swap(xNew);
    }
}

// Build extension (no reference counterpart): multi-colour Gauss-Seidel smoothing
// of a level, colours from buildMeshColoring (memory_manger.cpp:334-409) of the
// level's graph, each colour updated in place in colour-major cell order:
// x_n = (b_n - sum AF x_nb) / AC_n, rows with |AC| < 1e-30 left unchanged (as
// jacobiRow leaves them).
void smoothenGS(MGLevel& L, int sweeps) {
    const int N = L.grid.nCells;
    if (N <= 0 || sweeps <= 0) return;
    if (L.colorStart.empty())
        buildColoring(N, L.grid.faceStart.data(), L.grid.faceNeighbor.data(), L.colorStart, L.cellOrder);
    const Coeff& c = L.coeff;
    const int nColors = (int)L.colorStart.size() - 1;
    for (int s = 0; s < sweeps; s++)
        for (int col = 0; col < nColors; col++)
            for (int t = L.colorStart[col]; t < L.colorStart[col + 1]; t++) {
                const int n = L.cellOrder[t];
                const double AC = c.AC[n];
                if (std::fabs(AC) < 1.0e-30) continue;
                double val = c.b[n];
                for (int k = c.faceStart[n]; k < c.faceStart[n + 1]; k++) {
                    const int nb = c.faceNeighbor[k];
                    if (nb >= 0) val -= c.AF[k] * L.x[nb];
                }
                L.x[n] = val / AC;
            }
}

// residualRaw(sign = true, no scaling), residuals.cuh:17-48
void residualSigned(const Coeff& c, const std::vector<double>& x, std::vector<double>& res) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < c.N; n++) {
        double Ax = c.AC[n] * x[n];
        for (int k = c.faceStart[n]; k < c.faceStart[n + 1]; k++) {
            const int nb = c.faceNeighbor[k];
            if (nb >= 0) Ax += c.AF[k] * x[nb];
        }
        res[n] = c.b[n] - Ax;
    }
}

// buildCoarseOperatorKernel, multigrid.cu:350-389 (serial scatter in fine-cell
// order; the reference's atomicAdd order is unspecified)
// Gathered per coarse cell over its fine cells in ascending order: every sum
// is added in the serial scatter's order (a fine cell's slots map into its
// coarse cell's row), so the result does not depend on the thread count.
void buildCoarseOperator(const MGLevel& fine, MGLevel& coarse) {
    const Coeff& f = fine.coeff;
    const Grid& g = fine.grid;
    Coeff& cc = coarse.coeff;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < cc.N; c++) {
        for (int k = cc.faceStart[c]; k < cc.faceStart[c + 1]; k++) cc.AF[k] = 0.0;
        double ac = 0.0;
        for (int j = g.memberStart[c]; j < g.memberStart[c + 1]; j++) {
            const int n = g.members[j];
            double diag = f.AC[n];
            for (int k = f.faceStart[n]; k < f.faceStart[n + 1]; k++) {
                const int nb = f.faceNeighbor[k];
                if (nb < 0) continue;
                const int slot = g.fineSlotToCoarseSlot[k];
                if (slot >= 0) cc.AF[slot] += f.AF[k];
                else diag += f.AF[k];
            }
            ac += diag;
        }
        cc.AC[c] = ac;
    }
}

struct Multigrid {
    std::vector<MGLevel> levels;
    MGConfig cfg;

    void init(const std::vector<int>& fs, const std::vector<int>& fn, int N, int maxLevels) {
        Grid g;
        g.nCells = N;
        g.faceStart = fs;
        g.faceNeighbor = fn;
        std::vector<Grid> grids = buildLevels(std::move(g), maxLevels);
        levels.clear();
        for (Grid& gr : grids) {
            MGLevel L;
            L.grid = std::move(gr);
            if (!L.grid.isCoarsest()) {
                Grid& q = L.grid;
                int nc = 0;
                for (int c : q.cellToCoarse) nc = std::max(nc, c + 1);
                q.memberStart.assign(nc + 1, 0);
                for (int c : q.cellToCoarse) q.memberStart[c + 1]++;
                for (int c = 0; c < nc; c++) q.memberStart[c + 1] += q.memberStart[c];
                q.members.resize(q.nCells);
                std::vector<int> fill(q.memberStart.begin(), q.memberStart.end() - 1);
                for (int n = 0; n < q.nCells; n++) q.members[fill[q.cellToCoarse[n]]++] = n;
            }
            const int n = L.grid.nCells;
            L.coeff.N = n;
            L.coeff.nRefs = L.grid.nRefs();
            L.coeff.AC.assign(n, 0.0);
            L.coeff.b.assign(n, 0.0);
            L.coeff.invAC.assign(n, 0.0);
            L.coeff.AF.assign(L.coeff.nRefs, 0.0);
            L.coeff.faceStart = L.grid.faceStart;
            L.coeff.faceNeighbor = L.grid.faceNeighbor;
            L.x.assign(n, 0.0);
            L.res.assign(n, 0.0);
            levels.push_back(std::move(L));
        }
    }

    // vCycle, multigrid.cu:607-662
    void vCycle(int l) {
        const int nL = (int)levels.size();
        if (l < 0 || l >= nL) return;
        MGLevel& L = levels[l];
        auto smooth = [&](int sweeps) {
            if (cfg.smoother == AXS_MG_SMOOTH_RBGS) smoothenGS(L, sweeps);
            else smoothen(L, cfg.weight, sweeps);
        };
        if (L.grid.isCoarsest()) {
            smooth(cfg.linearSweep);
            return;
        }
        smooth(cfg.pre);
        residualSigned(L.coeff, L.x, L.res);
        MGLevel& C = levels[l + 1];
        // buildRestriction, multigrid.cu:398-410, 568-582
        // (gathered per coarse cell in ascending fine order, as a serial scatter adds)
        const Grid& g = L.grid;
#pragma omp parallel for schedule(static)
        for (int c = 0; c < C.grid.nCells; c++) {
            double b = 0.0;
            for (int j = g.memberStart[c]; j < g.memberStart[c + 1]; j++) b += L.res[g.members[j]];
            C.coeff.b[c] = b;
            C.x[c] = 0.0;
        }
        vCycle(l + 1);
        // buildProlongationKernel, multigrid.cu:415-427
#pragma omp parallel for schedule(static)
        for (int n = 0; n < g.nCells; n++) L.x[n] += C.x[g.cellToCoarse[n]];
        smooth(cfg.post);
    }

    // enqueueRun, multigrid.cu:766-811
    void run(const Coeff& A, std::vector<double>& x) {
        MGLevel& L0 = levels[0];
        L0.x = x;
        L0.coeff.AC = A.AC;
        L0.coeff.b = A.b;
        L0.coeff.invAC = A.invAC;
        L0.coeff.AF = A.AF;
        for (int l = 1; l < (int)levels.size(); l++) buildCoarseOperator(levels[l - 1], levels[l]);
        const int nCycles = cfg.maxIter > 1 ? cfg.maxIter : 1;
        for (int c = 0; c < nCycles; c++) vCycle(0);
        x = L0.x;
    }
};

// ---------------------------------------------------------------------------
// Face helpers (solver/header/solver_util.cuh:56-865)
// ---------------------------------------------------------------------------
struct Fields {
    std::vector<double> DU, DV, p, pp, u, v;
    // scalar transport: temperature and concentration, old levels, gradients
    std::vector<double> temp, conc, tempOld, concOld, tempOld2, concOld2, gradTZ, gradTR, gradCZ, gradCR;
    std::vector<double> uOld, vOld, uOld2, vOld2;
    std::vector<double> gradPZ, gradPR, gradUZ, gradUR, gradVZ, gradVR;
    std::vector<double> mDot;
    double alphaU = 0.7, alphaP = 0.3;
};

inline void outwardNormal(const Mesh& m, int c, int f, double& nz, double& nr) {  // :56-71
    nz = m.nZ[f];
    nr = m.nR[f];
    if (m.neighbor[f] == c) {
        nz = -nz;
        nr = -nr;
    }
}
inline double faceWeight(const Mesh& m, int c, int f) {  // :79-88
    const double w = m.wP[f];
    return (m.owner[f] == c) ? w : 1.0 - w;
}
inline double distCellToFace(const Mesh& m, int c, int f, double nz, double nr) {  // :90-105
    return std::fabs((m.fZ[f] - m.cZ[c]) * nz + (m.fR[f] - m.cR[c]) * nr);
}
inline double normalCorr(int c, int f, const Mesh& m, const Fields& s) {  // :107-124
    const double nz = m.nZ[f], nr = m.nR[f];
    return s.DU[c] * nz * nz + s.DV[c] * nr * nr;
}
inline double normalCorrToFace(int c, int f, const Mesh& m, const Fields& s) {  // :126-161
    const int ow = m.owner[f], nb = m.neighbor[f];
    const double DP = normalCorr(c, f, m, s);
    if (nb < 0) return DP;
    const int o = (ow == c) ? nb : ow;
    const double DN = normalCorr(o, f, m, s);
    const double w = faceWeight(m, c, f);
    return w * DP + (1.0 - w) * DN;
}
inline double faceAlong(const Mesh& m, int f) {  // :193-206
    const double nz = m.nZ[f], nr = m.nR[f];
    return m.fR[f] * nz * nz + m.fZ[f] * nr * nr;
}
inline double prescribed(const Mesh& m, int f, uint8_t t, double v, double L) {  // :223-238
    if (t != AXS_BC_FULLY_DEVELOPED) return v;
    const double x = faceAlong(m, f) / L;
    return v * (1.0 - x * x);
}
inline double interpToFace(int c, int f, const Mesh& m, const BC& bc,
                           const std::vector<double>& phi) {  // :240-297
    const int ow = m.owner[f], nb = m.neighbor[f];
    const double phiP = phi[c];
    if (nb >= 0) {
        const int o = (ow == c) ? nb : ow;
        const double w = faceWeight(m, c, f);
        return w * phiP + (1.0 - w) * phi[o];
    }
    const int g = m.group[f];
    if (g < 0 || g >= bc.nGroups) return phiP;
    const uint8_t t = bc.type[g];
    const double v = bc.value[g];
    if (t == AXS_BC_DIRICHLET || t == AXS_BC_FULLY_DEVELOPED) return prescribed(m, f, t, v, bc.length[g]);
    if (t == AXS_BC_NEUMANN) return phiP + v * m.dPB[f];
    return phiP;
}
inline double nonOrthoFlux(int c, int f, const Mesh& m, const std::vector<double>& gz,
                           const std::vector<double>& gr, double gamma) {  // :383-430
    const int ow = m.owner[f], nb = m.neighbor[f];
    if (nb < 0) return 0.0;
    const int o = (ow == c) ? nb : ow;
    double nz, nr;
    outwardNormal(m, c, f, nz, nr);
    const double area = m.area[f];
    const double dz = m.cZ[o] - m.cZ[c];
    const double dr = m.cR[o] - m.cR[c];
    const double invD = m.invDPN[f];
    const double aOverNd = area * invD;
    const double Tz = area * nz - aOverNd * dz;
    const double Tr = area * nr - aOverNd * dr;
    const double w = faceWeight(m, c, f);
    const double gfz = w * gz[c] + (1.0 - w) * gz[o];
    const double gfr = w * gr[c] + (1.0 - w) * gr[o];
    return gamma * (Tz * gfz + Tr * gfr);
}

// phiGradientGreenGauss, solver_util.cuh:458-532
void gradGG(int c, const Mesh& m, const BC& bc, const std::vector<double>& phi, double& gZ,
            double& gR) {
    double gz = 0.0, gr = 0.0, clz = 0.0, clr = 0.0;
    bool axis = false;
    double axisPhi = 0.0;
    for (int k = m.faceStart[c]; k < m.faceStart[c + 1]; k++) {
        const int f = m.faceIDs[k];
        double nz = 0.0, nr = 0.0;
        outwardNormal(m, c, f, nz, nr);
        const double phiF = interpToFace(c, f, m, bc, phi);
        const double rf = m.fR[f];
        if (rf > 1.0e-30) {
            const double L = m.length2D[f];
            gz += phiF * nz * L;
            gr += phiF * nr * L;
            clz += nz * L;
            clr += nr * L;
        } else {
            axis = true;
            axisPhi = phiF;
        }
    }
    if (axis) {
        gz += axisPhi * (-clz);
        gr += axisPhi * (-clr);
    }
    const double inv = m.invA2D[c];
    gZ = gz * inv;
    gR = gr * inv;
}

// phiGradientLeastSquare, solver_util.cuh:534-588
void gradLSQ(int c, const Mesh& m, const BC& bc, const std::vector<double>& phi, double& gZ,
             double& gR) {
    const double phiP = phi[c];
    double bz = 0.0, br = 0.0;
    for (int k = m.faceStart[c]; k < m.faceStart[c + 1]; k++) {
        const int f = m.faceIDs[k];
        const int ow = m.owner[f], nb = m.neighbor[f];
        double dphi;
        if (nb >= 0) {
            const bool isOwner = (ow == c);
            const double phiNb = phi[isOwner ? nb : ow];
            dphi = isOwner ? (phiNb - phiP) : (phiP - phiNb);
        } else {
            dphi = interpToFace(c, f, m, bc, phi) - phiP;
        }
        bz += m.lsqWZ[f] * dphi;
        br += m.lsqWR[f] * dphi;
    }
    const double izr = m.lsqInvZR[c];
    gZ = m.lsqInvZZ[c] * bz + izr * br;
    gR = izr * bz + m.lsqInvRR[c] * br;
}

// computeGradient, solver_util.cu:146-169
void computeGradient(const Mesh& m, const BC& bc, const std::vector<double>& phi,
                     std::vector<double>& gZ, std::vector<double>& gR, int scheme) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        gZ[n] = 0.0;
        gR[n] = 0.0;
        if (scheme == AXS_GRAD_GREEN_GAUSS) gradGG(n, m, bc, phi, gZ[n], gR[n]);
        else if (scheme == AXS_GRAD_LSQ) gradLSQ(n, m, bc, phi, gZ[n], gR[n]);
    }
}

// higherOrderFaceValue, solver_util.cuh:724-768
double higherOrder(int n, int nb, int f, const Mesh& m, const BC& bc, const std::vector<double>& phi,
                   const std::vector<double>& gz, const std::vector<double>& gr, int scheme,
                   double F, double phiUD) {
    if (scheme == AXS_CONV_CENTRAL) return interpToFace(n, f, m, bc, phi);
    if (scheme == AXS_CONV_SOU) {
        const int up = (F > 0.0) ? n : nb;
        const double dz = m.fZ[f] - m.cZ[up];
        const double dr = m.fR[f] - m.cR[up];
        const double ho = phi[up] + gz[up] * dz + gr[up] * dr;
        const double lo = std::fmin(phi[n], phi[nb]);
        const double hi = std::fmax(phi[n], phi[nb]);
        return std::fmin(std::fmax(ho, lo), hi);
    }
    return phiUD;
}

// ---- wall kinetics, solver/source/concentration_equation.cu:6-160 ----
// Rate laws of group g at concentration c and their derivatives: Michaelis-
// Menten Vmax c / (Km + c), Hill Vmax c^n / (Km^n + c^n), and the optional
// substrate-inhibition factor 1 - V2 c^m / (K2^m + c^m) (V2 = 0: factor 1).
struct Kinetics {
    double vmax, km, n, m, v2, kmN, k2M;
    Kinetics(const BC& bc, int g)
        : vmax(bc.vmax[g]), km(bc.km[g]), n(bc.hillN[g]), m(bc.inhibM[g]), v2(bc.v2[g]), kmN(bc.kmN[g]),
          k2M(bc.k2M[g]) {}
    double mm(double c) const { return (vmax * c) / (km + c); }                        // :6-14
    double dmm(double c) const {                                                       // :16-26
        const double inv = 1.0 / (km + c);
        return vmax * km * inv * inv;
    }
    double inhib(double c) const {                                                     // :28-41
        if (v2 == 0.0) return 1.0;
        const double cm = std::pow(c, m);
        return (1.0 - (v2 * cm / (k2M + cm)));
    }
    double dinhib(double c) const {                                                    // :43-61
        if (v2 == 0.0) return 0.0;
        const double cm = std::pow(c, m), cm1 = std::pow(c, m - 1.0);
        return (-(m * v2 * k2M) * (cm1) / ((k2M + cm) * (k2M + cm)));
    }
    double hill(double c) const {                                                      // :63-74
        const double cn = std::pow(c, n);
        return (vmax * cn) / (kmN + cn);
    }
    double dhill(double c) const {                                                     // :76-88
        const double cn = std::pow(c, n), cn1 = std::pow(c, n - 1.0);
        return ((n * vmax * kmN) * (cn1) / ((kmN + cn) * (kmN + cn)));
    }
};

// wallConcentrationMichaelisMenten / wallConcentrationHill (:91-160): Newton
// on h (c_P - c_w) = J(c_w) from c_w = c_P, c_w kept >= 0, at most 100 steps,
// stopping once |step| < 1e-14 + 1e-10 |c_w|.
double wallConcentration(const Kinetics& k, bool hill, double cp, double h) {
    double cw = cp;
    for (int i = 0; i < 100; i++) {
        const double rate = hill ? k.hill(cw) : k.mm(cw);
        const double inh = k.inhib(cw);
        const double J = rate * inh;
        const double dJ = rate * k.dinhib(cw) + (hill ? k.dhill(cw) : k.dmm(cw)) * inh;
        const double F = h * (cp - cw) - J;
        const double dF = -h - dJ;
        const double step = F / dF;
        cw -= step;
        cw = std::fmax(0.0, cw);
        if (std::fabs(step) < 1e-14 + 1e-10 * std::fabs(cw)) break;
    }
    return cw;
}

// addTransportCoefficients, solver_util.cu:218-448 (+ addConvectionContribution,
// solver_util.cuh:770-865).  wall: the reactive faces' state (null when the
// BCs have no kinetics type).
void addTransport(const Mesh& m, const Fields& s, Coeff& co, const BC& bc,
                  const std::vector<double>& phi, const std::vector<double>& gz,
                  const std::vector<double>& gr, int nonOrtho, double gamma, int conv, int scheme,
                  double fluxScale, Wall* wall = nullptr) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        double netF = 0.0;
        for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
            const int f = m.faceIDs[k];
            const int ow = m.owner[f], nb = m.neighbor[f];
            const double area = m.area[f];
            double F = 0.0;
            if (conv) {
                const double Fo = s.mDot[f] * fluxScale;
                if (ow == n) F = Fo;
                else if (nb == n) F = -Fo;
                if (std::fabs(F) <= 1.0e-30) F = 0.0;
                netF += F;
            }
            if (nb >= 0) {
                const double K = gamma * area * m.invDPN[f];
                co.AC[n] += K;
                co.AF[k] -= K;
                if (nonOrtho) co.b[n] += nonOrthoFlux(n, f, m, gz, gr, gamma);
                if (F != 0.0) {
                    const int o = (ow == n) ? nb : ow;
                    co.AC[n] += std::fmax(F, 0.0);
                    co.AF[k] += std::fmin(F, 0.0);
                    if (scheme != AXS_CONV_UPWIND) {
                        const double phiUD = (F > 0.0) ? phi[n] : phi[o];
                        const double ho = higherOrder(n, o, f, m, bc, phi, gz, gr, scheme, F, phiUD);
                        co.b[n] -= F * (ho - phiUD);
                    }
                }
            } else {
                const int g = m.group[f];
                const double nz = m.nZ[f], nr = m.nR[f];
                if (g >= 0 && g < bc.nGroups) {
                    const uint8_t t = bc.type[g];
                    const double v = bc.value[g];
                    const double dPF = distCellToFace(m, n, f, nz, nr);
                    const double K = gamma * area / dPF;
                    if (t == AXS_BC_DIRICHLET) {
                        co.AC[n] += K;
                        co.b[n] += K * v;
                    } else if (t == AXS_BC_NEUMANN) {
                        co.b[n] += gamma * area * v;
                    } else if (t == AXS_BC_FULLY_DEVELOPED) {
                        co.AC[n] += K;
                        co.b[n] += K * prescribed(m, f, t, v, bc.length[g]);
                    } else if (t == AXS_BC_MICHAELIS_MENTEN || t == AXS_BC_HILL) {
                        // reactive wall: conductance h through the near-wall
                        // half cell and the wall layers, wall value by Newton
                        const double Rtot = (dPF / gamma) + bc.rtot[g];
                        const double h = 1 / Rtot;
                        const Kinetics kin(bc, g);
                        const bool hill = t == AXS_BC_HILL;
                        const double cw = wallConcentration(kin, hill, phi[n], h);
                        if (wall) {
                            wall->cw[f] = cw;
                            wall->ocr[f] = area * (hill ? kin.hill(cw) : kin.mm(cw)) * kin.inhib(cw);
                        }
                        co.AC[n] += area * h;
                        co.b[n] += area * h * cw;
                    }
                }
                if (F != 0.0) {
                    if (g < 0 || g >= bc.nGroups) {
                        co.AC[n] += F;
                    } else {
                        const uint8_t t = bc.type[g];
                        const double v = bc.value[g];
                        if (t == AXS_BC_DIRICHLET) {
                            if (F < 0.0) co.b[n] += -F * v;
                            else co.AC[n] += F;
                        } else if (t == AXS_BC_NEUMANN) {
                            co.AC[n] += F;
                        } else if (t == AXS_BC_FULLY_DEVELOPED) {
                            if (F < 0.0) co.b[n] += -F * prescribed(m, f, t, v, bc.length[g]);
                            else co.AC[n] += F;
                        }
                    }
                }
            }
        }
        if (conv) co.AC[n] -= netF;
    }
}

// addRadialMomentumCylindricalSource, solver_util.cu:454-470
void radialSource(const Mesh& m, Coeff& co, double scalar) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        const double r = m.cR[n];
        co.AC[n] += scalar * m.vol[n] / (r * r);
    }
}

// addTransientCoefficient, solver_util.cu:536-578
void transientTerm(const Mesh& m, Coeff& co, const std::vector<double>* old,
                   const std::vector<double>* old2, double cap, double dt) {
    if (!old || dt <= 0.0) return;
    for (int n = 0; n < m.N; n++) {
        const double cV = cap * m.vol[n];
        if (old2) {
            const double a = cV / (2.0 * dt);
            co.AC[n] += 3.0 * a;
            co.b[n] += a * (4.0 * (*old)[n] - (*old2)[n]);
            continue;
        }
        const double a = cV / dt;
        co.AC[n] += a;
        co.b[n] += a * (*old)[n];
    }
}

// underRelaxEquation, solver_util.cu:582-600
void underRelax(Coeff& co, const std::vector<double>& x, double alpha) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < co.N; n++) {
        const double AC0 = co.AC[n];
        if (std::fabs(AC0) < 1.0e-30) continue;
        co.AC[n] = AC0 / alpha;
        co.b[n] += ((1.0 - alpha) / alpha) * AC0 * x[n];
    }
}

// buildInverseDiagonal, solver_util.cu:602-612
void inverseDiagonal(Coeff& co) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < co.N; n++) {
        const double AC = co.AC[n];
        co.invAC[n] = (std::fabs(AC) < 1.0e-30) ? 0.0 : 1.0 / AC;
    }
}

// createPPCoeff, simple.cu:7-88
void createPPCoeff(double rho, const Mesh& m, Coeff& co, const Fields& s, const BC& pbc) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
            const int f = m.faceIDs[k];
            const int nb = m.neighbor[f];
            const double area = m.area[f];
            const double Df = normalCorrToFace(n, f, m, s);
            if (nb >= 0) {
                const double K = rho * area * Df * m.invDPN[f];
                co.AC[n] += K;
                co.AF[k] -= K;
            } else {
                const int g = m.group[f];
                if (g < 0 || g >= pbc.nGroups) continue;
                if (pbc.type[g] == AXS_BC_DIRICHLET) {
                    const double K = rho * area * Df / m.dPB[f];
                    co.AC[n] += K;
                }
            }
        }
    }
}

// rhieChowNormalVelocityToFace, solver_util.cuh:590-713
double rhieChow(int c, int f, const Mesh& m, const Fields& s, const BC& bu, const BC& bv,
                const BC& bp) {
    const int ow = m.owner[f], nb = m.neighbor[f];
    double nz = 0.0, nr = 0.0;
    outwardNormal(m, c, f, nz, nr);
    const double uF = interpToFace(c, f, m, bu, s.u);
    const double vF = interpToFace(c, f, m, bv, s.v);
    const double unLin = uF * nz + vF * nr;
    if (nb < 0) {
        const int g = m.group[f];
        if (g < 0 || g >= bp.nGroups) return unLin;
        if (bp.type[g] != AXS_BC_DIRICHLET) return unLin;
        const double dPB = m.dPB[f];
        const double DfB = normalCorrToFace(c, f, m, s);
        const double compact = (bp.value[g] - s.p[c]) / dPB;
        const double cellN = s.gradPZ[c] * nz + s.gradPR[c] * nr;
        return unLin - DfB * (compact - cellN);
    }
    const int o = (ow == c) ? nb : ow;
    const double invD = m.invDPN[f];
    const double gPN = (s.p[o] - s.p[c]) * invD;
    const double w = faceWeight(m, c, f);
    const double gzF = w * s.gradPZ[c] + (1.0 - w) * s.gradPZ[o];
    const double grF = w * s.gradPR[c] + (1.0 - w) * s.gradPR[o];
    const double gFN = gzF * nz + grF * nr;
    const double Df = normalCorrToFace(c, f, m, s);
    return unLin - Df * (gPN - gFN);
}

// computeFaceMassFluxRhieChow, solver_util.cu:171-203
void faceMassFlux(double rho, const Mesh& m, Fields& s, const BC& bu, const BC& bv, const BC& bp) {
#pragma omp parallel for schedule(static)
    for (int f = 0; f < m.F; f++) {
        const double area = m.area[f];
        if (area <= 1.0e-30) {
            s.mDot[f] = 0.0;
            continue;
        }
        const double un = rhieChow(m.owner[f], f, m, s, bu, bv, bp);
        s.mDot[f] = rho * un * area;
    }
}

// createPPRhs, simple.cu:90-142
void createPPRhs(double rho, const Mesh& m, Coeff& co, const Fields& s, int nonOrtho) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        co.b[n] = 0.0;
        double imb = 0.0, cross = 0.0;
        for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
            const int f = m.faceIDs[k];
            const double md = s.mDot[f];
            if (m.owner[f] == n) imb += md;
            else if (m.neighbor[f] == n) imb -= md;
            if (nonOrtho) {
                const double Df = normalCorrToFace(n, f, m, s);
                cross += nonOrthoFlux(n, f, m, s.gradPZ, s.gradPR, rho * Df);
            }
        }
        co.b[n] = -imb + cross;
    }
}

// createMomentumPressureRhs, simple.cu:144-164
void momentumPressureRhs(const Mesh& m, Coeff& uc, Coeff& vc, const Fields& s) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        const double V = m.vol[n];
        uc.b[n] += -V * s.gradPZ[n];
        vc.b[n] += -V * s.gradPR[n];
    }
}

// getCorrectionCoefficients, solver_util.cu:127-144
void correctionCoeffs(const Mesh& m, const Coeff& uc, const Coeff& vc, Fields& s) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        const double V = m.vol[n];
        s.DU[n] = V * uc.invAC[n];
        s.DV[n] = V * vc.invAC[n];
    }
}

// updateMomentumPressure, simple.cu:166-179
void updateMomentumPressure(const Mesh& m, Fields& s) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        s.u[n] -= s.DU[n] * s.gradPZ[n];
        s.v[n] -= s.DV[n] * s.gradPR[n];
        s.p[n] += s.alphaP * s.pp[n];
    }
}

// updateMassFlux, simple.cu:181-258
void updateMassFlux(double rho, const Mesh& m, Fields& s, const BC& pbc, int nonOrtho) {
#pragma omp parallel for schedule(static)
    for (int f = 0; f < m.F; f++) {
        const int ow = m.owner[f], nb = m.neighbor[f];
        const double area = m.area[f];
        if (area <= 1.0e-30) {
            s.mDot[f] = 0.0;
            continue;
        }
        const double nz = m.nZ[f], nr = m.nR[f];
        const double Df = normalCorrToFace(ow, f, m, s);
        if (nb >= 0) {
            const double invD = m.invDPN[f];
            const double ppP = s.pp[ow], ppN = s.pp[nb];
            s.mDot[f] -= rho * area * Df * (ppN - ppP) * invD;
            if (nonOrtho) {
                const double Df2 = normalCorrToFace(ow, f, m, s);
                s.mDot[f] -= nonOrthoFlux(ow, f, m, s.gradPZ, s.gradPR, rho * Df2);
            }
            continue;
        }
        const int g = m.group[f];
        if (g < 0 || g >= pbc.nGroups) continue;
        if (pbc.type[g] == AXS_BC_DIRICHLET) {
            const double dPB = distCellToFace(m, ow, f, nz, nr);
            const double ppB = 0.0, ppP = s.pp[ow];
            s.mDot[f] -= rho * area * Df * (ppB - ppP) / dPB;
        }
    }
}

// jacobi, linear_solver.cu:9-33
void jacobiSweep(const Coeff& co, const std::vector<double>& xOld, std::vector<double>& xNew) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < co.N; n++) {
        const double inv = co.invAC[n];
        double val = co.b[n];
        for (int k = co.faceStart[n]; k < co.faceStart[n + 1]; k++) {
            const int nb = co.faceNeighbor[k];
            if (nb >= 0) val -= co.AF[k] * xOld[nb];
        }
        xNew[n] = (inv != 0.0) ? val * inv : xOld[n];
    }
}

// gaussSeidelColorSweep, linear_solver.cu:42-72
void gsColorSweep(const Coeff& co, std::vector<double>& x, const std::vector<int>& order, int begin,
                  int count) {
    for (int t = 0; t < count; t++) {
        const int n = order[begin + t];
        const double inv = co.invAC[n];
        double val = co.b[n];
        for (int k = co.faceStart[n]; k < co.faceStart[n + 1]; k++) {
            const int nb = co.faceNeighbor[k];
            if (nb >= 0) val -= co.AF[k] * x[nb];
        }
        x[n] = (inv != 0.0) ? val * inv : x[n];
    }
}

// residualRaw (sign = false), residuals.cuh:17-48
void residualAbs(const Coeff& co, const std::vector<double>& x, int scaleType,
                 std::vector<double>& res, std::vector<double>& scale) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < co.N; n++) {
        double Ax = co.AC[n] * x[n];
        if (scaleType == AXS_SCALE_DIAGONAL) scale[n] = Ax;
        for (int k = co.faceStart[n]; k < co.faceStart[n + 1]; k++) {
            const int nb = co.faceNeighbor[k];
            if (nb >= 0) Ax += co.AF[k] * x[nb];
        }
        const double r = co.b[n] - Ax;
        res[n] = std::fabs(r);
    }
}

// continuityResidual, residuals.cu:5-35
void continuity(const Mesh& m, const Fields& s, std::vector<double>& res) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < m.N; n++) {
        double imb = 0.0;
        for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
            const int f = m.faceIDs[k];
            if (m.owner[f] == n) imb += s.mDot[f];
            else if (m.neighbor[f] == n) imb -= s.mDot[f];
        }
        res[n] = imb;
    }
}

struct ResCfg {
    axs_residual_config cfg{};
    std::vector<double> res, scale;
    double resVal = 0.0, scaleVal = 0.0;
};

}  // namespace

// ---------------------------------------------------------------------------
// residualAllHost for one residual, residuals.cu:37-100.  The sums run in a
// fixed order of chunks (the reference reduces with a 256-wide tree).
// ---------------------------------------------------------------------------

// Sums of term(i) over [0, N): sequential within chunks of 32768 terms, the
// chunks' partial sums (computed in parallel) added in chunk order, so the
// value is the same for every thread count.
template <class Term>
double orderedSum(int N, Term term) {
    constexpr int kChunk = 1 << 15;
    const int nChunks = (N + kChunk - 1) / kChunk;
    std::vector<double> part(nChunks > 0 ? nChunks : 1, 0.0);
#pragma omp parallel for schedule(static) if (nChunks > 1)
    for (int c = 0; c < nChunks; c++) {
        double s = 0.0;
        const int end = std::min(N, (c + 1) * kChunk);
        for (int i = c * kChunk; i < end; i++) s += term(i);
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < nChunks; c++) s += part[c];
    return s;
}

extern "C" void orc_residual_value(int isContinuity, int norm, int scaling, const double* res,
                                   const double* scale, int N, int scaleIteration,
                                   double* scaleVal, double* resVal) {
    auto sumAbs = [&](const double* a) { return orderedSum(N, [a](int i) { return std::fabs(a[i]); }); };
    if (isContinuity) {
        *resVal = sumAbs(res);
    } else {
        if (norm == AXS_NORM_L1) *resVal = sumAbs(res);
        else if (norm == AXS_NORM_L2) {
            *resVal = orderedSum(N, [res](int i) { return res[i] * res[i]; });
        } else {
            double mx = -INFINITY;
            for (int i = 0; i < N; i++) mx = std::fmax(mx, std::fabs(res[i]));
            *resVal = mx;
        }
        if (scaling == AXS_SCALE_NONE) *scaleVal = 1.0;
        else if (scaling == AXS_SCALE_N) *scaleVal = N;
        else if (scaling == AXS_SCALE_SQRT_N) *scaleVal = std::sqrt((double)N);
        else if (scaling == AXS_SCALE_DIAGONAL) *scaleVal = sumAbs(scale);
    }
    if (isContinuity && scaleIteration < 5) *scaleVal = std::max(*resVal, 0.0);
    if (*scaleVal == 0.0) return;
    if (norm == AXS_NORM_L2) *resVal = std::sqrt(*resVal);
    *resVal /= *scaleVal;
}

namespace {

// ---------------------------------------------------------------------------
// Solver::runSimple, solver/source/solver.cpp:753-1396 (u, v, p' only)
// ---------------------------------------------------------------------------
struct Oracle {
    Mesh m;
    axs_config cfg{};
    BC bcU, bcV, bcP, ppBC, bcT, bcC;
    Coeff uC, vC, ppC, tC, cC;
    Wall wall;
    Fields s;
    Multigrid mg;
    std::vector<int> colorStart, cellOrder;
    int nColors = 0;
    ResCfg rc[AXS_RES_COUNT];
    std::vector<double> hist[AXS_RES_COUNT];
    int currentIteration = 0;
    bool converged = false;
    bool stopRequested = false;
    int stopAtIteration = -1;   // test hook: request a stop before this iteration
    bool allocated = false;
    // Solver::TimeFrame / timeFrames / maxTimeFrames (solver/header/solver.h:96-108)
    struct Frame {
        double time = 0.0;
        std::vector<double> u, v, p, t, c;
    };
    std::vector<Frame> frames;

    // captureTimeFrame, solver.cpp:265-296 (the cap stops capture, not the solve)
    void captureFrame(double time) {
        const int cap = cfg.max_time_frames > 0 ? cfg.max_time_frames : 500;
        if ((int)frames.size() >= cap) return;
        frames.push_back(Frame{time, s.u, s.v, s.p, cfg.solve_energy ? s.temp : std::vector<double>{},
                               cfg.solve_concentration ? s.conc : std::vector<double>{}});
    }

    void allocCoeff(Coeff& c, const std::vector<int>& fs, const std::vector<int>& fn) {
        c.N = m.N;
        c.nRefs = (int)fn.size();
        c.AC.assign(m.N, 0.0);
        c.b.assign(m.N, 0.0);
        c.invAC.assign(m.N, 0.0);
        c.AF.assign(c.nRefs, 0.0);
        c.faceStart = fs;
        c.faceNeighbor = fn;
    }

    void allocate() {
        std::vector<int> fs, fn;
        buildCSR(m, fs, fn);
        allocCoeff(uC, fs, fn);
        allocCoeff(vC, fs, fn);
        allocCoeff(ppC, fs, fn);
        allocCoeff(tC, fs, fn);
        allocCoeff(cC, fs, fn);
        wall.cw.assign(m.F, 0.0);
        wall.ocr.assign(m.F, 0.0);
        const int N = m.N;
        for (auto* a : {&s.DU, &s.DV, &s.p, &s.pp, &s.u, &s.v, &s.uOld, &s.vOld, &s.uOld2, &s.vOld2,
                        &s.gradPZ, &s.gradPR, &s.gradUZ, &s.gradUR, &s.gradVZ, &s.gradVR, &s.temp, &s.conc,
                        &s.tempOld, &s.concOld, &s.tempOld2, &s.concOld2, &s.gradTZ, &s.gradTR, &s.gradCZ,
                        &s.gradCR})
            a->assign(N, 0.0);
        s.mDot.assign(m.F, 0.0);
        for (auto& r : rc) {
            r.res.assign(N, 0.0);
            r.scale.assign(N, 0.0);
            r.resVal = 0.0;
            r.scaleVal = 0.0;
        }
        nColors = 0;
        colorStart.clear();
        cellOrder.clear();
        if (cfg.linear_solver == AXS_LINEAR_GS_RB)
            nColors = buildColoring(N, fs.data(), fn.data(), colorStart, cellOrder);
        currentIteration = 0;
        for (auto& h : hist) h.clear();
        allocated = true;
    }

    void linearSolve(const Coeff& co, std::vector<double>& x) {  // linear_solver.cu:91-148
        const bool colored = nColors > 0;
        if (cfg.linear_solver == AXS_LINEAR_GS_RB && colored) {
            for (int it = 0; it < cfg.linear_max_iter; it++)
                for (int c = 0; c < nColors; c++) {
                    const int begin = colorStart[c], count = colorStart[c + 1] - begin;
                    if (count <= 0) continue;
                    gsColorSweep(co, x, cellOrder, begin, count);
                }
            return;
        }
        std::vector<double> xt(x.size());
        for (int it = 0; it < cfg.linear_max_iter; it++) {
            jacobiSweep(co, x, xt);
            x.swap(xt);
        }
    }

    // updatePulsatileBoundaryValues, solver.cpp:375-424 with
    // BoundaryCondition::valueAtTime, structs/boundary_struct.h:184-192
    static void updatePulsatile(BC& bc, double t) {
        for (int g = 0; g < bc.nGroups; g++) {
            if (!bc.pulsatile[g]) continue;
            bc.value[g] = bc.mean[g] * (1.0 + bc.amplitude[g] * std::sin(2.0 * PI * bc.frequency[g] * t));
        }
    }

    void residualCheck(int scaleIteration) {  // solver.cpp:1243-1273
        const int types[2] = {rc[AXS_RES_U].cfg.scaling, rc[AXS_RES_V].cfg.scaling};
        residualAbs(uC, s.u, types[0], rc[AXS_RES_U].res, rc[AXS_RES_U].scale);
        residualAbs(vC, s.v, types[1], rc[AXS_RES_V].res, rc[AXS_RES_V].scale);
        continuity(m, s, rc[AXS_RES_CONTINUITY].res);
        if (cfg.solve_energy)
            residualAbs(tC, s.temp, rc[AXS_RES_TEMPERATURE].cfg.scaling, rc[AXS_RES_TEMPERATURE].res,
                        rc[AXS_RES_TEMPERATURE].scale);
        if (cfg.solve_concentration)
            residualAbs(cC, s.conc, rc[AXS_RES_CONCENTRATION].cfg.scaling, rc[AXS_RES_CONCENTRATION].res,
                        rc[AXS_RES_CONCENTRATION].scale);
        for (int i = 0; i < AXS_RES_COUNT; i++) {
            ResCfg& r = rc[i];
            if (!r.cfg.enabled) continue;
            orc_residual_value(i == AXS_RES_CONTINUITY, r.cfg.norm, r.cfg.scaling, r.res.data(),
                               r.scale.data(), m.N, scaleIteration, &r.scaleVal, &r.resVal);
        }
        for (int i = 0; i < AXS_RES_COUNT; i++) hist[i].push_back(rc[i].resVal);
    }

    bool residualsConverged() const {  // solver.cpp:59-72
        bool any = false;
        for (const ResCfg& r : rc) {
            if (!r.cfg.enabled) continue;
            any = true;
            if (!(r.resVal <= r.cfg.tol)) return false;
        }
        return any;
    }

    void outerIteration(bool transient, const std::vector<double>* uOld2,
                        const std::vector<double>* vOld2) {  // solver.cpp:1116-1217
        const int nonOrtho = cfg.use_non_orth ? 1 : 0;
        const bool needGrad = cfg.add_convection &&
                              (cfg.convection_scheme == AXS_CONV_SOU || cfg.convection_scheme == AXS_CONV_QUICK);
        // QUICK is clamped to upwind on load (file_manager.cpp:273-275) and
        // falls to upwind inside the kernel (solver_util.cuh:765-766).
        const int scheme = cfg.convection_scheme == AXS_CONV_QUICK ? AXS_CONV_UPWIND : cfg.convection_scheme;
        for (Coeff* c : {&uC, &vC, &ppC, &tC, &cC}) {  // clearAllCoefficients, solver_util.cuh:976-997
            std::fill(c->AC.begin(), c->AC.end(), 0.0);
            std::fill(c->b.begin(), c->b.end(), 0.0);
            std::fill(c->AF.begin(), c->AF.end(), 0.0);
        }
        if (nonOrtho || needGrad) {
            computeGradient(m, bcU, s.u, s.gradUZ, s.gradUR, cfg.gradient_scheme);
            computeGradient(m, bcV, s.v, s.gradVZ, s.gradVR, cfg.gradient_scheme);
        }
        addTransport(m, s, uC, bcU, s.u, s.gradUZ, s.gradUR, nonOrtho, cfg.mu, cfg.add_convection ? 1 : 0,
                     scheme, 1.0);
        addTransport(m, s, vC, bcV, s.v, s.gradVZ, s.gradVR, nonOrtho, cfg.mu, cfg.add_convection ? 1 : 0,
                     scheme, 1.0);
        radialSource(m, vC, cfg.mu);
        if (transient) {
            transientTerm(m, uC, &s.uOld, uOld2, cfg.rho, cfg.dt);
            transientTerm(m, vC, &s.vOld, vOld2, cfg.rho, cfg.dt);
        }
        computeGradient(m, bcP, s.p, s.gradPZ, s.gradPR, cfg.gradient_scheme);
        momentumPressureRhs(m, uC, vC, s);
        underRelax(uC, s.u, s.alphaU);
        underRelax(vC, s.v, s.alphaU);
        inverseDiagonal(uC);
        inverseDiagonal(vC);
        correctionCoeffs(m, uC, vC, s);
        linearSolve(uC, s.u);
        linearSolve(vC, s.v);
        createPPCoeff(cfg.rho, m, ppC, s, bcP);
        inverseDiagonal(ppC);
        faceMassFlux(cfg.rho, m, s, bcU, bcV, bcP);
        std::fill(s.pp.begin(), s.pp.end(), 0.0);
        for (int corr = 0; corr <= nonOrtho; corr++) {
            const int cross = corr > 0 ? 1 : 0;
            if (cross) computeGradient(m, ppBC, s.pp, s.gradPZ, s.gradPR, cfg.gradient_scheme);
            createPPRhs(cfg.rho, m, ppC, s, cross);
            if (cfg.use_multigrid) mg.run(ppC, s.pp);
            else linearSolve(ppC, s.pp);
        }
        computeGradient(m, ppBC, s.pp, s.gradPZ, s.gradPR, cfg.gradient_scheme);
        updateMomentumPressure(m, s);
        updateMassFlux(cfg.rho, m, s, bcP, nonOrtho);
        // energy and concentration (solver.cpp:976-1009, 1222-1237): the
        // momentum assembly with diffusivity k / (rho cp) or D, the volumetric
        // flux, capacity 1; temperature under-relaxes with alpha_u, species not
        for (int sc = 0; sc < 2; sc++) {
            if (sc == 0 ? !cfg.solve_energy : !cfg.solve_concentration) continue;
            Coeff& co = sc == 0 ? tC : cC;
            const BC& bc = sc == 0 ? bcT : bcC;
            std::vector<double>& phi = sc == 0 ? s.temp : s.conc;
            std::vector<double>& gz = sc == 0 ? s.gradTZ : s.gradCZ;
            std::vector<double>& gr = sc == 0 ? s.gradTR : s.gradCR;
            const double gamma = sc == 0 ? cfg.k / (cfg.rho * cfg.cp) : cfg.diffusivity;
            if (nonOrtho || needGrad) computeGradient(m, bc, phi, gz, gr, cfg.gradient_scheme);
            addTransport(m, s, co, bc, phi, gz, gr, nonOrtho, gamma, cfg.add_convection ? 1 : 0, scheme,
                         1.0 / cfg.rho, &wall);
            if (transient)
                transientTerm(m, co, sc == 0 ? &s.tempOld : &s.concOld,
                              uOld2 ? (sc == 0 ? &s.tempOld2 : &s.concOld2) : nullptr, 1.0, cfg.dt);
            underRelax(co, phi, sc == 0 ? s.alphaU : 1.0);
            inverseDiagonal(co);
            linearSolve(co, phi);
        }
    }

    int run(int continueRun) {
        const bool canCont = continueRun && allocated;
        if (!canCont) allocate();
        if (cfg.max_iter < 1) cfg.max_iter = 50;  // runCheck, solver.cpp:537-547
        if (cfg.check_conv < 1) cfg.check_conv = 1;
        if (cfg.linear_max_iter < 1) cfg.linear_max_iter = 20;
        if (cfg.structured_mesh) cfg.use_non_orth = 0;
        if (cfg.save_key_frame_iter < 1) cfg.save_key_frame_iter = 1;  // solver.cpp:557-559
        if (!(cfg.rho > 0.0) || !(cfg.mu > 0.0)) return AXS_ERR_INVALID;
        if (cfg.transient && (!(cfg.dt > 0.0) || !(cfg.t_end > 0.0))) return AXS_ERR_INVALID;
        s.alphaU = cfg.alpha_u;
        s.alphaP = cfg.alpha_p;
        ppBC = bcP;  // solver.cpp:788-795
        std::fill(ppBC.value.begin(), ppBC.value.end(), 0.0);
        if (cfg.use_multigrid) {
            std::vector<int> fs, fn;
            buildCSR(m, fs, fn);
            mg.init(fs, fn, m.N, cfg.mg_max_levels);
            mg.cfg.maxIter = cfg.mg_max_iter;
            mg.cfg.weight = cfg.mg_weight;
            mg.cfg.linearSweep = cfg.mg_coarse_sweeps;
            mg.cfg.pre = cfg.mg_pre_sweeps;
            mg.cfg.post = cfg.mg_post_sweeps;
            mg.cfg.smoother = cfg.mg_smoother;
        }
        const bool transient = cfg.transient != 0;
        const bool bdf2 = transient && cfg.time_scheme == AXS_TIME_BDF2;
        const int numSteps = transient ? (int)std::ceil(cfg.t_end / cfg.dt) : 1;
        bcU.value = bcU.mean;  // valueAtTime(0.0), memory_manger.cpp:102-103
        bcV.value = bcV.mean;
        converged = false;
        // frames are kept only by a continued transient run; the continued run
        // carries the time (and the pulsatile phase) on from the last frame
        // (solver.cpp:890-900); the first frame is the state the run starts from
        // (solver.cpp:1028-1031)
        if (!transient || !canCont) frames.clear();
        const double timeOffset = frames.empty() ? 0.0 : frames.back().time;
        if (transient && frames.empty()) captureFrame(timeOffset);
        for (int t = 0; t < numSteps; t++) {
            if (transient) {
                const double stepTime = timeOffset + (double)(t + 1) * cfg.dt;
                updatePulsatile(bcU, stepTime);
                updatePulsatile(bcV, stepTime);
                if (bdf2) {
                    s.uOld2.swap(s.uOld);
                    s.vOld2.swap(s.vOld);
                }
                s.uOld = s.u;
                if (bdf2) {
                    s.tempOld2.swap(s.tempOld);
                    s.concOld2.swap(s.concOld);
                }
                s.tempOld = s.temp;
                s.concOld = s.conc;
                s.vOld = s.v;
            }
            const bool bdf2Step = bdf2 && t > 0;
            converged = false;
            for (int k = 0; k < cfg.max_iter; k++) {
                if (stopAtIteration >= 0 && currentIteration >= stopAtIteration) stopRequested = true;
                if (stopRequested) break;
                outerIteration(transient, bdf2Step ? &s.uOld2 : nullptr, bdf2Step ? &s.vOld2 : nullptr);
                if (k % cfg.check_conv == 0) {
                    residualCheck(transient ? k : currentIteration);
                    converged = residualsConverged();
                }
                currentIteration++;
                if (converged) break;
            }
            // keyframe interval and the last (or stopped) step, solver.cpp:1286-1294
            const bool lastStep = (t == numSteps - 1) || stopRequested;
            if (transient && (t % cfg.save_key_frame_iter == 0 || lastStep))
                captureFrame(timeOffset + (double)(t + 1) * cfg.dt);
            if (stopRequested) break;
        }
        // output gradients, solver.cpp:1330-1336
        computeGradient(m, bcU, s.u, s.gradUZ, s.gradUR, cfg.gradient_scheme);
        computeGradient(m, bcV, s.v, s.gradVZ, s.gradVR, cfg.gradient_scheme);
        computeGradient(m, bcP, s.p, s.gradPZ, s.gradPR, cfg.gradient_scheme);
        if (cfg.solve_energy) computeGradient(m, bcT, s.temp, s.gradTZ, s.gradTR, cfg.gradient_scheme);
        if (cfg.solve_concentration) computeGradient(m, bcC, s.conc, s.gradCZ, s.gradCR, cfg.gradient_scheme);
        return AXS_OK;
    }
};

void loadBC(BC& bc, const axs_bc_view* v) {
    bc = BC{};
    if (!v) return;
    bc.nGroups = v->n_groups;
    bc.type.assign(bc.nGroups, AXS_BC_NONE);
    bc.value.assign(bc.nGroups, 0.0);
    bc.length.assign(bc.nGroups, 0.0);
    bc.amplitude.assign(bc.nGroups, 0.0);
    bc.frequency.assign(bc.nGroups, 0.0);
    bc.pulsatile.assign(bc.nGroups, 0);
    for (int g = 0; g < bc.nGroups; g++) {
        const uint8_t t = v->type[g];
        bc.pulsatile[g] = (t == AXS_BC_PULSATILE);
        bc.type[g] = (t == AXS_BC_PULSATILE) ? (uint8_t)AXS_BC_DIRICHLET : t;  // memory_manger.cpp:96-100
        bc.value[g] = v->value ? v->value[g] : 0.0;
        bc.length[g] = v->length ? v->length[g] : 0.0;
        if (v->amplitude) bc.amplitude[g] = v->amplitude[g];
        if (v->frequency) bc.frequency[g] = v->frequency[g];
    }
    bc.mean = bc.value;
    auto col = [&](const double* src, double dflt) {
        std::vector<double> a(bc.nGroups, dflt);
        if (src)
            for (int g = 0; g < bc.nGroups; g++) a[g] = src[g];
        return a;
    };
    bc.vmax = col(v->vmax, 0.0);
    bc.km = col(v->km, 0.0);
    bc.hillN = col(v->hill_n, 1.0);
    bc.inhibM = col(v->inhib_m, 1.0);
    bc.k2 = col(v->inhib_k2, 0.0);
    bc.v2 = col(v->inhib_v2, 0.0);
    bc.rtot = col(v->wall_resistance, 0.0);
    bc.kmN.assign(bc.nGroups, 0.0);
    bc.k2M.assign(bc.nGroups, 0.0);
    for (int g = 0; g < bc.nGroups; g++) {  // memory_manger.cpp:175-184
        bc.kmN[g] = std::pow(bc.km[g], bc.hillN[g]);
        bc.k2M[g] = std::pow(bc.k2[g], bc.inhibM[g]);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface used by the tests
// ---------------------------------------------------------------------------
struct orc_solver {
    Oracle o;
};

extern "C" {

orc_solver* orc_create(const axs_mesh_view* v, const axs_config* cfg) {
    if (!v || !cfg || v->n_cells <= 0 || v->n_faces <= 0) return nullptr;
    orc_solver* h = new orc_solver();
    Mesh& m = h->o.m;
    m.N = v->n_cells;
    m.F = v->n_faces;
    const int S = v->cell_face_start[m.N];
    m.cZ.assign(v->cell_center_z, v->cell_center_z + m.N);
    m.cR.assign(v->cell_center_r, v->cell_center_r + m.N);
    m.vol.assign(v->cell_volume, v->cell_volume + m.N);
    m.area2D.assign(v->cell_area2d, v->cell_area2d + m.N);
    m.faceStart.assign(v->cell_face_start, v->cell_face_start + m.N + 1);
    m.faceIDs.assign(v->cell_face_ids, v->cell_face_ids + S);
    m.owner.assign(v->face_owner, v->face_owner + m.F);
    m.neighbor.assign(v->face_neighbor, v->face_neighbor + m.F);
    m.group.assign(v->face_group, v->face_group + m.F);
    m.nZ.assign(v->face_normal_z, v->face_normal_z + m.F);
    m.nR.assign(v->face_normal_r, v->face_normal_r + m.F);
    m.fZ.assign(v->face_center_z, v->face_center_z + m.F);
    m.fR.assign(v->face_center_r, v->face_center_r + m.F);
    m.area.assign(v->face_area, v->face_area + m.F);
    preCompute(m);
    h->o.cfg = *cfg;
    for (int i = 0; i < AXS_RES_COUNT; i++) h->o.rc[i].cfg = cfg->residual[i];
    return h;
}

void orc_destroy(orc_solver* h) { delete h; }
// threads of the OpenMP loops (n >= 1); the results do not depend on it
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads() { return omp_get_max_threads(); }

int orc_set_bcs(orc_solver* h, int var, const axs_bc_view* v) {
    if (var == AXS_VAR_U) loadBC(h->o.bcU, v);
    else if (var == AXS_VAR_V) loadBC(h->o.bcV, v);
    else if (var == AXS_VAR_P) loadBC(h->o.bcP, v);
    else if (var == AXS_VAR_T) loadBC(h->o.bcT, v);
    else if (var == AXS_VAR_C) loadBC(h->o.bcC, v);
    else return AXS_ERR_INVALID;
    return AXS_OK;
}

int orc_run(orc_solver* h, int continue_run) { return h->o.run(continue_run); }

// Replace the configuration between runs (the GUI writes Solver's config
// members directly between runs; a continued run picks them up).
void orc_configure(orc_solver* h, const axs_config* cfg) {
    h->o.cfg = *cfg;
    for (int i = 0; i < AXS_RES_COUNT; i++) h->o.rc[i].cfg = cfg->residual[i];
}
void orc_set_stop_at(orc_solver* h, int iteration) {
    h->o.stopAtIteration = iteration;
    h->o.stopRequested = false;
}
int orc_iterations(const orc_solver* h) { return h->o.currentIteration; }
int orc_converged(const orc_solver* h) { return h->o.converged ? 1 : 0; }

int orc_get_residual_history(const orc_solver* h, int res, double* out, int max_n) {
    if (res < 0 || res >= AXS_RES_COUNT) return -1;
    const auto& v = h->o.hist[res];
    const int n = std::min<int>(max_n, (int)v.size());
    for (int i = 0; i < n; i++) out[i] = v[i];
    return (int)v.size();
}

int orc_get_field(const orc_solver* h, int field, double* out, size_t n) {
    const Oracle& o = h->o;
    const Mesh& m = o.m;
    const std::vector<double>* src = nullptr;
    std::vector<double> tmp;
    switch (field) {
    case AXS_FIELD_U: src = &o.s.u; break;
    case AXS_FIELD_V: src = &o.s.v; break;
    case AXS_FIELD_P: src = &o.s.p; break;
    case AXS_FIELD_PP: src = &o.s.pp; break;
    case AXS_FIELD_GRAD_UZ: src = &o.s.gradUZ; break;
    case AXS_FIELD_GRAD_UR: src = &o.s.gradUR; break;
    case AXS_FIELD_GRAD_VZ: src = &o.s.gradVZ; break;
    case AXS_FIELD_GRAD_VR: src = &o.s.gradVR; break;
    case AXS_FIELD_GRAD_PZ: src = &o.s.gradPZ; break;
    case AXS_FIELD_GRAD_PR: src = &o.s.gradPR; break;
    case AXS_FIELD_MDOT: src = &o.s.mDot; break;
    case AXS_FIELD_DU: src = &o.s.DU; break;
    case AXS_FIELD_DV: src = &o.s.DV; break;
    case AXS_FIELD_CONTINUITY:  // getMassImbalance, solver.cpp:298-331
        tmp.assign(m.N, 0.0);
        continuity(m, o.s, tmp);
        src = &tmp;
        break;
    case AXS_FIELD_VMAG:
    case AXS_FIELD_CELL_RE:  // createDerivedVelocitySolutions, solver.cpp:221-263
        tmp.assign(m.N, 0.0);
        for (int c = 0; c < m.N; c++) {
            const double sp = std::sqrt(o.s.u[c] * o.s.u[c] + o.s.v[c] * o.s.v[c]);
            if (field == AXS_FIELD_VMAG) tmp[c] = sp;
            else {
                const double hh = std::sqrt(std::max(m.area2D[c], 0.0));
                tmp[c] = (o.cfg.mu > 0.0) ? o.cfg.rho * sp * hh / o.cfg.mu : 0.0;
            }
        }
        src = &tmp;
        break;
    case AXS_FIELD_U_AC: src = &o.uC.AC; break;
    case AXS_FIELD_U_B: src = &o.uC.b; break;
    case AXS_FIELD_U_INVAC: src = &o.uC.invAC; break;
    case AXS_FIELD_U_AF: src = &o.uC.AF; break;
    case AXS_FIELD_V_AC: src = &o.vC.AC; break;
    case AXS_FIELD_V_B: src = &o.vC.b; break;
    case AXS_FIELD_V_INVAC: src = &o.vC.invAC; break;
    case AXS_FIELD_V_AF: src = &o.vC.AF; break;
    case AXS_FIELD_PP_AC: src = &o.ppC.AC; break;
    case AXS_FIELD_PP_B: src = &o.ppC.b; break;
    case AXS_FIELD_PP_INVAC: src = &o.ppC.invAC; break;
    case AXS_FIELD_PP_AF: src = &o.ppC.AF; break;
    case AXS_FIELD_U_OLD: src = &o.s.uOld; break;
    case AXS_FIELD_V_OLD: src = &o.s.vOld; break;
    case AXS_FIELD_U_OLD2: src = &o.s.uOld2; break;
    case AXS_FIELD_V_OLD2: src = &o.s.vOld2; break;
    case AXS_FIELD_T: src = &o.s.temp; break;
    case AXS_FIELD_C: src = &o.s.conc; break;
    case AXS_FIELD_GRAD_TZ: src = &o.s.gradTZ; break;
    case AXS_FIELD_GRAD_TR: src = &o.s.gradTR; break;
    case AXS_FIELD_GRAD_CZ: src = &o.s.gradCZ; break;
    case AXS_FIELD_GRAD_CR: src = &o.s.gradCR; break;
    case AXS_FIELD_T_AC: src = &o.tC.AC; break;
    case AXS_FIELD_T_B: src = &o.tC.b; break;
    case AXS_FIELD_T_INVAC: src = &o.tC.invAC; break;
    case AXS_FIELD_T_AF: src = &o.tC.AF; break;
    case AXS_FIELD_C_AC: src = &o.cC.AC; break;
    case AXS_FIELD_C_B: src = &o.cC.b; break;
    case AXS_FIELD_C_INVAC: src = &o.cC.invAC; break;
    case AXS_FIELD_C_AF: src = &o.cC.AF; break;
    case AXS_FIELD_T_OLD: src = &o.s.tempOld; break;
    case AXS_FIELD_C_OLD: src = &o.s.concOld; break;
    case AXS_FIELD_T_OLD2: src = &o.s.tempOld2; break;
    case AXS_FIELD_C_OLD2: src = &o.s.concOld2; break;
    case AXS_FIELD_WALL_CW: src = &o.wall.cw; break;
    case AXS_FIELD_WALL_OCR: src = &o.wall.ocr; break;
    case AXS_GEOM_INV_DPN: src = &m.invDPN; break;
    case AXS_GEOM_WP: src = &m.wP; break;
    case AXS_GEOM_DPB: src = &m.dPB; break;
    case AXS_GEOM_INV_A2D: src = &m.invA2D; break;
    case AXS_GEOM_LSQ_WZ: src = &m.lsqWZ; break;
    case AXS_GEOM_LSQ_WR: src = &m.lsqWR; break;
    case AXS_GEOM_LSQ_INV_ZZ: src = &m.lsqInvZZ; break;
    case AXS_GEOM_LSQ_INV_ZR: src = &m.lsqInvZR; break;
    case AXS_GEOM_LSQ_INV_RR: src = &m.lsqInvRR; break;
    case AXS_GEOM_LENGTH2D: src = &m.length2D; break;
    default: return AXS_ERR_INVALID;
    }
    if (src->size() != n) return AXS_ERR_RANGE;
    std::memcpy(out, src->data(), n * sizeof(double));
    return AXS_OK;
}

// ---- single kernels on caller data (closed-form and brute-force pins) ----

// computeGradient for an arbitrary field with the solver's BCs of `var`.
int orc_gradient(const orc_solver* h, int var, int scheme, const double* phi, double* gz, double* gr) {
    const Oracle& o = h->o;
    const BC& bc = var == AXS_VAR_U ? o.bcU : var == AXS_VAR_V ? o.bcV : o.bcP;
    std::vector<double> p(phi, phi + o.m.N), z(o.m.N), r(o.m.N);
    computeGradient(o.m, bc, p, z, r, scheme);
    std::memcpy(gz, z.data(), o.m.N * sizeof(double));
    std::memcpy(gr, r.data(), o.m.N * sizeof(double));
    return AXS_OK;
}

static const BC& bcOf(const Oracle& o, int var) {
    return var == AXS_VAR_U ? o.bcU : var == AXS_VAR_V ? o.bcV : var == AXS_VAR_T ? o.bcT : var == AXS_VAR_C ? o.bcC
                                                                                                       : o.bcP;
}
static std::vector<double> vec(const double* a, int n) {
    return a ? std::vector<double>(a, a + n) : std::vector<double>(n, 0.0);
}

// interpolateFieldToFace (solver_util.cuh:240-297) seen from each face's owner:
// out[f] = phi_f for every face, boundary faces from the BCs of `var`.
int orc_probe_face_values(const orc_solver* h, int var, const double* phi, double* out) {
    const Oracle& o = h->o;
    const std::vector<double> p = vec(phi, o.m.N);
    for (int f = 0; f < o.m.F; f++) out[f] = interpToFace(o.m.owner[f], f, o.m, bcOf(o, var), p);
    return AXS_OK;
}

// addTransportCoefficients (solver_util.cu:218-448) of one field into zeroed
// coefficients: AC[N], AF[S], b[N].
int orc_probe_transport(const orc_solver* h, int var, const double* phi, const double* gz, const double* gr,
                        const double* mdot, int nonOrtho, double gamma, int conv, int scheme, double fluxScale,
                        double* AC, double* AF, double* b) {
    const Oracle& o = h->o;
    const Mesh& m = o.m;
    Fields s;
    s.mDot = vec(mdot, m.F);
    Coeff co;
    co.N = m.N;
    co.AC.assign(m.N, 0.0);
    co.b.assign(m.N, 0.0);
    co.AF.assign(m.faceIDs.size(), 0.0);
    addTransport(m, s, co, bcOf(o, var), vec(phi, m.N), vec(gz, m.N), vec(gr, m.N), nonOrtho, gamma, conv, scheme,
                 fluxScale);
    std::memcpy(AC, co.AC.data(), m.N * sizeof(double));
    std::memcpy(b, co.b.data(), m.N * sizeof(double));
    std::memcpy(AF, co.AF.data(), co.AF.size() * sizeof(double));
    return AXS_OK;
}

// addTransientCoefficient (solver_util.cu:536-578) into zeroed AC, b; old2 NULL = BDF1.
int orc_probe_transient(const orc_solver* h, const double* old, const double* old2, double cap, double dt,
                        double* AC, double* b) {
    const Mesh& m = h->o.m;
    Coeff co;
    co.N = m.N;
    co.AC.assign(m.N, 0.0);
    co.b.assign(m.N, 0.0);
    const std::vector<double> o1 = vec(old, m.N), o2 = vec(old2, m.N);
    transientTerm(m, co, &o1, old2 ? &o2 : nullptr, cap, dt);
    std::memcpy(AC, co.AC.data(), m.N * sizeof(double));
    std::memcpy(b, co.b.data(), m.N * sizeof(double));
    return AXS_OK;
}

// Fields of the pressure-correction kernels from caller arrays.
static Fields probeFields(const Mesh& m, const double* u, const double* v, const double* p, const double* pp,
                          const double* gpz, const double* gpr, const double* DU, const double* DV,
                          const double* mdot) {
    Fields s;
    s.u = vec(u, m.N);
    s.v = vec(v, m.N);
    s.p = vec(p, m.N);
    s.pp = vec(pp, m.N);
    s.gradPZ = vec(gpz, m.N);
    s.gradPR = vec(gpr, m.N);
    s.DU = vec(DU, m.N);
    s.DV = vec(DV, m.N);
    s.mDot = vec(mdot, m.F);
    return s;
}

// computeFaceMassFluxRhieChow (solver_util.cu:171-203) with the solver's u, v, p BCs.
int orc_probe_face_flux(const orc_solver* h, double rho, const double* u, const double* v, const double* p,
                        const double* gpz, const double* gpr, const double* DU, const double* DV, double* mdot) {
    const Oracle& o = h->o;
    Fields s = probeFields(o.m, u, v, p, nullptr, gpz, gpr, DU, DV, nullptr);
    faceMassFlux(rho, o.m, s, o.bcU, o.bcV, o.bcP);
    std::memcpy(mdot, s.mDot.data(), o.m.F * sizeof(double));
    return AXS_OK;
}

// createPPCoeff (simple.cu:7-88) into zeroed AC[N], AF[S] with the solver's p BCs.
int orc_probe_pp_coeff(const orc_solver* h, double rho, const double* DU, const double* DV, double* AC,
                       double* AF) {
    const Oracle& o = h->o;
    const Mesh& m = o.m;
    Fields s = probeFields(m, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, DU, DV, nullptr);
    Coeff co;
    co.N = m.N;
    co.AC.assign(m.N, 0.0);
    co.b.assign(m.N, 0.0);
    co.AF.assign(m.faceIDs.size(), 0.0);
    createPPCoeff(rho, m, co, s, o.bcP);
    std::memcpy(AC, co.AC.data(), m.N * sizeof(double));
    std::memcpy(AF, co.AF.data(), co.AF.size() * sizeof(double));
    return AXS_OK;
}

// createPPRhs (simple.cu:90-142); gpz/gpr hold grad p' for the cross term.
int orc_probe_pp_rhs(const orc_solver* h, double rho, const double* mdot, const double* DU, const double* DV,
                     const double* gpz, const double* gpr, int cross, double* b) {
    const Oracle& o = h->o;
    const Mesh& m = o.m;
    Fields s = probeFields(m, nullptr, nullptr, nullptr, nullptr, gpz, gpr, DU, DV, mdot);
    Coeff co;
    co.N = m.N;
    co.b.assign(m.N, 0.0);
    createPPRhs(rho, m, co, s, cross);
    std::memcpy(b, co.b.data(), m.N * sizeof(double));
    return AXS_OK;
}

// updateMassFlux (simple.cu:181-258) of mdot in place, p BCs of the solver.
int orc_probe_mass_flux_update(const orc_solver* h, double rho, const double* pp, const double* DU,
                               const double* DV, const double* gpz, const double* gpr, int nonOrtho,
                               double* mdot) {
    const Oracle& o = h->o;
    const Mesh& m = o.m;
    Fields s = probeFields(m, nullptr, nullptr, nullptr, pp, gpz, gpr, DU, DV, mdot);
    updateMassFlux(rho, m, s, o.bcP, nonOrtho);
    std::memcpy(mdot, s.mDot.data(), m.F * sizeof(double));
    return AXS_OK;
}

// wallConsumptionDiagnostic (solver_util.cu:480-531) over the last run's
// concentration, summed in face order (out[10], as axs_get_wall_diagnostic).
int orc_wall_diagnostic(const orc_solver* h, double* out) {
    const Oracle& o = h->o;
    const Mesh& m = o.m;
    const BC& bc = o.bcC;
    for (int i = 0; i < 10; i++) out[i] = 0.0;
    if (!o.allocated) return AXS_ERR_STATE;
    for (int f = 0; f < m.F; f++) {
        if (m.neighbor[f] >= 0) continue;
        const int ow = m.owner[f], g = m.group[f];
        if (g < 0 || g >= bc.nGroups) continue;
        const uint8_t t = bc.type[g];
        const double F = o.s.mDot[f];
        if (t == AXS_BC_DIRICHLET && F < 0.0) out[1] += -F * bc.value[g];
        if (!(t == AXS_BC_MICHAELIS_MENTEN || t == AXS_BC_HILL)) continue;
        double nz, nr;
        outwardNormal(m, ow, f, nz, nr);
        const double dPF = distCellToFace(m, ow, f, nz, nr);
        const double hh = 1.0 / (dPF / o.cfg.diffusivity + bc.rtot[g]);
        const double area = m.area[f], cp = o.s.conc[ow], cw = o.wall.cw[f];
        out[0] += o.wall.ocr[f];
        out[2] += area * hh * cp;
        out[3] += 1.0;
        out[4] += cw;
        out[5] += dPF;
        out[6] += hh;
        out[7] += cp;
        out[8] += bc.vmax[g];
        out[9] += bc.km[g];
    }
    return AXS_OK;
}

// The wall Newton solve alone for given kinetics (hill: Hill law, else
// Michaelis-Menten), cell value cp and conductance h.
double orc_probe_wall_concentration(int hill, double vmax, double km, double n, double m, double k2, double v2,
                                    double cp, double h) {
    BC bc;
    bc.nGroups = 1;
    bc.vmax = {vmax};
    bc.km = {km};
    bc.hillN = {n};
    bc.inhibM = {m};
    bc.k2 = {k2};
    bc.v2 = {v2};
    bc.kmN = {std::pow(km, n)};
    bc.k2M = {std::pow(k2, m)};
    return wallConcentration(Kinetics(bc, 0), hill != 0, cp, h);
}

// Transient frames: count, and one frame's time and field (AXS_FIELD_U/V/P/T/C).
int orc_frame_count(const orc_solver* h) { return (int)h->o.frames.size(); }
int orc_get_frame(const orc_solver* h, int index, int field, double* time, double* out, size_t n) {
    const auto& fr = h->o.frames;
    if (index < 0 || index >= (int)fr.size()) return AXS_ERR_RANGE;
    const auto& f = fr[index];
    const std::vector<double>* src = field == AXS_FIELD_U ? &f.u : field == AXS_FIELD_V ? &f.v
                                   : field == AXS_FIELD_P ? &f.p : field == AXS_FIELD_T ? &f.t
                                   : field == AXS_FIELD_C ? &f.c : nullptr;
    if (time) *time = f.time;
    if (!out) return AXS_OK;
    if (!src) return AXS_ERR_INVALID;
    if (src->size() != n) return AXS_ERR_RANGE;
    std::memcpy(out, src->data(), n * sizeof(double));
    return AXS_OK;
}

// Value of group `group` of `var` as the last run left it on the device
// (updatePulsatileBoundaryValues at the last step's time, solver.cpp:1039-1056).
double orc_bc_current(const orc_solver* h, int var, int group) {
    const BC& bc = bcOf(h->o, var);
    if (group < 0 || group >= bc.nGroups) return NAN;
    return bc.value[group];
}

// BoundaryCondition::valueAtTime (structs/boundary_struct.h:184-192) of one
// group of `var` as the run loop prescribes it.
double orc_probe_bc_value(const orc_solver* h, int var, int group, double t) {
    BC bc = bcOf(h->o, var);
    if (group < 0 || group >= bc.nGroups) return NAN;
    Oracle::updatePulsatile(bc, t);
    return bc.value[group];
}

// Cell -> face CSR of the mesh (buildCellFaceCSR).
int orc_build_csr(const orc_solver* h, int* face_start, int* face_neighbor) {
    std::vector<int> fs, fn;
    buildCSR(h->o.m, fs, fn);
    std::memcpy(face_start, fs.data(), fs.size() * sizeof(int));
    std::memcpy(face_neighbor, fn.data(), fn.size() * sizeof(int));
    return AXS_OK;
}

int orc_build_coloring(int N, const int* fs, const int* fn, int* color_start, int* cell_order) {
    std::vector<int> cs, co;
    const int nc = buildColoring(N, fs, fn, cs, co);
    std::memcpy(color_start, cs.data(), cs.size() * sizeof(int));
    std::memcpy(cell_order, co.data(), co.size() * sizeof(int));
    return nc;
}

// ---- multigrid ----
struct orc_mg {
    Multigrid mg;
};

orc_mg* orc_mg_create(int N, const int* fs, const int* fn, int max_levels) {
    orc_mg* h = new orc_mg();
    std::vector<int> s(fs, fs + N + 1), nb(fn, fn + fs[N]);
    h->mg.init(s, nb, N, max_levels);
    return h;
}
void orc_mg_destroy(orc_mg* h) { delete h; }
int orc_mg_levels(const orc_mg* h) { return (int)h->mg.levels.size(); }
int orc_mg_level_info(const orc_mg* h, int l, int* n_cells, int* n_refs) {
    if (l < 0 || l >= (int)h->mg.levels.size()) return AXS_ERR_RANGE;
    *n_cells = h->mg.levels[l].grid.nCells;
    *n_refs = h->mg.levels[l].grid.nRefs();
    return AXS_OK;
}
int orc_mg_get_maps(const orc_mg* h, int l, int* fs, int* fn, int* c2c, int* s2s) {
    if (l < 0 || l >= (int)h->mg.levels.size()) return AXS_ERR_RANGE;
    const Grid& g = h->mg.levels[l].grid;
    if (fs) std::copy(g.faceStart.begin(), g.faceStart.end(), fs);
    if (fn) std::copy(g.faceNeighbor.begin(), g.faceNeighbor.end(), fn);
    if (c2c) std::copy(g.cellToCoarse.begin(), g.cellToCoarse.end(), c2c);
    if (s2s) std::copy(g.fineSlotToCoarseSlot.begin(), g.fineSlotToCoarseSlot.end(), s2s);
    return AXS_OK;
}
int orc_mg_solve(orc_mg* h, const double* AC, const double* AF, const double* b, double* x, int cycles,
                 double weight, int pre, int post, int coarse_sweeps, int smoother) {
    Multigrid& mg = h->mg;
    const int N = mg.levels[0].grid.nCells, S = mg.levels[0].grid.nRefs();
    Coeff A;
    A.N = N;
    A.nRefs = S;
    A.AC.assign(AC, AC + N);
    A.AF.assign(AF, AF + S);
    A.b.assign(b, b + N);
    A.invAC.assign(N, 0.0);
    std::vector<double> xv(x, x + N);
    mg.cfg.maxIter = cycles;
    mg.cfg.weight = weight;
    mg.cfg.pre = pre;
    mg.cfg.post = post;
    mg.cfg.linearSweep = coarse_sweeps;
    mg.cfg.smoother = smoother;
    mg.run(A, xv);
    std::memcpy(x, xv.data(), N * sizeof(double));
    return AXS_OK;
}
// Operator of level l after the last solve (Galerkin product of level l-1).
int orc_mg_get_operator(const orc_mg* h, int l, double* AC, double* AF) {
    if (l < 0 || l >= (int)h->mg.levels.size()) return AXS_ERR_RANGE;
    const Coeff& c = h->mg.levels[l].coeff;
    std::copy(c.AC.begin(), c.AC.end(), AC);
    std::copy(c.AF.begin(), c.AF.end(), AF);
    return AXS_OK;
}

// ---------------------------------------------------------------------------
// Quality::buildQuality, solver/source/quality.cpp:160-166: calculateAspectRatio
// (129-158), calculateOrthogonality (61-105), calculateSkewness (7-59), per
// cell, on the mesh view and the cells' corner rings (cornerStart: n_cells + 1
// offsets into the coordinate lists pz, pr).
// ---------------------------------------------------------------------------
int orc_mesh_quality(const axs_mesh_view* v, const int* cornerStart, const double* pz, const double* pr,
                     double* aspect, double* orth, double* skew) {
    const int N = v->n_cells;
    for (int c = 0; c < N; c++) {
        double dMin = std::numeric_limits<double>::max(), dMax = 0.0;
        for (int k = v->cell_face_start[c]; k < v->cell_face_start[c + 1]; k++) {
            const int f = v->cell_face_ids[k];
            const double fz = v->face_center_z[f] - v->cell_center_z[c];
            const double fr = v->face_center_r[f] - v->cell_center_r[c];
            const double d = std::abs(fz * v->face_normal_z[f] + fr * v->face_normal_r[f]);
            dMin = std::min(dMin, d);
            dMax = std::max(dMax, d);
        }
        aspect[c] = (dMin > 1e-30) ? dMax / dMin : 0.0;
    }
    for (int c = 0; c < N; c++) {
        double o = 1.0;
        for (int k = v->cell_face_start[c]; k < v->cell_face_start[c + 1]; k++) {
            const int f = v->cell_face_ids[k];
            const double fz = v->face_center_z[f] - v->cell_center_z[c];
            const double fr = v->face_center_r[f] - v->cell_center_r[c];
            const double fDot = std::abs(fz * v->face_normal_z[f] + fr * v->face_normal_r[f]);
            const double fMag = std::sqrt(fz * fz + fr * fr);
            o = std::min(o, fDot / fMag);
            if (v->face_neighbor[f] < 0) continue;
            const int P = v->face_owner[f], Nb = v->face_neighbor[f];
            const double dz = v->cell_center_z[Nb] - v->cell_center_z[P];
            const double dr = v->cell_center_r[Nb] - v->cell_center_r[P];
            const double dDot = std::abs(dz * v->face_normal_z[f] + dr * v->face_normal_r[f]);
            const double dMag = std::sqrt(dz * dz + dr * dr);
            o = std::min(o, dDot / dMag);
        }
        orth[c] = o;
    }
    constexpr double radToDeg = 57.29577951308232;
    for (int c = 0; c < N; c++) {
        skew[c] = 0.0;
        const int begin = cornerStart[c], n = cornerStart[c + 1] - begin;
        if (n < 3) continue;
        double tMin = 180.0, tMax = 0.0;
        for (int k = 0; k < n; ++k) {
            const int ip = begin + (k + n - 1) % n, ih = begin + k, in = begin + (k + 1) % n;
            const double az = pz[ip] - pz[ih], ar = pr[ip] - pr[ih];
            const double bz = pz[in] - pz[ih], br = pr[in] - pr[ih];
            const double aMag = std::sqrt(az * az + ar * ar);
            const double bMag = std::sqrt(bz * bz + br * br);
            if (aMag < 1e-30 || bMag < 1e-30) continue;
            const double cosT = std::clamp((az * bz + ar * br) / (aMag * bMag), -1.0, 1.0);
            const double theta = std::acos(cosT) * radToDeg;
            tMin = std::min(tMin, theta);
            tMax = std::max(tMax, theta);
        }
        const double tE = 180.0 - 360.0 / (double)n;
        skew[c] = std::max((tMax - tE) / (180 - tE), (tE - tMin) / tE);
    }
    return AXS_OK;
}

}  // extern "C"
