// kernels.hip -- FP64 SIMPLE kernels for gfx950.
//
// One thread per cell (or face), 256-thread workgroups (four wave64s), CSR
// walks over each cell's face slots.  Every value is computed with the same
// operations in the same order as the reference kernel it replaces, so the FP64
// path reproduces the reference discretisation exactly up to the order of the
// residual-norm sums.  What changes is the schedule: per-cell steps that the
// reference runs as separate launches are fused (see each kernel), the u and v
// systems share one off-diagonal array (they are bit-identical in the
// reference: same diffusion and same upwind flux), and the multigrid transfer
// operators are deterministic gathers instead of atomicAdd scatters.
//
// Every kernel launched from the per-iteration graph first reads ctl->done and
// returns when the outer loop has already ended, so surplus graph launches
// queued ahead of the host's convergence poll do nothing.
#include <cmath>
#include <stdexcept>

#include "device.hpp"
#include "mixed.hpp"
#include "rows.hpp"

namespace axs {

namespace {



// ---- face helpers (solver/header/solver_util.cuh:56-430) ----

__device__ __forceinline__ double faceWeight(const DMesh& m, int c, int f) {
    const double w = m.wP[f];
    return (m.owner[f] == c) ? w : 1.0 - w;
}

__device__ __forceinline__ double normalCorr(const DState& s, int c, double nz, double nr) {
    return s.DU[c] * nz * nz + s.DV[c] * nr * nr;
}

// interpolateNormalCorrectionCoeffToFace, solver_util.cuh:126-161
__device__ __forceinline__ double normalCorrToFace(const DMesh& m, const DState& s, int c, int f) {
    const double nz = m.nZ[f], nr = m.nR[f];
    const double DP = normalCorr(s, c, nz, nr);
    const int ow = m.owner[f], nb = m.neighbor[f];
    if (nb < 0) return DP;
    const int o = (ow == c) ? nb : ow;
    const double DN = normalCorr(s, o, nz, nr);
    const double w = faceWeight(m, c, f);
    return w * DP + (1.0 - w) * DN;
}

// prescribedBoundaryFaceValue, solver_util.cuh:223-238
__device__ __forceinline__ double prescribed(const DMesh& m, int f, uint8_t t, double v, double L) {
    if (t != 2 /* FULLY_DEVELOPED */) return v;
    const double nz = m.nZ[f], nr = m.nR[f];
    const double along = m.fR[f] * nz * nz + m.fZ[f] * nr * nr;
    const double x = along / L;
    return v * (1.0 - x * x);
}

// boundary branch of interpolateFieldToFace, solver_util.cuh:263-297
__device__ __forceinline__ double boundaryFaceValue(const DMesh& m, const DBC& bc, double phiP, int f) {
    const int g = m.group[f];
    if (g < 0 || g >= bc.nGroups) return phiP;
    const uint8_t t = bc.type[g];
    const double v = bc.value[g];
    if (t == 0 || t == 2) return prescribed(m, f, t, v, bc.length[g]);
    if (t == 1) return phiP + v * m.dPB[f];
    return phiP;
}

// interpolateFieldToFace, solver_util.cuh:240-297
__device__ __forceinline__ double interpToFace(const DMesh& m, const DBC& bc, const double* phi, int c,
                                               int f) {
    const int ow = m.owner[f], nb = m.neighbor[f];
    const double phiP = phi[c];
    if (nb >= 0) {
        const int o = (ow == c) ? nb : ow;
        const double w = faceWeight(m, c, f);
        return w * phiP + (1.0 - w) * phi[o];
    }
    return boundaryFaceValue(m, bc, phiP, f);
}

// nonOrthoScalarDiffusionFlux, solver_util.cuh:383-430
__device__ __forceinline__ double nonOrthoFlux(const DMesh& m, int c, int f, const double* gz, const double* gr,
                                               double gamma) {
    const int ow = m.owner[f], nb = m.neighbor[f];
    if (nb < 0) return 0.0;
    const int o = (ow == c) ? nb : ow;
    double nz = m.nZ[f], nr = m.nR[f];
    if (nb == c) {
        nz = -nz;
        nr = -nr;
    }
    const double area = m.area[f];
    const double dz = m.cZ[o] - m.cZ[c];
    const double dr = m.cR[o] - m.cR[c];
    const double aOverNd = area * m.invDPN[f];
    const double Tz = area * nz - aOverNd * dz;
    const double Tr = area * nr - aOverNd * dr;
    const double w = faceWeight(m, c, f);
    const double gfz = w * gz[c] + (1.0 - w) * gz[o];
    const double gfr = w * gr[c] + (1.0 - w) * gr[o];
    return gamma * (Tz * gfz + Tr * gfr);
}

// ---- the same helpers over the slots of cell c (k: slot, o = slotNb[k]) ----
// The cell loops read the face data in slot order, oriented for c (sNZ/sNR =
// the normal pointing out of c, sW = c's interpolation weight, sTZ/sTR = the
// non-orthogonal vector A n - (A/|d|) d towards o, sLsqW* signed so that
// dphi = phi_o - phi_c, sFlags bit 0 = c owns the face), which the host derives
// from the face arrays with the reference's own expressions.  Each result is
// bit-identical to the face form (the squares of the normal in
// normalCorrToFace do not see its orientation); only boundary slots still read
// face-indexed data.


// nonOrthoFlux from loaded values (T, w, the gradients at c and o)
__device__ __forceinline__ double nonOrthoVal(double Tz, double Tr, double w, double gzc, double gzo, double grc,
                                              double gro, double gamma) {
    const double gfz = w * gzc + (1.0 - w) * gzo;
    const double gfr = w * grc + (1.0 - w) * gro;
    return gamma * (Tz * gfz + Tr * gfr);
}

// normalCorrToFace from loaded values (o < 0: boundary slot)
__device__ __forceinline__ double normalCorrVal(double DUc, double DVc, double DUo, double DVo, double nz, double nr,
                                                double w, int o) {
    const double DP = DUc * nz * nz + DVc * nr * nr;
    if (o < 0) return DP;
    const double DN = DUo * nz * nz + DVo * nr * nr;
    return w * DP + (1.0 - w) * DN;
}

// phiGradientGreenGauss / phiGradientLeastSquare, solver_util.cuh:458-588, of
// NF fields of cell c in one walk over its slots (the geometry is read once
// for all of them; each field's sums run in the reference's order).  In the
// least-squares form dphi = phi_o - phi_c with the weight negated where c is
// the face's neighbour: (-w)(a - b) == w(b - a) exactly.
template <int NF, int B>
__device__ void cellGradients(const DMesh& m, const DBC* bc, const double* const* phi, int c, int scheme,
                              double* gZ, double* gR) {
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, c, k0, ns);
    const bool batch = ns > 0 && ns <= B;
    double phiP[NF];
#pragma unroll
    for (int i = 0; i < NF; i++) phiP[i] = phi[i][c];
    if (scheme == 0) {
        double gz[NF], gr[NF], axisPhi[NF];
        double clz = 0.0, clr = 0.0;
#pragma unroll
        for (int i = 0; i < NF; i++) gz[i] = gr[i] = axisPhi[i] = 0.0;
        bool axis = false;
        auto step = [&](int k, int o, double nz, double nr, double w, uint8_t fl, double L, const double* phiO) {
            double phiF[NF];
            if (o >= 0) {
#pragma unroll
                for (int i = 0; i < NF; i++) phiF[i] = w * phiP[i] + (1.0 - w) * phiO[i];
            } else {
                const int f = m.faceIDs[k];
#pragma unroll
                for (int i = 0; i < NF; i++) phiF[i] = boundaryFaceValue(m, bc[i], phiP[i], f);
            }
            if (!(fl & 2)) {
#pragma unroll
                for (int i = 0; i < NF; i++) {
                    gz[i] += phiF[i] * nz * L;
                    gr[i] += phiF[i] * nr * L;
                }
                clz += nz * L;
                clr += nr * L;
            } else {
                axis = true;
#pragma unroll
                for (int i = 0; i < NF; i++) axisPhi[i] = phiF[i];
            }
        };
        if (batch) {
            int o[B];
            uint8_t fl[B];
            double nz[B], nr[B], w[B], L[B], po[B][NF];
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int k = slotAt(k0, ns, q);
                o[q] = m.slotNb[k];
                nz[q] = m.sNZ[k];
                nr[q] = m.sNR[k];
                w[q] = m.sW[k];
                fl[q] = m.sFlags[k];
                L[q] = m.sLen[k];
            }
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int oc = o[q] >= 0 ? o[q] : c;
#pragma unroll
                for (int i = 0; i < NF; i++) po[q][i] = phi[i][oc];
            }
#pragma unroll
            for (int q = 0; q < B; q++)
                if (q < ns) step(k0 + q, o[q], nz[q], nr[q], w[q], fl[q], L[q], po[q]);
        } else {
            for (int k = k0; k < k0 + ns; k++) {
                const int o = m.slotNb[k];
                const int oc = o >= 0 ? o : c;
                double po[NF];
#pragma unroll
                for (int i = 0; i < NF; i++) po[i] = phi[i][oc];
                step(k, o, m.sNZ[k], m.sNR[k], m.sW[k], m.sFlags[k], m.sLen[k], po);
            }
        }
        const double inv = m.invA2D[c];
#pragma unroll
        for (int i = 0; i < NF; i++) {
            if (axis) {
                gz[i] += axisPhi[i] * (-clz);
                gr[i] += axisPhi[i] * (-clr);
            }
            gZ[i] = gz[i] * inv;
            gR[i] = gr[i] * inv;
        }
    } else {
        double bz[NF], br[NF];
#pragma unroll
        for (int i = 0; i < NF; i++) bz[i] = br[i] = 0.0;
        auto step = [&](int k, int o, double wz, double wr, const double* phiO) {
            double dphi[NF];
            if (o >= 0) {
#pragma unroll
                for (int i = 0; i < NF; i++) dphi[i] = phiO[i] - phiP[i];
            } else {
                const int f = m.faceIDs[k];
#pragma unroll
                for (int i = 0; i < NF; i++) dphi[i] = boundaryFaceValue(m, bc[i], phiP[i], f) - phiP[i];
            }
#pragma unroll
            for (int i = 0; i < NF; i++) {
                bz[i] += wz * dphi[i];
                br[i] += wr * dphi[i];
            }
        };
        if (batch) {
            int o[B];
            double wz[B], wr[B], po[B][NF];
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int k = slotAt(k0, ns, q);
                o[q] = m.slotNb[k];
                wz[q] = m.sLsqWZ[k];
                wr[q] = m.sLsqWR[k];
            }
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int oc = o[q] >= 0 ? o[q] : c;
#pragma unroll
                for (int i = 0; i < NF; i++) po[q][i] = phi[i][oc];
            }
#pragma unroll
            for (int q = 0; q < B; q++)
                if (q < ns) step(k0 + q, o[q], wz[q], wr[q], po[q]);
        } else {
            for (int k = k0; k < k0 + ns; k++) {
                const int o = m.slotNb[k];
                const int oc = o >= 0 ? o : c;
                double po[NF];
#pragma unroll
                for (int i = 0; i < NF; i++) po[i] = phi[i][oc];
                step(k, o, m.sLsqWZ[k], m.sLsqWR[k], po);
            }
        }
        const double izz = m.lsqInvZZ[c], izr = m.lsqInvZR[c], irr = m.lsqInvRR[c];
#pragma unroll
        for (int i = 0; i < NF; i++) {
            gZ[i] = izz * bz[i] + izr * br[i];
            gR[i] = izr * bz[i] + irr * br[i];
        }
    }
}

template <int B>
__device__ __forceinline__ void cellGradient(const DMesh& m, const DBC& bc, const double* phi, int c, int scheme,
                                             double& gZ, double& gR) {
    double z[1], r[1];
    cellGradients<1, B>(m, &bc, &phi, c, scheme, z, r);
    gZ = z[0];
    gR = r[0];
}

// higherOrderFaceValue (solver_util.cuh:724-768) on the interior slot k of cell n
// (neighbour nb, c's weight w, phi at n and nb loaded)
__device__ __forceinline__ double higherOrderVal(const DMesh& m, const double* phi, const double* gz,
                                                 const double* gr, int scheme, int n, int nb, int k, double F,
                                                 double phiUD, double w, double phiN, double phiO) {
    if (scheme == 1) return w * phiN + (1.0 - w) * phiO;
    if (scheme == 2) {
        const int f = m.faceIDs[k];
        const int up = (F > 0.0) ? n : nb;
        const double dz = m.fZ[f] - m.cZ[up];
        const double dr = m.fR[f] - m.cR[up];
        const double ho = phi[up] + gz[up] * dz + gr[up] * dr;
        const double lo = fmin(phiN, phiO);
        const double hi = fmax(phiN, phiO);
        return fmin(fmax(ho, lo), hi);
    }
    return phiUD;
}

// Boundary half of addTransportCoefficients (solver_util.cu:340-436) and the
// boundary branch of addConvectionContribution (solver_util.cuh:817-864).
__device__ __forceinline__ void boundaryTransport(const DMesh& m, const DBC& bc, int n, int f, double area,
                                                  double gamma, double F, double& AC, double& b) {
    const int g = m.group[f];
    if (g >= 0 && g < bc.nGroups) {
        const uint8_t t = bc.type[g];
        const double v = bc.value[g];
        const double dPF = fabs((m.fZ[f] - m.cZ[n]) * m.nZ[f] + (m.fR[f] - m.cR[n]) * m.nR[f]);
        const double K = gamma * area / dPF;
        if (t == 0) {
            AC += K;
            b += K * v;
        } else if (t == 1) {
            b += gamma * area * v;
        } else if (t == 2) {
            AC += K;
            b += K * prescribed(m, f, t, v, bc.length[g]);
        }
    }
    if (F != 0.0) {
        if (g < 0 || g >= bc.nGroups) {
            AC += F;
        } else {
            const uint8_t t = bc.type[g];
            const double v = bc.value[g];
            if (t == 0) {
                if (F < 0.0) b += -F * v;
                else AC += F;
            } else if (t == 1) {
                AC += F;
            } else if (t == 2) {
                if (F < 0.0) b += -F * prescribed(m, f, t, v, bc.length[g]);
                else AC += F;
            }
        }
    }
}

// rhieChowNormalVelocityToFace, solver_util.cuh:590-713 (cellID = owner)
__device__ __forceinline__ double rhieChow(const DMesh& m, const DBC& bu, const DBC& bv, const DBC& bp,
                                          const DState& s, int c, int f) {
    const int nb = m.neighbor[f];
    const double nz = m.nZ[f], nr = m.nR[f];
    const double uF = interpToFace(m, bu, s.u, c, f);
    const double vF = interpToFace(m, bv, s.v, c, f);
    const double unLin = uF * nz + vF * nr;
    if (nb < 0) {
        const int g = m.group[f];
        if (g < 0 || g >= bp.nGroups) return unLin;
        if (bp.type[g] != 0) return unLin;
        const double dPB = m.dPB[f];
        const double DfB = normalCorrToFace(m, s, c, f);
        const double compact = (bp.value[g] - s.p[c]) / dPB;
        const double cellN = s.gPZ[c] * nz + s.gPR[c] * nr;
        return unLin - DfB * (compact - cellN);
    }
    const double gPN = (s.p[nb] - s.p[c]) * m.invDPN[f];
    const double w = m.wP[f];
    const double gzF = w * s.gPZ[c] + (1.0 - w) * s.gPZ[nb];
    const double grF = w * s.gPR[c] + (1.0 - w) * s.gPR[nb];
    const double gFN = gzF * nz + grF * nr;
    const double Df = normalCorrToFace(m, s, c, f);
    return unLin - Df * (gPN - gFN);
}

// ---- wave64 reductions by DPP (deterministic: a fixed tree) ----
// v moved across lanes by the DPP control `ctrl` (both 32-bit halves); lanes
// whose source is outside the enabled rows / banks, or outside the row, keep
// `identity`
template <int ctrl, int rowMask>
__device__ __forceinline__ double dppMove(double v, double identity) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(identity), __double2loint(v), ctrl, rowMask, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(identity), __double2hiint(v), ctrl, rowMask, 0xf, false);
    return __hiloint2double(hi, lo);
}
// Inclusive scan of a wave's values with VALU moves only (no LDS traffic):
// row_shr 1, 2, 4, 8 within each 16-lane row, then row_bcast 15 (rows 1, 3
// add the row before) and row_bcast 31 (rows 2, 3 add rows 0-1).  Lane 63
// ends with the wave's sum (MAX: maximum; the values are >= 0, so 0 is the
// identity of both).
template <bool MAX>
__device__ __forceinline__ double waveScan(double v) {
    auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
    v = op(v, dppMove<0x111, 0xf>(v, 0.0));  // row_shr:1
    v = op(v, dppMove<0x112, 0xf>(v, 0.0));  // row_shr:2
    v = op(v, dppMove<0x114, 0xf>(v, 0.0));  // row_shr:4
    v = op(v, dppMove<0x118, 0xf>(v, 0.0));  // row_shr:8
    v = op(v, dppMove<0x142, 0xa>(v, 0.0));  // row_bcast:15
    v = op(v, dppMove<0x143, 0xc>(v, 0.0));  // row_bcast:31
    return v;
}

}  // namespace

// ===========================================================================
// Gradients (computeGradient, solver_util.cu:146-169)
// ===========================================================================
// P: the pressure gradient too (then kMomentum reads it instead of computing
// it); the three fields share one walk over the slots and their geometry.
template <int B, bool P>
__global__ void __launch_bounds__(kBlock) kGradUV(DMesh m, DBC bu, DBC bv, DBC bp, DState s, int scheme) {
    if (finished(s.ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    constexpr int NF = P ? 3 : 2;
    const DBC bc[3] = {bu, bv, bp};
    const double* phi[3] = {s.u, s.v, s.p};
    double gz[NF], gr[NF];
    cellGradients<NF, B>(m, bc, phi, n, scheme, gz, gr);
    s.gUZ[n] = gz[0];
    s.gUR[n] = gr[0];
    s.gVZ[n] = gz[1];
    s.gVR[n] = gr[1];
    if (P) {
        s.gPZ[n] = gz[NF - 1];
        s.gPR[n] = gr[NF - 1];
    }
}

template <int B>
__global__ void __launch_bounds__(kBlock) kGradField(DMesh m, DBC bc, const double* phi, double* gzo,
                                                     double* gro, int scheme, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double gz, gr;
    cellGradient<B>(m, bc, phi, n, scheme, gz, gr);
    gzo[n] = gz;
    gro[n] = gr;
}

// ===========================================================================
// Momentum: one pass per cell doing, in the reference's order,
//   clearAllCoefficients            solver_util.cuh:976-997   (zero accumulators)
//   addTransportCoefficients u, v   solver_util.cu:218-448
//   addRadialMomentumCylindrical... solver_util.cu:454-470
//   addTransientCoefficient u, v    solver_util.cu:536-578    (transient only)
//   computeGradient(p)              solver_util.cu:146-169
//   createMomentumPressureRhs       simple.cu:144-164
//   underRelaxEquation u, v         solver_util.cu:582-600
//   buildInverseDiagonal u, v       solver_util.cu:602-612
//   getCorrectionCoefficients       solver_util.cu:127-144
// Every step is per-cell, so one launch replaces nine.
// ===========================================================================
// HO: higher-order convection (the neighbour's u, v are read); the upwind
// instance holds fewer registers.  PG: the pressure gradient is computed
// here (else kGradUV has already written it, from the same p).
// DEF: also the FP64 defect b - A u0 into s.momR32 (the mixed path's fused
// FP32 sweeps start from it)
template <int B, bool HO, bool PG, bool DEF>
__global__ void __launch_bounds__(kBlock) kMomentum(DMesh m, DBC bu, DBC bv, DBC bp, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    const double mu = ph.mu;
    double acU = 0.0, bU = 0.0, acV = 0.0, bV = 0.0, netF = 0.0;
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
    const bool needW = ph.nonOrtho || ph.scheme == 1;
    // the neighbours' u, v: for higher-order convection, and for the FP64
    // defect of the mixed path's first sweep
    const bool needUV = HO || DEF;
    const double uN = s.u[n], vN = s.v[n];
    const double guzN = ph.nonOrtho ? s.gUZ[n] : 0.0, gurN = ph.nonOrtho ? s.gUR[n] : 0.0;
    const double gvzN = ph.nonOrtho ? s.gVZ[n] : 0.0, gvrN = ph.nonOrtho ? s.gVR[n] : 0.0;
    // one slot (the values at the neighbour o already loaded); returns AF
    // noU / noV: the slot's non-orthogonal diffusion terms (nonOrthoVal),
    // evaluated before the walk and added at the same point of it
    auto step = [&](int k, int o, uint8_t fl, double area, double invDPN, double md, double noU, double noV,
                    double w, double uO, double vO) -> double {
        double F = 0.0;
        if (ph.conv) {
            // a boundary face is always owned by its one cell
            const double Fo = md * 1.0;
            F = (fl & 1) ? Fo : -Fo;
            if (fabs(F) <= 1.0e-30) F = 0.0;
            netF += F;
        }
        if (o >= 0) {
            const double K = mu * area * invDPN;
            double af = 0.0;
            acU += K;
            af -= K;
            if (ph.nonOrtho) bU += noU;
            if (F != 0.0) {
                acU += fmax(F, 0.0);
                af += fmin(F, 0.0);
                if (HO) {
                    const double ud = (F > 0.0) ? uN : uO;
                    const double ho = higherOrderVal(m, s.u, s.gUZ, s.gUR, ph.scheme, n, o, k, F, ud, w, uN, uO);
                    bU -= F * (ho - ud);
                }
            }
            acV += K;
            if (ph.nonOrtho) bV += noV;
            if (F != 0.0) {
                acV += fmax(F, 0.0);
                if (HO) {
                    const double ud = (F > 0.0) ? vN : vO;
                    const double ho = higherOrderVal(m, s.v, s.gVZ, s.gVR, ph.scheme, n, o, k, F, ud, w, vN, vO);
                    bV -= F * (ho - ud);
                }
            }
            return af;
        }
        const int f = m.faceIDs[k];
        boundaryTransport(m, bu, n, f, area, mu, F, acU, bU);
        boundaryTransport(m, bv, n, f, area, mu, F, acV, bV);
        return 0.0;
    };
    // kept for the defect at the end (batched rows)
    int o[B];
    double uo[B], vo[B], af[B];
    if (ns > 0 && ns <= B) {
        int fid[B];
        uint8_t fl[B];
        double area[B], inv[B], w[B], md[B];
        double noU[B], noV[B];
#pragma unroll
        for (int q = 0; q < B; q++) o[q] = m.slotNb[slotAt(k0, ns, q)];
        if (ph.nonOrtho) {
            // first pass: the non-orthogonal terms, so that their geometry and
            // the neighbours' gradients are not held through the main walk
            double tz[B], tr[B], wn[B], guz[B], gur[B], gvz[B], gvr[B];
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int k = slotAt(k0, ns, q);
                const int oc = o[q] >= 0 ? o[q] : n;
                tz[q] = m.sTZ[k];
                tr[q] = m.sTR[k];
                wn[q] = m.sW[k];
                guz[q] = s.gUZ[oc];
                gur[q] = s.gUR[oc];
                gvz[q] = s.gVZ[oc];
                gvr[q] = s.gVR[oc];
            }
#pragma unroll
            for (int q = 0; q < B; q++) {
                noU[q] = nonOrthoVal(tz[q], tr[q], wn[q], guzN, guz[q], gurN, gur[q], mu);
                noV[q] = nonOrthoVal(tz[q], tr[q], wn[q], gvzN, gvz[q], gvrN, gvr[q], mu);
            }
        } else {
#pragma unroll
            for (int q = 0; q < B; q++) noU[q] = noV[q] = 0.0;
        }
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            fid[q] = m.faceIDs[k];
            fl[q] = m.sFlags[k];
            area[q] = m.sArea[k];
            inv[q] = m.sInvDPN[k];
            w[q] = HO ? m.sW[k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int oc = o[q] >= 0 ? o[q] : n;
            md[q] = ph.conv ? s.mDot[fid[q]] : 0.0;
            uo[q] = needUV ? s.u[oc] : 0.0;
            vo[q] = needUV ? s.v[oc] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns) af[q] = step(k0 + q, o[q], fl[q], area[q], inv[q], md[q], noU[q], noV[q], w[q], uo[q], vo[q]);
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns) {
                s.AFuv[k0 + q] = af[q];
                if (s.AF32) s.AF32[k0 + q] = (float)af[q];
            }
    } else {
        for (int k = k0; k < k0 + ns; k++) {
            const int o = m.slotNb[k];
            const int oc = o >= 0 ? o : n;
            const double w = needW ? m.sW[k] : 0.0;
            const double noU = ph.nonOrtho ? nonOrthoVal(m.sTZ[k], m.sTR[k], w, guzN, s.gUZ[oc], gurN, s.gUR[oc], mu) : 0.0;
            const double noV = ph.nonOrtho ? nonOrthoVal(m.sTZ[k], m.sTR[k], w, gvzN, s.gVZ[oc], gvrN, s.gVR[oc], mu) : 0.0;
            const double a = step(k, o, m.sFlags[k], m.sArea[k], m.sInvDPN[k], ph.conv ? s.mDot[m.faceIDs[k]] : 0.0,
                                  noU, noV, w, needUV ? s.u[oc] : 0.0, needUV ? s.v[oc] : 0.0);
            s.AFuv[k] = a;
            if (s.AF32) s.AF32[k] = (float)a;
        }
    }
    if (ph.conv) {
        acU -= netF;
        acV -= netF;
    }
    const double V = m.vol[n];
    const double r = m.cR[n];
    acV += mu * V / (r * r);
    if (ph.transient && ph.dt > 0.0) {
        const double cV = ph.rho * V;
        if (s.ctl != nullptr && s.ctl->bdf2Step) {
            const double a = cV / (2.0 * ph.dt);
            acU += 3.0 * a;
            bU += a * (4.0 * s.uOld[n] - s.uOld2[n]);
            acV += 3.0 * a;
            bV += a * (4.0 * s.vOld[n] - s.vOld2[n]);
        } else {
            const double a = cV / ph.dt;
            acU += a;
            bU += a * s.uOld[n];
            acV += a;
            bV += a * s.vOld[n];
        }
    }
    double gz, gr;
    if (PG) {
        cellGradient<B>(m, bp, s.p, n, ph.gradScheme, gz, gr);
        s.gPZ[n] = gz;
        s.gPR[n] = gr;
    } else {
        gz = s.gPZ[n];
        gr = s.gPR[n];
    }
    bU += -V * gz;
    bV += -V * gr;
    const double al = ph.alphaU;
    if (!(fabs(acU) < 1.0e-30)) {
        const double a0 = acU;
        acU = a0 / al;
        bU += ((1.0 - al) / al) * a0 * s.u[n];
    }
    if (!(fabs(acV) < 1.0e-30)) {
        const double a0 = acV;
        acV = a0 / al;
        bV += ((1.0 - al) / al) * a0 * s.v[n];
    }
    const double iU = (fabs(acU) < 1.0e-30) ? 0.0 : 1.0 / acU;
    const double iV = (fabs(acV) < 1.0e-30) ? 0.0 : 1.0 / acV;
    s.uAC[n] = acU;
    s.ub[n] = bU;
    s.uInv[n] = iU;
    s.vAC[n] = acV;
    s.vb[n] = bV;
    s.vInv[n] = iV;
    s.DU[n] = V * iU;
    s.DV[n] = V * iV;
    if (s.uvInv32) s.uvInv32[n] = make_float2((float)iU, (float)iV);
    if (DEF) {
        // the FP64 defect b - A u0 of the mixed path's FP32 sweeps, as
        // kMomDefect forms it (diagonal term first, then the slots in order)
        double au = acU * uN, av = acV * vN;
        if (ns > 0 && ns <= B) {
#pragma unroll
            for (int q = 0; q < B; q++)
                if (q < ns && o[q] >= 0) {
                    au += af[q] * uo[q];
                    av += af[q] * vo[q];
                }
        } else {
            for (int k = k0; k < k0 + ns; k++) {
                const int nb = m.slotNb[k];
                if (nb >= 0) {
                    const double a = s.AFuv[k];
                    au += a * s.u[nb];
                    av += a * s.v[nb];
                }
            }
        }
        s.momR32[n] = make_float2((float)(bU - au), (float)(bV - av));
    }
}

// ===========================================================================
// Momentum smoothers: one Jacobi sweep of u and v together (jacobi,
// linear_solver.cu:9-33); the two systems share CSR and off-diagonals.
// ===========================================================================
// The u and v iterates either as two arrays or interleaved as (u, v) pairs.
// The sweeps in between the first and the last keep them interleaved, so a
// neighbour's u and v arrive in one 16-byte load from one cache line instead
// of two loads from two lines (the gathers are what the sweep over-fetches).
struct UVIn {
    const double *u, *v;
    const double2* uv;
    __device__ __forceinline__ double2 at(int i) const { return uv ? uv[i] : make_double2(u[i], v[i]); }
};
struct UVOut {
    double *u, *v;
    double2* uv;
    __device__ __forceinline__ void put(int i, double a, double b) const {
        if (uv) uv[i] = make_double2(a, b);
        else {
            u[i] = a;
            v[i] = b;
        }
    }
    __device__ __forceinline__ void put(int i, double2 x) const { put(i, x.x, x.y); }
};

template <int B>
__global__ void __launch_bounds__(kBlock) kJacobiUV(DMesh m, DState s, UVIn xo, UVOut xn) {
    if (finished(s.ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double au = s.ub[n], av = s.vb[n];
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
    if (ns > 0 && ns <= B) {
        int o[B];
        double a[B];
        double2 xv[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = m.slotNb[k];
            a[q] = s.AFuv[k];
        }
#pragma unroll
        for (int q = 0; q < B; q++) xv[q] = xo.at(o[q] >= 0 ? o[q] : n);
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns && o[q] >= 0) {
                au -= a[q] * xv[q].x;
                av -= a[q] * xv[q].y;
            }
    } else {
        for (int k = k0; k < k0 + ns; k++) {
            const int nb = m.slotNb[k];
            if (nb >= 0) {
                const double a = s.AFuv[k];
                const double2 x = xo.at(nb);
                au -= a * x.x;
                av -= a * x.y;
            }
        }
    }
    const double iu = s.uInv[n], iv = s.vInv[n];
    const double2 x = xo.at(n);
    xn.put(n, (iu != 0.0) ? au * iu : x.x, (iv != 0.0) ? av * iv : x.y);
}

// S Jacobi sweeps of u and v in one launch (temporal blocking).  Workgroup b
// owns C cells (the index range [bC, bC + C), or a breadth-first patch the
// host lists); sweep j updates them and the halo rings 1 .. S - j, the
// iterates of sweeps 1 .. S - 1 live in LDS, and only sweep S writes global
// memory.  Thread t holds two rows in registers for the whole launch: block
// cell t and halo entry t (a halo has at most sweepHalo<C>() cells), each as
// b, 1/AC and up to kSweepRegSlots off-diagonals with their
// neighbours, all loaded up front, so the CSR and coefficients are read once
// per S sweeps and no sweep waits on a chain of global loads.  Each slot's
// neighbour LDS slot comes from the host-built tables (int16, loaded with the
// row).  Every row is the same expression as in kJacobiUV, so the result is
// bit-identical to S launches of it.
// R: register slots per row (3 on triangle meshes, else kSweepRegSlots)
template <int R>
struct SweepRow {
    double af[R];
    int nb[R];   // neighbour cell, -1: none (boundary or past the row's slots)
    int loc[R];  // the neighbour's LDS slot, -1: none (loaded after the first sweep)
    int k0, ns;  // the row's slots
    double bu, bv, iu, iv;
};

template <int R>
__device__ __forceinline__ void loadSweepRow(SweepRow<R>& r, const DMesh& m, const DState& s, int c) {
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, c, k0, ns);
    r.k0 = k0;
    r.ns = ns;
    // slots read at indices clamped into the row (index 0 for an empty row),
    // then masked: unconditional loads go out together
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int k = ns > 0 ? k0 + min(q, ns - 1) : 0;
        const int nb = m.slotNb[k];
        const double af = s.AFuv[k];
        r.nb[q] = (q < ns) ? nb : -1;
        r.af[q] = (q < ns) ? af : 0.0;
    }
    r.bu = s.ub[c];
    r.bv = s.vb[c];
    r.iu = s.uInv[c];
    r.iv = s.vInv[c];
}

// jacobi's row (kJacobiUV) from registers; xat(q) = the neighbour of slot q
// (for a slot without one, any valid entry: it is read and not used, so that
// every read is unconditional and they all go out together); LDS: the
// neighbours are the ones with an LDS slot (the same ones)
template <int R, bool LDS = false, typename XAt>
__device__ __forceinline__ double2 sweepRow(const SweepRow<R>& r, XAt xat, double2 self) {
    double2 xv[R];
#pragma unroll
    for (int q = 0; q < R; q++) xv[q] = xat(q);
    double au = r.bu, av = r.bv;
#pragma unroll
    for (int q = 0; q < R; q++) {
        if (LDS ? r.loc[q] >= 0 : r.nb[q] >= 0) {
            au -= r.af[q] * xv[q].x;
            av -= r.af[q] * xv[q].y;
        }
    }
    return make_double2((r.iu != 0.0) ? au * r.iu : self.x, (r.iv != 0.0) ? av * r.iv : self.y);
}

// Triangle rows (R = 3) fit 64 VGPRs, i.e. two 1024-thread blocks per CU;
// the attribute holds the compiler to that (longer rows take what they need).
template <int C, int S, int R>
__global__ void __launch_bounds__(C) __attribute__((amdgpu_waves_per_eu(R <= 3 ? 8 : 1)))
kJacobiUVBlock(DMesh m, DState s, DSweepBlocks sb, UVIn xo, UVOut xn) {
    constexpr int H = sweepHalo<C>();
    constexpr int kMask = S > 2 ? 1 : 0;  // sweep j < S writes LDS buffer (j - 1) & kMask
    __shared__ double2 buf[kMask + 1][C + H];
    if (finished(s.ctl)) return;
    const int t = threadIdx.x, blk = xcdBlock();
    const int a = blk * C, nOwn = min(C, m.N - a);
    const int h0 = sb.haloStart[blk];
    // halo entries of rings 1 .. S - 1: the ones sweep 1 updates; of rings
    // 1 .. S - 2: the ones sweeps 2 .. S - 1 update again, from LDS
    const int nHalo1 = (S > 1) ? sb.ringEnd[blk * sb.rings + S - 2] : 0;
    const int nHaloLds = (S > 2) ? sb.ringEnd[blk * sb.rings + (S > 2 ? S - 3 : 0)] : 0;
    // block cell t: a + t, or the t-th cell of the block's patch
    const int n = sb.own ? sb.own[(size_t)blk * C + t] : a + t;
    const bool mine = sb.own ? n >= 0 : t < nOwn, haloRow = t < nHalo1;
    const int hc = haloRow ? sb.halo[h0 + t] : 0;
    SweepRow<R> own, hal;
    if (mine) loadSweepRow(own, m, s, n);
    if (haloRow) loadSweepRow(hal, m, s, hc);
    // sweep 1: global iterate in, LDS out
    double2* out = buf[0];
    if (mine) out[t] = sweepRow(own, [&](int q) { return xo.at(own.nb[q] >= 0 ? own.nb[q] : n); }, xo.at(n));
    if (haloRow)
        out[C + t] = sweepRow(hal, [&](int q) { return xo.at(hal.nb[q] >= 0 ? hal.nb[q] : hc); }, xo.at(hc));
    // the LDS slots the later sweeps read, loaded while the barrier waits
    if (mine) loadLoc(own, sb.locOwn + own.k0, R);
    if (t < nHaloLds) loadLoc(hal, sb.locHalo + (size_t)(h0 + t) * kSweepRegSlots, R);
    __syncthreads();
#pragma unroll
    for (int j = 2; j <= S; j++) {
        const double2* in = buf[(j - 2) & kMask];
        if (j == S) {
            if (mine) xn.put(n, sweepRow<R, true>(own, [&](int q) { return in[own.loc[q] >= 0 ? own.loc[q] : t]; }, in[t]));
        } else {
            double2* o = buf[(j - 1) & kMask];
            const int nHaloJ = sb.ringEnd[blk * sb.rings + (S - j) - 1];
            if (mine) o[t] = sweepRow<R, true>(own, [&](int q) { return in[own.loc[q] >= 0 ? own.loc[q] : t]; }, in[t]);
            if (t < nHaloJ)
                o[C + t] = sweepRow<R, true>(
                    hal, [&](int q) { return in[hal.loc[q] >= 0 ? hal.loc[q] : C + t]; }, in[C + t]);
            __syncthreads();
        }
    }
}

// gaussSeidelColorSweep, linear_solver.cu:42-72, u and v together
__global__ void __launch_bounds__(kBlock) kGSColorUV(DMesh m, DState s, const int* order, int begin, int count) {
    if (finished(s.ctl)) return;
    const int t = xcdBlock() * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int n = order[begin + t];
    double au = s.ub[n], av = s.vb[n];
    for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
        const int nb = m.slotNb[k];
        if (nb >= 0) {
            const double a = s.AFuv[k];
            au -= a * s.u[nb];
            av -= a * s.v[nb];
        }
    }
    const double iu = s.uInv[n], iv = s.vInv[n];
    s.u[n] = (iu != 0.0) ? au * iu : s.u[n];
    s.v[n] = (iv != 0.0) ? av * iv : s.v[n];
}

// jacobi / gaussSeidelColorSweep on one system (the p' solve without multigrid)
__global__ void __launch_bounds__(kBlock) kJacobiSys(DMesh m, const double* b, const double* inv, const double* AF,
                                                     const double* __restrict__ xo, double* __restrict__ xn,
                                                     const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double a = b[n];
    for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
        const int nb = m.slotNb[k];
        if (nb >= 0) a -= AF[k] * xo[nb];
    }
    const double i = inv[n];
    xn[n] = (i != 0.0) ? a * i : xo[n];
}

__global__ void __launch_bounds__(kBlock) kGSColorSys(DMesh m, const double* b, const double* inv, const double* AF,
                                                      double* x, const int* order, int begin, int count,
                                                      const Ctl* ctl) {
    if (finished(ctl)) return;
    const int t = xcdBlock() * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int n = order[begin + t];
    double a = b[n];
    for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
        const int nb = m.slotNb[k];
        if (nb >= 0) a -= AF[k] * x[nb];
    }
    const double i = inv[n];
    x[n] = (i != 0.0) ? a * i : x[n];
}

// ===========================================================================
// Pressure correction matrix (createPPCoeff, simple.cu:7-88) + its inverse
// diagonal (buildInverseDiagonal on ppCoeff, solver.cpp:1178)
// ===========================================================================
template <int B>
__global__ void __launch_bounds__(kBlock) kPPCoeff(DMesh m, DBC bp, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double ac = 0.0;
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
    const double DUn = s.DU[n], DVn = s.DV[n];
    auto step = [&](int k, int o, double area, double nz, double nr, double w, double invDPN, double DUo,
                    double DVo) -> double {
        const double Df = normalCorrVal(DUn, DVn, DUo, DVo, nz, nr, w, o);
        double af = 0.0;
        if (o >= 0) {
            const double K = ph.rho * area * Df * invDPN;
            ac += K;
            af -= K;
        } else {
            const int f = m.faceIDs[k];
            const int g = m.group[f];
            if (g >= 0 && g < bp.nGroups && bp.type[g] == 0) {
                const double K = ph.rho * area * Df / m.dPB[f];
                ac += K;
            }
        }
        return af;
    };
    if (ns > 0 && ns <= B) {
        int o[B];
        double area[B], nz[B], nr[B], w[B], inv[B];
        double du[B], dv[B], af[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = m.slotNb[k];
            area[q] = m.sArea[k];
            nz[q] = m.sNZ[k];
            nr[q] = m.sNR[k];
            w[q] = m.sW[k];
            inv[q] = m.sInvDPN[k];
        }
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int oc = o[q] >= 0 ? o[q] : n;
            du[q] = s.DU[oc];
            dv[q] = s.DV[oc];
        }
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns) af[q] = step(k0 + q, o[q], area[q], nz[q], nr[q], w[q], inv[q], du[q], dv[q]);
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns) {
                s.ppAF[k0 + q] = af[q];
                if (s.ppAF32) s.ppAF32[k0 + q] = (float)af[q];
            }
    } else {
        for (int k = k0; k < k0 + ns; k++) {
            const int o = m.slotNb[k];
            const int oc = o >= 0 ? o : n;
            const double af =
                step(k, o, m.sArea[k], m.sNZ[k], m.sNR[k], m.sW[k], m.sInvDPN[k], s.DU[oc], s.DV[oc]);
            s.ppAF[k] = af;
            if (s.ppAF32) s.ppAF32[k] = (float)af;
        }
    }
    s.ppAC[n] = ac;
    if (s.ppAC32) s.ppAC32[n] = (float)ac;
    s.ppInv[n] = (fabs(ac) < 1.0e-30) ? 0.0 : 1.0 / ac;
}

// computeFaceMassFluxRhieChow, solver_util.cu:171-203.  An interior face's
// data (the face's geometry, both cells' u, v, p, grad p, DU, DV) are loaded
// in one round, unconditionally (a boundary face reads its owner twice), and
// rhieChow's interior arithmetic runs on them in its order; boundary faces
// take rhieChow itself.
__global__ void __launch_bounds__(kBlock) kFaceFlux(DMesh m, DBC bu, DBC bv, DBC bp, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int f = xcdBlock() * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const int c = m.owner[f], nb = m.neighbor[f];
    const int o = nb >= 0 ? nb : c;
    const double area = m.area[f], nz = m.nZ[f], nr = m.nR[f], w = m.wP[f], inv = m.invDPN[f];
    const double uc = s.u[c], vc = s.v[c], pc = s.p[c], gzc = s.gPZ[c], grc = s.gPR[c], duc = s.DU[c], dvc = s.DV[c];
    const double uo = s.u[o], vo = s.v[o], po = s.p[o], gzo = s.gPZ[o], gro = s.gPR[o], duo = s.DU[o], dvo = s.DV[o];
    if (area <= 1.0e-30) {
        s.mDot[f] = 0.0;
        return;
    }
    if (nb < 0) {
        s.mDot[f] = ph.rho * rhieChow(m, bu, bv, bp, s, c, f) * area;
        return;
    }
    // interpToFace / normalCorrToFace seen from the owner: its weight is wP
    const double uF = w * uc + (1.0 - w) * uo;
    const double vF = w * vc + (1.0 - w) * vo;
    const double unLin = uF * nz + vF * nr;
    const double gPN = (po - pc) * inv;
    const double gzF = w * gzc + (1.0 - w) * gzo;
    const double grF = w * grc + (1.0 - w) * gro;
    const double gFN = gzF * nz + grF * nr;
    const double DP = duc * nz * nz + dvc * nr * nr;
    const double DN = duo * nz * nz + dvo * nr * nr;
    const double Df = w * DP + (1.0 - w) * DN;
    const double un = unLin - Df * (gPN - gFN);
    s.mDot[f] = ph.rho * un * area;
}

// createPPRhs, simple.cu:90-142
// CROSS: with the non-orthogonal cross term (the corrector pass).
// zero: when set, also p' = 0 there (the solve's initial guess,
// the memset at solver.cpp:1193), saving a fill launch.
template <int B, bool CROSS>
__global__ void __launch_bounds__(kBlock) kPPRhs(DMesh m, DState s, Physics ph, double* __restrict__ zero) {
    constexpr bool cross = CROSS;
    if (finished(s.ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    if (zero) zero[n] = 0.0;
    double imb = 0.0, cs = 0.0;
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
    const double DUn = cross ? s.DU[n] : 0.0, DVn = cross ? s.DV[n] : 0.0;
    const double gzN = cross ? s.gPZ[n] : 0.0, grN = cross ? s.gPR[n] : 0.0;
    auto step = [&](int o, uint8_t fl, double md, double nz, double nr, double w, double tz, double tr, double DUo,
                    double DVo, double gzO, double grO) {
        if (fl & 1) imb += md;
        else imb -= md;
        if (cross) {
            const double Df = normalCorrVal(DUn, DVn, DUo, DVo, nz, nr, w, o);
            cs += (o >= 0) ? nonOrthoVal(tz, tr, w, gzN, gzO, grN, grO, ph.rho * Df) : 0.0;
        }
    };
    if (ns > 0 && ns <= B) {
        int o[B], fid[B];
        uint8_t fl[B];
        double md[B], nz[B], nr[B], w[B], tz[B], tr[B];
        double du[B], dv[B], gz[B], gr[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            fid[q] = m.faceIDs[k];
            fl[q] = m.sFlags[k];
            o[q] = cross ? m.slotNb[k] : -1;
            nz[q] = cross ? m.sNZ[k] : 0.0;
            nr[q] = cross ? m.sNR[k] : 0.0;
            w[q] = cross ? m.sW[k] : 0.0;
            tz[q] = cross ? m.sTZ[k] : 0.0;
            tr[q] = cross ? m.sTR[k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int oc = o[q] >= 0 ? o[q] : n;
            md[q] = s.mDot[fid[q]];
            du[q] = cross ? s.DU[oc] : 0.0;
            dv[q] = cross ? s.DV[oc] : 0.0;
            gz[q] = cross ? s.gPZ[oc] : 0.0;
            gr[q] = cross ? s.gPR[oc] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns) step(o[q], fl[q], md[q], nz[q], nr[q], w[q], tz[q], tr[q], du[q], dv[q], gz[q], gr[q]);
    } else {
        for (int k = k0; k < k0 + ns; k++) {
            const int o = cross ? m.slotNb[k] : -1;
            const int oc = o >= 0 ? o : n;
            step(o, m.sFlags[k], s.mDot[m.faceIDs[k]], cross ? m.sNZ[k] : 0.0, cross ? m.sNR[k] : 0.0,
                 cross ? m.sW[k] : 0.0, cross ? m.sTZ[k] : 0.0, cross ? m.sTR[k] : 0.0, cross ? s.DU[oc] : 0.0,
                 cross ? s.DV[oc] : 0.0, cross ? s.gPZ[oc] : 0.0, cross ? s.gPR[oc] : 0.0);
        }
    }
    s.ppb[n] = -imb + cs;
    if (zero && s.ppb32) s.ppb32[n] = (float)(-imb + cs);
}

// computeGradient(p', ppBC) fused with updateMomentumPressure (simple.cu:166-179)
template <int B>
__global__ void __launch_bounds__(kBlock) kGradPPUpdate(DMesh m, DBC ppbc, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double gz, gr;
    cellGradient<B>(m, ppbc, s.pp, n, ph.gradScheme, gz, gr);
    s.gPZ[n] = gz;
    s.gPR[n] = gr;
    s.u[n] -= s.DU[n] * gz;
    s.v[n] -= s.DV[n] * gr;
    s.p[n] += ph.alphaP * s.pp[n];
}

// updateMassFlux, simple.cu:181-258
// An interior face's data are loaded in one round, unconditionally (a
// boundary face reads its owner twice), and the update runs on them in the
// order of normalCorrToFace and nonOrthoFlux seen from the owner (weight wP,
// the normal as stored, d = x_nb - x_owner); boundary faces take those helpers.
__global__ void __launch_bounds__(kBlock) kMassFluxUpdate(DMesh m, DBC bp, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int f = xcdBlock() * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const int ow = m.owner[f], nb = m.neighbor[f];
    const int o = nb >= 0 ? nb : ow;
    const double area = m.area[f], nz = m.nZ[f], nr = m.nR[f], w = m.wP[f], inv = m.invDPN[f];
    const double mdf = s.mDot[f];
    const double duc = s.DU[ow], dvc = s.DV[ow], ppc = s.pp[ow];
    const double duo = s.DU[o], dvo = s.DV[o], ppo = s.pp[o];
    const bool no = ph.nonOrtho != 0;
    const double zc = no ? m.cZ[ow] : 0.0, rc = no ? m.cR[ow] : 0.0, zo = no ? m.cZ[o] : 0.0, ro = no ? m.cR[o] : 0.0;
    const double gzc = no ? s.gPZ[ow] : 0.0, grc = no ? s.gPR[ow] : 0.0;
    const double gzo = no ? s.gPZ[o] : 0.0, gro = no ? s.gPR[o] : 0.0;
    if (area <= 1.0e-30) {
        s.mDot[f] = 0.0;
        return;
    }
    if (nb >= 0) {
        const double DP = duc * nz * nz + dvc * nr * nr;
        const double DN = duo * nz * nz + dvo * nr * nr;
        const double Df = w * DP + (1.0 - w) * DN;
        double md = mdf;
        md -= ph.rho * area * Df * (ppo - ppc) * inv;
        if (no) {
            const double dz = zo - zc, dr = ro - rc;
            const double aOverNd = area * inv;
            const double Tz = area * nz - aOverNd * dz;
            const double Tr = area * nr - aOverNd * dr;
            const double gfz = w * gzc + (1.0 - w) * gzo;
            const double gfr = w * grc + (1.0 - w) * gro;
            md -= (ph.rho * Df) * (Tz * gfz + Tr * gfr);
        }
        s.mDot[f] = md;
        return;
    }
    const double Df = normalCorrToFace(m, s, ow, f);
    const int g = m.group[f];
    if (g < 0 || g >= bp.nGroups) return;
    if (bp.type[g] == 0) {
        const double dPB = fabs((m.fZ[f] - m.cZ[ow]) * m.nZ[f] + (m.fR[f] - m.cR[ow]) * m.nR[f]);
        s.mDot[f] -= ph.rho * area * Df * (0.0 - s.pp[ow]) / dPB;
    }
}

// ===========================================================================
// Residuals: residualAll(u, v) (residuals.cuh:17-60) + continuityResidual
// (residuals.cu:5-35) + the first pass of the norm reductions, per workgroup.
// partial[b*9 + j]: 0 sum|ru| 1 sum ru^2 2 max|ru| 3 sum|AC_u u|
//                   4 sum|rv| 5 sum rv^2 6 max|rv| 7 sum|AC_v v|  8 sum|imbalance|
// ===========================================================================
// blockCells (batched runs): the cell range of each block, so that a block never
// straddles two cases and each case's blocks start at its first cell.
template <int B>
__global__ void __launch_bounds__(kBlock) kResidual(DMesh m, DState s, const int2* blockCells) {
    if (finished(s.ctl)) return;
    if (s.ctl->k % s.ctl->checkConv != 0) return;
    int n = xcdBlock() * blockDim.x + threadIdx.x, nEnd = m.N;
    if (blockCells) {
        const int2 r = blockCells[xcdBlock()];
        n = r.x + threadIdx.x;
        nEnd = r.y;
    }
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (n < nEnd) {
        double axu = s.uAC[n] * s.u[n];
        double axv = s.vAC[n] * s.v[n];
        v[3] = fabs(axu);
        v[7] = fabs(axv);
        double imb = 0.0;
        int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
        auto step = [&](int o, double a, double uo, double vo, uint8_t fl, double md) {
            if (o >= 0) {
                axu += a * uo;
                axv += a * vo;
            }
            if (fl & 1) imb += md;
            else imb -= md;
        };
        if (ns > 0 && ns <= B) {
            int o[B], fid[B];
            uint8_t fl[B];
            double a[B], uo[B], vo[B], md[B];
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int k = slotAt(k0, ns, q);
                o[q] = m.slotNb[k];
                a[q] = s.AFuv[k];
                fid[q] = m.faceIDs[k];
                fl[q] = m.sFlags[k];
            }
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int oc = o[q] >= 0 ? o[q] : n;
                uo[q] = s.u[oc];
                vo[q] = s.v[oc];
                md[q] = s.mDot[fid[q]];
            }
#pragma unroll
            for (int q = 0; q < B; q++)
                if (q < ns) step(o[q], a[q], uo[q], vo[q], fl[q], md[q]);
        } else {
            for (int k = k0; k < k0 + ns; k++) {
                const int o = m.slotNb[k];
                const int oc = o >= 0 ? o : n;
                step(o, s.AFuv[k], s.u[oc], s.v[oc], m.sFlags[k], s.mDot[m.faceIDs[k]]);
            }
        }
        const double ru = fabs(s.ub[n] - axu), rv = fabs(s.vb[n] - axv);
        v[0] = ru;
        v[1] = ru * ru;
        v[2] = ru;
        v[4] = rv;
        v[5] = rv * rv;
        v[6] = rv;
        v[8] = fabs(imb);
    }
    __shared__ double red[kBlock / 64][9];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j < 9; j++) {
        const double w = (j == 2 || j == 6) ? waveScan<true>(v[j]) : waveScan<false>(v[j]);
        if (lane == 63) red[wave][j] = w;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        const int j = threadIdx.x;
        double a = red[0][j];
        for (int w = 1; w < kBlock / 64; w++) a = (j == 2 || j == 6) ? fmax(a, red[w][j]) : a + red[w][j];
        s.partial[xcdBlock() * 9 + j] = a;
    }
}

// Final reduction + residualAllHost (residuals.cu:37-100) + residualsConverged
// (solver.cpp:59-72) + the loop bookkeeping of solver.cpp:1243-1283, on device.
// One workgroup of kFinishThreads: each thread loads the partials of four
// blocks in one round (indices clamped, extra loads unused), then each wave
// reduces by DPP and one thread per quantity combines the waves.
// SC: the scalars' partials too (8 more per block after the 9 * blocks of
// momentum and continuity: 9 sum|rT| 10 sum rT^2 11 max|rT| 12 sum|AC_T T|,
// 13-16 the same of the concentration).
constexpr int kFinishThreads = 1024;
__device__ __forceinline__ constexpr bool isMaxQuantity(int j) { return j == 2 || j == 6 || j == 11 || j == 15; }
// The partials of blocks [b0 .. b0 + blocks) reduced to tot[0 .. 16] by one
// workgroup of kFinishThreads: each thread loads the partials of four blocks in
// one round (indices clamped, extra loads unused), each wave reduces by DPP and
// one thread per quantity combines the waves.  Every workgroup thread calls it.
template <bool SC>
__device__ void reducePartials(const double* partial, const double* ps, int b0, int blocks, double* tot) {
    constexpr int Q = SC ? 17 : 9;
    constexpr int R = SC ? 1 : 2;  // blocks per thread and round (R * Q doubles in registers, no spill)
    __shared__ double red[kFinishThreads / 64][Q];
    const int t = threadIdx.x;
    double a[Q];
#pragma unroll
    for (int j = 0; j < Q; j++) a[j] = 0.0;
    for (int bb = 0; bb < blocks; bb += R * kFinishThreads) {
        double x[R][Q];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int b = b0 + min(bb + r * kFinishThreads + t, blocks - 1);
#pragma unroll
            for (int j = 0; j < 9; j++) x[r][j] = partial[(size_t)b * 9 + j];
#pragma unroll
            for (int j = 9; j < Q; j++) x[r][j] = ps[(size_t)b * 8 + (j - 9)];
        }
#pragma unroll
        for (int r = 0; r < R; r++)
            if (bb + r * kFinishThreads + t < blocks)
#pragma unroll
                for (int j = 0; j < Q; j++) a[j] = isMaxQuantity(j) ? fmax(a[j], x[r][j]) : a[j] + x[r][j];
    }
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        const double w = isMaxQuantity(j) ? waveScan<true>(a[j]) : waveScan<false>(a[j]);
        if (lane == 63) red[wave][j] = w;
    }
    __syncthreads();
    if (t < Q) {
        double v = red[0][t];
        for (int w = 1; w < kFinishThreads / 64; w++) v = isMaxQuantity(t) ? fmax(v, red[w][t]) : v + red[w][t];
        tot[t] = v;
    }
    if (!SC && t >= 9 && t < 17) tot[t] = 0.0;
    __syncthreads();
}

// residualAllHost (residuals.cu:37-100) for the enabled residuals from the
// reduced sums, N cells; scaleVal keeps the continuity scale captured while
// scaleIter < 5; then residualsConverged (solver.cpp:59-72)
__device__ bool normaliseResiduals(const Ctl& cc, const double* tot, int N, int scaleIter, double* resVal,
                                   double* scaleVal) {
    for (int i = 0; i < 5; i++) {
        if (!cc.enabled[i]) continue;
        double res, scale = scaleVal[i];
        if (i == 2) {
            res = tot[8];
        } else {
            const int o = i < 2 ? i * 4 : 9 + (i - 3) * 4;
            res = cc.norm[i] == 0 ? tot[o] : cc.norm[i] == 1 ? tot[o + 1] : tot[o + 2];
            switch (cc.scaling[i]) {
            case 0: scale = 1.0; break;
            case 1: scale = (double)N; break;
            case 2: scale = sqrt((double)N); break;
            case 3: scale = tot[o + 3]; break;
            default: break;
            }
        }
        if (i == 2 && scaleIter < 5) scale = fmax(res, 0.0);
        scaleVal[i] = scale;
        if (scale != 0.0) {
            if (cc.norm[i] == 1) res = sqrt(res);
            res /= scale;
        }
        resVal[i] = res;
    }
    bool any = false, ok = true;
    for (int i = 0; i < 5; i++) {
        if (!cc.enabled[i]) continue;
        any = true;
        if (!(resVal[i] <= cc.tol[i])) ok = false;
    }
    return any && ok;
}

// The control block into LDS, one 32-bit word per thread (visible after the
// caller's next barrier).
__device__ __forceinline__ void copyCtl(Ctl& dst, const Ctl* src) {
    static_assert(sizeof(Ctl) % 4 == 0, "the control block is copied in 32-bit words");
    constexpr int W = (int)(sizeof(Ctl) / 4);
    for (int i = threadIdx.x; i < W; i += blockDim.x) reinterpret_cast<int*>(&dst)[i] = reinterpret_cast<const int*>(src)[i];
}

template <bool SC>
__global__ void __launch_bounds__(kFinishThreads) kFinish(DState s, int blocks) {
    Ctl* c = s.ctl;
    if (c->done) return;
    const bool check = (c->k % c->checkConv) == 0;
    // the control block copied into LDS by the workgroup (a copy in every
    // thread's registers spilled); thread 0 works on it after the reduction's
    // barriers and stores back the fields it changes
    __shared__ Ctl sc;
    __shared__ double tot[17];
    copyCtl(sc, c);
    if (check) reducePartials<SC>(s.partial, s.partial + (size_t)blocks * 9, 0, blocks, tot);
    else __syncthreads();
    if (threadIdx.x != 0) return;
    Ctl& cc = sc;
    if (check) {
        const int scaleIter = cc.transient ? cc.k : cc.iter;
        cc.converged = normaliseResiduals(cc, tot, cc.N, scaleIter, cc.resVal, cc.scaleVal) ? 1 : 0;
        if (cc.nHist < cc.histCap) {
            for (int i = 0; i < 5; i++) s.hist[cc.nHist * kHistStride + i] = cc.resVal[i];
            s.hist[cc.nHist * kHistStride + 5] = (double)cc.iter;
            cc.nHist++;
        }
    }
    cc.iter++;
    cc.k++;
    if (cc.converged || cc.k >= cc.maxIter) cc.done = 1;
    for (int i = 0; i < 5; i++) {
        c->resVal[i] = cc.resVal[i];
        c->scaleVal[i] = cc.scaleVal[i];
    }
    c->nHist = cc.nHist;
    c->converged = cc.converged;
    c->iter = cc.iter;
    c->k = cc.k;
    c->done = cc.done;
}

// Batched runs, one workgroup per case: the case's residual blocks reduced as
// a run of that case alone reduces them (its blocks start at its first cell),
// normalised with its own cell count and its own continuity scale, and its
// own convergence test.  A case already done this time step is left as it is.
template <bool SC>
__global__ void __launch_bounds__(kFinishThreads) kFinishCases(DState s, DCases cs, int blocks) {
    if (s.ctl->done) return;
    const int ci = blockIdx.x;
    CaseCtl& k = cs.ctl[ci];
    if (k.done) return;  // uniform over the workgroup
    const bool check = (s.ctl->k % s.ctl->checkConv) == 0;
    __shared__ Ctl sc;  // as in kFinish
    __shared__ double tot[17];
    copyCtl(sc, s.ctl);
    const int b0 = cs.blockStart[ci], nb = cs.blockStart[ci + 1] - b0;
    if (check) reducePartials<SC>(s.partial, s.partial + (size_t)blocks * 9, b0, nb, tot);
    else __syncthreads();
    if (threadIdx.x != 0) return;
    const Ctl& cc = sc;
    if (check) {
        const int scaleIter = cc.transient ? cc.k : cc.iter;
        double res[5], scale[5];
        for (int i = 0; i < 5; i++) {
            res[i] = k.resVal[i];
            scale[i] = k.scaleVal[i];
        }
        k.converged = normaliseResiduals(cc, tot, cs.cellCount[ci], scaleIter, res, scale) ? 1 : 0;
        for (int i = 0; i < 5; i++) {
            k.resVal[i] = res[i];
            k.scaleVal[i] = scale[i];
        }
    }
    k.iters++;
    if (k.converged || cc.k + 1 >= cc.maxIter) {
        k.done = 1;
        k.stopIter = cc.iter + 1;
    }
}

// After kFinishCases: the run's bookkeeping.  The history records, per check,
// the largest of each residual over the cases that ran this iteration.
__global__ void kFinishRun(DState s, DCases cs) {
    Ctl* c = s.ctl;
    Ctl cc = *c;
    if (cc.done) return;
    const bool check = (cc.k % cc.checkConv) == 0;
    bool all = true, allConverged = true;
    double worst[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < cs.n; i++) {
        const CaseCtl& k = cs.ctl[i];
        all = all && k.done;
        allConverged = allConverged && k.converged;
        if (!k.frozen)
            for (int j = 0; j < 5; j++) worst[j] = fmax(worst[j], k.resVal[j]);
    }
    if (check && cc.nHist < cc.histCap) {
        for (int j = 0; j < 5; j++) s.hist[cc.nHist * kHistStride + j] = worst[j];
        s.hist[cc.nHist * kHistStride + 5] = (double)cc.iter;
        c->nHist = cc.nHist + 1;
    }
    for (int j = 0; j < 5; j++) c->resVal[j] = worst[j];
    c->converged = allConverged ? 1 : 0;
    c->iter = cc.iter + 1;
    c->k = cc.k + 1;
    c->done = (all || cc.k + 1 >= cc.maxIter) ? 1 : 0;
}

// Cases done keep their fields: the frozen ones get them back from the
// snapshot, the ones done this iteration have them copied into it.  Runs in
// every iteration (also after the run's last one, where it changes nothing).
__global__ void __launch_bounds__(kBlock) kFreeze(const int* caseOf, const CaseCtl* ctl, int n, FreezeSet set) {
    const int i = xcdBlock() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const CaseCtl& k = ctl[caseOf[i]];
    if (k.frozen) {
        for (int a = 0; a < set.n; a++) set.field[a][i] = set.snap[a][i];
    } else if (k.done) {
        for (int a = 0; a < set.n; a++) set.snap[a][i] = set.field[a][i];
    }
}

__global__ void kMarkFrozen(CaseCtl* ctl, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ctl[i].done) ctl[i].frozen = 1;
}

// Start of a time step: freeze the old levels (solver.cpp:1058-1095) and reset
// the per-step loop state.  Copies instead of the reference's pointer swap
// keep the graph's addresses fixed; the values are the same.
__global__ void __launch_bounds__(kBlock) kStepBegin(DMesh m, DState s, int bdf2) {
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n < m.N) {
        if (bdf2) {
            s.uOld2[n] = s.uOld[n];
            s.vOld2[n] = s.vOld[n];
        }
        s.uOld[n] = s.u[n];
        s.vOld[n] = s.v[n];
    }
}

// ===========================================================================
// Scalar transport (energy, concentration): one pass per cell doing, in the
// reference's order, for one field,
//   addTransportCoefficients     solver_util.cu:218-448 (+ the wall kinetics
//                                of the Michaelis-Menten / Hill faces)
//   addTransientCoefficient      solver_util.cu:536-578 (capacity 1)
//   underRelaxEquation           solver_util.cu:582-600
//   buildInverseDiagonal         solver_util.cu:602-612
// (solver.cpp:1222-1235).  The system is the field's own (its diffusivity and
// the volumetric flux differ from the momentum's).
// ===========================================================================
namespace {

// concentration_equation.cu:6-88: rate laws of group g and their derivatives
struct WallKinetics {
    double vmax, km, n, m, v2, kmN, k2M;
    __device__ WallKinetics(const DBC& bc, int g)
        : vmax(bc.vmax[g]), km(bc.km[g]), n(bc.hillN[g]), m(bc.inhibM[g]), v2(bc.v2[g]), kmN(bc.kmN[g]),
          k2M(bc.k2M[g]) {}
    __device__ double mm(double c) const { return (vmax * c) / (km + c); }
    __device__ double dmm(double c) const {
        const double inv = 1.0 / (km + c);
        return vmax * km * inv * inv;
    }
    __device__ double inhib(double c) const {
        if (v2 == 0.0) return 1.0;
        const double cm = pow(c, m);
        return (1.0 - (v2 * cm / (k2M + cm)));
    }
    __device__ double dinhib(double c) const {
        if (v2 == 0.0) return 0.0;
        const double cm = pow(c, m), cm1 = pow(c, m - 1.0);
        return (-(m * v2 * k2M) * (cm1) / ((k2M + cm) * (k2M + cm)));
    }
    __device__ double hill(double c) const {
        const double cn = pow(c, n);
        return (vmax * cn) / (kmN + cn);
    }
    __device__ double dhill(double c) const {
        const double cn = pow(c, n), cn1 = pow(c, n - 1.0);
        return ((n * vmax * kmN) * (cn1) / ((kmN + cn) * (kmN + cn)));
    }
};

// wallConcentrationMichaelisMenten / Hill, concentration_equation.cu:91-160:
// Newton on h (c_P - c_w) = J(c_w) from c_w = c_P, at most 100 steps
__device__ double wallNewton(const WallKinetics& k, bool hill, double cp, double h) {
    double cw = cp;
#pragma unroll 1
    for (int i = 0; i < 100; i++) {
        const double rate = hill ? k.hill(cw) : k.mm(cw);
        const double inh = k.inhib(cw);
        const double J = rate * inh;
        const double dJ = rate * k.dinhib(cw) + (hill ? k.dhill(cw) : k.dmm(cw)) * inh;
        const double F = h * (cp - cw) - J;
        const double dF = -h - dJ;
        const double step = F / dF;
        cw -= step;
        cw = fmax(0.0, cw);
        if (fabs(step) < 1e-14 + 1e-10 * fabs(cw)) break;
    }
    return cw;
}

// boundary half of addTransportCoefficients for a scalar: boundaryTransport's
// types plus the reactive walls (solver_util.cu:383-412)
__device__ void boundaryTransportScalar(const DMesh& m, const DBC& bc, DWall wall, int n, int f, double area,
                                        double gamma, double F, double phiP, double& AC, double& b) {
    const int g = m.group[f];
    if (g >= 0 && g < bc.nGroups && (bc.type[g] == 3 || bc.type[g] == 4)) {
        const double dPF = fabs((m.fZ[f] - m.cZ[n]) * m.nZ[f] + (m.fR[f] - m.cR[n]) * m.nR[f]);
        const double Rtot = (dPF / gamma) + bc.rtot[g];
        const double h = 1 / Rtot;
        const WallKinetics kin(bc, g);
        const bool hill = bc.type[g] == 4;
        const double cw = wallNewton(kin, hill, phiP, h);
        wall.cw[f] = cw;
        wall.ocr[f] = area * (hill ? kin.hill(cw) : kin.mm(cw)) * kin.inhib(cw);
        AC += area * h;
        b += area * h * cw;
        return;  // kinetics walls add no convection (addConvectionContribution, solver_util.cuh:817-864)
    }
    boundaryTransport(m, bc, n, f, area, gamma, F, AC, b);
}

}  // namespace

template <int B, bool HO>
__global__ void __launch_bounds__(kBlock) kScalar(DMesh m, DBC bc, DScalar sc, DWall wall, const double* mDot,
                                                  ScalarPhysics ph, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    const double gamma = ph.gamma;
    double ac = 0.0, b = 0.0, netF = 0.0;
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
    const double pN = sc.phi[n];
    const double gzN = ph.nonOrtho ? sc.gZ[n] : 0.0, grN = ph.nonOrtho ? sc.gR[n] : 0.0;
    auto step = [&](int k, int o, uint8_t fl, double area, double invDPN, double md, double no, double w,
                    double pO) -> double {
        double F = 0.0;
        if (ph.conv) {
            const double Fo = md * ph.fluxScale;
            F = (fl & 1) ? Fo : -Fo;
            if (fabs(F) <= 1.0e-30) F = 0.0;
            netF += F;
        }
        if (o >= 0) {
            const double K = gamma * area * invDPN;
            double af = 0.0;
            ac += K;
            af -= K;
            if (ph.nonOrtho) b += no;
            if (F != 0.0) {
                ac += fmax(F, 0.0);
                af += fmin(F, 0.0);
                if (HO) {
                    const double ud = (F > 0.0) ? pN : pO;
                    const double ho = higherOrderVal(m, sc.phi, sc.gZ, sc.gR, ph.scheme, n, o, k, F, ud, w, pN, pO);
                    b -= F * (ho - ud);
                }
            }
            return af;
        }
        boundaryTransportScalar(m, bc, wall, n, m.faceIDs[k], area, gamma, F, pN, ac, b);
        return 0.0;
    };
    if (ns > 0 && ns <= B) {
        int o[B], fid[B];
        uint8_t fl[B];
        double area[B], inv[B], w[B], md[B], po[B], af[B], no[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = m.slotNb[k];
            fid[q] = m.faceIDs[k];
            fl[q] = m.sFlags[k];
            area[q] = m.sArea[k];
            inv[q] = m.sInvDPN[k];
            w[q] = (HO || ph.nonOrtho) ? m.sW[k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            const int oc = o[q] >= 0 ? o[q] : n;
            md[q] = ph.conv ? mDot[fid[q]] : 0.0;
            po[q] = HO ? sc.phi[oc] : 0.0;
            no[q] = ph.nonOrtho ? nonOrthoVal(m.sTZ[k], m.sTR[k], w[q], gzN, sc.gZ[oc], grN, sc.gR[oc], gamma) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns) af[q] = step(k0 + q, o[q], fl[q], area[q], inv[q], md[q], no[q], w[q], po[q]);
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns) sc.AF[k0 + q] = af[q];
    } else {
        for (int k = k0; k < k0 + ns; k++) {
            const int o = m.slotNb[k];
            const int oc = o >= 0 ? o : n;
            const double w = (HO || ph.nonOrtho) ? m.sW[k] : 0.0;
            const double no = ph.nonOrtho ? nonOrthoVal(m.sTZ[k], m.sTR[k], w, gzN, sc.gZ[oc], grN, sc.gR[oc], gamma)
                                          : 0.0;
            sc.AF[k] = step(k, o, m.sFlags[k], m.sArea[k], m.sInvDPN[k], ph.conv ? mDot[m.faceIDs[k]] : 0.0, no, w,
                            HO ? sc.phi[oc] : 0.0);
        }
    }
    if (ph.conv) ac -= netF;
    if (ph.transient && ph.dt > 0.0) {
        const double cV = 1.0 * m.vol[n];
        if (ctl != nullptr && ctl->bdf2Step) {
            const double a = cV / (2.0 * ph.dt);
            ac += 3.0 * a;
            b += a * (4.0 * sc.old[n] - sc.old2[n]);
        } else {
            const double a = cV / ph.dt;
            ac += a;
            b += a * sc.old[n];
        }
    }
    const double al = ph.alpha;
    if (!(fabs(ac) < 1.0e-30)) {
        const double a0 = ac;
        ac = a0 / al;
        b += ((1.0 - al) / al) * a0 * pN;
    }
    sc.AC[n] = ac;
    sc.b[n] = b;
    sc.inv[n] = (fabs(ac) < 1.0e-30) ? 0.0 : 1.0 / ac;
}

// residualRaw (residuals.cuh:17-48) of the solved scalars' systems and the first
// pass of their norms: per workgroup 0 sum|r| 1 sum r^2 2 max|r| 3 sum|AC x|, of
// the temperature, then 4-7 of the concentration (zero for a field not solved)
template <int B>
__global__ void __launch_bounds__(kBlock) kResidualScalars(DMesh m, DScalar t, DScalar c, double* partial,
                                                           const Ctl* ctl, const int2* blockCells) {
    if (finished(ctl)) return;
    if (ctl->k % ctl->checkConv != 0) return;
    int n = xcdBlock() * blockDim.x + threadIdx.x, nEnd = m.N;
    if (blockCells) {
        const int2 r = blockCells[xcdBlock()];
        n = r.x + threadIdx.x;
        nEnd = r.y;
    }
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (n < nEnd) {
        int k0, ns;
        rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
        const DScalar* sys[2] = {&t, &c};
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const DScalar& x = *sys[i];
            if (!x.phi) continue;
            double ax = x.AC[n] * x.phi[n];
            v[i * 4 + 3] = fabs(ax);
            for (int k = k0; k < k0 + ns; k++) {
                const int o = m.slotNb[k];
                if (o >= 0) ax += x.AF[k] * x.phi[o];
            }
            const double r = fabs(x.b[n] - ax);
            v[i * 4 + 0] = r;
            v[i * 4 + 1] = r * r;
            v[i * 4 + 2] = r;
        }
    }
    __shared__ double red[kBlock / 64][8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j < 8; j++) {
        const double w = (j == 2 || j == 6) ? waveScan<true>(v[j]) : waveScan<false>(v[j]);
        if (lane == 63) red[wave][j] = w;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int j = threadIdx.x;
        double a = red[0][j];
        for (int w = 1; w < kBlock / 64; w++) a = (j == 2 || j == 6) ? fmax(a, red[w][j]) : a + red[w][j];
        partial[xcdBlock() * 8 + j] = a;
    }
}

// Old levels of a scalar at the start of a time step (solver.cpp:1058-1095).
__global__ void __launch_bounds__(kBlock) kRotate(int n, double* old2, double* old, const double* phi, int bdf2) {
    const int i = xcdBlock() * blockDim.x + threadIdx.x;
    if (i < n) {
        if (bdf2) old2[i] = old[i];
        old[i] = phi[i];
    }
}

__global__ void __launch_bounds__(kBlock) kCopy(double* dst, const double* src, int n, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = xcdBlock() * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void __launch_bounds__(kBlock) kFill(double* dst, double value, int n, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = xcdBlock() * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = value;
}

// Streams n doubles through the caches (a benchmark's cache flush): each
// thread sums a grid-stride share; a sum equal to `mark` (never, for the
// filled buffer the caller passes) is stored, so the loads stay.
__global__ void __launch_bounds__(kBlock) kStreamRead(const double* __restrict__ src, size_t n, double* out,
                                                      double mark) {
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc += src[i];
    if (acc == mark) out[blockIdx.x] = acc;
}

// ===========================================================================
// Agglomeration multigrid, FP64 (multigrid.cu:350-541)
// ===========================================================================

// Galerkin coarse operator (buildCoarseOperator, multigrid.cu:350-389) by
// gather, in two coalesced passes: (1) one thread per fine cell forms
// diag_n = AC[n] + its internal off-diagonals (slot order) into f.res; (2) one
// thread per coarse cell sums diag over its members (ascending fine index) and
// one thread per coarse slot sums its fine slots (ascending fine slot).  The
// same additions as the reference's per-fine-row scatter, in a fixed order.
// (One thread per coarse cell doing both passes is a chain of dependent loads
// per member and runs several times slower even on small levels.)
__global__ void __launch_bounds__(kBlock) kMGCoarseDiag(DLevel f, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= f.n) return;
    double diag = f.AC[n];
    for (int k = f.start[n]; k < f.start[n + 1]; k++) {
        if (f.nb[k] < 0) continue;
        if (f.slotToCoarseSlot[k] < 0) diag += f.AF[k];
    }
    f.res[n] = diag;
}

__global__ void __launch_bounds__(kBlock) kMGCoarseGather(DLevel f, DLevel c, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // gathers by coarse slot: no XCD ranges
    if (i < c.refs) {
        double a = 0.0;
        for (int j = f.cslotStart[i]; j < f.cslotStart[i + 1]; j++) a += f.AF[f.cslotFine[j]];
        c.AF[i] = a;
        if (c.AF32) c.AF32[i] = (float)a;
    } else if (i < c.refs + c.n) {
        const int r = i - c.refs;
        double acc = 0.0;
        for (int j = f.memberStart[r]; j < f.memberStart[r + 1]; j++) acc += f.res[f.members[j]];
        c.AC[r] = acc;
        if (c.AC32) c.AC32[r] = (float)acc;
    }
}

// A level's values in FP64 or in its FP32 copies (the mixed-precision cycle):
// the per-cell multigrid kernels below are templates on the value type.
namespace {
template <class T>
struct LV {
    const T *AC, *AF;
    T *b, *x, *xNew, *res;
};
template <class T>
__device__ __forceinline__ LV<T> view(const DLevel& L);
template <>
__device__ __forceinline__ LV<double> view<double>(const DLevel& L) {
    return LV<double>{L.AC, L.AF, L.b, L.x, L.xNew, L.res};
}
template <>
__device__ __forceinline__ LV<float> view<float>(const DLevel& L) {
    return LV<float>{L.AC32, L.AF32, L.b32, L.x32, L.xNew32, L.res32};
}
template <class T>
__device__ __forceinline__ T tabs(T v) {
    return v < T(0) ? -v : v;
}
// jacobiRow (multigrid.cu:447-484) applied to x = 0: every product with x is a
// (signed) zero, b - Ax == b, and 0 + w*b/AC is the reference's value bit for
// bit (a zero result comes out +0 either way).
template <class T>
__device__ __forceinline__ T sweepFromZero(T b, T AC, T w) {
    if (tabs(AC) < T(1.0e-30)) return T(0);
    return T(0) + w * b / AC;
}
}  // namespace

// kMGCoarseDiag and kMGCoarseGather in one launch, over the same workgroups
// of whole aggregates as kMGResidualRestrict (f.aggBlocks / f.aggThreads):
// the workgroup's threads stride over the coarse slots [cs0, cs1) of its
// coarse cells (kMGCoarseGather's sums); thread t forms the diagonal of the
// run's member t (kMGCoarseDiag's sum) into LDS, and after the barrier sums
// coarse cell j0 + t's members from LDS.
template <int B>
__global__ void __launch_bounds__(kBlock) kMGCoarseOperator(DLevel f, DLevel c, const Ctl* ctl) {
    __shared__ double diag[kBlock];
    if (finished(ctl)) return;
    const int blk = xcdBlock(), t = threadIdx.x;
    const AggThread me = f.aggThreads[(size_t)blk * kBlock + t];
    const AggBlock ab = f.aggBlocks[blk];
    // The coarse off-diagonals first: they do not read the diagonals, so their
    // loads overlap the diagonal phase below.  Two coarse slots per thread and
    // round, the first four fine slots of each loaded at once (indices clamped
    // into range), summed in ascending fine slot order, longer sums continued
    // by the loop after.
    for (int i = ab.cs0 + t; i < ab.cs1; i += 2 * kBlock) {
        const bool two = i + kBlock < ab.cs1;
        const int ic[2] = {i, two ? i + kBlock : i};
        int q0[2], nq[2], fs[2][4];
        double v[2][4], a[2];
#pragma unroll
        for (int p = 0; p < 2; p++) {
            q0[p] = f.cslotStart[ic[p]];
            nq[p] = f.cslotStart[ic[p] + 1] - q0[p];
        }
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int q = 0; q < 4; q++) fs[p][q] = f.cslotFine[nq[p] > 0 ? slotAt(q0[p], nq[p], q) : 0];
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int q = 0; q < 4; q++) v[p][q] = f.AF[fs[p][q]];
#pragma unroll
        for (int p = 0; p < 2; p++) {
            a[p] = 0.0;
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (q < nq[p]) a[p] += v[p][q];
            for (int q = q0[p] + 4; q < q0[p] + nq[p]; q++) a[p] += f.AF[f.cslotFine[q]];
        }
        c.AF[i] = a[0];
        if (c.AF32) c.AF32[i] = (float)a[0];
        if (two) {
            c.AF[i + kBlock] = a[1];
            if (c.AF32) c.AF32[i + kBlock] = (float)a[1];
        }
    }
    if (me.m >= 0) {
        const int n = me.m, k0 = me.k0, ns = me.ns;
        double d = f.AC[n];
        if (B > 0 && ns > 0) {
            // the row's first B slots in one round (clamped), summed in slot
            // order; a longer row continues slot by slot
            int o[B > 0 ? B : 1], cs[B > 0 ? B : 1];
            double a[B > 0 ? B : 1];
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int k = slotAt(k0, ns, q);
                o[q] = f.nb[k];
                cs[q] = f.slotToCoarseSlot[k];
                a[q] = f.AF[k];
            }
#pragma unroll
            for (int q = 0; q < B; q++)
                if (q < ns && o[q] >= 0 && cs[q] < 0) d += a[q];
            for (int k = k0 + B; k < k0 + ns; k++)
                if (f.nb[k] >= 0 && f.slotToCoarseSlot[k] < 0) d += f.AF[k];
        } else {
            for (int k = k0; k < k0 + ns; k++) {
                if (f.nb[k] < 0) continue;
                if (f.slotToCoarseSlot[k] < 0) d += f.AF[k];
            }
        }
        diag[t] = d;
    }
    __syncthreads();
    if (me.sum >= 0) {
        const int j = ab.j0 + t;
        const int a = me.sum & 0xffff, e = me.sum >> 16;
        double acc = 0.0;
        for (int q = a; q < e; q++) acc += diag[q];
        c.AC[j] = acc;
        if (c.AC32) c.AC32[j] = (float)acc;
    }
}

// computeResidual (residualAll sign = true, residuals.cuh:17-60) into f.res,
// one thread per fine cell; kMGRestrictSum then gathers it per coarse cell.
// B > 0: a row's first B slots, then their neighbours, load in one round each
// (indices clamped into the row), and a longer row continues slot by slot
// (the multigrid kernels take B = 3, 4, 7 or 8 by the level's longest row:
// coarse rows average 5-6 slots and rarely exceed 8); B = 0: the plain loop.  The sums run in slot order either way.
template <class T, int B>
__device__ __forceinline__ T mgResidualRowAt(const DLevel& f, const LV<T>& v, int n, int k0, int ns) {
    T Ax = v.AC[n] * v.x[n];
    if (B > 0 && ns > 0) {
        int o[B > 0 ? B : 1];
        T a[B > 0 ? B : 1], xo[B > 0 ? B : 1];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = f.nb[k];
            a[q] = v.AF[k];
        }
#pragma unroll
        for (int q = 0; q < B; q++) xo[q] = v.x[o[q] >= 0 ? o[q] : n];
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns && o[q] >= 0) Ax += a[q] * xo[q];
        for (int k = k0 + B; k < k0 + ns; k++) {  // slots past the batch, in order
            const int nb = f.nb[k];
            if (nb >= 0) Ax += v.AF[k] * v.x[nb];
        }
    } else {
        for (int k = k0; k < k0 + ns; k++) {
            const int nb = f.nb[k];
            if (nb >= 0) Ax += v.AF[k] * v.x[nb];
        }
    }
    return v.b[n] - Ax;
}
template <class T, int B>
__device__ __forceinline__ T mgResidualRow(const DLevel& f, const LV<T>& v, int n) {
    int k0, ns;
    rowRange(f.start, f.uniformRow, n, k0, ns);
    return mgResidualRowAt<T, B>(f, v, n, k0, ns);
}

template <class T, int B>
__global__ void __launch_bounds__(kBlock) kMGResidual(DLevel f, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= f.n) return;
    const LV<T> v = view<T>(f);
    v.res[n] = mgResidualRow<T, B>(f, v, n);
}

// jacobiFused, multigrid.cu:486-493 (B as in kMGResidual)
template <class T, int B>
__global__ void __launch_bounds__(kBlock) kMGJacobi(DLevel L, T w, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= L.n) return;
    const LV<T> v = view<T>(L);
    int k0, ns;
    rowRange(L.start, L.uniformRow, n, k0, ns);
    const T AC = v.AC[n], xn = v.x[n], bn = v.b[n];
    T r = xn;
    if (B > 0 && ns > 0) {
        int o[B > 0 ? B : 1];
        T a[B > 0 ? B : 1], xo[B > 0 ? B : 1];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = L.nb[k];
            a[q] = v.AF[k];
        }
#pragma unroll
        for (int q = 0; q < B; q++) xo[q] = v.x[o[q] >= 0 ? o[q] : n];
        if (!(tabs(AC) < T(1.0e-30))) {
            T Ax = AC * xn;
#pragma unroll
            for (int q = 0; q < B; q++)
                if (q < ns && o[q] != -1) Ax += a[q] * xo[q];
            for (int k = k0 + B; k < k0 + ns; k++) {  // slots past the batch, in order
                const int nb = L.nb[k];
                if (nb != -1) Ax += v.AF[k] * v.x[nb];
            }
            r = xn + w * (bn - Ax) / AC;
        }
    } else if (!(tabs(AC) < T(1.0e-30))) {
        T Ax = AC * xn;
        for (int k = k0; k < k0 + ns; k++) {
            const int nb = L.nb[k];
            if (nb != -1) Ax += v.AF[k] * v.x[nb];
        }
        r = xn + w * (bn - Ax) / AC;
    }
    v.xNew[n] = r;
}

// jacobiSingleBlock, multigrid.cu:501-541: all sweeps of a small level in one
// workgroup, x ping-ponged in LDS (up to 4096 cells)
__global__ void __launch_bounds__(1024) kMGSmoothBlock(DLevel L, double w, int sweeps, const Ctl* ctl) {
    if (finished(ctl)) return;
    extern __shared__ double sx[];
    double* xa = sx;
    double* xb = sx + L.n;
    for (int n = threadIdx.x; n < L.n; n += blockDim.x) xa[n] = L.x[n];
    __syncthreads();
    for (int it = 0; it < sweeps; it++) {
        for (int n = threadIdx.x; n < L.n; n += blockDim.x) {
            const double AC = L.AC[n];
            double r;
            if (fabs(AC) < 1.0e-30) {
                r = xa[n];
            } else {
                double Ax = AC * xa[n];
                for (int k = L.start[n]; k < L.start[n + 1]; k++) {
                    const int nb = L.nb[k];
                    if (nb != -1) Ax += L.AF[k] * xa[nb];
                }
                r = xa[n] + w * (L.b[n] - Ax) / AC;
            }
            xb[n] = r;
        }
        __syncthreads();
        double* t = xa;
        xa = xb;
        xb = t;
    }
    for (int n = threadIdx.x; n < L.n; n += blockDim.x) L.x[n] = xa[n];
}

// buildProlongationKernel (multigrid.cu:415-427) fused into the first
// post-smoothing sweep: x + P x_c is formed on the fly for the cell and its
// neighbours (the same sums the prolongation would store).
template <class T, int B>
__global__ void __launch_bounds__(kBlock) kMGProlongSmooth(DLevel f, DLevel c, T w, double* __restrict__ apply,
                                                           const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= f.n) return;
    const LV<T> v = view<T>(f);
    const LV<T> cv = view<T>(c);
    // the new iterate; with apply (the mixed cycle's level 0), also x += it in FP64
    auto put = [&](T r) {
        v.xNew[n] = r;
        if (apply) apply[n] += (double)r;
    };
    int k0, ns;
    rowRange(f.start, f.uniformRow, n, k0, ns);
    if (B > 0 && ns > 0) {
        int o[B > 0 ? B : 1], tc[B > 0 ? B : 1];
        T a[B > 0 ? B : 1], xo[B > 0 ? B : 1], co[B > 0 ? B : 1];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = f.nb[k];
            a[q] = v.AF[k];
        }
        const int tn = f.toCoarse[n];
        const T AC = v.AC[n], bn = v.b[n], xf = v.x[n];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int oc = o[q] >= 0 ? o[q] : n;
            xo[q] = v.x[oc];
            tc[q] = f.toCoarse[oc];
        }
        const T xc = cv.x[tn];
#pragma unroll
        for (int q = 0; q < B; q++) co[q] = cv.x[tc[q]];
        const T xn = xf + xc;
        if (tabs(AC) < T(1.0e-30)) {
            put(xn);
            return;
        }
        T Ax = AC * xn;
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns && o[q] != -1) Ax += a[q] * (xo[q] + co[q]);
        for (int k = k0 + B; k < k0 + ns; k++) {  // slots past the batch, in order
            const int nb = f.nb[k];
            if (nb != -1) Ax += v.AF[k] * (v.x[nb] + cv.x[f.toCoarse[nb]]);
        }
        put(xn + w * (bn - Ax) / AC);
        return;
    }
    const T xn = v.x[n] + cv.x[f.toCoarse[n]];
    const T AC = v.AC[n];
    if (tabs(AC) < T(1.0e-30)) {
        put(xn);
        return;
    }
    T Ax = AC * xn;
    for (int k = k0; k < k0 + ns; k++) {
        const int nb = f.nb[k];
        if (nb != -1) Ax += v.AF[k] * (v.x[nb] + cv.x[f.toCoarse[nb]]);
    }
    put(xn + w * (v.b[n] - Ax) / AC);
}

// Level-0 sweeps of the V-cycles fused per launch (temporal blocking, as in
// kJacobiUVBlock): workgroup b owns the C cells [bC, bC + C) of level 0 and
// runs, with the halo rings of the block recomputed and the iterates in LDS,
//   P: the prolongation x + P x_c as the first sweep's input (the post-sweep
//      that ends a cycle, as in kMGProlongSmooth),
//   S weighted-Jacobi sweeps (jacobiRow, multigrid.cu:447-484),
//   R: then the residual b - A x of the block's cells into f.res (kMGResidual).
// The final x goes to f.xNew.  One launch replaces the post-sweep of cycle c,
// the pre-sweep of cycle c + 1 and its residual, each row evaluated as in the
// one-pass kernels, so the result is bit-identical.
template <class T, int R>
struct MGRowReg {
    T af[R];
    int nb[R];
    int loc[R];  // LDS slots as in SweepRow
    int k0, ns;
    T ac, b;
};

template <class T, int R>
__device__ __forceinline__ void loadMGRow(MGRowReg<T, R>& r, const DLevel& L, const LV<T>& v, int c) {
    int k0, ns;
    rowRange(L.start, L.uniformRow, c, k0, ns);
    r.k0 = k0;
    r.ns = ns;
    // clamped, unconditional reads as in loadSweepRow
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int k = ns > 0 ? k0 + min(q, ns - 1) : 0;
        const int nb = L.nb[k];
        const T af = v.AF[k];
        r.nb[q] = (q < ns) ? nb : -1;
        r.af[q] = (q < ns) ? af : T(0);
    }
    r.ac = v.AC[c];
    r.b = v.b[c];
}

template <class T, int R, bool LDS = false, typename XAt>
__device__ __forceinline__ T mgJacobiReg(const MGRowReg<T, R>& r, XAt xat, T xn, T w) {
    if (tabs(r.ac) < T(1.0e-30)) return xn;
    T xv[R];
#pragma unroll
    for (int q = 0; q < R; q++) xv[q] = xat(q);  // slots without a neighbour: read, not used (as in sweepRow)
    T Ax = r.ac * xn;
#pragma unroll
    for (int q = 0; q < R; q++)
        if (LDS ? r.loc[q] >= 0 : r.nb[q] != -1) Ax += r.af[q] * xv[q];
    return xn + w * (r.b - Ax) / r.ac;
}

// The first launch of a solve refined once (FP32 cycles on the correction e,
// e = 0 at the start) takes its right-hand side in one of two ways:
//   kL0Zero:   x = 0 and b32 already holds b (written with the right-hand side);
//              no iterate is read, every x is the zero it would read;
//   kL0Defect: the rows' right-hand side is the FP64 defect b - A x, formed from
//              the FP64 row and x as kMGDefect forms it, rounded to FP32 (and
//              written to b32 for the block's own cells); the FP32 row is the
//              FP64 one rounded, as kMGCast writes it.
enum { kL0X = 0, kL0Zero = 1, kL0Defect = 2 };

template <int R>
__device__ __forceinline__ void loadMGRowDefect(MGRowReg<float, R>& r, const DLevel& L, int c) {
    int k0, ns;
    rowRange(L.start, L.uniformRow, c, k0, ns);
    r.k0 = k0;
    r.ns = ns;
    double a[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int k = ns > 0 ? k0 + min(q, ns - 1) : 0;
        const int nb = L.nb[k];
        const double af = L.AF[k];
        r.nb[q] = (q < ns) ? nb : -1;
        a[q] = (q < ns) ? af : 0.0;
    }
    const double ac = L.AC[c];
    double xo[R];
#pragma unroll
    for (int q = 0; q < R; q++) xo[q] = L.x[r.nb[q] >= 0 ? r.nb[q] : c];
    double Ax = ac * L.x[c];
#pragma unroll
    for (int q = 0; q < R; q++)
        if (r.nb[q] >= 0) Ax += a[q] * xo[q];
    r.b = (float)(L.b[c] - Ax);
    r.ac = (float)ac;
#pragma unroll
    for (int q = 0; q < R; q++) r.af[q] = (float)a[q];
}

// held to 64 VGPRs (two 1024-thread blocks per CU), which these rows fit
// without spilling.  T: FP64, or the FP32 copies (the mixed path's cycles).
template <class T, int C, int S, bool P, bool RES, int R, int X0 = kL0X>
__global__ void __launch_bounds__(C) __attribute__((amdgpu_waves_per_eu(8)))
kMGLevel0Block(DLevel f, DLevel c, DSweepBlocks sb, T w, const Ctl* ctl) {
    static_assert(X0 == kL0X || (!P && sizeof(T) == 4), "zero / defect starts are FP32 first launches");
    constexpr int H = sweepHalo<C>();
    constexpr int kSets = (S - 1 + (RES ? 1 : 0)) > 1 ? 2 : 1;  // LDS iterate buffers in use
    __shared__ T buf[kSets][C + H];
    const LV<T> v = view<T>(f);
    const LV<T> cv = view<T>(c);
    if (finished(ctl)) return;
    const int t = threadIdx.x, blk = xcdBlock();
    const int a = blk * C, nOwn = min(C, f.n - a);
    const int h0 = sb.haloStart[blk];
    // halo entries the first sweep updates: rings 1 .. S - 1 + RES; of those,
    // rings 1 .. S - 2 + RES are read from LDS by the later sweeps
    constexpr int kRings1 = S - 1 + (RES ? 1 : 0);
    const int nHalo1 = kRings1 > 0 ? sb.ringEnd[blk * sb.rings + kRings1 - 1] : 0;
    const int nHaloLds = (kRings1 > 1) ? sb.ringEnd[blk * sb.rings + (kRings1 > 1 ? kRings1 - 2 : 0)] : 0;
    const int n = sb.own ? sb.own[(size_t)blk * C + t] : a + t;
    const bool mine = sb.own ? n >= 0 : t < nOwn, haloRow = t < nHalo1;
    const int hc = haloRow ? sb.halo[h0 + t] : 0;
    MGRowReg<T, R> own, hal;
    if constexpr (X0 == kL0Defect) {
        if (mine) {
            loadMGRowDefect(own, f, n);
            v.b[n] = own.b;
        }
        if (haloRow) loadMGRowDefect(hal, f, hc);
    } else {
        if (mine) loadMGRow(own, f, v, n);
        if (haloRow) loadMGRow(hal, f, v, hc);
    }
    // the first sweep's input at any cell: x, plus the prolonged correction
    auto x0 = [&](int cell) -> T {
        if constexpr (X0 != kL0X) return T(0);
        else return P ? v.x[cell] + cv.x[f.toCoarse[cell]] : v.x[cell];
    };
    // sweep j (1-based) writes LDS buffer (j - 1) % kSets; the last one without
    // a residual writes f.xNew directly
    T lastOwn = T(0);
    {
        const T vo = mine ? mgJacobiReg(own, [&](int q) { return x0(own.nb[q] != -1 ? own.nb[q] : n); }, x0(n), w)
                          : T(0);
        const T vh =
            haloRow ? mgJacobiReg(hal, [&](int q) { return x0(hal.nb[q] != -1 ? hal.nb[q] : hc); }, x0(hc), w)
                    : T(0);
        if (S == 1 && !RES) {
            if (mine) v.xNew[n] = vo;
            return;
        }
        if (mine) buf[0][t] = vo;
        if (haloRow) buf[0][C + t] = vh;
        lastOwn = vo;
    }
    // the LDS slots the later sweeps and the residual read, loaded while the
    // barrier waits
    if (mine) loadLoc(own, sb.locOwn + own.k0, R);
    if (t < nHaloLds) loadLoc(hal, sb.locHalo + (size_t)(h0 + t) * kSweepRegSlots, R);
    __syncthreads();
#pragma unroll
    for (int j = 2; j <= S; j++) {
        const T* in = buf[(j - 2) % kSets];
        T* out = buf[(j - 1) % kSets];
        const T vo =
            mine ? mgJacobiReg<T, R, true>(own, [&](int q) { return in[own.loc[q] >= 0 ? own.loc[q] : t]; }, in[t], w)
                 : T(0);
        const int nHaloJ = (S - j + (RES ? 1 : 0)) > 0 ? sb.ringEnd[blk * sb.rings + (S - j + (RES ? 1 : 0)) - 1] : 0;
        const T vh = (t < nHaloJ) ? mgJacobiReg<T, R, true>(
                                        hal, [&](int q) { return in[hal.loc[q] >= 0 ? hal.loc[q] : C + t]; },
                                        in[C + t], w)
                                  : T(0);
        if (j == S && !RES) {
            if (mine) v.xNew[n] = vo;
            return;
        }
        if (mine) out[t] = vo;
        if (t < nHaloJ) out[C + t] = vh;
        lastOwn = vo;
        __syncthreads();
    }
    // residual (kMGResidual) from the last sweep's iterate
    if (RES && mine) {
        const T* xs = buf[(S - 1) % kSets];
        v.xNew[n] = lastOwn;
        T xv[R];
#pragma unroll
        for (int q = 0; q < R; q++) xv[q] = xs[own.loc[q] >= 0 ? own.loc[q] : t];
        T Ax = own.ac * lastOwn;
#pragma unroll
        for (int q = 0; q < R; q++)
            if (own.loc[q] >= 0) Ax += own.af[q] * xv[q];
        v.res[n] = own.b - Ax;
    }
}

__global__ void __launch_bounds__(kBlock) kMGProlong(DLevel f, DLevel c, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= f.n) return;
    f.x[n] += c.x[f.toCoarse[n]];
}

// ===========================================================================
// Restriction, FP64 and FP32 (the mixed-precision cycle) from one template.
// ===========================================================================

// buildRestrictionKernel (multigrid.cu:398-410) as a gather in ascending fine
// order of the residuals kMGResidual left in f.res, then the coarse x memset
// (multigrid.cu:650-655) -- or, with sweep1, the coarse level's first
// pre-smoothing sweep, which starts from that zero.
template <class T, int B>
__global__ void __launch_bounds__(kBlock) kMGRestrictSum(DLevel f, DLevel c, double w, int sweep1, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = xcdBlock() * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const LV<T> fv = view<T>(f);
    const LV<T> cv = view<T>(c);
    T acc = T(0);
    const int j0 = f.memberStart[i], nm = f.memberStart[i + 1] - j0;
    if (B > 0 && nm > 0 && nm <= B) {
        // members, then their residuals, in one round each; summed in order
        int mb[B > 0 ? B : 1];
        T r[B > 0 ? B : 1];
#pragma unroll
        for (int q = 0; q < B; q++) mb[q] = f.members[slotAt(j0, nm, q)];
#pragma unroll
        for (int q = 0; q < B; q++) r[q] = fv.res[mb[q]];
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < nm) acc += r[q];
    } else {
        for (int j = j0; j < j0 + nm; j++) acc += fv.res[f.members[j]];
    }
    cv.b[i] = acc;
    cv.x[i] = sweep1 ? sweepFromZero(acc, cv.AC[i], (T)w) : T(0);
}

// kMGResidual and kMGRestrictSum in one launch.  Workgroup b covers the coarse
// cells [j0, j1) of f.aggBlocks[b]; their members (at most kBlock, one run of
// the member list) are its fine cells.  Thread t computes the residual of the
// run's member t (its row range from f.aggThreads) with kMGResidual's
// expression into LDS; after the barrier thread t sums coarse cell j0 + t's
// members from LDS in list order, as kMGRestrictSum does.  The fine residual
// never goes to memory.
template <class T, int B>
__global__ void __launch_bounds__(kBlock) kMGResidualRestrict(DLevel f, DLevel c, double w, int sweep1,
                                                              const Ctl* ctl) {
    __shared__ T r[kBlock];
    if (finished(ctl)) return;
    const int blk = xcdBlock(), t = threadIdx.x;
    const AggThread me = f.aggThreads[(size_t)blk * kBlock + t];
    const int j = f.aggBlocks[blk].j0 + t;
    const LV<T> fv = view<T>(f);
    const LV<T> cv = view<T>(c);
    const T acj = (sweep1 && me.sum >= 0) ? cv.AC[j] : T(0);  // read ahead of the barrier
    if (me.m >= 0) r[t] = mgResidualRowAt<T, B>(f, fv, me.m, me.k0, me.ns);
    __syncthreads();
    if (me.sum >= 0) {
        const int a = me.sum & 0xffff, e = me.sum >> 16;
        T acc = T(0);
        for (int q = a; q < e; q++) acc += r[q];
        cv.b[j] = acc;
        cv.x[j] = sweep1 ? sweepFromZero(acc, acj, (T)w) : T(0);
    }
}

// kMGRestrictSum into level 1 (the level-0 residuals gathered per level-1
// cell in member order, then the first level-1 sweep from zero) and level 1's
// kMGResidualRestrict into level 2, in one launch over level 1's aggregate
// workgroups (f = level 1).  Thread t forms b and x of its level-1 cell (as
// kMGRestrictSum does, and writes them) and of its halo cell, one of the
// level-1 cells the workgroup's rows read (t.halo); after the barrier the
// residual of its row reads the x it needs from LDS through the host-built
// slot table (t.loc), in slot order as in kMGResidual; after the next barrier
// the level-2 sums run as in kMGResidualRestrict.  Same values, bit for bit.
template <class T, int B>
__device__ __forceinline__ T memberSum(const DLevel& g, const T* res, int i) {
    const int j0 = g.memberStart[i], nm = g.memberStart[i + 1] - j0;
    int mb[B];
    T r[B];
#pragma unroll
    for (int q = 0; q < B; q++) mb[q] = g.members[nm > 0 ? slotAt(j0, nm, q) : 0];
#pragma unroll
    for (int q = 0; q < B; q++) r[q] = res[mb[q]];
    T acc = T(0);
#pragma unroll
    for (int q = 0; q < B; q++)
        if (q < nm) acc += r[q];
    for (int j = j0 + B; j < j0 + nm; j++) acc += res[g.members[j]];
    return acc;
}

template <class T>
__global__ void __launch_bounds__(kBlock) kMGRestrictFused01(DLevel g, DLevel f, DLevel c, L1Fused tb, T w,
                                                             int sweep1, int sweep2, const Ctl* ctl) {
    __shared__ T xs[2 * kBlock];
    __shared__ T r[kBlock];
    if (finished(ctl)) return;
    const LV<T> gv = view<T>(g);
    const LV<T> fv = view<T>(f);
    const LV<T> cv = view<T>(c);
    const int blk = xcdBlock(), t = threadIdx.x;
    const AggThread me = f.aggThreads[(size_t)blk * kBlock + t];
    const int hc = tb.halo[(size_t)blk * kBlock + t];
    const int j = f.aggBlocks[blk].j0 + t;
    T bOwn = T(0), xOwn = T(0);
    if (me.m >= 0) {
        bOwn = memberSum<T, 6>(g, gv.res, me.m);
        xOwn = sweep1 ? sweepFromZero(bOwn, fv.AC[me.m], w) : T(0);
        fv.b[me.m] = bOwn;
        fv.x[me.m] = xOwn;
        xs[t] = xOwn;
    }
    if (hc >= 0) {
        const T bh = memberSum<T, 6>(g, gv.res, hc);
        xs[kBlock + t] = sweep1 ? sweepFromZero(bh, fv.AC[hc], w) : T(0);
    }
    const T acj = (sweep2 && me.sum >= 0) ? cv.AC[j] : T(0);
    __syncthreads();
    if (me.m >= 0) {
        constexpr int K = kL1Slots;
        const int k0 = me.k0, ns = me.ns;
        int lc[K];
        T a[K], xv[K];
        const int16_t* l = tb.loc + ((size_t)blk * kBlock + t) * K;
#pragma unroll
        for (int q = 0; q < K; q++) {
            lc[q] = l[q];
            a[q] = fv.AF[ns > 0 ? slotAt(k0, ns, q) : 0];
        }
#pragma unroll
        for (int q = 0; q < K; q++) xv[q] = xs[lc[q] >= 0 ? lc[q] : t];
        T Ax = fv.AC[me.m] * xOwn;
#pragma unroll
        for (int q = 0; q < K; q++)
            if (q < ns && lc[q] >= 0) Ax += a[q] * xv[q];
        r[t] = bOwn - Ax;
    }
    __syncthreads();
    if (me.sum >= 0) {
        const int a = me.sum & 0xffff, e = me.sum >> 16;
        T acc = T(0);
        for (int q = a; q < e; q++) acc += r[q];
        cv.b[j] = acc;
        cv.x[j] = sweep2 ? sweepFromZero(acc, acj, w) : T(0);
    }
}

// ===========================================================================
// Multi-colour Gauss-Seidel smoothing of a level (build extension; the
// reference smooths with weighted Jacobi only).  Cells of one colour share no
// face, so each colour is a parallel in-place update:
//   x_n = (b_n - sum AF x_nb) / AC_n   (rows with |AC| < 1e-30 keep x_n)
// and the result does not depend on the order within a colour.
// ===========================================================================
namespace {
template <class T>
__device__ __forceinline__ T gsRow(const DLevel& L, const LV<T>& v, const T* x, int n) {
    T val = v.b[n];
    for (int k = L.start[n]; k < L.start[n + 1]; k++) {
        const int nb = L.nb[k];
        if (nb >= 0) val -= v.AF[k] * x[nb];
    }
    return val / v.AC[n];
}
}  // namespace

template <class T>
__global__ void __launch_bounds__(kBlock) kMGGSColor(DLevel L, const int* __restrict__ order, int begin, int count,
                                                     const Ctl* ctl) {
    if (finished(ctl)) return;
    const int t = xcdBlock() * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const LV<T> v = view<T>(L);
    const int n = order[begin + t];
    if (tabs(v.AC[n]) < T(1.0e-30)) return;
    v.x[n] = gsRow(L, v, v.x, n);
}

// every sweep and colour of a small level in one workgroup, x held in LDS
template <class T>
__global__ void __launch_bounds__(1024) kMGSmoothBlockGS(DLevel L, const int* __restrict__ order,
                                                         const int* __restrict__ colorStart, int nColors,
                                                         int sweeps, const Ctl* ctl) {
    if (finished(ctl)) return;
    extern __shared__ __align__(16) unsigned char smemGS[];
    T* x = (T*)smemGS;
    const LV<T> v = view<T>(L);
    for (int n = threadIdx.x; n < L.n; n += blockDim.x) x[n] = v.x[n];
    __syncthreads();
    for (int s = 0; s < sweeps; s++)
        for (int c = 0; c < nColors; c++) {
            const int c0 = colorStart[c], c1 = colorStart[c + 1];
            for (int t = c0 + (int)threadIdx.x; t < c1; t += blockDim.x) {
                const int n = order[t];
                if (!(tabs(v.AC[n]) < T(1.0e-30))) x[n] = gsRow(L, v, x, n);
            }
            __syncthreads();
        }
    for (int n = threadIdx.x; n < L.n; n += blockDim.x) v.x[n] = x[n];
}

// ===========================================================================
// Tail of the V-cycle: the coarsest levels, from `lt` down, in ONE workgroup
// with every one of them staged in LDS (operator, CSR, transfer maps, b and
// the x ping-pong).  Coarse levels are tiny and launch-latency bound; here a
// sweep is an LDS pass + barrier instead of a kernel launch.  Thread t owns
// cell t of every level (the host launches at least as many threads as the
// largest tail level has cells) and keeps that cell's row in registers, so a
// sweep waits for one round of independent LDS reads instead of a chain per
// slot.  Same arithmetic, in the same order, as kMGJacobi / kMGResidual +
// kMGRestrictSum / kMGProlongSmooth / the coarsest solve.  Level lt's x and b
// come in from its global buffers (x after `preDone` sweeps) and its answer
// goes back to its x buffer.
// ===========================================================================
namespace {
constexpr int kRowK = 10;     // slots a row keeps in registers (longer rows loop over LDS)
constexpr int kMemberK = 8;   // aggregate members gathered without a loop
constexpr int kCopyR = 8;     // staged 16-byte vectors each thread holds per round

template <class T>
struct TRow {
    T ac, b;
    int k0, deg;
    int nb[kRowK];
    T af[kRowK];
};

template <class T>
__device__ __forceinline__ void loadRow(TRow<T>& r, const T* AC, const T* b, const T* AF, const int* st,
                                        const int* nbs, int n) {
    r.ac = AC[n];
    r.b = b[n];
    r.k0 = st[n];
    r.deg = st[n + 1] - r.k0;
    // every slot read at an index clamped into the row, then masked: the reads
    // stay unconditional and go out together instead of one branch each
    int nv[kRowK];
    T av[kRowK];
#pragma unroll
    for (int k = 0; k < kRowK; k++) {
        const int kk = r.k0 + (k < r.deg ? k : 0);
        nv[k] = nbs[kk];
        av[k] = AF[kk];
    }
#pragma unroll
    for (int k = 0; k < kRowK; k++) {
        const bool in = k < r.deg;
        r.nb[k] = in ? nv[k] : -1;
        r.af[k] = in ? av[k] : T(0);
    }
}

// Ax of the row: AC*xn + AF*x(nb) over the slots with a neighbour, in slot
// order.  xat(j) gives the (possibly prolonged) x of cell j; self is the row's
// own cell, read (and not used) in place of missing neighbours so that every
// read is unconditional.
template <class T, class XAt>
__device__ __forceinline__ T rowAx(const TRow<T>& r, T xn, XAt xat, const T* AF, const int* nbs, int self) {
    T Ax = r.ac * xn;
    if (r.deg <= kRowK) {
        T xv[kRowK];
#pragma unroll
        for (int k = 0; k < kRowK; k++) xv[k] = xat(r.nb[k] != -1 ? r.nb[k] : self);
#pragma unroll
        for (int k = 0; k < kRowK; k++)
            if (r.nb[k] != -1) Ax += r.af[k] * xv[k];
    } else {
        for (int k = r.k0; k < r.k0 + r.deg; k++) {
            const int nb = nbs[k];
            if (nb != -1) Ax += AF[k] * xat(nb);
        }
    }
    return Ax;
}
}  // namespace

// jacobiRow of a register row with at most K slots, its neighbours read from
// x in one round (missing ones read the row's own entry and are not used)
template <int K, class T>
__device__ __forceinline__ T waveRowSweep(const TRow<T>& r, const T* x, T xn, T w, int self) {
    T xv[K];
#pragma unroll
    for (int k = 0; k < K; k++) xv[k] = x[r.nb[k] != -1 ? r.nb[k] : self];
    T Ax = r.ac * xn;
#pragma unroll
    for (int k = 0; k < K; k++)
        if (r.nb[k] != -1) Ax += r.af[k] * xv[k];
    return xn + w * (r.b - Ax) / r.ac;
}

template <class T>
__global__ void __launch_bounds__(1024) kMGTail(const TailLevel* __restrict__ tl, uint4* __restrict__ image,
                                                int vecs, int nl, int intBase, double wd, int preDone, int pre,
                                                int post, int coarseSweeps, const double* inv, const Ctl* ctl) {
    if (finished(ctl)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    T* sT = (T*)smem;
    int* sI = (int*)(smem + intBase);
    uint4* sV = (uint4*)smem;
    const T w = (T)wd;
    const int t = threadIdx.x, NT = blockDim.x;

    // ---- stage: the image is the LDS layout itself; each round loads all its
    // vectors before storing any, so they are in flight together
    for (int base = 0; base < vecs; base += kCopyR * NT) {
        uint4 v[kCopyR];
        // indices clamped into the image, for the loads and the stores alike:
        // both are unconditional, so the loads issue back to back (a guarded
        // load waits before the next one).  Lanes past the end store the last
        // vector again, with the value it already has.
#pragma unroll
        for (int r = 0; r < kCopyR; r++) v[r] = image[min(base + r * NT + t, vecs - 1)];
#pragma unroll
        for (int r = 0; r < kCopyR; r++) sV[min(base + r * NT + t, vecs - 1)] = v[r];
    }
    __syncthreads();

    unsigned flip = 0;  // bit q set: level q's current x is its x1 array
    auto X = [&](const TailLevel& S, int q) { return sT + (((flip >> q) & 1u) ? S.x1 : S.x0); };
    auto Y = [&](const TailLevel& S, int q) { return sT + (((flip >> q) & 1u) ? S.x0 : S.x1); };
    // jacobiRow sweeps of level q, the row held in registers across them
    auto sweeps = [&](int q, int count) {
        if (count <= 0) return;
        const TailLevel S = tl[q];
        const T* AF = sT + S.AF;
        const int* nbs = sI + S.nb;
        const bool own = t < S.n;
        // thread t sweeps cell perm[t] where the level has an order (the
        // coarsest, above one wave: cells by slot count), else cell t
        const int cell = (own && S.perm >= 0) ? sI[S.perm + t] : t;
        TRow<T> r;
        if (own) loadRow(r, sT + S.AC, sT + S.b, AF, sI + S.start, nbs, cell);
        if (S.n <= 64) {
            // a level of at most one wave's cells (the coarsest, usually):
            // wave 0 sweeps alone; a wave's LDS accesses complete in program
            // order, so its sweeps need no workgroup barrier in between
            if (t < 64) {
                unsigned f = flip;
                // rows of at most kWaveK slots (the coarsest level's are) sweep
                // with a shorter register row: fewer dependent steps per sweep
                // (6 or 8 slots, whichever covers the wave's longest row)
                constexpr int kWaveK = 8;
                static_assert(kWaveK <= kRowK, "a short row is a prefix of the register row");
                const bool shortRows = __all(!own || r.deg <= kWaveK);
                const bool shorterRows = __all(!own || r.deg <= 6);
                const bool keep = own && tabs(r.ac) < T(1.0e-30);
                for (int it = 0; it < count; it++) {
                    const T* x = sT + (((f >> q) & 1u) ? S.x1 : S.x0);
                    T* y = sT + (((f >> q) & 1u) ? S.x0 : S.x1);
                    if (own) {
                        const T xn = x[t];
                        T res;
                        if (keep) {
                            res = xn;
                        } else if (shorterRows) {
                            res = waveRowSweep<6>(r, x, xn, w, t);
                        } else if (shortRows) {
                            res = waveRowSweep<kWaveK>(r, x, xn, w, t);
                        } else {
                            const T Ax = rowAx(r, xn, [&](int j) { return x[j]; }, AF, nbs, t);
                            res = xn + w * (r.b - Ax) / r.ac;
                        }
                        y[t] = res;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    f ^= 1u << q;
                }
            }
            if (count & 1) flip ^= 1u << q;
            __syncthreads();
            return;
        }
        // Each wave reads as many slots as its longest row has (4, 6, 8 or
        // kRowK; a longer row loops), so short rows do not pay for long ones;
        // the row's arithmetic is the same whichever thread does it.
        const TRow<T>& rc = r;
        const bool keep = own && tabs(rc.ac) < T(1.0e-30);
        const bool le4 = __all(!own || rc.deg <= 4), le6 = __all(!own || rc.deg <= 6);
        const bool le8 = __all(!own || rc.deg <= 8), leK = __all(!own || rc.deg <= kRowK);
        for (int it = 0; it < count; it++) {
            const T* x = X(S, q);
            T* y = Y(S, q);
            if (own) {
                const T xn = x[cell];
                T res;
                if (keep) {
                    res = xn;
                } else if (le4) {
                    res = waveRowSweep<4>(rc, x, xn, w, cell);
                } else if (le6) {
                    res = waveRowSweep<6>(rc, x, xn, w, cell);
                } else if (le8) {
                    res = waveRowSweep<8>(rc, x, xn, w, cell);
                } else if (leK) {
                    res = waveRowSweep<kRowK>(rc, x, xn, w, cell);
                } else {
                    const T Ax = rowAx(rc, xn, [&](int j) { return x[j]; }, AF, nbs, cell);
                    res = xn + w * (rc.b - Ax) / rc.ac;
                }
                y[cell] = res;
            }
            __syncthreads();
            flip ^= 1u << q;
        }
    };

    // ---- descent
    int done = preDone;
    for (int q = 0; q < nl - 1; q++) {
        sweeps(q, pre - done);
        const TailLevel F = tl[q];
        const TailLevel C = tl[q + 1];
        const bool coarsest = q + 1 == nl - 1;
        const int s1 = coarsest ? (inv == nullptr && coarseSweeps >= 1) : (pre >= 1);
        // residual of each fine cell into the free x buffer of level q ...
        {
            const T* x = X(F, q);
            T* res = Y(F, q);
            const T* AF = sT + F.AF;
            const int* nbs = sI + F.nb;
            if (t < F.n) {
                TRow<T> r;
                loadRow(r, sT + F.AC, sT + F.b, AF, sI + F.start, nbs, t);
                const T Ax = rowAx(r, x[t], [&](int j) { return x[j]; }, AF, nbs, t);
                res[t] = r.b - Ax;
            }
        }
        __syncthreads();
        // ... then gathered per coarse cell in ascending fine order
        {
            const T* res = Y(F, q);
            const int* mS = sI + F.mS;
            const int* mem = sI + F.mem;
            T* xc = X(C, q + 1);
            if (t < C.n) {
                const int j0 = mS[t], cnt = mS[t + 1] - j0;
                T acc = T(0);
                if (cnt <= kMemberK) {
                    T rv[kMemberK];
                    int mv[kMemberK];
#pragma unroll
                    for (int j = 0; j < kMemberK; j++) mv[j] = mem[j0 + (j < cnt ? j : 0)];
#pragma unroll
                    for (int j = 0; j < kMemberK; j++) rv[j] = res[cnt > 0 ? mv[j] : t];
#pragma unroll
                    for (int j = 0; j < kMemberK; j++)
                        if (j < cnt) acc += rv[j];
                } else {
                    for (int j = j0; j < j0 + cnt; j++) acc += res[mem[j]];
                }
                sT[C.b + t] = acc;
                xc[t] = s1 ? sweepFromZero(acc, sT[C.AC + t], w) : T(0);
            }
        }
        __syncthreads();
        done = s1;
    }

    // ---- coarsest
    {
        const int q = nl - 1;
        const TailLevel S = tl[q];
        if (inv != nullptr) {
            const T* b = sT + S.b;
            T* x = X(S, q);
            if (t < S.n) {
                double acc = 0.0;
                for (int c = 0; c < S.n; c++) acc += inv[t * S.n + c] * (double)b[c];
                x[t] = (T)acc;
            }
            __syncthreads();
        } else {
            sweeps(q, coarseSweeps - done);
        }
    }

    // ---- ascent: buildProlongationKernel fused into the first post sweep
    for (int q = nl - 2; q >= 0; q--) {
        const TailLevel F = tl[q];
        const TailLevel C = tl[q + 1];
        const T* xc = X(C, q + 1);
        const int* toC = sI + F.toC;
        if (post >= 1) {
            const T* x = X(F, q);
            T* y = Y(F, q);
            const T* AF = sT + F.AF;
            const int* nbs = sI + F.nb;
            if (t < F.n) {
                TRow<T> r;
                loadRow(r, sT + F.AC, sT + F.b, AF, sI + F.start, nbs, t);
                const T xn = x[t] + xc[toC[t]];
                if (tabs(r.ac) < T(1.0e-30)) {
                    y[t] = xn;
                } else {
                    const T Ax = rowAx(r, xn, [&](int j) { return x[j] + xc[toC[j]]; }, AF, nbs, t);
                    y[t] = xn + w * (r.b - Ax) / r.ac;
                }
            }
            __syncthreads();
            flip ^= 1u << q;
            sweeps(q, post - 1);
        } else {
            T* x = X(F, q);
            if (t < F.n) x[t] += xc[toC[t]];
            __syncthreads();
        }
    }

    // ---- level lt's answer back to its x buffer (x0 of the global image)
    const TailLevel top = tl[0];
    const T* x = X(top, 0);
    T* out = (T*)image + top.x0;
    if (t < top.n) out[t] = x[t];
}

// ===========================================================================
// launchers
// ===========================================================================
int tailLdsLimit() {
    int dev = 0, limit = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    if (hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || limit <= 0)
        return 64 * 1024;
    if (limit > 64 * 1024) {
        // dynamic LDS beyond 64 KiB has to be opted into per kernel; the
        // kernel's static LDS comes out of the same budget
        const int dyn = limit - kTailStaticLds;
        if (hipFuncSetAttribute((const void*)kMGTail<double>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn) !=
                hipSuccess ||
            hipFuncSetAttribute((const void*)kMGTail<float>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn) !=
                hipSuccess) {
            (void)hipGetLastError();
            return 64 * 1024;
        }
    }
    return limit;
}
void launchMGTail(hipStream_t s, const TailLaunch& t, double w, int preDone, int pre, int post, int coarseSweeps,
                  bool fp32, const double* inv, const Ctl* ctl) {
    if (fp32)
        kMGTail<float><<<1, t.threads, t.ldsBytes, s>>>(t.plan, (uint4*)t.image, t.vecs, t.nl, t.intBase, w, preDone,
                                                        pre, post, coarseSweeps, inv, ctl);
    else
        kMGTail<double><<<1, t.threads, t.ldsBytes, s>>>(t.plan, (uint4*)t.image, t.vecs, t.nl, t.intBase, w, preDone,
                                                         pre, post, coarseSweeps, inv, ctl);
}
void launchMGCoarseDiag(hipStream_t s, const DLevel& f, const Ctl* ctl) {
    kMGCoarseDiag<<<blocksFor(f.n), kBlock, 0, s>>>(f, ctl);
}
void launchMGCoarseOperator(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    if (f.nAggBlocks <= 0) throw std::logic_error("level has no aggregate blocks");
    if (f.maxRow <= 3) kMGCoarseOperator<3><<<f.nAggBlocks, kBlock, 0, s>>>(f, c, ctl);
    else if (f.maxRow <= 4) kMGCoarseOperator<4><<<f.nAggBlocks, kBlock, 0, s>>>(f, c, ctl);
    else if (f.maxRow <= 7) kMGCoarseOperator<7><<<f.nAggBlocks, kBlock, 0, s>>>(f, c, ctl);
    else kMGCoarseOperator<8><<<f.nAggBlocks, kBlock, 0, s>>>(f, c, ctl);
}
void launchMGCoarseGather(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    kMGCoarseGather<<<blocksFor((long)c.refs + c.n), kBlock, 0, s>>>(f, c, ctl);
}
// the row batch of a level: its longest row rounded up to 4, 8 or 16 (0: loop)
template <class T>
static void launchMGResidualT(hipStream_t s, const DLevel& f, const Ctl* ctl) {
    const dim3 g(blocksFor(f.n, f.block)), b(f.block);
    if (f.maxRow <= 4) kMGResidual<T, 4><<<g, b, 0, s>>>(f, ctl);
    else kMGResidual<T, 8><<<g, b, 0, s>>>(f, ctl);
}
template <class T>
static void launchMGJacobiT(hipStream_t s, const DLevel& L, T w, const Ctl* ctl) {
    const dim3 g(blocksFor(L.n, L.block)), b(L.block);
    if (L.maxRow <= 4) kMGJacobi<T, 4><<<g, b, 0, s>>>(L, w, ctl);
    else kMGJacobi<T, 8><<<g, b, 0, s>>>(L, w, ctl);
}
template <class T>
static void launchMGProlongSmoothT(hipStream_t s, const DLevel& f, const DLevel& c, T w, const Ctl* ctl,
                                   double* apply = nullptr) {
    const dim3 g(blocksFor(f.n, f.block)), b(f.block);
    if (f.maxRow <= 3) kMGProlongSmooth<T, 3><<<g, b, 0, s>>>(f, c, w, apply, ctl);
    else if (f.maxRow <= 4) kMGProlongSmooth<T, 4><<<g, b, 0, s>>>(f, c, w, apply, ctl);
    else if (f.maxRow <= 7) kMGProlongSmooth<T, 7><<<g, b, 0, s>>>(f, c, w, apply, ctl);
    else kMGProlongSmooth<T, 8><<<g, b, 0, s>>>(f, c, w, apply, ctl);
}
void launchMGResidual(hipStream_t s, const DLevel& f, const Ctl* ctl) { launchMGResidualT<double>(s, f, ctl); }
void launchMGResidual32(hipStream_t s, const DLevel& f, const Ctl* ctl) { launchMGResidualT<float>(s, f, ctl); }
void launchMGJacobi32(hipStream_t s, const DLevel& L, float w, const Ctl* ctl) { launchMGJacobiT<float>(s, L, w, ctl); }
void launchMGProlongSmooth32(hipStream_t s, const DLevel& f, const DLevel& c, float w, const Ctl* ctl,
                             double* apply) {
    launchMGProlongSmoothT<float>(s, f, c, w, ctl, apply);
}
template <class T>
static void launchMGResidualRestrictT(hipStream_t s, const DLevel& f, const DLevel& c, double w, int sweep1,
                                      const Ctl* ctl) {
    const dim3 g(f.nAggBlocks), b(kBlock);
    if (f.maxRow <= 4) kMGResidualRestrict<T, 4><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
    else if (f.maxRow <= 7) kMGResidualRestrict<T, 7><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
    else kMGResidualRestrict<T, 8><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
}
void launchMGResidualRestrict(hipStream_t s, const DLevel& f, const DLevel& c, double w, int sweep1, bool fp32,
                              const Ctl* ctl) {
    if (f.nAggBlocks <= 0) throw std::logic_error("level has no aggregate blocks");
    if (fp32) launchMGResidualRestrictT<float>(s, f, c, w, sweep1, ctl);
    else launchMGResidualRestrictT<double>(s, f, c, w, sweep1, ctl);
}
void launchMGRestrictFused01(hipStream_t s, const DLevel& l0, const DLevel& l1, const DLevel& l2, L1Fused t,
                             double w, int sweep1, int sweep2, bool fp32, const Ctl* ctl) {
    if (l1.nAggBlocks <= 0 || !t.halo || !t.loc || l1.maxRow > kL1Slots)
        throw std::logic_error("level 1 has no fused restriction tables");
    if (fp32) kMGRestrictFused01<float><<<l1.nAggBlocks, kBlock, 0, s>>>(l0, l1, l2, t, (float)w, sweep1, sweep2, ctl);
    else kMGRestrictFused01<double><<<l1.nAggBlocks, kBlock, 0, s>>>(l0, l1, l2, t, w, sweep1, sweep2, ctl);
}
void launchMGRestrictSum(hipStream_t s, const DLevel& f, const DLevel& c, double w, int sweep1, bool fp32,
                         const Ctl* ctl) {
    const dim3 g(blocksFor(c.n, c.block)), b(c.block);
    if (fp32) {
        if (f.maxMembers <= 4) kMGRestrictSum<float, 4><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
        else if (f.maxMembers <= 6) kMGRestrictSum<float, 6><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
        else if (f.maxMembers <= 8) kMGRestrictSum<float, 8><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
        else kMGRestrictSum<float, 0><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
    } else {
        if (f.maxMembers <= 4) kMGRestrictSum<double, 4><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
        else if (f.maxMembers <= 6) kMGRestrictSum<double, 6><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
        else if (f.maxMembers <= 8) kMGRestrictSum<double, 8><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
        else kMGRestrictSum<double, 0><<<g, b, 0, s>>>(f, c, w, sweep1, ctl);
    }
}
void launchGradUV(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DBC bp, DState st, int scheme, bool withP) {
    const dim3 g(blocksFor(m.N)), b(kBlock);
    if (m.maxSlots <= 3) {
        if (withP) kGradUV<3, true><<<g, b, 0, s>>>(m, bu, bv, bp, st, scheme);
        else kGradUV<3, false><<<g, b, 0, s>>>(m, bu, bv, bp, st, scheme);
    } else {
        if (withP) kGradUV<4, true><<<g, b, 0, s>>>(m, bu, bv, bp, st, scheme);
        else kGradUV<4, false><<<g, b, 0, s>>>(m, bu, bv, bp, st, scheme);
    }
}
void launchGradField(hipStream_t s, const DMesh& m, DBC bc, const double* phi, double* gz, double* gr,
                     int scheme, const Ctl* ctl) {
    if (m.maxSlots <= 3) kGradField<3><<<blocksFor(m.N), kBlock, 0, s>>>(m, bc, phi, gz, gr, scheme, ctl);
    else kGradField<4><<<blocksFor(m.N), kBlock, 0, s>>>(m, bc, phi, gz, gr, scheme, ctl);
}
void launchMomentum(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DBC bp, DState st, Physics ph, bool gradPDone) {
    const dim3 g(blocksFor(m.N)), b(kBlock);
    const bool ho = ph.scheme != 0;
#define AXS_MOMENTUM(BB, HO, PG)                                              \
    (st.momR32 ? kMomentum<BB, HO, PG, true><<<g, b, 0, s>>>(m, bu, bv, bp, st, ph) \
               : kMomentum<BB, HO, PG, false><<<g, b, 0, s>>>(m, bu, bv, bp, st, ph))
    if (m.maxSlots <= 3) {
        if (ho) gradPDone ? AXS_MOMENTUM(3, true, false) : AXS_MOMENTUM(3, true, true);
        else gradPDone ? AXS_MOMENTUM(3, false, false) : AXS_MOMENTUM(3, false, true);
    } else {
        if (ho) gradPDone ? AXS_MOMENTUM(4, true, false) : AXS_MOMENTUM(4, true, true);
        else gradPDone ? AXS_MOMENTUM(4, false, false) : AXS_MOMENTUM(4, false, true);
    }
#undef AXS_MOMENTUM
}
void launchJacobiUV(hipStream_t s, const DMesh& m, DState st, const double* uo, const double* vo, double* un,
                    double* vn) {
    launchJacobiUVPacked(s, m, st, uo, vo, nullptr, un, vn, nullptr);
}
void launchJacobiUVPacked(hipStream_t s, const DMesh& m, DState st, const double* uo, const double* vo,
                          const double2* uvo, double* un, double* vn, double2* uvn) {
    const dim3 g(blocksFor(m.N, m.sweepBlock)), b(m.sweepBlock);
    if (m.maxSlots <= 3) kJacobiUV<3><<<g, b, 0, s>>>(m, st, UVIn{uo, vo, uvo}, UVOut{un, vn, uvn});
    else kJacobiUV<4><<<g, b, 0, s>>>(m, st, UVIn{uo, vo, uvo}, UVOut{un, vn, uvn});
}
int sweepHaloMax(int cells) {
    if (cells == 512) return sweepHalo<512>();
    if (cells == 1024) return sweepHalo<1024>();
    return 0;
}
namespace {
template <int C, int S>
void launchJacobiUVBlockCS(hipStream_t s, const DMesh& m, DState st, const DSweepBlocks& sb, UVIn in, UVOut out) {
    if (m.maxSlots <= 3) kJacobiUVBlock<C, S, 3><<<sb.nBlocks, C, 0, s>>>(m, st, sb, in, out);
    else kJacobiUVBlock<C, S, kSweepRegSlots><<<sb.nBlocks, C, 0, s>>>(m, st, sb, in, out);
}
template <int C>
void launchJacobiUVBlockC(hipStream_t s, const DMesh& m, DState st, const DSweepBlocks& sb, int sweeps, UVIn in,
                          UVOut out) {
    if (sweeps == 2) launchJacobiUVBlockCS<C, 2>(s, m, st, sb, in, out);
    else if (sweeps == 3) launchJacobiUVBlockCS<C, 3>(s, m, st, sb, in, out);
    else if (sweeps == 4) launchJacobiUVBlockCS<C, 4>(s, m, st, sb, in, out);
    else launchJacobiUVBlockCS<C, 5>(s, m, st, sb, in, out);
}
}  // namespace
void launchJacobiUVBlock(hipStream_t s, const DMesh& m, DState st, const DSweepBlocks& sb, int sweeps,
                         const double* uo, const double* vo, const double2* uvo, double* un, double* vn,
                         double2* uvn) {
    if (sweeps < 2 || sweeps > kMaxFusedSweeps || sweeps > sb.rings + 1)
        throw std::invalid_argument("bad fused sweep count");
    const UVIn in{uo, vo, uvo};
    const UVOut out{un, vn, uvn};
    if (sb.cells == 512) launchJacobiUVBlockC<512>(s, m, st, sb, sweeps, in, out);
    else if (sb.cells == 1024) launchJacobiUVBlockC<1024>(s, m, st, sb, sweeps, in, out);
    else throw std::invalid_argument("fused sweeps take blocks of 512 or 1024 cells");
}
void launchGSColorUV(hipStream_t s, const DMesh& m, DState st, const int* order, int begin, int count) {
    kGSColorUV<<<blocksFor(count), kBlock, 0, s>>>(m, st, order, begin, count);
}
void launchJacobiSys(hipStream_t s, const DMesh& m, const double* b, const double* inv, const double* AF,
                     const double* xo, double* xn, const Ctl* ctl) {
    kJacobiSys<<<blocksFor(m.N), kBlock, 0, s>>>(m, b, inv, AF, xo, xn, ctl);
}
void launchGSColorSys(hipStream_t s, const DMesh& m, const double* b, const double* inv, const double* AF, double* x,
                      const int* order, int begin, int count, const Ctl* ctl) {
    kGSColorSys<<<blocksFor(count), kBlock, 0, s>>>(m, b, inv, AF, x, order, begin, count, ctl);
}
void launchPPCoeff(hipStream_t s, const DMesh& m, DBC bp, DState st, Physics ph) {
    if (m.maxSlots <= 3) kPPCoeff<3><<<blocksFor(m.N), kBlock, 0, s>>>(m, bp, st, ph);
    else kPPCoeff<4><<<blocksFor(m.N), kBlock, 0, s>>>(m, bp, st, ph);
}
void launchFaceFlux(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DBC bp, DState st, Physics ph) {
    kFaceFlux<<<blocksFor(m.F), kBlock, 0, s>>>(m, bu, bv, bp, st, ph);
}
void launchPPRhs(hipStream_t s, const DMesh& m, DState st, Physics ph, int cross, double* zero) {
    const dim3 g(blocksFor(m.N)), b(kBlock);
    if (m.maxSlots <= 3) {
        if (cross) kPPRhs<3, true><<<g, b, 0, s>>>(m, st, ph, zero);
        else kPPRhs<3, false><<<g, b, 0, s>>>(m, st, ph, zero);
    } else {
        if (cross) kPPRhs<4, true><<<g, b, 0, s>>>(m, st, ph, zero);
        else kPPRhs<4, false><<<g, b, 0, s>>>(m, st, ph, zero);
    }
}
void launchGradPPUpdate(hipStream_t s, const DMesh& m, DBC ppbc, DState st, Physics ph) {
    if (m.maxSlots <= 3) kGradPPUpdate<3><<<blocksFor(m.N), kBlock, 0, s>>>(m, ppbc, st, ph);
    else kGradPPUpdate<4><<<blocksFor(m.N), kBlock, 0, s>>>(m, ppbc, st, ph);
}
void launchMassFluxUpdate(hipStream_t s, const DMesh& m, DBC bp, DState st, Physics ph) {
    kMassFluxUpdate<<<blocksFor(m.F), kBlock, 0, s>>>(m, bp, st, ph);
}
void launchResidual(hipStream_t s, const DMesh& m, DState st, int blocks, const int2* blockCells) {
    if (m.maxSlots <= 3) kResidual<3><<<blocks, kBlock, 0, s>>>(m, st, blockCells);
    else kResidual<4><<<blocks, kBlock, 0, s>>>(m, st, blockCells);
}
void launchFinish(hipStream_t s, DState st, int blocks, bool scalars) {
    if (scalars) kFinish<true><<<1, kFinishThreads, 0, s>>>(st, blocks);
    else kFinish<false><<<1, kFinishThreads, 0, s>>>(st, blocks);
}
void launchFinishCases(hipStream_t s, DState st, DCases cs, int blocks, bool scalars) {
    if (scalars) kFinishCases<true><<<cs.n, kFinishThreads, 0, s>>>(st, cs, blocks);
    else kFinishCases<false><<<cs.n, kFinishThreads, 0, s>>>(st, cs, blocks);
    kFinishRun<<<1, 1, 0, s>>>(st, cs);
}
void launchFreeze(hipStream_t s, DCases cs, int nCells, FreezeSet cells, int nFaces, FreezeSet faces) {
    if (cells.n > 0) kFreeze<<<blocksFor(nCells), kBlock, 0, s>>>(cs.caseOfCell, cs.ctl, nCells, cells);
    if (faces.n > 0) kFreeze<<<blocksFor(nFaces), kBlock, 0, s>>>(cs.caseOfFace, cs.ctl, nFaces, faces);
    kMarkFrozen<<<blocksFor(cs.n), kBlock, 0, s>>>(cs.ctl, cs.n);
}
void launchResidualScalars(hipStream_t s, const DMesh& m, DState st, DScalar t, DScalar c, int blocks,
                           const int2* blockCells) {
    double* part = st.partial + (size_t)blocks * 9;
    if (m.maxSlots <= 3) kResidualScalars<3><<<blocks, kBlock, 0, s>>>(m, t, c, part, st.ctl, blockCells);
    else kResidualScalars<4><<<blocks, kBlock, 0, s>>>(m, t, c, part, st.ctl, blockCells);
}
void launchScalar(hipStream_t s, const DMesh& m, DBC bc, DScalar sc, DWall wall, const double* mDot, ScalarPhysics ph,
                  const Ctl* ctl) {
    const int blocks = blocksFor(m.N);
    const bool ho = ph.conv && ph.scheme != 0;
    if (m.maxSlots <= 3) {
        if (ho) kScalar<3, true><<<blocks, kBlock, 0, s>>>(m, bc, sc, wall, mDot, ph, ctl);
        else kScalar<3, false><<<blocks, kBlock, 0, s>>>(m, bc, sc, wall, mDot, ph, ctl);
    } else {
        if (ho) kScalar<4, true><<<blocks, kBlock, 0, s>>>(m, bc, sc, wall, mDot, ph, ctl);
        else kScalar<4, false><<<blocks, kBlock, 0, s>>>(m, bc, sc, wall, mDot, ph, ctl);
    }
}
void launchRotate(hipStream_t s, int n, double* old2, double* old, const double* phi, int bdf2) {
    kRotate<<<blocksFor(n), kBlock, 0, s>>>(n, old2, old, phi, bdf2);
}
void launchStepBegin(hipStream_t s, const DMesh& m, DState st, int bdf2) {
    kStepBegin<<<blocksFor(m.N), kBlock, 0, s>>>(m, st, bdf2);
}
void launchCopy(hipStream_t s, double* dst, const double* src, int n, const Ctl* ctl) {
    kCopy<<<blocksFor(n), kBlock, 0, s>>>(dst, src, n, ctl);
}
void launchFill(hipStream_t s, double* dst, double value, int n, const Ctl* ctl) {
    kFill<<<blocksFor(n), kBlock, 0, s>>>(dst, value, n, ctl);
}
void launchStreamRead(hipStream_t s, const double* src, size_t n, double* out, int blocks, double mark) {
    kStreamRead<<<blocks, kBlock, 0, s>>>(src, n, out, mark);
}
void launchMGJacobi(hipStream_t s, const DLevel& L, double w, const Ctl* ctl) { launchMGJacobiT<double>(s, L, w, ctl); }
void launchMGSmoothBlock(hipStream_t s, const DLevel& L, double w, int sweeps, const Ctl* ctl) {
    const int threads = L.n >= 1024 ? 1024 : ((L.n + 63) / 64) * 64;
    kMGSmoothBlock<<<1, threads, 2 * (size_t)L.n * sizeof(double), s>>>(L, w, sweeps, ctl);
}
void launchMGProlongSmooth(hipStream_t s, const DLevel& f, const DLevel& c, double w, const Ctl* ctl) {
    launchMGProlongSmoothT<double>(s, f, c, w, ctl);
}
namespace {
template <class T, int C, int S, bool P, int X0 = kL0X>
void launchMGLevel0BlockT(hipStream_t s, const DLevel& f, const DLevel& c, const DSweepBlocks& sb, T w,
                          const Ctl* ctl) {
    if (f.maxRow <= 3) kMGLevel0Block<T, C, S, P, true, 3, X0><<<sb.nBlocks, C, 0, s>>>(f, c, sb, w, ctl);
    else kMGLevel0Block<T, C, S, P, true, kSweepRegSlots, X0><<<sb.nBlocks, C, 0, s>>>(f, c, sb, w, ctl);
}
template <class T, int S, bool P>
void launchMGLevel0BlockS(hipStream_t s, const DLevel& f, const DLevel& c, const DSweepBlocks& sb, T w,
                          const Ctl* ctl) {
    if (sb.cells == 512) launchMGLevel0BlockT<T, 512, S, P>(s, f, c, sb, w, ctl);
    else launchMGLevel0BlockT<T, 1024, S, P>(s, f, c, sb, w, ctl);
}
template <class T>
void launchMGLevel0BlockV(hipStream_t s, const DLevel& f, const DLevel& c, const DSweepBlocks& sb, int sweeps,
                          bool prolong, T w, const Ctl* ctl) {
    if (sweeps == 1 && !prolong) {
        if (sb.rings < 1) throw std::invalid_argument("fused level-0 sweeps need one halo ring");
        launchMGLevel0BlockS<T, 1, false>(s, f, c, sb, w, ctl);
    } else if (sweeps == 2 && prolong) {
        if (sb.rings < 2) throw std::invalid_argument("fused level-0 sweeps need two halo rings");
        launchMGLevel0BlockS<T, 2, true>(s, f, c, sb, w, ctl);
    } else {
        throw std::invalid_argument("unsupported fused level-0 sweep combination");
    }
}
}  // namespace
void launchMGLevel0Block(hipStream_t s, const DLevel& f, const DLevel& c, const DSweepBlocks& sb, int sweeps,
                         bool prolong, double w, bool fp32, const Ctl* ctl) {
    if (sb.cells != 512 && sb.cells != 1024) throw std::invalid_argument("fused level-0 sweeps take 512 or 1024 cells");
    if (fp32) launchMGLevel0BlockV<float>(s, f, c, sb, sweeps, prolong, (float)w, ctl);
    else launchMGLevel0BlockV<double>(s, f, c, sb, sweeps, prolong, w, ctl);
}
void launchMGLevel0BlockStart32(hipStream_t s, const DLevel& f, const DLevel& c, const DSweepBlocks& sb, double w,
                                bool defect, const Ctl* ctl) {
    if (sb.cells != 512 && sb.cells != 1024) throw std::invalid_argument("fused level-0 sweeps take 512 or 1024 cells");
    if (sb.rings < 1) throw std::invalid_argument("fused level-0 sweeps need one halo ring");
    const float w32 = (float)w;
    if (defect) {
        if (sb.cells == 512) launchMGLevel0BlockT<float, 512, 1, false, kL0Defect>(s, f, c, sb, w32, ctl);
        else launchMGLevel0BlockT<float, 1024, 1, false, kL0Defect>(s, f, c, sb, w32, ctl);
    } else {
        if (sb.cells == 512) launchMGLevel0BlockT<float, 512, 1, false, kL0Zero>(s, f, c, sb, w32, ctl);
        else launchMGLevel0BlockT<float, 1024, 1, false, kL0Zero>(s, f, c, sb, w32, ctl);
    }
}
void launchMGProlong(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    kMGProlong<<<blocksFor(f.n, f.block), f.block, 0, s>>>(f, c, ctl);
}
void launchMGGSColor(hipStream_t s, const DLevel& L, const int* order, int begin, int count, bool fp32,
                     const Ctl* ctl) {
    if (count <= 0) return;
    if (fp32) kMGGSColor<float><<<blocksFor(count, L.block), L.block, 0, s>>>(L, order, begin, count, ctl);
    else kMGGSColor<double><<<blocksFor(count, L.block), L.block, 0, s>>>(L, order, begin, count, ctl);
}
void launchMGSmoothBlockGS(hipStream_t s, const DLevel& L, const int* order, const int* colorStart, int nColors,
                           int sweeps, bool fp32, const Ctl* ctl) {
    const int threads = L.n >= 1024 ? 1024 : ((L.n + 63) / 64) * 64;
    if (fp32)
        kMGSmoothBlockGS<float><<<1, threads, (size_t)L.n * sizeof(float), s>>>(L, order, colorStart, nColors,
                                                                               sweeps, ctl);
    else
        kMGSmoothBlockGS<double><<<1, threads, (size_t)L.n * sizeof(double), s>>>(L, order, colorStart, nColors,
                                                                                 sweeps, ctl);
}

}  // namespace axs
