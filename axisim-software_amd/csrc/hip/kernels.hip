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

#include "device.hpp"

namespace axs {

namespace {

__device__ __forceinline__ bool finished(const Ctl* ctl) { return ctl != nullptr && ctl->done != 0; }

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
    const int g = m.group[f];
    if (g < 0 || g >= bc.nGroups) return phiP;
    const uint8_t t = bc.type[g];
    const double v = bc.value[g];
    if (t == 0 || t == 2) return prescribed(m, f, t, v, bc.length[g]);
    if (t == 1) return phiP + v * m.dPB[f];
    return phiP;
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

// phiGradientGreenGauss / phiGradientLeastSquare, solver_util.cuh:458-588
__device__ void cellGradient(const DMesh& m, const DBC& bc, const double* phi, int c, int scheme, double& gZ,
                             double& gR) {
    const int k0 = m.faceStart[c], k1 = m.faceStart[c + 1];
    if (scheme == 0) {
        double gz = 0.0, gr = 0.0, clz = 0.0, clr = 0.0, axisPhi = 0.0;
        bool axis = false;
        for (int k = k0; k < k1; k++) {
            const int f = m.faceIDs[k];
            double nz = m.nZ[f], nr = m.nR[f];
            if (m.neighbor[f] == c) {
                nz = -nz;
                nr = -nr;
            }
            const double phiF = interpToFace(m, bc, phi, c, f);
            if (m.fR[f] > 1.0e-30) {
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
    } else {
        const double phiP = phi[c];
        double bz = 0.0, br = 0.0;
        for (int k = k0; k < k1; k++) {
            const int f = m.faceIDs[k];
            const int ow = m.owner[f], nb = m.neighbor[f];
            double dphi;
            if (nb >= 0) {
                const bool own = (ow == c);
                const double phiNb = phi[own ? nb : ow];
                dphi = own ? (phiNb - phiP) : (phiP - phiNb);
            } else {
                dphi = interpToFace(m, bc, phi, c, f) - phiP;
            }
            bz += m.lsqWZ[f] * dphi;
            br += m.lsqWR[f] * dphi;
        }
        const double izr = m.lsqInvZR[c];
        gZ = m.lsqInvZZ[c] * bz + izr * br;
        gR = izr * bz + m.lsqInvRR[c] * br;
    }
}

// higherOrderFaceValue, solver_util.cuh:724-768
__device__ __forceinline__ double higherOrder(const DMesh& m, const DBC& bc, const double* phi, const double* gz,
                                              const double* gr, int scheme, int n, int nb, int f, double F,
                                              double phiUD) {
    if (scheme == 1) return interpToFace(m, bc, phi, n, f);
    if (scheme == 2) {
        const int up = (F > 0.0) ? n : nb;
        const double dz = m.fZ[f] - m.cZ[up];
        const double dr = m.fR[f] - m.cR[up];
        const double ho = phi[up] + gz[up] * dz + gr[up] * dr;
        const double lo = fmin(phi[n], phi[nb]);
        const double hi = fmax(phi[n], phi[nb]);
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

// ---- wave64 / workgroup reductions (deterministic for a fixed launch) ----
__device__ __forceinline__ double waveSum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double waveMax(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace

// ===========================================================================
// Gradients (computeGradient, solver_util.cu:146-169)
// ===========================================================================
__global__ void __launch_bounds__(kBlock) kGradUV(DMesh m, DBC bu, DBC bv, DState s, int scheme) {
    if (finished(s.ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double gz, gr;
    cellGradient(m, bu, s.u, n, scheme, gz, gr);
    s.gUZ[n] = gz;
    s.gUR[n] = gr;
    cellGradient(m, bv, s.v, n, scheme, gz, gr);
    s.gVZ[n] = gz;
    s.gVR[n] = gr;
}

__global__ void __launch_bounds__(kBlock) kGradField(DMesh m, DBC bc, const double* phi, double* gzo,
                                                     double* gro, int scheme, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double gz, gr;
    cellGradient(m, bc, phi, n, scheme, gz, gr);
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
__global__ void __launch_bounds__(kBlock) kMomentum(DMesh m, DBC bu, DBC bv, DBC bp, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    const double mu = ph.mu;
    double acU = 0.0, bU = 0.0, acV = 0.0, bV = 0.0, netF = 0.0;
    const int k0 = m.faceStart[n], k1 = m.faceStart[n + 1];
    for (int k = k0; k < k1; k++) {
        const int f = m.faceIDs[k];
        const int ow = m.owner[f], nb = m.neighbor[f];
        const double area = m.area[f];
        double F = 0.0;
        if (ph.conv) {
            const double Fo = s.mDot[f] * 1.0;
            if (ow == n) F = Fo;
            else if (nb == n) F = -Fo;
            if (fabs(F) <= 1.0e-30) F = 0.0;
            netF += F;
        }
        if (nb >= 0) {
            const int o = (ow == n) ? nb : ow;
            const double K = mu * area * m.invDPN[f];
            double af = 0.0;
            acU += K;
            af -= K;
            if (ph.nonOrtho) bU += nonOrthoFlux(m, n, f, s.gUZ, s.gUR, mu);
            if (F != 0.0) {
                acU += fmax(F, 0.0);
                af += fmin(F, 0.0);
                if (ph.scheme != 0) {
                    const double ud = (F > 0.0) ? s.u[n] : s.u[o];
                    const double ho = higherOrder(m, bu, s.u, s.gUZ, s.gUR, ph.scheme, n, o, f, F, ud);
                    bU -= F * (ho - ud);
                }
            }
            acV += K;
            if (ph.nonOrtho) bV += nonOrthoFlux(m, n, f, s.gVZ, s.gVR, mu);
            if (F != 0.0) {
                acV += fmax(F, 0.0);
                if (ph.scheme != 0) {
                    const double ud = (F > 0.0) ? s.v[n] : s.v[o];
                    const double ho = higherOrder(m, bv, s.v, s.gVZ, s.gVR, ph.scheme, n, o, f, F, ud);
                    bV -= F * (ho - ud);
                }
            }
            s.AFuv[k] = af;
            if (s.AF32) s.AF32[k] = (float)af;
        } else {
            boundaryTransport(m, bu, n, f, area, mu, F, acU, bU);
            boundaryTransport(m, bv, n, f, area, mu, F, acV, bV);
            s.AFuv[k] = 0.0;
            if (s.AF32) s.AF32[k] = 0.0f;
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
    cellGradient(m, bp, s.p, n, ph.gradScheme, gz, gr);
    s.gPZ[n] = gz;
    s.gPR[n] = gr;
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
    if (s.uInv32) {
        s.uInv32[n] = (float)iU;
        s.vInv32[n] = (float)iV;
    }
}

// ===========================================================================
// Momentum smoothers: one Jacobi sweep of u and v together (jacobi,
// linear_solver.cu:9-33); the two systems share CSR and off-diagonals.
// ===========================================================================
__global__ void __launch_bounds__(kBlock) kJacobiUV(DMesh m, DState s, const double* __restrict__ uo,
                                                    const double* __restrict__ vo, double* __restrict__ un,
                                                    double* __restrict__ vn) {
    if (finished(s.ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double au = s.ub[n], av = s.vb[n];
    for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
        const int nb = m.slotNb[k];
        if (nb >= 0) {
            const double a = s.AFuv[k];
            au -= a * uo[nb];
            av -= a * vo[nb];
        }
    }
    const double iu = s.uInv[n], iv = s.vInv[n];
    un[n] = (iu != 0.0) ? au * iu : uo[n];
    vn[n] = (iv != 0.0) ? av * iv : vo[n];
}

// gaussSeidelColorSweep, linear_solver.cu:42-72, u and v together
__global__ void __launch_bounds__(kBlock) kGSColorUV(DMesh m, DState s, const int* order, int begin, int count) {
    if (finished(s.ctl)) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
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
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
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
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ void __launch_bounds__(kBlock) kPPCoeff(DMesh m, DBC bp, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double ac = 0.0;
    for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
        const int f = m.faceIDs[k];
        const int nb = m.neighbor[f];
        const double area = m.area[f];
        const double Df = normalCorrToFace(m, s, n, f);
        double af = 0.0;
        if (nb >= 0) {
            const double K = ph.rho * area * Df * m.invDPN[f];
            ac += K;
            af -= K;
        } else {
            const int g = m.group[f];
            if (g >= 0 && g < bp.nGroups && bp.type[g] == 0) {
                const double K = ph.rho * area * Df / m.dPB[f];
                ac += K;
            }
        }
        s.ppAF[k] = af;
    }
    s.ppAC[n] = ac;
    s.ppInv[n] = (fabs(ac) < 1.0e-30) ? 0.0 : 1.0 / ac;
}

// computeFaceMassFluxRhieChow, solver_util.cu:171-203
__global__ void __launch_bounds__(kBlock) kFaceFlux(DMesh m, DBC bu, DBC bv, DBC bp, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const double area = m.area[f];
    if (area <= 1.0e-30) {
        s.mDot[f] = 0.0;
        return;
    }
    const double un = rhieChow(m, bu, bv, bp, s, m.owner[f], f);
    s.mDot[f] = ph.rho * un * area;
}

// createPPRhs, simple.cu:90-142
__global__ void __launch_bounds__(kBlock) kPPRhs(DMesh m, DState s, Physics ph, int cross) {
    if (finished(s.ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double imb = 0.0, cs = 0.0;
    for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
        const int f = m.faceIDs[k];
        const double md = s.mDot[f];
        if (m.owner[f] == n) imb += md;
        else if (m.neighbor[f] == n) imb -= md;
        if (cross) {
            const double Df = normalCorrToFace(m, s, n, f);
            cs += nonOrthoFlux(m, n, f, s.gPZ, s.gPR, ph.rho * Df);
        }
    }
    s.ppb[n] = -imb + cs;
}

// computeGradient(p', ppBC) fused with updateMomentumPressure (simple.cu:166-179)
__global__ void __launch_bounds__(kBlock) kGradPPUpdate(DMesh m, DBC ppbc, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double gz, gr;
    cellGradient(m, ppbc, s.pp, n, ph.gradScheme, gz, gr);
    s.gPZ[n] = gz;
    s.gPR[n] = gr;
    s.u[n] -= s.DU[n] * gz;
    s.v[n] -= s.DV[n] * gr;
    s.p[n] += ph.alphaP * s.pp[n];
}

// updateMassFlux, simple.cu:181-258
__global__ void __launch_bounds__(kBlock) kMassFluxUpdate(DMesh m, DBC bp, DState s, Physics ph) {
    if (finished(s.ctl)) return;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const int ow = m.owner[f], nb = m.neighbor[f];
    const double area = m.area[f];
    if (area <= 1.0e-30) {
        s.mDot[f] = 0.0;
        return;
    }
    const double Df = normalCorrToFace(m, s, ow, f);
    if (nb >= 0) {
        double md = s.mDot[f];
        md -= ph.rho * area * Df * (s.pp[nb] - s.pp[ow]) * m.invDPN[f];
        if (ph.nonOrtho) md -= nonOrthoFlux(m, ow, f, s.gPZ, s.gPR, ph.rho * Df);
        s.mDot[f] = md;
        return;
    }
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
__global__ void __launch_bounds__(kBlock) kResidual(DMesh m, DState s) {
    if (finished(s.ctl)) return;
    if (s.ctl->k % s.ctl->checkConv != 0) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (n < m.N) {
        double axu = s.uAC[n] * s.u[n];
        double axv = s.vAC[n] * s.v[n];
        v[3] = fabs(axu);
        v[7] = fabs(axv);
        double imb = 0.0;
        for (int k = m.faceStart[n]; k < m.faceStart[n + 1]; k++) {
            const int nb = m.slotNb[k];
            if (nb >= 0) {
                const double a = s.AFuv[k];
                axu += a * s.u[nb];
                axv += a * s.v[nb];
            }
            const int f = m.faceIDs[k];
            if (m.owner[f] == n) imb += s.mDot[f];
            else if (m.neighbor[f] == n) imb -= s.mDot[f];
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
        const double w = (j == 2 || j == 6) ? waveMax(v[j]) : waveSum(v[j]);
        if (lane == 0) red[wave][j] = w;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        const int j = threadIdx.x;
        double a = red[0][j];
        for (int w = 1; w < kBlock / 64; w++) a = (j == 2 || j == 6) ? fmax(a, red[w][j]) : a + red[w][j];
        s.partial[blockIdx.x * 9 + j] = a;
    }
}

// Final reduction + residualAllHost (residuals.cu:37-100) + residualsConverged
// (solver.cpp:59-72) + the loop bookkeeping of solver.cpp:1243-1283, on device.
__global__ void __launch_bounds__(kBlock) kFinish(DState s, int blocks) {
    Ctl* c = s.ctl;
    if (c->done) return;
    const bool check = (c->k % c->checkConv) == 0;
    __shared__ double red[kBlock][9];
    if (check) {
        double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int b = threadIdx.x; b < blocks; b += kBlock)
            for (int j = 0; j < 9; j++) {
                const double x = s.partial[b * 9 + j];
                a[j] = (j == 2 || j == 6) ? fmax(a[j], x) : a[j] + x;
            }
        for (int j = 0; j < 9; j++) red[threadIdx.x][j] = a[j];
        __syncthreads();
        for (int st = kBlock / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st)
                for (int j = 0; j < 9; j++) {
                    const double x = red[threadIdx.x + st][j];
                    red[threadIdx.x][j] = (j == 2 || j == 6) ? fmax(red[threadIdx.x][j], x) : red[threadIdx.x][j] + x;
                }
            __syncthreads();
        }
    }
    if (threadIdx.x != 0) return;
    if (check) {
        const int scaleIter = c->transient ? c->k : c->iter;
        for (int i = 0; i < 3; i++) {
            if (!c->enabled[i]) continue;
            double res, scale = c->scaleVal[i];
            if (i == 2) {
                res = red[0][8];
            } else {
                const int o = i * 4;
                res = c->norm[i] == 0 ? red[0][o] : c->norm[i] == 1 ? red[0][o + 1] : red[0][o + 2];
                switch (c->scaling[i]) {
                case 0: scale = 1.0; break;
                case 1: scale = (double)c->N; break;
                case 2: scale = sqrt((double)c->N); break;
                case 3: scale = red[0][o + 3]; break;
                default: break;
                }
            }
            if (i == 2 && scaleIter < 5) scale = fmax(res, 0.0);
            c->scaleVal[i] = scale;
            if (scale != 0.0) {
                if (c->norm[i] == 1) res = sqrt(res);
                res /= scale;
            }
            c->resVal[i] = res;
        }
        if (c->nHist < c->histCap) {
            for (int i = 0; i < 3; i++) s.hist[c->nHist * 4 + i] = c->resVal[i];
            s.hist[c->nHist * 4 + 3] = (double)c->iter;
            c->nHist++;
        }
        bool any = false, ok = true;
        for (int i = 0; i < 3; i++) {
            if (!c->enabled[i]) continue;
            any = true;
            if (!(c->resVal[i] <= c->tol[i])) ok = false;
        }
        c->converged = (any && ok) ? 1 : 0;
    }
    c->iter++;
    c->k++;
    if (c->converged || c->k >= c->maxIter) c->done = 1;
}

// Start of a time step: freeze the old levels (solver.cpp:1058-1095) and reset
// the per-step loop state.  Copies instead of the reference's pointer swap
// keep the graph's addresses fixed; the values are the same.
__global__ void __launch_bounds__(kBlock) kStepBegin(DMesh m, DState s, int bdf2) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < m.N) {
        if (bdf2) {
            s.uOld2[n] = s.uOld[n];
            s.vOld2[n] = s.vOld[n];
        }
        s.uOld[n] = s.u[n];
        s.vOld[n] = s.v[n];
    }
}

__global__ void __launch_bounds__(kBlock) kCopy(double* dst, const double* src, int n, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void __launch_bounds__(kBlock) kFill(double* dst, double value, int n, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = value;
}

// ===========================================================================
// Agglomeration multigrid, FP64 (multigrid.cu:350-541)
// ===========================================================================

// Galerkin coarse operator by gather: one thread per coarse cell sums its
// members' diagonal + internal faces (ascending fine index) and its coarse
// row's off-diagonals from their fine slots (ascending fine slot).  Same
// additions as the reference's per-fine-row scatter, in a fixed order.
__global__ void __launch_bounds__(kBlock) kMGCoarseOp(DLevel f, DLevel c, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    double acc = 0.0;
    for (int j = f.memberStart[i]; j < f.memberStart[i + 1]; j++) {
        const int n = f.members[j];
        double diag = f.AC[n];
        for (int k = f.start[n]; k < f.start[n + 1]; k++) {
            if (f.nb[k] < 0) continue;
            if (f.slotToCoarseSlot[k] < 0) diag += f.AF[k];
        }
        acc += diag;
    }
    c.AC[i] = acc;
    for (int q = c.start[i]; q < c.start[i + 1]; q++) {
        double a = 0.0;
        for (int j = f.cslotStart[q]; j < f.cslotStart[q + 1]; j++) a += f.AF[f.cslotFine[j]];
        c.AF[q] = a;
    }
}

// The same Galerkin product in two coalesced passes: (1) one thread per fine
// cell forms diag_n = AC[n] + its internal off-diagonals (slot order) into
// f.res; (2) one thread per coarse cell sums diag over its members (ascending)
// and one thread per coarse slot sums its fine slots (ascending).  Same sums in
// the same order as kMGCoarseOp.
__global__ void __launch_bounds__(kBlock) kMGCoarseDiag(DLevel f, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
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
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
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

// computeResidual (residualAll sign = true, residuals.cuh:17-60) into f.res,
// one thread per fine cell; kMGRestrictSum then gathers it per coarse cell.
__global__ void __launch_bounds__(kBlock) kMGResidual(DLevel f, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= f.n) return;
    double Ax = f.AC[n] * f.x[n];
    for (int k = f.start[n]; k < f.start[n + 1]; k++) {
        const int nb = f.nb[k];
        if (nb >= 0) Ax += f.AF[k] * f.x[nb];
    }
    f.res[n] = f.b[n] - Ax;
}

// buildRestrictionKernel (multigrid.cu:398-410) as a gather in ascending fine
// order + the coarse x memset (multigrid.cu:650-655)
__global__ void __launch_bounds__(kBlock) kMGRestrictSum(DLevel f, DLevel c, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    double acc = 0.0;
    for (int j = f.memberStart[i]; j < f.memberStart[i + 1]; j++) acc += f.res[f.members[j]];
    c.b[i] = acc;
    c.x[i] = 0.0;
}

// jacobiRow, multigrid.cu:447-484
__device__ __forceinline__ double jacobiRow(const DLevel& L, const double* x, int n, double w) {
    const double AC = L.AC[n];
    if (fabs(AC) < 1.0e-30) return x[n];
    double Ax = AC * x[n];
    for (int k = L.start[n]; k < L.start[n + 1]; k++) {
        const int nb = L.nb[k];
        if (nb != -1) Ax += L.AF[k] * x[nb];
    }
    return x[n] + w * (L.b[n] - Ax) / AC;
}

// jacobiFused, multigrid.cu:486-493
__global__ void __launch_bounds__(kBlock) kMGJacobi(DLevel L, double w, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= L.n) return;
    L.xNew[n] = jacobiRow(L, L.x, n, w);
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

// computeResidual (residualAll sign=true) + buildRestrictionKernel
// (multigrid.cu:398-410) + the coarse x memset (multigrid.cu:650-655): per
// coarse cell, the residuals of its members summed in ascending fine order.
__global__ void __launch_bounds__(kBlock) kMGRestrict(DLevel f, DLevel c, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    double acc = 0.0;
    for (int j = f.memberStart[i]; j < f.memberStart[i + 1]; j++) {
        const int n = f.members[j];
        double Ax = f.AC[n] * f.x[n];
        for (int k = f.start[n]; k < f.start[n + 1]; k++) {
            const int nb = f.nb[k];
            if (nb >= 0) Ax += f.AF[k] * f.x[nb];
        }
        acc += f.b[n] - Ax;
    }
    c.b[i] = acc;
    c.x[i] = 0.0;
}

// buildProlongationKernel (multigrid.cu:415-427) fused into the first
// post-smoothing sweep: x + P x_c is formed on the fly for the cell and its
// neighbours (the same sums the prolongation would store).
__global__ void __launch_bounds__(kBlock) kMGProlongSmooth(DLevel f, DLevel c, double w, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= f.n) return;
    const double xn = f.x[n] + c.x[f.toCoarse[n]];
    const double AC = f.AC[n];
    if (fabs(AC) < 1.0e-30) {
        f.xNew[n] = xn;
        return;
    }
    double Ax = AC * xn;
    for (int k = f.start[n]; k < f.start[n + 1]; k++) {
        const int nb = f.nb[k];
        if (nb != -1) Ax += f.AF[k] * (f.x[nb] + c.x[f.toCoarse[nb]]);
    }
    f.xNew[n] = xn + w * (f.b[n] - Ax) / AC;
}

__global__ void __launch_bounds__(kBlock) kMGProlong(DLevel f, DLevel c, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= f.n) return;
    f.x[n] += c.x[f.toCoarse[n]];
}

// ===========================================================================
// Tail of the V-cycle: every level from `lt` down, in ONE workgroup.
// Coarse levels are tiny and launch-latency bound; here each sweep is a
// workgroup-wide loop + barrier instead of a launch.  Same arithmetic, in the
// same order, as kMGJacobi / kMGRestrict / kMGProlongSmooth / the coarsest
// solve; the x ping-pong is tracked per level in registers, and the answer of
// level lt is left in that level's own x buffer for the prolongation above.
// ===========================================================================
namespace {
template <class T>
struct LV {
    const T *AC, *AF;
    T *b, *x, *xNew;
};
template <class T>
__device__ __forceinline__ LV<T> view(const DLevel& L);
template <>
__device__ __forceinline__ LV<double> view<double>(const DLevel& L) {
    return LV<double>{L.AC, L.AF, L.b, L.x, L.xNew};
}
template <>
__device__ __forceinline__ LV<float> view<float>(const DLevel& L) {
    return LV<float>{L.AC32, L.AF32, L.b32, L.x32, L.xNew32};
}
template <class T>
__device__ __forceinline__ T tabs(T v) {
    return v < T(0) ? -v : v;
}
}  // namespace

template <class T>
__global__ void __launch_bounds__(1024) kMGTail(const DLevel* __restrict__ Ls, int lt, int L, double wd, int pre,
                                                int post, int coarseSweeps, const double* inv, int stage,
                                                const Ctl* ctl) {
    if (finished(ctl)) return;
    constexpr int kMax = 32;  // levels lt..L-1, L - lt <= kMax (checked by the host)
    const T w = (T)wd;
    const int t = threadIdx.x, T_ = blockDim.x;
    T* xa[kMax];
    T* xb[kMax];
    for (int l = lt; l < L; l++) {
        const LV<T> v = view<T>(Ls[l]);
        xa[l - lt] = v.x;
        xb[l - lt] = v.xNew;
    }
    auto sweeps = [&](int l, int count) {
        const DLevel& D = Ls[l];
        const LV<T> v = view<T>(D);
        for (int it = 0; it < count; it++) {
            const T* x = xa[l - lt];
            T* y = xb[l - lt];
            for (int n = t; n < D.n; n += T_) {
                const T AC = v.AC[n];
                T r;
                if (tabs(AC) < T(1.0e-30)) {
                    r = x[n];
                } else {
                    T Ax = AC * x[n];
                    for (int k = D.start[n]; k < D.start[n + 1]; k++) {
                        const int nb = D.nb[k];
                        if (nb != -1) Ax += v.AF[k] * x[nb];
                    }
                    r = x[n] + w * (v.b[n] - Ax) / AC;
                }
                y[n] = r;
            }
            __syncthreads();
            T* tmp = xa[l - lt];
            xa[l - lt] = xb[l - lt];
            xb[l - lt] = tmp;
        }
    };
    // descent
    for (int l = lt; l < L - 1; l++) {
        sweeps(l, pre);
        const DLevel& F = Ls[l];
        const DLevel& C = Ls[l + 1];
        const LV<T> f = view<T>(F);
        const LV<T> c = view<T>(C);
        const T* x = xa[l - lt];
        for (int i = t; i < C.n; i += T_) {
            T acc = T(0);
            for (int j = F.memberStart[i]; j < F.memberStart[i + 1]; j++) {
                const int n = F.members[j];
                T Ax = f.AC[n] * x[n];
                for (int k = F.start[n]; k < F.start[n + 1]; k++) {
                    const int nb = F.nb[k];
                    if (nb >= 0) Ax += f.AF[k] * x[nb];
                }
                acc += f.b[n] - Ax;
            }
            c.b[i] = acc;
            xa[l + 1 - lt][i] = T(0);
        }
        __syncthreads();
    }
    // coarsest
    {
        const DLevel& D = Ls[L - 1];
        if (inv != nullptr) {
            const LV<T> v = view<T>(D);
            for (int r = t; r < D.n; r += T_) {
                double acc = 0.0;
                for (int c = 0; c < D.n; c++) acc += inv[r * D.n + c] * (double)v.b[c];
                xa[L - 1 - lt][r] = (T)acc;
            }
            __syncthreads();
        } else if (stage) {
            // coarsest level staged in LDS for its sweeps: operator, CSR and the
            // x ping-pong never leave the CU
            extern __shared__ unsigned char smem[];
            const LV<T> v = view<T>(D);
            const int nC = D.n, rC = D.start[nC];
            T* sx = (T*)smem;
            T* sy = sx + nC;
            T* sAC = sy + nC;
            T* sb = sAC + nC;
            T* sAF = sb + nC;
            int* sst = (int*)(sAF + rC);
            int* snb = sst + nC + 1;
            const T* x0 = xa[L - 1 - lt];
            for (int n = t; n < nC; n += T_) {
                sx[n] = x0[n];
                sAC[n] = v.AC[n];
                sb[n] = v.b[n];
            }
            for (int n = t; n <= nC; n += T_) sst[n] = D.start[n];
            for (int k = t; k < rC; k += T_) {
                sAF[k] = v.AF[k];
                snb[k] = D.nb[k];
            }
            __syncthreads();
            for (int it = 0; it < coarseSweeps; it++) {
                for (int n = t; n < nC; n += T_) {
                    const T AC = sAC[n];
                    T r;
                    if (tabs(AC) < T(1.0e-30)) {
                        r = sx[n];
                    } else {
                        T Ax = AC * sx[n];
                        for (int k = sst[n]; k < sst[n + 1]; k++) {
                            const int nb = snb[k];
                            if (nb != -1) Ax += sAF[k] * sx[nb];
                        }
                        r = sx[n] + w * (sb[n] - Ax) / AC;
                    }
                    sy[n] = r;
                }
                __syncthreads();
                T* tmp = sx;
                sx = sy;
                sy = tmp;
            }
            T* xo = xa[L - 1 - lt];
            for (int n = t; n < nC; n += T_) xo[n] = sx[n];
            __syncthreads();
        } else {
            sweeps(L - 1, coarseSweeps);
        }
    }
    // ascent
    for (int l = L - 2; l >= lt; l--) {
        const DLevel& F = Ls[l];
        const LV<T> f = view<T>(F);
        const T* xc = xa[l + 1 - lt];
        if (post >= 1) {
            const T* x = xa[l - lt];
            T* y = xb[l - lt];
            for (int n = t; n < F.n; n += T_) {
                const T xn = x[n] + xc[F.toCoarse[n]];
                const T AC = f.AC[n];
                if (tabs(AC) < T(1.0e-30)) {
                    y[n] = xn;
                    continue;
                }
                T Ax = AC * xn;
                for (int k = F.start[n]; k < F.start[n + 1]; k++) {
                    const int nb = F.nb[k];
                    if (nb != -1) Ax += f.AF[k] * (x[nb] + xc[F.toCoarse[nb]]);
                }
                y[n] = xn + w * (f.b[n] - Ax) / AC;
            }
            __syncthreads();
            T* tmp = xa[l - lt];
            xa[l - lt] = xb[l - lt];
            xb[l - lt] = tmp;
            sweeps(l, post - 1);
        } else {
            T* x = xa[l - lt];
            for (int n = t; n < F.n; n += T_) x[n] += xc[F.toCoarse[n]];
            __syncthreads();
        }
    }
    // leave level lt's answer in its own x buffer
    const LV<T> top = view<T>(Ls[lt]);
    if (xa[0] != top.x)
        for (int n = t; n < Ls[lt].n; n += T_) top.x[n] = xa[0][n];
}

// ===========================================================================
// launchers
// ===========================================================================
void launchMGTail(hipStream_t s, const DLevel* levels, int lt, int L, double w, int pre, int post, int coarseSweeps,
                  bool fp32, const double* inv, size_t stageBytes, const Ctl* ctl) {
    const int stage = stageBytes > 0 ? 1 : 0;
    if (fp32) kMGTail<float><<<1, 1024, stageBytes, s>>>(levels, lt, L, w, pre, post, coarseSweeps, inv, stage, ctl);
    else kMGTail<double><<<1, 1024, stageBytes, s>>>(levels, lt, L, w, pre, post, coarseSweeps, inv, stage, ctl);
}
void launchMGCoarseDiag(hipStream_t s, const DLevel& f, const Ctl* ctl) {
    kMGCoarseDiag<<<blocksFor(f.n), kBlock, 0, s>>>(f, ctl);
}
void launchMGCoarseGather(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    kMGCoarseGather<<<blocksFor((long)c.refs + c.n), kBlock, 0, s>>>(f, c, ctl);
}
void launchMGResidual(hipStream_t s, const DLevel& f, const Ctl* ctl) {
    kMGResidual<<<blocksFor(f.n), kBlock, 0, s>>>(f, ctl);
}
void launchMGRestrictSum(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    kMGRestrictSum<<<blocksFor(c.n), kBlock, 0, s>>>(f, c, ctl);
}
void launchGradUV(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DState st, int scheme) {
    kGradUV<<<blocksFor(m.N), kBlock, 0, s>>>(m, bu, bv, st, scheme);
}
void launchGradField(hipStream_t s, const DMesh& m, DBC bc, const double* phi, double* gz, double* gr,
                     int scheme, const Ctl* ctl) {
    kGradField<<<blocksFor(m.N), kBlock, 0, s>>>(m, bc, phi, gz, gr, scheme, ctl);
}
void launchMomentum(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DBC bp, DState st, Physics ph) {
    kMomentum<<<blocksFor(m.N), kBlock, 0, s>>>(m, bu, bv, bp, st, ph);
}
void launchJacobiUV(hipStream_t s, const DMesh& m, DState st, const double* uo, const double* vo, double* un,
                    double* vn) {
    kJacobiUV<<<blocksFor(m.N), kBlock, 0, s>>>(m, st, uo, vo, un, vn);
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
    kPPCoeff<<<blocksFor(m.N), kBlock, 0, s>>>(m, bp, st, ph);
}
void launchFaceFlux(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DBC bp, DState st, Physics ph) {
    kFaceFlux<<<blocksFor(m.F), kBlock, 0, s>>>(m, bu, bv, bp, st, ph);
}
void launchPPRhs(hipStream_t s, const DMesh& m, DState st, Physics ph, int cross) {
    kPPRhs<<<blocksFor(m.N), kBlock, 0, s>>>(m, st, ph, cross);
}
void launchGradPPUpdate(hipStream_t s, const DMesh& m, DBC ppbc, DState st, Physics ph) {
    kGradPPUpdate<<<blocksFor(m.N), kBlock, 0, s>>>(m, ppbc, st, ph);
}
void launchMassFluxUpdate(hipStream_t s, const DMesh& m, DBC bp, DState st, Physics ph) {
    kMassFluxUpdate<<<blocksFor(m.F), kBlock, 0, s>>>(m, bp, st, ph);
}
void launchResidual(hipStream_t s, const DMesh& m, DState st, int blocks) {
    kResidual<<<blocks, kBlock, 0, s>>>(m, st);
}
void launchFinish(hipStream_t s, DState st, int blocks) { kFinish<<<1, kBlock, 0, s>>>(st, blocks); }
void launchStepBegin(hipStream_t s, const DMesh& m, DState st, int bdf2) {
    kStepBegin<<<blocksFor(m.N), kBlock, 0, s>>>(m, st, bdf2);
}
void launchCopy(hipStream_t s, double* dst, const double* src, int n, const Ctl* ctl) {
    kCopy<<<blocksFor(n), kBlock, 0, s>>>(dst, src, n, ctl);
}
void launchFill(hipStream_t s, double* dst, double value, int n, const Ctl* ctl) {
    kFill<<<blocksFor(n), kBlock, 0, s>>>(dst, value, n, ctl);
}
void launchMGCoarseOp(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    kMGCoarseOp<<<blocksFor(c.n), kBlock, 0, s>>>(f, c, ctl);
}
void launchMGJacobi(hipStream_t s, const DLevel& L, double w, const Ctl* ctl) {
    kMGJacobi<<<blocksFor(L.n), kBlock, 0, s>>>(L, w, ctl);
}
void launchMGSmoothBlock(hipStream_t s, const DLevel& L, double w, int sweeps, const Ctl* ctl) {
    const int threads = L.n >= 1024 ? 1024 : ((L.n + 63) / 64) * 64;
    kMGSmoothBlock<<<1, threads, 2 * (size_t)L.n * sizeof(double), s>>>(L, w, sweeps, ctl);
}
void launchMGRestrict(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    kMGRestrict<<<blocksFor(c.n), kBlock, 0, s>>>(f, c, ctl);
}
void launchMGProlongSmooth(hipStream_t s, const DLevel& f, const DLevel& c, double w, const Ctl* ctl) {
    kMGProlongSmooth<<<blocksFor(f.n), kBlock, 0, s>>>(f, c, w, ctl);
}
void launchMGProlong(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    kMGProlong<<<blocksFor(f.n), kBlock, 0, s>>>(f, c, ctl);
}

}  // namespace axs
