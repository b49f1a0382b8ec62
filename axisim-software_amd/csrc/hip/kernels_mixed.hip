// kernels_mixed.hip -- FP32 inner solves with FP64 defect correction.
//
// Build extension with no reference counterpart.  The momentum and pressure
// correction systems are still assembled in FP64 (kMomentum, kPPCoeff); only
// the inner iterations run in FP32, on the defect of the FP64 system:
//
//   momentum:  r0 = b - A u0 (FP64), then the reference's Jacobi sweeps on
//              delta (delta_0 = 0, delta_{k+1} = D^-1 (r0 - sum AF delta_k)) in
//              FP32, and u = u0 + delta in FP64.  In exact arithmetic this is the
//              same iterate as sweeping u itself.
//   p':        every V-cycle is x += B (b - A x) with the FP64 residual and an
//              FP32 V-cycle B started from zero; a linear multigrid cycle is
//              exactly that affine map, so again only the rounding differs.
//              The coarsest level is solved directly in FP64 (Gauss-Jordan
//              inverse, refreshed once per p' solve).
// Rounding errors of the FP32 parts scale with the defect, which goes to zero
// as SIMPLE converges, so the converged fields keep FP64 accuracy.
#include <cmath>
#include <stdexcept>

#include "device.hpp"
#include "mixed.hpp"
#include "rows.hpp"

namespace axs {

// ---------------------------------------------------------------------------
// momentum
// ---------------------------------------------------------------------------

// The row walks below batch a cell's slots as the FP64 kernels do (rows.hpp):
// B = 3 or 4 slots loaded in one round, then the neighbours in one round.

// r0 = b - A u0 for u and v (FP64), stored FP32, and the first sweep from zero:
// delta_1 = invAC * r0.
template <int B>
__global__ void __launch_bounds__(kBlock) kMomDefect(DMesh m, DState s, MixedState x) {
    if (finished(s.ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    double au = s.uAC[n] * s.u[n], av = s.vAC[n] * s.v[n];
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
    if (ns > 0 && ns <= B) {
        int o[B];
        double a[B], uo[B], vo[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = m.slotNb[k];
            a[q] = s.AFuv[k];
        }
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int oc = o[q] >= 0 ? o[q] : n;
            uo[q] = s.u[oc];
            vo[q] = s.v[oc];
        }
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns && o[q] >= 0) {
                au += a[q] * uo[q];
                av += a[q] * vo[q];
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
    const float ru = (float)(s.ub[n] - au), rv = (float)(s.vb[n] - av);
    const float2 iv = x.inv[n];
    x.r[n] = make_float2(ru, rv);
    x.d[n] = make_float2(iv.x * ru, iv.y * rv);
}

// one FP32 Jacobi sweep of the (u, v) defect equations
template <int B>
__device__ __forceinline__ float2 momSweep32(const DMesh& m, const MixedState& x, const float2* __restrict__ dO,
                                             int n) {
    const float2 r = x.r[n];
    float au = r.x, av = r.y;
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, n, k0, ns);
    if (ns > 0 && ns <= B) {
        int o[B];
        float a[B];
        float2 dv[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = m.slotNb[k];
            a[q] = x.AF32[k];
        }
#pragma unroll
        for (int q = 0; q < B; q++) dv[q] = dO[o[q] >= 0 ? o[q] : n];
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns && o[q] >= 0) {
                au -= a[q] * dv[q].x;
                av -= a[q] * dv[q].y;
            }
    } else {
        for (int k = k0; k < k0 + ns; k++) {
            const int nb = m.slotNb[k];
            if (nb >= 0) {
                const float a = x.AF32[k];
                const float2 d = dO[nb];
                au -= a * d.x;
                av -= a * d.y;
            }
        }
    }
    const float2 iv = x.inv[n];
    const float2 d = dO[n];
    return make_float2((iv.x != 0.0f) ? au * iv.x : d.x, (iv.y != 0.0f) ? av * iv.y : d.y);
}

template <int B>
__global__ void __launch_bounds__(kBlock) kMomSweep32(DMesh m, const Ctl* ctl, MixedState x,
                                                     const float2* __restrict__ dO, float2* __restrict__ dN) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    dN[n] = momSweep32<B>(m, x, dO, n);
}

// last FP32 sweep fused with u += delta, v += delta (FP64)
// Several FP32 sweeps in one launch: kJacobiUVBlock's temporal blocking
// (the same blocks, halo rings and LDS-slot tables) on the defect iterates,
// each row evaluated as momSweep32 evaluates it, so S sweeps here are
// bit-identical to S launches of kMomSweep32.
template <int R>
struct SweepRow32 {
    float af[R];
    int nb[R];   // neighbour cell, -1: none (boundary or past the row's slots)
    int loc[R];  // the neighbour's LDS slot, -1: none (loaded after the first sweep)
    int k0, ns;
    float2 r, inv;
};

template <int R>
__device__ __forceinline__ void loadSweepRow32(SweepRow32<R>& w, const DMesh& m, const MixedState& x, int c) {
    int k0, ns;
    rowRange(m.faceStart, m.uniformSlots, c, k0, ns);
    w.k0 = k0;
    w.ns = ns;
#pragma unroll
    for (int q = 0; q < R; q++) {
        const int k = ns > 0 ? k0 + min(q, ns - 1) : 0;  // clamped, unconditional reads
        const int nb = m.slotNb[k];
        const float af = x.AF32[k];
        w.nb[q] = (q < ns) ? nb : -1;
        w.af[q] = (q < ns) ? af : 0.0f;
    }
    w.r = x.r[c];
    w.inv = x.inv[c];
}

// momSweep32's row from registers; dat(q) = the neighbour of slot q (any valid
// entry where there is none: read, not used); LDS: the neighbours with an LDS slot
template <int R, bool LDS = false, typename DAt>
__device__ __forceinline__ float2 sweepRow32(const SweepRow32<R>& w, DAt dat, float2 self) {
    float2 dv[R];
#pragma unroll
    for (int q = 0; q < R; q++) dv[q] = dat(q);
    float au = w.r.x, av = w.r.y;
#pragma unroll
    for (int q = 0; q < R; q++)
        if (LDS ? w.loc[q] >= 0 : w.nb[q] >= 0) {
            au -= w.af[q] * dv[q].x;
            av -= w.af[q] * dv[q].y;
        }
    return make_float2((w.inv.x != 0.0f) ? au * w.inv.x : self.x, (w.inv.y != 0.0f) ? av * w.inv.y : self.y);
}

// The last launch of a solve can end in u += delta, v += delta (APPLY:
// kMomSweepApply's update after the last sweep).  The first launch starts from
// the FP64 defect r0 = b - A u0 that kMomentum wrote into x.r (DState::momR32):
// its first sweep is the pointwise delta_1 = r0 / AC (ZERO, kMomDefect's first
// sweep).
enum { kSweepPlain = 0, kSweepApply = 2, kSweepZero = 3 };

template <int C, int S, int R, int MODE>
__global__ void __launch_bounds__(C) kMomSweep32Block(DMesh m, DState st, MixedState x, DSweepBlocks sb,
                                                      const float2* __restrict__ dO, float2* __restrict__ dN) {
    constexpr int H = sweepHalo<C>();
    constexpr int kMask = S > 2 ? 1 : 0;  // sweep j < S writes LDS buffer (j - 1) & kMask
    __shared__ float2 buf[kMask + 1][C + H];
    if (finished(st.ctl)) return;
    const int t = threadIdx.x, blk = xcdBlock();
    const int a = blk * C, nOwn = min(C, m.N - a);
    const int h0 = sb.haloStart[blk];
    // halo rings 1 .. S - 1 are swept first; rings 1 .. S - 2 again from LDS
    const int nHalo1 = sb.ringEnd[blk * sb.rings + S - 2];
    const int nHaloLds = (S > 2) ? sb.ringEnd[blk * sb.rings + (S > 2 ? S - 3 : 0)] : 0;
    const int n = sb.own ? sb.own[(size_t)blk * C + t] : a + t;
    const bool mine = sb.own ? n >= 0 : t < nOwn, haloRow = t < nHalo1;
    const int hc = haloRow ? sb.halo[h0 + t] : 0;
    SweepRow32<R> own, hal;
    float2* out = buf[0];
    if (MODE == kSweepZero) {
        if (mine) {
            loadSweepRow32(own, m, x, n);
            out[t] = make_float2(own.inv.x * own.r.x, own.inv.y * own.r.y);
        }
        if (haloRow) {
            loadSweepRow32(hal, m, x, hc);
            out[C + t] = make_float2(hal.inv.x * hal.r.x, hal.inv.y * hal.r.y);
        }
    } else {
        if (mine) loadSweepRow32(own, m, x, n);
        if (haloRow) loadSweepRow32(hal, m, x, hc);
        if (mine) out[t] = sweepRow32(own, [&](int q) { return dO[own.nb[q] >= 0 ? own.nb[q] : n]; }, dO[n]);
        if (haloRow) out[C + t] = sweepRow32(hal, [&](int q) { return dO[hal.nb[q] >= 0 ? hal.nb[q] : hc]; }, dO[hc]);
    }
    if (mine) loadLoc(own, sb.locOwn + own.k0, R);
    if (t < nHaloLds) loadLoc(hal, sb.locHalo + (size_t)(h0 + t) * kSweepRegSlots, R);
    __syncthreads();
#pragma unroll
    for (int j = 2; j <= S; j++) {
        const float2* in = buf[(j - 2) & kMask];
        if (j == S) {
            if (!mine) return;
            const float2 d = sweepRow32<R, true>(own, [&](int q) { return in[own.loc[q] >= 0 ? own.loc[q] : t]; }, in[t]);
            if (MODE == kSweepApply) {
                st.u[n] += (double)d.x;
                st.v[n] += (double)d.y;
            } else {
                dN[n] = d;
            }
        } else {
            float2* o = buf[(j - 1) & kMask];
            const int nHaloJ = sb.ringEnd[blk * sb.rings + (S - j) - 1];
            if (mine) o[t] = sweepRow32<R, true>(own, [&](int q) { return in[own.loc[q] >= 0 ? own.loc[q] : t]; }, in[t]);
            if (t < nHaloJ)
                o[C + t] = sweepRow32<R, true>(
                    hal, [&](int q) { return in[hal.loc[q] >= 0 ? hal.loc[q] : C + t]; }, in[C + t]);
            __syncthreads();
        }
    }
}

template <int B>
__global__ void __launch_bounds__(kBlock) kMomSweepApply(DMesh m, DState s, MixedState x,
                                                        const float2* __restrict__ dO, int sweep) {
    if (finished(s.ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= m.N) return;
    const float2 d = sweep ? momSweep32<B>(m, x, dO, n) : dO[n];
    s.u[n] += (double)d.x;
    s.v[n] += (double)d.y;
}

// ---------------------------------------------------------------------------
// multigrid, FP32 levels
// ---------------------------------------------------------------------------

// Level-0 defect r = b - A x (FP64) -> FP32 right-hand side of the correction
// cycle; with pre >= 1 also the first weighted-Jacobi sweep from e = 0.
template <int B>
__global__ void __launch_bounds__(kBlock) kMGDefect(DLevel L, double w, int firstSweep, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= L.n) return;
    double Ax = L.AC[n] * L.x[n];
    int k0, ns;
    rowRange(L.start, L.uniformRow, n, k0, ns);
    if (ns > 0 && ns <= B) {
        int o[B];
        double a[B], xo[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int k = slotAt(k0, ns, q);
            o[q] = L.nb[k];
            a[q] = L.AF[k];
        }
#pragma unroll
        for (int q = 0; q < B; q++) xo[q] = L.x[o[q] >= 0 ? o[q] : n];
#pragma unroll
        for (int q = 0; q < B; q++)
            if (q < ns && o[q] >= 0) Ax += a[q] * xo[q];
    } else {
        for (int k = k0; k < k0 + ns; k++) {
            const int nb = L.nb[k];
            if (nb >= 0) Ax += L.AF[k] * L.x[nb];
        }
    }
    const float r = (float)(L.b[n] - Ax);
    L.b32[n] = r;
    if (firstSweep) {
        const float ac = L.AC32[n];
        L.x32[n] = (fabsf(ac) < 1.0e-30f) ? 0.0f : (float)w * r / ac;
    } else {
        L.x32[n] = 0.0f;
    }
}

__global__ void __launch_bounds__(1024) kMGSmoothBlock32(DLevel L, float w, int sweeps, const Ctl* ctl) {
    if (finished(ctl)) return;
    extern __shared__ float sx32[];
    float* xa = sx32;
    float* xb = sx32 + L.n;
    for (int n = threadIdx.x; n < L.n; n += blockDim.x) xa[n] = L.x32[n];
    __syncthreads();
    for (int it = 0; it < sweeps; it++) {
        for (int n = threadIdx.x; n < L.n; n += blockDim.x) {
            const float AC = L.AC32[n];
            float r;
            if (fabsf(AC) < 1.0e-30f) {
                r = xa[n];
            } else {
                float Ax = AC * xa[n];
                for (int k = L.start[n]; k < L.start[n + 1]; k++) {
                    const int nb = L.nb[k];
                    if (nb != -1) Ax += L.AF32[k] * xa[nb];
                }
                r = xa[n] + w * (L.b32[n] - Ax) / AC;
            }
            xb[n] = r;
        }
        __syncthreads();
        float* t = xa;
        xa = xb;
        xb = t;
    }
    for (int n = threadIdx.x; n < L.n; n += blockDim.x) L.x32[n] = xa[n];
}

__global__ void __launch_bounds__(kBlock) kMGProlong32(DLevel f, DLevel c, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= f.n) return;
    f.x32[n] += c.x32[f.toCoarse[n]];
}

// x (FP64) += e (FP32): the refinement step closing each correction cycle
__global__ void __launch_bounds__(kBlock) kMGApply(DLevel L, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n >= L.n) return;
    L.x[n] += (double)L.x32[n];
}

// FP32 copies of a level operator (level 0 each p' solve; coarse levels are
// cast inside kMGCoarseOp64)
__global__ void __launch_bounds__(kBlock) kMGCast(DLevel L, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = xcdBlock() * blockDim.x + threadIdx.x;
    if (n < L.n) L.AC32[n] = (float)L.AC[n];
    for (int k = n; k < L.refs; k += gridDim.x * blockDim.x) L.AF32[k] = (float)L.AF[k];
}

// ---------------------------------------------------------------------------
// coarsest level: FP64 dense inverse (Gauss-Jordan, partial pivoting), one
// workgroup; n <= kDirectOneBlock.  inv is n x n row-major scratch, aug 2 n x n.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) kCoarseInvert(DLevel L, double* aug, double* inv, int* status,
                                                     const Ctl* ctl) {
    if (finished(ctl)) return;
    const int n = L.n, W = 2 * n, t = threadIdx.x, T = blockDim.x;
    __shared__ int pivRow;
    __shared__ double pivVal;
    __shared__ double redV[1024];
    __shared__ int redI[1024];
    for (int e = t; e < n * W; e += T) {
        const int r = e / W, col = e % W;
        aug[e] = (col == n + r) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int r = t; r < n; r += T) {
        aug[r * W + r] = L.AC[r];
        for (int k = L.start[r]; k < L.start[r + 1]; k++)
            if (L.nb[k] >= 0) aug[r * W + L.nb[k]] += L.AF[k];
    }
    __syncthreads();
    int singular = 0;
    for (int p = 0; p < n; p++) {
        double best = -1.0;
        int bi = p;
        for (int r = p + t; r < n; r += T) {
            const double v = fabs(aug[r * W + p]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        redV[t] = best;
        redI[t] = bi;
        __syncthreads();
        for (int st = T / 2; st > 0; st >>= 1) {
            if (t < st && (redV[t + st] > redV[t] || (redV[t + st] == redV[t] && redI[t + st] < redI[t]))) {
                redV[t] = redV[t + st];
                redI[t] = redI[t + st];
            }
            __syncthreads();
        }
        if (t == 0) {
            pivRow = redI[0];
            pivVal = aug[redI[0] * W + p];
        }
        __syncthreads();
        if (pivVal == 0.0) {
            singular = 1;
            break;
        }
        if (pivRow != p)
            for (int col = t; col < W; col += T) {
                const double a = aug[p * W + col];
                aug[p * W + col] = aug[pivRow * W + col];
                aug[pivRow * W + col] = a;
            }
        __syncthreads();
        const double inv0 = 1.0 / pivVal;
        for (int col = t; col < W; col += T) aug[p * W + col] *= inv0;
        __syncthreads();
        for (int e = t; e < n * W; e += T) {
            const int r = e / W, col = e % W;
            if (r == p || col == p) continue;
            aug[e] -= aug[r * W + p] * aug[p * W + col];
        }
        __syncthreads();
        for (int r = t; r < n; r += T)
            if (r != p) aug[r * W + p] = 0.0;
        __syncthreads();
    }
    for (int e = t; e < n * n; e += T) inv[e] = aug[(e / n) * W + n + (e % n)];
    if (t == 0) *status = singular;
}

// Larger coarsest levels: blocked in-place Gauss-Jordan inversion of the
// dense operator, padded to N (a multiple of kGJ) with the identity, in tiles
// of kGJ x kGJ.  No pivoting: the coarsest operator is symmetric positive
// definite (a Galerkin product of the p' matrix), and so is its padding.
// Block step k with P = (M_kk)^-1:
//   M_kk <- P,  M_kj <- P M_kj,  M_ik <- -M_ik P,
//   M_ij <- M_ij - M_ik (P M_kj)   (i, j != k)
// after the last step M holds the inverse.  One launch per step, one wave per
// tile, reading M and writing its copy (no tile is read after it is
// rewritten); the tile (k+1, k+1) also inverts its new value for the next step.
constexpr int kGJ = 32;

// the dense, padded operator: row r of M (ld N) per workgroup
__global__ void __launch_bounds__(kBlock) kCoarseDense(DLevel L, double* M, int N, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int r = blockIdx.x, n = L.n;
    for (int c = threadIdx.x; c < N; c += blockDim.x)
        M[(size_t)r * N + c] = (c == r) ? (r < n ? L.AC[r] : 1.0) : 0.0;
    __syncthreads();
    if (threadIdx.x == 0 && r < n)
        for (int k = L.start[r]; k < L.start[r + 1]; k++)
            if (L.nb[k] >= 0) M[(size_t)r * N + L.nb[k]] += L.AF[k];
}

// One workgroup of kGJThreads per tile; thread t holds the 2 x 2 block of rows
// 2 (t / 16) + {0, 1}, columns 2 (t % 16) + {0, 1}.  Tiles sit in LDS with a
// row stride of kGJS doubles (16-byte aligned rows); a left operand is stored
// transposed, so that both operands of a product step are read as pairs.
constexpr int kGJThreads = 256;
constexpr int kGJS = kGJ + 2;

// global tiles (row stride ld) -> LDS, as they are or transposed: this
// thread's elements of a tile are loaded into registers first (gjFetch), so
// that the loads of every tile a launch needs are in flight together, then
// stored (gjStore)
constexpr int kGJQ = kGJ * kGJ / kGJThreads;  // elements of a tile per thread
__device__ __forceinline__ void gjFetch(double v[kGJQ], const double* src, int ld) {
#pragma unroll
    for (int q = 0; q < kGJQ; q++) {
        const int e = threadIdx.x + kGJThreads * q;
        v[q] = src[(size_t)(e / kGJ) * ld + e % kGJ];
    }
}
template <bool TRANSPOSE>
__device__ __forceinline__ void gjStore(double* T, const double v[kGJQ]) {
#pragma unroll
    for (int q = 0; q < kGJQ; q++) {
        const int e = threadIdx.x + kGJThreads * q, r = e / kGJ, c = e % kGJ;
        T[TRANSPOSE ? c * kGJS + r : r * kGJS + c] = v[q];
    }
}
template <bool TRANSPOSE>
__device__ __forceinline__ void gjLoad(double* T, const double* src, int ld) {
    double v[kGJQ];
    gjFetch(v, src, ld);
    gjStore<TRANSPOSE>(T, v);
}

// acc = A B for this thread's 2 x 2 block (At: A transposed); the sums run
// over t = 0 .. kGJ - 1 in order
__device__ __forceinline__ void gjProduct(const double* At, const double* B, double acc[2][2]) {
    const int r0 = 2 * (threadIdx.x >> 4), c0 = 2 * (threadIdx.x & 15);
    acc[0][0] = acc[0][1] = acc[1][0] = acc[1][1] = 0.0;
#pragma unroll 8
    for (int t = 0; t < kGJ; t++) {
        const double2 x = *(const double2*)&At[t * kGJS + r0];
        const double2 y = *(const double2*)&B[t * kGJS + c0];
        acc[0][0] += x.x * y.x;
        acc[0][1] += x.x * y.y;
        acc[1][0] += x.y * y.x;
        acc[1][1] += x.y * y.y;
    }
}

// 4 x 4 in-place Gauss-Jordan inverse in registers (no pivoting)
__device__ __forceinline__ void inv4(double q[4][4]) {
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const double d = 1.0 / q[p][p];
        double r[4], c[4];
#pragma unroll
        for (int j = 0; j < 4; j++) r[j] = q[p][j];
#pragma unroll
        for (int i = 0; i < 4; i++) c[i] = q[i][p];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (i == p) q[i][j] = (j == p) ? d : r[j] * d;
                else if (j == p) q[i][j] = -(c[i] * d);
                else q[i][j] = q[i][j] - c[i] * (r[j] * d);
            }
    }
}

// In-place blocked Gauss-Jordan inverse of the LDS tile T (row stride kGJS)
// by the whole workgroup, with 4 x 4 pivot blocks: each pivot block is
// inverted by every thread in registers, then each thread updates its 2 x 2
// block from the pivot rows and columns it read before the barrier.
__device__ void gjInvertTile(double* T) {
    const int r0 = 2 * (threadIdx.x >> 4), c0 = 2 * (threadIdx.x & 15);
    double own[2][2];  // this thread's block, current between pivot blocks
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) own[i][j] = T[(r0 + i) * kGJS + c0 + j];
    for (int p0 = 0; p0 < kGJ; p0 += 4) {
        double q[4][4], tr[2][4], tc[4][2];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) q[i][j] = T[(p0 + i) * kGJS + p0 + j];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int s2 = 0; s2 < 4; s2++) tr[i][s2] = T[(r0 + i) * kGJS + p0 + s2];
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++)
#pragma unroll
            for (int j = 0; j < 2; j++) tc[s2][j] = T[(p0 + s2) * kGJS + c0 + j];
        inv4(q);
        const bool rin = r0 >= p0 && r0 < p0 + 4, cin = c0 >= p0 && c0 < p0 + 4;
        const bool rlo = r0 == p0, clo = c0 == p0;  // which half of the pivot block, where inside
        // this thread's rows of the pivot block's inverse (where rin) and its
        // columns (where cin), picked by value so that q stays in registers
        double qrow[2][4], qcol[4][2];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int u = 0; u < 4; u++) qrow[i][u] = rlo ? q[i][u] : q[2 + i][u];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int j = 0; j < 2; j++) qcol[u][j] = clo ? q[u][j] : q[u][2 + j];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                double v;
                if (rin && cin) {
                    v = clo ? qrow[i][j] : qrow[i][2 + j];
                } else if (rin) {
                    v = 0.0;
#pragma unroll
                    for (int u = 0; u < 4; u++) v += qrow[i][u] * tc[u][j];
                } else if (cin) {
                    v = 0.0;
#pragma unroll
                    for (int u = 0; u < 4; u++) v += tr[i][u] * qcol[u][j];
                    v = -v;
                } else {
                    double w = 0.0;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        double ws = 0.0;
#pragma unroll
                        for (int x = 0; x < 4; x++) ws += q[u][x] * tc[x][j];
                        w += tr[i][u] * ws;
                    }
                    v = own[i][j] - w;
                }
                own[i][j] = v;
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) T[(r0 + i) * kGJS + c0 + j] = own[i][j];
        __syncthreads();
    }
}

// P0 = (M_00)^-1
__global__ void __launch_bounds__(kGJThreads) kGJPivot0(const double* M, int N, double* P, const Ctl* ctl) {
    __shared__ __align__(16) double T[kGJ * kGJS];
    if (finished(ctl)) return;
    gjLoad<false>(T, M, N);
    __syncthreads();
    gjInvertTile(T);
    const int r0 = 2 * (threadIdx.x >> 4), c0 = 2 * (threadIdx.x & 15);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) P[(r0 + i) * kGJ + c0 + j] = T[(r0 + i) * kGJS + c0 + j];
}

__global__ void __launch_bounds__(kGJThreads) kGJStep(const double* __restrict__ M, double* __restrict__ Mn, int N,
                                                      int k, const double* __restrict__ P, double* __restrict__ Pn,
                                                      const Ctl* ctl) {
    __shared__ __align__(16) double A[kGJ * kGJS], B[kGJ * kGJS], C[kGJ * kGJS], U[kGJ * kGJS];
    if (finished(ctl)) return;
    // workgroup 0 takes the tile (k + 1, k + 1), whose launch is the longest
    // (it also inverts the next pivot block), so that it starts first
    const int K = N / kGJ, kn = k + 1 < K ? k + 1 : 0;
    const int id = blockIdx.x == 0 ? kn * K + kn : (blockIdx.x == kn * K + kn ? 0 : (int)blockIdx.x);
    const int ti = id / K, tj = id % K;
    const int r0 = 2 * (threadIdx.x >> 4), c0 = 2 * (threadIdx.x & 15);
    double* out = Mn + (size_t)ti * kGJ * N + (size_t)tj * kGJ;
    const double* Mij = M + (size_t)ti * kGJ * N + (size_t)tj * kGJ;
    const double* Mkj = M + (size_t)k * kGJ * N + (size_t)tj * kGJ;
    const double* Mik = M + (size_t)ti * kGJ * N + (size_t)k * kGJ;
    auto put = [&](double v[2][2], double sg) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) out[(size_t)(r0 + i) * N + c0 + j] = sg * v[i][j];
    };
    double acc[2][2];
    if (ti == k && tj == k) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = P[(r0 + i) * kGJ + c0 + j];
        put(acc, 1.0);
        return;
    }
    double va[kGJQ], vb[kGJQ], vc[kGJQ];
    if (ti == k) {  // P M_kj
        gjFetch(va, P, kGJ);
        gjFetch(vb, Mkj, N);
        gjStore<true>(A, va);
        gjStore<false>(B, vb);
        __syncthreads();
        gjProduct(A, B, acc);
        put(acc, 1.0);
        return;
    }
    if (tj == k) {  // -M_ik P
        gjFetch(va, Mik, N);
        gjFetch(vb, P, kGJ);
        gjStore<true>(A, va);
        gjStore<false>(B, vb);
        __syncthreads();
        gjProduct(A, B, acc);
        put(acc, -1.0);
        return;
    }
    // M_ij - M_ik (P M_kj)
    double mij[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) mij[i][j] = Mij[(size_t)(r0 + i) * N + c0 + j];
    gjFetch(va, P, kGJ);
    gjFetch(vb, Mkj, N);
    gjFetch(vc, Mik, N);
    gjStore<true>(A, va);
    gjStore<false>(B, vb);
    gjStore<true>(C, vc);
    __syncthreads();
    gjProduct(A, B, acc);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) U[(r0 + i) * kGJS + c0 + j] = acc[i][j];
    __syncthreads();
    gjProduct(C, U, acc);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) mij[i][j] -= acc[i][j];
    put(mij, 1.0);
    if (!(ti == k + 1 && tj == k + 1)) return;
    // the next step's pivot block, inverted here (B is free again)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) B[(r0 + i) * kGJS + c0 + j] = mij[i][j];
    __syncthreads();
    gjInvertTile(B);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) Pn[(r0 + i) * kGJ + c0 + j] = B[(r0 + i) * kGJS + c0 + j];
}

// x_c = A_c^-1 b_c with one wave per row (FP64 products summed in a fixed
// order: lane l takes columns l, l + 64, ..., then a fixed shuffle tree);
// inv has row stride ld.  NC >= ceil(n / 64) column rounds per lane, all
// loaded at once (indices clamped into the row, the extra ones not summed).
// T: the level's value type (float for the FP32 cycles, double for FP64 ones).
template <int NC, typename T>
__global__ void __launch_bounds__(kBlock) kCoarseApplyWide(int n, const double* inv, int ld, const T* __restrict__ b,
                                                           T* __restrict__ xOut, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int lane = threadIdx.x & 63, r = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (r >= n) return;  // the whole wave
    const double* row = inv + (size_t)r * ld;
    double a[NC];
    T x[NC];
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const int c = min(lane + 64 * i, n - 1);
        a[i] = row[c];
        x[i] = b[c];
    }
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < NC; i++)
        if (lane + 64 * i < n) acc += a[i] * (double)x[i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) xOut[r] = (T)acc;
}

// x_c = A_c^-1 b_c (FP64 product, FP32 in / out)
__global__ void __launch_bounds__(kBlock) kCoarseApply(DLevel L, const double* inv, const Ctl* ctl) {
    if (finished(ctl)) return;
    const int r = xcdBlock() * blockDim.x + threadIdx.x;
    if (r >= L.n) return;
    double acc = 0.0;
    for (int c = 0; c < L.n; c++) acc += inv[r * L.n + c] * (double)L.b32[c];
    L.x32[r] = (float)acc;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launchMomDefect(hipStream_t s, const DMesh& m, DState st, MixedState x) {
    if (m.maxSlots <= 3) kMomDefect<3><<<blocksFor(m.N), kBlock, 0, s>>>(m, st, x);
    else kMomDefect<4><<<blocksFor(m.N), kBlock, 0, s>>>(m, st, x);
}
void launchMomSweep32(hipStream_t s, const DMesh& m, const Ctl* ctl, MixedState x, const float2* dO, float2* dN) {
    if (m.maxSlots <= 3) kMomSweep32<3><<<blocksFor(m.N), kBlock, 0, s>>>(m, ctl, x, dO, dN);
    else kMomSweep32<4><<<blocksFor(m.N), kBlock, 0, s>>>(m, ctl, x, dO, dN);
}
void launchMomSweepApply(hipStream_t s, const DMesh& m, DState st, MixedState x, const float2* dO, int sweep) {
    if (m.maxSlots <= 3) kMomSweepApply<3><<<blocksFor(m.N), kBlock, 0, s>>>(m, st, x, dO, sweep);
    else kMomSweepApply<4><<<blocksFor(m.N), kBlock, 0, s>>>(m, st, x, dO, sweep);
}
void launchMGDefect(hipStream_t s, const DLevel& L, double w, int firstSweep, const Ctl* ctl) {
    const dim3 g(blocksFor(L.n, L.block)), b(L.block);
    if (L.maxRow <= 4) kMGDefect<4><<<g, b, 0, s>>>(L, w, firstSweep, ctl);
    else kMGDefect<8><<<g, b, 0, s>>>(L, w, firstSweep, ctl);
}
void launchMGSmoothBlock32(hipStream_t s, const DLevel& L, float w, int sweeps, const Ctl* ctl) {
    const int threads = L.n >= 1024 ? 1024 : ((L.n + 63) / 64) * 64;
    kMGSmoothBlock32<<<1, threads, 2 * (size_t)L.n * sizeof(float), s>>>(L, w, sweeps, ctl);
}
void launchMGProlong32(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl) {
    kMGProlong32<<<blocksFor(f.n, f.block), f.block, 0, s>>>(f, c, ctl);
}
void launchMGApply(hipStream_t s, const DLevel& L, const Ctl* ctl) {
    kMGApply<<<blocksFor(L.n, L.block), L.block, 0, s>>>(L, ctl);
}
void launchMGCast(hipStream_t s, const DLevel& L, const Ctl* ctl) {
    kMGCast<<<blocksFor(L.n, L.block), L.block, 0, s>>>(L, ctl);
}
int coarseInverseLd(int n) { return n <= kDirectOneBlock ? n : (n + kGJ - 1) / kGJ * kGJ; }
size_t coarseInverseWork(int n) {
    const size_t N = (size_t)coarseInverseLd(n);
    return n <= kDirectOneBlock ? 2 * N * N : N * N + 2 * (size_t)kGJ * kGJ;
}
void launchCoarseInvert(hipStream_t s, const DLevel& L, double* aug, double* inv, int* status, const Ctl* ctl) {
    if (L.n <= kDirectOneBlock) {
        kCoarseInvert<<<1, 1024, 0, s>>>(L, aug, inv, status, ctl);
        return;
    }
    // aug: the dense operator (N x N) and two pivot inverses; the steps
    // alternate between it and inv, so that the last one writes inv
    const int N = coarseInverseLd(L.n), K = N / kGJ;
    double* P[2] = {aug + (size_t)N * N, aug + (size_t)N * N + kGJ * kGJ};
    double* m[2] = {K % 2 == 0 ? inv : aug, K % 2 == 0 ? aug : inv};
    kCoarseDense<<<N, kBlock, 0, s>>>(L, m[0], N, ctl);
    kGJPivot0<<<1, kGJThreads, 0, s>>>(m[0], N, P[0], ctl);
    for (int k = 0; k < K; k++)
        kGJStep<<<K * K, kGJThreads, 0, s>>>(m[k & 1], m[(k + 1) & 1], N, k, P[k & 1], P[(k + 1) & 1], ctl);
}
static_assert(kDirectMax <= 32 * 64, "kCoarseApplyWide<32> covers at most 2048 columns");
namespace {
template <typename T>
void coarseApply(hipStream_t s, int n, const double* inv, const T* b, T* x, const Ctl* ctl) {
    if (n > kDirectMax) throw std::invalid_argument("coarsest level too large for the dense FP64 solve");
    const dim3 g((n + kBlock / 64 - 1) / (kBlock / 64)), blk(kBlock);
    const int ld = coarseInverseLd(n);
    if (n <= 256) kCoarseApplyWide<4, T><<<g, blk, 0, s>>>(n, inv, ld, b, x, ctl);
    else if (n <= 512) kCoarseApplyWide<8, T><<<g, blk, 0, s>>>(n, inv, ld, b, x, ctl);
    else if (n <= 1024) kCoarseApplyWide<16, T><<<g, blk, 0, s>>>(n, inv, ld, b, x, ctl);
    else kCoarseApplyWide<32, T><<<g, blk, 0, s>>>(n, inv, ld, b, x, ctl);
}
}  // namespace
void launchCoarseApply(hipStream_t s, const DLevel& L, const double* inv, const Ctl* ctl) {
    coarseApply<float>(s, L.n, inv, L.b32, L.x32, ctl);
}
void launchCoarseApply64(hipStream_t s, const DLevel& L, const double* inv, const Ctl* ctl) {
    coarseApply<double>(s, L.n, inv, L.b, L.x, ctl);
}

namespace {
template <int C, int S, int MODE>
void launchMomSweep32BlockCS(hipStream_t s, const DMesh& m, const DState& st, MixedState x, const DSweepBlocks& sb,
                             const float2* dO, float2* dN) {
    if (m.maxSlots <= 3) kMomSweep32Block<C, S, 3, MODE><<<sb.nBlocks, C, 0, s>>>(m, st, x, sb, dO, dN);
    else kMomSweep32Block<C, S, kSweepRegSlots, MODE><<<sb.nBlocks, C, 0, s>>>(m, st, x, sb, dO, dN);
}
template <int C, int MODE>
void launchMomSweep32BlockC(hipStream_t s, const DMesh& m, const DState& st, MixedState x, const DSweepBlocks& sb,
                            int sweeps, const float2* dO, float2* dN) {
    if (sweeps == 2) launchMomSweep32BlockCS<C, 2, MODE>(s, m, st, x, sb, dO, dN);
    else if (sweeps == 3) launchMomSweep32BlockCS<C, 3, MODE>(s, m, st, x, sb, dO, dN);
    else if (sweeps == 4) launchMomSweep32BlockCS<C, 4, MODE>(s, m, st, x, sb, dO, dN);
    else launchMomSweep32BlockCS<C, 5, MODE>(s, m, st, x, sb, dO, dN);
}
template <int MODE>
void launchMomSweep32BlockM(hipStream_t s, const DMesh& m, const DState& st, MixedState x, const DSweepBlocks& sb,
                            int sweeps, const float2* dO, float2* dN) {
    if (sweeps < 2 || sweeps > kMaxFusedSweeps || sweeps > sb.rings + 1)
        throw std::invalid_argument("bad fused sweep count");
    if (sb.cells == 512) launchMomSweep32BlockC<512, MODE>(s, m, st, x, sb, sweeps, dO, dN);
    else if (sb.cells == 1024) launchMomSweep32BlockC<1024, MODE>(s, m, st, x, sb, sweeps, dO, dN);
    else throw std::invalid_argument("fused sweeps take blocks of 512 or 1024 cells");
}
}  // namespace
void launchMomSweep32Block(hipStream_t s, const DMesh& m, const DState& st, MixedState x, const DSweepBlocks& sb,
                           int sweeps, const float2* dO, float2* dN) {
    launchMomSweep32BlockM<kSweepPlain>(s, m, st, x, sb, sweeps, dO, dN);
}
void launchMomSweep32BlockApply(hipStream_t s, const DMesh& m, const DState& st, MixedState x,
                                const DSweepBlocks& sb, int sweeps, const float2* dO) {
    launchMomSweep32BlockM<kSweepApply>(s, m, st, x, sb, sweeps, dO, nullptr);
}
void launchMomSweep32BlockFromDefect(hipStream_t s, const DMesh& m, const DState& st, MixedState x,
                                     const DSweepBlocks& sb, int sweeps, float2* dN) {
    launchMomSweep32BlockM<kSweepZero>(s, m, st, x, sb, sweeps, nullptr, dN);
}

}  // namespace axs
