// mpnn.hip -- message-passing network on a cell graph: forward and backward
// kernels (build extension with no reference counterpart).
//
// The network maps the right-hand side of a linear system A x = b on a cell
// graph to a guess of x.  Graph weights are the Jacobi-scaled off-diagonals
// w_ik = -AF_ik / AC_i.  Layer l:
//     A_l = [ H_{l-1} | M_{l-1} ],  M_i = sum_k w_ik H_{l-1, nb(k)}
//     H_l = ReLU(A_l W_l)
// then y_i = H_L,i . w_out and x_i = sigma * y_i / AC_i, with sigma the RMS of
// b/AC (every layer is positively homogeneous, so x scales with b).  Inputs
// are [relu(r), relu(-r), relu(m), relu(-m)] with r = b/(AC sigma), m = sum w r.
//
// Storage: activations and weights BF16, every product accumulated in FP32 on
// the matrix cores (v_mfma_f32_16x16x16_bf16), gradients FP32.  Rows are cells,
// padded to a multiple of 16; padded rows are zero and stay zero.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mpnn.hpp"

namespace axs {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (NaN kept quiet)
__device__ __forceinline__ uint16_t toBF16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float fromBF16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ double blockSum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
    __syncthreads();
    return s;  // valid in thread 0
}

}  // namespace

// ---------------------------------------------------------------------------
// graph weights: w_ik = -AF_ik / AC_i (0 on boundary slots and empty rows), and
// the transposed weights wT at slot k of cell j: the weight of j in the row of
// nb(k) (rev[k] = that slot)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kNetWeights(NetGraph g, const double* __restrict__ AC,
                                                   const double* __restrict__ AF) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const double ac = AC[i];
    for (int k = g.start[i]; k < g.start[i + 1]; k++) {
        const bool ok = g.nb[k] >= 0 && ac != 0.0;
        g.w[k] = ok ? (float)(-AF[k] / ac) : 0.0f;
    }
}
__global__ void __launch_bounds__(256) kNetWeightsT(NetGraph g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.refs) return;
    const int r = g.rev[k];
    g.wT[k] = r >= 0 ? g.w[r] : 0.0f;
}

// per-block partial sums of (b/AC)^2 (sigma^2 = mean)
__global__ void __launch_bounds__(256) kNetScalePartial(int n, const double* __restrict__ AC,
                                                        const double* __restrict__ b, double* partial) {
    __shared__ double red[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n && AC[i] != 0.0) {
        const double r = b[i] / AC[i];
        v = r * r;
    }
    const double s = blockSum(v, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
// sigma = sqrt(sum / n) (1 when the right-hand side is zero), or with n = 0
// the plain sum; the partials are summed in block order
__global__ void __launch_bounds__(64) kNetScaleFinish(int n, int blocks, const double* partial, double* out) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < blocks; i++) s += partial[i];
    if (n == 0) {
        out[0] = s;
        return;
    }
    const double v = sqrt(s / (double)n);
    out[0] = v > 0.0 ? v : 1.0;
}

// r_i = b_i / (AC_i sigma)
__global__ void __launch_bounds__(256) kNetScaledRhs(int n, const double* __restrict__ AC,
                                                     const double* __restrict__ b, const double* sigma, float* r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    r[i] = AC[i] != 0.0 ? (float)(b[i] / (AC[i] * sigma[0])) : 0.0f;
}

// input features [relu(r), relu(-r), relu(m), relu(-m), 0 ...] (kNetIn wide)
__global__ void __launch_bounds__(256) kNetFeatures(NetGraph g, const float* __restrict__ r, uint16_t* H0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nPad) return;
    uint16_t* row = H0 + (size_t)i * kNetIn;
    float ri = 0.0f, m = 0.0f;
    if (i < g.n) {
        ri = r[i];
        for (int k = g.start[i]; k < g.start[i + 1]; k++) {
            const int j = g.nb[k];
            if (j >= 0) m += g.w[k] * r[j];
        }
    }
    row[0] = toBF16(fmaxf(ri, 0.0f));
    row[1] = toBF16(fmaxf(-ri, 0.0f));
    row[2] = toBF16(fmaxf(m, 0.0f));
    row[3] = toBF16(fmaxf(-m, 0.0f));
    for (int f = 4; f < kNetIn; f++) row[f] = 0;
}

// A = [H | M], M_i = sum_k w_ik H_nb(k) (FP32 sum in slot order, stored BF16);
// one thread per (cell, feature)
__global__ void __launch_bounds__(256) kNetAggregate(NetGraph g, const uint16_t* __restrict__ H, int D,
                                                     uint16_t* A) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.nPad * D) return;
    const int i = (int)(t / D), f = (int)(t % D);
    uint16_t* row = A + (size_t)i * 2 * D;
    row[f] = H[(size_t)i * D + f];
    float m = 0.0f;
    if (i < g.n)
        for (int k = g.start[i]; k < g.start[i + 1]; k++) {
            const int j = g.nb[k];
            if (j >= 0) m += g.w[k] * fromBF16(H[(size_t)j * D + f]);
        }
    row[D + f] = toBF16(m);
}

// C[M x N] = A[M x K] B[K x N], BF16 in, FP32 accumulate on the matrix cores.
// One wave per 16 x 16 tile of C; a 256-thread workgroup covers 16 rows x 64
// columns.  mode 0: C FP32; mode 1: relu(C) as BF16.  M, K multiples of 16,
// N a multiple of 64.
__global__ void __launch_bounds__(256) kNetGemm(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                int M, int N, int K, int mode, float* Cf, uint16_t* Cb) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int m0 = blockIdx.x * 16;
    const int n0 = (blockIdx.y * 4 + wave) * 16;
    if (m0 >= M || n0 >= N) return;
    const int lr = l & 15, lk = (l >> 4) * 4;
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        v4s a, b;
        const uint16_t* ap = A + (size_t)(m0 + lr) * K + k0 + lk;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            a[i] = (short)ap[i];
            b[i] = (short)B[(size_t)(k0 + lk + i) * N + n0 + lr];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const size_t o = (size_t)(m0 + lk + i) * N + n0 + lr;
        if (mode == 0) Cf[o] = acc[i];
        else Cb[o] = toBF16(fmaxf(acc[i], 0.0f));
    }
}

// Partial dW = A^T dZ over one chunk of kNetChunk cells: C[K x N] with K = the
// features of A, N = 64, the cells the reduction dimension.  One workgroup per
// chunk; each wave walks the 16 x 16 output tiles tile = wave, wave + 4, ...
__global__ void __launch_bounds__(256) kNetGemmAtB(const uint16_t* __restrict__ A, const uint16_t* __restrict__ dZ,
                                                   int rows, int K, int N, float* partial) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int c0 = blockIdx.x * kNetChunk;
    const int c1 = min(rows, c0 + kNetChunk);
    const int lr = l & 15, lk = (l >> 4) * 4;
    const int tilesM = K / 16, tilesN = N / 16;
    float* out = partial + (size_t)blockIdx.x * K * N;
    for (int tile = wave; tile < tilesM * tilesN; tile += 4) {
        const int f0 = (tile / tilesN) * 16, n0 = (tile % tilesN) * 16;
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int c = c0; c < c1; c += 16) {
            v4s a, b;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int cell = c + lk + i;  // rows are padded to 16, chunks are multiples of 16
                a[i] = (short)A[(size_t)cell * K + f0 + lr];
                b[i] = (short)dZ[(size_t)cell * N + n0 + lr];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) out[(size_t)(f0 + lk + i) * N + n0 + lr] = acc[i];
    }
}

// sum of the chunk partials in chunk order (deterministic), added to grad
__global__ void __launch_bounds__(256) kNetReducePartials(const float* __restrict__ partial, int chunks, int count,
                                                          float* grad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; c++) s += partial[(size_t)c * count + t];
    grad[t] += s;
}

// y_i = H_i . w_out (FP32)
__global__ void __launch_bounds__(256) kNetReadout(int n, const uint16_t* __restrict__ H, const float* __restrict__ wo,
                                                   float* y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t* row = H + (size_t)i * kNetHidden;
    float s = 0.0f;
    for (int f = 0; f < kNetHidden; f++) s += fromBF16(row[f]) * wo[f];
    y[i] = s;
}

// x_i = sigma y_i / AC_i into the double output (guess of the solution)
__global__ void __launch_bounds__(256) kNetOutput(int n, const float* __restrict__ y, const double* __restrict__ AC,
                                                  const double* sigma, double* x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = AC[i] != 0.0 ? sigma[0] * (double)y[i] / AC[i] : 0.0;
}

// Energy of the guess direction x on its graph: (A x)_i = AC_i x_i + sum AF x_nb
// into Ax, and per-block partial sums of x.b and x.Ax (two per block)
__global__ void __launch_bounds__(256) kNetEnergyPartial(NetGraph g, const double* __restrict__ AC,
                                                         const double* __restrict__ AF, const double* __restrict__ b,
                                                         const double* __restrict__ x, double* Ax, double* partial) {
    __shared__ double red[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double xb = 0.0, xax = 0.0;
    if (i < g.n) {
        double a = AC[i] * x[i];
        for (int k = g.start[i]; k < g.start[i + 1]; k++) {
            const int j = g.nb[k];
            if (j >= 0) a += AF[k] * x[j];
        }
        Ax[i] = a;
        xb = x[i] * b[i];
        xax = x[i] * a;
    }
    const double s1 = blockSum(xb, red);
    const double s2 = blockSum(xax, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s1;
        partial[2 * blockIdx.x + 1] = s2;
    }
}
// scal = [x.b, x.Ax, alpha = x.b / x.Ax (0 when x.Ax <= 0)], partials in block order
__global__ void __launch_bounds__(64) kNetEnergyFinish(int blocks, const double* partial, double* scal) {
    if (threadIdx.x != 0) return;
    double xb = 0.0, xax = 0.0;
    for (int i = 0; i < blocks; i++) {
        xb += partial[2 * i];
        xax += partial[2 * i + 1];
    }
    scal[0] = xb;
    scal[1] = xax;
    scal[2] = xax > 0.0 ? xb / xax : 0.0;
}
// x *= alpha (the energy-optimal multiple of the guess direction)
__global__ void __launch_bounds__(256) kNetScaleVec(int n, double* x, const double* scal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= scal[2];
}
// Loss = 1 - (x.b)^2 / (x.Ax * E*): the share of the solution's energy
// E* = b.x* = |x*|_A^2 the best multiple of x leaves uncaptured.  Its gradient
// in x is -2 (x.b)/(x.Ax E*) (b - alpha A x); through x = sigma y / AC in y.
__global__ void __launch_bounds__(256) kNetEnergyGrad(int n, const double* __restrict__ b,
                                                      const double* __restrict__ Ax, const double* __restrict__ AC,
                                                      const double* sigma, const double* scal, double estar, float* dy,
                                                      double* loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double xb = scal[0], xax = scal[1], alpha = scal[2];
    const bool ok = xax > 0.0 && estar > 0.0;
    if (i == 0) loss[0] = ok ? 1.0 - xb * xb / (xax * estar) : 1.0;
    if (i >= n) return;
    const double c = ok ? -2.0 * xb / (xax * estar) : 0.0;
    const double gx = c * (b[i] - alpha * Ax[i]);
    dy[i] = AC[i] != 0.0 ? (float)(gx * sigma[0] / AC[i]) : 0.0f;
}

// readout backward: dH_i = dy_i w_out (FP32 rows, padded rows zero); per-block
// partial dw_out (64 per block)
__global__ void __launch_bounds__(256) kNetReadoutBack(int n, int nPad, const float* __restrict__ dy,
                                                       const uint16_t* __restrict__ H, const float* __restrict__ wo,
                                                       float* dH, float* partial) {
    __shared__ float acc[kNetHidden];
    for (int f = threadIdx.x; f < kNetHidden; f += blockDim.x) acc[f] = 0.0f;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nPad) {
        const float d = i < n ? dy[i] : 0.0f;
        for (int f = 0; f < kNetHidden; f++) dH[(size_t)i * kNetHidden + f] = d * wo[f];
    }
    __syncthreads();
    // column sums of dy_i H_i over the block's cells, in cell order
    for (int f = threadIdx.x; f < kNetHidden; f += blockDim.x) {
        float s = 0.0f;
        const int c0 = blockIdx.x * blockDim.x;
        for (int c = c0; c < min(n, c0 + (int)blockDim.x); c++) s += dy[c] * fromBF16(H[(size_t)c * kNetHidden + f]);
        partial[(size_t)blockIdx.x * kNetHidden + f] = s;
    }
}

// dZ = dH where the layer's output is positive (ReLU'), as BF16
__global__ void __launch_bounds__(256) kNetReluBack(long count, const float* __restrict__ dH,
                                                    const uint16_t* __restrict__ H, uint16_t* dZ) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    dZ[t] = fromBF16(H[t]) > 0.0f ? toBF16(dH[t]) : (uint16_t)0;
}

// dH_{l-1} from dA = dZ W^T (FP32, rows of 2D): the left half directly, the
// right half (dM) through the transposed aggregation sum_k wT_jk dM_nb(k)
__global__ void __launch_bounds__(256) kNetAggregateBack(NetGraph g, const float* __restrict__ dA, int D, float* dH) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.nPad * D) return;
    const int j = (int)(t / D), f = (int)(t % D);
    float s = dA[(size_t)j * 2 * D + f];
    if (j < g.n)
        for (int k = g.start[j]; k < g.start[j + 1]; k++) {
            const int i = g.nb[k];
            if (i >= 0) s += g.wT[k] * dA[(size_t)i * 2 * D + D + f];
        }
    dH[t] = s;
}

// Adam on the FP32 master parameters
__global__ void __launch_bounds__(256) kNetAdam(int count, float* p, const float* __restrict__ g, float* m, float* v,
                                                float lr, float b1, float b2, float eps, float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// BF16 copies of a weight matrix W [K x N] and of its transpose [N x K]
__global__ void __launch_bounds__(256) kNetCastWeights(const float* __restrict__ W, int K, int N, uint16_t* Wb,
                                                       uint16_t* WTb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K * N) return;
    const int k = t / N, n = t % N;
    const uint16_t h = toBF16(W[t]);
    Wb[t] = h;
    WTb[(size_t)n * K + k] = h;
}

__global__ void __launch_bounds__(256) kNetZero(long count, float* p) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = 0.0f;
}

// prolongation of a level's vector to the level above: out_n = in_toCoarse(n)
__global__ void __launch_bounds__(256) kNetProlong(int n, const int* __restrict__ toC, const double* __restrict__ in,
                                                   double* out, const Ctl* ctl) {
    if (ctl != nullptr && ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[toC[i]];
}
// restriction: out_c = sum over the members of c (ascending) of in
__global__ void __launch_bounds__(256) kNetRestrict(int nc, const int* __restrict__ mStart,
                                                    const int* __restrict__ members, const double* __restrict__ in,
                                                    double* out, const Ctl* ctl) {
    if (ctl != nullptr && ctl->done) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double s = 0.0;
    for (int j = mStart[c]; j < mStart[c + 1]; j++) s += in[members[j]];
    out[c] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
inline int gridFor(long n, int t = 256) { return (int)((n + t - 1) / t); }
}

void netLaunchWeights(hipStream_t s, const NetGraph& g, const double* AC, const double* AF) {
    kNetWeights<<<gridFor(g.n), 256, 0, s>>>(g, AC, AF);
    if (g.refs > 0) kNetWeightsT<<<gridFor(g.refs), 256, 0, s>>>(g);
}
void netLaunchScale(hipStream_t s, int n, const double* AC, const double* b, double* partial, double* sigma) {
    const int blocks = gridFor(n);
    kNetScalePartial<<<blocks, 256, 0, s>>>(n, AC, b, partial);
    kNetScaleFinish<<<1, 64, 0, s>>>(n, blocks, partial, sigma);
}
void netLaunchFeatures(hipStream_t s, const NetGraph& g, const double* AC, const double* b, const double* sigma,
                       float* r, uint16_t* H0) {
    kNetScaledRhs<<<gridFor(g.n), 256, 0, s>>>(g.n, AC, b, sigma, r);
    kNetFeatures<<<gridFor(g.nPad), 256, 0, s>>>(g, r, H0);
}
void netLaunchAggregate(hipStream_t s, const NetGraph& g, const uint16_t* H, int D, uint16_t* A) {
    kNetAggregate<<<gridFor((long)g.nPad * D), 256, 0, s>>>(g, H, D, A);
}
void netLaunchGemm(hipStream_t s, const uint16_t* A, const uint16_t* B, int M, int N, int K, float* Cf,
                   uint16_t* Cb) {
    dim3 grid(M / 16, N / 64);
    kNetGemm<<<grid, 256, 0, s>>>(A, B, M, N, K, Cb ? 1 : 0, Cf, Cb);
}
void netLaunchGemmAtB(hipStream_t s, const uint16_t* A, const uint16_t* dZ, int rows, int K, int N, float* partial,
                      float* grad) {
    const int chunks = (rows + kNetChunk - 1) / kNetChunk;
    kNetGemmAtB<<<chunks, 256, 0, s>>>(A, dZ, rows, K, N, partial);
    kNetReducePartials<<<gridFor((long)K * N), 256, 0, s>>>(partial, chunks, K * N, grad);
}
void netLaunchReadout(hipStream_t s, int n, const uint16_t* H, const float* wo, float* y) {
    kNetReadout<<<gridFor(n), 256, 0, s>>>(n, H, wo, y);
}
void netLaunchOutput(hipStream_t s, int n, const float* y, const double* AC, const double* sigma, double* x) {
    kNetOutput<<<gridFor(n), 256, 0, s>>>(n, y, AC, sigma, x);
}
void netLaunchEnergy(hipStream_t s, const NetGraph& g, const double* AC, const double* AF, const double* b,
                     const double* x, double* Ax, double* partial, double* scal) {
    const int blocks = gridFor(g.n);
    kNetEnergyPartial<<<blocks, 256, 0, s>>>(g, AC, AF, b, x, Ax, partial);
    kNetEnergyFinish<<<1, 64, 0, s>>>(blocks, partial, scal);
}
void netLaunchScaleVec(hipStream_t s, int n, double* x, const double* scal) {
    kNetScaleVec<<<gridFor(n), 256, 0, s>>>(n, x, scal);
}
void netLaunchEnergyGrad(hipStream_t s, int n, const double* b, const double* Ax, const double* AC,
                         const double* sigma, const double* scal, double estar, float* dy, double* loss) {
    kNetEnergyGrad<<<gridFor(n), 256, 0, s>>>(n, b, Ax, AC, sigma, scal, estar, dy, loss);
}
void netLaunchReadoutBack(hipStream_t s, int n, int nPad, const float* dy, const uint16_t* H, const float* wo,
                          float* dH, float* partial, float* gradWo) {
    const int blocks = gridFor(nPad);
    kNetReadoutBack<<<blocks, 256, 0, s>>>(n, nPad, dy, H, wo, dH, partial);
    kNetReducePartials<<<1, 256, 0, s>>>(partial, blocks, kNetHidden, gradWo);
}
void netLaunchReluBack(hipStream_t s, long count, const float* dH, const uint16_t* H, uint16_t* dZ) {
    kNetReluBack<<<gridFor(count), 256, 0, s>>>(count, dH, H, dZ);
}
void netLaunchAggregateBack(hipStream_t s, const NetGraph& g, const float* dA, int D, float* dH) {
    kNetAggregateBack<<<gridFor((long)g.nPad * D), 256, 0, s>>>(g, dA, D, dH);
}
void netLaunchAdam(hipStream_t s, int count, float* p, const float* g, float* m, float* v, float lr, float b1,
                   float b2, float eps, int step) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    kNetAdam<<<gridFor(count), 256, 0, s>>>(count, p, g, m, v, lr, b1, b2, eps, c1, c2);
}
void netLaunchCastWeights(hipStream_t s, const float* W, int K, int N, uint16_t* Wb, uint16_t* WTb) {
    kNetCastWeights<<<gridFor((long)K * N), 256, 0, s>>>(W, K, N, Wb, WTb);
}
void netLaunchZero(hipStream_t s, long count, float* p) { kNetZero<<<gridFor(count), 256, 0, s>>>(count, p); }
void netLaunchProlong(hipStream_t s, int n, const int* toC, const double* in, double* out, const Ctl* ctl) {
    kNetProlong<<<gridFor(n), 256, 0, s>>>(n, toC, in, out, ctl);
}
void netLaunchRestrict(hipStream_t s, int nc, const int* mStart, const int* members, const double* in, double* out,
                       const Ctl* ctl) {
    kNetRestrict<<<gridFor(nc), 256, 0, s>>>(nc, mStart, members, in, out, ctl);
}

}  // namespace axs
