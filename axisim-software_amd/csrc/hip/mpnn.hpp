// mpnn.hpp -- message-passing network on a cell graph: device data and the
// kernel launchers of mpnn.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace axs {

constexpr int kNetIn = 16;      // input features per cell (4 used, padded for the matrix cores)
constexpr int kNetHidden = 64;  // hidden width
constexpr int kNetChunk = 256;  // cells per workgroup in the weight-gradient reduction

// A cell graph on the device: CSR rows (nb = -1 on boundary slots), the
// Jacobi-scaled weights w and, per slot k of cell j, wT[k] = the weight of j in
// the row of nb(k) (rev[k] = that slot, -1 on boundary slots).  Rows padded
// to nPad (a multiple of 16).
struct NetGraph {
    int n = 0, nPad = 0, refs = 0;
    const int* start = nullptr;
    const int* nb = nullptr;
    const int* rev = nullptr;
    float* w = nullptr;
    float* wT = nullptr;
};

void netLaunchWeights(hipStream_t s, const NetGraph& g, const double* AC, const double* AF);
void netLaunchScale(hipStream_t s, int n, const double* AC, const double* b, double* partial, double* sigma);
void netLaunchFeatures(hipStream_t s, const NetGraph& g, const double* AC, const double* b, const double* sigma,
                       float* r, uint16_t* H0);
void netLaunchAggregate(hipStream_t s, const NetGraph& g, const uint16_t* H, int D, uint16_t* A);
// C = A B (BF16 in, FP32 accumulate): Cf FP32, or relu(C) as BF16 when Cb is set
void netLaunchGemm(hipStream_t s, const uint16_t* A, const uint16_t* B, int M, int N, int K, float* Cf,
                   uint16_t* Cb);
// grad += A^T dZ over `rows` cells (A: rows x K, dZ: rows x N)
void netLaunchGemmAtB(hipStream_t s, const uint16_t* A, const uint16_t* dZ, int rows, int K, int N, float* partial,
                      float* grad);
void netLaunchReadout(hipStream_t s, int n, const uint16_t* H, const float* wo, float* y);
void netLaunchOutput(hipStream_t s, int n, const float* y, const double* AC, const double* sigma, double* x);
// x.b, x.Ax and alpha = x.b / x.Ax into scal[0..2]; A x into Ax
void netLaunchEnergy(hipStream_t s, const NetGraph& g, const double* AC, const double* AF, const double* b,
                     const double* x, double* Ax, double* partial, double* scal);
// x *= scal[2]
void netLaunchScaleVec(hipStream_t s, int n, double* x, const double* scal);
// loss = 1 - (x.b)^2 / (x.Ax E*) and its gradient in the readout output y
void netLaunchEnergyGrad(hipStream_t s, int n, const double* b, const double* Ax, const double* AC,
                         const double* sigma, const double* scal, double estar, float* dy, double* loss);
void netLaunchReadoutBack(hipStream_t s, int n, int nPad, const float* dy, const uint16_t* H, const float* wo,
                          float* dH, float* partial, float* gradWo);
void netLaunchReluBack(hipStream_t s, long count, const float* dH, const uint16_t* H, uint16_t* dZ);
void netLaunchAggregateBack(hipStream_t s, const NetGraph& g, const float* dA, int D, float* dH);
void netLaunchAdam(hipStream_t s, int count, float* p, const float* g, float* m, float* v, float lr, float b1,
                   float b2, float eps, int step);
void netLaunchCastWeights(hipStream_t s, const float* W, int K, int N, uint16_t* Wb, uint16_t* WTb);
void netLaunchZero(hipStream_t s, long count, float* p);
// transfers between multigrid levels for the guess (ctl: skip once the outer loop ended)
void netLaunchProlong(hipStream_t s, int n, const int* toC, const double* in, double* out, const Ctl* ctl);
void netLaunchRestrict(hipStream_t s, int nc, const int* mStart, const int* members, const double* in, double* out,
                       const Ctl* ctl);

}  // namespace axs
