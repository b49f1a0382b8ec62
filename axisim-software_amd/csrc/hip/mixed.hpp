// mixed.hpp -- FP32 defect-correction buffers and launchers (kernels_mixed.hip).
#pragma once

#include "device.hpp"

namespace axs {

constexpr int kDirectMax = 2048;      // largest coarsest level solved by the dense FP64 inverse
constexpr int kDirectOneBlock = 256;  // up to this size the inverse is formed by one workgroup
// the coarsest inverse's row stride (n, or n padded for the blocked inverse)
// and the work buffer (doubles) launchCoarseInvert needs beside the inverse
int coarseInverseLd(int n);
size_t coarseInverseWork(int n);

// FP32 copies of the momentum operator and the defect iterates.
// u and v interleaved as (u, v) pairs, so a neighbour's pair is one 8-byte load.
struct MixedState {
    float* AF32 = nullptr;       // shared u / v off-diagonals (S)
    float2* inv = nullptr;       // (1/AC_u, 1/AC_v)
    float2* r = nullptr;         // FP64 defects rounded to FP32
    float2* d = nullptr;         // delta iterates (ping)
    float2* d2 = nullptr;        // (pong)
};

void launchMomDefect(hipStream_t s, const DMesh& m, DState st, MixedState x);
void launchMomSweep32(hipStream_t s, const DMesh& m, const Ctl* ctl, MixedState x, const float2* dO, float2* dN);
// `sweeps` (2 .. kMaxFusedSweeps, at most sb.rings + 1) FP32 sweeps in one launch
// over the fused momentum blocks; bit-identical to as many launchMomSweep32
void launchMomSweep32Block(hipStream_t s, const DMesh& m, const DState& st, MixedState x, const DSweepBlocks& sb,
                           int sweeps, const float2* dO, float2* dN);
// the same, its last sweep ending in u += delta, v += delta
// (launchMomSweepApply) / its first sweep from the FP64 defect kMomentum wrote
// into x.r (DState::momR32; launchMomDefect's first sweep); bit-identical to
// those launches
void launchMomSweep32BlockApply(hipStream_t s, const DMesh& m, const DState& st, MixedState x,
                                const DSweepBlocks& sb, int sweeps, const float2* dO);
void launchMomSweep32BlockFromDefect(hipStream_t s, const DMesh& m, const DState& st, MixedState x,
                                     const DSweepBlocks& sb, int sweeps, float2* dN);
void launchMomSweepApply(hipStream_t s, const DMesh& m, DState st, MixedState x, const float2* dO, int sweep);
void launchMGDefect(hipStream_t s, const DLevel& L, double w, int firstSweep, const Ctl* ctl);
void launchMGJacobi32(hipStream_t s, const DLevel& L, float w, const Ctl* ctl);
void launchMGSmoothBlock32(hipStream_t s, const DLevel& L, float w, int sweeps, const Ctl* ctl);
void launchMGResidual32(hipStream_t s, const DLevel& f, const Ctl* ctl);
// apply: when set, also apply[n] += the new iterate (in FP64) -- the mixed
// cycle's x += e at level 0, fused into its last sweep
void launchMGProlongSmooth32(hipStream_t s, const DLevel& f, const DLevel& c, float w, const Ctl* ctl,
                             double* apply = nullptr);
void launchMGProlong32(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl);
void launchMGApply(hipStream_t s, const DLevel& L, const Ctl* ctl);
void launchMGCast(hipStream_t s, const DLevel& L, const Ctl* ctl);
void launchCoarseInvert(hipStream_t s, const DLevel& L, double* aug, double* inv, int* status, const Ctl* ctl);
void launchCoarseApply(hipStream_t s, const DLevel& L, const double* inv, const Ctl* ctl);
// the same on the level's FP64 b and x (the FP64 cycles)
void launchCoarseApply64(hipStream_t s, const DLevel& L, const double* inv, const Ctl* ctl);

}  // namespace axs
