// mixed.hpp -- FP32 defect-correction buffers and launchers (kernels_mixed.hip).
#pragma once

#include "device.hpp"

namespace axs {

constexpr int kDirectMax = 256;  // largest coarsest level solved by the dense FP64 inverse

// FP32 copies of the momentum operator and the defect iterates.
struct MixedState {
    float* AF32 = nullptr;           // shared u / v off-diagonals (S)
    float *invU = nullptr, *invV = nullptr;
    float *ru = nullptr, *rv = nullptr;      // FP64 defects rounded to FP32
    float *du = nullptr, *dv = nullptr;      // delta iterates (ping)
    float *du2 = nullptr, *dv2 = nullptr;    // (pong)
};

void launchMomDefect(hipStream_t s, const DMesh& m, DState st, MixedState x);
void launchMomSweep32(hipStream_t s, const DMesh& m, const Ctl* ctl, MixedState x, const float* uo, const float* vo,
                      float* un, float* vn);
void launchMomSweepApply(hipStream_t s, const DMesh& m, DState st, MixedState x, const float* uo, const float* vo,
                         int sweep);
void launchMGDefect(hipStream_t s, const DLevel& L, double w, int firstSweep, const Ctl* ctl);
void launchMGJacobi32(hipStream_t s, const DLevel& L, float w, const Ctl* ctl);
void launchMGSmoothBlock32(hipStream_t s, const DLevel& L, float w, int sweeps, const Ctl* ctl);
void launchMGRestrict32(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl);
void launchMGResidual32(hipStream_t s, const DLevel& f, const Ctl* ctl);
void launchMGRestrictSum32(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl);
void launchMGProlongSmooth32(hipStream_t s, const DLevel& f, const DLevel& c, float w, const Ctl* ctl);
void launchMGProlong32(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl);
void launchMGApply(hipStream_t s, const DLevel& L, const Ctl* ctl);
void launchMGCast(hipStream_t s, const DLevel& L, const Ctl* ctl);
void launchMGCoarseOp64(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl);
void launchCoarseInvert(hipStream_t s, const DLevel& L, double* aug, double* inv, int* status, const Ctl* ctl);
void launchCoarseApply(hipStream_t s, const DLevel& L, const double* inv, const Ctl* ctl);

}  // namespace axs
