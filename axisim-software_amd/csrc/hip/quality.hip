// Mesh quality metrics (Quality::buildQuality, solver/source/quality.cpp:160-166):
// per cell, one thread, the three metrics the mesh inspector shows.
//
//   aspect ratio          max / min over the cell's faces of |(x_f - x_c) . n_f|
//                         (0 when the smallest is not above 1e-30), quality.cpp:129-158
//   non-orthogonality     min over the cell's faces of |(x_f - x_c) . n_f| / |x_f - x_c|
//                         and, for interior faces, |(x_N - x_P) . n_f| / |x_N - x_P|
//                         (owner P, neighbour N of the face), starting from 1,
//                         quality.cpp:61-105
//   skewness              from the polygon's interior angles (degrees): max of
//                         (theta_max - theta_e) / (180 - theta_e) and
//                         (theta_e - theta_min) / theta_e with theta_e = 180 - 360 / n;
//                         corners closer than 1e-30 skipped, 0 for fewer than 3
//                         corners, quality.cpp:7-59
//
// Every expression is the reference's, in its order, with the minimum taken as
// the reference's std::min (the first operand unless the second is smaller, so
// a NaN ratio leaves the running value).  Only acos differs from the host's
// last bit; the rest is correctly rounded on both sides.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace axs {

namespace {
__device__ __forceinline__ double minRef(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double maxRef(double a, double b) { return (a < b) ? b : a; }
}  // namespace

__global__ void __launch_bounds__(kBlock) kMeshQuality(QualityMesh m, double* __restrict__ aspect,
                                                       double* __restrict__ orth, double* __restrict__ skew) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const double cz = m.cz[c], cr = m.cr[c];
    double dMin = 1.7976931348623157e308, dMax = 0.0, o = 1.0;
    for (int k = m.faceStart[c]; k < m.faceStart[c + 1]; k++) {
        const int f = m.faceIds[k];
        const double fz = m.fz[f] - cz, fr = m.fr[f] - cr;
        const double nz = m.nz[f], nr = m.nr[f];
        const double d = fabs(fz * nz + fr * nr);
        dMin = minRef(dMin, d);
        dMax = maxRef(dMax, d);
        const double fMag = sqrt(fz * fz + fr * fr);
        o = minRef(o, d / fMag);
        const int nb = m.neighbor[f];
        if (nb < 0) continue;
        const int ow = m.owner[f];
        const double dz = m.cz[nb] - m.cz[ow], dr = m.cr[nb] - m.cr[ow];
        const double dDot = fabs(dz * nz + dr * nr);
        const double dMag = sqrt(dz * dz + dr * dr);
        o = minRef(o, dDot / dMag);
    }
    aspect[c] = (dMin > 1e-30) ? dMax / dMin : 0.0;
    orth[c] = o;

    const int b = m.cornerStart[c], n = m.cornerStart[c + 1] - b;
    double s = 0.0;
    if (n >= 3) {
        constexpr double radToDeg = 57.29577951308232;
        double tMin = 180.0, tMax = 0.0;
        for (int k = 0; k < n; k++) {
            const int ip = b + (k + n - 1) % n, ih = b + k, in = b + (k + 1) % n;
            const double az = m.pz[ip] - m.pz[ih], ar = m.pr[ip] - m.pr[ih];
            const double bz = m.pz[in] - m.pz[ih], br = m.pr[in] - m.pr[ih];
            const double aMag = sqrt(az * az + ar * ar);
            const double bMag = sqrt(bz * bz + br * br);
            if (aMag < 1e-30 || bMag < 1e-30) continue;
            double cosT = (az * bz + ar * br) / (aMag * bMag);
            cosT = cosT < -1.0 ? -1.0 : (1.0 < cosT ? 1.0 : cosT);  // std::clamp(v, -1, 1)
            const double theta = acos(cosT) * radToDeg;
            tMin = minRef(tMin, theta);
            tMax = maxRef(tMax, theta);
        }
        const double tE = 180.0 - 360.0 / (double)n;
        s = maxRef((tMax - tE) / (180 - tE), (tE - tMin) / tE);
    }
    skew[c] = s;
}

void launchMeshQuality(hipStream_t s, const QualityMesh& m, double* aspect, double* orth, double* skew) {
    if (m.n > 0) kMeshQuality<<<blocksFor(m.n), kBlock, 0, s>>>(m, aspect, orth, skew);
}

}  // namespace axs
