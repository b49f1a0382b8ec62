// noalloc.hip -- axs_mg_solve_noalloc (include/axs.h), the entry point of
// libaxs.so that allocates no device memory.
//
// It runs the same launches as axs_mg_solve on caller-owned device arrays.  The
// V-cycle work arrays (every coarse level's operator, right-hand side, iterates
// and residual, level 0's second iterate and residual, the LDS tail's image)
// live in `scratch`, a static device array of AXS_MG_NOALLOC_SCRATCH_BYTES that
// every GPU the library runs on holds until the process ends.  In host code the
// name `scratch` is only a stand-in for the symbol: its device address on each
// GPU is looked up once, with that GPU current, by hipGetSymbolAddress, and only
// that address is handed to kernels and copies.  Every call on one GPU shares the
// array; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "axs.h"
#include "handles.hpp"

namespace {

__device__ __align__(256) unsigned char scratch[AXS_MG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::once_flag gOnce[kMaxDevices];
unsigned char* gScratch[kMaxDevices] = {};
hipError_t gStatus[kMaxDevices] = {};

// This GPU's address of `scratch`; `device` must be current.
unsigned char* scratchOn(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::out_of_range("device index beyond the scratch table");
    std::call_once(gOnce[device], [device] {
        void* p = nullptr;
        gStatus[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        gScratch[device] = gStatus[device] == hipSuccess ? (unsigned char*)p : nullptr;
    });
    if (gStatus[device] != hipSuccess || !gScratch[device])
        throw axs::HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(gStatus[device]));
    return gScratch[device];
}

}  // namespace

extern "C" axs_status axs_mg_solve_noalloc(axs_mg* h, const double* AC, const double* AF, const double* b, double* x,
                                           int cycles, double weight, int pre, int post, int coarse, void* stream) {
    if (!h || !AC || !AF || !b || !x) return AXS_ERR_INVALID;
    return axs::guarded(nullptr, [&] {
        if (pre < 0 || post < 0 || coarse < 0) throw std::invalid_argument("negative sweep count");
        axs::MGParams p;
        p.cycles = cycles;
        p.weight = weight;
        p.pre = pre;
        p.post = post;
        p.coarseSweeps = coarse;
        const size_t need = h->mg.workBytes();
        if (need > (size_t)AXS_MG_NOALLOC_SCRATCH_BYTES)
            throw std::out_of_range("the hierarchy's V-cycle work arrays need " + std::to_string(need) +
                                    " bytes; the no-allocation solve has " +
                                    std::to_string((size_t)AXS_MG_NOALLOC_SCRATCH_BYTES));
        AXS_HIP(hipSetDevice(h->device));
        unsigned char* base = scratchOn(h->device);
        hipStream_t s = stream ? (hipStream_t)stream : h->stream;
        // the previous call's contents are cleared (a fill kernel, no memset node)
        axs::launchFill(s, (double*)base, 0.0, (int)(need / sizeof(double)), nullptr);
        const axs::MGDevice::Bound saved = h->mg.bindWork(base, AC, b, AF, x, s);
        try {
            h->mg.enqueueSolve(s, p, nullptr, true);
        } catch (...) {
            h->mg.unbindWork(saved);
            throw;
        }
        h->mg.unbindWork(saved);
        AXS_HIP(hipGetLastError());
    });
}
