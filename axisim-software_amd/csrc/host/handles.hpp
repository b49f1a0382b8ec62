// handles.hpp -- the C ABI's opaque handle types that more than one source
// file of libaxs.so opens.
#pragma once

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

#include "axs.h"
#include "solver.hpp"

namespace axs {

// The calling thread's error text for entry points without a solver handle
// (axs_last_error(NULL)).
std::string& threadError();

// Run f, mapping exceptions to axs_status codes; the message goes to *err (a
// handle's text) or, when err is null, to the thread's text.
template <class F>
axs_status guarded(std::string* err, F&& f) {
    std::string& out = err ? *err : threadError();
    try {
        f();
        if (err) err->clear();
        return AXS_OK;
    } catch (const HipError& e) {
        out = e.what();
        return AXS_ERR_DEVICE;
    } catch (const std::invalid_argument& e) {
        out = e.what();
        return AXS_ERR_INVALID;
    } catch (const std::out_of_range& e) {
        out = e.what();
        return AXS_ERR_RANGE;
    } catch (const std::logic_error& e) {
        // domain_error (unsupported option) derives from logic_error
        out = e.what();
        return dynamic_cast<const std::domain_error*>(&e) ? AXS_ERR_UNSUPPORTED : AXS_ERR_STATE;
    } catch (const std::bad_alloc&) {
        out = "out of host memory";
        return AXS_ERR_DEVICE;
    } catch (const std::exception& e) {
        out = e.what();
        return AXS_ERR_DEVICE;
    }
}

}  // namespace axs

// One solver instance (axs_create).
struct axs_solver {
    axs::Solver* s = nullptr;
};

// A stand-alone multigrid hierarchy on one device (axs_mg_*): the levels and
// their tables, the level-0 system and iterates the host-array entry points
// copy into, and the stream every call runs on.
struct axs_mg {
    axs::MGDevice mg;
    int device = 0;
    hipStream_t stream = nullptr;
    axs::Arena arena;
    double *AC = nullptr, *b = nullptr, *AF = nullptr, *x = nullptr, *xNew = nullptr;
    int n = 0, refs = 0;
    int refine = AXS_REFINE_CYCLE;  // axs_mg_set_mixed
};
