// capi.cpp -- extern "C" entry points of libaxs.so (declared in include/axs.h).
//
// Every entry point catches C++ exceptions and maps them to axs_status codes;
// the message is kept on the handle (or thread-locally for handle-less calls)
// for axs_last_error.  The reference aborts on a CUDA error (CUDA_CHECK,
// manager/header/gpu_utils.h:5-11) and prints validation failures to its
// console (Solver::runCheck, solver.cpp:530-661); here both come back as codes.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>

#include "axs.h"
#include "cdt.hpp"
#include "mesh_types.hpp"
#include "net.hpp"
#include "solver.hpp"
#include "topology.hpp"
#include "handles.hpp"

using namespace axs;



struct axs_mesh {
    HostMesh m;
};

struct axs_net {
    Mpnn* net = nullptr;
};

struct axs_cdt {
    axs::cdt::Result r;
};

namespace axs {
std::string& threadError() {
    thread_local std::string text;
    return text;
}
}  // namespace axs


extern "C" {

void axs_config_defaults(axs_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->convection_scheme = AXS_CONV_UPWIND;   // solver_struct.h:206
    c->gradient_scheme = AXS_GRAD_LSQ;        // :207
    c->linear_solver = AXS_LINEAR_JACOBI;     // :208
    c->time_scheme = AXS_TIME_BDF1;           // :209
    c->linear_max_iter = 20;                  // :212
    c->use_multigrid = 1;                     // :213
    c->add_convection = 1;                    // :215
    c->transient = 0;                         // :216
    c->use_non_orth = 0;                      // :226
    c->dt = 0.1;                              // :228
    c->t_end = 2.0;                           // :229
    c->max_iter = 50;                         // :159
    c->check_conv = 1;                        // :160
    c->mg_max_iter = 3;                       // :173
    c->mg_weight = 0.6;                       // :178
    c->mg_coarse_sweeps = 30;                 // :183
    c->mg_pre_sweeps = 1;                     // :187
    c->mg_post_sweeps = 1;                    // :188
    c->alpha_u = 0.7;                         // :312
    c->alpha_p = 0.3;                         // :313
    c->rho = 998.0;                           // src/header/setting.cuh:63
    c->mu = 0.0010518;                        // src/header/setting.cuh:64
    for (int i = 0; i < AXS_RES_COUNT; i++) { // initConfigResiduals, solver.cpp:87-101
        c->residual[i].enabled = i < AXS_RES_TEMPERATURE;  // the scalars' only when switched on
        c->residual[i].norm = AXS_NORM_L1;
        c->residual[i].scaling = AXS_SCALE_DIAGONAL;
        c->residual[i].tol = 0.001;
    }
    c->structured_mesh = 0;
    c->mg_max_levels = 0;
    c->mg_smoother = AXS_MG_SMOOTH_JACOBI;
    c->precision = AXS_PREC_FP64;
    c->coarse_direct = 0;
    c->save_key_frame_iter = 2;               // solver_struct.h:230
    c->max_time_frames = 500;                 // solver/header/solver.h:108
    c->solve_energy = 0;                      // solver_struct.h:10
    c->solve_concentration = 0;               // solver_struct.h:11
    c->cp = 4180.0;                           // src/header/setting.cuh:66
    c->k = 0.6;                               // :67
    c->diffusivity = 3.0277e-9;               // :69
    c->mixed_refine = AXS_REFINE_SOLVE;
}

axs_status axs_create(axs_solver** out, int device) {
    if (!out) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        int count = 0;
        AXS_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw std::invalid_argument("no such device");
        axs_solver* h = new axs_solver();
        try {
            h->s = new Solver(device);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void axs_destroy(axs_solver* h) {
    if (!h) return;
    delete h->s;
    delete h;
}

const char* axs_last_error(const axs_solver* h) {
    if (!h || !h->s) return axs::threadError().c_str();
    return h->s->lastError.c_str();
}

axs_status axs_set_mesh(axs_solver* h, const axs_mesh_view* v) {
    if (!h || !v) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->setMesh(*v); });
}

axs_status axs_set_bcs(axs_solver* h, int var, const axs_bc_view* v) {
    if (!h || !v) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->setBC(var, *v); });
}

axs_status axs_configure(axs_solver* h, const axs_config* c) {
    if (!h || !c) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->configure(*c); });
}

axs_status axs_run(axs_solver* h, axs_progress_cb cb, void* user, int continue_run) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->run(cb, user, continue_run != 0); });
}

axs_status axs_run_timed(axs_solver* h, int n, double* ms) {
    if (!h || !ms) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->runTimed(n, ms); });
}

axs_status axs_bench_kernel(axs_solver* h, int kernel, int repeats, double* ms, double* bytes) {
    if (!h || !ms || !bytes) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->benchKernel(kernel, repeats, ms, bytes); });
}

axs_status axs_bench_kernel_cold(axs_solver* h, int kernel, int repeats, double* ms, double* bytes) {
    if (!h || !ms || !bytes) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->benchKernel(kernel, repeats, ms, bytes, true); });
}

axs_status axs_autotune(axs_solver* h, int repeats, double* ms_default, double* ms_tuned) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->autotune(repeats, ms_default, ms_tuned); });
}

axs_status axs_set_sweep_fusion(axs_solver* h, int sweeps, int block_cells) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->setSweepFusion(sweeps, block_cells); });
}

void axs_request_stop(axs_solver* h) {
    if (h && h->s) h->s->requestStop();
}

int axs_can_continue(const axs_solver* h, char* reason, size_t len) {
    if (!h) return 0;
    std::string why;
    const bool ok = h->s->canContinue(why);
    if (reason && len > 0) {
        std::strncpy(reason, why.c_str(), len - 1);
        reason[len - 1] = 0;
    }
    return ok ? 1 : 0;
}

axs_status axs_get_field(axs_solver* h, int field, double* out, size_t n) {
    if (!h || !out) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->getField(field, out, n); });
}

size_t axs_field_length(const axs_solver* h, int field) { return h ? h->s->fieldLength(field) : 0; }

axs_status axs_get_residual_history(axs_solver* h, int res, double* out, int max_n, int* count) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] {
        const int c = h->s->residualHistory(res, out, max_n);
        if (count) *count = c;
    });
}

axs_status axs_get_stats(axs_solver* h, axs_stats* out) {
    if (!h || !out) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->stats(*out); });
}

int axs_frame_count(const axs_solver* h) { return h ? h->s->frameCount() : 0; }

axs_status axs_get_frame(axs_solver* h, int index, int field, double* time, double* out, size_t n) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] {
        const Solver::Frame& f = h->s->frame(index);
        if (time) *time = f.time;
        if (!out) return;
        const std::vector<double>* src = field == AXS_FIELD_U ? &f.u : field == AXS_FIELD_V ? &f.v
                                       : field == AXS_FIELD_P ? &f.p : field == AXS_FIELD_T ? &f.t
                                       : field == AXS_FIELD_C ? &f.c : nullptr;
        if (!src) throw std::invalid_argument("frames hold u, v, p, T and C");
        if (src->empty()) throw std::logic_error("that field was not solved");
        if (n != src->size()) throw std::out_of_range("output length does not match the frame");
        std::copy(src->begin(), src->end(), out);
    });
}

axs_status axs_set_cases(axs_solver* h, int count, const int* cell_start, const int* face_start) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->setCases(count, cell_start, face_start); });
}

axs_status axs_get_case_stats(axs_solver* h, int c, axs_case_stats* out) {
    if (!h || !out) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] {
        std::memset(out, 0, sizeof(*out));
        h->s->caseStats(c, out->iterations, out->converged, out->last_residual, out->stop_iteration);
    });
}

axs_status axs_get_wall_diagnostic(axs_solver* h, double out[10]) {
    if (!h || !out) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->wallDiagnostic(out); });
}

// ---------------- multigrid, stand-alone ----------------
axs_status axs_mg_create(axs_mg** out, int n_cells, const int* face_start, const int* face_neighbor, int max_levels) {
    if (!out || n_cells <= 0 || !face_start || !face_neighbor) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        if (face_start[0] != 0) throw std::invalid_argument("face_start[0] must be 0");
        const int S = face_start[n_cells];
        for (int k = 0; k < S; k++)
            if (face_neighbor[k] < -1 || face_neighbor[k] >= n_cells) throw std::invalid_argument("neighbour out of range");
        std::vector<int> st(face_start, face_start + n_cells + 1), nb(face_neighbor, face_neighbor + S);
        axs_mg* h = new axs_mg();
        try {
            AXS_HIP(hipGetDevice(&h->device));
            AXS_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
            h->mg.build(n_cells, st, nb, max_levels);
            h->n = n_cells;
            h->refs = S;
            h->AC = h->arena.zeros<double>(n_cells);
            h->b = h->arena.zeros<double>(n_cells);
            h->AF = h->arena.zeros<double>(S);
            h->x = h->arena.zeros<double>(n_cells);
            h->xNew = h->arena.zeros<double>(n_cells);
            h->mg.bindLevel0(h->AC, h->b, h->AF, h->x, h->xNew);
        } catch (...) {
            if (h->stream) (void)hipStreamDestroy(h->stream);
            delete h;
            throw;
        }
        *out = h;
    });
}

void axs_mg_destroy(axs_mg* h) {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    h->arena.release();
    h->mg.arena.release();
    (void)hipStreamDestroy(h->stream);
    delete h;
}

int axs_mg_levels(const axs_mg* h) { return h ? (int)h->mg.host.size() : 0; }

axs_status axs_mg_level_info(const axs_mg* h, int level, int* n_cells, int* n_refs) {
    if (!h) return AXS_ERR_INVALID;
    if (level < 0 || level >= (int)h->mg.host.size()) return AXS_ERR_RANGE;
    if (n_cells) *n_cells = h->mg.host[level].n;
    if (n_refs) *n_refs = h->mg.host[level].refs();
    return AXS_OK;
}

axs_status axs_mg_tail_info(const axs_mg* h, int precision, int* first_level, int* lds_bytes) {
    if (!h) return AXS_ERR_INVALID;
    if (precision != AXS_PREC_FP64 && precision != AXS_PREC_MIXED) return AXS_ERR_INVALID;
    const bool fp32 = precision == AXS_PREC_MIXED;
    if (first_level) *first_level = h->mg.tailLevel(fp32);
    if (lds_bytes) *lds_bytes = (int)h->mg.tailBytes(fp32);
    return AXS_OK;
}

axs_status axs_mg_get_maps(const axs_mg* h, int level, int* fs, int* fn, int* c2c, int* s2s) {
    if (!h) return AXS_ERR_INVALID;
    if (level < 0 || level >= (int)h->mg.host.size()) return AXS_ERR_RANGE;
    const Level& L = h->mg.host[level];
    if (fs) std::copy(L.start.begin(), L.start.end(), fs);
    if (fn) std::copy(L.nb.begin(), L.nb.end(), fn);
    if (c2c) std::copy(L.toCoarse.begin(), L.toCoarse.end(), c2c);
    if (s2s) std::copy(L.slotToCoarseSlot.begin(), L.slotToCoarseSlot.end(), s2s);
    return AXS_OK;
}

static void mgUpload(axs_mg* h, const double* AC, const double* AF, const double* b, const double* x) {
    AXS_HIP(hipMemcpy(h->AC, AC, h->n * sizeof(double), hipMemcpyHostToDevice));
    AXS_HIP(hipMemcpy(h->AF, AF, std::max(h->refs, 0) * sizeof(double), hipMemcpyHostToDevice));
    AXS_HIP(hipMemcpy(h->b, b, h->n * sizeof(double), hipMemcpyHostToDevice));
    if (x) AXS_HIP(hipMemcpy(h->x, x, h->n * sizeof(double), hipMemcpyHostToDevice));
    else AXS_HIP(hipMemset(h->x, 0, h->n * sizeof(double)));
}

static MGParams mgParams(int cycles, double weight, int pre, int post, int coarse, int precision, int direct,
                         int smoother) {
    if (precision != AXS_PREC_FP64 && precision != AXS_PREC_MIXED) throw std::invalid_argument("bad precision");
    if (smoother != AXS_MG_SMOOTH_JACOBI && smoother != AXS_MG_SMOOTH_RBGS)
        throw std::invalid_argument("bad multigrid smoother");
    if (direct < 0) throw std::invalid_argument("coarse_direct must be 0 or 1");
    if (pre < 0 || post < 0 || coarse < 0) throw std::invalid_argument("negative sweep count");
    MGParams p;
    p.cycles = cycles;
    p.weight = weight;
    p.pre = pre;
    p.post = post;
    p.coarseSweeps = coarse;
    p.precision = precision;
    p.coarseDirect = direct;
    p.smoother = smoother;
    return p;
}

axs_status axs_mg_set_mixed(axs_mg* h, int refine, int fuse_level0) {
    if (!h) return AXS_ERR_INVALID;
    if (refine != AXS_REFINE_CYCLE && refine != AXS_REFINE_SOLVE) return AXS_ERR_INVALID;
    h->refine = refine;
    h->mg.setFuseLevel0(fuse_level0 != 0);
    return AXS_OK;
}

axs_status axs_mg_solve(axs_mg* h, const double* AC, const double* AF, const double* b, double* x, int cycles,
                        double weight, int pre, int post, int coarse, int precision, int direct, int smoother) {
    if (!h || !AC || !AF || !b || !x) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] {
        MGParams p = mgParams(cycles, weight, pre, post, coarse, precision, direct, smoother);
        p.refine = h->refine;
        AXS_HIP(hipSetDevice(h->device));
        mgUpload(h, AC, AF, b, x);
        h->mg.enqueueSolve(h->stream, p, nullptr, true);
        AXS_HIP(hipStreamSynchronize(h->stream));
        AXS_HIP(hipGetLastError());
        AXS_HIP(hipMemcpy(x, h->x, h->n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

// One solve (operators + cycles) captured in a graph, as the solver's
// iteration graph runs it, and `repeats` replays timed with HIP events on the
// hierarchy's stream.  cold: a kernel streams 512 MiB (more than the L2 of
// each XCD and the 256 MB last-level cache hold) before each replay, and the
// solve's time is what `repeats` (flush, replay) pairs take beyond as many
// flushes alone.
static void mgBench(axs_mg* h, const double* AC, const double* AF, const double* b, const MGParams& p, int repeats,
                    bool cold, double* ms, double* bytes, double* x_out) {
    AXS_HIP(hipSetDevice(h->device));
    mgUpload(h, AC, AF, b, nullptr);
    h->mg.enqueueSolve(h->stream, p, nullptr, true);  // warm-up
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    AXS_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    try {
        h->mg.enqueueSolve(h->stream, p, nullptr, true);
    } catch (...) {
        (void)hipStreamEndCapture(h->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw;
    }
    AXS_HIP(hipStreamEndCapture(h->stream, &g));
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    AXS_HIP(ie);
    const size_t flushN = cold ? ((size_t)64 << 20) : 0;  // doubles
    double *flush = nullptr, *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float f = 0.0f;
    hipError_t err = hipGraphLaunch(ge, h->stream);  // warm-up of the graph
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (cold && err == hipSuccess) err = hipMalloc((void**)&flush, flushN * sizeof(double));
    if (cold && err == hipSuccess) err = hipMalloc((void**)&sink, 2048 * sizeof(double));
    auto timed = [&](bool solve, bool flushFirst) {
        float t = 0.0f;
        if (flushFirst) launchStreamRead(h->stream, flush, flushN, sink, 2048, -1.0);
        if (err == hipSuccess) err = hipEventRecord(e0, h->stream);
        for (int r = 0; r < repeats && err == hipSuccess; r++) {
            if (flushFirst) launchStreamRead(h->stream, flush, flushN, sink, 2048, -1.0);
            if (solve) err = hipGraphLaunch(ge, h->stream);
        }
        if (err == hipSuccess) err = hipEventRecord(e1, h->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
        return t;
    };
    if (!cold) {
        f = timed(true, false);
    } else if (err == hipSuccess) {
        launchFill(h->stream, flush, 0.0, (int)flushN, nullptr);
        const float flushOnly = timed(false, true);
        f = std::max(0.0f, timed(true, true) - flushOnly);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (flush) (void)hipFree(flush);
    if (sink) (void)hipFree(sink);
    (void)hipGraphExecDestroy(ge);
    AXS_HIP(err);
    AXS_HIP(hipGetLastError());
    *ms = f / repeats;
    *bytes = h->mg.bytesPerSolve(p, true);
    if (x_out) AXS_HIP(hipMemcpy(x_out, h->x, h->n * sizeof(double), hipMemcpyDeviceToHost));
}

axs_status axs_mg_bench(axs_mg* h, const double* AC, const double* AF, const double* b, int cycles, double weight,
                        int pre, int post, int coarse, int precision, int direct, int smoother, int repeats,
                        double* ms, double* bytes, double* x_out) {
    if (!h || !AC || !AF || !b || !ms || !bytes || repeats <= 0) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] {
        MGParams p = mgParams(cycles, weight, pre, post, coarse, precision, direct, smoother);
        p.refine = h->refine;
        mgBench(h, AC, AF, b, p, repeats, false, ms, bytes, x_out);
    });
}

axs_status axs_mg_bench_cold(axs_mg* h, const double* AC, const double* AF, const double* b, int cycles,
                             double weight, int pre, int post, int coarse, int precision, int direct, int smoother,
                             int repeats, double* ms, double* bytes) {
    if (!h || !AC || !AF || !b || !ms || !bytes || repeats <= 0) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] {
        MGParams p = mgParams(cycles, weight, pre, post, coarse, precision, direct, smoother);
        p.refine = h->refine;
        mgBench(h, AC, AF, b, p, repeats, true, ms, bytes, nullptr);
    });
}

axs_status axs_build_coloring(int n, const int* fs, const int* fn, int* color_start, int* cell_order, int* n_colors) {
    if (n < 0 || !fs || !fn || !color_start || !cell_order || !n_colors) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] {
        std::vector<int> cs, order;
        *n_colors = greedyColoring(n, fs, fn, cs, order);
        std::copy(cs.begin(), cs.end(), color_start);
        std::copy(order.begin(), order.end(), cell_order);
    });
}

// ---------------- case generator ----------------
axs_status axs_mesh_multiblock(axs_mesh** out, int n_blocks, const axs_block_spec* blocks) {
    if (!out) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        axs_mesh* h = new axs_mesh();
        const std::string err = buildMultiblock(h->m, n_blocks, blocks);
        if (!err.empty()) {
            delete h;
            throw std::invalid_argument(err);
        }
        *out = h;
    });
}

axs_status axs_mesh_triangles(axs_mesh** out, int n_points, const double* pz, const double* pr, int n_tris,
                              const int* tris, int n_seg, const double* sz0, const double* sr0, const double* sz1,
                              const double* sr1, const int* sg, double tol) {
    if (!out || n_points <= 0 || n_tris <= 0 || !pz || !pr || !tris) return AXS_ERR_INVALID;
    if (n_seg > 0 && (!sz0 || !sr0 || !sz1 || !sr1 || !sg)) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        axs_mesh* h = new axs_mesh();
        std::vector<double> z(pz, pz + n_points), r(pr, pr + n_points);
        std::vector<int> t(tris, tris + 3 * (size_t)n_tris);
        const std::string err = buildTriangles(h->m, z, r, t, n_seg, sz0, sr0, sz1, sr1, sg, tol);
        if (!err.empty()) {
            delete h;
            throw std::invalid_argument(err);
        }
        *out = h;
    });
}

axs_status axs_mesh_nozzle(axs_mesh** out, const axs_nozzle_spec* spec) {
    if (!out || !spec) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        axs_mesh* h = new axs_mesh();
        const std::string err = buildNozzle(h->m, *spec);
        if (!err.empty()) {
            delete h;
            throw std::invalid_argument(err);
        }
        *out = h;
    });
}

axs_status axs_mesh_concat(axs_mesh** out, int count, const axs_mesh_view* meshes, int groups_per_mesh) {
    if (!out || !meshes || count <= 0) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        axs_mesh* h = new axs_mesh();
        const std::string err = concatMeshes(h->m, count, meshes, groups_per_mesh);
        if (!err.empty()) {
            delete h;
            throw std::invalid_argument(err);
        }
        *out = h;
    });
}

axs_status axs_mesh_get_view(const axs_mesh* h, axs_mesh_view* v) {
    if (!h || !v) return AXS_ERR_INVALID;
    *v = h->m.view();
    return AXS_OK;
}

axs_status axs_mesh_get_triangles(const axs_mesh* h, int* n_points, double* pz, double* pr, int* n_tris, int* tris) {
    if (!h) return AXS_ERR_INVALID;
    if (n_points) *n_points = (int)h->m.pointZ.size();
    if (n_tris) *n_tris = (int)h->m.tris.size() / 3;
    if (pz) std::copy(h->m.pointZ.begin(), h->m.pointZ.end(), pz);
    if (pr) std::copy(h->m.pointR.begin(), h->m.pointR.end(), pr);
    if (tris) std::copy(h->m.tris.begin(), h->m.tris.end(), tris);
    return AXS_OK;
}

axs_status axs_mesh_get_quads(const axs_mesh* h, int* n_cells, double* z4, double* r4) {
    if (!h) return AXS_ERR_INVALID;
    if (n_cells) *n_cells = (int)h->m.quadZ.size() / 4;
    if (z4) std::copy(h->m.quadZ.begin(), h->m.quadZ.end(), z4);
    if (r4) std::copy(h->m.quadR.begin(), h->m.quadR.end(), r4);
    return AXS_OK;
}

void axs_mesh_destroy(axs_mesh* h) { delete h; }

// ---------------- unstructured mesher ----------------
void axs_cdt_params_defaults(axs_cdt_params* p) {
    if (!p) return;
    const axs::cdt::Params d;
    p->B = d.skinny;
    p->beta = d.grading;
    p->classify_tol = d.accept;
    p->max_points = d.maxPoints;
}

axs_status axs_cdt_generate(axs_cdt** out, int n, const double* x, const double* y, int nSeg, const int* seg,
                            const axs_cdt_params* params) {
    if (!out || n <= 0 || !x || !y || nSeg < 0 || (nSeg > 0 && !seg)) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        axs::cdt::Params prm;
        if (params) {
            prm.skinny = params->B;
            prm.grading = params->beta;
            prm.accept = params->classify_tol;
            if (params->max_points > 0) prm.maxPoints = params->max_points;
        }
        std::vector<axs::cdt::Pt> pts(n);
        for (int i = 0; i < n; i++) pts[i] = {x[i], y[i]};
        std::vector<axs::cdt::Seg> segs(nSeg);
        for (int k = 0; k < nSeg; k++) segs[k] = {seg[2 * k], seg[2 * k + 1]};
        axs_cdt* h = new axs_cdt();
        try {
            h->r = axs::cdt::generate(pts, segs, prm);
        } catch (const axs::cdt::MeshError& e) {
            delete h;
            throw std::invalid_argument(e.what());
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

axs_status axs_cdt_insert_points(axs_cdt** out, int n, const double* x, const double* y) {
    if (!out || n <= 0 || !x || !y) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        axs_cdt* h = new axs_cdt();
        h->r.points.resize(n);
        for (int i = 0; i < n; i++) h->r.points[i] = {x[i], y[i]};
        try {
            h->r.tris = axs::cdt::insertAll(h->r.points);
        } catch (const axs::cdt::MeshError& e) {
            delete h;
            throw std::invalid_argument(e.what());
        }
        *out = h;
    });
}

axs_status axs_cdt_counts(const axs_cdt* h, int* nPoints, int* nTris, int* nSeg) {
    if (!h) return AXS_ERR_INVALID;
    if (nPoints) *nPoints = (int)h->r.points.size();
    if (nTris) *nTris = (int)h->r.tris.size();
    if (nSeg) *nSeg = (int)h->r.segs.size();
    return AXS_OK;
}

axs_status axs_cdt_get(const axs_cdt* h, double* x, double* y, int* triV, int* triAdj, int* seg) {
    if (!h) return AXS_ERR_INVALID;
    const auto& r = h->r;
    for (size_t i = 0; i < r.points.size(); i++) {
        if (x) x[i] = r.points[i].x;
        if (y) y[i] = r.points[i].y;
    }
    for (size_t t = 0; t < r.tris.size(); t++)
        for (int k = 0; k < 3; k++) {
            if (triV) triV[3 * t + k] = r.tris[t].v[k];
            if (triAdj) triAdj[3 * t + k] = r.tris[t].adj[k];
        }
    if (seg)
        for (size_t k = 0; k < r.segs.size(); k++) {
            seg[2 * k] = r.segs[k].a;
            seg[2 * k + 1] = r.segs[k].b;
        }
    return AXS_OK;
}

void axs_cdt_destroy(axs_cdt* h) { delete h; }

// Quality::buildQuality (solver/source/quality.cpp:160-166) on the GPU: the mesh
// arrays and the cells' corner rings go up, one thread per cell computes the
// three metrics, the results come back.
axs_status axs_mesh_quality(const axs_mesh_view* v, const axs_cell_polygons* poly, int device, double* aspect,
                            double* non_orth, double* skew) {
    if (!v || !aspect || !non_orth || !skew || v->n_cells < 0 || v->n_faces < 0) return AXS_ERR_INVALID;
    if (v->n_cells > 0 && (!v->cell_center_z || !v->cell_center_r || !v->cell_face_start || !v->cell_face_ids ||
                           !v->face_owner || !v->face_neighbor || !v->face_center_z || !v->face_center_r ||
                           !v->face_normal_z || !v->face_normal_r))
        return AXS_ERR_INVALID;
    return guarded(nullptr, [&] {
        int count = 0;
        AXS_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw std::invalid_argument("no such device");
        AXS_HIP(hipSetDevice(device));
        const int N = v->n_cells, F = v->n_faces;
        const int S = N > 0 ? v->cell_face_start[N] : 0;
        // the corner rings as one coordinate list (quads: 4 per cell in ring
        // order; triangles: their three points); no polygons: empty rings
        std::vector<int> cs(N + 1, 0);
        std::vector<double> pz, pr;
        const int corners = poly ? poly->corners : 0;
        if (corners == 4) {
            if (!poly->quad_z || !poly->quad_r) throw std::invalid_argument("quad corners missing");
            pz.assign(poly->quad_z, poly->quad_z + 4 * (size_t)N);
            pr.assign(poly->quad_r, poly->quad_r + 4 * (size_t)N);
            for (int c = 0; c <= N; c++) cs[c] = 4 * c;
        } else if (corners == 3) {
            if (!poly->tris || !poly->point_z || !poly->point_r) throw std::invalid_argument("triangles missing");
            pz.resize(3 * (size_t)N);
            pr.resize(3 * (size_t)N);
            for (int c = 0; c < N; c++)
                for (int j = 0; j < 3; j++) {
                    const int p = poly->tris[3 * (size_t)c + j];
                    if (p < 0 || p >= poly->n_points) throw std::out_of_range("triangle corner out of range");
                    pz[3 * (size_t)c + j] = poly->point_z[p];
                    pr[3 * (size_t)c + j] = poly->point_r[p];
                }
            for (int c = 0; c <= N; c++) cs[c] = 3 * c;
        } else if (corners != 0) {
            throw std::invalid_argument("corners must be 0, 3 or 4");
        }
        for (int k = 0; k < S; k++)
            if (v->cell_face_ids[k] < 0 || v->cell_face_ids[k] >= F) throw std::out_of_range("face id out of range");
        for (int f = 0; f < F; f++)
            if (v->face_owner[f] < 0 || v->face_owner[f] >= N || v->face_neighbor[f] >= N)
                throw std::out_of_range("face cell out of range");
        axs::Arena a;
        auto up = [&](const auto* src, size_t n) {
            using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
            T* d = a.alloc<T>(n);
            if (n) AXS_HIP(hipMemcpy(d, src, n * sizeof(T), hipMemcpyHostToDevice));
            return (const T*)d;
        };
        axs::QualityMesh m{};
        m.n = N;
        m.cz = up(v->cell_center_z, N);
        m.cr = up(v->cell_center_r, N);
        m.faceStart = up(v->cell_face_start, (size_t)N + 1);
        m.faceIds = up(v->cell_face_ids, S);
        m.owner = up(v->face_owner, F);
        m.neighbor = up(v->face_neighbor, F);
        m.fz = up(v->face_center_z, F);
        m.fr = up(v->face_center_r, F);
        m.nz = up(v->face_normal_z, F);
        m.nr = up(v->face_normal_r, F);
        m.cornerStart = up(cs.data(), cs.size());
        m.pz = up(pz.data(), pz.size());
        m.pr = up(pr.data(), pr.size());
        double* out = a.alloc<double>(3 * (size_t)N);
        axs::launchMeshQuality(nullptr, m, out, out + N, out + 2 * (size_t)N);
        AXS_HIP(hipGetLastError());
        AXS_HIP(hipDeviceSynchronize());
        AXS_HIP(hipMemcpy(aspect, out, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
        AXS_HIP(hipMemcpy(non_orth, out + N, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
        AXS_HIP(hipMemcpy(skew, out + 2 * (size_t)N, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
    });
}

// ---------------- learned initial guess ----------------
axs_status axs_net_create(axs_net** out, int layers, uint64_t seed, int device) {
    if (!out) return AXS_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        int count = 0;
        AXS_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw std::invalid_argument("no such device");
        auto* h = new axs_net();
        try {
            h->net = new Mpnn(layers, seed, device);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void axs_net_destroy(axs_net* h) {
    if (!h) return;
    delete h->net;
    delete h;
}

int axs_net_num_params(const axs_net* h) { return h ? h->net->numParams() : 0; }

axs_status axs_net_get_params(const axs_net* h, float* out) {
    if (!h || !out) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] { h->net->getParams(out); });
}

axs_status axs_net_set_params(axs_net* h, const float* in) {
    if (!h || !in) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] { h->net->setParams(in); });
}

axs_status axs_net_add_sample(axs_net* h, int n, const int* fs, const int* fn, const double* AC, const double* AF,
                              const double* b, const double* x, int* index) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] {
        const int i = h->net->addSample(n, fs, fn, AC, AF, b, x);
        if (index) *index = i;
    });
}

axs_status axs_net_clear_samples(axs_net* h) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] { h->net->clearSamples(); });
}

axs_status axs_net_loss_grad(axs_net* h, int sample, double* loss, float* grad) {
    if (!h || !loss) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] { *loss = h->net->lossGrad(sample, grad); });
}

axs_status axs_net_train(axs_net* h, int epochs, int batch, double lr, uint64_t seed, double* loss_per_epoch) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] { h->net->train(epochs, batch, lr, seed, loss_per_epoch); });
}

axs_status axs_net_predict(axs_net* h, int n, const int* fs, const int* fn, const double* AC, const double* AF,
                           const double* b, double* x) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(nullptr, [&] { h->net->predict(n, fs, fn, AC, AF, b, x); });
}

axs_status axs_set_learned_guess(axs_solver* h, axs_net* net, int level) {
    if (!h) return AXS_ERR_INVALID;
    return guarded(&h->s->lastError, [&] { h->s->setLearnedGuess(net ? net->net : nullptr, level); });
}

}  // extern "C"
