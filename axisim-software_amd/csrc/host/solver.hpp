// solver.hpp -- the MI355X SIMPLE solver behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hip/device.hpp"
#include "../hip/mixed.hpp"
#include "../hip/mpnn.hpp"
#include "axs.h"
#include "mesh_types.hpp"
#include "topology.hpp"

namespace axs {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define AXS_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::axs::HipError(std::string(#call) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// Owns device allocations; frees them on destruction.
class Arena {
public:
    Arena() = default;
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    ~Arena() { release(); }
    template <class T>
    T* alloc(size_t n) {
        void* p = nullptr;
        AXS_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        ptrs_.push_back(p);
        return (T*)p;
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(v.size());
        if (!v.empty()) AXS_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    template <class T>
    T* zeros(size_t n) {
        T* p = alloc<T>(n);
        AXS_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
        return p;
    }
    void release();

private:
    std::vector<void*> ptrs_;
};

// Multigrid hierarchy resident on the device.
struct MGParams {
    int cycles = 3;
    double weight = 0.6;
    int pre = 1, post = 1, coarseSweeps = 30;
    int precision = AXS_PREC_FP64;
    int coarseDirect = 0;
    int smoother = AXS_MG_SMOOTH_JACOBI;
    int refine = AXS_REFINE_CYCLE;  // mixed precision: FP64 defect per cycle or per solve
    // level 0's FP32 operator copies are written by the p' assembly (no cast)
    bool l0Copies = false;
    // x = 0 and level 0's b32 holds b (written with the right-hand side)
    bool startZero = false;
    // coarse_direct: the coarsest inverse is formed with the operators (else
    // the one formed last is applied)
    bool formInverse = true;
};

class MGDevice {
public:
    std::vector<Level> host;
    std::vector<DLevel> dev;
    Arena arena;
    Arena arena32;  // FP32 copies and the coarsest inverse (mixed precision)
    // Coarsest-level FP64 inverse (coarse_direct): Gauss-Jordan work + result.
    double* aug = nullptr;
    double* inv = nullptr;
    int* invStatus = nullptr;
    bool mixedReady = false;

    void build(int n, const std::vector<int>& start, const std::vector<int>& nb, int maxLevels);
    // Level 0 operator and x buffers are owned by the caller.
    void bindLevel0(double* AC, double* b, double* AF, double* x, double* xNew);
    // Allocate the FP32 buffers of every level (idempotent).
    void ensureMixed();
    // the coarsest inverse's arrays (coarse_direct on the FP64 cycles);
    // allocates, so it runs before any capture
    void ensureInverse();
    // Colour every level for the Gauss-Seidel smoother (idempotent; allocates,
    // so it runs before any graph capture).
    void ensureColoring();
    // Enqueue one solve (enqueueRun, multigrid.cu:766-811).  Leaves the answer
    // in the level-0 buffer that was bound as x.
    void enqueueSolve(hipStream_t s, const MGParams& p, const Ctl* ctl, bool buildOperators);
    // Only the operator part of a solve: Galerkin operators of every level (and
    // their FP32 copies and the coarsest inverse in mixed precision).
    void enqueueOperators(hipStream_t s, const MGParams& p, const Ctl* ctl);
    double bytesPerSolve(const MGParams& p, bool buildOperators) const;
    double restrictBytes(int l, double v) const;
    int coarsestCells() const { return host.empty() ? 0 : host.back().n; }
    // Work arrays of one FP64 weighted-Jacobi solve placed in an external
    // region: level 0's second iterate and residual, every coarser level's
    // operator, right-hand side, iterates and residual, and the LDS tail's
    // image.  workBytes() is the region's size (256-byte aligned pieces);
    // bindWork() points the hierarchy at the region (level 0 at the caller's
    // AC, b, AF, x) and enqueues the copy of the tail image's integer maps
    // into it; unbindWork() restores the hierarchy's own arrays.  Maps and
    // tables stay where build() put them.
    size_t workBytes() const;
    struct Bound {
        std::vector<DLevel> dev;
        unsigned char* tailImage = nullptr;
        double *x0Home = nullptr, *x0Other = nullptr;
    };
    Bound bindWork(unsigned char* base, const double* AC, const double* b, const double* AF, double* x,
                   hipStream_t s);
    void unbindWork(const Bound& saved);
    // first level of the LDS-resident tail (-1: no tail) and its LDS bytes
    int tailLevel(bool fp32) const { return (fp32 ? tail32_ : tail64_).lt; }
    // run level 0 as one-pass launches even where the fused cycles apply
    // (the fused cycles' yardstick in tests)
    void setFuseLevel0(bool on) { fuseL0Off_ = !on; }
    size_t tailBytes(bool fp32) const { return (fp32 ? tail32_ : tail64_).bytes; }

private:
    // The tail kernel's LDS image of levels lt..lt+nl-1 (value type of size vsize)
    struct TailPlan {
        int lt = -1, nl = 0, intBase = 0, threads = 0;
        size_t bytes = 0;   // the whole image (the kernel's LDS)
        size_t staged = 0;  // its prefix a launch copies in (values read first, integers)
        std::vector<TailLevel> lv;
        TailLevel* dev = nullptr;        // device copy of lv
        unsigned char* image = nullptr;  // the tail levels' arrays, laid out as in LDS
        bool holds(int l) const { return lt >= 1 && l >= lt; }
        TailLaunch launch() const;
    };
    TailPlan planTail(int lt, size_t vsize) const;
    TailPlan selectTail(size_t vsize) const;
    void placeInts(const TailPlan& t, int l, DLevel& d);
    double tailTrafficBytes(const TailPlan& t, size_t vsize) const;
    void vcycle(hipStream_t s, int l, int preDone, const MGParams& p, const Ctl* ctl);
    void smooth(hipStream_t s, int l, int sweeps, const MGParams& p, const Ctl* ctl);
    void restrictTo(hipStream_t s, int l, int sweep1, bool fp32, const MGParams& p, const Ctl* ctl);
    void enqueueSolveMixed(hipStream_t s, const MGParams& p, const Ctl* ctl, bool buildOperators);
    void vcycle32(hipStream_t s, int l, int preDone, const MGParams& p, const Ctl* ctl,
                  double* apply = nullptr);
    void smooth32(hipStream_t s, int l, int sweeps, const MGParams& p, const Ctl* ctl);
    void buildOperators(hipStream_t s, const Ctl* ctl);
    // multi-colour Gauss-Seidel smoother (AXS_MG_SMOOTH_RBGS)
    void smoothGS(hipStream_t s, int l, int sweeps, bool fp32, const Ctl* ctl);
    void vcycleGS(hipStream_t s, int l, bool fp32, const MGParams& p, const Ctl* ctl);
    std::vector<std::vector<int>> colorStart_;
    std::vector<int*> dOrder_, dColorStart_;
    void swapX(int l);
    void swapX32(int l);
    double* x0Home_ = nullptr;
    double* x0Other_ = nullptr;
    TailPlan tail64_, tail32_;
    int ldsLimit_ = 0;
    // level 0's blocks and halo rings for the fused level-0 sweeps (FP64
    // Jacobi V-cycles with one pre- and one post-sweep)
    SweepBlocks l0Blocks_;
    DSweepBlocks dl0_{};
    bool fuseL0_ = false;
    bool fuseL0Off_ = false;
    // the restriction into level 1 and level 1's residual + restriction in one
    // launch (fused level-0 cycles): per level-1 aggregate workgroup, the halo
    // of level-1 cells its rows read and each row slot's LDS index
    L1Fused l1f_{};
    bool fuseL1_ = false;
    std::vector<std::vector<AggThread>> aggThreadsHost_;
    static constexpr int kFuseL0MinCells = 16384;
    bool fusesLevel0(const MGParams& p) const;
    void cyclesFused(hipStream_t s, int cycles, const MGParams& p, const Ctl* ctl, bool fp32 = false,
                     double* apply = nullptr);
};

class Solver {
public:
    explicit Solver(int device);
    ~Solver();

    void setMesh(const axs_mesh_view& v);
    void setBC(int var, const axs_bc_view& v);
    void configure(const axs_config& c);
    void run(axs_progress_cb cb, void* user, bool continueRun);
    void runTimed(int n, double* ms);
    void benchKernel(int kernel, int repeats, double* ms, double* bytes, bool cold = false);
    // Workgroup sizes of the momentum sweep and of every multigrid level's
    // per-cell kernels, each chosen from 64 / 128 / 256 threads by timing them on
    // the current state (after a run).  *msDefault / *msTuned: summed kernel
    // times at 256 threads and at the chosen sizes.
    void autotune(int repeats, double* msDefault, double* msTuned);
    int sweepBlock() const { return dm_.sweepBlock; }
    void requestStop() { stop_ = true; }
    bool canContinue(std::string& reason) const;
    void getField(int field, double* out, size_t n);
    size_t fieldLength(int field) const;
    int residualHistory(int res, double* out, int maxN) const;
    void stats(axs_stats& s) const;
    // Transient frames (Solver::TimeFrame / timeFrames, solver/header/solver.h:96-108)
    struct Frame {
        double time = 0.0;
        std::vector<double> u, v, p, t, c;  // t, c: when those fields are solved
    };
    // wallConsumptionDiagnostic of the last run (axs_get_wall_diagnostic)
    void wallDiagnostic(double out[10]);
    int frameCount() const { return (int)frames_.size(); }
    const HostMesh& hostMesh() const { return mesh_; }
    // Batched runs: the union mesh's cases as contiguous cell / face ranges
    // (count 0: one case).  Each case then keeps its own residuals, continuity
    // scale and convergence, and stops on them.
    void setCases(int count, const int* cellStart, const int* faceStart);
    int caseCount() const { return nCases_; }
    void caseStats(int c, int& iterations, int& converged, double* lastResidual, int& stopIteration) const;
    const axs_config& config() const { return cfg_; }
    const Frame& frame(int index) const;
    // Learned initial guess of the first p' solve of every outer iteration
    // (net null: off).  level: multigrid level the net runs on, -1 = the first
    // one with at most kGuessCells cells.
    void setLearnedGuess(class Mpnn* net, int level);
    static constexpr int kGuessCells = 16384;
    // Momentum Jacobi sweeps fused per launch (1 = one sweep per launch, 2 or
    // 3) and the cells per fused block (512 or 1024; 0: 1024 where its halos
    // fit, else 512).  The next run uses the largest count <= sweeps whose
    // halos fit; results are unchanged.
    void setSweepFusion(int sweeps, int blockCells);
    int sweepsPerLaunch() const { return fuse_; }

    std::string lastError;

private:
    struct HostBC {
        int nGroups = 0;
        std::vector<uint8_t> type;        // device type (PULSATILE -> DIRICHLET)
        std::vector<uint8_t> pulsatile;
        std::vector<double> value, length, amplitude, frequency;
        // kinetics (scalars): Vmax, Km, n, m, K2, V2, wall resistance, Km^n, K2^m
        std::vector<double> vmax, km, hillN, inhibM, k2, v2, rtot, kmN, k2M;
        bool kinetics = false;
    };

    void uploadMesh();
    void allocateState();
    void uploadBCs(double time);
    void buildGraph();
    void enqueueIteration(hipStream_t s);
    Physics physics() const;
    void launchBatch(int n);
    void pullControl(axs_progress_cb cb, void* user);
    double bytesPerIteration() const;

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    HostMesh mesh_;
    bool haveMesh_ = false;
    Geometry geom_;
    std::vector<int> csrStart_, csrNb_;
    axs_config cfg_{};
    bool haveCfg_ = false;
    HostBC bc_[AXS_VAR_COUNT];

    Arena meshArena_, stateArena_;
    DMesh dm_{};
    DState ds_{};
    DBC dbc_[AXS_VAR_COUNT]{}, ppbc_{};
    double* bcValueDev_[AXS_VAR_COUNT] = {};
    Arena bcArena_;
    bool meshUploaded_ = false, stateAllocated_ = false;
    int partialBlocks_ = 0;
    int histCap_ = 0;

    MixedState mx_{};
    Arena mixedArena_;
    bool mixedAllocated_ = false;
    void allocateMixed();

    MGDevice mg_;
    // learned guess: net, its level, the level's graph, work buffers and the
    // restricted right-hand sides / prolonged guesses of levels 1..level
    class Mpnn* net_ = nullptr;
    int netLevelReq_ = -1, netLevel_ = -1;
    struct NetSetup;
    NetSetup* netSetup_ = nullptr;
    void setupGuess();
    void enqueueGuess(hipStream_t s, const MGParams& mp);
    bool mgBuilt_ = false;
    int mgBuiltLevels_ = -1;
    // fused momentum sweeps: requested and in-use sweeps per launch, the blocks
    // and halo rings (host and device copies, built for fuseCellsReq_ and
    // fuse_ - 1 rings)
    int fuseReq_ = kMaxFusedSweeps, fuseCellsReq_ = 0, fuse_ = 1;
    bool mgFuseL0_ = true;  // the caller's choice only (the autotuner leaves it)
    int graphNodes_ = 0;  // nodes of the captured iteration graph
    SweepBlocks sweepBlocks_;
    DSweepBlocks dsb_{};
    Arena sweepArena_;
    void prepareSweepBlocks();
    void dropGraph();
    void captureIteration(hipGraph_t* graph, hipGraphExec_t* exec);
    std::vector<int> sweepChunks(int sweeps) const;
    std::vector<int> colorStart_;
    int* dOrder_ = nullptr;
    Arena colorArena_;
    int nColors_ = 0;

    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    // coarse_direct = K > 1: the iteration without the coarsest inverse's
    // formation, launched on all but every K-th outer iteration
    hipGraph_t graphLag_ = nullptr;
    hipGraphExec_t execLag_ = nullptr;
    long long launches_ = 0;     // graph launches since the graphs were built
    bool formInverse_ = true;    // read by enqueueIteration while capturing

    Ctl* hctl_ = nullptr;   // pinned mirror
    double* hhist_ = nullptr;
    int hhistCap_ = 0;

    std::atomic<bool> stop_{false};
    std::atomic<bool> running_{false};
    int currentIteration_ = 0;
    bool converged_ = false, stopped_ = false, ready_ = false;
    std::vector<double> history_[AXS_RES_COUNT];
    // scalar transport: temperature and concentration state (allocated when
    // solved), the reactive walls' per-face state
    // batched cases (setCases): host tables, their device copies, the per-case
    // control blocks with a host mirror, the residual partials and the
    // snapshots of the fields of cases done
    int nCases_ = 0;
    std::vector<int> caseCellStart_, caseFaceStart_;
    DCases dcases_{};
    Arena caseArena_;
    int caseBlocks_ = 0;
    double* casePartial_ = nullptr;
    std::vector<CaseCtl> caseHost_;
    FreezeSet freezeCells_{}, freezeFaces_{};
    bool casesReady_ = false;
    void prepareCases();
    void resetCases();
    DScalar sc_[2]{};
    DWall wall_{};
    bool scAlloc_[2] = {false, false};
    bool scalarsOn() const { return cfg_.solve_energy || cfg_.solve_concentration; }
    void enqueueScalars(hipStream_t s);
    ScalarPhysics scalarPhysics(int which) const;
    double setupMs_ = 0.0, solveMs_ = 0.0;
    std::vector<Frame> frames_;
    void captureFrame(double time);
    int contCells_ = -1, contFaces_ = -1, contRefs_ = -1;
};

}  // namespace axs
