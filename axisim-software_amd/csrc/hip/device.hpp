// device.hpp -- device data views and the kernel launch interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace axs {

constexpr int kBlock = 256;   // threads per workgroup (4 wave64)

// Mesh on the device: SoA cells + faces (FVMeshDevice, structs/solver_struct.h:333-397).
struct DMesh {
    int N = 0, F = 0, S = 0;
    int sweepBlock = kBlock;  // workgroup size of the momentum sweep (autotuned)
    int maxSlots = 0;         // most slots of a cell (the cell loops batch up to 3 or 4)
    int uniformSlots = 0;     // > 0: every cell has this many slots (faceStart[n] = uniformSlots * n)
    const double *cZ, *cR, *vol, *invA2D, *lsqInvZZ, *lsqInvZR, *lsqInvRR;
    const int *faceStart, *faceIDs;
    const int *slotNb;      // neighbour cell of each slot, -1 on a boundary face
    const int *owner, *neighbor, *group;
    const double *nZ, *nR, *fZ, *fR, *area, *invDPN, *wP, *dPB, *lsqWZ, *lsqWR, *length2D;
    // face data in slot order, oriented for the slot's cell (SlotGeometry)
    const double *sArea, *sInvDPN, *sNZ, *sNR, *sW, *sTZ, *sTR, *sLsqWZ, *sLsqWR, *sLen;
    const uint8_t* sFlags;  // bit 0: the cell owns the face; bit 1: face on the axis
};

// Blocks of consecutive cells with their halo rings (SweepBlocks, device copy):
// the fused momentum sweeps run `sweeps` Jacobi sweeps per launch, one
// workgroup of `cells` threads per block, the intermediate iterates in LDS.
struct DSweepBlocks {
    int cells = 0, rings = 0, nBlocks = 0;
    const int* haloStart = nullptr;
    const int* ringEnd = nullptr;
    const int* halo = nullptr;
    const int* own = nullptr;          // non-null: block b's cells are own[b*cells ..] (-1: none)
    const int16_t* locOwn = nullptr;   // LDS slot of each slot's neighbour (block rows)
    const int16_t* locHalo = nullptr;  // the same for halo rows, kSweepRegSlots per entry
};
constexpr int kSweepRegSlots = 4;   // slots of a block cell kept in registers across the sweeps
constexpr int kMaxFusedSweeps = 5;  // momentum Jacobi sweeps one fused launch may run

// Boundary conditions of one variable (BoundaryFieldDevice, solver_struct.h:431-454).
struct DBC {
    int nGroups = 0;
    const uint8_t* type = nullptr;
    const double* value = nullptr;
    const double* length = nullptr;
    // Michaelis-Menten / Hill kinetics per group (scalars only; null otherwise):
    // Vmax, Km, Hill n, inhibition m and V2, wall-layer resistance, Km^n, K2^m
    const double *vmax = nullptr, *km = nullptr, *hillN = nullptr, *inhibM = nullptr, *v2 = nullptr;
    const double *rtot = nullptr, *kmN = nullptr, *k2M = nullptr;
};

// One transported scalar (temperature or concentration): field, Jacobi
// partner, old levels, gradient and its linear system (VariablesSimple temp /
// conc, Coefficients "Temperature" / "Concentration").
struct DScalar {
    double *phi = nullptr, *tmp = nullptr, *old = nullptr, *old2 = nullptr, *gZ = nullptr, *gR = nullptr;
    double *AC = nullptr, *b = nullptr, *inv = nullptr, *AF = nullptr;
};
// Reactive walls (FVFaceDevice cw / ocrWall, per face).
struct DWall {
    double *cw = nullptr, *ocr = nullptr;
};
// What the scalar assembly needs besides the mesh and the BCs.
struct ScalarPhysics {
    double gamma;      // k / (rho cp) or D
    double fluxScale;  // 1 / rho: the volumetric flux
    double alpha;      // under-relaxation (alpha_u for temperature, 1 for species)
    double dt;
    int conv, scheme, nonOrtho, transient;
};

// Run control shared by the host and the per-iteration graph.
struct Ctl {
    int done;          // set once the outer loop must end: every kernel returns at once
    int iter;          // currentIteration
    int k;             // iteration index within the current time step
    int maxIter;
    int checkConv;
    int transient;
    int bdf2Step;      // BDF2 term active in this time step
    int nHist;         // residual checks recorded
    int histCap;       // capacity of DState::hist in checks
    int converged;
    int N;
    int enabled[5], norm[5], scaling[5];   // U, V, Continuity, Temperature, Concentration
    double tol[5];
    double resVal[5], scaleVal[5];
};
constexpr int kHistStride = 6;  // per residual check in DState::hist: five residuals, then the iteration

// One case of a batched run (many cases solved as one union mesh): its own
// convergence state, as a run of that case alone keeps it (solver.cpp:1243-1283).
struct CaseCtl {
    int done;       // converged, or out of iterations, in the current time step
    int frozen;     // its fields are restored from the snapshot after every iteration
    int converged;
    int iters;      // outer iterations the case has run (all time steps)
    int stopIter;   // the run's iteration count when the case last stopped
    double resVal[5], scaleVal[5];
};
// Device view of the cases: residual blocks per case and the cell / face maps
// the freeze kernels read.
struct DCases {
    int n = 0;                      // cases (0: not a batched run)
    const int* blockStart = nullptr;  // n + 1: the case's residual blocks
    const int2* blockCells = nullptr;  // per residual block: [first cell, end)
    const int* cellCount = nullptr;  // n: cells of each case (N of its scalings)
    const int* caseOfCell = nullptr;
    const int* caseOfFace = nullptr;
    CaseCtl* ctl = nullptr;          // n
};

// Everything the SIMPLE kernels touch (VariablesSimple + Coefficients).
struct DState {
    double *u, *v, *p, *pp, *uTmp, *vTmp, *ppTmp;
    double *DU, *DV, *gPZ, *gPR, *gUZ, *gUR, *gVZ, *gVR, *mDot;
    double *uOld, *vOld, *uOld2, *vOld2;
    double *uAC, *ub, *uInv, *vAC, *vb, *vInv, *AFuv;
    double *ppAC, *ppb, *ppInv, *ppAF;
    double2 *uvA, *uvB;  // interleaved (u, v) iterates of the momentum Jacobi sweeps
    // FP32 copies written by kMomentum in mixed precision (null otherwise)
    float* AF32;
    float2* uvInv32;  // (1/AC_u, 1/AC_v) interleaved
    // FP32 copies of the p' system for the mixed cycles refined once per solve
    // (null otherwise): kPPCoeff writes AC and AF, kPPRhs b when it zeroes p'
    float *ppAC32, *ppAF32, *ppb32;
    // the momentum defect b - A u0 (u, v) in FP32, written by kMomentum for the
    // mixed path's fused FP32 sweeps (null otherwise)
    float2* momR32;
    double *partial;   // per-block residual partial sums (9 per block, then 8 of the scalars)
    double *hist;      // residual history: kHistStride values per check
    Ctl* ctl;
};

struct Physics {
    double rho, mu, alphaU, alphaP, dt;
    int conv, scheme, gradScheme, nonOrtho;
    int transient;   // add the unsteady term (BDF1, or BDF2 when ctl->bdf2Step)
};

// One multigrid level on the device (MultigridLevel, multigrid.cuh:60-106)
// Per-thread entry of the aggregate workgroups below: member m of the run
// (-1: none) with its row's slot range, and, for coarse cell j0 + t of the
// run, where its members' values sit in the workgroup's LDS (sum = a | e << 16,
// the range [a, e); -1: none).  Built once on the host so that a thread reaches
// its row in one load instead of a chain through the run and member lists.
struct alignas(16) AggThread {
    int m, k0, ns, sum;
};
// Per-workgroup entry: coarse cells [j0, j1), their coarse slots [cs0, cs1).
struct alignas(16) AggBlock {
    int j0, j1, cs0, cs1;
};

struct DLevel {
    int n = 0, refs = 0, nCoarse = 0;
    int block = kBlock;  // workgroup size of this level's per-cell kernels (autotuned)
    int maxRow = 0;      // most off-diagonal slots of a row
    int maxMembers = 0;  // most fine cells of one coarse cell (levels with a coarser one)
    int uniformRow = 0;  // > 0: every row has this many slots (start[n] = uniformRow * n)
    // workgroups of kMGResidualRestrict and kMGCoarseOperator: runs of whole
    // coarse cells whose members fit one workgroup of kBlock threads
    // (nAggBlocks entries, nAggBlocks * kBlock thread entries; 0: not built)
    const AggBlock* aggBlocks = nullptr;
    const AggThread* aggThreads = nullptr;
    int nAggBlocks = 0;
    int *start = nullptr, *nb = nullptr;
    double *AC = nullptr, *b = nullptr, *AF = nullptr;
    double *x = nullptr, *xNew = nullptr;
    double* res = nullptr;  // per-cell residual / diagonal scratch (levels with a coarser one)
    int *toCoarse = nullptr, *slotToCoarseSlot = nullptr;
    int *memberStart = nullptr, *members = nullptr;
    int *cslotStart = nullptr, *cslotFine = nullptr;
    // FP32 copies for the mixed-precision V-cycle
    float *AC32 = nullptr, *b32 = nullptr, *AF32 = nullptr, *x32 = nullptr, *xNew32 = nullptr;
    float* res32 = nullptr;
};

// Where one level of the V-cycle tail lives in the tail kernel's LDS image:
// offsets in elements of the value type (AC, b, x0, x1, AF) and of int
// (start, nb, toC, mS, mem).  Built by the host, read by the kernel.
struct TailLevel {
    int n, refs, nc;
    int AC, b, x0, x1, AF;
    int start, nb, toC, mS, mem;
    int perm;  // the coarsest level: cells in ascending slot count (thread t sweeps cell perm[t]); -1: none
};
constexpr int kTailLevels = 8;         // levels the tail kernel may hold
constexpr int kTailStaticLds = 2048;   // LDS reserved next to the tail's dynamic image
// Everything a tail launch needs: the plan (device copy) and the global image
// of the tail levels' arrays (16-byte vectors), staged into LDS as it is.
struct TailLaunch {
    const TailLevel* plan = nullptr;
    unsigned char* image = nullptr;
    int vecs = 0, lt = -1, nl = 0, intBase = 0, threads = 0;
    size_t ldsBytes = 0;
};

inline int blocksFor(long n, int t = kBlock) { return (int)((n + t - 1) / t); }

// ---- launchers (kernels.hip) ----
// withP: also the pressure gradient (then launchMomentum with gradPDone)
void launchGradUV(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DBC bp, DState st, int scheme, bool withP);
void launchGradField(hipStream_t s, const DMesh& m, DBC bc, const double* phi, double* gz, double* gr,
                     int scheme, const Ctl* ctl);
void launchMomentum(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DBC bp, DState st, Physics ph, bool gradPDone);
void launchJacobiUV(hipStream_t s, const DMesh& m, DState st, const double* uo, const double* vo, double* un,
                    double* vn);
// the same sweep with either side interleaved as (u, v) pairs: a non-null uvo
// (uvn) replaces uo/vo (un/vn)
void launchJacobiUVPacked(hipStream_t s, const DMesh& m, DState st, const double* uo, const double* vo,
                          const double2* uvo, double* un, double* vn, double2* uvn);
// `sweeps` (2 or 3, at most sb.rings + 1) Jacobi sweeps in one launch; the
// mesh's cells have at most kSweepRegSlots slots and every block's halo fits
// sweepHaloMax(sb.cells)
int sweepHaloMax(int cells);
void launchJacobiUVBlock(hipStream_t s, const DMesh& m, DState st, const DSweepBlocks& sb, int sweeps,
                         const double* uo, const double* vo, const double2* uvo, double* un, double* vn,
                         double2* uvn);
void launchGSColorUV(hipStream_t s, const DMesh& m, DState st, const int* order, int begin, int count);
void launchJacobiSys(hipStream_t s, const DMesh& m, const double* b, const double* inv, const double* AF,
                     const double* xo, double* xn, const Ctl* ctl);
void launchGSColorSys(hipStream_t s, const DMesh& m, const double* b, const double* inv, const double* AF, double* x,
                      const int* order, int begin, int count, const Ctl* ctl);
void launchPPCoeff(hipStream_t s, const DMesh& m, DBC bp, DState st, Physics ph);
void launchFaceFlux(hipStream_t s, const DMesh& m, DBC bu, DBC bv, DBC bp, DState st, Physics ph);
// zero: when non-null, also zero[n] = 0 for every cell (p' before the solve)
void launchPPRhs(hipStream_t s, const DMesh& m, DState st, Physics ph, int cross, double* zero = nullptr);
void launchGradPPUpdate(hipStream_t s, const DMesh& m, DBC ppbc, DState st, Physics ph);
void launchMassFluxUpdate(hipStream_t s, const DMesh& m, DBC bp, DState st, Physics ph);
void launchResidual(hipStream_t s, const DMesh& m, DState st, int blocks, const int2* blockCells = nullptr);
// scalars: non-null phi of each solved scalar; their residual partials go
// after the momentum ones (8 per block from st.partial + 9 * blocks)
void launchResidualScalars(hipStream_t s, const DMesh& m, DState st, DScalar t, DScalar c, int blocks,
                           const int2* blockCells = nullptr);
void launchFinish(hipStream_t s, DState st, int blocks, bool scalars);
// batched run: each case's residuals, scale capture and convergence in one
// workgroup per case (blocks: residual blocks of all cases), then the run's
// bookkeeping; then the cases done keep their fields (restore the frozen ones
// from the snapshot, snapshot the ones just done).  snapCells / fieldsCells: nc
// cell arrays; snapFaces / fieldsFaces: nf face arrays.
void launchFinishCases(hipStream_t s, DState st, DCases cs, int blocks, bool scalars);
constexpr int kFreezeMax = 8;
struct FreezeSet {
    int n = 0;
    double* field[kFreezeMax] = {};
    double* snap[kFreezeMax] = {};
};
void launchFreeze(hipStream_t s, DCases cs, int nCells, FreezeSet cells, int nFaces, FreezeSet faces);
// the scalar equation of one field: transport, wall kinetics, unsteady term,
// under-relaxation and inverse diagonal (solver.cpp:1222-1235)
void launchScalar(hipStream_t s, const DMesh& m, DBC bc, DScalar sc, DWall wall, const double* mDot, ScalarPhysics ph,
                  const Ctl* ctl);
// old levels of a scalar at the start of a time step (old2 <- old first when bdf2)
void launchRotate(hipStream_t s, int n, double* old2, double* old, const double* phi, int bdf2);
void launchStepBegin(hipStream_t s, const DMesh& m, DState st, int bdf2);
void launchCopy(hipStream_t s, double* dst, const double* src, int n, const Ctl* ctl);
// Mesh quality metrics (quality.hip): the mesh arrays one thread per cell reads.
struct QualityMesh {
    int n;
    const double *cz, *cr;            // cell centres
    const int *faceStart, *faceIds;   // the cell's faces
    const int *owner, *neighbor;
    const double *fz, *fr, *nz, *nr;  // face centres, unit normals
    const int* cornerStart;           // n + 1: the cell's corners, in ring order
    const double *pz, *pr;            // corner coordinates, one entry per corner
};
void launchMeshQuality(hipStream_t s, const QualityMesh& m, double* aspect, double* orth, double* skew);
void launchFill(hipStream_t s, double* dst, double value, int n, const Ctl* ctl);
// streams src[0 .. n) through the caches; out[b] gets workgroup b's partial
// sum only where it equals `mark`
void launchStreamRead(hipStream_t s, const double* src, size_t n, double* out, int blocks, double mark);

// multigrid (FP64)
void launchMGJacobi(hipStream_t s, const DLevel& L, double w, const Ctl* ctl);
void launchMGSmoothBlock(hipStream_t s, const DLevel& L, double w, int sweeps, const Ctl* ctl);
void launchMGProlongSmooth(hipStream_t s, const DLevel& f, const DLevel& c, double w, const Ctl* ctl);
void launchMGProlong(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl);
// level-0 sweeps fused per launch over the blocks sb (512 cells, halo rings):
// `sweeps` weighted-Jacobi sweeps (after the prolongation when `prolong`), then
// the residual into f.res; the final x in f.xNew.  Supported: 1 sweep without
// the prolongation (a cycle's start), 2 sweeps after it (one cycle's end and
// the next one's start).  fp32: on the FP32 copies (x32, b32, res32, ...).
void launchMGLevel0Block(hipStream_t s, const DLevel& f, const DLevel& c, const DSweepBlocks& sb, int sweeps,
                         bool prolong, double w, bool fp32, const Ctl* ctl);
// the first FP32 launch of a solve refined once (1 sweep from e = 0, then the
// residual): defect: its right-hand side is the FP64 defect b - A x formed by
// the launch (and written to b32), else b32 holds it already and x = 0
void launchMGLevel0BlockStart32(hipStream_t s, const DLevel& f, const DLevel& c, const DSweepBlocks& sb, double w,
                                bool defect, const Ctl* ctl);
// multi-colour Gauss-Seidel smoothing: one colour (order[begin .. begin+count)),
// or every sweep and colour of a level of at most kSmoothBlockGSMax cells in one
// workgroup (colorStart: device copy, nColors + 1 entries)
constexpr int kSmoothBlockGSMax = 4096;
void launchMGGSColor(hipStream_t s, const DLevel& L, const int* order, int begin, int count, bool fp32,
                     const Ctl* ctl);
void launchMGSmoothBlockGS(hipStream_t s, const DLevel& L, const int* order, const int* colorStart, int nColors,
                           int sweeps, bool fp32, const Ctl* ctl);
// V-cycle of levels lt..lt+nl-1 in one workgroup, all of them staged in LDS
void launchMGTail(hipStream_t s, const TailLaunch& t, double w, int preDone, int pre, int post, int coarseSweeps,
                  bool fp32, const double* inv, const Ctl* ctl);
// LDS one workgroup may use on the current device (opting the tail kernel in
// beyond 64 KiB where the device allows it)
int tailLdsLimit();
// Galerkin product in two coalesced passes (fine diagonals, then gathers)
void launchMGCoarseDiag(hipStream_t s, const DLevel& f, const Ctl* ctl);
void launchMGCoarseGather(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl);
// both in one launch over f.aggBlocks (the fine diagonals stay in LDS)
void launchMGCoarseOperator(hipStream_t s, const DLevel& f, const DLevel& c, const Ctl* ctl);
// kMGRestrictSum into level 1 followed by level 1's kMGResidualRestrict into
// level 2, in one launch over level 1's aggregate workgroups: each workgroup
// forms b and x of its level-1 cells and of the halo cells their rows read.
constexpr int kL1Slots = 8;  // row slots of a level-1 cell the LDS-index table holds
struct L1Fused {
    const int* halo = nullptr;     // nAggBlocks * kBlock: halo cell of (block, thread), -1: none
    const int16_t* loc = nullptr;  // nAggBlocks * kBlock * kL1Slots: LDS index of each slot's neighbour, -1: none
};
void launchMGRestrictFused01(hipStream_t s, const DLevel& l0, const DLevel& l1, const DLevel& l2, L1Fused t,
                             double w, int sweep1, int sweep2, bool fp32, const Ctl* ctl);
// residual per fine cell + ordered gather per coarse cell (sweep1: also the
// coarse level's first smoothing sweep from zero); fp32: the FP32 copies
void launchMGResidual(hipStream_t s, const DLevel& f, const Ctl* ctl);
void launchMGRestrictSum(hipStream_t s, const DLevel& f, const DLevel& c, double w, int sweep1, bool fp32,
                         const Ctl* ctl);
// the two above in one launch over f.aggBlocks (the fine residual stays in LDS)
void launchMGResidualRestrict(hipStream_t s, const DLevel& f, const DLevel& c, double w, int sweep1, bool fp32,
                              const Ctl* ctl);

}  // namespace axs
