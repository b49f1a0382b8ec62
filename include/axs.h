/*
 * axs.h -- C ABI of the MI355X SIMPLE / multigrid solver (libaxs.so).
 *
 * Plain C, POD views, one opaque handle per solver.  Every pointer handed in
 * is host memory and is copied before the call returns unless the entry point
 * says "device".  All entry points return an axs_status; axs_last_error()
 * gives the message of the last failure on that handle (thread-local text for
 * entry points without a handle, read with axs_last_error(NULL)).
 *
 * What each entry point replaces in the reference (Luippe/AxiSim-Software):
 *   axs_create / axs_destroy      Solver::Solver(Config&), ~Solver           solver/header/solver.h:21-24, solver/source/solver.cpp:104-114
 *   axs_set_mesh                  Mesh::getFVMesh() copy + createFVMeshDevice solver/source/solver.cpp:774-785, manager/source/memory_manger.cpp:621-681
 *   axs_set_bcs                   createBoundarySolverDevice(boundaryGroups)  manager/source/memory_manger.cpp:51-237
 *   axs_configure                 Solver::configSolver/configSimple/configMultigrid/cfg/simple.*Relaxation/f
 *                                 (written directly by the GUI)               solver/header/solver.h:92-139, structs/solver_struct.h:158-263
 *   axs_run                       Solver::run(Mesh&) -> runSimple (blocking form; the C++ shim owns the thread)
 *                                                                             solver/source/solver.cpp:146-173, 753-1396
 *   axs_request_stop              Solver::requestStop()                       solver/source/solver.cpp:175-184
 *   axs_can_continue              Solver::canContinue(mesh, &reason)          solver/source/solver.cpp:469-528
 *   axs_get_field                 Solver::solutions[name].field               solver/source/solver.cpp:186-263
 *   axs_get_residual_history      ResidualPlot::add(iteration, cfg) feed      graphics/gui/source/residual_plot.cpp:113-130
 *   axs_frame_count / get_frame   Solver::timeFrames                          solver/header/solver.h:96-108, solver/source/solver.cpp:265-296
 *   axs_get_stats                 console lines: hierarchy, colours, timing   solver/source/multigrid.cu:727-764, solver/source/solver.cpp:836-841, 1318-1323
 *   axs_mg_*                      MultigridSolver (stand-alone p' solve)      solver/source/multigrid.cu:23-31, 298-325, 766-811
 *   axs_build_coloring            buildMeshColoring                           manager/source/memory_manger.cpp:334-409
 *   axs_mesh_*                    FVMesh producers: toPackedMesh / createUnstructuredMesh
 *                                                                             structs/multiblock.h:59-444, solver/source/mesh.cpp:1835-2002
 *   axs_cdt_*                     AxiMesh::generateMesh / insertPoints        aximesh/include/aximesh/aximesh.h:87-100, aximesh/source/aximesh.cpp
 */
#ifndef AXS_H
#define AXS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int axs_status;
enum {
    AXS_OK = 0,
    AXS_ERR_INVALID = 1,      /* bad argument or inconsistent input            */
    AXS_ERR_DEVICE = 2,       /* HIP runtime error (the reference aborts here) */
    AXS_ERR_STATE = 3,        /* call out of order (no mesh, still running ...)*/
    AXS_ERR_UNSUPPORTED = 4,  /* option the GPU path does not implement        */
    AXS_ERR_RANGE = 5         /* size or index out of range                    */
};

/* Enum values equal the reference's enum order (structs/boundary_struct.h:88-97,
 * 406-412; structs/solver_struct.h:19-63, 193-196). */
enum { AXS_BC_DIRICHLET = 0, AXS_BC_NEUMANN = 1, AXS_BC_FULLY_DEVELOPED = 2,
       AXS_BC_MICHAELIS_MENTEN = 3, AXS_BC_HILL = 4, AXS_BC_NONE = 5, AXS_BC_PULSATILE = 6 };
enum { AXS_BOUNDARY_WALL = 0, AXS_BOUNDARY_VELOCITY_INLET = 1, AXS_BOUNDARY_PRESSURE_OUTLET = 2,
       AXS_BOUNDARY_SYMMETRY = 3, AXS_BOUNDARY_FAR_FIELD = 4 };
enum { AXS_CONV_UPWIND = 0, AXS_CONV_CENTRAL = 1, AXS_CONV_SOU = 2, AXS_CONV_QUICK = 3 };
enum { AXS_GRAD_GREEN_GAUSS = 0, AXS_GRAD_LSQ = 1 };
enum { AXS_LINEAR_JACOBI = 0, AXS_LINEAR_GS_RB = 1 };
enum { AXS_TIME_BDF1 = 0, AXS_TIME_BDF2 = 1 };
enum { AXS_NORM_L1 = 0, AXS_NORM_L2 = 1, AXS_NORM_LINF = 2 };
enum { AXS_SCALE_NONE = 0, AXS_SCALE_N = 1, AXS_SCALE_SQRT_N = 2, AXS_SCALE_DIAGONAL = 3,
       AXS_SCALE_CONTINUITY = 4 };
/* Build extensions with no reference counterpart. */
enum { AXS_PREC_FP64 = 0,      /* reference arithmetic and schedule                  */
       AXS_PREC_MIXED = 1 };   /* FP32 smoothers + V-cycles, FP64 residual/flux/coarse */
enum { AXS_MG_SMOOTH_JACOBI = 0, AXS_MG_SMOOTH_RBGS = 1 };
/* Mixed precision, pressure correction: the FP64 defect b - A x before every
 * FP32 V-cycle (x += e after each), or once per solve (all cycles FP32 on the
 * correction e from 0, with the level-0 work fused across cycles as on the
 * FP64 path, x += e at the end; the outer SIMPLE loop, whose continuity
 * imbalance and fluxes are FP64, refines). */
enum { AXS_REFINE_CYCLE = 0, AXS_REFINE_SOLVE = 1 };
/* Variables with boundary conditions (BoundaryVariable, structs/boundary_struct.h:88-97
 * order U, V, P, StaticTemperature, Concentration). */
enum { AXS_VAR_U = 0, AXS_VAR_V = 1, AXS_VAR_P = 2, AXS_VAR_T = 3, AXS_VAR_C = 4, AXS_VAR_COUNT = 5 };
/* Residuals in the order of Solver::residualPlotType (solver/header/solver.h:35). */
enum { AXS_RES_U = 0, AXS_RES_V = 1, AXS_RES_CONTINUITY = 2, AXS_RES_TEMPERATURE = 3, AXS_RES_CONCENTRATION = 4,
       AXS_RES_COUNT = 5 };

/* Face-based mesh, structure of arrays (FVMeshHostPacked, structs/solver_struct.h:400-429).
 * Slot k of cell c is cell_face_ids[cell_face_start[c] + k]. */
typedef struct axs_mesh_view {
    int n_cells;
    int n_faces;
    const double* cell_center_z;
    const double* cell_center_r;
    const double* cell_volume;     /* revolved volume 2*pi*r_c*A2D           */
    const double* cell_area2d;     /* planar cross-section                   */
    const int* cell_face_start;    /* n_cells + 1                            */
    const int* cell_face_ids;      /* cell_face_start[n_cells]               */
    const int* face_owner;
    const int* face_neighbor;      /* -1 on a boundary face                  */
    const double* face_normal_z;   /* unit, owner -> neighbour / outward     */
    const double* face_normal_r;
    const double* face_center_z;
    const double* face_center_r;
    const double* face_area;       /* revolved area 2*pi*r_f*L               */
    const int* face_group;         /* boundary group id, -1 if none          */
} axs_mesh_view;

/* Boundary conditions of one variable, indexed by boundary group id
 * (BoundaryFieldHost, structs/solver_struct.h:464-476).  PULSATILE groups
 * are Dirichlet on the device with value(t) = value*(1 + amplitude*sin(2 pi f t))
 * (structs/boundary_struct.h:184-192). */
typedef struct axs_bc_view {
    int n_groups;
    const uint8_t* type;           /* AXS_BC_*                                     */
    const uint8_t* boundary_type;  /* AXS_BOUNDARY_*, may be NULL (WALL)          */
    const double* value;           /* mean / prescribed value                     */
    const double* length;          /* group length, fully-developed profile       */
    const double* amplitude;       /* pulsatile only, may be NULL                 */
    const double* frequency;       /* pulsatile only (Hz), may be NULL            */
    /* Michaelis-Menten / Hill wall kinetics (MichaelisMentenParams, HillParams,
     * structs/boundary_struct.h:112-113; createBoundaryFieldHost,
     * manager/source/memory_manger.cpp:51-152): per group Vmax, Km, the Hill
     * exponent n (1 for Michaelis-Menten), the substrate-inhibition m, K2, V2
     * (V2 = 0: no inhibition) and the wall layers' total resistance sum d/k.
     * Each may be NULL (defaults 0, except n = m = 1). */
    const double* vmax;
    const double* km;
    const double* hill_n;
    const double* inhib_m;
    const double* inhib_k2;
    const double* inhib_v2;
    const double* wall_resistance;
} axs_bc_view;

typedef struct axs_residual_config {
    int enabled;
    int norm;       /* AXS_NORM_*    */
    int scaling;    /* AXS_SCALE_*   */
    double tol;
} axs_residual_config;

typedef struct axs_config {
    /* ConfigSolver (structs/solver_struct.h:203-231) */
    int convection_scheme;
    int gradient_scheme;
    int linear_solver;
    int time_scheme;
    int linear_max_iter;
    int use_multigrid;
    int add_convection;
    int transient;
    int use_non_orth;
    double dt;
    double t_end;
    /* ConfigSimple (solver_struct.h:158-163) */
    int max_iter;
    int check_conv;
    /* ConfigMultigrid (solver_struct.h:167-190) */
    int mg_max_iter;
    double mg_weight;
    int mg_coarse_sweeps;
    int mg_pre_sweeps;
    int mg_post_sweeps;
    /* VariablesSimple relaxation (solver_struct.h:312-313) and FluidPropertyConfig */
    double alpha_u;
    double alpha_p;
    double rho;
    double mu;
    /* ConfigResidual for U, V, Continuity, Temperature, Concentration
     * (solver_struct.h:240-263; initConfigResiduals, solver.cpp:87-101) */
    axs_residual_config residual[AXS_RES_COUNT];
    /* mesh type: a structured mesh forces the non-orthogonal corrector off
     * (Solver::runCheck, solver/source/solver.cpp:553-555) */
    int structured_mesh;
    /* ---- build extensions (no reference counterpart) ---- */
    int mg_max_levels;    /* 0 = reference rule (coarsen to <= 64 cells)        */
    int mg_smoother;      /* AXS_MG_SMOOTH_*                                    */
    int precision;        /* AXS_PREC_*                                         */
    int coarse_direct;    /* K >= 1: FP64 inverse on the coarsest level (at most
                           * 2048 cells), formed on every K-th outer iteration
                           * and applied in between; 0 = the coarsest sweeps */
    /* Transient frames: ConfigSolver::saveKeyFrameIter (structs/solver_struct.h:230,
     * clamped to >= 1 by runCheck, solver/source/solver.cpp:557-559) and
     * Solver::maxTimeFrames (solver/header/solver.h:108) */
    int save_key_frame_iter;
    int max_time_frames;
    /* Scalar transport: SolverFieldOption (structs/solver_struct.h:9-12) and
     * FluidPropertyConfig cp, k, D (src/header/setting.cuh:61-71).  Energy
     * diffuses with k / (rho cp), concentration with D; both convect with the
     * volumetric flux (solver/source/solver.cpp:976-1009). */
    int solve_energy;
    int solve_concentration;
    double cp;
    double k;
    double diffusivity;
    /* ---- build extension ---- */
    int mixed_refine;     /* AXS_REFINE_* (precision = AXS_PREC_MIXED)          */
} axs_config;

/* Reference defaults (structs/solver_struct.h, src/header/setting.cuh:61-71,
 * solver/source/solver.cpp:87-101). */
void axs_config_defaults(axs_config* cfg);

typedef struct axs_solver axs_solver;

/* Progress callback, called after every residual check on the solving thread
 * (the feed of ResidualPlot::add and the console line, solver.cpp:1245-1266).
 * values[i] is the normalised residual of names[i]. */
typedef void (*axs_progress_cb)(void* user, int iteration, int n, const char* const* names,
                                const double* values);

axs_status axs_create(axs_solver** out, int device);
void axs_destroy(axs_solver* s);
const char* axs_last_error(const axs_solver* s);

axs_status axs_set_mesh(axs_solver* s, const axs_mesh_view* mesh);
axs_status axs_set_bcs(axs_solver* s, int var, const axs_bc_view* bc);
axs_status axs_configure(axs_solver* s, const axs_config* cfg);

/* Blocking solve.  continue_run != 0 keeps fields and the iteration counter
 * when the mesh counts match (Continue Solver). */
axs_status axs_run(axs_solver* s, axs_progress_cb cb, void* user, int continue_run);
void axs_request_stop(axs_solver* s);
/* 1 if a run may continue on the current mesh, else 0 with a reason. */
int axs_can_continue(const axs_solver* s, char* reason, size_t reason_len);

/* Output fields (host copy).  n must equal the field's length. */
enum {
    AXS_FIELD_U = 0, AXS_FIELD_V = 1, AXS_FIELD_P = 2, AXS_FIELD_PP = 3,
    AXS_FIELD_GRAD_UZ = 4, AXS_FIELD_GRAD_UR = 5, AXS_FIELD_GRAD_VZ = 6, AXS_FIELD_GRAD_VR = 7,
    AXS_FIELD_GRAD_PZ = 8, AXS_FIELD_GRAD_PR = 9,
    AXS_FIELD_CONTINUITY = 10, AXS_FIELD_VMAG = 11, AXS_FIELD_CELL_RE = 12,
    AXS_FIELD_MDOT = 13,            /* length n_faces */
    AXS_FIELD_DU = 14, AXS_FIELD_DV = 15,
    /* internal state after the last iteration, for parity checks */
    AXS_FIELD_U_AC = 16, AXS_FIELD_U_B = 17, AXS_FIELD_U_INVAC = 18, AXS_FIELD_U_AF = 19,
    AXS_FIELD_V_AC = 20, AXS_FIELD_V_B = 21, AXS_FIELD_V_INVAC = 22, AXS_FIELD_V_AF = 23,
    AXS_FIELD_PP_AC = 24, AXS_FIELD_PP_B = 25, AXS_FIELD_PP_INVAC = 26, AXS_FIELD_PP_AF = 27,
    /* transient old levels n and n-1 (VariablesSimple uOld/vOld/uOld2/vOld2,
     * structs/solver_struct.h:265-321) */
    AXS_FIELD_U_OLD = 28, AXS_FIELD_V_OLD = 29, AXS_FIELD_U_OLD2 = 30, AXS_FIELD_V_OLD2 = 31,
    /* scalar transport (VariablesSimple temp / conc and their gradients) */
    AXS_FIELD_T = 32, AXS_FIELD_C = 33, AXS_FIELD_GRAD_TZ = 34, AXS_FIELD_GRAD_TR = 35,
    AXS_FIELD_GRAD_CZ = 36, AXS_FIELD_GRAD_CR = 37,
    /* precomputed geometry (preComputeVariables, memory_manger.cpp:512-617) */
    AXS_GEOM_INV_DPN = 40, AXS_GEOM_WP = 41, AXS_GEOM_DPB = 42, AXS_GEOM_INV_A2D = 43,
    AXS_GEOM_LSQ_WZ = 44, AXS_GEOM_LSQ_WR = 45, AXS_GEOM_LSQ_INV_ZZ = 46,
    AXS_GEOM_LSQ_INV_ZR = 47, AXS_GEOM_LSQ_INV_RR = 48, AXS_GEOM_LENGTH2D = 49,
    /* scalar systems and old levels after the last iteration, for parity checks */
    AXS_FIELD_T_AC = 50, AXS_FIELD_T_B = 51, AXS_FIELD_T_INVAC = 52, AXS_FIELD_T_AF = 53,
    AXS_FIELD_C_AC = 54, AXS_FIELD_C_B = 55, AXS_FIELD_C_INVAC = 56, AXS_FIELD_C_AF = 57,
    AXS_FIELD_T_OLD = 58, AXS_FIELD_C_OLD = 59, AXS_FIELD_T_OLD2 = 60, AXS_FIELD_C_OLD2 = 61,
    /* reactive walls, length n_faces (FVFaceDevice cw / ocrWall): the wall
     * concentration of the last Newton solve and area * rate there */
    AXS_FIELD_WALL_CW = 62, AXS_FIELD_WALL_OCR = 63
};
axs_status axs_get_field(axs_solver* s, int field, double* out, size_t n);
/* Number of elements field `field` holds (0 if unknown). */
size_t axs_field_length(const axs_solver* s, int field);

/* Normalised residual history of one residual (AXS_RES_*): one entry per
 * residual check.  Returns the number of entries written through *count. */
axs_status axs_get_residual_history(axs_solver* s, int res, double* out, int max_n, int* count);

/* Transient frames (Solver::TimeFrame / timeFrames, solver/header/solver.h:96-108;
 * captureTimeFrame, solver/source/solver.cpp:265-296): a transient run captures
 * the state it starts from, then every save_key_frame_iter steps and the last
 * (or stopped) step, at most max_time_frames; a continued transient run appends
 * and carries time and the pulsatile phase on from the last frame
 * (solver.cpp:890-900, 1036-1056).  axs_get_frame gives frame `index`'s time
 * (may be NULL) and copies its field AXS_FIELD_U, V or P into out (NULL: time
 * only; n = cell count). */
int axs_frame_count(const axs_solver* s);
/* Frames also hold AXS_FIELD_T / AXS_FIELD_C when those are solved. */
axs_status axs_get_frame(axs_solver* s, int index, int field, double* time, double* out, size_t n);

/* Batched runs (build extension; the reference solves one case at a time,
 * solver.cpp:1243-1283): the union mesh of axs_mesh_concat as `count` cases,
 * case c = cells cell_start[c] .. cell_start[c+1]-1 and faces face_start[c] ..
 * face_start[c+1]-1 (each case's faces connect only its cells).  Every case then
 * keeps its own residual norms (summed over blocks from its first cell, its own
 * cell count in the scalings), continuity scale and convergence test, and stops
 * on them: a case done keeps its fields while the others iterate, and the run
 * ends when every case is done (or at max_iter), per time step in a transient
 * run.  Without multigrid a case then runs exactly as it runs alone.  count 0:
 * one case (the default); set after axs_set_mesh. */
axs_status axs_set_cases(axs_solver* s, int count, const int* cell_start, const int* face_start);
typedef struct axs_case_stats {
    int iterations;                       /* outer iterations the case ran (all steps) */
    int converged;                        /* in its last time step                   */
    double last_residual[AXS_RES_COUNT];  /* at its last residual check              */
    int stop_iteration;                   /* the run's iteration count (axs_stats
                                             iterations) when it last stopped    */
} axs_case_stats;
axs_status axs_get_case_stats(axs_solver* s, int c, axs_case_stats* out);

/* Reactive-wall consumption diagnostic of the last run's concentration field
 * (wallConsumptionDiagnostic, solver/source/solver_util.cu:480-531; reported by
 * solver.cpp:1338-1385): out[0] total wall OCR, [1] substrate carried in through
 * Dirichlet inflow faces, [2] mass-transfer ceiling sum A h c_P, [3] reactive
 * face count, then sums over the reactive faces of [4] c_w, [5] d_PF, [6] h,
 * [7] c_P, [8] Vmax, [9] Km.  Summed in face order. */
axs_status axs_get_wall_diagnostic(axs_solver* s, double out[10]);

typedef struct axs_stats {
    int iterations;        /* outer iterations done (currentIteration)           */
    int converged;
    int stopped;
    int n_cells, n_faces, n_face_refs;
    int n_colors;          /* greedy colours, 0 if not built                     */
    int mg_levels;
    int mg_level_cells[32];
    double setup_ms;       /* upload + hierarchy + graph capture                 */
    double solve_ms;       /* outer loop wall time                               */
    double last_residual[AXS_RES_COUNT];
    double bytes_per_iteration;  /* algorithmic bytes of one outer iteration       */
    double vcycle_bytes;         /* algorithmic bytes of one p' multigrid solve    */
    int sweeps_per_launch;       /* momentum Jacobi sweeps fused per launch        */
    int sweep_block_cells;       /* cells per fused block (0: not fused)          */
    int graph_nodes;             /* nodes of the outer-iteration graph (launches
                                    and copies per iteration; 0 before a run)     */
} axs_stats;
axs_status axs_get_stats(axs_solver* s, axs_stats* out);

/* Run exactly `n` outer iterations from the current state without touching the
 * convergence bookkeeping flags, timing them on the GPU with HIP events (one
 * graph launch per iteration).  For benchmarks; *ms receives the elapsed time. */
axs_status axs_run_timed(axs_solver* s, int n, double* ms);

/* Time one kernel of the outer iteration: `repeats` eager launches on the
 * solver's stream between two HIP events, from the current state (run the
 * solver first).  *ms = average launch duration, *bytes = algorithmic bytes of
 * one launch.  No reference counterpart: a benchmark hook. */
enum {
    AXS_KERNEL_JACOBI_UV = 0,   /* one u+v momentum Jacobi sweep             */
    AXS_KERNEL_MOMENTUM = 1,    /* fused momentum assembly                   */
    AXS_KERNEL_MG_SMOOTH = 2,   /* one weighted-Jacobi sweep on the p' level 0 */
    AXS_KERNEL_FACE_FLUX = 3,   /* Rhie-Chow face mass flux                  */
    AXS_KERNEL_PP_COEFF = 4,    /* pressure-correction matrix                */
    AXS_KERNEL_JACOBI_UV_FUSED = 5  /* the fused launch of several u+v sweeps  */
};
axs_status axs_bench_kernel(axs_solver* s, int kernel, int repeats, double* ms, double* bytes);
/* The same kernel timed with cold caches: before each launch a kernel reads
 * 512 MiB (more than the L2s and the last-level cache hold); *ms = the time
 * `repeats` (flush, launch) pairs take beyond `repeats` flushes, divided by
 * `repeats`.  Allocates the 512 MiB for the call. */
axs_status axs_bench_kernel_cold(axs_solver* s, int kernel, int repeats, double* ms, double* bytes);
/* Autotuner (build extension): choose the workgroup size (64, 128 or 256
 * threads) of the momentum sweep and of each multigrid level's per-cell kernels
 * by timing them on the solver's current state; call after a run, the next run
 * uses the choice.  *ms_default / *ms_tuned (may be NULL): summed kernel times
 * of one pass at 256 threads and at the chosen sizes. */
axs_status axs_autotune(axs_solver* s, int repeats, double* ms_default, double* ms_tuned);
/* Momentum Jacobi sweeps per launch (build extension): `sweeps` (1 to 5)
 * sweeps of u and v run in one launch, one workgroup per block of
 * `block_cells` (512 or 1024; 0: 1024 where the halos fit, else 512)
 * consecutive cells with the intermediate iterates in LDS; the results are
 * bit-identical to one sweep per launch.  The next run uses the largest count
 * <= sweeps that the mesh's halos allow (axs_stats::sweeps_per_launch,
 * sweep_block_cells).  Default 5 sweeps, block size 0; axs_autotune also picks
 * them.  sweeps = 1 also runs the multigrid's level 0 one pass per launch
 * where its cycles would otherwise be fused (the same results). */
axs_status axs_set_sweep_fusion(axs_solver* s, int sweeps, int block_cells);

/* ---------------- Agglomeration multigrid, stand-alone ---------------- */
typedef struct axs_mg axs_mg;
/* Build the hierarchy of the cell graph (buildLevels, multigrid.cu:298-325).
 * max_levels 0 = reference rule. */
axs_status axs_mg_create(axs_mg** out, int n_cells, const int* face_start,
                         const int* face_neighbor, int max_levels);
void axs_mg_destroy(axs_mg* mg);
int axs_mg_levels(const axs_mg* mg);
axs_status axs_mg_level_info(const axs_mg* mg, int level, int* n_cells, int* n_refs);
/* Where the V-cycle switches to its single-workgroup, LDS-resident tail for the
 * given precision: *first_level = -1 when no tail fits; *lds_bytes = its LDS
 * image.  Build extension: the reference launches every level separately
 * (multigrid.cu:607-662) and smooths only the coarsest in one block (501-541). */
axs_status axs_mg_tail_info(const axs_mg* mg, int precision, int* first_level, int* lds_bytes);
/* Host copies of one level's maps; any pointer may be NULL.  cell_to_coarse and
 * fine_slot_to_coarse_slot are empty (untouched) on the coarsest level. */
axs_status axs_mg_get_maps(const axs_mg* mg, int level, int* face_start, int* face_neighbor,
                           int* cell_to_coarse, int* fine_slot_to_coarse_slot);
/* Solve A x = b with `cycles` V-cycles from the given x (host arrays, level-0
 * CSR order).  AC, AF, b, x as in Coefficients (solver_struct.h:103-120).
 * smoother: AXS_MG_SMOOTH_* (the reference smooths with weighted Jacobi).
 * coarse_direct = 1: the coarsest level solved with its FP64 inverse (formed
 * once per call; at most 2048 cells, else AXS_ERR_UNSUPPORTED) instead of
 * coarse_sweeps, in either precision. */
axs_status axs_mg_solve(axs_mg* mg, const double* AC, const double* AF, const double* b,
                        double* x, int cycles, double weight, int pre_sweeps, int post_sweeps,
                        int coarse_sweeps, int precision, int coarse_direct, int smoother);
/* The solve of axs_mg_solve (FP64, weighted-Jacobi smoothing, no coarsest
 * direct solve) without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  AC, AF, b
 * and x are DEVICE arrays of the current process in the level-0 CSR order of
 * `mg` (x in: start, out: answer).  The V-cycles' work arrays live in one
 * static device array of AXS_MG_NOALLOC_SCRATCH_BYTES per GPU, cleared at the
 * start of every call; a hierarchy that needs more (axs_mg_create on more than
 * about 1.4M triangle cells) is rejected with AXS_ERR_RANGE.  The solve is
 * enqueued on `stream` (a hipStream_t; NULL: the hierarchy's own stream) and
 * the call returns without waiting.  Calls that overlap on one GPU share that
 * array and race.  Results are bit-identical to axs_mg_solve's. */
#define AXS_MG_NOALLOC_SCRATCH_BYTES (64u << 20)
axs_status axs_mg_solve_noalloc(axs_mg* mg, const double* AC, const double* AF, const double* b, double* x,
                                int cycles, double weight, int pre_sweeps, int post_sweeps, int coarse_sweeps,
                                void* stream);
/* Mixed-precision options of the solves below (build extension): refine =
 * AXS_REFINE_* (default AXS_REFINE_CYCLE); fuse_level0 = 0 runs level 0 as
 * one-pass launches where the fused level-0 cycles would apply (default 1;
 * the results are the same bit for bit). */
axs_status axs_mg_set_mixed(axs_mg* mg, int refine, int fuse_level0);
/* Benchmark form: the system is uploaded once (x = 0), one solve of `cycles`
 * V-cycles (coarse operators included) is captured in a HIP graph, as the
 * solver's iteration graph runs it, and `repeats` replays are timed with HIP
 * events; each solve continues from the x the previous one left (the work does
 * not depend on x).  *ms = average time of one solve; *bytes = algorithmic
 * bytes of one solve. */
axs_status axs_mg_bench(axs_mg* mg, const double* AC, const double* AF, const double* b,
                        int cycles, double weight, int pre_sweeps, int post_sweeps,
                        int coarse_sweeps, int precision, int coarse_direct, int smoother,
                        int repeats, double* ms, double* bytes, double* x_out);
/* The same with cold caches: a kernel streams 512 MiB (beyond the L2s and the
 * 256 MB last-level cache) before each replay; *ms = the replays' time beyond
 * that of the flushes alone, per solve. */
axs_status axs_mg_bench_cold(axs_mg* mg, const double* AC, const double* AF, const double* b,
                             int cycles, double weight, int pre_sweeps, int post_sweeps,
                             int coarse_sweeps, int precision, int coarse_direct, int smoother,
                             int repeats, double* ms, double* bytes);

/* ---------------- Learned initial guess (build extension) ----------------
 * No counterpart in the reference.  A message-passing network on a cell graph
 * (Jacobi-scaled operator weights, hidden width 64, BF16 weights, FP32
 * accumulate on the matrix cores) maps the right-hand side b of A x = b to a
 * guess of x.  It is trained on the GPU by its own forward and backward kernels
 * on (system, solution) pairs, and in the solver it replaces p' = 0 before the
 * first pressure-correction solve of every outer iteration, run on one coarse
 * multigrid level and prolonged to the mesh.  The guess is alpha x with
 * alpha = x.b / x.Ax, the multiple of the network's direction x closest to the
 * solution in the energy norm, so it never starts the solve further from it
 * than zero does; the multigrid cycles then run as configured, and the answer
 * is still the multigrid's.  Parameters are FP32 in
 * the order W_1 (32 x 64), W_2 .. W_L (128 x 64, row-major, input-major), w_out (64). */
typedef struct axs_net axs_net;
axs_status axs_net_create(axs_net** out, int layers, uint64_t seed, int device);
void axs_net_destroy(axs_net* net);
int axs_net_num_params(const axs_net* net);
axs_status axs_net_get_params(const axs_net* net, float* out);
axs_status axs_net_set_params(axs_net* net, const float* in);
/* A training pair, stored on the device: CSR graph (face_neighbor -1 on boundary
 * slots), operator AC/AF, right-hand side b and the solution x of A x = b.
 * *index (may be NULL) receives its sample number. */
axs_status axs_net_add_sample(axs_net* net, int n_cells, const int* face_start, const int* face_neighbor,
                              const double* AC, const double* AF, const double* b, const double* x,
                              int* index);
axs_status axs_net_clear_samples(axs_net* net);
/* Loss of one sample, 1 - (x.b)^2 / (x.Ax * b.x*): the share of the solution's
 * energy |x*|_A^2 = b.x* the guess alpha x leaves uncaptured; and its gradient
 * (num_params FP32, may be NULL). */
axs_status axs_net_loss_grad(axs_net* net, int sample, double* loss, float* grad);
/* Adam over all samples, `batch` per step; loss_per_epoch (may be NULL) gets
 * the mean sample loss of each epoch. */
axs_status axs_net_train(axs_net* net, int epochs, int batch, double lr, uint64_t seed, double* loss_per_epoch);
/* The guess x for one system (host arrays). */
axs_status axs_net_predict(axs_net* net, int n_cells, const int* face_start, const int* face_neighbor,
                           const double* AC, const double* AF, const double* b, double* x);
/* Use `net` for the first p' solve of every outer iteration (NULL: off, the
 * default).  level: multigrid level it runs on, -1 = the first with at most
 * 16384 cells.  The net must outlive its use by the solver. */
axs_status axs_set_learned_guess(axs_solver* s, axs_net* net, int level);

/* Greedy colouring in cell order + colour-major cell order
 * (buildMeshColoring, memory_manger.cpp:334-409).  color_start needs
 * n_cells + 1 entries; *n_colors receives the colour count. */
axs_status axs_build_coloring(int n_cells, const int* face_start, const int* face_neighbor,
                              int* color_start, int* cell_order, int* n_colors);

/* ---------------- Case generator (input producer) ---------------- */
typedef struct axs_mesh axs_mesh;

/* One rectangular block of a conformal multiblock grid (makeRectBlock,
 * structs/multiblock.h:118-136).  grading 0 = uniform, 1 = geometric
 * progression with ratio coef.  edge_group: West, East, South, North. */
typedef struct axs_block_spec {
    double z0, r0;
    double axial_length;  int axial_cells;  int axial_grading;  double axial_coef;
    double radial_length; int radial_cells; int radial_grading; double radial_coef;
    int edge_group[4];
} axs_block_spec;
axs_status axs_mesh_multiblock(axs_mesh** out, int n_blocks, const axs_block_spec* blocks);

/* Triangles -> FVMesh (createUnstructuredMesh, mesh.cpp:1848-2002); boundary
 * faces take the group of the nearest tagged boundary segment within tol
 * (groupAtPoint, manager/source/math_func.cpp:175-195). */
axs_status axs_mesh_triangles(axs_mesh** out, int n_points, const double* point_z,
                              const double* point_r, int n_tris, const int* tris,
                              int n_segments, const double* seg_z0, const double* seg_r0,
                              const double* seg_z1, const double* seg_r1, const int* seg_group,
                              double tol);

/* Seeded Poisson-disk + jitter + Delaunay mesh of an axisymmetric nozzle:
 * inlet pipe (radius r_in, length l_in), cone of half-angle cone_deg down to
 * r_throat, throat of length l_throat, diffuser of half-angle diff_deg back to
 * r_out, outlet pipe of length l_out.  spacing = Poisson-disk radius,
 * jitter = fraction of spacing.  Groups: 0 inlet, 1 outlet, 2 wall, 3 axis. */
typedef struct axs_nozzle_spec {
    double r_in, l_in, cone_deg, r_throat, l_throat, diff_deg, r_out, l_out;
    double spacing, jitter;
    uint64_t seed_points, seed_jitter;
} axs_nozzle_spec;
axs_status axs_mesh_nozzle(axs_mesh** out, const axs_nozzle_spec* spec);

/* Batched runs (build extension; the reference solves one case at a time):
 * the disjoint union of `count` meshes, each case's cells, faces and slots in
 * order after the previous ones and its boundary group g renumbered
 * g + c * groups_per_mesh, so one solver (one launch per kernel) advances every
 * case, each exactly as it would alone apart from the shared multigrid
 * hierarchy's coarsest levels and the pooled residual norms. */
axs_status axs_mesh_concat(axs_mesh** out, int count, const axs_mesh_view* meshes, int groups_per_mesh);
axs_status axs_mesh_get_view(const axs_mesh* m, axs_mesh_view* view);
/* Triangle mesh only: point and triangle arrays (NULL to query counts). */
axs_status axs_mesh_get_triangles(const axs_mesh* m, int* n_points, double* point_z,
                                  double* point_r, int* n_tris, int* tris);
/* Multiblock mesh only: each cell's corner quad in cell order, 4 (z, r) per cell
 * (Mesh::multiBlockCellCorners, solver/source/mesh.cpp:2514-2537); *n_cells = 0
 * for a mesh without them.  NULL arrays query the count. */
axs_status axs_mesh_get_quads(const axs_mesh* m, int* n_cells, double* z4, double* r4);
void axs_mesh_destroy(axs_mesh* m);

/* ---------------- Unstructured mesher (aximesh) ---------------- */
/* AxiMesh::Params (aximesh/include/aximesh/aximesh.h:56-61).  max_points is
 * not in the reference: refinement stops with AXS_ERR_INVALID when the mesh
 * would pass that many points (0: 2^25) instead of exhausting memory. */
typedef struct axs_cdt_params {
    double B;             /* skinny bound of the reference's disabled Ruppert pass; unused */
    double beta;          /* size-field grading: |h(v) - h(w)| <= beta |vw| */
    double classify_tol;  /* a triangle is accepted when sqrt(3) R <= classify_tol * mean h */
    long long max_points;
} axs_cdt_params;
/* The reference defaults: B 1.4, beta 0.2, classify_tol 1.4. */
void axs_cdt_params_defaults(axs_cdt_params* p);

typedef struct axs_cdt axs_cdt;
/* AxiMesh::generateMesh(points, segments, size, params)
 * (aximesh/source/aximesh.cpp:1105-1175; called by
 * Mesh::runAxiMeshTriangulation, solver/source/mesh.cpp:1208-1251): the
 * constrained Delaunay triangulation of n_points points with the n_segments
 * segments segments[2k] - segments[2k+1] (caller indices), the outside removed
 * by a parity flood fill across the segments, refined by Rebay's frontal
 * method to the size field of the segment lengths.  The result is the
 * reference's, index for index and bit for bit: points in bin order followed
 * by the refinement points, triangles (counter-clockwise, adj[i] across edge
 * v[i] -> v[i+1], -1 on the boundary), segments renumbered.  Failures the
 * reference throws (a point search or size field that does not settle:
 * "point search has stalled") and inputs on which the reference's behaviour
 * is undefined (no area, a segment from a point to itself, non-finite
 * coordinates) return AXS_ERR_INVALID, indices out of range AXS_ERR_RANGE;
 * the text is in axs_last_error(NULL).  params NULL: the defaults. */
axs_status axs_cdt_generate(axs_cdt** out, int n_points, const double* x, const double* y,
                            int n_segments, const int* segments, const axs_cdt_params* params);
/* AxiMesh::insertPoints(points, size) (aximesh.cpp:466-489): the super
 * triangle (-100,-100) (100,-100) (0,100), appended as points n..n+2, then
 * each point inserted in the given order with no scaling or sorting; 2n + 1
 * triangles. */
axs_status axs_cdt_insert_points(axs_cdt** out, int n_points, const double* x, const double* y);
/* Sizes of a result (any pointer may be NULL). */
axs_status axs_cdt_counts(const axs_cdt* m, int* n_points, int* n_triangles, int* n_segments);
/* Copies of a result: x, y n_points each; tri_v, tri_adj 3 per triangle;
 * segments 2 per segment.  NULL arrays are skipped. */
axs_status axs_cdt_get(const axs_cdt* m, double* x, double* y, int* tri_v, int* tri_adj, int* segments);
void axs_cdt_destroy(axs_cdt* m);

/* ---------------- Output formats (manager/source/file_manager.cpp) ---------------- */
/* The cell polygons a solution export writes beside the fields: corners 4 (quads
 * in cell order, welded into one point list as the exporter does), 3 (triangles
 * indexing points, cell c = triangle c) or 0 (none). */
typedef struct axs_cell_polygons {
    int corners;
    const double* quad_z;      /* corners == 4: 4 per cell                       */
    const double* quad_r;
    int n_points;              /* corners == 3                                    */
    const double* point_z;
    const double* point_r;
    const int* tris;           /* corners == 3: 3 per cell                        */
    int structured;            /* meta "meshType": 1 Structured, 0 Unstructured   */
    int multiblock;            /* meta "multiBlock"                               */
} axs_cell_polygons;
/* Mesh quality metrics (Quality::buildQuality, solver/source/quality.cpp:160-166),
 * per cell, computed on GPU `device`: aspect_ratio = max/min over the cell's
 * faces of |(x_f - x_c).n_f| (0 when the smallest is not above 1e-30,
 * quality.cpp:129-158); non_orthogonality = the smallest of
 * |(x_f - x_c).n_f|/|x_f - x_c| over the faces and |(x_N - x_P).n_f|/|x_N - x_P|
 * over the interior ones, from 1 (quality.cpp:61-105); skewness from the corner
 * ring's interior angles (quality.cpp:7-59; 0 without corners).  `poly` gives
 * the corner rings (axs_cell_polygons, corners 4 or 3; NULL or 0: none).  Each
 * output holds n_cells values.  Allocates device memory for the call. */
axs_status axs_mesh_quality(const axs_mesh_view* mesh, const axs_cell_polygons* poly, int device,
                            double* aspect_ratio, double* non_orthogonality, double* skewness);
/* saveSolutionNpy (file_manager.cpp:2074-2309): into directory `dir` (created),
 * solution.npy (cells x [z, r, volume, area2D, then the solved fields in the
 * results panel's order]), points.npy + cells.npy when `poly` has corners,
 * frame_NNNN.npy per transient frame (the same field columns; only the frame's
 * primary fields are non-zero) and meta.json; float64 / int32, little-endian,
 * NPY format 1.0. */
axs_status axs_save_solution_npy(axs_solver* s, const axs_cell_polygons* poly, const char* dir);

/* The solver settings file (.axislv solver payload, format version 5, and the
 * version-4 payload loadFromPathSolver still reads (readSolverPayloadLegacy,
 * file_manager.cpp:321-373);
 * saveFromPathSolver / loadFromPathSolver, file_manager.cpp:1482-1549): magic
 * "AXVS", version, VariableUnits, SolverFieldOption, ConfigSolver, the fluid
 * properties, ConfigSimple, then type / norm / scaling / enabled / tolerance of
 * the residuals U, V, Continuity, Temperature, Concentration.  What the file
 * holds beyond axs_config: */
typedef struct axs_solver_file {
    uint8_t var_units[15];            /* VariableUnits display-unit indices        */
    int velocity_solver;              /* VelocitySolverType (0: SIMPLE)            */
    double mom_tol, pp_tol;           /* ConfigSimple tolerances                   */
    int residual_type[AXS_RES_COUNT]; /* ResidualType per residual                 */
} axs_solver_file;
/* Write cfg's settings (and `extra`, NULL: defaults) after the reference's
 * sanitizeSolverConfig clamps; load reads a version-5 file into cfg (its other
 * members untouched) and `extra` (may be NULL), clamped the same way.  Files of
 * another magic or version: AXS_ERR_UNSUPPORTED; short files: AXS_ERR_INVALID. */
axs_status axs_save_solver_file(const char* path, const axs_config* cfg, const axs_solver_file* extra);
axs_status axs_load_solver_file(const char* path, axs_config* cfg, axs_solver_file* extra);

#ifdef __cplusplus
}
#endif
#endif /* AXS_H */
