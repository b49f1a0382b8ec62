// compat/solver.hpp -- C++ drop-in for AxiSim's `class Solver`
// (solver/header/solver.h:18-185) on top of the libaxs.so C ABI.
//
// The GUI, mesher and Results talk to the reference solver through this class:
// they write the config members directly, call run(mesh) / requestStop() /
// canContinue(), and read `solutions` once `solverRunning` drops.  This shim
// keeps those names, defaults and threading (one worker std::thread per run,
// atomics for the running / stop flags), and forwards the solve to axs_run().
//
// The types mirror the reference's (structs/solver_struct.h:9-263,
// structs/boundary_struct.h:88-193, src/header/setting.cuh:61-71) field by
// field so existing call sites compile unchanged; only what the SIMPLE path
// reads is carried.
#pragma once

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "axs.h"

// The GUI's console and residual plot (graphics/gui/header/console.h,
// residual_plot.h).  The solver only holds pointers to them; it reaches their
// methods through the hooks in gui_hooks.hpp, which the GUI's build compiles
// against its own headers (compat/gui_hooks.cpp).
class Console;
class ResidualPlot;

namespace axs::compat {

enum class ResidualType : uint8_t { RESIDUAL_SCALED = 0, RESIDUAL_RAW = 1, RESIDUAL_RMS = 2 };
enum class ResidualNormType : uint8_t { RESIDUAL_L1 = 0, RESIDUAL_L2 = 1, RESIDUAL_LINF = 2 };
enum class ResidualScalingType : uint8_t {
    RESIDUAL_SCALING_NONE = 0, RESIDUAL_SCALING_N = 1, RESIDUAL_SCALING_SQRT_N = 2,
    RESIDUAL_SCALING_DIAGONAL = 3, RESIDUAL_SCALING_CONTINUITY = 4
};
enum class LinearSolverType : uint8_t { LINEAR_JACOBI = 0, LINEAR_GS_RB = 1 };
enum class VelocitySolverType : uint8_t { SOLVER_SIMPLE = 0 };
enum class ConvectionScheme : uint8_t { CONV_UPWIND, CONV_CENTRAL, CONV_SECOND_ORDER_UPWIND, CONV_QUICK };
enum class GradientScheme : uint8_t { GRAD_GREEN_GAUSS = 0, GRAD_LSQ = 1 };
enum class TimeScheme : uint8_t { TIME_FIRST_ORDER = 0, TIME_SECOND_ORDER = 1 };
enum class BCType : uint8_t { DIRICHLET, NEUMANN, FULLY_DEVELOPED, MICHAELIS_MENTEN, HILL, NONE, PULSATILE };
enum class BoundaryType { WALL, VELOCITY_INLET, PRESSURE_OUTLET, SYMMETRY, FAR_FIELD };
enum class BoundaryVariable : uint8_t { UVelocity, VVelocity, Pressure, StaticTemperature, Concentration, None };
enum class MeshType { Structured, Unstructured };

struct ConfigSimple {
    int maxIter = 50;
    int checkConv = 1;
    double momTol = 1e-8;
    double ppTol = 1e-5;
};

struct ConfigMultigrid {
    int maxIter = 3;
    double weight = 0.6;
    int linearSweep = 30;
    int linearPreSweep = 1;
    int linearPostSweep = 1;
    // build extensions (no reference counterpart)
    int maxLevels = 0;
    int precision = AXS_PREC_FP64;
    int coarseDirect = 0;  // K >= 1: FP64 coarsest inverse formed every K-th outer iteration
};

struct ConfigSolver {
    VelocitySolverType velocitySolver = VelocitySolverType::SOLVER_SIMPLE;
    ConvectionScheme convectionScheme = ConvectionScheme::CONV_UPWIND;
    GradientScheme gradientScheme = GradientScheme::GRAD_LSQ;
    LinearSolverType type = LinearSolverType::LINEAR_JACOBI;
    TimeScheme timeScheme = TimeScheme::TIME_FIRST_ORDER;
    int linearMaxIter = 20;
    bool useMultigrid = true;
    bool addConvectionTerm = true;
    bool transient = false;
    bool useNonOrthCorrector = false;
    double dt = 0.1;
    double tEnd = 2.0;
    int saveKeyFrameIter = 2;
};

// ConfigResidual (structs/solver_struct.h:240-263): *resVal / *scaleVal are the
// last check's normalised residual and scale, host memory the Solver owns (the
// residual plot reads *resVal); res / scale stay null (the per-cell arrays
// live on the device inside libaxs.so).
struct ConfigResidual {
    ResidualType type = ResidualType::RESIDUAL_SCALED;
    ResidualNormType normType = ResidualNormType::RESIDUAL_L1;
    ResidualScalingType scaleType = ResidualScalingType::RESIDUAL_SCALING_DIAGONAL;
    bool enabled = false;
    double* res = nullptr;
    double* scale = nullptr;
    double* resVal = nullptr;
    double* scaleVal = nullptr;
    double tol = 0.001;
};

struct FluidPropertyConfig {  // src/header/setting.cuh:61-71
    double rho = 998.0;
    double mu = 0.0010518;
    double cp = 4180.0;
    double k = 0.6;
    double D = 3.0277e-9;
};

struct GridConfig {  // src/header/setting.cuh:14-37
    int nr = 170;
    int nz = 100;
    double R = 0.0017;
    double L = 0.01;
    std::vector<double> dz, dr, r, z, rFace, zFace, Az, Ar, Vcell;
    double zBias = 1.0;
    double rBias = 1.0;
    int n_cell = 0;
    double A_tot = 0.0;
};

struct VariableUnits {  // manager/header/unit_manager.h:6-31 (display-unit indices)
    std::uint8_t LUnit = 0, RUnit = 0;
    std::uint8_t axialUnit = 0, radialUnit = 0, pressureUnit = 0, temperatureUnit = 0, concentrationUnit = 0;
    std::uint8_t rhoUnit = 0, muUnit = 0, DUnit = 0;
    std::uint8_t specificHeatUnit = 0, heatCondUnit = 0;
    std::uint8_t VmaxUnit = 0;
    std::uint8_t layerKUnit = 0, layerDUnit = 0;
};

// Config (structs/solver_struct.h:325-331): the project-owned configuration the
// Solver is constructed on and keeps references into.
struct Config {
    FluidPropertyConfig f;
    GridConfig g;
    VariableUnits varUnits;
};

struct VariablesSimple {  // only the relaxation factors are configuration
    double momentumRelaxation = 0.7;
    double pressureRelaxation = 0.3;
};

struct SolverFieldOption {
    bool solveEnergy = false;
    bool solveConcentration = false;
};

struct Vec2 {
    double z = 0.0;
    double r = 0.0;
};

struct FVCell {
    Vec2 center;
    double area2D = 0.0;
    double volume = 0.0;
    std::vector<int> faceIDs;
};

struct FVFace {
    int owner = -1;
    int neighbor = -1;
    Vec2 normal;
    Vec2 center;
    double length2D = 0.0;
    double area = 0.0;
    int boundaryGroupID = -1;
};

struct FVMesh {
    std::vector<FVCell> cells;
    std::vector<FVFace> faces;
    int numCells() const { return (int)cells.size(); }
    int numFaces() const { return (int)faces.size(); }
};

// One boundary condition (BoundaryCondition, boundary_struct.h:134-193):
// the type plus the scalars the SIMPLE path reads.
struct BoundaryCondition {
    BCType type = BCType::NONE;
    double value = 0.0;
    double amplitude = 0.1;  // pulsatile only
    double frequency = 1.0;  // pulsatile only (Hz)
    // Michaelis-Menten / Hill kinetics (MichaelisMentenParams, HillParams,
    // boundary_struct.h:112-113): n is the Hill exponent; the substrate
    // inhibition m, K2, V2 applies when `inhibition` is set
    double Vmax = 0.0, Km = 0.0, n = 1.0;
    bool inhibition = false;
    double m = 1.0, K2 = 0.0, V2 = 0.0;
};

// One wall layer (Layer, boundary_struct.h:24-41): conductivity k, thickness d.
struct Layer {
    double k = 0.0;
    double d = 0.0;
    double R = 0.0;
    double resistance() const { return (k != 0.0) ? d / k : 0.0; }
};

struct BoundarySegmentGroup {
    int id = -1;
    std::string name;
    BoundaryType type = BoundaryType::WALL;
    float totalLength = 0.0f;
    std::unordered_map<int, BoundaryCondition> bcs;  // keyed by (int)BoundaryVariable
    // wall layers per variable (temperature, concentration), a series resistance
    std::unordered_map<int, std::vector<Layer>> layers;  // keyed by (int)BoundaryVariable
};

// What Solver::run reads from Mesh (mesh.h:226-245).
struct Mesh {
    bool isReady = false;
    MeshType currentMeshType = MeshType::Structured;
    FVMesh fvMesh;
    std::vector<BoundarySegmentGroup> boundaryGroups;
    const FVMesh& getFVMesh() const { return fvMesh; }
    void refreshFVMesh() {}  // the mesher rebuilds its cached FVMesh here
};

struct SolutionField {  // structs/boundary_struct.h:537-541
    std::vector<double> field;
    std::vector<double> dr, dz;
    BoundaryVariable boundaryVariable = BoundaryVariable::None;
};

// getDefaultBCType + getEffectiveBC (manager/source/boundary_func.cpp:104-197)
BCType defaultBCType(BoundaryType t, BoundaryVariable v);
BoundaryCondition effectiveBC(const BoundarySegmentGroup& g, BoundaryVariable v);

class Solver {
public:
    explicit Solver(Config& config);  // solver/source/solver.cpp:104-114
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    // combo-box labels the settings panels index with the enum values
    // (solver/header/solver.h:23-35)
    const char* boundaryType[5] = {"Wall", "Velocity Inlet", "Pressure Outlet", "Symmetry", "Far Field"};
    const char* residualType[4] = {"Scaled", "Raw Residual", "RMS", "Custom Residual"};
    const char* residualNormType[3] = {"L1 Norm", "L2 Norm", "Linf Norm"};
    const char* residualScalingType[4] = {"None", "N", "sqrt(N)", "Diagonal"};
    const char* linearSolverType[2] = {"Jacobi", "Red Black Gauss Seidel"};
    const char* velocitySolverType[1] = {"SIMPLE"};
    const char* bcTypeNames[2] = {"Constant", "Flux"};
    const char* bcInletTypeNames[3]{"Constant", "Flux", "Fully Developed"};
    const char* convectionDiscretizationType[3] = {"First Order Upwind", "Central Difference", "Second Order Upwind"};
    const char* gradientSchemeType[2] = {"Green-Gauss", "Least Squares"};
    const char* timeSchemeType[2] = {"First Order Implicit", "Second Order Implicit"};
    const char* residualPlotType[5] = {"U", "V", "Continuity", "Temperature", "Concentration"};

    // the fields a run produces, in the results panel's order (addFieldType,
    // solver.cpp:333-362)
    std::vector<std::string> fieldType;

    bool continueSolver = false;
    bool isReady = false;
    std::atomic<bool> solverRunning{false};
    std::atomic<bool> stopRequested{false};

    // Starts the solve on a worker thread and returns (solver.cpp:146-173).
    void run(Mesh& mesh);
    void requestStop();
    bool canContinue(const Mesh& mesh, std::string* reason = nullptr) const;
    void shutdown();  // join the worker
    void reset();

    // configuration, written directly by the settings panels; config / g / f /
    // varUnits refer into the project's Config
    Config& config;
    GridConfig& g;
    FluidPropertyConfig& f;
    VariableUnits& varUnits;
    SolverFieldOption fieldOption;

    // GUI classes (set by the window, graphics/window/source/gui.cpp:307-313)
    Console* console = nullptr;
    ResidualPlot* residualPlot = nullptr;
    ConfigSimple configSimple;
    ConfigSolver configSolver;
    ConfigMultigrid configMultigrid;
    VariablesSimple simple;
    std::unordered_map<std::string, ConfigResidual> cfg;

    // results, read once solverRunning drops
    std::unordered_map<std::string, SolutionField> solutions;
    FVMesh fvMesh;

    // transient frames (solver/header/solver.h:96-108): the start state, every
    // saveKeyFrameIter steps and the last step, at most maxTimeFrames; a
    // continued transient run appends
    struct TimeFrame {
        double time = 0.0;
        std::unordered_map<std::string, std::vector<double>> fields;
    };
    std::vector<TimeFrame> timeFrames;
    int maxTimeFrames = 500;
    int currentIteration() const { return currentIteration_; }
    std::string lastMessage() const;

    // ResidualPlot::add / console feed: called on the worker thread after
    // every residual check with (iteration, {"U", "V", "Continuity"} values).
    std::function<void(int, const std::map<std::string, double>&)> onResidual;

private:
    void runSimple(Mesh mesh);
    bool runCheck(const Mesh& mesh);
    void addFieldType();
    void say(const std::string& line);
    void bindResidualStorage();
    static void progressThunk(void* user, int iteration, int n, const char* const* names, const double* values);

    // host storage behind cfg[...].resVal / scaleVal (five residuals)
    double resStorage_[5] = {}, scaleStorage_[5] = {};

    axs_solver* handle_ = nullptr;
    std::thread worker_;
    mutable std::mutex msgMutex_;
    std::string message_;
    int currentIteration_ = 0;
    bool structured_ = true;
};

}  // namespace axs::compat

// The GUI names these types unqualified (global in the reference's headers,
// which the drop-in replaces).
using axs::compat::BCType;
using axs::compat::BoundaryCondition;
using axs::compat::BoundarySegmentGroup;
using axs::compat::BoundaryType;
using axs::compat::BoundaryVariable;
using axs::compat::Config;
using axs::compat::ConfigMultigrid;
using axs::compat::ConfigResidual;
using axs::compat::ConfigSimple;
using axs::compat::ConfigSolver;
using axs::compat::ConvectionScheme;
using axs::compat::FluidPropertyConfig;
using axs::compat::FVCell;
using axs::compat::FVFace;
using axs::compat::FVMesh;
using axs::compat::GradientScheme;
using axs::compat::GridConfig;
using axs::compat::LinearSolverType;
using axs::compat::Mesh;
using axs::compat::MeshType;
using axs::compat::ResidualNormType;
using axs::compat::ResidualScalingType;
using axs::compat::ResidualType;
using axs::compat::SolutionField;
using axs::compat::Solver;
using axs::compat::SolverFieldOption;
using axs::compat::TimeScheme;
using axs::compat::VariableUnits;
using axs::compat::VariablesSimple;
