// compat/solver.cpp -- the Solver drop-in (see solver.hpp).
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

#include "gui_hooks.hpp"

namespace axs::compat {

namespace {

void initConfigResiduals(std::unordered_map<std::string, ConfigResidual>& cfg) {  // solver.cpp:87-101
    cfg.clear();
    for (const char* n : {"U", "V", "Continuity", "Temperature", "Concentration"}) cfg.emplace(n, ConfigResidual{});
    cfg.at("U").enabled = true;
    cfg.at("V").enabled = true;
    cfg.at("Continuity").enabled = true;
}

struct PackedMesh {  // packFVMeshForDevice, memory_manger.cpp:434-509
    std::vector<double> cz, cr, vol, a2d, nz, nr, fz, fr, area;
    std::vector<int> fs, ids, own, nb, grp;
    axs_mesh_view view() const {
        axs_mesh_view v{};
        v.n_cells = (int)cz.size();
        v.n_faces = (int)own.size();
        v.cell_center_z = cz.data();
        v.cell_center_r = cr.data();
        v.cell_volume = vol.data();
        v.cell_area2d = a2d.data();
        v.cell_face_start = fs.data();
        v.cell_face_ids = ids.data();
        v.face_owner = own.data();
        v.face_neighbor = nb.data();
        v.face_normal_z = nz.data();
        v.face_normal_r = nr.data();
        v.face_center_z = fz.data();
        v.face_center_r = fr.data();
        v.face_area = area.data();
        v.face_group = grp.data();
        return v;
    }
};

PackedMesh pack(const FVMesh& m) {
    PackedMesh p;
    for (const FVCell& c : m.cells) {
        p.cz.push_back(c.center.z);
        p.cr.push_back(c.center.r);
        p.vol.push_back(c.volume);
        p.a2d.push_back(c.area2D);
    }
    p.fs.push_back(0);
    for (const FVCell& c : m.cells) {
        for (int f : c.faceIDs) p.ids.push_back(f);
        p.fs.push_back((int)p.ids.size());
    }
    for (const FVFace& f : m.faces) {
        p.own.push_back(f.owner);
        p.nb.push_back(f.neighbor);
        p.nz.push_back(f.normal.z);
        p.nr.push_back(f.normal.r);
        p.fz.push_back(f.center.z);
        p.fr.push_back(f.center.r);
        p.area.push_back(f.area);
        p.grp.push_back(f.boundaryGroupID);
    }
    return p;
}

struct BCArrays {  // createBoundaryFieldHost, memory_manger.cpp:51-152
    std::vector<uint8_t> type, btype;
    std::vector<double> value, length, amp, freq, vmax, km, n, m, k2, v2, rtot;
    axs_bc_view view() const {
        axs_bc_view v{};
        v.n_groups = (int)type.size();
        v.type = type.data();
        v.boundary_type = btype.data();
        v.value = value.data();
        v.length = length.data();
        v.amplitude = amp.data();
        v.frequency = freq.data();
        v.vmax = vmax.data();
        v.km = km.data();
        v.hill_n = n.data();
        v.inhib_m = m.data();
        v.inhib_k2 = k2.data();
        v.inhib_v2 = v2.data();
        v.wall_resistance = rtot.data();
        return v;
    }
};

BCArrays bcArrays(const std::vector<BoundarySegmentGroup>& groups, BoundaryVariable var) {
    int maxId = -1;
    for (const auto& g : groups) maxId = std::max(maxId, g.id);
    BCArrays a;
    const int n = maxId + 1;
    a.type.assign(n, (uint8_t)BCType::NONE);
    a.btype.assign(n, (uint8_t)BoundaryType::WALL);
    a.value.assign(n, 0.0);
    a.length.assign(n, 0.0);
    a.amp.assign(n, 0.0);
    a.freq.assign(n, 0.0);
    a.vmax.assign(n, 0.0);
    a.km.assign(n, 0.0);
    a.n.assign(n, 1.0);
    a.m.assign(n, 1.0);
    a.k2.assign(n, 0.0);
    a.v2.assign(n, 0.0);
    a.rtot.assign(n, 0.0);
    for (const auto& g : groups) {
        if (g.id < 0) continue;
        const BoundaryCondition bc = effectiveBC(g, var);
        a.type[g.id] = (uint8_t)bc.type;
        a.btype[g.id] = (uint8_t)g.type;
        a.value[g.id] = bc.value;
        a.length[g.id] = g.totalLength;
        if (bc.type == BCType::PULSATILE) {
            a.amp[g.id] = bc.amplitude;
            a.freq[g.id] = bc.frequency;
        }
        if (bc.type == BCType::MICHAELIS_MENTEN || bc.type == BCType::HILL) {
            a.vmax[g.id] = bc.Vmax;
            a.km[g.id] = bc.Km;
            if (bc.type == BCType::HILL) a.n[g.id] = bc.n;
            if (bc.inhibition) {
                a.m[g.id] = bc.m;
                a.k2[g.id] = bc.K2;
                a.v2[g.id] = bc.V2;
            }
        }
        auto layers = g.layers.find((int)var);
        if (layers != g.layers.end()) {
            double r = 0.0;
            for (const Layer& l : layers->second) r += l.resistance();
            a.rtot[g.id] = r;
        }
    }
    return a;
}

}  // namespace

BCType defaultBCType(BoundaryType t, BoundaryVariable v) {
    const bool vel = v == BoundaryVariable::UVelocity || v == BoundaryVariable::VVelocity;
    const bool scalar = v == BoundaryVariable::StaticTemperature || v == BoundaryVariable::Concentration;
    switch (t) {
    case BoundaryType::VELOCITY_INLET:
        if (vel || scalar) return BCType::DIRICHLET;
        return v == BoundaryVariable::Pressure ? BCType::NEUMANN : BCType::NONE;
    case BoundaryType::PRESSURE_OUTLET:
        if (v == BoundaryVariable::Pressure) return BCType::DIRICHLET;
        return (vel || scalar) ? BCType::NEUMANN : BCType::NONE;
    case BoundaryType::WALL:
    case BoundaryType::FAR_FIELD:
        if (vel) return BCType::DIRICHLET;
        return (v == BoundaryVariable::Pressure || scalar) ? BCType::NEUMANN : BCType::NONE;
    case BoundaryType::SYMMETRY:
        if (v == BoundaryVariable::VVelocity) return BCType::DIRICHLET;
        return (v == BoundaryVariable::UVelocity || v == BoundaryVariable::Pressure || scalar) ? BCType::NEUMANN
                                                                                                 : BCType::NONE;
    }
    return BCType::NONE;
}

BoundaryCondition effectiveBC(const BoundarySegmentGroup& g, BoundaryVariable v) {
    auto it = g.bcs.find((int)v);
    if (it != g.bcs.end()) return it->second;
    BoundaryCondition bc;
    bc.type = defaultBCType(g.type, v);
    bc.value = 0.0;
    return bc;
}

Solver::Solver(Config& cfgRef) : config(cfgRef), g(cfgRef.g), f(cfgRef.f), varUnits(cfgRef.varUnits) {
    initConfigResiduals(cfg);
    bindResidualStorage();
    if (axs_create(&handle_, 0) != AXS_OK) {
        message_ = std::string("no GPU: ") + axs_last_error(nullptr);
        handle_ = nullptr;
    }
}

// resVal / scaleVal of every residual point at this Solver's host storage
// (allocateResiduals gives each its pinned pair, memory_manger.cpp:270-285)
void Solver::bindResidualStorage() {
    const char* names[5] = {"U", "V", "Continuity", "Temperature", "Concentration"};
    for (int i = 0; i < 5; i++) {
        ConfigResidual& r = cfg.at(names[i]);
        resStorage_[i] = scaleStorage_[i] = 0.0;
        r.resVal = &resStorage_[i];
        r.scaleVal = &scaleStorage_[i];
    }
}

void Solver::say(const std::string& line) {
    {
        std::lock_guard<std::mutex> lock(msgMutex_);
        message_ = line;
    }
    consoleAddLine(console, line);
}

// addFieldType, solver.cpp:333-362
void Solver::addFieldType() {
    fieldType.clear();
    for (const char* n : {"Axial Velocity", "Radial Velocity", "Velocity Magnitude", "Pressure"}) fieldType.push_back(n);
    if (fieldOption.solveEnergy)
        for (const char* n : {"Temperature", "dT/dz", "dT/dr"}) fieldType.push_back(n);
    if (fieldOption.solveConcentration)
        for (const char* n : {"Concentration", "dC/dz", "dC/dr"}) fieldType.push_back(n);
    for (const char* n : {"Continuity", "Cell Reynolds Number", "dU/dz", "dU/dr", "dV/dz", "dV/dr"})
        fieldType.push_back(n);
}

Solver::~Solver() {
    shutdown();
    axs_destroy(handle_);
}

void Solver::shutdown() {
    if (worker_.joinable()) worker_.join();
}

void Solver::reset() {
    shutdown();
    solverRunning = false;
    stopRequested = false;
    continueSolver = false;
    isReady = false;
    solutions.clear();
    timeFrames.clear();
    fieldType.clear();
    fvMesh = FVMesh{};
    currentIteration_ = 0;
    fieldOption = SolverFieldOption{};
    configSolver = ConfigSolver{};
    configSimple = ConfigSimple{};
    configMultigrid = ConfigMultigrid{};
    initConfigResiduals(cfg);
    bindResidualStorage();
}

std::string Solver::lastMessage() const {
    std::lock_guard<std::mutex> lock(msgMutex_);
    return message_;
}

void Solver::requestStop() {  // solver.cpp:175-184
    if (!solverRunning) return;
    stopRequested = true;
    if (handle_) axs_request_stop(handle_);
}

bool Solver::canContinue(const Mesh& mesh, std::string* reason) const {  // solver.cpp:469-528
    auto fail = [&](const char* why) {
        if (reason) *reason = why;
        return false;
    };
    if (solverRunning) return fail("Solver is already running.");
    if (!isReady || !handle_) return fail("Run the solver once before continuing.");
    if (!mesh.isReady) return fail("Generate a mesh first.");
    const FVMesh& m = mesh.getFVMesh();
    if (m.numCells() != fvMesh.numCells()) return fail("Mesh cell count changed.");
    if (m.numFaces() != fvMesh.numFaces()) return fail("Mesh face count changed.");
    size_t a = 0, b = 0;
    for (const auto& c : m.cells) a += c.faceIDs.size();
    for (const auto& c : fvMesh.cells) b += c.faceIDs.size();
    if (a != b) return fail("Mesh connectivity changed.");
    return true;
}

bool Solver::runCheck(const Mesh& mesh) {  // solver.cpp:530-661
    auto fail = [&](const std::string& why) {
        say(why);
        return false;
    };
    if (solverRunning) return fail("Solver still running");
    if (!handle_) return fail("no GPU available");
    if (configSimple.maxIter < 1) configSimple.maxIter = 50;
    if (configSimple.checkConv < 1) configSimple.checkConv = 1;
    if (configSolver.linearMaxIter < 1) configSolver.linearMaxIter = 20;
    if (mesh.currentMeshType == MeshType::Structured) configSolver.useNonOrthCorrector = false;
    if (configSolver.saveKeyFrameIter < 1) configSolver.saveKeyFrameIter = 1;
    if (configSolver.transient) {
        if (!std::isfinite(configSolver.dt) || configSolver.dt <= 0.0)
            return fail("Transient run needs a positive time step.\n");
        if (!std::isfinite(configSolver.tEnd) || configSolver.tEnd <= 0.0)
            return fail("Transient run needs a positive end time.\n");
    }
    bool pulsatile = false;
    for (const auto& g : mesh.boundaryGroups) {
        if (g.type != BoundaryType::VELOCITY_INLET) continue;
        for (BoundaryVariable v : {BoundaryVariable::UVelocity, BoundaryVariable::VVelocity}) {
            const BoundaryCondition bc = effectiveBC(g, v);
            if (bc.type != BCType::PULSATILE) continue;
            pulsatile = true;
            if (!std::isfinite(bc.value) || !std::isfinite(bc.amplitude) || bc.amplitude < 0.0 ||
                !std::isfinite(bc.frequency) || bc.frequency <= 0.0)
                return fail("Pulsatile inlet '" + g.name + "' needs a finite mean velocity, A >= 0, and f > 0 Hz.\n");
        }
    }
    if (pulsatile && !configSolver.transient) return fail("Pulsatile inlet requires Transient Solver to be enabled.\n");
    if (!std::isfinite(f.rho) || !std::isfinite(f.mu) || f.rho <= 0.0 || f.mu <= 0.0)
        return fail("Solver needs positive density and viscosity.\n");
    if (continueSolver) {
        std::string why;
        if (!canContinue(mesh, &why)) {
            continueSolver = false;
            say("Continue Solver disabled: " + why + " Starting from a fresh solve.\n");
        }
    }
    return true;
}

void Solver::run(Mesh& mesh) {  // solver.cpp:146-173
    mesh.refreshFVMesh();
    if (!runCheck(mesh)) return;
    if (!continueSolver) residualPlotResetState(residualPlot);
    shutdown();
    addFieldType();
    stopRequested = false;
    solverRunning = true;
    worker_ = std::thread([this, m = mesh]() mutable {
        runSimple(std::move(m));
        solverRunning = false;
    });
}

void Solver::progressThunk(void* user, int iteration, int n, const char* const* names, const double* values) {
    Solver* self = static_cast<Solver*>(user);
    std::map<std::string, double> vals;
    for (int i = 0; i < n; i++) {
        vals[names[i]] = values[i];
        auto it = self->cfg.find(names[i]);
        if (it != self->cfg.end() && it->second.resVal) *it->second.resVal = values[i];
    }
    self->currentIteration_ = iteration + 1;
    // residualPlot->add + printResidualConsole (solver.cpp:1260-1261)
    residualPlotAdd(self->residualPlot, iteration, self->cfg);
    if (self->console) {
        std::ostringstream line;
        line << "Iteration " << iteration;
        for (int i = 0; i < n; i++) line << "  " << names[i] << ": " << values[i];
        consoleAddLine(self->console, line.str() + "\n");
    }
    if (self->onResidual) self->onResidual(iteration, vals);
}

void Solver::runSimple(Mesh mesh) {
    const bool sameMesh = continueSolver && isReady;
    fvMesh = mesh.getFVMesh();
    const PackedMesh pm = pack(fvMesh);
    const axs_mesh_view mv = pm.view();
    auto report = [&](axs_status st) {
        if (st == AXS_OK) return true;
        say(std::string(axs_last_error(handle_)) + "\n");
        return false;
    };
    if (!sameMesh && !report(axs_set_mesh(handle_, &mv))) return;
    const BoundaryVariable vars[5] = {BoundaryVariable::UVelocity, BoundaryVariable::VVelocity,
                                      BoundaryVariable::Pressure, BoundaryVariable::StaticTemperature,
                                      BoundaryVariable::Concentration};
    for (int i = 0; i < 5; i++) {
        const BCArrays a = bcArrays(mesh.boundaryGroups, vars[i]);
        const axs_bc_view v = a.view();
        if (!report(axs_set_bcs(handle_, i, &v))) return;
    }
    axs_config c;
    axs_config_defaults(&c);
    c.convection_scheme = (int)configSolver.convectionScheme;
    c.gradient_scheme = (int)configSolver.gradientScheme;
    c.linear_solver = (int)configSolver.type;
    c.time_scheme = (int)configSolver.timeScheme;
    c.linear_max_iter = configSolver.linearMaxIter;
    c.use_multigrid = configSolver.useMultigrid;
    c.add_convection = configSolver.addConvectionTerm;
    c.transient = configSolver.transient;
    c.use_non_orth = configSolver.useNonOrthCorrector;
    c.dt = configSolver.dt;
    c.t_end = configSolver.tEnd;
    c.save_key_frame_iter = configSolver.saveKeyFrameIter;
    c.max_time_frames = maxTimeFrames;
    c.max_iter = configSimple.maxIter;
    c.check_conv = configSimple.checkConv;
    c.mg_max_iter = configMultigrid.maxIter;
    c.mg_weight = configMultigrid.weight;
    c.mg_coarse_sweeps = configMultigrid.linearSweep;
    c.mg_pre_sweeps = configMultigrid.linearPreSweep;
    c.mg_post_sweeps = configMultigrid.linearPostSweep;
    c.mg_max_levels = configMultigrid.maxLevels;
    c.precision = configMultigrid.precision;
    c.coarse_direct = configMultigrid.coarseDirect;
    c.alpha_u = simple.momentumRelaxation;
    c.alpha_p = simple.pressureRelaxation;
    c.rho = f.rho;
    c.mu = f.mu;
    c.structured_mesh = mesh.currentMeshType == MeshType::Structured;
    c.solve_energy = fieldOption.solveEnergy;
    c.solve_concentration = fieldOption.solveConcentration;
    c.cp = f.cp;
    c.k = f.k;
    c.diffusivity = f.D;
    const char* names[5] = {"U", "V", "Continuity", "Temperature", "Concentration"};
    for (int i = 0; i < 5; i++) {
        const ConfigResidual& r = cfg.at(names[i]);
        c.residual[i].enabled = r.enabled;
        c.residual[i].norm = (int)r.normType;
        c.residual[i].scaling = (int)r.scaleType;
        c.residual[i].tol = r.tol;
    }
    if (!report(axs_configure(handle_, &c))) return;
    if (!report(axs_run(handle_, &Solver::progressThunk, this, continueSolver ? 1 : 0))) return;

    // createSolutions / createDerivedVelocitySolutions, solver.cpp:186-263
    const std::pair<const char*, int> fields[] = {
        {"Axial Velocity", AXS_FIELD_U},   {"Radial Velocity", AXS_FIELD_V}, {"Pressure", AXS_FIELD_P},
        {"dU/dz", AXS_FIELD_GRAD_UZ},      {"dU/dr", AXS_FIELD_GRAD_UR},     {"dV/dz", AXS_FIELD_GRAD_VZ},
        {"dV/dr", AXS_FIELD_GRAD_VR},      {"dP/dz", AXS_FIELD_GRAD_PZ},     {"dP/dr", AXS_FIELD_GRAD_PR},
        {"Continuity", AXS_FIELD_CONTINUITY}, {"Velocity Magnitude", AXS_FIELD_VMAG},
        {"Cell Reynolds Number", AXS_FIELD_CELL_RE}, {"Temperature", AXS_FIELD_T}, {"dT/dz", AXS_FIELD_GRAD_TZ},
        {"dT/dr", AXS_FIELD_GRAD_TR}, {"Concentration", AXS_FIELD_C}, {"dC/dz", AXS_FIELD_GRAD_CZ},
        {"dC/dr", AXS_FIELD_GRAD_CR}};
    solutions.clear();
    for (const auto& [name, id] : fields) {
        const bool isT = id == AXS_FIELD_T || id == AXS_FIELD_GRAD_TZ || id == AXS_FIELD_GRAD_TR;
        const bool isC = id == AXS_FIELD_C || id == AXS_FIELD_GRAD_CZ || id == AXS_FIELD_GRAD_CR;
        if ((isT && !fieldOption.solveEnergy) || (isC && !fieldOption.solveConcentration)) continue;
        std::vector<double> v(axs_field_length(handle_, id));
        if (!report(axs_get_field(handle_, id, v.data(), v.size()))) return;
        BoundaryVariable var = BoundaryVariable::None;
        if (id == AXS_FIELD_U || id == AXS_FIELD_GRAD_UZ || id == AXS_FIELD_GRAD_UR) var = BoundaryVariable::UVelocity;
        if (id == AXS_FIELD_V || id == AXS_FIELD_GRAD_VZ || id == AXS_FIELD_GRAD_VR) var = BoundaryVariable::VVelocity;
        if (id == AXS_FIELD_P || id == AXS_FIELD_GRAD_PZ || id == AXS_FIELD_GRAD_PR) var = BoundaryVariable::Pressure;
        if (isT) var = BoundaryVariable::StaticTemperature;
        if (isC) var = BoundaryVariable::Concentration;
        solutions[name] = SolutionField{std::move(v), g.dr, g.dz, var};
    }
    // timeFrames: the library keeps a transient run's frames (appending on a
    // continued run, dropping them on a fresh or steady one); mirror its list
    timeFrames.clear();
    const int nFrames = axs_frame_count(handle_);
    const std::pair<const char*, int> frameFields[] = {{"Axial Velocity", AXS_FIELD_U},
                                                       {"Radial Velocity", AXS_FIELD_V},
                                                       {"Pressure", AXS_FIELD_P},
                                                       {"Temperature", AXS_FIELD_T},
                                                       {"Concentration", AXS_FIELD_C}};
    for (int i = 0; i < nFrames; i++) {
        TimeFrame fr;
        for (const auto& [name, id] : frameFields) {
            if ((id == AXS_FIELD_T && !fieldOption.solveEnergy) || (id == AXS_FIELD_C && !fieldOption.solveConcentration))
                continue;
            std::vector<double> v(fvMesh.numCells());
            if (!report(axs_get_frame(handle_, i, id, &fr.time, v.data(), v.size()))) return;
            fr.fields[name] = std::move(v);
        }
        timeFrames.push_back(std::move(fr));
    }
    axs_stats st;
    if (axs_get_stats(handle_, &st) == AXS_OK) currentIteration_ = st.iterations;
    isReady = true;
}

}  // namespace axs::compat
