// compat_demo.cpp -- plays AxiSim's GUI against the Solver drop-in: constructs
// it on the project Config, hands it the console and residual plot, fills Mesh
// (FVMesh + boundary groups), writes the config members and labels the
// settings the way the panels do (name arrays, varUnits), run()s on the worker
// thread, polls solverRunning, reads `solutions`, `fieldType` and, for a
// transient run, `timeFrames` (Results builds its animation from them).
//
// usage: compat_demo [max_iter]
//   1. steady pipe to convergence: prints "iterations=... u_rms_err=..."
//   2. transient pulsatile pipe, 5 steps, keyframes every 2 steps, then a
//      continued run of 3 steps: prints "frames=... times=..."
// exit 0 when both check out.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>

#include "console.h"
#include "residual_plot.h"
#include "solver.hpp"

namespace {

Mesh pipeMesh() {
    // z in [0, 0.2] x 64 cells, r in [0, 0.01] x 16 cells
    axs_block_spec blk{};
    blk.axial_length = 0.2;
    blk.axial_cells = 64;
    blk.radial_length = 0.01;
    blk.radial_cells = 16;
    blk.axial_coef = blk.radial_coef = 1.0;
    blk.edge_group[0] = 0;  // west: inlet
    blk.edge_group[1] = 1;  // east: outlet
    blk.edge_group[2] = 3;  // south: axis
    blk.edge_group[3] = 2;  // north: wall
    axs_mesh* hm = nullptr;
    if (axs_mesh_multiblock(&hm, 1, &blk) != AXS_OK) std::exit(3);
    axs_mesh_view v;
    axs_mesh_get_view(hm, &v);
    Mesh mesh;
    mesh.isReady = true;
    mesh.currentMeshType = MeshType::Structured;
    for (int c = 0; c < v.n_cells; c++) {
        FVCell cell;
        cell.center = {v.cell_center_z[c], v.cell_center_r[c]};
        cell.volume = v.cell_volume[c];
        cell.area2D = v.cell_area2d[c];
        for (int k = v.cell_face_start[c]; k < v.cell_face_start[c + 1]; k++) cell.faceIDs.push_back(v.cell_face_ids[k]);
        mesh.fvMesh.cells.push_back(cell);
    }
    for (int f = 0; f < v.n_faces; f++) {
        FVFace face;
        face.owner = v.face_owner[f];
        face.neighbor = v.face_neighbor[f];
        face.normal = {v.face_normal_z[f], v.face_normal_r[f]};
        face.center = {v.face_center_z[f], v.face_center_r[f]};
        face.area = v.face_area[f];
        face.boundaryGroupID = v.face_group[f];
        mesh.fvMesh.faces.push_back(face);
    }
    axs_mesh_destroy(hm);
    const BoundaryType types[4] = {BoundaryType::VELOCITY_INLET, BoundaryType::PRESSURE_OUTLET, BoundaryType::WALL,
                                   BoundaryType::SYMMETRY};
    const char* names[4] = {"inlet", "outlet", "wall", "axis"};
    for (int g = 0; g < 4; g++) {
        BoundarySegmentGroup grp;
        grp.id = g;
        grp.name = names[g];
        grp.type = types[g];
        mesh.boundaryGroups.push_back(grp);
    }
    BoundaryCondition inlet;
    inlet.type = BCType::DIRICHLET;
    inlet.value = 0.005;
    mesh.boundaryGroups[0].bcs[(int)BoundaryVariable::UVelocity] = inlet;
    return mesh;
}

void waitFor(Solver& solver) {
    while (solver.solverRunning) std::this_thread::sleep_for(std::chrono::milliseconds(5));
    solver.shutdown();
}

}  // namespace

int main(int argc, char** argv) {
    Config project;            // the project's configuration (Project::config)
    project.f.rho = 1000.0;    // solver_gui.cpp:771-783 write these through solver.f
    Solver solver(project);
    Console console;
    ResidualPlot plot;
    solver.console = &console;  // graphics/window/source/gui.cpp:307-313
    solver.residualPlot = &plot;
    solver.f.mu = 1e-3;
    solver.varUnits.rhoUnit = 0;
    if (&solver.f != &project.f || &solver.g != &project.g || &solver.varUnits != &project.varUnits ||
        &solver.config != &project)
        return 4;
    // the settings panels label their combos from the name arrays
    solver.configSolver.gradientScheme = GradientScheme::GRAD_GREEN_GAUSS;
    std::printf("scheme: %s / %s, residuals:", solver.convectionDiscretizationType[(int)solver.configSolver.convectionScheme],
                solver.gradientSchemeType[(int)solver.configSolver.gradientScheme]);
    for (const char* name : solver.residualPlotType) std::printf(" %s", name);
    std::printf(" | %s %s %s %s %s %s %s %s %s\n", solver.boundaryType[1], solver.residualType[0],
                solver.residualNormType[0], solver.residualScalingType[3], solver.linearSolverType[0],
                solver.velocitySolverType[0], solver.bcTypeNames[0], solver.bcInletTypeNames[2],
                solver.timeSchemeType[1]);

    // 1. steady pipe
    Mesh mesh = pipeMesh();
    solver.configSimple.maxIter = argc > 1 ? std::atoi(argv[1]) : 20000;
    for (auto& [name, r] : solver.cfg) r.tol = 1e-6;
    int checks = 0;
    solver.onResidual = [&](int, const std::map<std::string, double>&) { checks++; };
    solver.run(mesh);
    waitFor(solver);
    if (!solver.isReady) {
        std::printf("run failed: %s\n", solver.lastMessage().c_str());
        return 2;
    }
    const auto& u = solver.solutions.at("Axial Velocity").field;
    double err = 0.0;
    int n = 0;
    for (size_t c = 0; c < mesh.fvMesh.cells.size(); c++) {
        const FVCell& cell = mesh.fvMesh.cells[c];
        if (cell.center.z < 0.15) continue;
        const double ua = 0.01 * (1.0 - (cell.center.r / 0.01) * (cell.center.r / 0.01));
        err += (u[c] - ua) * (u[c] - ua);
        n++;
    }
    err = std::sqrt(err / n) / 0.01;
    const double cont = *solver.cfg.at("Continuity").resVal;
    bool fieldsOk = solver.fieldType.size() == 10;
    for (const std::string& name : solver.fieldType) fieldsOk = fieldsOk && solver.solutions.count(name) == 1;
    const bool plotOk = (int)plot.iterations.size() == checks && plot.values.size() == 3 &&
                        plot.values.at("Continuity").back() == cont;
    std::printf("iterations=%d checks=%d continuity=%.3e u_rms_err=%.4f fields=%d plot=%d console=%zu frames=%zu\n",
                solver.currentIteration(), checks, cont, err, fieldsOk, plotOk, console.lines.size(),
                solver.timeFrames.size());
    const bool steadyOk = cont <= 1e-6 && err < 0.01 && fieldsOk && plotOk && !console.lines.empty() &&
                          solver.timeFrames.empty();

    // 2. transient pulsatile inlet: 5 steps of dt = 0.0625 s, a frame every
    // 2 steps plus the last, then a continued run of 3 steps appends
    BoundaryCondition pulse;
    pulse.type = BCType::PULSATILE;
    pulse.value = 0.005;
    pulse.amplitude = 0.3;
    pulse.frequency = 2.0;
    mesh.boundaryGroups[0].bcs[(int)BoundaryVariable::UVelocity] = pulse;
    solver.configSolver.transient = true;
    solver.configSolver.timeScheme = TimeScheme::TIME_SECOND_ORDER;
    solver.configSolver.dt = 0.0625;
    solver.configSolver.tEnd = 5 * 0.0625;
    solver.configSolver.saveKeyFrameIter = 2;
    solver.configSimple.maxIter = 5;
    solver.maxTimeFrames = 500;
    solver.continueSolver = false;
    solver.run(mesh);
    waitFor(solver);
    const std::vector<double> first = {0, 1, 3, 5};
    bool framesOk = solver.isReady && solver.timeFrames.size() == first.size();
    for (size_t i = 0; framesOk && i < first.size(); i++)
        framesOk = std::fabs(solver.timeFrames[i].time - first[i] * 0.0625) < 1e-12 &&
                   solver.timeFrames[i].fields.at("Axial Velocity").size() == mesh.fvMesh.cells.size();
    solver.continueSolver = true;
    solver.configSolver.tEnd = 3 * 0.0625;
    solver.run(mesh);
    waitFor(solver);
    const std::vector<double> all = {0, 1, 3, 5, 6, 8};
    framesOk = framesOk && solver.timeFrames.size() == all.size();
    for (size_t i = 0; framesOk && i < all.size(); i++)
        framesOk = std::fabs(solver.timeFrames[i].time - all[i] * 0.0625) < 1e-12;
    std::printf("frames=%zu times=", solver.timeFrames.size());
    for (const auto& fr : solver.timeFrames) std::printf("%g ", fr.time);
    std::printf("ok=%d\n", framesOk);
    return (steadyOk && framesOk) ? 0 : 1;
}
