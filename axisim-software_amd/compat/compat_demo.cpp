// compat_demo.cpp -- drives the Solver drop-in the way AxiSim's GUI does:
// fill Mesh (FVMesh + boundary groups), write the config members, run() on the
// worker thread, poll solverRunning, read `solutions`.
//
// usage: compat_demo [max_iter]   -> prints "iterations=... u_rms_err=..."; exit 0 when converged
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>

#include "solver.hpp"

using namespace axs::compat;

int main(int argc, char** argv) {
    // pipe: z in [0, 0.2] x 64 cells, r in [0, 0.01] x 16 cells
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
    if (axs_mesh_multiblock(&hm, 1, &blk) != AXS_OK) return 3;
    axs_mesh_view v;
    axs_mesh_get_view(hm, &v);

    Mesh mesh;
    mesh.isReady = true;
    mesh.currentMeshType = MeshType::Structured;
    for (int c = 0; c < v.n_cells; c++) {
        FVCell cell;
        cell.center = Vec2{v.cell_center_z[c], v.cell_center_r[c]};
        cell.volume = v.cell_volume[c];
        cell.area2D = v.cell_area2d[c];
        for (int k = v.cell_face_start[c]; k < v.cell_face_start[c + 1]; k++) cell.faceIDs.push_back(v.cell_face_ids[k]);
        mesh.fvMesh.cells.push_back(cell);
    }
    for (int f = 0; f < v.n_faces; f++) {
        FVFace face;
        face.owner = v.face_owner[f];
        face.neighbor = v.face_neighbor[f];
        face.normal = Vec2{v.face_normal_z[f], v.face_normal_r[f]};
        face.center = Vec2{v.face_center_z[f], v.face_center_r[f]};
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

    Solver solver;
    solver.f.rho = 1000.0;
    solver.f.mu = 1e-3;
    solver.configSolver.gradientScheme = GradientScheme::GRAD_GREEN_GAUSS;
    solver.configSimple.maxIter = argc > 1 ? std::atoi(argv[1]) : 20000;
    for (auto& [name, r] : solver.cfg) r.tol = 1e-6;
    int checks = 0;
    solver.onResidual = [&](int, const std::map<std::string, double>&) { checks++; };
    solver.run(mesh);
    while (solver.solverRunning) std::this_thread::sleep_for(std::chrono::milliseconds(5));
    solver.shutdown();
    if (!solver.isReady) {
        std::printf("run failed: %s\n", solver.lastMessage().c_str());
        return 2;
    }
    const auto& u = solver.solutions.at("Axial Velocity").field;
    double err = 0.0;
    int n = 0;
    for (int c = 0; c < v.n_cells; c++) {
        const FVCell& cell = mesh.fvMesh.cells[c];
        if (cell.center.z < 0.15) continue;
        const double ua = 0.01 * (1.0 - (cell.center.r / 0.01) * (cell.center.r / 0.01));
        err += (u[c] - ua) * (u[c] - ua);
        n++;
    }
    err = std::sqrt(err / n) / 0.01;
    const double cont = solver.cfg.at("Continuity").resValue;
    std::printf("iterations=%d checks=%d continuity=%.3e u_rms_err=%.4f\n", solver.currentIteration(), checks, cont,
                err);
    return (cont <= 1e-6 && err < 0.01) ? 0 : 1;
}
