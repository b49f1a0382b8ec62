// mesh_types.hpp -- host-side face-based mesh (structure of arrays).
//
// Mirrors FVMeshHostPacked (structs/solver_struct.h:400-429): the single
// representation both case generators emit and the solver uploads.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "axs.h"

namespace axs {

struct HostMesh {
    int nCells = 0;
    int nFaces = 0;
    std::vector<double> cellCenterZ, cellCenterR, cellVolume, cellArea2D;
    std::vector<int> cellFaceStart, cellFaceIDs;
    std::vector<int> faceOwner, faceNeighbor, faceGroup;
    std::vector<double> faceNormalZ, faceNormalR, faceCenterZ, faceCenterR, faceArea;
    // triangle meshes keep their node list (for export / invariants)
    std::vector<double> pointZ, pointR;
    std::vector<int> tris;
    // multiblock meshes keep each cell's corner quad, 4 per cell in cell order
    // (Mesh::multiBlockCellCorners, solver/source/mesh.cpp:2514-2537)
    std::vector<double> quadZ, quadR;

    axs_mesh_view view() const {
        axs_mesh_view v{};
        v.n_cells = nCells;
        v.n_faces = nFaces;
        v.cell_center_z = cellCenterZ.data();
        v.cell_center_r = cellCenterR.data();
        v.cell_volume = cellVolume.data();
        v.cell_area2d = cellArea2D.data();
        v.cell_face_start = cellFaceStart.data();
        v.cell_face_ids = cellFaceIDs.data();
        v.face_owner = faceOwner.data();
        v.face_neighbor = faceNeighbor.data();
        v.face_normal_z = faceNormalZ.data();
        v.face_normal_r = faceNormalR.data();
        v.face_center_z = faceCenterZ.data();
        v.face_center_r = faceCenterR.data();
        v.face_area = faceArea.data();
        v.face_group = faceGroup.data();
        return v;
    }
};

// Builders (casegen.cpp).  Return an empty string on success, else the reason.
std::string buildMultiblock(HostMesh& out, int nBlocks, const axs_block_spec* blocks);
std::string buildTriangles(HostMesh& out, const std::vector<double>& pz, const std::vector<double>& pr,
                           const std::vector<int>& tris, int nSeg, const double* sz0, const double* sr0,
                           const double* sz1, const double* sr1, const int* sgroup, double tol);
std::string buildNozzle(HostMesh& out, const axs_nozzle_spec& spec);
// Disjoint union of meshes for a batched run: cells, faces and slots of mesh c
// follow those of meshes 0..c-1 in order, and its boundary group g becomes
// g + c * groupsPerMesh.  Every per-cell and per-face step of the solver then
// treats each case exactly as it would alone.
std::string concatMeshes(HostMesh& out, int count, const axs_mesh_view* meshes, int groupsPerMesh);

}  // namespace axs
