// topology.hpp -- host-side set-up builders of the solver (run once per mesh).
#pragma once

#include <vector>

#include "mesh_types.hpp"

namespace axs {

// Geometry derived from the mesh (preComputeVariables, memory_manger.cpp:512-617).
struct Geometry {
    std::vector<double> invDPN, wP, dPB, invA2D, lsqWZ, lsqWR, lsqInvZZ, lsqInvZR, lsqInvRR, length2D;
};
void computeGeometry(const HostMesh& m, Geometry& g);

// Cell -> neighbour CSR in cell.faceIDs order (buildCellFaceCSR, memory_manger.cpp:287-323).
void cellNeighbourCSR(const HostMesh& m, std::vector<int>& start, std::vector<int>& nb);

// Greedy smallest-free colouring in cell order, colour-major order
// (buildMeshColoring, memory_manger.cpp:334-409).  Returns the colour count.
int greedyColoring(int n, const int* start, const int* nb, std::vector<int>& colorStart,
                   std::vector<int>& cellOrder);

// One level of the agglomeration hierarchy (GridLevel, multigrid.cuh:29-58)
// plus the gather tables the GPU uses instead of atomics.
struct Level {
    int n = 0;
    std::vector<int> start, nb;                     // CSR of this level
    std::vector<int> toCoarse, slotToCoarseSlot;    // empty on the coarsest level
    // gather tables toward the next coarser level (empty on the coarsest):
    std::vector<int> memberStart, members;          // coarse cell -> fine members, ascending
    std::vector<int> cslotStart, cslotFine;         // coarse slot -> fine slots, ascending
    int refs() const { return (int)nb.size(); }
    bool coarsest() const { return toCoarse.empty(); }
};

// buildLevels, multigrid.cu:298-325; maxLevels 0 = reference rule.
std::vector<Level> buildHierarchy(int n, const std::vector<int>& start, const std::vector<int>& nb,
                                  int maxLevels);

}  // namespace axs
