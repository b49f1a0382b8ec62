// topology.hpp -- host-side set-up builders of the solver (run once per mesh).
#pragma once

#include <cstdint>
#include <vector>

#include "mesh_types.hpp"

namespace axs {

// Geometry derived from the mesh (preComputeVariables, memory_manger.cpp:512-617).
struct Geometry {
    std::vector<double> invDPN, wP, dPB, invA2D, lsqWZ, lsqWR, lsqInvZZ, lsqInvZR, lsqInvRR, length2D;
};
void computeGeometry(const HostMesh& m, Geometry& g);

// The face data the cell loops read, copied into slot order (slot k of cell c
// = face cellFaceIDs[k]) and oriented for c.  Every entry is the face value,
// its negation, 1 - w, or the reference's own expression for the non-orthogonal
// vector evaluated once here, so the kernels reading it compute bit-identical
// results while streaming slots instead of gathering faces from both sides.
struct SlotGeometry {
    enum : uint8_t { kOwner = 1, kAxis = 2 };
    std::vector<double> area, invDPN;   // face area, 1/|d_PN|
    std::vector<double> nz, nr;         // normal, negated where c is the neighbour
    std::vector<double> w;              // interpolation weight of c: wP, or 1 - wP for the neighbour
    std::vector<double> tz, tr;         // A n - (A / |d_PN|) d_cN (interior faces)
    std::vector<double> lsqWZ, lsqWR;   // LSQ weights, negated where c is the neighbour
    std::vector<double> len2D;          // planar face length
    std::vector<uint8_t> flags;         // kOwner: c owns the face; kAxis: face centre on the axis
};
void computeSlotGeometry(const HostMesh& m, const Geometry& g, SlotGeometry& s);

// Cell -> neighbour CSR in cell.faceIDs order (buildCellFaceCSR, memory_manger.cpp:287-323).
void cellNeighbourCSR(const HostMesh& m, std::vector<int>& start, std::vector<int>& nb);

// Blocks of `cells` consecutive cells and their halo rings, for running several
// Jacobi sweeps in one workgroup: ring r of a block holds the cells at graph
// distance r from it.  Sweep j of s fused sweeps updates the block and rings
// 1 .. s - j, so s sweeps need rings 1 .. s - 1.  Per block, the halo lists
// ring 1's cells, then ring 2's, ..., each ring ascending.  In the workgroup's
// LDS, block cell a + i sits at slot i and halo entry i at slot cells + i;
// the tables give, per slot of a row that a later sweep reads from LDS, its
// neighbour's LDS slot (-1: boundary slot or past the row's slots).
struct SweepBlocks {
    int cells = 0, rings = 0, nBlocks = 0, rowSlots = 0;
    // patches: the blocks are breadth-first patches, not index ranges; own
    // lists each block's cells (cells entries per block, ascending, -1 padded)
    bool patches = false;
    std::vector<int> own;
    int maxHalo = 0;   // largest halo of any block
    int maxSlots = 0;  // most slots of any cell
    std::vector<int> haloStart;  // nBlocks + 1 offsets into halo
    std::vector<int> ringEnd;    // [block * rings + r - 1]: halo entries of rings 1..r
    std::vector<int> halo;
    std::vector<int16_t> locOwn;   // per slot of every cell (the mesh's slot order)
    std::vector<int16_t> locHalo;  // per halo entry, rowSlots slots (rings 1 .. rings - 1)
};
// rowSlots: slots per halo row in locHalo (cells with more slots leave the
// fused path; maxSlots tells).  patches: grow each block breadth-first from
// the lowest-numbered cell not yet in a block (for meshes whose index ranges
// are long thin strips, e.g. structured grids numbered row by row), instead
// of taking consecutive cells.
void buildSweepBlocks(int nCells, const std::vector<int>& start, const std::vector<int>& nb, int cells, int rings,
                      int rowSlots, SweepBlocks& out, bool patches = false);

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
