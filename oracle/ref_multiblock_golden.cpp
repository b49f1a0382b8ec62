// ref_multiblock_golden.cpp -- golden-vector writer built against the
// reference's own header-only multiblock packer (structs/multiblock.h).
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile from the reference
// sources where they lie (never copied); the binary goes to oracle/_ref/.  It
// writes the packed face-based mesh of a block layout given on the command line
// so tests can compare the case generator's output with the reference's.
//
// usage: ref_multiblock_golden OUT.bin nblocks {z0 r0 al ac acoef rl rc rcoef gW gE gS gN}*
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "multiblock.h"

struct Packed {  // the members toPackedMesh writes (structs/solver_struct.h:400-429)
    int nCells = 0, nFaces = 0;
    std::vector<int> faceOwner, faceNeighbor, faceBoundaryGroupID, cellFaceStart, cellFaceIDs;
    std::vector<double> faceNormalZ, faceNormalR, faceCenterZ, faceCenterR, faceArea;
    std::vector<double> cellCenterZ, cellCenterR, cellArea2D, cellVolume;
};

template <class T>
static void put(FILE* f, const std::vector<T>& v) {
    const long long n = (long long)v.size();
    fwrite(&n, sizeof(n), 1, f);
    if (n) fwrite(v.data(), sizeof(T), v.size(), f);
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const int nb = std::atoi(argv[2]);
    if (argc != 3 + 12 * nb) return 2;
    MultiBlockMesh m;
    for (int b = 0; b < nb; b++) {
        const char* const* a = argv + 3 + 12 * b;
        MeshZone az{std::atof(a[2]), std::atoi(a[3])};
        MeshZone rz{std::atof(a[5]), std::atoi(a[6])};
        const double acoef = std::atof(a[4]), rcoef = std::atof(a[7]);
        if (acoef != 1.0) { az.grading = Grading::Progression; az.coef = acoef; }
        if (rcoef != 1.0) { rz.grading = Grading::Progression; rz.coef = rcoef; }
        Block blk = makeRectBlock(b, std::atof(a[0]), std::atof(a[1]), az, rz);
        for (int e = 0; e < 4; e++) blk.edgeGroup[e] = std::atoi(a[8 + e]);
        m.blocks.push_back(blk);
    }
    m.assignGlobalNumbering();
    std::string reason;
    if (!autoDetectInterfaces(m, reason)) {
        std::fprintf(stderr, "%s\n", reason.c_str());
        return 1;
    }
    Packed p;
    toPackedMesh(m, p);
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 1;
    put(f, p.cellCenterZ); put(f, p.cellCenterR); put(f, p.cellVolume); put(f, p.cellArea2D);
    put(f, p.cellFaceStart); put(f, p.cellFaceIDs);
    put(f, p.faceOwner); put(f, p.faceNeighbor); put(f, p.faceNormalZ); put(f, p.faceNormalR);
    put(f, p.faceCenterZ); put(f, p.faceCenterR); put(f, p.faceArea); put(f, p.faceBoundaryGroupID);
    std::fclose(f);
    return 0;
}
