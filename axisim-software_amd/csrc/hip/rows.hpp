// rows.hpp -- device helpers shared by the kernel files: run control, the
// XCD-aware block numbering and the batched walk over a row's slots.
#pragma once

#include "device.hpp"

namespace axs {
namespace {

__device__ __forceinline__ bool finished(const Ctl* ctl) { return ctl != nullptr && ctl->done != 0; }

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2.  Every kernel here numbers its blocks
// so that XCD x runs one contiguous range of them: the neighbour gathers
// across a block's edge then mostly hit lines its neighbour block, on the same
// XCD, has just brought in.  (Only the placement changes; each block still
// covers the same cells.)
__device__ __forceinline__ int xcdBlock() {
    const int g = gridDim.x, b = blockIdx.x;
    const int q = g >> 3, r = g & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

// Cells with at most B slots (B = 3 on triangle meshes, 4 otherwise) walk
// their slots in three dependent rounds of loads instead of one round per slot:
// every slot's data (indices clamped into the cell's range, so every load is in
// bounds and can issue at once), then every neighbour gather, then the
// arithmetic slot by slot in the reference's order.  Larger cells take the
// plain loop.
__device__ __forceinline__ int slotAt(int k0, int ns, int q) { return k0 + (q < ns ? q : ns - 1); }

// The slot range of row n.  When every row has the same number of slots
// (uniform > 0: all-triangle or all-quad meshes) it is uniform * n, with no load.
__device__ __forceinline__ void rowRange(const int* start, int uniform, int n, int& k0, int& ns) {
    if (uniform > 0) {
        k0 = uniform * n;
        ns = uniform;
    } else {
        k0 = start[n];
        ns = start[n + 1] - k0;
    }
}

// ---- the fused (temporally blocked) sweep kernels' shared pieces ----

// halo entries a block may have: at most one per thread (thread t holds halo
// row t), and the LDS of two iterate buffers within the static 64 KiB
template <int C>
constexpr int sweepHalo() { return C <= 512 ? C : 5 * C / 8; }

// the neighbours' LDS slots of a row from its table entries l[0 .. ns)
template <class Row>
__device__ __forceinline__ void loadLoc(Row& r, const int16_t* l, int R) {
    for (int q = 0; q < R; q++) r.loc[q] = (q < r.ns) ? (int)l[q] : -1;
}

}  // namespace
}  // namespace axs
