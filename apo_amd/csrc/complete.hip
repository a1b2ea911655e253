// complete.hip -- consistency rounds of the completion mode (no counterpart
// in the reference).
//
// A vertex of the Apollonius diagram is tangent to four sites and belongs to
// the four cells.  The reference builds each cell on its own and can leave a
// cell without a vertex its three other cells hold (for instance when a
// contributor search skips the site that would split one of the cell's
// edges).  A round:
//  * check: every record (i; j, k, l) of a live cell, with j, k, l real
//    sites, asks cells j, k and l for as many records of the quadruple as
//    cell i holds.  A cell m holding fewer wants the quadruple sites it does
//    not know; a device-wide atomicMin keeps the smallest one per cell, so
//    the round's result does not depend on the order of the requests.  A cell
//    that knows all three sites yet holds fewer records cannot be repaired
//    by an update and is counted.  Buried cells and cells with a capacity
//    flag take no part;
//  * fix: each cell with a wanted site is updated with it (updateCell, the
//    full vertex list as in vertex validation) and goes back to validation;
//    the vertex + edge validation rounds that follow validate its new
//    vertices and edges.
// A round after one that updated no cell does nothing, so every round can be
// enqueued (and captured in a graph) without a host round trip.
#include "apo_device.h"
#include "apo_internal.h"
#include "cell_update.h"

namespace apo_hip {

namespace {

// Records of cell m (global memory) whose three sites are (a, b, d).
__device__ __forceinline__ uint32_t recordCount(const CellCtx& c, uint32_t m, uint32_t a, uint32_t b, uint32_t d) {
    const uint4* gv = c.vtx + size_t(m) * kVertexCap;
    const uint32_t V = min(c.vnb[m], kVertexCap);
    uint32_t count = 0;
    for (uint32_t v = 0; v < V; v++) {
        const uint4 r = gv[v];
        count += r.x == a && r.y == b && r.z == d;
    }
    return count;
}

// Site j is an endpoint of one of cell m's edges (global memory).
__device__ __forceinline__ bool cellKnows(const CellCtx& c, uint32_t m, uint32_t j) {
    const uint4* ge = c.edg + size_t(m) * kEdgeCap;
    const uint32_t E = min(c.enb[m], kEdgeCap);
    for (uint32_t e = 0; e < E; e++) {
        const uint4 r = ge[e];
        if (r.x == j || r.y == j) return true;
    }
    return false;
}

__device__ __forceinline__ bool takesPart(const CellCtx& c, uint32_t m) {
    return c.status[m] != kStatusBuried && c.flags[m] == 0u;
}

}  // namespace

// One wave per cell, grid-stride; the cell's records sit in LDS for the
// counts of its own quadruples.
__global__ __launch_bounds__(64) void k_consistency_check(CellCtx c, uint32_t* __restrict__ want,
                                                          uint32_t* counters, uint32_t round) {
    __shared__ uint4 recs[kVertexCap];
    if (round > 0u && __builtin_amdgcn_readfirstlane(counters[kCompletionUpdated + round - 1u]) == 0u) return;
    const uint32_t lane = laneId();
    const uint32_t n = c.n;
    uint32_t stuck = 0;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        if (!takesPart(c, i)) continue;
        const uint32_t V = min(c.vnb[i], kVertexCap);
        const uint4* gv = c.vtx + size_t(i) * kVertexCap;
        for (uint32_t v = lane; v < V; v += 64u) recs[v] = gv[v];
        waveSync();
        for (uint32_t v = lane; v < V; v += 64u) {
            const uint4 r = recs[v];
            if (r.z >= n) continue;  // j < k < l: an artificial site is last
            uint32_t own = 0;
            for (uint32_t u = 0; u < V; u++) own += recs[u].x == r.x && recs[u].y == r.y && recs[u].z == r.z;
            const uint32_t q[4] = {i, r.x, r.y, r.z};
#pragma unroll
            for (int t = 1; t < 4; t++) {
                const uint32_t m = q[t];
                if (!takesPart(c, m)) continue;
                uint32_t o[3];
                int k = 0;
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (u != t) o[k++] = q[u];
                sort3(o[0], o[1], o[2]);
                if (recordCount(c, m, o[0], o[1], o[2]) >= own) continue;
                bool any = false;
#pragma unroll
                for (int u = 0; u < 3; u++)
                    if (!cellKnows(c, m, o[u])) {
                        atomicMin(&want[m], o[u]);
                        any = true;
                    }
                if (!any) stuck++;
            }
        }
        waveSync();
    }
    const uint32_t total = waveSumU(stuck);
    if (lane == 0 && total) atomicAdd(&counters[kCompletionStuck], total);
}

// One wave per cell with a wanted site: update, store, back to validation.
__global__ __launch_bounds__(64) void k_consistency_fix(CellCtx c, uint32_t* __restrict__ want, uint32_t* counters,
                                                        uint32_t round) {
    __shared__ CellLds L;
    const uint32_t lane = laneId();
    uint32_t updated = 0;
    for (uint32_t m = blockIdx.x; m < c.n; m += gridDim.x) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(want[m]);
        if (j == kNone) continue;
        WaveCell w;
        loadCell(c, L, w, m, true);
        const D4 sm = loadSite(c, m);
        const bool changed = updateCell(c, L, w, m, sm, j, loadSite(c, j));
        storeCell(c, L, w, m);
        if (lane == 0) {
            want[m] = kNone;  // ready for the next round
            if (changed) c.status[m] = kStatusInValidation;
        }
        updated += changed ? 1u : 0u;
        waveSync();
    }
    if (lane == 0 && updated) atomicAdd(&counters[kCompletionUpdated + round], updated);
}

hipError_t consistencyLaunch(const CellCtx& c, uint32_t* want, uint32_t* counters, uint32_t round, uint32_t blocks,
                             hipStream_t st) {
    if (round >= kMaxCompletionRounds) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_consistency_check, dim3(blocks), dim3(64), 0, st, c, want, counters, round);
    hipLaunchKernelGGL(k_consistency_fix, dim3(blocks), dim3(64), 0, st, c, want, counters, round);
    return hipGetLastError();
}

}  // namespace apo_hip
