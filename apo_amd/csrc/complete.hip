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
//    that knows all three sites yet holds fewer records is stuck: the order
//    in which its sites were inserted lost an edge (a trisector the update
//    never created), and no further update creates the vertex.  The record
//    goes to the round's stuck list.  Buried cells and cells with a capacity
//    flag take no part;
//  * fix: each cell with a wanted site is updated with it (updateCell, the
//    full vertex list as in vertex validation); a stuck cell with no wanted
//    site is rebuilt from its octahedron, first with the sites of its stuck
//    records (ascending), which creates their vertices at once, then with
//    its other neighbours nearest first (FP32 gap, ties by index).  An empty
//    vertex is never invalidated by a later update, so it survives.  A cell
//    is rebuilt at most once per build, and a round whose stuck list
//    overflows rebuilds no cell (the list's order is not deterministic, its
//    content is).  Changed cells go back to validation; the vertex + edge
//    validation rounds that follow validate their new vertices and edges.
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
                                                          uint4* __restrict__ stuck, uint32_t* counters,
                                                          uint32_t round) {
    __shared__ uint4 recs[kVertexCap];
    if (round > 0u && __builtin_amdgcn_readfirstlane(counters[kCompletionUpdated + round - 1u]) == 0u) return;
    const uint32_t lane = laneId();
    const uint32_t n = c.n;
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
                if (!any) {
                    const uint32_t idx = atomicAdd(&counters[kCompletionStuck + round], 1u);
                    if (idx < kStuckCap) stuck[idx] = make_uint4(m, o[0], o[1], o[2]);
                    atomicMin(&want[m], kWantRebuild);
                }
            }
        }
        waveSync();
    }
}

struct RebuildLds {
    CellLds cell;
    uint32_t order[kNeighbourCap];  // insertion order of the rebuild
    uint32_t ids[kNeighbourCap];    // stuck-record sites, then the other neighbours
    float key[kNeighbourCap];       // FP32 gap of the other neighbours
};

// Rebuild stuck cell m (see the file comment); the stuck list holds count
// <= kStuckCap records.  Returns false, leaving the cell as it was, when its
// stuck sites exceed the LDS list.
__device__ bool rebuildCell(const CellCtx& c, RebuildLds& S, WaveCell& w, uint32_t m, const uint4* stuck,
                            uint32_t count) {
    CellLds& L = S.cell;
    const uint32_t lane = laneId();
    const uint32_t n = c.n;
    // the sites of m's stuck records, with repeats
    uint32_t nAll = 0;
    for (uint32_t b = 0; b < count; b += 64u) {
        const uint32_t e = b + lane;
        uint4 r = make_uint4(kNone, 0u, 0u, 0u);
        if (e < count) r = stuck[e];
        const bool mine = r.x == m;
        const uint64_t hm = ballot(mine);
        const uint32_t pos = nAll + 3u * countBelow(hm);
        if (mine && pos + 3u <= kNeighbourCap) {
            S.ids[pos] = r.y;
            S.ids[pos + 1u] = r.z;
            S.ids[pos + 2u] = r.w;
        }
        nAll += 3u * popc64(hm);
    }
    if (nAll > kNeighbourCap) return false;
    waveSync();
    // distinct sites, ascending: first occurrences flagged, then ranked
    for (uint32_t e = lane; e < nAll; e += 64u) {
        const uint32_t x = S.ids[e];
        bool first = true;
        for (uint32_t u = 0; u < e; u++) first = first && S.ids[u] != x;
        S.key[e] = first ? 1.f : 0.f;
    }
    waveSync();
    uint32_t nFirst = 0;
    for (uint32_t b = 0; b < nAll; b += 64u) {
        const uint32_t e = b + lane;
        const bool keep = e < nAll && S.key[e] != 0.f;
        if (keep) {
            const uint32_t x = S.ids[e];
            uint32_t rank = 0;
            for (uint32_t u = 0; u < nAll; u++) rank += (S.key[u] != 0.f && S.ids[u] < x) ? 1u : 0u;
            S.order[rank] = x;
        }
        nFirst += popc64(ballot(keep));
    }
    waveSync();
    // the other neighbours: real, not buried, not among the stuck sites
    buildNeighbours(c, L, w);
    const F4 fm = c.sorted[m];
    uint32_t nRest = 0;
    for (uint32_t b = 0; b < w.F; b += 64u) {
        const uint32_t q = b + lane;
        bool take = false;
        uint32_t id = kNone;
        if (q < w.F) {
            id = L.nb[q];
            take = id < n && c.status[id] != kStatusBuried;
            for (uint32_t u = 0; u < nFirst && take; u++) take = S.order[u] != id;
        }
        const uint64_t tm = ballot(take);
        if (take) {
            S.ids[nRest + countBelow(tm)] = id;
            S.key[nRest + countBelow(tm)] = sphereDistF(fm, c.sorted[id]);
        }
        nRest += popc64(tm);
    }
    waveSync();
    for (uint32_t b = 0; b < nRest; b += 64u) {
        const uint32_t e = b + lane;
        if (e < nRest) {
            const float k = S.key[e];
            const uint32_t x = S.ids[e];
            uint32_t rank = 0;
            for (uint32_t u = 0; u < nRest; u++) {
                const float k2 = S.key[u];
                const uint32_t y = S.ids[u];
                rank += (k2 < k || (k2 == k && y < x)) ? 1u : 0u;
            }
            S.order[nFirst + rank] = x;
        }
    }
    waveSync();
    initCell(L, w, artificialBase(c, cellMolecule(c, m)));
    const D4 sm = loadSite(c, m);
    const uint32_t total = nFirst + nRest;
    for (uint32_t t = 0; t < total; t++) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(S.order[t]);
        updateCell(c, L, w, m, sm, j, loadSite(c, j));
    }
    return true;
}

// One wave per cell with a wanted site (update) or stuck (rebuild): store,
// back to validation.
__global__ __launch_bounds__(64) void k_consistency_fix(CellCtx c, uint32_t* __restrict__ want,
                                                        const uint4* __restrict__ stuck, uint32_t* counters,
                                                        uint32_t round) {
    __shared__ RebuildLds S;
    CellLds& L = S.cell;
    const uint32_t lane = laneId();
    uint32_t updated = 0, rebuilt = 0;
    const uint32_t count = __builtin_amdgcn_readfirstlane(counters[kCompletionStuck + round]);
    for (uint32_t m = blockIdx.x; m < c.n; m += gridDim.x) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(want[m]);
        if (j == kNone || j == kWantRebuilt) continue;
        WaveCell w;
        if (j == kWantRebuild) {
            bool done = false;
            if (count <= kStuckCap) {
                loadCell(c, L, w, m, true);
                done = rebuildCell(c, S, w, m, stuck, count);
                if (done) storeCell(c, L, w, m);
            }
            if (lane == 0) {
                want[m] = done ? kWantRebuilt : kNone;
                if (done) c.status[m] = kStatusInValidation;
            }
            updated += done ? 1u : 0u;
            rebuilt += done ? 1u : 0u;
            waveSync();
            continue;
        }
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
    if (lane == 0 && rebuilt) atomicAdd(&counters[kCompletionRebuilt], rebuilt);
}

hipError_t consistencyLaunch(const CellCtx& c, uint32_t* want, uint4* stuck, uint32_t* counters, uint32_t round,
                             uint32_t blocks, hipStream_t st) {
    if (round >= kMaxCompletionRounds) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_consistency_check, dim3(blocks), dim3(64), 0, st, c, want, stuck, counters, round);
    hipLaunchKernelGGL(k_consistency_fix, dim3(blocks), dim3(64), 0, st, c, want, stuck, counters, round);
    return hipGetLastError();
}

}  // namespace apo_hip
