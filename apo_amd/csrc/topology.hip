// topology.hip -- per-site topology of the built diagram on gfx950.
//
// Reference: AlgorithmGPU::toTopology (apo/include/apo/gpu/algorithm.inl:
// 1732-1821), which walks every cell on the host with std::set.  Here one
// wave per cell (sorted space) gathers the cell's candidate bisectors
// {i, j > i}, trisectors {i, j, k > i} and quadrisectors {i, j, k, l > i}
// (artificial sites excluded), keeps the first occurrence of each, ranks
// them in ascending order in LDS and writes them, mapped to original ids,
// at the cell's offset: the reference's output, element for element.
//
// The per-site cell records (apo_get_cell_topology) are gathered by
// k_cell_records into original-site order with original ids.
#include "apo_device.h"
#include "apo_internal.h"

namespace apo_hip {

struct TopoLds {
    uint32_t bis[2 * kEdgeCap];
    uint2 tri[kEdgeCap];
    uint4 quad[kVertexCap];
    uint8_t firstB[2 * kEdgeCap];
    uint8_t firstT[kEdgeCap];
    uint8_t firstQ[kVertexCap];
};

__device__ __forceinline__ bool lessT(uint2 a, uint2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); }
__device__ __forceinline__ bool eqT(uint2 a, uint2 b) { return a.x == b.x && a.y == b.y; }
__device__ __forceinline__ bool lessQ(uint4 a, uint4 b) {
    return a.x < b.x || (a.x == b.x && (a.y < b.y || (a.y == b.y && a.z < b.z)));
}
__device__ __forceinline__ bool eqQ(uint4 a, uint4 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Write == false: counts[3 * i + {0,1,2}] = distinct bisectors, trisectors,
// quadrisectors of cell i.  Write == true: the lists at offsets[3 * i + ...].
template <bool Write>
__global__ __launch_bounds__(64) void k_topology(CellCtx c, uint32_t* counts, const uint64_t* offsets,
                                                  const uint32_t* __restrict__ indices, uint32_t* bis, uint32_t* tri,
                                                  uint32_t* quad) {
    __shared__ TopoLds T;
    const uint32_t lane = laneId();
    const uint32_t i = blockIdx.x;
    const uint32_t n = c.n;
    const uint32_t E = min(c.enb[i], kEdgeCap), V = min(c.vnb[i], kVertexCap);
    const uint4* ge = c.edg + size_t(i) * kEdgeCap;
    const uint4* gv = c.vtx + size_t(i) * kVertexCap;
    // candidates in record order (the std::set keeps one of each)
    uint32_t nb = 0, nt = 0, nq = 0;
    for (uint32_t b = 0; b < E; b += 64) {
        const uint32_t e = b + lane;
        uint4 ed = {0, 0, 0, 0};
        if (e < E) ed = ge[e];
        const bool bj = e < E && i < ed.x && ed.x < n, bk = e < E && i < ed.y && ed.y < n;
        const bool t = bj && bk;
        const uint64_t mj = ballot(bj), mk = ballot(bk), mt = ballot(t);
        const uint32_t o = nb + countBelow(mj) + countBelow(mk);
        if (bj) T.bis[o] = ed.x;
        if (bk) T.bis[o + (bj ? 1u : 0u)] = ed.y;
        if (t) T.tri[nt + countBelow(mt)] = make_uint2(ed.x, ed.y);
        nb += popc64(mj) + popc64(mk);
        nt += popc64(mt);
    }
    for (uint32_t b = 0; b < V; b += 64) {
        const uint32_t v = b + lane;
        uint4 r = {0, 0, 0, 0};
        if (v < V) r = gv[v];
        const bool q = v < V && i < r.x && i < r.y && i < r.z && r.x < n && r.y < n && r.z < n;
        const uint64_t mq = ballot(q);
        if (q) T.quad[nq + countBelow(mq)] = make_uint4(r.x, r.y, r.z, 0u);
        nq += popc64(mq);
    }
    waveSync();
    // first occurrences
    uint32_t db = 0, dt = 0, dq = 0;
    for (uint32_t b = 0; b < nb; b += 64) {
        const uint32_t a = b + lane;
        bool f = a < nb;
        if (f)
            for (uint32_t a2 = 0; a2 < a; a2++)
                if (T.bis[a2] == T.bis[a]) {
                    f = false;
                    break;
                }
        if (a < nb) T.firstB[a] = f ? 1 : 0;
        db += popc64(ballot(f));
    }
    for (uint32_t b = 0; b < nt; b += 64) {
        const uint32_t a = b + lane;
        bool f = a < nt;
        if (f)
            for (uint32_t a2 = 0; a2 < a; a2++)
                if (eqT(T.tri[a2], T.tri[a])) {
                    f = false;
                    break;
                }
        if (a < nt) T.firstT[a] = f ? 1 : 0;
        dt += popc64(ballot(f));
    }
    for (uint32_t b = 0; b < nq; b += 64) {
        const uint32_t a = b + lane;
        bool f = a < nq;
        if (f)
            for (uint32_t a2 = 0; a2 < a; a2++)
                if (eqQ(T.quad[a2], T.quad[a])) {
                    f = false;
                    break;
                }
        if (a < nq) T.firstQ[a] = f ? 1 : 0;
        dq += popc64(ballot(f));
    }
    waveSync();
    if (!Write) {
        if (lane == 0) {
            counts[3 * size_t(i)] = db;
            counts[3 * size_t(i) + 1] = dt;
            counts[3 * size_t(i) + 2] = dq;
        }
        return;
    }
    // ascending rank among the distinct entries
    const uint64_t ob = offsets[3 * size_t(i)], ot = offsets[3 * size_t(i) + 1], oq = offsets[3 * size_t(i) + 2];
    const uint32_t oi = indices[i];
    for (uint32_t b = 0; b < nb; b += 64) {
        const uint32_t a = b + lane;
        if (a < nb && T.firstB[a]) {
            const uint32_t x = T.bis[a];
            uint32_t rank = 0;
            for (uint32_t a2 = 0; a2 < nb; a2++) rank += (T.firstB[a2] && T.bis[a2] < x) ? 1u : 0u;
            const size_t o = 2 * (ob + rank);
            bis[o] = oi;
            bis[o + 1] = indices[x];
        }
    }
    for (uint32_t b = 0; b < nt; b += 64) {
        const uint32_t a = b + lane;
        if (a < nt && T.firstT[a]) {
            const uint2 x = T.tri[a];
            uint32_t rank = 0;
            for (uint32_t a2 = 0; a2 < nt; a2++) rank += (T.firstT[a2] && lessT(T.tri[a2], x)) ? 1u : 0u;
            const size_t o = 3 * (ot + rank);
            tri[o] = oi;
            tri[o + 1] = indices[x.x];
            tri[o + 2] = indices[x.y];
        }
    }
    for (uint32_t b = 0; b < nq; b += 64) {
        const uint32_t a = b + lane;
        if (a < nq && T.firstQ[a]) {
            const uint4 x = T.quad[a];
            uint32_t rank = 0;
            for (uint32_t a2 = 0; a2 < nq; a2++) rank += (T.firstQ[a2] && lessQ(T.quad[a2], x)) ? 1u : 0u;
            const size_t o = 4 * (oq + rank);
            quad[o] = oi;
            quad[o + 1] = indices[x.x];
            quad[o + 2] = indices[x.y];
            quad[o + 3] = indices[x.z];
        }
    }
}

// Records of sorted cell i at the offsets of original site indices[i]:
// vertices (j, k, l, type) and edges (j, k, vertexNb, 0) with site ids in
// original numbering (artificial sites keep ids n..n+5).  Buried sites have
// no cell and no records.
__global__ __launch_bounds__(64) void k_cell_records(CellCtx c, const uint32_t* __restrict__ indices,
                                                     const uint64_t* __restrict__ vOff, const uint64_t* __restrict__ eOff,
                                                     uint4* vOut, uint4* eOut) {
    const uint32_t lane = laneId();
    const uint32_t i = blockIdx.x;
    const uint32_t n = c.n;
    if (c.status[i] == kStatusBuried) return;
    const uint32_t o = indices[i];
    auto map = [&](uint32_t id) { return id < n ? indices[id] : id; };
    const uint32_t V = min(c.vnb[i], kVertexCap), E = min(c.enb[i], kEdgeCap);
    const uint4* gv = c.vtx + size_t(i) * kVertexCap;
    const uint4* ge = c.edg + size_t(i) * kEdgeCap;
    for (uint32_t v = lane; v < V; v += 64) {
        const uint4 r = gv[v];
        vOut[vOff[o] + v] = make_uint4(map(r.x), map(r.y), map(r.z), r.w & 0xffu);
    }
    for (uint32_t e = lane; e < E; e += 64) {
        const uint4 r = ge[e];
        eOut[eOff[o] + e] = make_uint4(map(r.x), map(r.y), r.z, 0u);
    }
}

hipError_t topologyCountLaunch(const CellCtx& c, uint32_t* counts, hipStream_t st) {
    hipLaunchKernelGGL(k_topology<false>, dim3(c.n), dim3(64), 0, st, c, counts, nullptr, nullptr, nullptr, nullptr,
                       nullptr);
    return hipGetLastError();
}

hipError_t topologyWriteLaunch(const CellCtx& c, const uint64_t* offsets, const uint32_t* indices, uint32_t* bis,
                               uint32_t* tri, uint32_t* quad, hipStream_t st) {
    hipLaunchKernelGGL(k_topology<true>, dim3(c.n), dim3(64), 0, st, c, nullptr, offsets, indices, bis, tri, quad);
    return hipGetLastError();
}

hipError_t cellRecordsLaunch(const CellCtx& c, const uint32_t* indices, const uint64_t* vOff, const uint64_t* eOff,
                             uint4* vOut, uint4* eOut, hipStream_t st) {
    hipLaunchKernelGGL(k_cell_records, dim3(c.n), dim3(64), 0, st, c, indices, vOff, eOff, vOut, eOut);
    return hipGetLastError();
}

}  // namespace apo_hip
