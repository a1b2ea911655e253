// cells.hip -- per-site cell construction on gfx950, one wave64 per cell.
//
// Reference: apo/include/apo/gpu/topology_update.inl:49-653 (updateCell,
// cellRadius, the initial octahedron) and apo/include/apo/gpu/algorithm.inl:101-206
// (getCells), 1235-1307 (vertex output).
//
// Design (wave64, LDS-resident cell):
//  * the cell's vertex and edge records live in LDS for the whole update
//    sequence; HBM is touched once to read the neighbour list and once to
//    store the final cell;
//  * the validation predicate "no site on a cell edge intersects the
//    candidate" is evaluated over the set of distinct sites on the edges
//    (rebuilt per update), which gives the same answer as the reference's
//    walk over both endpoints of every edge with a third of the tests;
//  * lanes own edges; the serial parts of the reference (ballot-ordered
//    forEach) become ballot + mbcnt prefix sums, visited lowest lane first;
//  * a persistent grid pulls cells from an atomic queue in Morton order.
#include "apo_device.h"
#include "apo_internal.h"
#include "cell_update.h"

namespace apo_hip {

// Known radius of a range query of bound b that found fewer sites than its
// list holds: every site whose FP32 gap is below rangeTop(b, r_i) was
// listed, so every site whose exact gap is below it, less the FP32 rounding
// of a gap (a margin of 2^-18 of the magnitudes involved), is known.
__device__ __forceinline__ double rangeKnownRadius(const CellCtx& c, uint32_t i, float b, const D4& si) {
    const double top = double(rangeTop(b, float(si.w)));
    const double scale = __builtin_fabs(top) + __builtin_fabs(si.x) + __builtin_fabs(si.y) + __builtin_fabs(si.z) +
                         __builtin_fabs(si.w);
    return top - scale * 0x1p-18;
}

// getCells (algorithm.inl:101-206) for cell i: K updates in kNN order (the
// first `use` of the K-long list), then the security-radius certification.
// The first pass starts from the octahedron; a resumed pass (a later kNN
// pass) continues the stored cell and leaves certified and buried cells as
// they are.  W lanes work on the cell (a wave, or half a wave with the other
// half on another cell); the list is read W entries per lane step.
template <uint32_t W>
__device__ __forceinline__ void cellTask(const CellCtx& c, CellLds& L, const uint32_t* __restrict__ knn, uint32_t K,
                                         uint32_t use, uint32_t i, bool resume, const float* __restrict__ bound) {
    using G = Lanes<W>;
    constexpr uint32_t kChunks = 64u / W;
    const uint32_t lane = G::lane();
    const uint32_t n = c.n;
    WaveCell w;
    const uint8_t st = c.status[i];
    const bool active = !(st == kStatusCertified || st == kStatusBuried || K == 0);
    if (resume) {
        if (!active) return;
        loadCell<W>(c, L, w, i, true);
    } else {
        initCell<W>(L, w, artificialBase(c, cellMolecule(c, i)));
    }
    float kr = -FLT_MAX;
    bool certified = false;
    uint64_t readBytes = 1;  // status
    if (active) {
        const D4 si = loadSite(c, i);
        uint32_t myJ[kChunks];
        uint64_t buriedMask[kChunks], noneMask[kChunks];
        uint32_t count = 64u;
#pragma unroll
        for (uint32_t ch = 0; ch < kChunks; ch++) {
            const uint32_t slot = ch * W + lane;
            uint32_t j = kNone;
            bool buried = false;
            if (slot < K) {
                j = knn[size_t(i) * K + slot];
                if (j != kNone && j >= n) {
                    atomicOr(c.err, kErrCellNeighbour);
                    j = kNone;
                }
                if (j != kNone) buried = c.status[j] == kStatusBuried;
            }
            myJ[ch] = j;
            buriedMask[ch] = G::vote(buried);
            noneMask[ch] = G::vote(j == kNone);
            // neighbours are used up to the first empty slot
            if (count == 64u && noneMask[ch]) count = ch * W + uint32_t(__builtin_ctzll(noneMask[ch]));
        }
        const uint32_t limit = count < use ? count : use;
        double knownRadius = -DBL_MAX;
        uint32_t used = 0;
        for (uint32_t t = 0; t < limit; t++) {
            const uint32_t ch = t / W, tl = t % W;
            const uint64_t bm = kChunks == 1u ? buriedMask[0] : (ch ? buriedMask[kChunks - 1u] : buriedMask[0]);
            if ((bm >> tl) & 1u) continue;
            const uint32_t src = kChunks == 1u ? myJ[0] : (ch ? myJ[kChunks - 1u] : myJ[0]);
            const uint32_t j = G::read(src, tl);
            const D4 sj = loadSite(c, j);
            updateCell<W, true>(c, L, w, i, si, j, sj);
            knownRadius = __builtin_fmax(knownRadius, sphereDistD(si, sj));
            used++;
        }
        if (bound && count < K) knownRadius = __builtin_fmax(knownRadius, rangeKnownRadius(c, i, bound[i], si));
        const double maxR = cellRadius<W>(c, L, w, si);
        certified = gtD(knownRadius, 2.0 * maxR);
        kr = float(knownRadius);
        // own site, K neighbour ids and their status, the sites used
        readBytes += 16u + 5u * uint64_t(K) + 16u * uint64_t(used);
    }
    storeCell<W>(c, L, w, i);
    if (lane == 0) {
        if (certified) c.status[i] = kStatusCertified;
        c.knownRadii[i] = kr;
        // records, vertex and edge counts, known radius, status
        countBytes(c, kBytesCells, readBytes + 16u * uint64_t(w.V + w.E) + 13u);
    }
}

// Neighbour count of cell i: the list length K, or with cls (the learned
// proposal) the count of the cell's class, at most K.
__device__ __forceinline__ uint32_t cellUse(const uint8_t* __restrict__ cls, uint32_t i, uint32_t K,
                                            uint32_t fallbackK) {
    if (!cls) return K;
    const uint32_t q = cls[i];
    const uint32_t use = q < kProposalClasses ? proposalK(q, fallbackK) : K;
    return use > K ? K : use;
}

// Consecutive cells a wave of the cell pass claims per queue atomic: up to 8,
// while every wave still makes at least 32 claims (short passes keep claims
// of one cell, so the last cells spread over the waves).
__device__ __forceinline__ uint32_t cellBatch(uint32_t cells) {
    const uint32_t perWave = cells / (gridDim.x * 32u);
    return perWave < 1u ? 1u : (perWave > 8u ? 8u : perWave);
}

// Persistent cell pass: each wave takes cells from the XCD-ranged queue
// (Morton order within each range) until every cell is done.
__global__ __launch_bounds__(64) void k_cells(CellCtx c, const uint32_t* __restrict__ knn, uint32_t K,
                                              uint32_t* queue, const uint8_t* __restrict__ cls, uint32_t fallbackK,
                                              uint32_t resume, const float* __restrict__ bound) {
    __shared__ CellLds L;
    WorkClaim claim;
    const uint32_t batch = cellBatch(c.n);
    for (;;) {
        const uint32_t i = popWorkBatch(queue, c.n, claim, batch);
        if (i == kNone) break;
        cellTask<64>(c, L, knn, K, cellUse(cls, i, K, fallbackK), i, resume != 0u, bound);
        waveSync();
    }
}

// The same pass with two cells per wave, one per half: a cell update keeps
// fewer than 32 records busy on most of its steps, so each half of the
// wave's instructions works on a cell of its own.  Each half takes its own
// cell from the queue; the records are those of k_cells.
__global__ __launch_bounds__(64) void k_cells_pair(CellCtx c, const uint32_t* __restrict__ knn, uint32_t K,
                                                   uint32_t* queue, const uint8_t* __restrict__ cls,
                                                   uint32_t fallbackK, uint32_t resume,
                                                   const float* __restrict__ bound) {
    __shared__ CellLds L[2];
    const uint32_t half = laneId() >> 5;
    WorkClaim claim;
    const uint32_t batch = cellBatch(c.n);
    for (;;) {
        const uint32_t i0 = popWorkBatch(queue, c.n, claim, batch);
        if (i0 == kNone) break;
        const uint32_t i1 = popWorkBatch(queue, c.n, claim, batch);
        const uint32_t i = half ? i1 : i0;
        if (i != kNone) cellTask<32>(c, L[half], knn, K, cellUse(cls, i, K, fallbackK), i, resume != 0u, bound);
        waveSync();
    }
}

// No kNN pass (knnValidationStep = 0, algorithm.inl:1526-1549): every list
// empty, every cell in validation with no flags (the cell launch that
// follows writes the octahedra and the lowest known radius).
__global__ void k_reset_cells(CellCtx c, uint32_t* __restrict__ knn, uint32_t K) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < c.n) {
        c.status[t] = kStatusInValidation;
        c.flags[t] = 0u;
    }
    if (t < size_t(c.n) * K) knn[t] = kNone;
}

hipError_t cellsResetLaunch(const CellCtx& c, uint32_t* knn, uint32_t k, hipStream_t st) {
    const size_t words = size_t(c.n) * (k > 0u ? k : 1u);
    hipLaunchKernelGGL(k_reset_cells, dim3(uint32_t((words + 255u) / 256u)), dim3(256), 0, st, c, knn, k);
    return hipGetLastError();
}

// The double rounded up to a float (never below it).
__device__ __forceinline__ float roundUpF(double d) {
    float f = float(d);
    if (double(f) < d) {
        const uint32_t b = __builtin_bit_cast(uint32_t, f);
        f = f == 0.f ? __builtin_bit_cast(float, 1u) : __builtin_bit_cast(float, f > 0.f ? b + 1u : b - 1u);
    }
    return f;
}

// cellRadius (topology_update.inl:595-653) of every stored cell, rounded up
// to FP32; -FLT_MAX for buried cells.  One wave per cell.
__global__ __launch_bounds__(64) void k_cell_radii(CellCtx c, float* __restrict__ out) {
    __shared__ CellLds L;
    for (uint32_t i = blockIdx.x; i < c.n; i += gridDim.x) {
        const uint8_t st = c.status[i];
        float r = -FLT_MAX;
        if (st != kStatusBuried) {
            WaveCell w;
            loadCell(c, L, w, i, true);
            r = roundUpF(cellRadius(c, L, w, loadSite(c, i)));
        }
        if (laneId() == 0) out[i] = r;
        waveSync();
    }
}

hipError_t cellRadiiLaunch(const CellCtx& c, float* out, uint32_t blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_cell_radii, dim3(blocks), dim3(64), 0, st, c, out);
    return hipGetLastError();
}

hipError_t cellsOccupancy(uint32_t* blocksPerCu, uint32_t* pairBlocksPerCu) {
    int b = 0, p = 0;
    hipError_t r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_cells, 64, 0);
    if (r == hipSuccess) r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, k_cells_pair, 64, 0);
    *blocksPerCu = uint32_t(b > 0 ? b : 1);
    *pairBlocksPerCu = uint32_t(p > 0 ? p : 1);
    return r;
}

hipError_t cellsLaunch(const CellCtx& c, const uint32_t* knn, uint32_t k, uint32_t* queue, uint32_t blocks,
                       hipStream_t st, const uint8_t* cls, uint32_t fallbackK, bool resume, const float* bound,
                       bool pair) {
    if (k > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, queue, kXcdQueueWords, 0u);
    if (pair)
        hipLaunchKernelGGL(k_cells_pair, dim3(blocks), dim3(64), 0, st, c, knn, k, queue, cls, fallbackK,
                           resume ? 1u : 0u, bound);
    else
        hipLaunchKernelGGL(k_cells, dim3(blocks), dim3(64), 0, st, c, knn, k, queue, cls, fallbackK,
                           resume ? 1u : 0u, bound);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Output: count -> scan -> write (algorithm.inl:1235-1307, 1596-1646)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool emitted(uint32_t i, uint4 r, uint32_t n) {
    return !(i > r.x || r.x >= n || r.y >= n || r.z >= n);
}

__global__ __launch_bounds__(256) void k_out_count(CellCtx c, uint32_t* cellCount, uint32_t* blockSums) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    uint32_t cnt = 0;
    if (i < c.n && c.status[i] != kStatusBuried) {
        const uint4* gv = c.vtx + size_t(i) * kVertexCap;
        const uint32_t V = c.vnb[i];
        for (uint32_t v = 0; v < V; v++) cnt += emitted(i, gv[v], c.n) ? 1u : 0u;
    }
    if (i < c.n) cellCount[i] = cnt;
    __shared__ uint32_t s[4];
    const uint32_t sum = waveSumU(cnt);
    if (laneId() == 0) s[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) blockSums[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ __launch_bounds__(1024) void k_out_scan(uint32_t* blockSums, uint32_t blocks, uint32_t* total) {
    __shared__ uint32_t s[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < blocks; b += 1024) {
        const uint32_t idx = b + threadIdx.x;
        const uint32_t v = idx < blocks ? blockSums[idx] : 0u;
        const uint32_t inc = waveInclusiveSum(v);
        if (laneId() == 63) s[threadIdx.x >> 6] = inc;
        __syncthreads();
        uint32_t off = carry;
        for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) off += s[w];
        if (idx < blocks) blockSums[idx] = off + inc - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = off + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_out_write(CellCtx c, const uint32_t* cellCount, const uint32_t* blockOffsets,
                                                   const uint32_t* __restrict__ indices, D4* xyzw, uint4* ids,
                                                   uint64_t capacity) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t cnt = i < c.n ? cellCount[i] : 0u;
    const uint32_t inc = waveInclusiveSum(cnt);
    __shared__ uint32_t s[4];
    if (laneId() == 63) s[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t off = blockOffsets[blockIdx.x];
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) off += s[w];
    off += inc - cnt;
    if (i >= c.n || cnt == 0) return;
    const D4 si = loadSite(c, i);
    const uint4* gv = c.vtx + size_t(i) * kVertexCap;
    const uint32_t V = c.vnb[i];
    for (uint32_t v = 0; v < V; v++) {
        const uint4 r = gv[v];
        if (!emitted(i, r, c.n)) continue;
        if (off < capacity) {
            xyzw[off] = vertexPosition(c, si, r);
            ids[off] = make_uint4(indices[i], indices[r.x], indices[r.y], indices[r.z]);
        }
        off++;
    }
}

hipError_t scanBlockSumsLaunch(uint32_t* blockSums, uint32_t blocks, uint32_t* total, hipStream_t st) {
    hipLaunchKernelGGL(k_out_scan, dim3(1), dim3(1024), 0, st, blockSums, blocks, total);
    return hipGetLastError();
}

hipError_t outputCountLaunch(const CellCtx& c, const OutputBuffers& o, hipStream_t st) {
    const uint32_t blocks = (c.n + 255u) / 256u;
    hipLaunchKernelGGL(k_out_count, dim3(blocks), dim3(256), 0, st, c, o.cellCount, o.blockSums);
    hipLaunchKernelGGL(k_out_scan, dim3(1), dim3(1024), 0, st, o.blockSums, blocks, o.total);
    return hipGetLastError();
}

hipError_t outputWriteLaunch(const CellCtx& c, const OutputBuffers& o, const uint32_t* indices, hipStream_t st) {
    const uint32_t blocks = (c.n + 255u) / 256u;
    hipLaunchKernelGGL(k_out_write, dim3(blocks), dim3(256), 0, st, c, o.cellCount, o.blockSums, indices, o.xyzw,
                       o.ids, o.capacity);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Closed trisector edges of the full diagram: count -> scan -> write
// (getWriteableClosedEdges / saveClosedEdges, algorithm.inl:1309-1383).  An
// edge record (j < k, kNone vertices) is a closed elliptic trisector; the
// cell with the smallest id emits it (i <= j, both real sites), at the
// extreme point of smaller empty-sphere radius.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool closedEmitted(uint32_t i, uint4 e, uint32_t n) {
    return !(e.z != kNone || i > e.x || e.x >= n || e.y >= n);
}

__global__ __launch_bounds__(256) void k_closed_count(CellCtx c, uint32_t* cellCount, uint32_t* blockSums) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    uint32_t cnt = 0;
    if (i < c.n && c.status[i] != kStatusBuried) {
        const uint4* ge = c.edg + size_t(i) * kEdgeCap;
        const uint32_t E = min(c.enb[i], kEdgeCap);
        for (uint32_t e = 0; e < E; e++) cnt += closedEmitted(i, ge[e], c.n) ? 1u : 0u;
    }
    if (i < c.n) cellCount[i] = cnt;
    __shared__ uint32_t s[4];
    const uint32_t sum = waveSumU(cnt);
    if (laneId() == 0) s[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) blockSums[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ __launch_bounds__(256) void k_closed_write(CellCtx c, const uint32_t* cellCount,
                                                      const uint32_t* blockOffsets,
                                                      const uint32_t* __restrict__ indices, D4* xyzw, uint4* ids,
                                                      uint64_t capacity) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t cnt = i < c.n ? cellCount[i] : 0u;
    const uint32_t inc = waveInclusiveSum(cnt);
    __shared__ uint32_t s[4];
    if (laneId() == 63) s[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t off = blockOffsets[blockIdx.x];
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) off += s[w];
    off += inc - cnt;
    if (i >= c.n || cnt == 0) return;
    const D4 si = loadSite(c, i);
    const uint4* ge = c.edg + size_t(i) * kEdgeCap;
    const uint32_t E = min(c.enb[i], kEdgeCap);
    for (uint32_t e = 0; e < E; e++) {
        const uint4 r = ge[e];
        if (!closedEmitted(i, r, c.n)) continue;
        // the reference leaves both extreme points uninitialised when the
        // solve has fewer than two; here they start at zero
        D4 x0 = {0.0, 0.0, 0.0, 0.0}, x1 = {0.0, 0.0, 0.0, 0.0};
        trisector(si, loadSite(c, r.x), loadSite(c, r.y), x0, x1);
        if (x1.w < x0.w) x0 = x1;
        if (off < capacity) {
            xyzw[off] = x0;
            ids[off] = make_uint4(indices[i], indices[r.x], indices[r.y], 0u);
        }
        off++;
    }
}

hipError_t closedEdgesCountLaunch(const CellCtx& c, const OutputBuffers& o, hipStream_t st) {
    const uint32_t blocks = (c.n + 255u) / 256u;
    hipLaunchKernelGGL(k_closed_count, dim3(blocks), dim3(256), 0, st, c, o.cellCount, o.blockSums);
    hipLaunchKernelGGL(k_out_scan, dim3(1), dim3(1024), 0, st, o.blockSums, blocks, o.total);
    return hipGetLastError();
}

hipError_t closedEdgesWriteLaunch(const CellCtx& c, const OutputBuffers& o, const uint32_t* indices,
                                  hipStream_t st) {
    const uint32_t blocks = (c.n + 255u) / 256u;
    hipLaunchKernelGGL(k_closed_write, dim3(blocks), dim3(256), 0, st, c, o.cellCount, o.blockSums, indices, o.xyzw,
                       o.ids, o.capacity);
    return hipGetLastError();
}

// Vertex offsets of the molecules of a batch: the exclusive prefix of the
// per-cell vertex counts at each molecule's first cell (after the scan,
// blockSums holds the exclusive prefix of each 256-cell block).
__global__ __launch_bounds__(64) void k_molecule_vertices(CellCtx c, const uint32_t* __restrict__ cellCount,
                                                          const uint32_t* __restrict__ blockOffsets,
                                                          const uint32_t* __restrict__ total, uint64_t* out) {
    const uint32_t m = blockIdx.x;
    if (m == c.molecules) {
        if (laneId() == 0) out[m] = *total;
        return;
    }
    const uint32_t first = c.offsets[m];
    const uint32_t b0 = first & ~255u;
    uint32_t part = 0;
    for (uint32_t i = b0 + laneId(); i < first; i += 64) part += cellCount[i];
    const uint32_t sum = waveSumU(part);
    if (laneId() == 0) out[m] = uint64_t(blockOffsets[first / 256u]) + sum;
}

hipError_t moleculeVerticesLaunch(const CellCtx& c, const OutputBuffers& o, uint64_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_molecule_vertices, dim3(c.molecules + 1), dim3(64), 0, st, c, o.cellCount, o.blockSums,
                       o.total, out);
    return hipGetLastError();
}

}  // namespace apo_hip
