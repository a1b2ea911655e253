// apo_internal.h -- buffers and launchers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "apo_device.h"

#ifndef DBL_MAX
#define DBL_MAX __DBL_MAX__
#endif
#ifndef FLT_MAX
#define FLT_MAX __FLT_MAX__
#endif

namespace apo_hip {

// Radix passes of the Morton sort: 8 over the 64-bit key (Morton code in
// bits 0-59, the molecule's low 4 bits in 60-63), then up to 2 over the
// molecule index's higher bits for batches of more than 16 molecules.
constexpr int kRadixPasses = 8;
constexpr int kMaxRadixPasses = 10;
constexpr uint32_t kMaxMolecules = 1u << 20;
constexpr uint32_t kBoxBlocks = 512;

// Per-cell capacities (apo/include/apo/gpu/algorithm.inl:25-30).
constexpr uint32_t kVertexCap = 152;
constexpr uint32_t kEdgeCap = 264;
constexpr uint32_t kNewEdgeCap = 128;
constexpr uint32_t kBisectorCap = 512;
constexpr uint32_t kNeighbourCap = 2 * kEdgeCap;
constexpr uint32_t kStackDepth = 32;

// Scene box + artificial sites, produced on the device.
struct Scene {
    double mn[3], mx[3];
    double rmin, rmax;
    float fmn[3], fmx[3];
    float scale[3];
    float pad;
    D4 art[6];
};

// Sort passes for a batch of m molecules.
inline int radixPassesFor(uint32_t m) { return m <= 16u ? 8 : m <= 4096u ? 9 : 10; }

// Molecule of a site index: the largest m with offsets[m] <= i.  Sites of a
// molecule are contiguous both in input order and in sorted order.
__host__ __device__ __forceinline__ uint32_t moleculeOf(const uint32_t* offsets, uint32_t molecules, uint32_t i) {
    uint32_t lo = 0, hi = molecules;  // offsets[lo] <= i < offsets[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Segments of a build: one molecule (offsets == nullptr) or a batch.  The
// nodes of molecule m are internal [offsets[m] - m, offsets[m+1] - m - 1)
// (its root first) and leaves internal + [offsets[m], offsets[m+1]), with
// internal = n - molecules; one molecule is the reference's numbering.
struct Segments {
    uint32_t molecules;
    const uint32_t* offsets;  // molecules + 1 (device), nullptr for one molecule
};

struct LbvhBuffers {
    Scene* scene;             // one per molecule
    double* boxPartials;   // kBoxBlocks * 8
    uint32_t* boxTicket;   // 1
    uint32_t* sortScratch; // hist [P*256] + tile counters [P] + look-back [P*tiles*256]
    uint64_t* codesA;
    uint64_t* codesB;
    uint32_t* idxA;        // sorted position -> original index (final)
    uint32_t* idxB;
    F4* sorted;            // sites in Morton order
    Node* nodes;           // n-1 internal nodes
    int32_t* parent;       // 2n-1
    uint32_t* refitCounter;// n-1
    float* rootBox;        // 7 floats: min, max, max radius
    uint32_t* err;         // device guard bits
    uint32_t* range;       // 2(n-1): first and last leaf of each internal node
    uint32_t* startLists;  // refitStartWords(n): climb starts of the later refit phases
};

// Refit phases (lbvh.hip): phase 0 over blocks of 2^kRefitShift leaves,
// phase p over blocks of 2^(kRefitShift + kRefitFanShift * p) leaves.
constexpr uint32_t kRefitShift = 9;
constexpr uint32_t kRefitLeaves = 1u << kRefitShift;
constexpr uint32_t kRefitFanShift = 6;
constexpr uint32_t kRefitLevels = 4;  // levels 0..3 cover every 32-bit leaf range

// Start lists of phases 1..kRefitLevels: level L holds one slot per block
// boundary of phase L (slot = boundary >> shift), ~0 when no climb starts
// there.
struct RefitStarts {
    uint32_t* list;
    uint32_t n;
};
__host__ __device__ inline uint32_t refitSlots(uint32_t n, uint32_t level) {
    return n > 1u ? ((n - 1u) >> (kRefitShift + kRefitFanShift * level)) + 1u : 1u;
}
__host__ __device__ inline uint32_t refitSlotOffset(uint32_t n, uint32_t level) {
    uint32_t o = 0;
    for (uint32_t l = 0; l < level; l++) o += refitSlots(n, l);
    return o;
}
__host__ __device__ inline uint32_t refitStartWords(uint32_t n) { return refitSlotOffset(n, kRefitLevels); }

// Phases after the first: enough that the last one's block holds every leaf.
inline uint32_t refitPhases(uint32_t n) {
    if (n < 2u || ((n - 1u) >> kRefitShift) == 0u) return 0u;
    const uint32_t hb = 31u - uint32_t(__builtin_clz(n - 1u));
    return (hb - kRefitShift) / kRefitFanShift + 1u;
}

size_t sortScratchWords(uint32_t n);
// Deepest leaf of the trees into *height (read by validation); diff and
// far: n words of scratch each, blocks: n / 4096 + 1 pairs of scratch.
hipError_t treeHeightLaunch(const uint32_t* range, uint32_t n, uint32_t internal, int32_t* diff, int32_t* far,
                            int2* blocks, uint32_t* height, hipStream_t st);
hipError_t lbvhBuild(const LbvhBuffers& b, const F4* sitesIn, uint32_t n, const Segments& seg, hipStream_t st);

// Device view of the construction state.
struct CellCtx {
    uint32_t n;             // real sites of all molecules: ids [0, n); artificial ids n + 6m + [0, 6)
    uint32_t internal;      // internal nodes (n - molecules); leaf j is node internal + j
    uint32_t molecules;
    const uint32_t* offsets;  // molecules + 1 site offsets (nullptr for one molecule)
    const F4* sorted;
    const Scene* scene;     // per molecule; artificial sites in scene[m].art
    const Node* nodes;
    const int32_t* parent;  // parent of every node (-1 at a root)
    const uint32_t* treeHeight;  // deepest leaf of the trees: a traversal stack never grows past it
    uint8_t* status;
    float* knownRadii;
    uint32_t* flags;
    uint4* vtx;             // n * kVertexCap records (j, k, l, type | status << 8)
    uint32_t* vnb;
    uint4* edg;             // n * kEdgeCap records (j, k, vertexNb, status)
    uint32_t* enb;
    uint32_t* err;          // device guard bits
    unsigned long long* bytes;  // algorithmic HBM bytes per kernel family, kByteSlots counters each
    float coverRatio;       // edge searches with a sphere above coverRatio x known radius use the edge cover
    uint32_t complete;      // completion mode: hyperbolic edge searches prune only boxes wholly outside a half-space
};

__device__ __forceinline__ uint32_t nodeCount(const CellCtx& c) { return c.n + c.internal; }
__device__ __forceinline__ uint32_t cellMolecule(const CellCtx& c, uint32_t i) {
    return c.molecules == 1u ? 0u : moleculeOf(c.offsets, c.molecules, i);
}
// root of molecule m's tree
__device__ __forceinline__ uint32_t rootNode(const CellCtx& c, uint32_t m) {
    return m == 0u ? 0u : c.offsets[m] - m;
}
// id of molecule m's first artificial site
__device__ __forceinline__ uint32_t artificialBase(const CellCtx& c, uint32_t m) { return c.n + 6u * m; }

// Byte counters: family * kByteSlots + (block & (kByteSlots - 1)), summed on the host.
constexpr uint32_t kByteSlots = 32;
enum ByteFamily : uint32_t { kBytesCells = 0, kBytesVertexVal = 1, kBytesEdgeVal = 2, kBytesBisectorVal = 3,
                             kByteFamilies = 4 };
__device__ __forceinline__ void countBytes(const CellCtx& c, uint32_t family, uint64_t b) {
    atomicAdd(&c.bytes[family * kByteSlots + (blockIdx.x & (kByteSlots - 1u))], static_cast<unsigned long long>(b));
}

// One kNN pass (k in 1..64 neighbours per site); pass > 0 continues the
// lists of the previous pass in place (knnValidationStep, algorithm.inl:43-60).
// bound (sorted order, may be null, pass 0 only): per-site upper bound on the
// gap of the listed sites (a range query of at most k sites).
hipError_t knnLaunch(const CellCtx& c, uint32_t k, uint32_t* knn, hipStream_t st, uint32_t pass = 0u,
                     const float* bound = nullptr);
// No kNN pass: empty lists (n * k words), status in validation, flags clear.
hipError_t cellsResetLaunch(const CellCtx& c, uint32_t* knn, uint32_t k, hipStream_t st);
// Blocks per CU of the cell pass with one cell per wave and with two.
hipError_t cellsOccupancy(uint32_t* blocksPerCu, uint32_t* pairBlocksPerCu);
// cellRadius of every stored cell (sorted order), rounded up to FP32; -FLT_MAX when buried.
hipError_t cellRadiiLaunch(const CellCtx& c, float* out, uint32_t blocks, hipStream_t st);
// Persistent cell pass over all sites; queue = kXcdQueueWords heads.
// cls (may be null): learned-proposal classes in sorted order; each cell then
// runs min(proposalK(class, fallbackK), k) neighbours.  resume: the cells
// still in validation continue from their stored records (a later kNN pass)
// instead of the octahedron, and the others are left untouched.
// bound (sorted order, may be null): the lists are range queries below these
// gaps (knnLaunch with bound); a cell whose list is not full knows every site
// below its bound, which then serves as its known radius.
// pair: two cells per wave (half a wave each), the same records.
hipError_t cellsLaunch(const CellCtx& c, const uint32_t* knn, uint32_t k, uint32_t* queue, uint32_t blocks,
                       hipStream_t st, const uint8_t* cls = nullptr, uint32_t fallbackK = 16u,
                       bool resume = false, const float* bound = nullptr, bool pair = false);

// Device counters of a build (words of one buffer): work-queue heads of the
// persistent kernels (8 per queue, one range per XCD), the task count of
// the current validation round and the task count of every round.
constexpr uint32_t kXcdQueueWords = 8;
constexpr uint32_t kCellQueue = 0;
constexpr uint32_t kTaskCount = 8;
constexpr uint32_t kVertexQueue = 16;
constexpr uint32_t kEdgeQueue = 24;
constexpr uint32_t kRoundWords = kEdgeQueue + kXcdQueueWords - kTaskCount;  // cleared per round
constexpr uint32_t kBisectorQueue = 32;
constexpr uint32_t kTreeHeight = 40;   // deepest leaf of the build's trees (root = 0)
constexpr uint32_t kRoundTasks = 64;
constexpr uint32_t kMaxRounds = 256;
// completion mode: cells updated by each consistency round, the missing
// records whose cell knows all their sites found by each round (the append
// index of the round's stuck-record list), then the task counts of its
// validation rounds (after the bisector stage, then after each consistency
// round: kMaxRounds each), then the cells rebuilt over all rounds
constexpr uint32_t kMaxCompletionRounds = 8;
constexpr uint32_t kCompletionUpdated = kRoundTasks + kMaxRounds;
constexpr uint32_t kCompletionStuck = kCompletionUpdated + kMaxCompletionRounds;
constexpr uint32_t kCompletionTasks = kCompletionStuck + kMaxCompletionRounds;
constexpr uint32_t kCompletionRebuilt = kCompletionTasks + (1 + kMaxCompletionRounds) * kMaxRounds;
// the bisector stage split by search: cells marked for the per-cell stage
constexpr uint32_t kBisMarked = kCompletionRebuilt + 1;
constexpr uint32_t kCounterWords = kBisMarked + 1;
// A round's list of stuck records (cell, the record's other three sites):
// a round that finds more rebuilds no cell.  want[] codes besides site ids:
constexpr uint32_t kStuckCap = 16384;
constexpr uint32_t kWantRebuilt = 0xfffffffdu;  // rebuilt in this build: not again
constexpr uint32_t kWantRebuild = 0xfffffffeu;  // stuck: rebuild in this round

static __global__ void k_fill_words(uint32_t* p, uint32_t count, uint32_t value) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

// Work queue over [0, count) for a persistent grid: 8 contiguous ranges,
// range x served first to the blocks of XCD x (the dispatcher hands blocks
// to the 8 XCDs round-robin), so each XCD's L2 works through one spatially
// coherent run of cells; a wave whose range is drained takes work from the
// following ranges.  Returns the same index to every lane, kNone when all
// ranges are drained.
__device__ __forceinline__ uint32_t popWork(uint32_t* heads, uint32_t count) {
    constexpr uint32_t kXcd = kXcdQueueWords;
    uint32_t r = kNone;
    if (laneId() == 0) {
        const uint32_t home = blockIdx.x % kXcd;
        const uint32_t base = count / kXcd, extra = count % kXcd;
        for (uint32_t k = 0; k < kXcd; k++) {
            const uint32_t x = (home + k) % kXcd;
            const uint32_t len = base + (x < extra ? 1u : 0u);
            if (__hip_atomic_load(&heads[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) continue;
            const uint32_t t = atomicAdd(&heads[x], 1u);
            if (t < len) {
                r = x * base + (x < extra ? x : extra) + t;
                break;
            }
        }
    }
    return __builtin_amdgcn_readfirstlane(r);
}

// A wave's claim on a work queue: items [cur, end) taken but not yet served.
struct WorkClaim {
    uint32_t cur = 0, end = 0;
};

// popWork taking up to `batch` consecutive items of a range per atomic: the
// queue's heads are device-scope atomics shared by every XCD, whose
// throughput, not the kernels, bounds a grid that takes one cell at a time
// (a per-cell launch of a million cheap cells spent most of its time here).
// Returns the next item of the wave's claim, claiming more when it is
// served; kNone when every range is drained.
__device__ __forceinline__ uint32_t popWorkBatch(uint32_t* heads, uint32_t count, WorkClaim& w, uint32_t batch) {
    if (w.cur < w.end) return w.cur++;
    constexpr uint32_t kXcd = kXcdQueueWords;
    uint32_t first = kNone, stop = kNone;
    if (laneId() == 0) {
        const uint32_t home = blockIdx.x % kXcd;
        const uint32_t base = count / kXcd, extra = count % kXcd;
        for (uint32_t k = 0; k < kXcd; k++) {
            const uint32_t x = (home + k) % kXcd;
            const uint32_t len = base + (x < extra ? 1u : 0u);
            if (__hip_atomic_load(&heads[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) continue;
            const uint32_t t = atomicAdd(&heads[x], batch);
            if (t < len) {
                const uint32_t start = x * base + (x < extra ? x : extra);
                first = start + t;
                stop = start + (len - t < batch ? len : t + batch);
                break;
            }
        }
    }
    first = __builtin_amdgcn_readfirstlane(first);
    stop = __builtin_amdgcn_readfirstlane(stop);
    if (first == kNone) return kNone;
    w.cur = first + 1u;
    w.end = stop;
    return first;
}

// Validation (algorithm.inl:208-1233).  Task lists live on the device.
struct ValidationBuffers {
    uint32_t* tasks;       // n
    uint32_t* counters;    // kCounterWords
    uint32_t* spill;       // spillBlocks * waveScratchWords(): per-wave deep stack slots and edge cover buffers
    uint32_t spillBlocks;  // the most blocks a vertex or edge validation grid may have
};
// Words of device memory per validation wave (deep stack slots of the vertex
// and edge searches, then the edge cover buffers).
uint32_t waveScratchWords();
// Persistent grid sizes (blocks of one wave); bisectorPlan and
// bisectorLanes: the grids of the split bisector stage (bisectorLanes 0: the
// stage runs per cell only).
struct ValidationGrid {
    uint32_t vertex, edge, bisector, bisectorPlan, bisectorLanes;
};
// Slots of the bisector stage split by search (validationLaunch), per cell:
// its searches' bisector sites, its distinct neighbours, the counts (searches
// | neighbours << 16) and a mark (1: the per-cell stage runs it).  A cell
// with more searches or neighbours than slots is marked.
constexpr uint32_t kBisSearchSlots = 32;
constexpr uint32_t kBisNeighbourSlots = 64;
struct BisectorSplit {
    uint32_t* search;  // n * kBisSearchSlots
    uint32_t* nb;      // n * kBisNeighbourSlots
    uint32_t* count;   // n
    uint8_t* mark;     // n
};
// Events around each vertex / edge validation launch (rounds beyond
// kTimedRounds run untimed).
constexpr uint32_t kTimedRounds = 16;
struct ValidationTiming {
    hipEvent_t ev[kTimedRounds][3];  // before vertex, between, after edge
    uint32_t rounds;                 // rounds with events
};
hipError_t validationOccupancy(uint32_t* vertexBlocks, uint32_t* edgeBlocks, uint32_t* bisectorBlocks,
                               uint32_t* bisectorPlanBlocks, uint32_t* bisectorLaneBlocks);
// roundTasks: one word per round for its task count (vertex + edge rounds).
// bisector: the bisector stage instead, split by search when `split` holds
// buffers and g.bisectorLanes is set, else one wave per cell.
hipError_t validationLaunch(const CellCtx& c, const ValidationBuffers& v, const ValidationGrid& g,
                            uint32_t iterations, uint32_t maxRestarts, bool bisector, hipStream_t st,
                            ValidationTiming* tm, uint32_t* roundTasks = nullptr,
                            const BisectorSplit* split = nullptr);
// One consistency round of the completion mode (complete.hip): want holds n
// words (kNone or kWantRebuilt between rounds, kNone at the build's start),
// stuck kStuckCap records; round < kMaxCompletionRounds.
hipError_t consistencyLaunch(const CellCtx& c, uint32_t* want, uint4* stuck, uint32_t* counters, uint32_t round,
                             uint32_t blocks, hipStream_t st);

// Output (algorithm.inl:1235-1307): count, scan, write.
struct OutputBuffers {
    uint32_t* cellCount;   // n
    uint32_t* blockSums;   // blocks
    uint32_t* total;       // 1
    D4* xyzw;              // capacity records
    uint4* ids;            // capacity records
    uint64_t capacity;
};
hipError_t outputCountLaunch(const CellCtx& c, const OutputBuffers& o, hipStream_t st);
hipError_t outputWriteLaunch(const CellCtx& c, const OutputBuffers& o, const uint32_t* indices, hipStream_t st);
// Closed trisector edges of the full diagram (algorithm.inl:1309-1383):
// count into cellCount / blockSums / total, then write at most capacity.
hipError_t closedEdgesCountLaunch(const CellCtx& c, const OutputBuffers& o, hipStream_t st);
hipError_t closedEdgesWriteLaunch(const CellCtx& c, const OutputBuffers& o, const uint32_t* indices,
                                  hipStream_t st);
// Vertex offsets of each molecule of a batch (molecules + 1 values).
hipError_t moleculeVerticesLaunch(const CellCtx& c, const OutputBuffers& o, uint64_t* out, hipStream_t st);

// Learned proposal (learned.hip): 8 neighbour-count class logits and one
// search-radius output, log2 of the radius over d16 (the centre distance of
// the 16th nearest neighbour)
constexpr uint32_t kFeatures = 52, kHidden = 64, kProposalClasses = 8;
constexpr uint32_t kProposalOutputs = kProposalClasses + 1;
constexpr uint32_t kRadiusOutput = kProposalClasses;
constexpr uint32_t kProposalParams = kFeatures * kHidden + kHidden + kHidden * kHidden + kHidden +
                                     kHidden * kProposalOutputs + kProposalOutputs;
// weight of the radius head's squared error beside the cross-entropy
constexpr float kRadiusLossWeight = 0.25f;
constexpr uint32_t kLabelBuried = 255;
enum CellsMode : uint32_t { kProposal = 0, kLabels = 1 };
// Neighbours a class asks for; the last class (not certified within 64)
// runs `fallback` (16 by default) and goes to validation.
__host__ __device__ constexpr uint32_t proposalK(uint32_t c, uint32_t fallback = 16u) {
    return c == 0 ? 20u : c == 1 ? 24u : c == 2 ? 28u : c == 3 ? 32u : c == 4 ? 40u : c == 5 ? 48u : c == 6 ? 64u
                                                                                                       : fallback;
}
// Neighbour-list length of the learned path: the most neighbours a class may run.
inline uint32_t proposalListK(uint32_t maxK, uint32_t fallbackK) { return maxK > fallbackK ? maxK : fallbackK; }
// Class of a label (neighbours used at certification, 0 = never within 64).
__host__ __device__ constexpr uint32_t proposalClass(uint32_t label) {
    if (label == 0) return kProposalClasses - 1;
    for (uint32_t c = 0; c + 1 < kProposalClasses; c++)
        if (label <= proposalK(c)) return c;
    return kProposalClasses - 1;
}
struct ProposalWeights {
    uint16_t *w1, *w2, *w3;  // BF16, transposed [out][64]
    float *b1, *b2, *b3;
};
// Training state: all features and classes on the device, per-wave partial
// gradients and losses, master weights (w), their BF16-rounded copy (wq,
// read by the backward pass), Adam moments and the packed forward weights.
struct TrainBuffers {
    const float* feat;
    const uint8_t* cls;
    const float* radius;  // radius-head targets per row (NaN: none), or null
    float *w, *wq, *m, *v;
    float* partial;      // maxWaves x kProposalParams
    float* lossPartial;  // maxWaves
    float* epochLoss;
    uint32_t maxWaves;
    ProposalWeights pw;
};
hipError_t cellsKLaunch(const CellCtx& c, const uint32_t* knn, uint32_t K, const uint8_t* cls, uint8_t* labels,
                        uint32_t mode, uint32_t fallbackK, hipStream_t st);
// d16 (may be null) receives each site's 16th-neighbour centre distance.
hipError_t featuresLaunch(const CellCtx& c, const uint32_t* knn, uint32_t K, float* feat, hipStream_t st,
                          float* d16 = nullptr);
// cls[i] = given[indices[i]]: caller-given classes (input order) in sorted order.
hipError_t gatherClassesLaunch(const uint32_t* indices, const uint8_t* given, uint32_t n, uint8_t* cls,
                               hipStream_t st);
// radii[i] = given[indices[i]]: caller-given search radii (input order) in sorted order.
hipError_t gatherRadiiLaunch(const uint32_t* indices, const float* given, uint32_t n, float* radii, hipStream_t st);
// Classes of n sites; classes asking for more than maxK neighbours are
// replaced by the last class (not certified within maxK).  logits (may be
// null): kProposalOutputs per site.  radii (may be null, with d16): the
// radius head's search radius d16 x 2^out per site.
hipError_t mlpInferLaunch(const float* feat, uint32_t n, const ProposalWeights& w, uint8_t* cls, float* logits,
                          float cost, uint32_t fallbackK, uint32_t maxK, hipStream_t st, const float* d16 = nullptr,
                          float* radii = nullptr);
hipError_t packWeightsLaunch(const float* params, const ProposalWeights& w, hipStream_t st);
// Waves (and partial gradients) of a batch of `rows`.
uint32_t trainWaves(uint32_t rows);
// Partial gradients of the mean loss over rows order[0..rows) (scaled by
// invBatch) with the current forward and backward weights.
hipError_t trainGradLaunch(const TrainBuffers& t, const uint32_t* order, uint32_t rows, float invBatch,
                           hipStream_t st);
hipError_t mlpTrainStep(const TrainBuffers& t, const uint32_t* order, uint32_t rows, float lr, uint32_t step,
                        hipStream_t st);

// Topology (topology.hip)
hipError_t topologyCountLaunch(const CellCtx& c, uint32_t* counts, hipStream_t st);
hipError_t topologyWriteLaunch(const CellCtx& c, const uint64_t* offsets, const uint32_t* indices, uint32_t* bis,
                               uint32_t* tri, uint32_t* quad, hipStream_t st);
hipError_t cellRecordsLaunch(const CellCtx& c, const uint32_t* indices, const uint64_t* vOff, const uint64_t* eOff,
                             uint4* vOut, uint4* eOut, hipStream_t st);

hipError_t checkEmptyLaunch(const CellCtx& c, const uint32_t* indices, uint32_t* rank, const D4* xyzw,
                            const uint4* ids, uint64_t count, double tol, unsigned long long* violations,
                            hipStream_t st);
// rank[indices[j]] = j: original site id -> sorted position.
hipError_t rankLaunch(const uint32_t* indices, uint32_t n, uint32_t* rank, hipStream_t st);
// Exclusive scan of `blocks` block sums in place; *total = their sum.
hipError_t scanBlockSumsLaunch(uint32_t* blockSums, uint32_t blocks, uint32_t* total, hipStream_t st);

// Cavity detection (cavity.hip, app/cavity/utils.cu): hit ratio of each
// vertex over `samples` rays (directions dirs), then the vertices whose
// ratio and radius pass the thresholds, compacted in order.
hipError_t hitRatiosLaunch(const CellCtx& c, const uint32_t* indices, uint32_t* rank, const D4* xyzw,
                           const uint4* ids, uint64_t count, const F3* dirs, uint32_t samples, float* ratios,
                           hipStream_t st);
hipError_t cavityCountLaunch(const D4* xyzw, const float* ratios, uint64_t count, float hitThreshold,
                             float radiusThreshold, uint32_t* blockSums, uint32_t* total, hipStream_t st);
hipError_t cavityWriteLaunch(const D4* xyzw, const float* ratios, uint64_t count, float hitThreshold,
                             float radiusThreshold, const uint32_t* blockOffsets, float4* out, uint64_t capacity,
                             hipStream_t st);

// Interval validator (app/precision/precision.cpp:14-163): counts2[0] +=
// points some other site is certainly closer to, counts2[1] += the others
// some site may be closer to.  arity 4: vertices, 3: closed-edge points.
hipError_t precisionLaunch(const CellCtx& c, const uint32_t* indices, uint32_t* rank, const D4* xyzw,
                           const uint4* ids, uint64_t count, uint32_t arity, unsigned long long* counts2,
                           hipStream_t st);

}  // namespace apo_hip
