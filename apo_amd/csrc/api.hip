// api.hip -- C ABI (include/apo_c.h) over the gfx950 kernels.
//
// Orchestration follows AlgorithmGPU<K>::build / toHost
// (apo/include/apo/gpu/algorithm.inl:1410-1646) with every step on one
// context-owned stream: no host round trip between the LBVH, the kNN pass,
// the cell kernel, the validation rounds (task lists are compacted on the
// device) and the vertex output.  The only synchronisation is at the end of
// a build, to read the vertex count.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <cmath>
#include <mutex>
#include <random>
#include <vector>

#include "../../include/apo_c.h"
#include "apo_device.h"
#include "apo_internal.h"

using namespace apo_hip;

namespace {

// kNN passes after the first whose kNN launch is timed (the rest count as cell time)
constexpr uint32_t kTimedPasses = 32;

// What a captured build graph depends on besides buffer addresses (a change
// of any buffer drops the graph).
struct GraphKey {
    const void* sites = nullptr;
    uint32_t n = 0, molecules = 0;
    bool learned = false, given = false, ranged = false;
    uint64_t outCap = 0;
    apo_config cfg{};
    bool operator==(const GraphKey& o) const {
        return sites == o.sites && n == o.n && molecules == o.molecules && learned == o.learned && given == o.given &&
               ranged == o.ranged &&
               outCap == o.outCap && std::memcmp(&cfg, &o.cfg, sizeof(cfg)) == 0;
    }
};

}  // namespace

// ---- working memory of apo_build_device_static: one array per device,
// held until the process ends.  Host code never uses its name as an
// address: the device address comes from hipGetSymbolAddress, once per
// device, at apo_create.
static __device__ __align__(256) unsigned char scratch[APO_STATIC_SCRATCH_BYTES];

namespace {

constexpr uint32_t kStaticMaxSites = APO_STATIC_MAX_SITES;
// vertex / edge validation waves of a static build (each uses its own
// device region of the array)
constexpr uint32_t kStaticWaveRegions = 1024;
constexpr int kStaticMaxDevices = 64;
std::mutex gScratchLock;
void* gScratchAddr[kStaticMaxDevices] = {};

// Device address of this device's copy of `scratch`, with the device current.
hipError_t staticScratchAddress(int device, void** out) {
    if (device < 0 || device >= kStaticMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gScratchLock);
    if (!gScratchAddr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        gScratchAddr[device] = p;
    }
    *out = gScratchAddr[device];
    return hipSuccess;
}

}  // namespace

struct apo_ctx {
    apo_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    int cus = 0;

    uint32_t cap = 0;      // site capacity of the buffers below
    uint32_t capK = 0;
    uint32_t n = 0;        // sites of the last build
    uint32_t molecules = 1;  // molecules of the last build
    bool built = false;

    // batches: site offsets of the molecules and the vertex offsets of each
    // molecule's output (device), molecule capacity of these and of `scene`
    uint32_t capM = 0;
    uint32_t* offsets = nullptr;
    uint64_t* moleculeVertices = nullptr;
    // blocks per CU of the persistent kernels at full occupancy
    uint32_t occCells = 1, occCellsPair = 1, occVertex = 1, occEdge = 1, occBisector = 1, occBisectorPlan = 1,
             occBisectorLanes = 1;
    // the captured build (config.graph) and what it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    GraphKey graphKey{};
    std::vector<uint32_t> roundTasks;  // task count of each validation round of the last build

    // input staging
    F4* sitesIn = nullptr;
    // LBVH
    Scene* scene = nullptr;
    double* boxPartials = nullptr;
    uint32_t* boxTicket = nullptr;
    uint32_t* sortScratch = nullptr;
    uint32_t* range = nullptr;
    uint32_t* startLists = nullptr;
    uint64_t* codesA = nullptr;
    uint64_t* codesB = nullptr;
    uint32_t* idxA = nullptr;
    uint32_t* idxB = nullptr;
    F4* sorted = nullptr;
    Node* nodes = nullptr;
    int32_t* parent = nullptr;
    uint32_t* refitCounter = nullptr;
    float* rootBox = nullptr;
    // cells
    uint8_t* status = nullptr;
    float* knownRadii = nullptr;
    uint32_t* flags = nullptr;
    uint4* vtx = nullptr;
    uint32_t* vnb = nullptr;
    uint4* edg = nullptr;
    uint32_t* enb = nullptr;
    uint32_t* knn = nullptr;
    uint32_t* counters = nullptr;   // [0] cell queue, [1..15] validation
    uint32_t* spill = nullptr;      // per-wave regions of the validation waves (deep stack slots, covers)
    uint32_t spillBlocks = 0;
    uint32_t* err = nullptr;        // device guard bits
    uint32_t* tasks = nullptr;
    // the bisector stage split by search (BisectorSplit); null in static builds
    uint32_t* bisSearch = nullptr;
    uint32_t* bisNb = nullptr;
    uint32_t* bisCount = nullptr;
    uint8_t* bisMark = nullptr;
    // output
    uint32_t* cellCount = nullptr;
    uint32_t* blockSums = nullptr;
    uint32_t* total = nullptr;
    D4* outXyzw = nullptr;
    uint4* outIds = nullptr;
    uint64_t outCap = 0;
    uint64_t vertexCount = 0;

    unsigned long long* bytes = nullptr;  // per-kernel algorithmic byte counters
    // learned neighbour-count proposal
    bool proposalOn = false;
    ProposalWeights pw{};       // inference layout (BF16 weights)
    float* params = nullptr;    // FP32 parameters
    float* feat = nullptr;      // sites x kFeatures
    uint8_t* cls = nullptr;     // predicted class per sorted site
    uint8_t* labels = nullptr;  // label pass output per sorted site
    uint32_t capP = 0;          // site capacity of feat / cls / labels
    uint32_t* want = nullptr;     // completion mode: smallest wanted site per cell (kNone: none)
    uint32_t capW = 0;
    uint4* stuck = nullptr;       // completion mode: a round's stuck records (kStuckCap)
    float* givenRadii = nullptr;  // search radii given by the caller (input order): range-query cell pass
    uint32_t givenRN = 0;
    uint32_t givenRCap = 0;
    float* radiiSorted = nullptr;  // the same in sorted order (or the radius head's), capacity capP
    float* d16 = nullptr;          // 16th-neighbour centre distance per sorted site, capacity capP
    uint8_t* givenCls = nullptr;  // classes given by the caller (input order), used instead of the MLP
    uint32_t givenN = 0;          // their site count (0: none given)
    uint32_t givenCap = 0;
    // cavity detection: hit ratios of the current build's vertices
    float* cavRatios = nullptr;
    uint64_t cavCap = 0;
    bool cavValid = false;
    void* staticScratch = nullptr;  // this device's copy of the static scratch array (apo_build_device_static)
    hipEvent_t ev[10] = {};
    hipEvent_t passEv[kTimedPasses][2] = {};  // around the kNN launch of passes 1..kTimedPasses
    ValidationTiming vt{};
    apo_stats stats{};
};

namespace {

apo_status fail(apo_ctx* c, apo_status s, const std::string& msg) {
    if (c) c->error = msg;
    return s;
}

apo_status hipFail(apo_ctx* c, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(c, e == hipErrorOutOfMemory ? APO_ERR_OUT_OF_MEMORY : APO_ERR_HIP, m);
}

#define APO_HIP(call)                                         \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return hipFail(ctx, e_, #call); \
    } while (0)

template <typename T>
hipError_t alloc(T*& p, size_t count) {
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        if (e != hipSuccess) return e;
    }
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
}

// Scoped device buffer of the introspection calls (freed on every return path).
template <typename T>
struct DeviceTemp {
    T* p = nullptr;
    hipError_t alloc(size_t count) { return hipMalloc(reinterpret_cast<void**>(&p), (count ? count : 1) * sizeof(T)); }
    ~DeviceTemp() {
        if (p) (void)hipFree(p);
    }
};

void dropGraph(apo_ctx* ctx) {
    if (ctx->graphExec) (void)hipGraphExecDestroy(ctx->graphExec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graphExec = nullptr;
    ctx->graph = nullptr;
}

apo_status ensureMolecules(apo_ctx* ctx, uint32_t molecules) {
    if (molecules <= ctx->capM) return APO_OK;
    dropGraph(ctx);
    APO_HIP(alloc(ctx->scene, molecules));
    APO_HIP(alloc(ctx->offsets, size_t(molecules) + 1));
    APO_HIP(alloc(ctx->moleculeVertices, size_t(molecules) + 1));
    ctx->capM = molecules;
    return APO_OK;
}

apo_status ensureCapacity(apo_ctx* ctx, uint32_t n, uint32_t K) {
    if (n <= ctx->cap && K <= ctx->capK) return APO_OK;
    dropGraph(ctx);
    const uint32_t cap = n;
    const size_t nodes = 2 * size_t(cap) - 1;
    APO_HIP(alloc(ctx->bytes, size_t(kByteFamilies) * kByteSlots));
    APO_HIP(alloc(ctx->sitesIn, cap));
    APO_HIP(alloc(ctx->boxPartials, size_t(kBoxBlocks) * 8));
    APO_HIP(alloc(ctx->boxTicket, 1));
    // the ticket is self-resetting; clear it once, ordered on the build stream
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, ctx->stream, ctx->boxTicket, 1u, 0u);
    APO_HIP(hipGetLastError());
    APO_HIP(alloc(ctx->sortScratch, sortScratchWords(cap)));
    APO_HIP(alloc(ctx->range, 2 * size_t(cap)));
    APO_HIP(alloc(ctx->startLists, refitStartWords(cap)));
    APO_HIP(alloc(ctx->codesA, cap));
    APO_HIP(alloc(ctx->codesB, cap));
    APO_HIP(alloc(ctx->idxA, cap));
    APO_HIP(alloc(ctx->idxB, cap));
    APO_HIP(alloc(ctx->sorted, cap));
    APO_HIP(alloc(ctx->nodes, cap));
    APO_HIP(alloc(ctx->parent, nodes));
    APO_HIP(alloc(ctx->refitCounter, cap));
    APO_HIP(alloc(ctx->rootBox, 8));
    APO_HIP(alloc(ctx->status, cap));
    APO_HIP(alloc(ctx->knownRadii, cap));
    APO_HIP(alloc(ctx->flags, cap));
    APO_HIP(alloc(ctx->vtx, size_t(cap) * kVertexCap));
    APO_HIP(alloc(ctx->vnb, cap));
    APO_HIP(alloc(ctx->edg, size_t(cap) * kEdgeCap));
    APO_HIP(alloc(ctx->enb, cap));
    APO_HIP(alloc(ctx->knn, size_t(cap) * K));
    APO_HIP(alloc(ctx->counters, kCounterWords));
    APO_HIP(alloc(ctx->err, 1));
    APO_HIP(alloc(ctx->tasks, cap));
    APO_HIP(alloc(ctx->bisSearch, size_t(cap) * kBisSearchSlots));
    APO_HIP(alloc(ctx->bisNb, size_t(cap) * kBisNeighbourSlots));
    APO_HIP(alloc(ctx->bisCount, cap));
    APO_HIP(alloc(ctx->bisMark, cap));
    APO_HIP(alloc(ctx->cellCount, cap));
    APO_HIP(alloc(ctx->blockSums, (cap + 255) / 256 + 2));
    APO_HIP(alloc(ctx->total, 1));
    ctx->outCap = uint64_t(cap) * 8 + 1024;
    APO_HIP(alloc(ctx->outXyzw, ctx->outCap));
    APO_HIP(alloc(ctx->outIds, ctx->outCap));
    ctx->cap = cap;
    ctx->capK = K;
    return APO_OK;
}

// The learned path runs with MLP weights or with classes given by the caller.
bool learnedPath(const apo_ctx* ctx) { return ctx->proposalOn || ctx->givenN != 0u || ctx->givenRN != 0u; }

apo_status ensureProposalBuffers(apo_ctx* ctx, uint32_t n) {
    if (n <= ctx->capP) return APO_OK;
    APO_HIP(alloc(ctx->feat, size_t(n) * kFeatures));
    APO_HIP(alloc(ctx->cls, n));
    APO_HIP(alloc(ctx->labels, n));
    APO_HIP(alloc(ctx->radiiSorted, n));
    APO_HIP(alloc(ctx->d16, n));
    ctx->capP = n;
    return APO_OK;
}


CellCtx cellCtx(apo_ctx* ctx) {
    CellCtx c;
    c.n = ctx->n;
    c.internal = ctx->n - ctx->molecules;
    c.molecules = ctx->molecules;
    c.offsets = ctx->molecules > 1u ? ctx->offsets : nullptr;
    c.sorted = ctx->sorted;
    c.scene = ctx->scene;
    c.nodes = ctx->nodes;
    c.parent = ctx->parent;
    c.treeHeight = ctx->counters + kTreeHeight;
    c.status = ctx->status;
    c.knownRadii = ctx->knownRadii;
    c.flags = ctx->flags;
    c.vtx = ctx->vtx;
    c.vnb = ctx->vnb;
    c.edg = ctx->edg;
    c.enb = ctx->enb;
    c.err = ctx->err;
    c.bytes = ctx->bytes;
    c.coverRatio = ctx->cfg.cover_ratio < 0.f ? __builtin_inff() : ctx->cfg.cover_ratio;
    c.complete = ctx->cfg.complete ? 1u : 0u;
    return c;
}

// Blocks of a persistent launch: the CUs times the kernel's own waves per
// CU setting, else config.waves_per_cu, else its full occupancy.
uint32_t persistentBlocks(apo_ctx* ctx, uint32_t own, uint32_t occupancy) {
    const uint32_t perCu = own ? own : ctx->cfg.waves_per_cu ? ctx->cfg.waves_per_cu : occupancy;
    return uint32_t(ctx->cus) * perCu;
}

// The cell pass runs two cells per wave when the config asks for it.
bool cellPairs(const apo_ctx* ctx) { return ctx->cfg.cells_per_wave == 2u; }

// Grid of the cell pass: the configured waves per CU, at most what fits a CU
// when a wave holds two cells (their records take twice the LDS).
uint32_t cellPassBlocks(apo_ctx* ctx) {
    if (!cellPairs(ctx)) return persistentBlocks(ctx, ctx->cfg.waves_cells, ctx->occCells);
    const uint32_t own = ctx->cfg.waves_cells ? ctx->cfg.waves_cells : ctx->cfg.waves_per_cu;
    const uint32_t perCu = own && own < ctx->occCellsPair ? own : ctx->occCellsPair;
    return uint32_t(ctx->cus) * perCu;
}

std::string guardText(uint32_t bits) {
    static const char* names[] = {"sort tile", "sort scatter", "sort look-back spin", "gather index",
                                  "refit parent", "refit child", "knn child", "cell site id",
                                  "cell neighbour id", "validation"};
    std::string out;
    for (int b = 0; b < 10; b++)
        if (bits & (1u << b)) out += std::string(out.empty() ? "" : ", ") + names[b];
    return out;
}

// With APO_SYNC_STAGES=1 every stage is followed by a synchronisation and a
// check of the runtime error state and of the device guard word, so a failure
// is attributed to the stage that produced it.
apo_status stageCheck(apo_ctx* ctx, const char* stage) {
    static const bool sync = [] {
        const char* e = std::getenv("APO_SYNC_STAGES");
        return e && e[0] == '1';
    }();
    if (!sync) return APO_OK;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hipFail(ctx, e, stage);
    uint32_t bits = 0;
    e = hipMemcpy(&bits, ctx->err, sizeof(bits), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipFail(ctx, e, stage);
    std::fprintf(stderr, "[apo] stage %s ok (guards 0x%x)\n", stage, bits);
    if (bits) return fail(ctx, APO_ERR_HIP, std::string("device guard tripped after ") + stage + ": " + guardText(bits));
    return APO_OK;
}

LbvhBuffers lbvhBuffers(apo_ctx* ctx) {
    return LbvhBuffers{ctx->scene,  ctx->boxPartials, ctx->boxTicket,    ctx->sortScratch, ctx->codesA,
                       ctx->codesB, ctx->idxA,        ctx->idxB,         ctx->sorted,      ctx->nodes,
                       ctx->parent, ctx->refitCounter, ctx->rootBox,     ctx->err,
                       ctx->range,  ctx->startLists};
}

Segments segments(apo_ctx* ctx) { return Segments{ctx->molecules, ctx->molecules > 1u ? ctx->offsets : nullptr}; }

// Clears the guard word, the byte counters and the build counters.
void resetCounters(apo_ctx* ctx, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, ctx->err, 1u, 0u);
    constexpr uint32_t byteWords = 2u * kByteFamilies * kByteSlots;
    hipLaunchKernelGGL(k_fill_words, dim3((byteWords + 63u) / 64u), dim3(64), 0, st,
                       reinterpret_cast<uint32_t*>(ctx->bytes), byteWords, 0u);
    hipLaunchKernelGGL(k_fill_words, dim3((kCounterWords + 63u) / 64u), dim3(64), 0, st, ctx->counters,
                       kCounterWords, 0u);
}

// Every device step of a build in stream order, with no host round trip:
// what config.graph captures.  With `timed`, events mark the stages and the
// validation rounds, and APO_SYNC_STAGES checks each stage.
apo_status enqueueBuild(apo_ctx* ctx, const F4* sitesDev, uint32_t n, bool learned, bool timed) {
    hipStream_t st = ctx->stream;
    const CellCtx cc = cellCtx(ctx);
    const uint32_t K = ctx->cfg.k;
    auto mark = [&](int e) { return timed ? hipEventRecord(ctx->ev[e], st) : hipSuccess; };
#define APO_TIMED_STAGE(name)                            \
    do {                                                 \
        if (timed) {                                     \
            apo_status s_ = stageCheck(ctx, name);       \
            if (s_ != APO_OK) return s_;                 \
        }                                                \
    } while (0)
    resetCounters(ctx, st);
    APO_HIP(mark(0));
    APO_HIP(lbvhBuild(lbvhBuffers(ctx), sitesDev, n, segments(ctx), st));
    APO_TIMED_STAGE("lbvh");
    APO_HIP(mark(1));
    // the learned path's lists hold the most neighbours any class may ask for
    const uint32_t Kk = learned ? proposalListK(ctx->cfg.proposal_max_k, ctx->cfg.proposal_fallback_k) : K;
    const uint32_t passes = learned ? 1u : ctx->cfg.knn_passes;
    // search radii given by the caller: every list is a range query below its radius
    const bool ranged = learned && ctx->givenRN != 0u;
    if (ranged) APO_HIP(gatherRadiiLaunch(ctx->idxA, ctx->givenRadii, n, ctx->radiiSorted, st));
    if (passes == 0u) {
        // no kNN or cell pass (knnValidationStep = 0): empty lists, every cell
        // in validation; the cell launch below writes the octahedra
        APO_HIP(cellsResetLaunch(cc, ctx->knn, Kk, st));
    } else {
        APO_HIP(knnLaunch(cc, Kk, ctx->knn, st, 0u, ranged ? ctx->radiiSorted : nullptr));
    }
    APO_TIMED_STAGE("knn");
    APO_HIP(mark(2));
    if (ctx->cfg.stage_mask & 8u) return APO_OK;  // stop after the kNN pass
    if (ranged && (ctx->cfg.stage_mask & 1u)) {
        // each cell updated with its whole range list, the known radius its
        // range when the list holds fewer sites than its capacity
        APO_HIP(mark(7));
        APO_HIP(cellsLaunch(cc, ctx->knn, Kk, ctx->counters + kCellQueue,
                            cellPassBlocks(ctx), st, nullptr,
                            ctx->cfg.proposal_fallback_k, false, ctx->radiiSorted, cellPairs(ctx)));
    } else if (learned && ctx->cfg.proposal_radius && !ctx->givenN && (ctx->cfg.stage_mask & 1u)) {
        // the radius head: features of the 16 nearest, each site's search
        // radius, then the range-query lists below it and the ranged cell pass
        APO_HIP(featuresLaunch(cc, ctx->knn, Kk, ctx->feat, st, ctx->d16));
        APO_HIP(mlpInferLaunch(ctx->feat, n, ctx->pw, ctx->cls, nullptr, ctx->cfg.proposal_cost,
                               ctx->cfg.proposal_fallback_k, ctx->cfg.proposal_max_k, st, ctx->d16,
                               ctx->radiiSorted));
        APO_HIP(knnLaunch(cc, Kk, ctx->knn, st, 0u, ctx->radiiSorted));
        APO_HIP(mark(7));
        APO_HIP(cellsLaunch(cc, ctx->knn, Kk, ctx->counters + kCellQueue,
                            cellPassBlocks(ctx), st, nullptr,
                            ctx->cfg.proposal_fallback_k, false, ctx->radiiSorted, cellPairs(ctx)));
    } else if (learned && (ctx->cfg.stage_mask & 1u)) {
        // features of the 16 nearest, proposal classes, then each cell with its
        // proposed neighbour count and the plain certification test
        if (ctx->givenN) {
            APO_HIP(gatherClassesLaunch(ctx->idxA, ctx->givenCls, n, ctx->cls, st));
        } else {
            APO_HIP(featuresLaunch(cc, ctx->knn, Kk, ctx->feat, st));
            APO_HIP(mlpInferLaunch(ctx->feat, n, ctx->pw, ctx->cls, nullptr, ctx->cfg.proposal_cost,
                                   ctx->cfg.proposal_fallback_k, ctx->cfg.proposal_max_k, st));
        }
        APO_HIP(mark(7));
        APO_HIP(cellsLaunch(cc, ctx->knn, Kk, ctx->counters + kCellQueue,
                            cellPassBlocks(ctx), st, ctx->cls,
                            ctx->cfg.proposal_fallback_k, false, nullptr, cellPairs(ctx)));
    } else {
        APO_HIP(mark(7));
        const uint32_t cellBlocks = cellPassBlocks(ctx);
        APO_HIP(cellsLaunch(cc, ctx->knn, (ctx->cfg.stage_mask & 1u) && passes > 0u ? K : 0u,
                            ctx->counters + kCellQueue, cellBlocks, st, nullptr, 16u, false, nullptr,
                            cellPairs(ctx)));
        // knnValidationStep > 1 (algorithm.inl:1535-1549): the next K neighbours
        // of every site, then the cells still in validation continue with them
        for (uint32_t s = 1; s < passes; s++) {
            const bool tp = timed && s <= kTimedPasses;
            if (tp) APO_HIP(hipEventRecord(ctx->passEv[s - 1][0], st));
            APO_HIP(knnLaunch(cc, K, ctx->knn, st, s));
            if (tp) APO_HIP(hipEventRecord(ctx->passEv[s - 1][1], st));
            if (ctx->cfg.stage_mask & 1u)
                APO_HIP(cellsLaunch(cc, ctx->knn, K, ctx->counters + kCellQueue, cellBlocks, st, nullptr, 16u, true,
                                    nullptr, cellPairs(ctx)));
        }
    }
    APO_TIMED_STAGE("cells");
    APO_HIP(mark(3));
    // tree height: lets vertex validation skip a search that cannot overflow
    // (the task list, cell counts and block sums are free until the validation
    // rounds and the output)
    APO_HIP(treeHeightLaunch(ctx->range, n, n - ctx->molecules, reinterpret_cast<int32_t*>(ctx->tasks),
                             reinterpret_cast<int32_t*>(ctx->cellCount), reinterpret_cast<int2*>(ctx->blockSums),
                             ctx->counters + kTreeHeight, st));
    const ValidationBuffers vb{ctx->tasks, ctx->counters, ctx->spill, ctx->spillBlocks};
    // validation waves each own a device region: at most spillBlocks of them
    const ValidationGrid vg{std::min(persistentBlocks(ctx, ctx->cfg.waves_vertex, ctx->occVertex), ctx->spillBlocks),
                            std::min(persistentBlocks(ctx, ctx->cfg.waves_edge, ctx->occEdge), ctx->spillBlocks),
                            std::min(persistentBlocks(ctx, ctx->cfg.waves_bisector, ctx->occBisector),
                                     ctx->spillBlocks),
                            persistentBlocks(ctx, 0u, ctx->occBisectorPlan),
                            std::min(persistentBlocks(ctx, 0u, ctx->occBisectorLanes), ctx->spillBlocks)};
    const BisectorSplit bsplit{ctx->bisSearch, ctx->bisNb, ctx->bisCount, ctx->bisMark};
    if (ctx->cfg.stage_mask & 2u)
        APO_HIP(validationLaunch(cc, vb, vg, ctx->cfg.max_iterations, ctx->cfg.max_restarts, false, st,
                                 timed ? &ctx->vt : nullptr));
    APO_TIMED_STAGE("vertex/edge validation");
    APO_HIP(mark(4));
    if (ctx->cfg.stage_mask & 4u)
        APO_HIP(validationLaunch(cc, vb, vg, 1, ctx->cfg.max_restarts, true, st, nullptr, nullptr,
                                 ctx->cfg.bisector_per_cell ? nullptr : &bsplit));
    APO_TIMED_STAGE("bisector validation");
    APO_HIP(mark(5));
    if (ctx->cfg.complete && (ctx->cfg.stage_mask & 2u)) {
        // completion mode: the cells the bisector stage changed (status 0) are
        // validated again, then the consistency rounds, each followed by
        // validation rounds over the cells it updated
        hipLaunchKernelGGL(k_fill_words, dim3((n + 255u) / 256u), dim3(256), 0, st, ctx->want, n, kNone);
        APO_HIP(validationLaunch(cc, vb, vg, ctx->cfg.max_iterations, ctx->cfg.max_restarts, false, st, nullptr,
                                 ctx->counters + kCompletionTasks));
        const uint32_t blocks = cellPassBlocks(ctx);
        for (uint32_t r = 0; r < ctx->cfg.complete; r++) {
            APO_HIP(consistencyLaunch(cc, ctx->want, ctx->stuck, ctx->counters, r, blocks, st));
            APO_HIP(validationLaunch(cc, vb, vg, ctx->cfg.max_iterations, ctx->cfg.max_restarts, false, st, nullptr,
                                     ctx->counters + kCompletionTasks + (r + 1u) * kMaxRounds));
        }
    }
    APO_TIMED_STAGE("completion");
    APO_HIP(mark(8));
    const OutputBuffers ob{ctx->cellCount, ctx->blockSums, ctx->total, ctx->outXyzw, ctx->outIds, ctx->outCap};
    APO_HIP(outputCountLaunch(cc, ob, st));
    APO_HIP(outputWriteLaunch(cc, ob, ctx->idxA, st));
    if (ctx->molecules > 1u) APO_HIP(moleculeVerticesLaunch(cc, ob, ctx->moleculeVertices, st));
    APO_TIMED_STAGE("output");
    APO_HIP(mark(6));
#undef APO_TIMED_STAGE
    return APO_OK;
}

// After the device work: vertex count, guard word, output growth and the
// statistics.  `timed` builds carry per-stage times; a graph replay only
// its total (events 0 and 6 around the launch).
apo_status finishBuild(apo_ctx* ctx, uint32_t n, bool learned, bool timed) {
    hipStream_t st = ctx->stream;
    uint32_t total = 0, guard = 0;
    std::vector<uint32_t> counters(kCounterWords);
    APO_HIP(hipMemcpyAsync(&total, ctx->total, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(&guard, ctx->err, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(counters.data(), ctx->counters, kCounterWords * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    if (guard) return fail(ctx, APO_ERR_HIP, "device guard tripped: " + guardText(guard));
    if (ctx->cfg.stage_mask & 8u) {
        // LBVH and kNN only: their stage times, no cells
        ctx->vertexCount = 0;
        apo_stats& S = ctx->stats;
        std::memset(&S, 0, sizeof(S));
        if (timed) {
            float a = 0.f, b = 0.f;
            APO_HIP(hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
            APO_HIP(hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
            S.ms_lbvh = a;
            S.ms_knn = b;
            S.ms_total = a + b;
        }
        S.sites = n;
        ctx->built = true;
        return APO_OK;
    }
    if (total > ctx->outCap) {
        // more vertices than the estimate: grow the output and write again
        dropGraph(ctx);
        ctx->outCap = uint64_t(total) + (uint64_t(total) >> 3) + 1024;
        APO_HIP(alloc(ctx->outXyzw, ctx->outCap));
        APO_HIP(alloc(ctx->outIds, ctx->outCap));
        const OutputBuffers ob2{ctx->cellCount, ctx->blockSums, ctx->total, ctx->outXyzw, ctx->outIds, ctx->outCap};
        APO_HIP(outputWriteLaunch(cellCtx(ctx), ob2, ctx->idxA, st));
        APO_HIP(hipStreamSynchronize(st));
    }
    ctx->vertexCount = total;
    apo_stats& S = ctx->stats;
    std::memset(&S, 0, sizeof(S));
    float all = 0;
    APO_HIP(hipEventElapsedTime(&all, ctx->ev[0], ctx->ev[6]));
    S.ms_total = all;
    if (timed) {
        float ms[6] = {};
        for (int q = 0; q < 6; q++) APO_HIP(hipEventElapsedTime(&ms[q], ctx->ev[q], ctx->ev[q + 1]));
        float msProposal = 0.f, msCellsOnly = 0.f;
        APO_HIP(hipEventElapsedTime(&msProposal, ctx->ev[2], ctx->ev[7]));
        APO_HIP(hipEventElapsedTime(&msCellsOnly, ctx->ev[7], ctx->ev[3]));
        // kNN launches of the later passes sit inside the cell interval
        float msLaterKnn = 0.f;
        const uint32_t passes = learned ? 1u : ctx->cfg.knn_passes;
        for (uint32_t s = 1; s < passes && s <= kTimedPasses; s++) {
            float a = 0.f;
            APO_HIP(hipEventElapsedTime(&a, ctx->passEv[s - 1][0], ctx->passEv[s - 1][1]));
            msLaterKnn += a;
        }
        S.ms_lbvh = ms[0];
        S.ms_knn = ms[1] + msLaterKnn;
        S.ms_cells = msCellsOnly - msLaterKnn;
        S.ms_proposal = msProposal;
        S.ms_validation = ms[3];
        float msCompletion = 0.f, msOutput = 0.f;
        APO_HIP(hipEventElapsedTime(&msCompletion, ctx->ev[5], ctx->ev[8]));
        APO_HIP(hipEventElapsedTime(&msOutput, ctx->ev[8], ctx->ev[6]));
        S.ms_bisector = ms[4];
        S.ms_completion = msCompletion;
        S.ms_output = msOutput;
        for (uint32_t r = 0; r < ctx->vt.rounds; r++) {
            float a = 0.f, b = 0.f;
            APO_HIP(hipEventElapsedTime(&a, ctx->vt.ev[r][0], ctx->vt.ev[r][1]));
            APO_HIP(hipEventElapsedTime(&b, ctx->vt.ev[r][1], ctx->vt.ev[r][2]));
            S.ms_vertex_validation_kernel += a;
            S.ms_edge_validation_kernel += b;
        }
    }
    S.proposal = learned ? 1u : 0u;
    S.sites = n;
    S.vertices = total;
    // rounds that had task cells
    ctx->roundTasks.clear();
    const uint32_t rounds = (ctx->cfg.stage_mask & 2u) ? ctx->cfg.max_iterations : 0u;
    for (uint32_t r = 0; r < rounds; r++) {
        const uint32_t t = counters[kRoundTasks + r];
        if (t == 0) break;
        ctx->roundTasks.push_back(t);
        S.validation_tasks += t;
    }
    S.validation_launches = uint32_t(ctx->roundTasks.size());
    if (ctx->cfg.complete && (ctx->cfg.stage_mask & 2u)) {
        // round r ran when r == 0 or round r - 1 updated a cell
        for (uint32_t r = 0; r < ctx->cfg.complete; r++) {
            if (r > 0 && counters[kCompletionUpdated + r - 1] == 0) break;
            S.completion_rounds++;
            S.completion_updated_cells += counters[kCompletionUpdated + r];
        }
        // records still missing when the rounds end: the last round's count
        if (S.completion_rounds) S.completion_unrepairable = counters[kCompletionStuck + S.completion_rounds - 1];
        S.completion_rebuilt_cells = counters[kCompletionRebuilt];
        for (uint32_t q = 0; q <= ctx->cfg.complete; q++)
            for (uint32_t r = 0; r < ctx->cfg.max_iterations; r++)
                S.completion_validation_tasks += counters[kCompletionTasks + q * kMaxRounds + r];
    }
    unsigned long long counts[kByteFamilies * kByteSlots] = {};
    APO_HIP(hipMemcpy(counts, ctx->bytes, sizeof(counts), hipMemcpyDeviceToHost));
    uint64_t fam[kByteFamilies] = {};
    for (uint32_t f = 0; f < kByteFamilies; f++)
        for (uint32_t q = 0; q < kByteSlots; q++) fam[f] += counts[f * kByteSlots + q];
    S.bytes_cells = fam[kBytesCells];
    S.bytes_vertex_validation = fam[kBytesVertexVal];
    S.bytes_edge_validation = fam[kBytesEdgeVal];
    S.bytes_bisector_validation = fam[kBytesBisectorVal];
    ctx->built = true;
    return APO_OK;
}

// One build of the sites at sitesDev (molecule offsets already on the
// device): eager with stage events, or as a replay of the captured graph.
apo_status runPipeline(apo_ctx* ctx, const F4* sitesDev, uint32_t n) {
    hipStream_t st = ctx->stream;
    ctx->n = n;
    const bool learned = learnedPath(ctx);
    const bool graph = ctx->cfg.graph != 0u && !(ctx->cfg.stage_mask & 8u);
    if (!graph) {
        apo_status s = enqueueBuild(ctx, sitesDev, n, learned, true);
        if (s != APO_OK) return s;
        return finishBuild(ctx, n, learned, true);
    }
    GraphKey key;
    key.sites = sitesDev;
    key.n = n;
    key.molecules = ctx->molecules;
    key.cfg = ctx->cfg;
    key.learned = learned;
    key.given = ctx->givenN != 0u;
    key.ranged = ctx->givenRN != 0u;
    key.outCap = ctx->outCap;
    if (!ctx->graphExec || !(key == ctx->graphKey)) {
        dropGraph(ctx);
        APO_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const apo_status s = enqueueBuild(ctx, sitesDev, n, learned, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (s != APO_OK) {
            if (g) (void)hipGraphDestroy(g);
            return s;
        }
        APO_HIP(e);
        ctx->graph = g;
        APO_HIP(hipGraphInstantiate(&ctx->graphExec, g, nullptr, nullptr, 0));
        ctx->graphKey = key;
    }
    APO_HIP(hipEventRecord(ctx->ev[0], st));
    APO_HIP(hipGraphLaunch(ctx->graphExec, st));
    APO_HIP(hipEventRecord(ctx->ev[6], st));
    return finishBuild(ctx, n, learned, false);
}

// Label pass of the learned proposal: LBVH, kNN with 64 neighbours,
// features and the cell pass that tests after every update, then stop.
apo_status runLabelPass(apo_ctx* ctx, const F4* sitesDev, uint32_t n) {
    hipStream_t st = ctx->stream;
    ctx->n = n;
    const CellCtx cc = cellCtx(ctx);
    resetCounters(ctx, st);
    APO_HIP(lbvhBuild(lbvhBuffers(ctx), sitesDev, n, segments(ctx), st));
    APO_HIP(knnLaunch(cc, 64u, ctx->knn, st));
    APO_HIP(featuresLaunch(cc, ctx->knn, 64u, ctx->feat, st));
    APO_HIP(cellsKLaunch(cc, ctx->knn, 64u, nullptr, ctx->labels, kLabels, 16u, st));
    uint32_t guard = 0;
    APO_HIP(hipMemcpyAsync(&guard, ctx->err, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    if (guard) return fail(ctx, APO_ERR_HIP, "device guard tripped: " + guardText(guard));
    ctx->built = false;  // cells stopped at certification: not a diagram
    return APO_OK;
}

// Ray directions of the cavity app: Halton points on the sphere (Wong et
// al., "Sampling with Hammersley and Halton points"; app/cavity/utils.cu:
// 201-242 with p2 = 3 and its truncated pi), each normalised.
std::vector<F3> cavityDirections(uint32_t samples) {
    std::vector<F3> out(samples);
    for (uint32_t k = 0; k < samples; k++) {
        float t = 0.f;
        float h = 0.5f;
        for (uint32_t b = k; b; h *= .5f, b >>= 1)
            if (b & 1u) t += h;
        t = 2.f * t - 1.f;
        const float st = std::sqrt(1.f - t * t);
        float phi = 0.f;
        const float ip = 1.f / 3.f;
        float p = ip;
        for (uint32_t b = k; b; p *= ip, b /= 3u) {
            const uint32_t a = b % 3u;
            if (a) phi += float(a) * p;
        }
        const float phirad = float(double(phi * 4.f) * 3.1415926535);
        F3 d{st * std::cos(phirad), st * std::sin(phirad), t};
        const float inv = 1.f / std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        out[k] = {d.x * inv, d.y * inv, d.z * inv};
    }
    return out;
}
}  // namespace

extern "C" {

void apo_default_config(apo_config* cfg) {
    if (!cfg) return;
    cfg->k = 16;
    cfg->max_iterations = 10;
    cfg->max_restarts = 100;
    cfg->stage_mask = 7;
    cfg->waves_per_cu = 0;
    cfg->device = 0;
    cfg->proposal_cost = 0.f;
    cfg->graph = 0;
    cfg->cover_ratio = 2.f;
    cfg->waves_cells = cfg->waves_vertex = cfg->waves_edge = cfg->waves_bisector = 0;
    cfg->proposal_fallback_k = 16;
    cfg->proposal_max_k = 64;
    cfg->knn_passes = 1;
    cfg->complete = 0;
    cfg->proposal_radius = 0;
    cfg->cells_per_wave = 0;
    cfg->bisector_per_cell = 0;
}

// pcg hash of apo/src/core/utils.cpp:6-12, scaled by 2^-32
static double joggleNoise(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return std::ldexp(double((word >> 22u) ^ word), -32);
}

void apo_joggle(double* xyzr, uint64_t values, double range) {
    if (!xyzr) return;
    const double a = 2.0 * range, b = -range;
    const uint64_t count = values / 4;
    for (uint32_t i = 0; i < count; i++)
        for (uint32_t j = 0; j < 4; j++) xyzr[size_t(i) * 4 + j] += joggleNoise(i * 4u + j) * a + b;
}

void apo_joggle_f32(float* xyzr, uint64_t values, double range) {
    if (!xyzr) return;
    const double a = 2.0 * range, b = -range;
    const uint64_t count = values / 4;
    for (uint32_t i = 0; i < count; i++)
        for (uint32_t j = 0; j < 4; j++) {
            float& v = xyzr[size_t(i) * 4 + j];
            v = float(double(v) + (joggleNoise(i * 4u + j) * a + b));
        }
}

apo_status apo_create(apo_ctx** out, const apo_config* cfg) {
    if (!out) return APO_ERR_INVALID;
    *out = nullptr;
    apo_ctx* ctx = new apo_ctx();
    if (cfg) ctx->cfg = *cfg;
    else apo_default_config(&ctx->cfg);
    const uint32_t k = ctx->cfg.k, fk = ctx->cfg.proposal_fallback_k, mk = ctx->cfg.proposal_max_k;
    auto listK = [](uint32_t v) { return v == 16 || v == 24 || v == 32 || v == 48 || v == 64; };
    if (k == 0 || k > 64 || ctx->cfg.max_iterations > kMaxRounds || !listK(fk) || !listK(mk) ||
        ctx->cfg.complete > kMaxCompletionRounds || ctx->cfg.cells_per_wave > 2 || ctx->cfg.bisector_per_cell > 1) {
        delete ctx;
        return APO_ERR_INVALID;
    }
    ctx->device = int(ctx->cfg.device);
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    if (e == hipSuccess) e = staticScratchAddress(ctx->device, &ctx->staticScratch);
    if (e == hipSuccess) e = cellsOccupancy(&ctx->occCells, &ctx->occCellsPair);
    if (e == hipSuccess)
        e = validationOccupancy(&ctx->occVertex, &ctx->occEdge, &ctx->occBisector, &ctx->occBisectorPlan,
                                &ctx->occBisectorLanes);
    for (int s = 0; s < 10 && e == hipSuccess; s++) e = hipEventCreate(&ctx->ev[s]);
    for (uint32_t r = 0; r < kTimedRounds && e == hipSuccess; r++)
        for (int q = 0; q < 3 && e == hipSuccess; q++) e = hipEventCreate(&ctx->vt.ev[r][q]);
    for (uint32_t p = 0; p < kTimedPasses && e == hipSuccess; p++)
        for (int q = 0; q < 2 && e == hipSuccess; q++) e = hipEventCreate(&ctx->passEv[p][q]);
    if (e != hipSuccess) {
        apo_destroy(ctx);
        return e == hipErrorOutOfMemory ? APO_ERR_OUT_OF_MEMORY : APO_ERR_HIP;
    }
    *out = ctx;
    return APO_OK;
}

void apo_destroy(apo_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    dropGraph(ctx);
    void* ptrs[] = {ctx->offsets, ctx->moleculeVertices, ctx->sitesIn, ctx->scene, ctx->boxPartials, ctx->boxTicket, ctx->sortScratch, ctx->codesA,
                    ctx->codesB, ctx->idxA, ctx->idxB, ctx->sorted, ctx->nodes, ctx->parent,
                    ctx->refitCounter, ctx->rootBox, ctx->status, ctx->knownRadii, ctx->flags, ctx->vtx, ctx->vnb,
                    ctx->edg, ctx->enb, ctx->knn, ctx->counters, ctx->tasks, ctx->cellCount, ctx->blockSums,
                    ctx->total, ctx->outXyzw, ctx->outIds, ctx->err, ctx->bytes, ctx->range, ctx->startLists,
                    ctx->bisSearch, ctx->bisNb, ctx->bisCount, ctx->bisMark,
                    ctx->params, ctx->feat, ctx->cls, ctx->labels, ctx->givenCls, ctx->cavRatios, ctx->spill, ctx->want, ctx->stuck, ctx->givenRadii, ctx->radiiSorted, ctx->d16, ctx->pw.w1, ctx->pw.w2, ctx->pw.w3,
                    ctx->pw.b1, ctx->pw.b2, ctx->pw.b3};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& round : ctx->vt.ev)
        for (auto& e : round)
            if (e) (void)hipEventDestroy(e);
    for (auto& pass : ctx->passEv)
        for (auto& e : pass)
            if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* apo_last_error(const apo_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

}  // extern "C"

namespace {

// Checks a batch's site offsets and stores them on the device.
apo_status setMolecules(apo_ctx* ctx, const uint32_t* offsets, uint32_t molecules, uint32_t n) {
    if (molecules == 0 || molecules > kMaxMolecules) return fail(ctx, APO_ERR_INVALID, "molecule count out of range");
    if (offsets[0] != 0u || offsets[molecules] != n) return fail(ctx, APO_ERR_INVALID, "offsets must run from 0 to the site count");
    for (uint32_t m = 0; m < molecules; m++)
        if (offsets[m + 1] < offsets[m] || offsets[m + 1] - offsets[m] < 2u)
            return fail(ctx, APO_ERR_INVALID, "every molecule needs at least 2 sites");
    apo_status s = ensureMolecules(ctx, molecules);
    if (s != APO_OK) return s;
    ctx->molecules = molecules;
    if (molecules > 1u)
        APO_HIP(hipMemcpyAsync(ctx->offsets, offsets, (size_t(molecules) + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
    return APO_OK;
}

// Shared front of every build entry: argument checks, buffers, molecules.
apo_status prepareBuild(apo_ctx* ctx, uint32_t n, const uint32_t* offsets, uint32_t molecules) {
    if (n < 2) return fail(ctx, APO_ERR_INVALID, "need at least 2 sites");
    if (n >= (1u << 28)) return fail(ctx, APO_ERR_INVALID, "site count must be below 2^28");
    APO_HIP(hipSetDevice(ctx->device));
    ctx->built = false;
    if (ctx->givenN && ctx->givenN != n)
        return fail(ctx, APO_ERR_INVALID, "proposal classes were given for another site count");
    if (ctx->givenRN && ctx->givenRN != n)
        return fail(ctx, APO_ERR_INVALID, "search radii were given for another site count");
    if (ctx->givenRN && offsets && molecules > 1u)
        return fail(ctx, APO_ERR_INVALID, "search radii apply to single builds");
    if (learnedPath(ctx) && ctx->cfg.knn_passes != 1u)
        return fail(ctx, APO_ERR_INVALID, "the learned proposal runs one kNN pass (knn_passes = 1)");
    ctx->cavValid = false;
    // per-wave device regions of the vertex, edge and bisector validation
    // grids (the kernels run one after the other on the stream and share them)
    const uint32_t vblocks = std::max(std::max(std::max(persistentBlocks(ctx, ctx->cfg.waves_vertex, ctx->occVertex),
                                                        persistentBlocks(ctx, ctx->cfg.waves_edge, ctx->occEdge)),
                                               persistentBlocks(ctx, ctx->cfg.waves_bisector, ctx->occBisector)),
                                      persistentBlocks(ctx, 0u, ctx->occBisectorLanes));
    if (vblocks > ctx->spillBlocks) {
        dropGraph(ctx);
        if (ctx->spill) APO_HIP(hipFree(ctx->spill));
        ctx->spill = nullptr;
        ctx->spillBlocks = 0;
        APO_HIP(alloc(ctx->spill, size_t(vblocks) * waveScratchWords()));
        ctx->spillBlocks = vblocks;
    }
    apo_status s = ensureCapacity(ctx, n, learnedPath(ctx) ? (ctx->cfg.k > 64u ? ctx->cfg.k : 64u) : ctx->cfg.k);
    if (s != APO_OK) return s;
    if (ctx->cfg.complete && n > ctx->capW) {
        dropGraph(ctx);
        APO_HIP(alloc(ctx->want, n));
        ctx->capW = n;
    }
    if (ctx->cfg.complete && !ctx->stuck) {
        dropGraph(ctx);
        APO_HIP(alloc(ctx->stuck, kStuckCap));
    }
    if (learnedPath(ctx)) {
        s = ensureProposalBuffers(ctx, n);
        if (s != APO_OK) return s;
    }
    const uint32_t one[2] = {0u, n};
    return setMolecules(ctx, offsets ? offsets : one, offsets ? molecules : 1u, n);
}

}  // namespace

extern "C" {

apo_status apo_build_device(apo_ctx* ctx, const float* xyzr_device, uint32_t n) {
    if (!ctx || !xyzr_device) return APO_ERR_INVALID;
    apo_status s = prepareBuild(ctx, n, nullptr, 1);
    if (s != APO_OK) return s;
    return runPipeline(ctx, reinterpret_cast<const F4*>(xyzr_device), n);
}

apo_status apo_build(apo_ctx* ctx, const float* xyzr, uint32_t n) {
    if (!ctx || !xyzr) return APO_ERR_INVALID;
    apo_status s = prepareBuild(ctx, n, nullptr, 1);
    if (s != APO_OK) return s;
    APO_HIP(hipMemcpyAsync(ctx->sitesIn, xyzr, size_t(n) * sizeof(F4), hipMemcpyHostToDevice, ctx->stream));
    return runPipeline(ctx, ctx->sitesIn, n);
}

apo_status apo_build_batched_device(apo_ctx* ctx, const float* xyzr_device, const uint32_t* offsets,
                                    uint32_t molecules) {
    if (!ctx || !xyzr_device || !offsets) return APO_ERR_INVALID;
    apo_status s = prepareBuild(ctx, offsets[molecules], offsets, molecules);
    if (s != APO_OK) return s;
    return runPipeline(ctx, reinterpret_cast<const F4*>(xyzr_device), offsets[molecules]);
}

apo_status apo_build_batched(apo_ctx* ctx, const float* xyzr, const uint32_t* offsets, uint32_t molecules) {
    if (!ctx || !xyzr || !offsets) return APO_ERR_INVALID;
    const uint32_t n = offsets[molecules];
    apo_status s = prepareBuild(ctx, n, offsets, molecules);
    if (s != APO_OK) return s;
    APO_HIP(hipMemcpyAsync(ctx->sitesIn, xyzr, size_t(n) * sizeof(F4), hipMemcpyHostToDevice, ctx->stream));
    return runPipeline(ctx, ctx->sitesIn, n);
}

}  // extern "C"

namespace {

// The device buffers of a build, by name: apo_build_device_static points
// them into the static scratch array for one call and restores the
// context's own afterwards.
#define APO_BUILD_BUFFERS(X)                                                                                     \
    X(scene) X(boxPartials) X(boxTicket) X(sortScratch) X(range) X(startLists) X(codesA) X(codesB) X(idxA) X(idxB) \
    X(sorted) X(nodes) X(parent) X(refitCounter) X(rootBox) X(status) X(knownRadii) X(flags) X(vtx)       \
    X(vnb) X(edg) X(enb) X(knn) X(counters) X(spill) X(err) X(tasks) X(cellCount) X(blockSums) X(total)            \
    X(outXyzw) X(outIds) X(bytes) X(want) X(stuck) X(bisSearch) X(bisNb) X(bisCount) X(bisMark)

struct BuildBuffers {
#define APO_DECLARE(f) decltype(apo_ctx::f) f;
    APO_BUILD_BUFFERS(APO_DECLARE)
#undef APO_DECLARE
    uint32_t spillBlocks;
    uint64_t outCap;
};

BuildBuffers takeBuffers(const apo_ctx* ctx) {
    BuildBuffers b;
#define APO_TAKE(f) b.f = ctx->f;
    APO_BUILD_BUFFERS(APO_TAKE)
#undef APO_TAKE
    b.spillBlocks = ctx->spillBlocks;
    b.outCap = ctx->outCap;
    return b;
}

void putBuffers(apo_ctx* ctx, const BuildBuffers& b) {
#define APO_PUT(f) ctx->f = b.f;
    APO_BUILD_BUFFERS(APO_PUT)
#undef APO_PUT
    ctx->spillBlocks = b.spillBlocks;
    ctx->outCap = b.outCap;
}

// Consecutive 256-byte aligned pieces of the static array; `fits` turns
// false when the pieces overrun it.
struct Carve {
    char* base;
    size_t used = 0;
    bool fits = true;
    template <typename T>
    void take(T*& p, size_t count) {
        used = (used + 255u) & ~size_t(255u);
        p = reinterpret_cast<T*>(base + used);
        used += (count ? count : 1u) * sizeof(T);
        if (used > size_t(APO_STATIC_SCRATCH_BYTES)) fits = false;
    }
};

// The buffers of one build of n <= kStaticMaxSites sites with K-long
// neighbour lists in the static array (device address `base`), the vertex
// output in the caller's buffers.
bool carveStatic(void* base, uint32_t n, uint32_t K, double* xyzw, uint32_t* ids, uint64_t capacity,
                 BuildBuffers& b) {
    Carve c{static_cast<char*>(base)};
    const size_t nodes = 2 * size_t(n) - 1;
    c.take(b.scene, 1);
    c.take(b.boxPartials, size_t(kBoxBlocks) * 8);
    c.take(b.boxTicket, 1);
    c.take(b.sortScratch, sortScratchWords(n));
    c.take(b.range, 2 * size_t(n));
    c.take(b.startLists, refitStartWords(n));
    c.take(b.codesA, n);
    c.take(b.codesB, n);
    c.take(b.idxA, n);
    c.take(b.idxB, n);
    c.take(b.sorted, n);
    c.take(b.nodes, n);
    c.take(b.parent, nodes);
    c.take(b.refitCounter, n);
    c.take(b.rootBox, 8);
    c.take(b.status, n);
    c.take(b.knownRadii, n);
    c.take(b.flags, n);
    c.take(b.vtx, size_t(n) * kVertexCap);
    c.take(b.vnb, n);
    c.take(b.edg, size_t(n) * kEdgeCap);
    c.take(b.enb, n);
    c.take(b.knn, size_t(n) * K);
    c.take(b.counters, kCounterWords);
    c.take(b.err, 1);
    c.take(b.tasks, n);
    c.take(b.cellCount, n);
    c.take(b.blockSums, (n + 255u) / 256u + 2u);
    c.take(b.total, 1);
    c.take(b.bytes, size_t(kByteFamilies) * kByteSlots);
    c.take(b.want, n);
    c.take(b.stuck, kStuckCap);
    c.take(b.spill, size_t(kStaticWaveRegions) * waveScratchWords());
    b.spillBlocks = kStaticWaveRegions;
    // the bisector stage runs per cell (no room for the split stage's lists)
    b.bisSearch = nullptr;
    b.bisNb = nullptr;
    b.bisCount = nullptr;
    b.bisMark = nullptr;
    b.outXyzw = reinterpret_cast<D4*>(xyzw);
    b.outIds = reinterpret_cast<uint4*>(ids);
    b.outCap = capacity;
    return c.fits;
}

}  // namespace

extern "C" {

apo_status apo_build_device_static(apo_ctx* ctx, const float* xyzr_device, uint32_t n, double* xyzw_device,
                                   uint32_t* ids_device, uint64_t capacity, uint64_t* count) {
    if (!ctx || !xyzr_device || !count || (capacity && (!xyzw_device || !ids_device))) return APO_ERR_INVALID;
    if (n < 2) return fail(ctx, APO_ERR_INVALID, "need at least 2 sites");
    if (n > kStaticMaxSites)
        return fail(ctx, APO_ERR_INVALID, "apo_build_device_static takes at most " +
                                              std::to_string(kStaticMaxSites) + " sites (APO_STATIC_MAX_SITES)");
    if (learnedPath(ctx)) return fail(ctx, APO_ERR_INVALID, "apo_build_device_static runs the plain path only");
    if (ctx->cfg.stage_mask & 8u)
        return fail(ctx, APO_ERR_INVALID, "apo_build_device_static runs the whole build (stage_mask bit 3 clear)");
    if (!ctx->staticScratch) return fail(ctx, APO_ERR_STATE, "no static scratch address for this device");
    APO_HIP(hipSetDevice(ctx->device));
    BuildBuffers mine{};
    if (!carveStatic(ctx->staticScratch, n, ctx->cfg.k, xyzw_device, ids_device, capacity, mine))
        return fail(ctx, APO_ERR_CAPACITY, "the build does not fit the static scratch array");
    // the context's own buffers and the state that describes them come back
    // on every return path
    struct Restore {
        apo_ctx* ctx;
        BuildBuffers own;
        uint32_t n, molecules;
        ~Restore() {
            putBuffers(ctx, own);
            ctx->n = n;
            ctx->molecules = molecules;
        }
    } restore{ctx, takeBuffers(ctx), ctx->n, ctx->molecules};
    ctx->built = false;  // the context's buffers no longer hold a complete build
    ctx->cavValid = false;
    putBuffers(ctx, mine);
    ctx->n = n;
    ctx->molecules = 1;
    hipStream_t st = ctx->stream;
    // the one word no kernel of the build clears
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, ctx->boxTicket, 1u, 0u);
    APO_HIP(hipGetLastError());
    APO_HIP(hipEventRecord(ctx->ev[0], st));
    const apo_status s = enqueueBuild(ctx, reinterpret_cast<const F4*>(xyzr_device), n, false, false);
    if (s != APO_OK) return s;
    APO_HIP(hipEventRecord(ctx->ev[6], st));
    uint32_t total = 0, guard = 0;
    APO_HIP(hipMemcpyAsync(&total, ctx->total, sizeof(total), hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(&guard, ctx->err, sizeof(guard), hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    if (guard) return fail(ctx, APO_ERR_HIP, "device guard tripped: " + guardText(guard));
    *count = total;
    if (total > capacity)
        return fail(ctx, APO_ERR_CAPACITY, "output buffers hold " + std::to_string(capacity) + " of " +
                                                std::to_string(total) + " vertices");
    return APO_OK;
}

apo_status apo_launch_occupancy(apo_ctx* ctx, uint32_t* out5) {
    if (!ctx || !out5) return APO_ERR_INVALID;
    out5[0] = ctx->occCells;
    out5[1] = ctx->occVertex;
    out5[2] = ctx->occEdge;
    out5[3] = ctx->occBisector;
    out5[4] = uint32_t(ctx->cus);
    return APO_OK;
}

apo_status apo_batch_vertex_offsets(apo_ctx* ctx, uint64_t* offsets) {
    if (!ctx || !offsets) return APO_ERR_INVALID;
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    if (ctx->molecules == 1u) {
        offsets[0] = 0;
        offsets[1] = ctx->vertexCount;
        return APO_OK;
    }
    APO_HIP(hipMemcpy(offsets, ctx->moleculeVertices, (size_t(ctx->molecules) + 1) * 8, hipMemcpyDeviceToHost));
    return APO_OK;
}

apo_status apo_vertex_count(apo_ctx* ctx, uint64_t* count) {
    if (!ctx || !count) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    *count = ctx->vertexCount;
    return APO_OK;
}

apo_status apo_get_vertices(apo_ctx* ctx, double* xyzw, uint32_t* ids, uint64_t capacity) {
    if (!ctx) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    if (capacity < ctx->vertexCount) return fail(ctx, APO_ERR_CAPACITY, "output buffer too small");
    const size_t v = ctx->vertexCount;
    if (xyzw) APO_HIP(hipMemcpy(xyzw, ctx->outXyzw, v * sizeof(D4), hipMemcpyDeviceToHost));
    if (ids) APO_HIP(hipMemcpy(ids, ctx->outIds, v * sizeof(uint4), hipMemcpyDeviceToHost));
    return APO_OK;
}

apo_status apo_get_device_vertices(apo_ctx* ctx, const double** xyzw, const uint32_t** ids) {
    if (!ctx) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    if (xyzw) *xyzw = reinterpret_cast<const double*>(ctx->outXyzw);
    if (ids) *ids = reinterpret_cast<const uint32_t*>(ctx->outIds);
    return APO_OK;
}

namespace {
// Closed trisector edges of the current build: count (and write, when xyzw
// or ids is given) on the context's stream with temporary buffers; the
// build's own output buffers are left untouched.
apo_status closedEdges(apo_ctx* ctx, uint64_t* count, double* xyzw, uint32_t* ids, uint64_t capacity) {
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t n = ctx->n;
    const uint32_t blocks = uint32_t((n + 255u) / 256u);
    DeviceTemp<uint32_t> cnt, sums, total;
    APO_HIP(cnt.alloc(n));
    APO_HIP(sums.alloc(blocks));
    APO_HIP(total.alloc(1));
    const CellCtx cc = cellCtx(ctx);
    OutputBuffers ob{cnt.p, sums.p, total.p, nullptr, nullptr, 0};
    APO_HIP(closedEdgesCountLaunch(cc, ob, st));
    uint32_t e = 0;
    APO_HIP(hipMemcpyAsync(&e, total.p, 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    *count = e;
    if (!xyzw && !ids) return APO_OK;
    if (capacity < e) return fail(ctx, APO_ERR_CAPACITY, "output buffer too small");
    DeviceTemp<D4> dx;
    DeviceTemp<uint4> di;
    APO_HIP(dx.alloc(e));
    APO_HIP(di.alloc(e));
    ob.xyzw = dx.p;
    ob.ids = di.p;
    ob.capacity = e;
    APO_HIP(closedEdgesWriteLaunch(cc, ob, ctx->idxA, st));
    if (xyzw) APO_HIP(hipMemcpyAsync(xyzw, dx.p, size_t(e) * sizeof(D4), hipMemcpyDeviceToHost, st));
    if (ids) APO_HIP(hipMemcpyAsync(ids, di.p, size_t(e) * sizeof(uint4), hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    return APO_OK;
}
}  // namespace

apo_status apo_closed_edge_count(apo_ctx* ctx, uint64_t* count) {
    if (!ctx || !count) return APO_ERR_INVALID;
    return closedEdges(ctx, count, nullptr, nullptr, 0);
}

apo_status apo_get_closed_edges(apo_ctx* ctx, double* xyzw, uint32_t* ids, uint64_t capacity) {
    if (!ctx || (!xyzw && !ids)) return APO_ERR_INVALID;
    uint64_t count = 0;
    return closedEdges(ctx, &count, xyzw, ids, capacity);
}

apo_status apo_get_stats(apo_ctx* ctx, apo_stats* stats) {
    if (!ctx || !stats) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    const uint32_t n = ctx->n;
    std::vector<uint8_t> st(n);
    std::vector<uint32_t> fl(n);
    APO_HIP(hipMemcpy(st.data(), ctx->status, n, hipMemcpyDeviceToHost));
    APO_HIP(hipMemcpy(fl.data(), ctx->flags, size_t(n) * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> vn(n), en(n);
    APO_HIP(hipMemcpy(vn.data(), ctx->vnb, size_t(n) * 4, hipMemcpyDeviceToHost));
    APO_HIP(hipMemcpy(en.data(), ctx->enb, size_t(n) * 4, hipMemcpyDeviceToHost));
    uint64_t buried = 0, uncertified = 0, flagged = 0, vr = 0, er = 0;
    for (uint32_t i = 0; i < n; i++) {
        buried += st[i] == kStatusBuried;
        uncertified += st[i] == kStatusInValidation;
        flagged += fl[i] != 0;
        vr += vn[i];
        er += en[i];
    }
    ctx->stats.buried = buried;
    ctx->stats.uncertified = uncertified;
    ctx->stats.flagged_cells = flagged;
    ctx->stats.vertex_records = vr;
    ctx->stats.edge_records = er;
    const uint64_t firstRound = ctx->roundTasks.empty() ? 0u : ctx->roundTasks[0];
    ctx->stats.cells_certified = n >= buried + firstRound ? n - buried - firstRound : 0u;
    *stats = ctx->stats;
    return APO_OK;
}

apo_status apo_check_empty_spheres(apo_ctx* ctx, const double* xyzw, const uint32_t* ids, uint64_t count, double tol,
                                   uint64_t* violations) {
    if (!ctx || !violations || (xyzw == nullptr) != (ids == nullptr)) return APO_ERR_INVALID;
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const D4* dx = ctx->outXyzw;
    const uint4* di = ctx->outIds;
    DeviceTemp<D4> ux;
    DeviceTemp<uint4> ui;
    if (xyzw) {
        APO_HIP(ux.alloc(count));
        APO_HIP(ui.alloc(count));
        APO_HIP(hipMemcpyAsync(ux.p, xyzw, count * sizeof(D4), hipMemcpyHostToDevice, st));
        APO_HIP(hipMemcpyAsync(ui.p, ids, count * sizeof(uint4), hipMemcpyHostToDevice, st));
        dx = ux.p;
        di = ui.p;
    } else {
        count = ctx->vertexCount;
    }
    DeviceTemp<unsigned long long> dv;
    APO_HIP(dv.alloc(1));
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(dv.p), 2u, 0u);
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, ctx->err, 1u, 0u);
    APO_HIP(checkEmptyLaunch(cellCtx(ctx), ctx->idxA, ctx->cellCount, dx, di, count, tol, dv.p, st));
    unsigned long long bad = 0;
    uint32_t guard = 0;
    APO_HIP(hipMemcpyAsync(&bad, dv.p, sizeof(bad), hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(&guard, ctx->err, sizeof(guard), hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    if (guard) return fail(ctx, APO_ERR_HIP, "device guard tripped in the empty-sphere check: " + guardText(guard));
    *violations = bad;
    return APO_OK;
}

namespace {
// The interval validator over device points; counts[0] bad, counts[1] uncertain.
apo_status precisionOnDevice(apo_ctx* ctx, const D4* dx, const uint4* di, uint64_t count, uint32_t arity,
                             uint64_t* bad, uint64_t* uncertain) {
    hipStream_t st = ctx->stream;
    DeviceTemp<unsigned long long> dc;
    APO_HIP(dc.alloc(2));
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(dc.p), 4u, 0u);
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, ctx->err, 1u, 0u);
    APO_HIP(precisionLaunch(cellCtx(ctx), ctx->idxA, ctx->cellCount, dx, di, count, arity, dc.p, st));
    unsigned long long h[2] = {0, 0};
    uint32_t guard = 0;
    APO_HIP(hipMemcpyAsync(h, dc.p, sizeof(h), hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(&guard, ctx->err, sizeof(guard), hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    if (guard) return fail(ctx, APO_ERR_HIP, "device guard tripped in the precision check: " + guardText(guard));
    *bad = h[0];
    *uncertain = h[1];
    return APO_OK;
}
}  // namespace

apo_status apo_precision_check_points(apo_ctx* ctx, const double* xyzw, const uint32_t* ids, uint64_t count,
                                      uint32_t arity, uint64_t* bad, uint64_t* uncertain) {
    if (!ctx || !bad || !uncertain || (count && (!xyzw || !ids)) || (arity != 3u && arity != 4u))
        return APO_ERR_INVALID;
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    APO_HIP(hipSetDevice(ctx->device));
    DeviceTemp<D4> dx;
    DeviceTemp<uint4> di;
    APO_HIP(dx.alloc(count));
    APO_HIP(di.alloc(count));
    if (count) {
        APO_HIP(hipMemcpyAsync(dx.p, xyzw, count * sizeof(D4), hipMemcpyHostToDevice, ctx->stream));
        APO_HIP(hipMemcpyAsync(di.p, ids, count * sizeof(uint4), hipMemcpyHostToDevice, ctx->stream));
    }
    return precisionOnDevice(ctx, dx.p, di.p, count, arity, bad, uncertain);
}

apo_status apo_precision_check(apo_ctx* ctx, apo_precision* out) {
    if (!ctx || !out) return APO_ERR_INVALID;
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    APO_HIP(hipSetDevice(ctx->device));
    apo_precision r{};
    r.quadrisectors = ctx->vertexCount;
    apo_status s = precisionOnDevice(ctx, ctx->outXyzw, ctx->outIds, ctx->vertexCount, 4u, &r.bad_quadrisectors,
                                     &r.uncertain_quadrisectors);
    if (s != APO_OK) return s;
    // closed trisector edges, written to temporary device buffers
    uint64_t e = 0;
    s = apo_closed_edge_count(ctx, &e);
    if (s != APO_OK) return s;
    r.trisectors = e;
    if (e) {
        const size_t n = ctx->n;
        const uint32_t blocks = uint32_t((n + 255u) / 256u);
        DeviceTemp<uint32_t> cnt, sums, total;
        DeviceTemp<D4> dx;
        DeviceTemp<uint4> di;
        APO_HIP(cnt.alloc(n));
        APO_HIP(sums.alloc(blocks));
        APO_HIP(total.alloc(1));
        APO_HIP(dx.alloc(e));
        APO_HIP(di.alloc(e));
        const CellCtx cc = cellCtx(ctx);
        const OutputBuffers ob{cnt.p, sums.p, total.p, dx.p, di.p, e};
        APO_HIP(closedEdgesCountLaunch(cc, ob, ctx->stream));
        APO_HIP(closedEdgesWriteLaunch(cc, ob, ctx->idxA, ctx->stream));
        s = precisionOnDevice(ctx, dx.p, di.p, e, 3u, &r.bad_trisectors, &r.uncertain_trisectors);
        if (s != APO_OK) return s;
    }
    *out = r;
    return APO_OK;
}


apo_status apo_cavity_hit_ratios(apo_ctx* ctx, uint32_t samples, float* hit_ratios) {
    if (!ctx || samples == 0 || samples > 4096u) return APO_ERR_INVALID;
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint64_t V = ctx->vertexCount;
    if (V > ctx->cavCap) {
        if (ctx->cavRatios) APO_HIP(hipFree(ctx->cavRatios));
        ctx->cavRatios = nullptr;
        ctx->cavCap = 0;
        APO_HIP(alloc(ctx->cavRatios, V));
        ctx->cavCap = V;
    }
    const std::vector<F3> dirs = cavityDirections(samples);
    DeviceTemp<F3> dd;
    APO_HIP(dd.alloc(samples));
    APO_HIP(hipMemcpyAsync(dd.p, dirs.data(), samples * sizeof(F3), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, ctx->err, 1u, 0u);
    APO_HIP(hitRatiosLaunch(cellCtx(ctx), ctx->idxA, ctx->cellCount, ctx->outXyzw, ctx->outIds, V, dd.p, samples,
                            ctx->cavRatios, st));
    uint32_t guard = 0;
    APO_HIP(hipMemcpyAsync(&guard, ctx->err, sizeof(guard), hipMemcpyDeviceToHost, st));
    if (hit_ratios && V) APO_HIP(hipMemcpyAsync(hit_ratios, ctx->cavRatios, V * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    if (guard) return fail(ctx, APO_ERR_HIP, "device guard tripped in the hit ratios: " + guardText(guard));
    ctx->cavValid = true;
    return APO_OK;
}

apo_status apo_cavity_vertices(apo_ctx* ctx, float hit_ratio_threshold, float radius_threshold, float* xyzw,
                               uint64_t capacity, uint64_t* count) {
    if (!ctx || !count) return APO_ERR_INVALID;
    if (!ctx->built || !ctx->cavValid) return fail(ctx, APO_ERR_STATE, "no hit ratios for the current build");
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const uint64_t V = ctx->vertexCount;
    const uint32_t blocks = uint32_t((V + 255u) / 256u);
    DeviceTemp<uint32_t> sums, total;
    APO_HIP(sums.alloc(blocks));
    APO_HIP(total.alloc(1));
    uint32_t kept = 0;
    if (V) {
        APO_HIP(cavityCountLaunch(ctx->outXyzw, ctx->cavRatios, V, hit_ratio_threshold, radius_threshold, sums.p,
                                  total.p, st));
        APO_HIP(hipMemcpyAsync(&kept, total.p, 4, hipMemcpyDeviceToHost, st));
        APO_HIP(hipStreamSynchronize(st));
    }
    *count = kept;
    if (!xyzw || kept == 0) return APO_OK;
    if (capacity < kept) return fail(ctx, APO_ERR_CAPACITY, "output buffer too small");
    DeviceTemp<float4> out;
    APO_HIP(out.alloc(kept));
    APO_HIP(cavityWriteLaunch(ctx->outXyzw, ctx->cavRatios, V, hit_ratio_threshold, radius_threshold, sums.p, out.p,
                              kept, st));
    APO_HIP(hipMemcpyAsync(xyzw, out.p, size_t(kept) * sizeof(float4), hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    return APO_OK;
}

apo_status apo_get_sorted_sites(apo_ctx* ctx, float* xyzr, uint32_t* indices) {
    if (!ctx) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    if (xyzr) APO_HIP(hipMemcpy(xyzr, ctx->sorted, size_t(ctx->n) * sizeof(F4), hipMemcpyDeviceToHost));
    if (indices) APO_HIP(hipMemcpy(indices, ctx->idxA, size_t(ctx->n) * 4, hipMemcpyDeviceToHost));
    return APO_OK;
}

apo_status apo_get_lbvh(apo_ctx* ctx, uint64_t* sorted_codes, int32_t* parent, int32_t* left, int32_t* right,
                        float* aabb, uint32_t* sizes, float* max_radius) {
    if (!ctx) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    const uint32_t n = ctx->n;
    const uint32_t internal = n - ctx->molecules;
    const size_t nodes = size_t(n) + internal;
    if (sorted_codes) {
        // the key's bits 60-63 carry the molecule index of a batch
        APO_HIP(hipMemcpy(sorted_codes, ctx->codesA, size_t(n) * 8, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) sorted_codes[i] &= (uint64_t(1) << 60) - 1u;
    }
    std::vector<int32_t> par(nodes);
    APO_HIP(hipMemcpy(par.data(), ctx->parent, nodes * 4, hipMemcpyDeviceToHost));
    if (parent) std::memcpy(parent, par.data(), nodes * 4);
    if (sizes) {
        // leaf counts: one per leaf, last - first + 1 per internal node
        std::vector<uint32_t> rg(size_t(internal) * 2);
        APO_HIP(hipMemcpy(rg.data(), ctx->range, rg.size() * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < internal; i++) sizes[i] = rg[2 * size_t(i) + 1] - rg[2 * size_t(i)] + 1u;
        for (size_t i = internal; i < nodes; i++) sizes[i] = 1u;
    }
    std::vector<Node> nd(internal);
    APO_HIP(hipMemcpy(nd.data(), ctx->nodes, size_t(internal) * sizeof(Node), hipMemcpyDeviceToHost));
    // unpack the child-in-parent records into per-node arrays; a root's box
    // is the union of its two child slots
    std::vector<float> box(nodes * 6, 0.f), mr(nodes, 0.f);
    std::vector<int32_t> L(nodes, 0), R(nodes, 0);
    for (uint32_t i = 0; i < internal; i++) {
        const float* f = reinterpret_cast<const float*>(&nd[i]);
        const int32_t lc = nd[i].d.x, rc = nd[i].d.y;
        L[i] = lc;
        R[i] = rc;
        for (int c = 0; c < 6; c++) {
            box[size_t(lc) * 6 + c] = f[c];
            box[size_t(rc) * 6 + c] = f[6 + c];
        }
        mr[lc] = f[14];
        mr[rc] = f[15];
        if (par[i] < 0) {
            for (int c = 0; c < 3; c++) {
                box[size_t(i) * 6 + c] = std::fmin(f[c], f[6 + c]);
                box[size_t(i) * 6 + 3 + c] = std::fmax(f[3 + c], f[9 + c]);
            }
            mr[i] = std::fmax(f[14], f[15]);
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        L[internal + i] = int32_t(i);
        R[internal + i] = int32_t(i) + 1;
    }
    if (left) std::memcpy(left, L.data(), nodes * 4);
    if (right) std::memcpy(right, R.data(), nodes * 4);
    if (aabb) std::memcpy(aabb, box.data(), nodes * 24);
    if (max_radius) std::memcpy(max_radius, mr.data(), nodes * 4);
    return APO_OK;
}

apo_status apo_get_artificial_sites(apo_ctx* ctx, double* xyzr24) {
    if (!ctx || !xyzr24) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    APO_HIP(hipMemcpy(xyzr24, &ctx->scene->art[0], 6 * sizeof(D4), hipMemcpyDeviceToHost));
    return APO_OK;
}

apo_status apo_get_knns(apo_ctx* ctx, uint32_t* knn) {
    if (!ctx || !knn) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    APO_HIP(hipMemcpy(knn, ctx->knn, size_t(ctx->n) * ctx->cfg.k * 4, hipMemcpyDeviceToHost));
    return APO_OK;
}

apo_status apo_get_cell_counts(apo_ctx* ctx, uint32_t* vertex_nb, uint32_t* edge_nb, uint8_t* status,
                               float* known_radius, uint32_t* flags) {
    if (!ctx) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    const size_t n = ctx->n;
    if (vertex_nb) APO_HIP(hipMemcpy(vertex_nb, ctx->vnb, n * 4, hipMemcpyDeviceToHost));
    if (edge_nb) APO_HIP(hipMemcpy(edge_nb, ctx->enb, n * 4, hipMemcpyDeviceToHost));
    if (status) APO_HIP(hipMemcpy(status, ctx->status, n, hipMemcpyDeviceToHost));
    if (known_radius) APO_HIP(hipMemcpy(known_radius, ctx->knownRadii, n * 4, hipMemcpyDeviceToHost));
    if (flags) APO_HIP(hipMemcpy(flags, ctx->flags, n * 4, hipMemcpyDeviceToHost));
    return APO_OK;
}

apo_status apo_get_all_cells(apo_ctx* ctx, uint32_t* vertices, uint32_t* edges) {
    if (!ctx) return APO_ERR_INVALID;
    if (!ctx->built) return fail(ctx, APO_ERR_STATE, "no build");
    const size_t n = ctx->n;
    std::vector<uint32_t> vn(n), en(n);
    APO_HIP(hipMemcpy(vn.data(), ctx->vnb, n * 4, hipMemcpyDeviceToHost));
    APO_HIP(hipMemcpy(en.data(), ctx->enb, n * 4, hipMemcpyDeviceToHost));
    std::vector<uint4> all;
    size_t vo = 0, eo = 0;
    if (vertices) {
        all.resize(n * kVertexCap);
        APO_HIP(hipMemcpy(all.data(), ctx->vtx, all.size() * sizeof(uint4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            std::memcpy(vertices + 4 * vo, all.data() + i * kVertexCap, vn[i] * sizeof(uint4));
            vo += vn[i];
        }
    }
    if (edges) {
        all.resize(n * kEdgeCap);
        APO_HIP(hipMemcpy(all.data(), ctx->edg, all.size() * sizeof(uint4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            std::memcpy(edges + 4 * eo, all.data() + i * kEdgeCap, en[i] * sizeof(uint4));
            eo += en[i];
        }
    }
    return APO_OK;
}

apo_status apo_get_topology_sets(apo_ctx* ctx, uint32_t* bisectors, uint64_t bisector_cap, uint32_t* trisectors,
                                 uint64_t trisector_cap, uint32_t* quadrisectors, uint64_t quadrisector_cap,
                                 uint64_t* counts3) {
    if (!ctx || !counts3) return APO_ERR_INVALID;
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t n = ctx->n;
    const CellCtx cc = cellCtx(ctx);
    DeviceTemp<uint32_t> cnt;
    APO_HIP(cnt.alloc(3 * n));
    APO_HIP(topologyCountLaunch(cc, cnt.p, st));
    std::vector<uint32_t> hc(3 * n);
    APO_HIP(hipMemcpyAsync(hc.data(), cnt.p, 3 * n * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> off(3 * n);
    uint64_t tot[3] = {0, 0, 0};
    for (size_t i = 0; i < n; i++)
        for (int t = 0; t < 3; t++) {
            off[3 * i + t] = tot[t];
            tot[t] += hc[3 * i + t];
        }
    for (int t = 0; t < 3; t++) counts3[t] = tot[t];
    if (!bisectors && !trisectors && !quadrisectors) return APO_OK;
    if (!bisectors || !trisectors || !quadrisectors) return fail(ctx, APO_ERR_INVALID, "pass all three arrays or none");
    if (bisector_cap < tot[0] || trisector_cap < tot[1] || quadrisector_cap < tot[2])
        return fail(ctx, APO_ERR_CAPACITY, "topology buffers too small");
    DeviceTemp<uint64_t> doff;
    DeviceTemp<uint32_t> b, t, q;
    APO_HIP(doff.alloc(3 * n));
    APO_HIP(b.alloc(2 * tot[0]));
    APO_HIP(t.alloc(3 * tot[1]));
    APO_HIP(q.alloc(4 * tot[2]));
    APO_HIP(hipMemcpyAsync(doff.p, off.data(), 3 * n * 8, hipMemcpyHostToDevice, st));
    APO_HIP(topologyWriteLaunch(cc, doff.p, ctx->idxA, b.p, t.p, q.p, st));
    APO_HIP(hipMemcpyAsync(bisectors, b.p, 2 * tot[0] * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(trisectors, t.p, 3 * tot[1] * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(quadrisectors, q.p, 4 * tot[2] * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    return APO_OK;
}

apo_status apo_get_cell_topology(apo_ctx* ctx, uint64_t* vertex_offsets, uint32_t* vertices, uint64_t vertex_cap,
                                 uint64_t* edge_offsets, uint32_t* edges, uint64_t edge_cap) {
    if (!ctx) return APO_ERR_INVALID;
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    if ((vertices || edges) && !(vertices && edges)) return fail(ctx, APO_ERR_INVALID, "pass both record arrays or none");
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t n = ctx->n;
    std::vector<uint32_t> vn(n), en(n), idx(n);
    std::vector<uint8_t> status(n);
    APO_HIP(hipMemcpyAsync(vn.data(), ctx->vnb, n * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(en.data(), ctx->enb, n * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(idx.data(), ctx->idxA, n * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(status.data(), ctx->status, n, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> vc(n + 1, 0), ec(n + 1, 0);
    for (size_t i = 0; i < n; i++) {
        if (status[i] == kStatusBuried) continue;
        vc[idx[i]] = vn[i] < kVertexCap ? vn[i] : kVertexCap;
        ec[idx[i]] = en[i] < kEdgeCap ? en[i] : kEdgeCap;
    }
    uint64_t vt = 0, et = 0;
    for (size_t o = 0; o <= n; o++) {
        const uint64_t a = vc[o], b = ec[o];
        vc[o] = vt;
        ec[o] = et;
        vt += a;
        et += b;
    }
    if (vertex_offsets) std::memcpy(vertex_offsets, vc.data(), (n + 1) * 8);
    if (edge_offsets) std::memcpy(edge_offsets, ec.data(), (n + 1) * 8);
    if (!vertices) return APO_OK;
    if (vertex_cap < vt || edge_cap < et) return fail(ctx, APO_ERR_CAPACITY, "record buffers too small");
    DeviceTemp<uint64_t> dvo, deo;
    DeviceTemp<uint4> dv, de;
    APO_HIP(dvo.alloc(n + 1));
    APO_HIP(deo.alloc(n + 1));
    APO_HIP(dv.alloc(vt));
    APO_HIP(de.alloc(et));
    APO_HIP(hipMemcpyAsync(dvo.p, vc.data(), (n + 1) * 8, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(deo.p, ec.data(), (n + 1) * 8, hipMemcpyHostToDevice, st));
    APO_HIP(cellRecordsLaunch(cellCtx(ctx), ctx->idxA, dvo.p, deo.p, dv.p, de.p, st));
    APO_HIP(hipMemcpyAsync(vertices, dv.p, vt * sizeof(uint4), hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(edges, de.p, et * sizeof(uint4), hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    return APO_OK;
}

apo_status apo_proposal_labels(apo_ctx* ctx, const float* xyzr, uint32_t n, float* features, uint8_t* labels) {
    if (!ctx || !xyzr || !labels) return APO_ERR_INVALID;
    if (n < 17) return fail(ctx, APO_ERR_INVALID, "the label pass needs at least 17 sites");
    if (n >= (1u << 28)) return fail(ctx, APO_ERR_INVALID, "site count must be below 2^28");
    APO_HIP(hipSetDevice(ctx->device));
    ctx->built = false;
    apo_status s = ensureCapacity(ctx, n, ctx->cfg.k > 64u ? ctx->cfg.k : 64u);
    if (s != APO_OK) return s;
    s = ensureProposalBuffers(ctx, n);
    if (s != APO_OK) return s;
    hipStream_t st = ctx->stream;
    APO_HIP(hipMemcpyAsync(ctx->sitesIn, xyzr, size_t(n) * sizeof(F4), hipMemcpyHostToDevice, st));
    s = ensureMolecules(ctx, 1);
    if (s != APO_OK) return s;
    ctx->molecules = 1;
    s = runLabelPass(ctx, ctx->sitesIn, n);
    if (s != APO_OK) return s;
    std::vector<uint32_t> idx(n);
    std::vector<uint8_t> lab(n);
    APO_HIP(hipMemcpy(idx.data(), ctx->idxA, size_t(n) * 4, hipMemcpyDeviceToHost));
    APO_HIP(hipMemcpy(lab.data(), ctx->labels, n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; i++) labels[idx[i]] = lab[i];
    if (features) {
        std::vector<float> f(size_t(n) * kFeatures);
        APO_HIP(hipMemcpy(f.data(), ctx->feat, f.size() * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; i++) std::memcpy(features + size_t(idx[i]) * kFeatures, &f[size_t(i) * kFeatures], kFeatures * 4);
    }
    return APO_OK;
}

apo_status apo_proposal_train(apo_ctx* ctx, const float* features, const uint8_t* labels,
                              const float* radius_targets, uint64_t count, uint32_t epochs, uint32_t batch, float lr,
                              uint32_t seed, const float* init, float* params, float* epoch_loss) {
    if (!ctx || !features || !labels || !params || count == 0 || batch == 0) return APO_ERR_INVALID;
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // buried sites carry no label; the rest become classes
    std::vector<uint32_t> keep;
    keep.reserve(count);
    for (uint64_t q = 0; q < count; q++)
        if (labels[q] != kLabelBuried) keep.push_back(uint32_t(q));
    if (keep.empty()) return fail(ctx, APO_ERR_INVALID, "no labelled site");
    const uint32_t m = uint32_t(keep.size());
    std::vector<float> f(size_t(m) * kFeatures);
    std::vector<uint8_t> y(m);
    std::vector<float> rt(radius_targets ? m : 0u);
    for (uint32_t q = 0; q < m; q++) {
        std::memcpy(&f[size_t(q) * kFeatures], features + size_t(keep[q]) * kFeatures, kFeatures * 4);
        y[q] = uint8_t(proposalClass(labels[keep[q]]));
        if (radius_targets) rt[q] = radius_targets[keep[q]];
    }
    // parameters: given, or uniform(+-sqrt(6 / fan_in)) from the seed
    std::vector<float> w(kProposalParams, 0.f);
    if (init) {
        std::memcpy(w.data(), init, kProposalParams * 4);
    } else {
        std::mt19937 gen(seed);
        auto fill = [&](size_t off, uint32_t fanIn, uint32_t cnt) {
            std::uniform_real_distribution<float> u(-std::sqrt(6.f / float(fanIn)), std::sqrt(6.f / float(fanIn)));
            for (uint32_t q = 0; q < cnt; q++) w[off + q] = u(gen);
        };
        size_t off = 0;
        fill(off, kFeatures, kFeatures * kHidden);
        off += kFeatures * kHidden + kHidden;
        fill(off, kHidden, kHidden * kHidden);
        off += kHidden * kHidden + kHidden;
        fill(off, kHidden, kHidden * kProposalOutputs);
    }
    std::vector<float> wq(kProposalParams);
    for (uint32_t q = 0; q < kProposalParams; q++) {
        uint32_t b;
        std::memcpy(&b, &w[q], 4);
        b = ((b + 0x7fffu + ((b >> 16) & 1u)) >> 16) << 16;
        std::memcpy(&wq[q], &b, 4);
    }
    const uint32_t maxWaves = trainWaves(batch < m ? batch : m);
    DeviceTemp<float> dF, dW, dWq, dM, dV, dPart, dLossPart, dLoss, dB1, dB2, dB3, dR;
    DeviceTemp<uint16_t> dW1, dW2, dW3;
    DeviceTemp<uint8_t> dY;
    DeviceTemp<uint32_t> dOrder;
    APO_HIP(dF.alloc(f.size()));
    if (radius_targets) {
        APO_HIP(dR.alloc(m));
        APO_HIP(hipMemcpyAsync(dR.p, rt.data(), size_t(m) * 4, hipMemcpyHostToDevice, st));
    }
    APO_HIP(dY.alloc(m));
    APO_HIP(dOrder.alloc(m));
    APO_HIP(dW.alloc(kProposalParams));
    APO_HIP(dWq.alloc(kProposalParams));
    APO_HIP(dM.alloc(kProposalParams));
    APO_HIP(dV.alloc(kProposalParams));
    APO_HIP(dPart.alloc(size_t(maxWaves) * kProposalParams));
    APO_HIP(dLossPart.alloc(maxWaves));
    APO_HIP(dLoss.alloc(1));
    APO_HIP(dW1.alloc(size_t(kHidden) * kHidden));
    APO_HIP(dW2.alloc(size_t(kHidden) * kHidden));
    APO_HIP(dW3.alloc(size_t(16) * kHidden));
    APO_HIP(dB1.alloc(kHidden));
    APO_HIP(dB2.alloc(kHidden));
    APO_HIP(dB3.alloc(16));
    APO_HIP(hipMemcpyAsync(dF.p, f.data(), f.size() * 4, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(dY.p, y.data(), m, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(dW.p, w.data(), kProposalParams * 4, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(dWq.p, wq.data(), kProposalParams * 4, hipMemcpyHostToDevice, st));
    std::vector<float> zeros(kProposalParams, 0.f);
    APO_HIP(hipMemcpyAsync(dM.p, zeros.data(), kProposalParams * 4, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(dV.p, zeros.data(), kProposalParams * 4, hipMemcpyHostToDevice, st));
    const TrainBuffers tb{dF.p, dY.p, radius_targets ? dR.p : nullptr, dW.p, dWq.p, dM.p, dV.p, dPart.p,
                          dLossPart.p, dLoss.p, maxWaves, ProposalWeights{dW1.p, dW2.p, dW3.p, dB1.p, dB2.p, dB3.p}};
    APO_HIP(packWeightsLaunch(dW.p, tb.pw, st));
    std::vector<uint32_t> order(m);
    for (uint32_t q = 0; q < m; q++) order[q] = q;
    std::mt19937 shuffler(seed ^ 0x9e3779b9u);
    uint32_t step = 0;
    for (uint32_t ep = 0; ep < epochs; ep++) {
        std::shuffle(order.begin(), order.end(), shuffler);
        APO_HIP(hipMemcpyAsync(dOrder.p, order.data(), size_t(m) * 4, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(dLoss.p), 1u, 0u);
        for (uint32_t b0 = 0; b0 < m; b0 += batch) {
            const uint32_t rows = (m - b0) < batch ? (m - b0) : batch;
            APO_HIP(mlpTrainStep(tb, dOrder.p + b0, rows, lr, step++, st));
        }
        float loss = 0.f;
        APO_HIP(hipMemcpyAsync(&loss, dLoss.p, 4, hipMemcpyDeviceToHost, st));
        APO_HIP(hipStreamSynchronize(st));
        if (epoch_loss) epoch_loss[ep] = loss / float(m);
    }
    APO_HIP(hipMemcpyAsync(params, dW.p, kProposalParams * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    return APO_OK;
}

apo_status apo_proposal_gradient(apo_ctx* ctx, const float* params, const float* features, const uint8_t* classes,
                                 const float* radius_targets, uint32_t count, float* grad, float* loss) {
    if (!ctx || !params || !features || !classes || !grad || count == 0) return APO_ERR_INVALID;
    for (uint32_t q = 0; q < count; q++)
        if (classes[q] >= kProposalClasses) return fail(ctx, APO_ERR_INVALID, "class out of range");
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    std::vector<float> wq(kProposalParams);
    for (uint32_t q = 0; q < kProposalParams; q++) {
        uint32_t b;
        std::memcpy(&b, &params[q], 4);
        b = ((b + 0x7fffu + ((b >> 16) & 1u)) >> 16) << 16;
        std::memcpy(&wq[q], &b, 4);
    }
    std::vector<uint32_t> order(count);
    for (uint32_t q = 0; q < count; q++) order[q] = q;
    const uint32_t waves = trainWaves(count);
    DeviceTemp<float> dF, dW, dWq, dPart, dLossPart, dB1, dB2, dB3, dR;
    DeviceTemp<uint16_t> dW1, dW2, dW3;
    DeviceTemp<uint8_t> dY;
    DeviceTemp<uint32_t> dOrder;
    APO_HIP(dF.alloc(size_t(count) * kFeatures));
    if (radius_targets) {
        APO_HIP(dR.alloc(count));
        APO_HIP(hipMemcpyAsync(dR.p, radius_targets, size_t(count) * 4, hipMemcpyHostToDevice, st));
    }
    APO_HIP(dY.alloc(count));
    APO_HIP(dOrder.alloc(count));
    APO_HIP(dW.alloc(kProposalParams));
    APO_HIP(dWq.alloc(kProposalParams));
    APO_HIP(dPart.alloc(size_t(waves) * kProposalParams));
    APO_HIP(dLossPart.alloc(waves));
    APO_HIP(dW1.alloc(size_t(kHidden) * kHidden));
    APO_HIP(dW2.alloc(size_t(kHidden) * kHidden));
    APO_HIP(dW3.alloc(size_t(16) * kHidden));
    APO_HIP(dB1.alloc(kHidden));
    APO_HIP(dB2.alloc(kHidden));
    APO_HIP(dB3.alloc(16));
    APO_HIP(hipMemcpyAsync(dF.p, features, size_t(count) * kFeatures * 4, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(dY.p, classes, count, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(dOrder.p, order.data(), size_t(count) * 4, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(dW.p, params, kProposalParams * 4, hipMemcpyHostToDevice, st));
    APO_HIP(hipMemcpyAsync(dWq.p, wq.data(), kProposalParams * 4, hipMemcpyHostToDevice, st));
    const TrainBuffers tb{dF.p, dY.p, radius_targets ? dR.p : nullptr, dW.p, dWq.p, nullptr, nullptr, dPart.p,
                          dLossPart.p, nullptr, waves, ProposalWeights{dW1.p, dW2.p, dW3.p, dB1.p, dB2.p, dB3.p}};
    APO_HIP(packWeightsLaunch(dW.p, tb.pw, st));
    APO_HIP(trainGradLaunch(tb, dOrder.p, count, 1.f / float(count), st));
    std::vector<float> part(size_t(waves) * kProposalParams), lp(waves);
    APO_HIP(hipMemcpyAsync(part.data(), dPart.p, part.size() * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipMemcpyAsync(lp.data(), dLossPart.p, waves * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    // the partials summed in the order the Adam kernel sums them
    for (uint32_t q = 0; q < kProposalParams; q++) {
        float g = 0.f;
        for (uint32_t wv = 0; wv < waves; wv++) g += part[size_t(wv) * kProposalParams + q];
        grad[q] = g;
    }
    if (loss) {
        float l = 0.f;
        for (uint32_t wv = 0; wv < waves; wv++) l += lp[wv];
        *loss = l / float(count);
    }
    return APO_OK;
}

apo_status apo_proposal_set(apo_ctx* ctx, const float* params) {
    if (!ctx) return APO_ERR_INVALID;
    APO_HIP(hipSetDevice(ctx->device));
    if (!params) {
        ctx->proposalOn = false;
        return APO_OK;
    }
    if (!ctx->params) {
        APO_HIP(alloc(ctx->params, kProposalParams));
        APO_HIP(alloc(ctx->pw.w1, size_t(kHidden) * kHidden));
        APO_HIP(alloc(ctx->pw.w2, size_t(kHidden) * kHidden));
        APO_HIP(alloc(ctx->pw.w3, size_t(16) * kHidden));
        APO_HIP(alloc(ctx->pw.b1, kHidden));
        APO_HIP(alloc(ctx->pw.b2, kHidden));
        APO_HIP(alloc(ctx->pw.b3, 16));
    }
    APO_HIP(hipMemcpyAsync(ctx->params, params, kProposalParams * 4, hipMemcpyHostToDevice, ctx->stream));
    APO_HIP(packWeightsLaunch(ctx->params, ctx->pw, ctx->stream));
    APO_HIP(hipStreamSynchronize(ctx->stream));
    ctx->proposalOn = true;
    return APO_OK;
}

apo_status apo_proposal_infer(apo_ctx* ctx, const float* features, uint32_t count, uint8_t* classes, float* logits) {
    if (!ctx || !features || !classes) return APO_ERR_INVALID;
    if (!ctx->proposalOn) return fail(ctx, APO_ERR_STATE, "no proposal weights set");
    APO_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DeviceTemp<float> dF, dL;
    DeviceTemp<uint8_t> dC;
    APO_HIP(dF.alloc(size_t(count) * kFeatures));
    APO_HIP(dC.alloc(count));
    if (logits) APO_HIP(dL.alloc(size_t(count) * kProposalOutputs));
    APO_HIP(hipMemcpyAsync(dF.p, features, size_t(count) * kFeatures * 4, hipMemcpyHostToDevice, st));
    APO_HIP(mlpInferLaunch(dF.p, count, ctx->pw, dC.p, logits ? dL.p : nullptr, ctx->cfg.proposal_cost,
                           ctx->cfg.proposal_fallback_k, ctx->cfg.proposal_max_k, st));
    APO_HIP(hipMemcpyAsync(classes, dC.p, count, hipMemcpyDeviceToHost, st));
    if (logits) APO_HIP(hipMemcpyAsync(logits, dL.p, size_t(count) * kProposalOutputs * 4, hipMemcpyDeviceToHost, st));
    APO_HIP(hipStreamSynchronize(st));
    return APO_OK;
}

apo_status apo_cell_radii(apo_ctx* ctx, float* radii) {
    if (!ctx || !radii) return APO_ERR_INVALID;
    if (!ctx->built || (ctx->cfg.stage_mask & 8u)) return fail(ctx, APO_ERR_STATE, "no complete build");
    APO_HIP(hipSetDevice(ctx->device));
    const uint32_t n = ctx->n;
    DeviceTemp<float> d;
    APO_HIP(d.alloc(n));
    APO_HIP(cellRadiiLaunch(cellCtx(ctx), d.p, cellPassBlocks(ctx), ctx->stream));
    std::vector<float> sorted(n);
    std::vector<uint32_t> idx(n);
    APO_HIP(hipMemcpyAsync(sorted.data(), d.p, size_t(n) * 4, hipMemcpyDeviceToHost, ctx->stream));
    APO_HIP(hipMemcpyAsync(idx.data(), ctx->idxA, size_t(n) * 4, hipMemcpyDeviceToHost, ctx->stream));
    APO_HIP(hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < n; i++) radii[idx[i]] = sorted[i];
    return APO_OK;
}

apo_status apo_proposal_set_radii(apo_ctx* ctx, const float* radii, uint32_t n) {
    if (!ctx) return APO_ERR_INVALID;
    APO_HIP(hipSetDevice(ctx->device));
    dropGraph(ctx);
    if (!radii) {
        ctx->givenRN = 0;
        return APO_OK;
    }
    if (n < 2) return fail(ctx, APO_ERR_INVALID, "need radii for at least 2 sites");
    for (uint32_t q = 0; q < n; q++)
        if (!(radii[q] == radii[q])) return fail(ctx, APO_ERR_INVALID, "a search radius is NaN");
    if (n > ctx->givenRCap) {
        if (ctx->givenRadii) APO_HIP(hipFree(ctx->givenRadii));
        ctx->givenRadii = nullptr;
        ctx->givenRCap = 0;
        APO_HIP(alloc(ctx->givenRadii, n));
        ctx->givenRCap = n;
    }
    APO_HIP(hipMemcpyAsync(ctx->givenRadii, radii, size_t(n) * 4, hipMemcpyHostToDevice, ctx->stream));
    APO_HIP(hipStreamSynchronize(ctx->stream));
    ctx->givenRN = n;
    return APO_OK;
}

apo_status apo_proposal_set_classes(apo_ctx* ctx, const uint8_t* classes, uint32_t n) {
    if (!ctx) return APO_ERR_INVALID;
    APO_HIP(hipSetDevice(ctx->device));
    if (!classes) {
        if (ctx->givenN) dropGraph(ctx);
        ctx->givenN = 0;
        return APO_OK;
    }
    if (n < 2 || n >= (1u << 28)) return fail(ctx, APO_ERR_INVALID, "site count out of range");
    for (uint32_t i = 0; i < n; i++)
        if (classes[i] >= kProposalClasses) return fail(ctx, APO_ERR_INVALID, "proposal class above 7");
    if (n > ctx->givenCap) {
        dropGraph(ctx);
        if (ctx->givenCls) APO_HIP(hipFree(ctx->givenCls));
        ctx->givenCls = nullptr;
        ctx->givenCap = 0;
        APO_HIP(alloc(ctx->givenCls, n));
        ctx->givenCap = n;
    }
    APO_HIP(hipMemcpyAsync(ctx->givenCls, classes, n, hipMemcpyHostToDevice, ctx->stream));
    APO_HIP(hipStreamSynchronize(ctx->stream));
    ctx->givenN = n;
    return APO_OK;
}

apo_status apo_proposal_radii(apo_ctx* ctx, float* radii) {
    if (!ctx || !radii) return APO_ERR_INVALID;
    if (!ctx->built || !ctx->stats.proposal || !(ctx->givenRN || (ctx->cfg.proposal_radius && !ctx->givenN)))
        return fail(ctx, APO_ERR_STATE, "no build with search radii");
    const size_t n = ctx->n;
    std::vector<uint32_t> idx(n);
    std::vector<float> r(n);
    APO_HIP(hipMemcpy(idx.data(), ctx->idxA, n * 4, hipMemcpyDeviceToHost));
    APO_HIP(hipMemcpy(r.data(), ctx->radiiSorted, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) radii[idx[i]] = r[i];
    return APO_OK;
}

apo_status apo_proposal_classes(apo_ctx* ctx, uint8_t* classes) {
    if (!ctx || !classes) return APO_ERR_INVALID;
    if (!ctx->built || !ctx->stats.proposal) return fail(ctx, APO_ERR_STATE, "no build with the proposal");
    const size_t n = ctx->n;
    std::vector<uint32_t> idx(n);
    std::vector<uint8_t> c(n);
    APO_HIP(hipMemcpy(idx.data(), ctx->idxA, n * 4, hipMemcpyDeviceToHost));
    APO_HIP(hipMemcpy(c.data(), ctx->cls, n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) classes[idx[i]] = c[i];
    return APO_OK;
}

}  // extern "C"
