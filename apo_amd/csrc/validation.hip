// validation.hip -- completion of uncertified cells on gfx950.
//
// Reference: apo/include/apo/gpu/algorithm.inl:208-476 (globalVertexValidation),
// 478-857 (globalEdgeValidation), 859-1233 (globalBisectorValidation) and the
// iteration loop of AlgorithmGPU::build, :1551-1593.
//
// One wave per task cell: a persistent grid of waves takes the round's
// task cells (compacted on the device) from a work queue, so a round needs
// no task count on the host and the whole build can be captured in a graph.
// Each lane runs its own BVH query (vertex / edge / bisector) with its stack
// in LDS ([slot][lane]); the contributors it finds are applied to the cell
// in ascending lane order with the wave-wide updateCell of cells.hip.  Steps
// are 64 records wide.
//
// Edge and bisector searches pause at each candidate leaf: every lane walks
// its own traversal (same order as the reference) to its next candidate,
// then the wave tests all pending candidates together (vertex solve and the
// empty-sphere test against the cell's neighbours, whose sites sit in LDS).  The expensive test thus runs once per round instead of once per
// lane that happens to reach a candidate.
#include "apo_device.h"
#include "apo_internal.h"
#include "cell_update.h"
#include "edge_bound.h"

namespace apo_hip {

// Traversal stack of the vertex and edge searches: the newest
// kVertexLdsDepth entries of every lane in an LDS ring, older ones evicted
// to a per-wave region of device memory (the reference's kStackDepth in
// all), so the kernels' LDS admits more waves per CU while pushes and pops
// at the top of a depth-first search stay in LDS.  A lane's slots [0, lo)
// live in the device region, slots [lo, sp) in the ring at slot %
// kVertexLdsDepth.  The ring occupies the cell's update scratch list, which
// holds nothing while the searches run (the cell is updated after them).
// The edge kernel keeps its cover buffers in the same device region, after
// the stack.
constexpr uint32_t kVertexLdsDepth = 8;
static_assert(kVertexLdsDepth * 64u <= kListCap, "the stack ring fits the cell's scratch list");
constexpr uint32_t kVertexSpillWords = (kStackDepth - kVertexLdsDepth) * 64;
constexpr uint32_t kWaveScratchWords = kVertexSpillWords + uint32_t(sizeof(CoverLds) / sizeof(uint32_t));
static_assert(kVertexSpillWords % 2u == 0u, "the cover buffers of a wave region are 8-byte aligned");
static_assert((kVertexLdsDepth & (kVertexLdsDepth - 1u)) == 0u, "the ring index is a mask");
struct VStack {
    uint32_t (*lds)[64];
    uint32_t* spill;  // kVertexSpillWords words of this wave
    // slot of a lane whose device-region part is [0, lo)
    __device__ __forceinline__ uint32_t get(uint32_t slot, uint32_t lane, uint32_t lo) const {
        return slot < lo ? spill[slot * 64u + lane] : lds[slot & (kVertexLdsDepth - 1u)][lane];
    }
};

// LDS of each kernel (one wave per block)
__device__ __forceinline__ uint32_t (*stackRing(CellLds& L))[64] {
    return reinterpret_cast<uint32_t (*)[64]>(L.list);
}

struct VertexLds {
    CellLds cell;
    float resD[64];       // per owner lane: smallest gap found by shared items
    uint32_t resN[64];    // ... and its site (earliest in traversal order on ties)
};
struct EdgeLds {
    CellLds cell;
    CoverOut coverOut;
    uint32_t res[64];     // per owner lane: earliest contributor found by a helper
};
struct BisectorLds {
    CellLds cell;
    uint32_t bis[kBisectorCap];
    uint32_t res[64];     // per owner lane: earliest contributor found by a helper
};

// Traversal state of one lane: next node and stack height.  With work
// sharing (edge searches) the stack occupies slots [sb, sp): the slots
// below sb were given to other lanes, and `base` is the depth the item's
// stack bottom has in the reference's single-stack traversal (its overflow
// test uses that depth, so a shared search overflows where the reference's
// would).
struct Trav {
    uint32_t node, sp;
    bool live;
    uint32_t sb = 0, base = 0;
    uint32_t lo = 0;  // slots below lo are in the device region (VStack)
};

// Lane of the k-th (0-based) set bit of a wave mask: the donor a receiving
// lane of rank k takes work from (binary search on prefix popcounts).
__device__ __forceinline__ uint32_t laneOfRank(uint64_t mask, uint32_t k) {
    uint32_t lo = 0, hi = 63;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (popc64(mask & ((uint64_t(2) << mid) - 1u)) > k) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}

// Take the top entry of a non-empty VStack (sp > sb).
__device__ __forceinline__ uint32_t topEntry(const VStack& st, Trav& t) {
    t.sp--;
    const uint32_t v = st.get(t.sp, laneId(), t.lo);
    if (t.sp < t.lo) t.lo = t.sp;
    return v;
}
// Push onto a VStack: when the ring is full its oldest entry moves to the
// device region first.
__device__ __forceinline__ void pushEntry(const VStack& st, Trav& t, uint32_t v) {
    const uint32_t lane = laneId();
    if (t.sp >= t.lo + kVertexLdsDepth) {
        st.spill[t.lo * 64u + lane] = st.lds[t.lo & (kVertexLdsDepth - 1u)][lane];
        t.lo++;
    }
    st.lds[t.sp & (kVertexLdsDepth - 1u)][lane] = v;
    t.sp++;
}
__device__ __forceinline__ void popNode(const VStack& st, Trav& t) {
    if (t.sp == t.sb) {
        t.live = false;
    } else {
        t.node = topEntry(st, t);
    }
}

// same test from one lane, over the neighbour set
__device__ __forceinline__ bool knownLane(const CellLds& L, uint32_t F, uint32_t j) {
    uint32_t q = 0;
    for (; q + 8u <= F; q += 8u) {
        uint32_t o[8];
#pragma unroll
        for (uint32_t u = 0; u < 8u; u++) o[u] = L.nb[q + u];
        bool hit = false;
#pragma unroll
        for (uint32_t u = 0; u < 8u; u++) hit |= o[u] == j;
        if (hit) return true;
    }
    for (; q < F; q++)
        if (L.nb[q] == j) return true;
    return false;
}

// ---------------------------------------------------------------------------
// Nearest contributing site to a vertex sphere (algorithm.inl:305-414)
// ---------------------------------------------------------------------------
// One step of a lane's nearest-site search (algorithm.inl:305-414): visit
// one internal node (descend to the closer child under the best gap so far,
// stack the other with its gap) or test one leaf; stacked nodes whose gap
// exceeds the best are skipped when popped.  The stack holds the parent of
// the stacked child and which child it is (bit 31); the gap is recomputed
// from the parent's copy of the child's box when the entry is popped, the
// same FP32 value the reference stacks, so no gap needs LDS.
__device__ __forceinline__ void vertexStep(const CellCtx& c, const CellLds& L, const VStack& stNode,
                                           uint32_t F, uint32_t i, uint4 vx, F4 cur, Trav& t,
                                           float& nd, uint32_t& nb, uint32_t& flags) {
    const uint32_t lane = laneId();
    const uint32_t n = c.n;
    const uint32_t node = t.node;
    if (node < c.internal) {
        const Node nn = c.nodes[node];
        Box lb, rb;
        nodeBoxes(nn, lb, rb);
        const uint32_t L0 = uint32_t(nn.d.x), R0 = uint32_t(nn.d.y);
        bool ef = false, ec = false;
        uint32_t farChild = 0, closeChild = 0;
        float minFar = 0.f;
        if (L0 >= nodeCount(c) || R0 >= nodeCount(c)) {
            atomicOr(c.err, kErrKnnChild);
        } else {
            const float dL = boxDistance(lb, cur), dR = boxDistance(rb, cur);
            const bool lc = dL < dR;
            minFar = lc ? dR : dL;
            farChild = lc ? R0 : L0;
            const float minClose = lc ? dL : dR;
            closeChild = lc ? L0 : R0;
            ef = minFar < nd;
            ec = minClose < nd;
        }
        if (ec || ef) {
            t.node = ec ? closeChild : farChild;
            if (ec && ef) {
                if (t.base + t.sp < kStackDepth) {
                    pushEntry(stNode, t, node | (farChild == R0 ? 0x80000000u : 0u));
                } else {
                    flags |= kFlagStackOverflow;
                }
            }
            return;
        }
    } else if (node < nodeCount(c)) {
        const uint32_t j = node - c.internal;
        if (j != i && j != vx.x && j != vx.y && j != vx.z && c.status[j] != kStatusBuried) {
            const float d = sphereDistF(c.sorted[j], cur);
            if (d < nd && !knownLane(L, F, j)) {
                nd = d;
                nb = j;
            }
        }
    }
    bool more = false;
    while (t.sp > t.sb) {
        const uint32_t e = topEntry(stNode, t);
        const Node pn = c.nodes[e & 0x7fffffffu];
        Box lb, rb;
        nodeBoxes(pn, lb, rb);
        const bool right = (e >> 31) != 0u;
        if (boxDistance(selBox(right, rb, lb), cur) > nd) continue;
        t.node = uint32_t(right ? pn.d.y : pn.d.x);
        more = true;
        break;
    }
    if (!more) t.live = false;
}

// Is leaf a visited before leaf b by a closest-child-first traversal for
// the search sphere bs?  Both leaves hang below their lowest common
// ancestor, whose closer child (by box distance to bs, the traversal's own
// order test) is visited first.
__device__ __forceinline__ bool visitedBefore(const CellCtx& c, uint32_t a, uint32_t b, F4 bs) {
    constexpr uint32_t kMaxDepth = 4096;  // bounds every climb, whatever the tree holds
    uint32_t na = c.internal + a, nb = c.internal + b;
    uint32_t da = 0, db = 0;
    for (int32_t p = c.parent[na]; p >= 0 && da < kMaxDepth; p = c.parent[p]) da++;
    for (int32_t p = c.parent[nb]; p >= 0 && db < kMaxDepth; p = c.parent[p]) db++;
    if (da >= kMaxDepth || db >= kMaxDepth) {
        atomicOr(c.err, kErrValidation);
        return a < b;
    }
    while (da > db) {
        na = uint32_t(c.parent[na]);
        da--;
    }
    while (db > da) {
        nb = uint32_t(c.parent[nb]);
        db--;
    }
    for (uint32_t guard = 0; guard <= kMaxDepth; guard++) {
        const int32_t pa = c.parent[na], pb = c.parent[nb];
        if (pa == pb || pa < 0 || pb < 0) {
            if (pa != pb || pa < 0) return a < b;  // different trees cannot happen for one query
            const Node nn = c.nodes[pa];
            Box lb, rb;
            nodeBoxes(nn, lb, rb);
            const bool leftFirst = boxDistance(lb, bs) < boxDistance(rb, bs);
            const bool aLeft = uint32_t(nn.d.x) == na;
            return aLeft == leftFirst;
        }
        na = uint32_t(pa);
        nb = uint32_t(pb);
    }
    atomicOr(c.err, kErrValidation);
    return a < b;
}

__device__ __forceinline__ F4 shflF4(F4 v, uint32_t src) {
    return {__shfl(v.x, int(src), 64), __shfl(v.y, int(src), 64), __shfl(v.z, int(src), 64), __shfl(v.w, int(src), 64)};
}
__device__ __forceinline__ F3 shflF3(F3 v, uint32_t src) {
    return {__shfl(v.x, int(src), 64), __shfl(v.y, int(src), 64), __shfl(v.z, int(src), 64)};
}
__device__ __forceinline__ D4 shflD4(D4 v, uint32_t src) {
    return {__shfl(v.x, int(src), 64), __shfl(v.y, int(src), 64), __shfl(v.z, int(src), 64), __shfl(v.w, int(src), 64)};
}

// Lanes with no search left take one when at least this many are idle.
constexpr uint32_t kShareIdle = 4;

// One step of a lane's search for the nearest site within its bound `nd`
// among the sites the reference's nearest-site search can return (not i,
// not the vertex's sites, not buried, not a cell neighbour).  Children are
// visited in the reference's order (vertexStep) but pruned soundly: a
// site's gap is at least its box's gap minus `slack` (FP32 rounding of boxes
// and gaps) when the vertex centre lies outside the box, and unbounded
// below when it lies inside (a centre inside a site gives a negative gap),
// so a box within `slack` of the centre is never pruned.  Leaves improve
// the bound strictly, so the first of equal gaps in this order is kept.  A
// leaf reached by a step's descent or pop is tested in the same step.
__device__ __forceinline__ void vertexNearStep(const CellCtx& c, const CellLds& L, const VStack& stNode,
                                               uint32_t F, uint32_t i, uint4 vx, F4 cur, float slack, Trav& t,
                                               float& nd, uint32_t& nb, uint32_t& flags) {
    const uint32_t lane = laneId();
    const uint32_t node = t.node;
    if (node < c.internal) {
        const Node nn = c.nodes[node];
        Box lb, rb;
        nodeBoxes(nn, lb, rb);
        const uint32_t L0 = uint32_t(nn.d.x), R0 = uint32_t(nn.d.y);
        bool ef = false, ec = false;
        uint32_t farChild = 0, closeChild = 0;
        if (L0 >= nodeCount(c) || R0 >= nodeCount(c)) {
            atomicOr(c.err, kErrKnnChild);
        } else {
            const bool lc = boxDistance(lb, cur) < boxDistance(rb, cur);
            const F3 p = xyzf(cur);
            const float eL = lenf(subf(p, closestPoint(lb, cur))), eR = lenf(subf(p, closestPoint(rb, cur)));
            const bool inL = eL <= slack || eL - cur.w - slack <= nd;
            const bool inR = eR <= slack || eR - cur.w - slack <= nd;
            farChild = lc ? R0 : L0;
            closeChild = lc ? L0 : R0;
            ef = lc ? inR : inL;
            ec = lc ? inL : inR;
        }
        if (ec || ef) {
            t.node = ec ? closeChild : farChild;
            if (ec && ef) {
                if (t.base + t.sp < kStackDepth) {
                    pushEntry(stNode, t, farChild);
                } else {
                    flags |= kFlagStackOverflow;
                }
            }
        } else {
            popNode(stNode, t);
        }
        if (!t.live || t.node < c.internal) return;
    }
    const uint32_t leaf = t.node;
    if (leaf < nodeCount(c)) {
        const uint32_t j = leaf - c.internal;
        if (j != i && j != vx.x && j != vx.y && j != vx.z && c.status[j] != kStatusBuried) {
            const float d = sphereDistF(c.sorted[j], cur);
            if (d < nd && !knownLane(L, F, j)) {
                nd = d;
                nb = j;
            }
        }
    }
    popNode(stNode, t);
}

// Does the reference's search (vertexStep) reach leaf j and keep it?  It
// descends into a node, or pops it from its stack, only when the node's box
// gap (as vertexStep computes it) is below its running best, which never
// drops below the smallest gap d over all eligible sites.  So when every
// node on the path from the root to leaf j has a box gap strictly below d
// -- the gap of j, the smallest, and the first of equal gaps in the
// traversal order -- the search reaches j with a best above d, keeps it,
// and nothing later improves on it.
__device__ __forceinline__ bool searchKeeps(const CellCtx& c, uint32_t j, float d, F4 cur) {
    uint32_t node = c.internal + j;
    for (uint32_t guard = 0; guard < nodeCount(c); guard++) {
        const int32_t p = c.parent[node];
        if (p < 0) return true;
        const Node nn = c.nodes[p];
        Box lb, rb;
        nodeBoxes(nn, lb, rb);
        const Box b = selBox(uint32_t(nn.d.x) == node, lb, rb);
        if (!(boxDistance(b, cur) < d)) return false;
        node = uint32_t(p);
    }
    return false;
}

// The nearest-site searches of all lanes' vertices (algorithm.inl:305-414).
//
// The caller uses one thing of a search: whether its site intersects the
// vertex sphere, and then which site.  Each vertex first looks, with the
// work sharing of the edge searches (idle lanes take the last subtree of
// busy lanes' stacks), for the nearest eligible site within a rounding
// margin of its sphere:
//  * none: the site the reference returns has an FP32 gap above the margin,
//    so its FP64 gap exceeds the intersection tolerance: no intersection,
//    whatever that site is;
//  * site j: if the reference's search provably reaches and keeps j
//    (searchKeeps), j is its answer; otherwise the vertex reruns the
//    reference's own traversal.
// Both shortcuts also need the reference's search not to overflow its stack
// (the overflow is recorded in the cell's flags): they are taken only when
// no leaf is deeper than the stack, which makes an overflow impossible.
// Returns the site the reference's search returns when it may intersect,
// else kNone.
__device__ __forceinline__ uint32_t vertexSearches(const CellCtx& c, VertexLds& S, const CellLds& L,
                                                   const WaveCell& w, uint32_t i, uint32_t root, bool need,
                                                   const uint4 vx0, const F4 cur0, uint32_t* spill, uint32_t& flags) {
    const uint32_t lane = laneId();
    const VStack st{stackRing(S.cell), spill};
    const float reach = 1e-5f * (__builtin_fabsf(cur0.x) + __builtin_fabsf(cur0.y) + __builtin_fabsf(cur0.z) +
                                 __builtin_fabsf(cur0.w)) + 1e-6f;
    bool exact = need;
    uint32_t site = kNone;
    if (__builtin_amdgcn_readfirstlane(*c.treeHeight) <= kStackDepth) {
        S.resN[lane] = kNone;
        S.resD[lane] = FLT_MAX;
        uint4 vx = vx0;
        F4 cur = cur0;
        float slack = reach;
        uint32_t owner = need ? lane : kNone, rflags = 0;
        bool own = need;
        float nd = reach, ownNd = reach;
        uint32_t nb = kNone, ownNb = kNone;
        Trav tr{root, 0u, need};
        waveSync();
        while (true) {
            const uint64_t live = ballot(tr.live);
            if (live == 0) break;
            const uint64_t idle = ~live;
            const uint64_t donors = ballot(tr.live && tr.sp > tr.sb);
            if (popc64(idle) >= kShareIdle && donors != 0) {
                const uint32_t m = popc64(idle) < popc64(donors) ? popc64(idle) : popc64(donors);
                const uint32_t dr = countBelow(donors), ir = countBelow(idle);
                const bool give = ((donors >> lane) & 1u) && dr < m;
                const bool take = ((idle >> lane) & 1u) && ir < m;
                waveSync();  // the donors' stack slots are visible to the takers
                const uint32_t src = take ? laneOfRank(donors, ir) : lane;
                const uint32_t dSb = __shfl(tr.sb, int(src), 64), dBase = __shfl(tr.base, int(src), 64);
                const uint32_t dLo = __shfl(tr.lo, int(src), 64);
                const uint32_t dOwner = __shfl(owner, int(src), 64);
                const float dNd = __shfl(nd, int(src), 64);
                const uint32_t node = take ? st.get(dSb, src, dLo) : 0u;
                cur = shflF4(cur, src);
                slack = __shfl(slack, int(src), 64);
                vx.x = __shfl(vx.x, int(src), 64);
                vx.y = __shfl(vx.y, int(src), 64);
                vx.z = __shfl(vx.z, int(src), 64);
                if (take) {
                    if (own) {
                        ownNd = nd;
                        ownNb = nb;
                    }
                    owner = dOwner;
                    own = false;
                    nd = dNd;
                    nb = kNone;
                    tr.node = node;
                    tr.sp = 0u;
                    tr.sb = 0u;
                    tr.lo = 0u;
                    tr.base = dBase + dSb;
                    tr.live = true;
                }
                if (give) tr.sb++;
                waveSync();
            }
            const bool wasLive = tr.live;
            if (tr.live) vertexNearStep(c, L, st, w.F, i, vx, cur, slack, tr, nd, nb, rflags);
            // a shared item that ended hands its nearest site to its owner
            const bool report = wasLive && !tr.live && !own && nb != kNone;
            uint64_t hm = ballot(report);
            while (hm) {
                const uint32_t t = uint32_t(__builtin_ctzll(hm));
                hm &= hm - 1;
                if (lane == t) {
                    const float bd = S.resD[owner];
                    const uint32_t bn = S.resN[owner];
                    if (bn == kNone || nd < bd || (nd == bd && visitedBefore(c, nb, bn, cur))) {
                        S.resD[owner] = nd;
                        S.resN[owner] = nb;
                    }
                }
                waveSync();
            }
        }
        waveSync();
        if (own) {
            ownNd = nd;
            ownNb = nb;
        }
        float best = ownNd;
        site = ownNb;
        // the lane's own item holds the earliest part of its search: it wins ties
        if (need && S.resN[lane] != kNone && (site == kNone || S.resD[lane] < best)) {
            best = S.resD[lane];
            site = S.resN[lane];
        }
        // an overflow cannot happen below the height bound; were it to, rerun
        exact = need && (rflags != 0u || (site != kNone && !searchKeeps(c, site, best, cur0)));
        if (!need || exact) site = kNone;
    }
    // the reference's traversal for the vertices that need it
    uint32_t nb2 = kNone;
    if (ballot(exact) != 0) {
        Trav t2{root, 0u, exact};
        float nd2 = FLT_MAX;
        while (ballot(t2.live) != 0) {
            if (t2.live) vertexStep(c, L, st, w.F, i, vx0, cur0, t2, nd2, nb2, flags);
        }
    }
    return exact ? nb2 : site;
}

__device__ __forceinline__ void vertexTask(const CellCtx& c, VertexLds& S, uint32_t i, uint32_t maxRestarts,
                                           uint32_t* spill) {
    CellLds& L = S.cell;
    const uint32_t lane = laneId();
    const uint8_t st = c.status[i];
    if (st == kStatusCertified || st == kStatusBuried) return;
    WaveCell w;
    loadCell(c, L, w, i, true);
    const uint32_t loaded = w.V + w.E;
    const float kr = c.knownRadii[i];
    const D4 si = loadSite(c, i);
    const uint32_t root = rootNode(c, cellMolecule(c, i));
    uint32_t vs = 0, restarts = 0;
    uint32_t flags = 0;
    while (vs < w.V / 64 + 1 && restarts < maxRestarts) {
        const uint32_t v = vs * 64 + lane;
        vs++;
        buildNeighbours(c, L, w);
        bool need = v < w.V && ((L.vtx[v].w >> 8) & 0xffu) == 0u;
        D4 cur{0, 0, 0, 0};
        uint32_t nb = kNone;
        if (need) {
            const uint4 r = L.vtx[v];
            cur = vertexPosition(c, si, r);
            if (2.0 * cur.w < double(kr)) need = false;
            L.vtx[v].w = (r.w & 0xffu) | (1u << 8);
        }
        // nearest non-neighbour site to the vertex sphere, all lanes' searches in step
        nb = vertexSearches(c, S, L, w, i, root, need, v < w.V ? L.vtx[v] : uint4{0, 0, 0, 0}, toF(cur), spill,
                            flags);
        bool invalid = false;
        if (nb != kNone) invalid = intersectD(cur, loadSite(c, nb));
        const uint64_t m = ballot(invalid);
        waveSync();
        uint64_t mm = m;
        while (mm) {
            const uint32_t t = uint32_t(__builtin_ctzll(mm));
            mm &= mm - 1;
            const uint32_t j = readLane(nb, t);
            if (knownSite(L, w, j)) continue;
            updateCell(c, L, w, i, si, j, loadSite(c, j));
        }
        if (m) {
            vs = 0;
            restarts++;
        }
    }
    if (restarts >= maxRestarts) flags |= kFlagRestartCap;
    uint32_t allFlags = flags;
    for (int o = 32; o >= 1; o >>= 1) allFlags |= __shfl_xor(allFlags, o, 64);
    w.flags |= allFlags;
    storeCell(c, L, w, i);
    // status, known radius, counts and site read; records read and written
    if (lane == 0) countBytes(c, kBytesVertexVal, 29u + 16u * uint64_t(loaded + w.V + w.E) + 8u);
}

// ---------------------------------------------------------------------------
// Contributor search around an edge (algorithm.inl:478-857)
// ---------------------------------------------------------------------------
struct EdgeQuery {
    F4 bs;
    bool hyperbola;
    F3 n1, n2, pp1, pp2;
    int cover;                // balls in `ball` (0: no pruning)
    F4 ball[kCoverBalls];     // every contributor j has |p_j - c| - r_j <= r for one ball (edge_bound.h)
};


// One step of a lane's contributor traversal (algorithm.inl:616-793): visit
// one internal node (descend to the closer accepted child, stack the other,
// or take the stack's top), then, when the node reached is a leaf, test it
// in the same step.  Returns the leaf's site when it is a candidate: other
// than i, k, l, not buried, touching the search sphere and, for a
// hyperbolic edge, on the right side of both tangent planes.  A lane walks
// its traversal in the reference's order; the wave never waits for lanes
// to reach leaves together.
__device__ __forceinline__ uint32_t edgeStep(const CellCtx& c, const VStack& stNode, uint32_t i, uint32_t k,
                                             uint32_t l, const EdgeQuery& q, Trav& t, uint32_t& flags) {
    const uint32_t lane = laneId();
    const uint32_t n = c.n;
    const uint32_t node = t.node;
    if (node < c.internal) {
        const Node nn = c.nodes[node];
        Box lb, rb;
        nodeBoxes(nn, lb, rb);
        const uint32_t L0 = uint32_t(nn.d.x), R0 = uint32_t(nn.d.y);
        bool ef = false, ec = false;
        uint32_t farChild = 0, closeChild = 0;
        if (L0 >= nodeCount(c) || R0 >= nodeCount(c)) {
            atomicOr(c.err, kErrKnnChild);
        } else {
            const F3 cpL = closestPoint(lb, q.bs), cpR = closestPoint(rb, q.bs);
            const float dL = boxDistance(lb, q.bs), dR = boxDistance(rb, q.bs);
            const bool lc = dL < dR;
            const float minFar = lc ? dR : dL;
            const F3 farPoint = selF3(lc, cpR, cpL);
            farChild = lc ? R0 : L0;
            const float minClose = lc ? dL : dR;
            const F3 closePoint = selF3(lc, cpL, cpR);
            closeChild = lc ? L0 : R0;
            ef = minFar < 0.f;
            ec = minClose < 0.f;
            if (q.hyperbola && c.complete) {
                // a subtree may hold a site inside both half-spaces wherever its
                // box reaches into them (the reference tests only the box point
                // closest to the search sphere's centre, algorithm.inl:660-670)
                const Box fb = selBox(lc, rb, lb), cb = selBox(lc, lb, rb);
                ef = ef && boxReachesHalfSpace(fb, q.n1, q.pp1) && boxReachesHalfSpace(fb, q.n2, q.pp2);
                ec = ec && boxReachesHalfSpace(cb, q.n1, q.pp1) && boxReachesHalfSpace(cb, q.n2, q.pp2);
            } else if (q.hyperbola) {
                if (ef) ef = dotf(q.n1, subf(farPoint, q.pp1)) < 0.f && dotf(q.n2, subf(farPoint, q.pp2)) < 0.f;
                if (ec) ec = dotf(q.n1, subf(closePoint, q.pp1)) < 0.f && dotf(q.n2, subf(closePoint, q.pp2)) < 0.f;
            }
            if (q.cover) {
                const float lr = __builtin_bit_cast(float, nn.d.z), rr = __builtin_bit_cast(float, nn.d.w);
                const Box fb = selBox(lc, rb, lb), cb = selBox(lc, lb, rb);
                const float fr = lc ? rr : lr, cr = lc ? lr : rr;
                bool tf = false, tc = false;
#pragma unroll
                for (int m = 0; m < kCoverBalls; m++) {
                    if (m < q.cover) {
                        tf = tf || boxNearBall(fb, q.ball[m], fr);
                        tc = tc || boxNearBall(cb, q.ball[m], cr);
                    }
                }
                ef = ef && tf;
                ec = ec && tc;
            }
        }
        if (ec || ef) {
            t.node = ec ? closeChild : farChild;
            if (ec && ef) {
                if (t.base + t.sp < kStackDepth) {
                    pushEntry(stNode, t, farChild);
                } else {
                    flags |= kFlagStackOverflow;
                }
            }
        } else {
            popNode(stNode, t);
        }
        if (!t.live || t.node < c.internal) return kNone;
    }
    const uint32_t leaf = t.node;
    uint32_t cand = kNone;
    if (leaf < nodeCount(c)) {
        const uint32_t j = leaf - c.internal;
        bool cc = j != i && j != k && j != l && c.status[j] != kStatusBuried;
        if (cc) {
            const F4 fsj = c.sorted[j];
            cc = intersectF(fsj, q.bs);
            if (q.hyperbola) {
                const F3 cj = xyzf(fsj);
                cc = cc && dotf(q.n1, subf(cj, q.pp1)) < -fsj.w && dotf(q.n2, subf(cj, q.pp2)) < -fsj.w;
            }
        }
        if (cc) cand = j;
    }
    popNode(stNode, t);
    return cand;
}

__device__ __forceinline__ uint32_t waveOr(uint32_t v) {
    for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

// Store the edges and append the vertices created during this kernel after
// the cell's stored vertices (algorithm.inl:840-855, 1216-1231).  In the
// completion mode the kernel updated the cell's whole vertex list, which
// replaces the stored one.
__device__ __forceinline__ void storeAppended(const CellCtx& c, const CellLds& L, WaveCell& w, uint32_t i,
                                              bool didChange, uint32_t family, uint32_t edgesLoaded) {
    const uint32_t lane = laneId();
    const uint32_t baseV = c.complete ? 0u : c.vnb[i];
    uint32_t total = baseV + w.V;
    if (total > kVertexCap) {
        total = kVertexCap;
        w.flags |= kFlagVertexOverflow;
    }
    uint4* gv = c.vtx + size_t(i) * kVertexCap;
    uint4* ge = c.edg + size_t(i) * kEdgeCap;
    for (uint32_t v = lane; baseV + v < total; v += 64) gv[baseV + v] = L.vtx[v];
    for (uint32_t e = lane; e < w.E; e += 64) ge[e] = L.edg[e];
    waveSync();
    if (lane == 0) {
        c.status[i] = didChange ? kStatusInValidation : kStatusCertified;
        c.enb[i] = w.E;
        c.vnb[i] = total;
        if (w.flags) c.flags[i] |= w.flags;
        // status, known radius, counts and site read, edges read; edges,
        // appended vertices, status and counts written
        countBytes(c, family, 29u + 16u * uint64_t(edgesLoaded + w.E + (total - baseV)) + 9u);
    }
}

// The contributor searches of all lanes' edges (algorithm.inl:616-793),
// each returning its first contributor in the reference's traversal order.
//
// Work sharing: a lane whose search is over takes the bottom entry of a
// busy lane's stack -- the subtree that lane would visit last -- together
// with a copy of that search's query, and searches it with a stack whose
// depth counts from that entry's depth (so overflow happens where the
// reference's would).  Everything a lane still holds precedes everything
// it gave away, so the owner lane's own item is always the earliest part of
// its search: a contributor found there is the answer and settles the
// search.  A contributor found in a shared item ends that item and is kept
// per owner when it precedes the one kept so far (visitedBefore); it is
// the answer if the owner's own item finds none.
__device__ __forceinline__ uint32_t edgeSearches(const CellCtx& c, EdgeLds& S, const CellLds& L, const WaveCell& w,
                                              uint32_t i, const D4& si, uint32_t root, bool need, uint32_t k,
                                              uint32_t l, const EdgeQuery& q, uint32_t* spill, uint32_t& flags) {
    const uint32_t lane = laneId();
    const VStack st{stackRing(S.cell), spill};
    S.res[lane] = kNone;
    uint64_t settled = 0;  // owner lanes whose own item found the contributor
    // the item this lane works on: the query it serves and its owner
    EdgeQuery wq = q;
    uint32_t wk = k, wl = l, owner = need ? lane : kNone;
    bool own = need;  // the owner's own (earliest) item
    D4 wsk{0, 0, 0, 0}, wsl{0, 0, 0, 0};
    if (need) {
        wsk = loadSite(c, k);
        wsl = loadSite(c, l);
    }
    Trav tr{root, 0u, need};
    uint32_t found = kNone;   // contributor of this lane's own item
    uint32_t cand = kNone;    // candidate waiting for the wave's test
    waveSync();
    while (true) {
        // a shared item of a settled search has nothing left to find
        if (tr.live && !own && owner != kNone && ((settled >> owner) & 1u)) tr.live = false;
        const uint64_t stepping = ballot(tr.live && cand == kNone), waiting = ballot(cand != kNone);
        if (stepping == 0 && waiting == 0) break;
        if (waiting != 0 && popc64(waiting) >= popc64(stepping)) {
            // test the waiting candidates together: not yet a neighbour,
            // two vertex solutions, one clear of the neighbours
            uint32_t shared = kNone;
            if (cand != kNone) {
                bool ok = !knownLane(L, w.F, cand);
                if (ok) {
                    D4 x0{0, 0, 0, 0}, x1{0, 0, 0, 0};
                    uint32_t mask = quadrisector(si, loadSite(c, cand), wsk, wsl, x0, x1);
                    if (mask != 3u) mask = 0u;
                    // two neighbours per step here: the search holds many live registers
                    const uint32_t valid = clearPairOfNeighbours<2>(c, L, w.F, x0, x1, mask, wk, wl);
                    ok = valid != 0u;
                }
                if (ok) {
                    if (own) found = cand;
                    else shared = cand;
                    tr.live = false;
                }
                cand = kNone;
            }
            settled |= ballot(own && found != kNone);
            // keep the earliest contributor of the shared items, per owner
            uint64_t hm = ballot(shared != kNone);
            while (hm) {
                const uint32_t t = uint32_t(__builtin_ctzll(hm));
                hm &= hm - 1;
                if (lane == t) {
                    const uint32_t cur = S.res[owner];
                    if (cur == kNone || visitedBefore(c, shared, cur, wq.bs)) S.res[owner] = shared;
                }
                waveSync();
            }
            waveSync();
            continue;
        }
        // idle lanes take the last subtree of busy lanes' stacks
        const uint64_t idle = ballot(!tr.live && cand == kNone);
        const uint64_t donors = ballot(tr.live && cand == kNone && tr.sp > tr.sb);
        if (popc64(idle) >= kShareIdle && donors != 0) {
            const uint32_t m = popc64(idle) < popc64(donors) ? popc64(idle) : popc64(donors);
            const uint32_t dr = countBelow(donors), ir = countBelow(idle);
            const bool give = ((donors >> lane) & 1u) && dr < m;
            const bool take = ((idle >> lane) & 1u) && ir < m;
            waveSync();  // the donors' stack slots are visible to the takers
            const uint32_t src = take ? laneOfRank(donors, ir) : lane;
            const uint32_t dSb = __shfl(tr.sb, int(src), 64), dBase = __shfl(tr.base, int(src), 64);
            const uint32_t dLo = __shfl(tr.lo, int(src), 64);
            const uint32_t dOwner = __shfl(owner, int(src), 64);
            const uint32_t node = take ? st.get(dSb, src, dLo) : 0u;
            wq.bs = shflF4(wq.bs, src);
            wq.hyperbola = __shfl(wq.hyperbola ? 1 : 0, int(src), 64) != 0;
            wq.n1 = shflF3(wq.n1, src);
            wq.n2 = shflF3(wq.n2, src);
            wq.pp1 = shflF3(wq.pp1, src);
            wq.pp2 = shflF3(wq.pp2, src);
            wq.cover = __shfl(wq.cover, int(src), 64);
#pragma unroll
            for (int b = 0; b < kCoverBalls; b++) wq.ball[b] = shflF4(wq.ball[b], src);
            wk = __shfl(wk, int(src), 64);
            wl = __shfl(wl, int(src), 64);
            wsk = shflD4(wsk, src);
            wsl = shflD4(wsl, src);
            if (take) {
                owner = dOwner;
                own = false;
                tr.node = node;
                tr.sp = 0u;
                tr.sb = 0u;
                tr.lo = 0u;
                tr.base = dBase + dSb;
                tr.live = true;
            }
            if (give) tr.sb++;
            waveSync();
        }
        if (tr.live && cand == kNone) cand = edgeStep(c, st, i, wk, wl, wq, tr, flags);
    }
    waveSync();
    return need && found == kNone ? S.res[lane] : found;
}

__device__ __forceinline__ void edgeTask(const CellCtx& c, EdgeLds& S, uint32_t i, uint32_t maxRestarts,
                                         uint32_t* scratch) {
    CoverLds& cover = *reinterpret_cast<CoverLds*>(scratch + kVertexSpillWords);
    CellLds& L = S.cell;
    const uint32_t lane = laneId();
    const uint8_t st = c.status[i];
    if (st == kStatusCertified || st == kStatusBuried) return;
    WaveCell w;
    // the reference's edge validation updates a vertex list that starts
    // empty (algorithm.inl:532), so a contributor never invalidates the
    // cell's stored vertices; the completion mode updates the whole list
    loadCell(c, L, w, i, c.complete != 0u);
    const uint32_t edgesLoaded = w.E;
    for (uint32_t e = lane; e < w.E; e += 64)
        if (L.edg[e].w != 1u) L.edg[e].w = 2u;
    waveSync();
    const float kr = c.knownRadii[i];
    const D4 si = loadSite(c, i);
    const uint32_t root = rootNode(c, cellMolecule(c, i));
    const F4 fsi = toF(si);
    uint32_t flags = 0;
    bool didChange = false;
    const uint32_t stepNb = w.E / 64 + 1;
    uint32_t es = 0, restarts = 0;
    while (es < stepNb) {
        buildNeighbours(c, L, w);
        const uint32_t e = es * 64 + lane;
        const bool inRange = e < w.E;
        bool need = inRange && L.edg[e].w == 2u;
        uint32_t k = kNone, l = kNone;
        if (inRange) {
            k = L.edg[e].x;
            l = L.edg[e].y;
        }
        EdgeQuery q;
        q.bs = {0, 0, 0, 0};
        q.hyperbola = false;
        q.n1 = q.n2 = q.pp1 = q.pp2 = {0, 0, 0};
        q.cover = 0;
#pragma unroll
        for (int m = 0; m < kCoverBalls; m++) q.ball[m] = {0, 0, 0, 0};
        if (need) {
            const D4 sk = loadSite(c, k), sl = loadSite(c, l);
            D4 ev0{0, 0, 0, 0}, ev1{0, 0, 0, 0};
            const uint32_t type = trisector(si, sk, sl, ev0, ev1);
            q.hyperbola = type == 1;
            if (type == 2) {
                const F4 v1 = toF(ev0), v2 = toF(ev1);
                const F3 c1 = xyzf(v1), c2 = xyzf(v2);
                const F3 center = mulf(addf(c1, c2), .5f);
                const float radius = lenf(subf(center, c1)) + v1.w;
                q.bs = {center.x, center.y, center.z, radius};
            } else {
                D3 nn1, nn2;
                tangentPlanes(si, sk, sl, nn1, nn2);
                q.n1 = toF3(nn1);
                q.n2 = toF3(nn2);
                const D3 p1 = add(xyz(si), mul(nn1, si.w));
                const D3 p2 = add(xyz(sk), mul(nn1, sk.w));
                const D3 p3 = add(xyz(sl), mul(nn1, sl.w));
                const D3 p4 = add(xyz(si), mul(nn2, si.w));
                const D3 p5 = add(xyz(sk), mul(nn2, sk.w));
                q.bs = toF(sphereFromPoints(p1, p2, p3, p5));
                q.pp1 = toF3(p1);
                q.pp2 = toF3(p4);
            }
        }
        const bool inSec = ltF(lenf(subf(xyzf(fsi), xyzf(q.bs))), kr - q.bs.w);
        if (inRange && inSec) {
            L.edg[e].w = 1u;
            need = false;
        }
        // searches whose sphere is large get a cover of the edge's empty
        // spheres, built by the whole wave one edge at a time
        uint64_t bm = ballot(need && q.bs.w > c.coverRatio * kr);
        while (bm) {
            const uint32_t t = uint32_t(__builtin_ctzll(bm));
            bm &= bm - 1;
            double* ldsRoots = reinterpret_cast<double*>(L.list);
            const CoverBuf buf = 2u * w.F + 1u <= kCoverLdsRoots ? CoverBuf{ldsRoots, ldsRoots + kCoverLdsRoots}
                                                                 : CoverBuf{cover.roots, cover.sorted};
            const int cnt = waveEdgeCover(c, L, buf, S.coverOut, w.F, si, readLane(k, t), readLane(l, t));
            const bool mine = cnt > 0 && lane == t;
            q.cover = mine ? cnt : q.cover;
#pragma unroll
            for (int m = 0; m < kCoverBalls; m++) q.ball[m] = selF4(mine, S.coverOut.ball[m], q.ball[m]);
            waveSync();
        }
        // the first contributor of each lane's traversal: a candidate with two
        // vertex solutions of which one is clear of the cell's neighbours
        const uint32_t found = edgeSearches(c, S, L, w, i, si, root, need, k, l, q, scratch, flags);
        if (need && found == kNone) L.edg[e].w = 1u;
        const uint64_t m = ballot(found != kNone);
        waveSync();
        bool changed = false;
        uint64_t mm = m;
        while (mm) {
            const uint32_t t = uint32_t(__builtin_ctzll(mm));
            mm &= mm - 1;
            const uint32_t j = readLane(found, t);
            if (knownSite(L, w, j)) continue;
            const bool ch = updateCell(c, L, w, i, si, j, loadSite(c, j));
            didChange = didChange || ch;
            changed = changed || ch;
        }
        if (changed && restarts < maxRestarts) {
            es = 1;
            restarts++;
        } else {
            if (changed) flags |= kFlagRestartCap;
            es++;
        }
    }
    w.flags |= waveOr(flags);
    storeAppended(c, L, w, i, didChange, kBytesEdgeVal, edgesLoaded);
}

// ---------------------------------------------------------------------------
// Contributor search along a bisector (algorithm.inl:859-1233)
// ---------------------------------------------------------------------------
// One step of a lane's bisector traversal (algorithm.inl:1019-1176); a leaf
// reached by the step's descent or pop is tested in the same step.
// Returns the leaf's site when it is a candidate: other than i, bk, inside
// the search sphere and the cylinder around the bisector axis.
// With axisPrune (no negative radius in the scene) a subtree whose box
// cannot hold a site passing the cylinder test is not entered (boxNearAxis):
// no candidate is lost and the others are visited in the same order, so the
// search returns the same site.
__device__ __forceinline__ uint32_t bisectorStep(const CellCtx& c, const VStack& stNode, uint32_t i, uint32_t bk,
                                                 F3 fci, F3 fck, F4 bs, float cylR, bool axisPrune, Trav& t,
                                                 uint32_t& flags) {
    const uint32_t lane = laneId();
    const uint32_t n = c.n;
    const uint32_t node = t.node;
    if (node < c.internal) {
        const Node nn = c.nodes[node];
        Box lb, rb;
        nodeBoxes(nn, lb, rb);
        const uint32_t L0 = uint32_t(nn.d.x), R0 = uint32_t(nn.d.y);
        bool ef = false, ec = false;
        uint32_t farChild = 0, closeChild = 0;
        if (L0 >= nodeCount(c) || R0 >= nodeCount(c)) {
            atomicOr(c.err, kErrKnnChild);
        } else {
            const float dL = boxDistance(lb, bs), dR = boxDistance(rb, bs);
            const bool lc = dL < dR;
            const float minFar = lc ? dR : dL;
            farChild = lc ? R0 : L0;
            const float minClose = lc ? dL : dR;
            closeChild = lc ? L0 : R0;
            ef = minFar < 0.f;
            ec = minClose < 0.f;
            if (axisPrune && (ef || ec)) {
                const F3 u = normalizef(subf(fck, fci));
                if (ef) ef = boxNearAxis(selBox(lc, rb, lb), fci, u, cylR);
                if (ec) ec = boxNearAxis(selBox(lc, lb, rb), fci, u, cylR);
            }
        }
        if (ec || ef) {
            t.node = ec ? closeChild : farChild;
            if (ec && ef) {
                if (t.base + t.sp < kStackDepth) {
                    pushEntry(stNode, t, farChild);
                } else {
                    flags |= kFlagStackOverflow;
                }
            }
        } else {
            popNode(stNode, t);
        }
        if (!t.live || t.node < c.internal) return kNone;
    }
    const uint32_t leaf = t.node;
    uint32_t cand = kNone;
    if (leaf < nodeCount(c)) {
        const uint32_t j = leaf - c.internal;
        bool cc = j != i && j != bk;
        if (cc) {
            const F4 fsj = c.sorted[j];
            const F3 fcj = xyzf(fsj);
            cc = bs.w >= fsj.w && (lenf(subf(xyzf(bs), fcj)) <= (bs.w - fsj.w));
            if (cc) {
                const F3 itoj = subf(fcj, fci);
                const float v = dotf(itoj, normalizef(subf(fck, fci)));
                const float d = __builtin_sqrtf(dotf(itoj, itoj) - v * v);
                cc = cc && (d + fsj.w) < cylR;
            }
        }
        if (cc) cand = j;
    }
    popNode(stNode, t);
    return cand;
}

// The contributor searches of all lanes' bisectors (algorithm.inl:1019-1176),
// with the work sharing of the edge searches: the owner's own item holds
// the earliest part of its search and settles it; contributors of shared
// items are kept per owner in traversal order.
__device__ __forceinline__ uint32_t bisectorSearches(const CellCtx& c, BisectorLds& S, const CellLds& L,
                                                     const WaveCell& w, uint32_t i, const D4& si, uint32_t root,
                                                     bool need, uint32_t bk, D4 sk, F3 fci, F3 fck, F4 bsph,
                                                     float cylR, bool axisPrune, uint32_t* spill, uint32_t& flags) {
    const uint32_t lane = laneId();
    const VStack st{stackRing(S.cell), spill};
    S.res[lane] = kNone;
    uint64_t settled = 0;  // owner lanes whose own item found the contributor
    uint32_t owner = need ? lane : kNone;
    bool own = need;
    Trav tr{root, 0u, need};
    uint32_t found = kNone, cand = kNone;
    waveSync();
    while (true) {
        if (tr.live && !own && owner != kNone && ((settled >> owner) & 1u)) tr.live = false;
        const uint64_t stepping = ballot(tr.live && cand == kNone), waiting = ballot(cand != kNone);
        if (stepping == 0 && waiting == 0) break;
        if (waiting != 0 && popc64(waiting) >= popc64(stepping)) {
            uint32_t shared = kNone;
            if (cand != kNone) {
                bool ok = !knownLane(L, w.F, cand);
                if (ok) {
                    D4 x0{0, 0, 0, 0}, x1{0, 0, 0, 0};
                    const uint32_t cnt = trisector(si, loadSite(c, cand), sk, x0, x1);
                    ok = cnt == 2u && clearOfNeighbours(c, L, w.F, x1, bk, bk);
                }
                if (ok) {
                    if (own) found = cand;
                    else shared = cand;
                    tr.live = false;
                }
                cand = kNone;
            }
            settled |= ballot(own && found != kNone);
            uint64_t hm = ballot(shared != kNone);
            while (hm) {
                const uint32_t t = uint32_t(__builtin_ctzll(hm));
                hm &= hm - 1;
                if (lane == t) {
                    const uint32_t cur = S.res[owner];
                    if (cur == kNone || visitedBefore(c, shared, cur, bsph)) S.res[owner] = shared;
                }
                waveSync();
            }
            waveSync();
            continue;
        }
        const uint64_t idle = ballot(!tr.live && cand == kNone);
        const uint64_t donors = ballot(tr.live && cand == kNone && tr.sp > tr.sb);
        if (popc64(idle) >= kShareIdle && donors != 0) {
            const uint32_t m = popc64(idle) < popc64(donors) ? popc64(idle) : popc64(donors);
            const uint32_t dr = countBelow(donors), ir = countBelow(idle);
            const bool give = ((donors >> lane) & 1u) && dr < m;
            const bool take = ((idle >> lane) & 1u) && ir < m;
            waveSync();  // the donors' stack slots are visible to the takers
            const uint32_t src = take ? laneOfRank(donors, ir) : lane;
            const uint32_t dSb = __shfl(tr.sb, int(src), 64), dBase = __shfl(tr.base, int(src), 64);
            const uint32_t dLo = __shfl(tr.lo, int(src), 64);
            const uint32_t dOwner = __shfl(owner, int(src), 64);
            const uint32_t node = take ? st.get(dSb, src, dLo) : 0u;
            bk = __shfl(bk, int(src), 64);
            sk = shflD4(sk, src);
            fck = shflF3(fck, src);
            bsph = shflF4(bsph, src);
            cylR = __shfl(cylR, int(src), 64);
            if (take) {
                owner = dOwner;
                own = false;
                tr.node = node;
                tr.sp = 0u;
                tr.sb = 0u;
                tr.lo = 0u;
                tr.base = dBase + dSb;
                tr.live = true;
            }
            if (give) tr.sb++;
            waveSync();
        }
        if (tr.live && cand == kNone)
            cand = bisectorStep(c, st, i, bk, fci, fck, bsph, cylR, axisPrune, tr, flags);
    }
    waveSync();
    return need && found == kNone ? S.res[lane] : found;
}

__device__ __forceinline__ void bisectorTask(const CellCtx& c, BisectorLds& S, uint32_t i, uint32_t maxRestarts,
                                             uint32_t* spill) {
    CellLds& L = S.cell;
    const uint32_t lane = laneId();
    const uint8_t st = c.status[i];
    if (st == kStatusBuried) return;
    WaveCell w;
    loadCell(c, L, w, i, c.complete != 0u);
    const uint32_t edgesLoaded = w.E;
    uint32_t flags = 0;
    // sorted distinct neighbours above i (algorithm.inl:903-968)
    buildNeighbours(c, L, w);
    const uint32_t F = w.F;
    uint32_t B = 0;
    for (uint32_t b = 0; b < F; b += 64) {
        const uint32_t q = b + lane;
        uint32_t id = kNone;
        bool keep = false;
        if (q < F) {
            id = L.nb[q];
            keep = id < c.n && id > i;
        }
        uint32_t rank = 0;
        if (keep)
            for (uint32_t q2 = 0; q2 < F; q2++) {
                const uint32_t o = L.nb[q2];
                if (o < c.n && o > i && o < id) rank++;
            }
        if (keep && rank < kBisectorCap) S.bis[rank] = id;
        B += popc64(ballot(keep));
    }
    if (B > kBisectorCap) {
        B = kBisectorCap;
        flags |= kFlagBisectorOverflow;
    }
    waveSync();
    // the vertex list of this kernel starts empty (algorithm.inl:970-971),
    // except in the completion mode, which updates the whole list
    if (!c.complete) w.V = 0;
    const float kr = c.knownRadii[i];
    const D4 si = loadSite(c, i);
    const uint32_t root = rootNode(c, cellMolecule(c, i));
    const bool axisPrune = c.scene[cellMolecule(c, i)].rmin >= 0.0;
    const F3 fci{float(si.x), float(si.y), float(si.z)};
    const float fri = float(si.w);
    bool didChange = false;
    const uint32_t stepNb = B / 64 + 1;
    uint32_t bsi = 0, restarts = 0;
    while (bsi < stepNb) {
        buildNeighbours(c, L, w);
        const uint32_t bid = bsi * 64 + lane;
        bool need = bid < B && S.bis[bid] != kNone;
        uint32_t found = kNone, bk = kNone;
        D4 sk{0, 0, 0, 0};
        F3 fck{0, 0, 0};
        F4 bsph{0, 0, 0, 0};
        float cylR = 0.f;
        if (need) {
            bk = S.bis[bid];
            sk = loadSite(c, bk);
            fck = F3{float(sk.x), float(sk.y), float(sk.z)};
            const float frk = float(sk.w);
            const float dist = lenf(subf(fci, fck)) + fri + frk;
            if (ltF(dist, kr)) {
                S.bis[bid] = kNone;
                need = false;
            }
            const F3 center = mulf(addf(fci, fck), .5f);
            const float radius = lenf(subf(center, fci)) + __builtin_fmaxf(fri, frk);
            bsph = F4{center.x, center.y, center.z, radius};
            cylR = __builtin_fmaxf(fri, frk);
        }
        // the first site of each lane's traversal whose trisector with i, bk
        // has two points, the second clear of the cell's neighbours
        found = bisectorSearches(c, S, L, w, i, si, root, need, bk, sk, fci, fck, bsph, cylR, axisPrune, spill, flags);
        if (need && found == kNone) S.bis[bid] = kNone;
        const uint64_t m = ballot(found != kNone);
        waveSync();
        bool changed = false;
        uint64_t mm = m;
        while (mm) {
            const uint32_t t = uint32_t(__builtin_ctzll(mm));
            mm &= mm - 1;
            const uint32_t j = readLane(found, t);
            if (knownSite(L, w, j)) continue;
            const bool ch = updateCell(c, L, w, i, si, j, loadSite(c, j));
            didChange = didChange || ch;
            changed = changed || ch;
        }
        if (changed && restarts < maxRestarts) {
            bsi = 1;
            restarts++;
        } else {
            if (changed) flags |= kFlagRestartCap;
            bsi++;
        }
    }
    w.flags |= waveOr(flags);
    storeAppended(c, L, w, i, didChange, kBytesBisectorVal, edgesLoaded);
}

// ---------------------------------------------------------------------------
// The bisector stage split by search.  A cell runs a handful of bisector
// searches (about five), so a wave that serves one cell keeps most of its
// lanes idle and waits on its few traversals.  The split stage gives each
// lane a search of its own, from many cells, and leaves the per-cell stage
// (bisectorTask) only the cells where a search finds a contributor:
//  * k_bisector_plan, a wave per cell: the cell's sorted distinct bisectors
//    above i (algorithm.inl:903-968), less those inside its known radius
//    (:993-1006), go to the cell's search slots, and its distinct
//    neighbours to its neighbour slots; the cell's status becomes validated
//    (a cell with more searches or neighbours than slots is marked);
//  * k_bisector_lanes, a search per lane: waves take runs of consecutive
//    cells and hand their searches to idle lanes; each runs the reference's
//    traversal (bisectorStep) and candidate test (not a neighbour, a
//    two-point trisector whose second point no neighbour other than bk
//    intersects) against the neighbour slots; a contributor marks its cell;
//  * the per-cell stage over the marked cells, updates included.
// Until a cell's first update, every search of the reference's stage sees
// the cell as stored, which is what the lanes search against; a cell none of
// whose searches finds a contributor is never updated, so the reference
// leaves it validated with the flags of its searches (stack overflow at the
// reference's depth: the lanes walk the same order) and of its bisector list.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bisectorPlanTask(const CellCtx& c, BisectorLds& S, uint32_t i,
                                                 const BisectorSplit& sp) {
    CellLds& L = S.cell;
    const uint32_t lane = laneId();
    if (c.status[i] == kStatusBuried) {
        if (lane == 0) {
            sp.count[i] = 0u;
            sp.mark[i] = 0u;
        }
        return;
    }
    WaveCell w;
    loadCell(c, L, w, i, false);
    uint32_t flags = 0;
    buildNeighbours(c, L, w);
    const uint32_t F = w.F;
    // sorted distinct neighbours above i, as bisectorTask lists them
    uint32_t B = 0;
    for (uint32_t b = 0; b < F; b += 64) {
        const uint32_t q = b + lane;
        uint32_t id = kNone;
        bool keep = false;
        if (q < F) {
            id = L.nb[q];
            keep = id < c.n && id > i;
        }
        uint32_t rank = 0;
        if (keep)
            for (uint32_t q2 = 0; q2 < F; q2++) {
                const uint32_t o = L.nb[q2];
                if (o < c.n && o > i && o < id) rank++;
            }
        if (keep && rank < kBisectorCap) S.bis[rank] = id;
        B += popc64(ballot(keep));
    }
    if (B > kBisectorCap) {
        B = kBisectorCap;
        flags |= kFlagBisectorOverflow;
    }
    waveSync();
    // the searches the reference runs: bisectors outside the known radius
    const float kr = c.knownRadii[i];
    const D4 si = loadSite(c, i);
    const F3 fci{float(si.x), float(si.y), float(si.z)};
    const float fri = float(si.w);
    uint32_t T = 0;
    for (uint32_t b = 0; b < B; b += 64) {
        const uint32_t q = b + lane;
        uint32_t bk = kNone;
        bool run = false;
        if (q < B) {
            bk = S.bis[q];
            const D4 sk = loadSite(c, bk);
            const F3 fck{float(sk.x), float(sk.y), float(sk.z)};
            const float dist = lenf(subf(fci, fck)) + fri + float(sk.w);
            run = !ltF(dist, kr);
        }
        const uint64_t m = ballot(run);
        waveSync();
        if (run) S.bis[T + countBelow(m)] = bk;
        T += popc64(m);
        waveSync();
    }
    const bool fits = T <= kBisSearchSlots && F <= kBisNeighbourSlots;
    if (fits) {
        if (lane < T) sp.search[size_t(i) * kBisSearchSlots + lane] = S.bis[lane];
        if (lane < F) sp.nb[size_t(i) * kBisNeighbourSlots + lane] = L.nb[lane];
    }
    if (lane == 0) {
        sp.count[i] = fits ? (T | (F << 16)) : 0u;
        sp.mark[i] = fits ? 0u : 1u;
        c.status[i] = kStatusCertified;
        if (flags) c.flags[i] |= flags;
        // status, known radius, counts and edges read; searches, neighbours,
        // counts, mark and status written
        countBytes(c, kBytesBisectorVal, 25u + 16u * uint64_t(w.E) + 4u * (T + F) + 10u);
    }
}

// Each wave plans a contiguous share of the cells (every cell costs about the
// same here, and a work queue's atomics would cost more than the planning).
__global__ __launch_bounds__(64) void k_bisector_plan(CellCtx c, BisectorSplit sp) {
    __shared__ BisectorLds S;
    const uint32_t per = (c.n + gridDim.x - 1u) / gridDim.x;
    const uint32_t lo = blockIdx.x * per, hi = lo + per < c.n ? lo + per : c.n;
    for (uint32_t i = lo; i < hi; i++) {
        bisectorPlanTask(c, S, i, sp);
        waveSync();
    }
}

// Neighbour id j is among the cell's F neighbour slots (four per step, their
// loads in flight together).
__device__ __forceinline__ bool slotKnown(const uint32_t* __restrict__ nb, uint32_t F, uint32_t j) {
    uint32_t q = 0;
    for (; q + 4u <= F; q += 4u) {
        const uint32_t a = nb[q], b = nb[q + 1u], d = nb[q + 2u], e = nb[q + 3u];
        if (a == j || b == j || d == j || e == j) return true;
    }
    for (; q < F; q++)
        if (nb[q] == j) return true;
    return false;
}

// No neighbour other than skip intersects the sphere x (clearOfNeighbours
// over the neighbour slots).
__device__ __forceinline__ bool slotClear(const CellCtx& c, const uint32_t* __restrict__ nb, uint32_t F, const D4& x,
                                          uint32_t skip) {
    for (uint32_t q = 0; q < F; q++) {
        const uint32_t id = nb[q];
        if (id == skip) continue;
        if (intersectD(loadSite(c, id), x)) return false;
    }
    return true;
}

// Cells per run a wave takes from the queue, and the idle lanes that make
// the wave hand out searches (or all lanes).
constexpr uint32_t kBisRunCells = 32;
constexpr uint32_t kLaneRefill = 16;

__global__ __launch_bounds__(64) void k_bisector_lanes(CellCtx c, uint32_t* counters, BisectorSplit sp,
                                                       uint32_t* spill) {
    __shared__ uint32_t ring[kVertexLdsDepth][64];
    __shared__ uint32_t runEnd[kBisRunCells];  // inclusive prefix of the run's search counts
    uint32_t* mySpill = spill + size_t(blockIdx.x) * kWaveScratchWords;
    const VStack st{ring, mySpill};
    const uint32_t lane = laneId();
    const uint32_t runs = (c.n + kBisRunCells - 1u) / kBisRunCells;
    // the wave's current run: cells [run0, run0 + 32), searches [next, total)
    uint32_t run0 = 0, next = 0, total = 0;
    bool drained = false;
    Trav tr{0u, 0u, false};
    uint32_t i = kNone, bk = kNone, cand = kNone, flags = 0, F = 0;
    const uint32_t* nb = nullptr;
    D4 si{0, 0, 0, 0}, sk{0, 0, 0, 0};
    F3 fci{0, 0, 0}, fck{0, 0, 0};
    F4 bsph{0, 0, 0, 0};
    float cylR = 0.f;
    bool axisPrune = false;
    for (;;) {
        uint64_t idle = ballot(!tr.live && cand == kNone);
        if (!drained && (popc64(idle) >= kLaneRefill || idle == ~uint64_t(0))) {
            if (next == total) {
                // the next run of cells and the prefix of their search counts
                const uint32_t r = popWork(counters + kBisectorQueue, runs);
                if (r == kNone) {
                    drained = true;
                } else {
                    run0 = r * kBisRunCells;
                    const uint32_t cell = run0 + lane;
                    const uint32_t cnt = lane < kBisRunCells && cell < c.n ? (sp.count[cell] & 0xffffu) : 0u;
                    const uint32_t incl = waveInclusiveSum(cnt);
                    if (lane < kBisRunCells) runEnd[lane] = incl;
                    total = readLane(incl, kBisRunCells - 1u);
                    next = 0;
                    waveSync();
                }
            }
            if (next < total) {
                // idle lanes take the run's next searches in order
                const uint32_t rank = countBelow(idle);
                const uint32_t k = next + rank;
                if (((idle >> lane) & 1u) && k < total) {
                    uint32_t lo = 0;
                    while (runEnd[lo] <= k) lo++;  // the run's cell holding search k
                    i = run0 + lo;
                    const uint32_t first = lo ? runEnd[lo - 1u] : 0u;
                    const uint32_t cnt = sp.count[i];
                    F = cnt >> 16;
                    nb = sp.nb + size_t(i) * kBisNeighbourSlots;
                    bk = sp.search[size_t(i) * kBisSearchSlots + (k - first)];
                    si = loadSite(c, i);
                    sk = loadSite(c, bk);
                    fci = F3{float(si.x), float(si.y), float(si.z)};
                    fck = F3{float(sk.x), float(sk.y), float(sk.z)};
                    const float fri = float(si.w), frk = float(sk.w);
                    const F3 center = mulf(addf(fci, fck), .5f);
                    const float radius = lenf(subf(center, fci)) + __builtin_fmaxf(fri, frk);
                    bsph = F4{center.x, center.y, center.z, radius};
                    cylR = __builtin_fmaxf(fri, frk);
                    const uint32_t mol = cellMolecule(c, i);
                    axisPrune = c.scene[mol].rmin >= 0.0;
                    tr = Trav{rootNode(c, mol), 0u, true};
                    flags = 0;
                }
                const uint32_t given = popc64(idle);
                next = next + given < total ? next + given : total;
            }
        }
        const uint64_t stepping = ballot(tr.live && cand == kNone), waiting = ballot(cand != kNone);
        if (stepping == 0 && waiting == 0) {
            if (drained) break;
            continue;
        }
        if (waiting != 0 && popc64(waiting) >= popc64(stepping)) {
            // the waiting candidates together (bisectorSearches' test)
            if (cand != kNone) {
                bool ok = !slotKnown(nb, F, cand);
                if (ok) {
                    D4 x0{0, 0, 0, 0}, x1{0, 0, 0, 0};
                    const uint32_t cnt = trisector(si, loadSite(c, cand), sk, x0, x1);
                    ok = cnt == 2u && slotClear(c, nb, F, x1, bk);
                }
                if (ok) {
                    sp.mark[i] = 1u;
                    tr.live = false;
                }
                cand = kNone;
                if (!tr.live && flags) {
                    atomicOr(&c.flags[i], flags);
                    flags = 0;
                }
            }
            continue;
        }
        if (tr.live && cand == kNone) {
            cand = bisectorStep(c, st, i, bk, fci, fck, bsph, cylR, axisPrune, tr, flags);
            if (!tr.live && cand == kNone && flags) {
                atomicOr(&c.flags[i], flags);
                flags = 0;
            }
        }
    }
}

// Cells marked for the per-cell stage, appended in wave order.
__global__ __launch_bounds__(256) void k_compact_marked(const uint8_t* __restrict__ mark, uint32_t n,
                                                        uint32_t* list, uint32_t* count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool take = i < n && mark[i] != 0u;
    const uint64_t m = ballot(take);
    uint32_t base = 0;
    if (laneId() == 0 && m) base = atomicAdd(count, popc64(m));
    base = __shfl(base, 0, 64);
    if (take) list[base + countBelow(m)] = i;
}

// Tasks a validation wave claims per queue atomic: one (the costs of cells
// vary, and neighbouring cells cost alike, so longer claims unbalance the
// waves).
constexpr uint32_t kTaskBatch = 1;

// ---------------------------------------------------------------------------
// Persistent launches: every wave takes cells from a device work queue
// until it is drained, so the grid is sized to the chip, not to the task
// count, and no launch depends on a count read back by the host.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_vertex_validation(CellCtx c, const uint32_t* __restrict__ tasks,
                                                          uint32_t* counters, uint32_t* roundTasks,
                                                          uint32_t maxRestarts, uint32_t* spill) {
    __shared__ VertexLds S;
    // the deep stack slots of this wave (the grid has at most one block per
    // wave region: validationLaunch checks it)
    uint32_t* mySpill = spill + size_t(blockIdx.x) * kWaveScratchWords;
    const uint32_t count = __builtin_amdgcn_readfirstlane(counters[kTaskCount]);
    if (blockIdx.x == 0 && laneId() == 0) *roundTasks = count;
    WorkClaim claim;
    for (;;) {
        const uint32_t t = popWorkBatch(counters + kVertexQueue, count, claim, kTaskBatch);
        if (t == kNone) break;
        vertexTask(c, S, tasks[t], maxRestarts, mySpill);
        waveSync();
    }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_edge_validation(CellCtx c, const uint32_t* __restrict__ tasks,
                                                        uint32_t* counters, uint32_t maxRestarts, uint32_t* scratch) {
    __shared__ EdgeLds S;
    uint32_t* myScratch = scratch + size_t(blockIdx.x) * kWaveScratchWords;
    const uint32_t count = __builtin_amdgcn_readfirstlane(counters[kTaskCount]);
    WorkClaim claim;
    for (;;) {
        const uint32_t t = popWorkBatch(counters + kEdgeQueue, count, claim, kTaskBatch);
        if (t == kNone) break;
        edgeTask(c, S, tasks[t], maxRestarts, myScratch);
        waveSync();
    }
}

// Every cell, or (list) the listed cells: the per-cell bisector stage.
__global__ __launch_bounds__(64) void k_bisector_validation(CellCtx c, uint32_t* counters, uint32_t maxRestarts,
                                                            uint32_t* spill, const uint32_t* __restrict__ list) {
    __shared__ BisectorLds S;
    uint32_t* mySpill = spill + size_t(blockIdx.x) * kWaveScratchWords;
    const uint32_t count = list ? readFirst(counters[kBisMarked]) : c.n;
    WorkClaim claim;
    for (;;) {
        const uint32_t t = popWorkBatch(counters + kBisectorQueue, count, claim, kTaskBatch);
        if (t == kNone) break;
        bisectorTask(c, S, list ? list[t] : t, maxRestarts, mySpill);
        waveSync();
    }
}

// Cells still in validation, appended in wave order (one atomic per wave).
__global__ __launch_bounds__(256) void k_compact_tasks(const uint8_t* __restrict__ status, uint32_t n, uint32_t* tasks,
                                                       uint32_t* count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool take = i < n && status[i] == kStatusInValidation;
    const uint64_t m = ballot(take);
    uint32_t base = 0;
    if (laneId() == 0 && m) base = atomicAdd(count, popc64(m));
    base = __shfl(base, 0, 64);
    if (take) tasks[base + countBelow(m)] = i;
}

// Clears the task count and the vertex / edge queue heads of a round.
__global__ void k_round_reset(uint32_t* counters) {
    const uint32_t t = threadIdx.x;
    if (t < kRoundWords) counters[kTaskCount + t] = 0u;
}

uint32_t waveScratchWords() { return kWaveScratchWords; }

hipError_t validationOccupancy(uint32_t* vertexBlocks, uint32_t* edgeBlocks, uint32_t* bisectorBlocks,
                               uint32_t* bisectorPlanBlocks, uint32_t* bisectorLaneBlocks) {
    int v = 0, e = 0, b = 0, p = 0, l = 0;
    hipError_t r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_vertex_validation, 64, 0);
    if (r == hipSuccess) r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&e, k_edge_validation, 64, 0);
    if (r == hipSuccess) r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_bisector_validation, 64, 0);
    if (r == hipSuccess) r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, k_bisector_plan, 64, 0);
    if (r == hipSuccess) r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&l, k_bisector_lanes, 64, 0);
    *vertexBlocks = uint32_t(v > 0 ? v : 1);
    *edgeBlocks = uint32_t(e > 0 ? e : 1);
    *bisectorBlocks = uint32_t(b > 0 ? b : 1);
    *bisectorPlanBlocks = uint32_t(p > 0 ? p : 1);
    *bisectorLaneBlocks = uint32_t(l > 0 ? l : 1);
    return r;
}

hipError_t validationLaunch(const CellCtx& c, const ValidationBuffers& v, const ValidationGrid& g,
                            uint32_t iterations, uint32_t maxRestarts, bool bisector, hipStream_t st,
                            ValidationTiming* tm, uint32_t* roundTasks, const BisectorSplit* split) {
    if (!roundTasks) roundTasks = v.counters + kRoundTasks;
    if (bisector) {
        if (g.bisector > v.spillBlocks || g.bisectorLanes > v.spillBlocks) return hipErrorInvalidValue;
        const uint32_t* list = nullptr;
        if (split && split->search && g.bisectorLanes) {
            // plan, search per lane, then the marked cells per cell (a list
            // in the validation task buffer, free until the next round)
            hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, v.counters + kBisectorQueue, kXcdQueueWords,
                               0u);
            hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, v.counters + kBisMarked, 1u, 0u);
            hipLaunchKernelGGL(k_bisector_plan, dim3(g.bisectorPlan), dim3(64), 0, st, c, *split);
            hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, v.counters + kBisectorQueue, kXcdQueueWords,
                               0u);
            hipLaunchKernelGGL(k_bisector_lanes, dim3(g.bisectorLanes), dim3(64), 0, st, c, v.counters, *split,
                               v.spill);
            hipLaunchKernelGGL(k_compact_marked, dim3((c.n + 255) / 256), dim3(256), 0, st, split->mark, c.n, v.tasks,
                               v.counters + kBisMarked);
            list = v.tasks;
        }
        hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, v.counters + kBisectorQueue, kXcdQueueWords, 0u);
        hipLaunchKernelGGL(k_bisector_validation, dim3(g.bisector), dim3(64), 0, st, c, v.counters, maxRestarts,
                           v.spill, list);
        return hipGetLastError();
    }
    if (iterations > kMaxRounds || g.vertex > v.spillBlocks || g.edge > v.spillBlocks) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    for (uint32_t it = 0; it < iterations && e == hipSuccess; it++) {
        hipLaunchKernelGGL(k_round_reset, dim3(1), dim3(64), 0, st, v.counters);
        hipLaunchKernelGGL(k_compact_tasks, dim3((c.n + 255) / 256), dim3(256), 0, st, c.status, c.n, v.tasks,
                           v.counters + kTaskCount);
        const bool timed = tm && it < kTimedRounds;
        if (timed) e = hipEventRecord(tm->ev[it][0], st);
        hipLaunchKernelGGL(k_vertex_validation, dim3(g.vertex), dim3(64), 0, st, c, v.tasks, v.counters,
                           roundTasks + it, maxRestarts, v.spill);
        if (timed && e == hipSuccess) e = hipEventRecord(tm->ev[it][1], st);
        hipLaunchKernelGGL(k_edge_validation, dim3(g.edge), dim3(64), 0, st, c, v.tasks, v.counters, maxRestarts,
                           v.spill);
        if (timed && e == hipSuccess) e = hipEventRecord(tm->ev[it][2], st);
        if (e == hipSuccess) e = hipGetLastError();
    }
    if (tm) tm->rounds = iterations < kTimedRounds ? iterations : kTimedRounds;
    return e;
}

}  // namespace apo_hip
