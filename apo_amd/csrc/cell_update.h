// cell_update.h -- cell update shared by the cell and validation kernels (a wave64, or a half wave, per cell).
//
// Reference: apo/include/apo/gpu/topology_update.inl:5-653 (updateCell,
// cellRadius, the initial octahedron).  The cell's vertex and edge records
// live in LDS; lanes own records; the reference's ballot-ordered serial
// steps become ballot + mbcnt prefix sums visited lowest lane first.  The
// validation predicate "no site on a cell edge intersects the candidate" is
// evaluated over the distinct sites of the edges, which answers the same
// question as the reference's walk over both endpoints of every edge.
#pragma once

#include "apo_device.h"
#include "apo_internal.h"

namespace apo_hip {

constexpr uint32_t kListCap = 2 * kEdgeCap;

struct CellLds {
    uint4 vtx[kVertexCap];
    uint4 edg[kEdgeCap];
    uint32_t nb[kNeighbourCap];
    uint32_t ne[kNewEdgeCap];
    uint32_t list[kListCap];
};

struct WaveCell {
    uint32_t V, E, F, flags;
};

// The lanes that work on one cell: the whole wave (W = 64) or one half of
// it (W = 32, lanes 32..63 the second half, each half with a cell of its
// own).  Every step of a cell update is written against these: a group's
// ballot holds its own lanes only (bit 0 = its first lane), `below` counts
// the group's set bits under the calling lane, `read` takes a value from a
// lane of the group and `uni` a value that every lane of the group holds
// (a scalar copy for the whole wave; the value itself for a half).
template <uint32_t W>
struct Lanes;
template <>
struct Lanes<64> {
    __device__ static __forceinline__ uint32_t lane() { return laneId(); }
    __device__ static __forceinline__ uint64_t vote(bool p) { return ballot(p); }
    __device__ static __forceinline__ uint32_t below(uint64_t m) { return countBelow(m); }
    __device__ static __forceinline__ uint32_t read(uint32_t v, uint32_t t) { return readLane(v, t); }
    __device__ static __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
    __device__ static __forceinline__ double maxD(double v) { return waveMaxD(v); }
};
template <>
struct Lanes<32> {
    __device__ static __forceinline__ uint32_t half() { return laneId() & 32u; }
    __device__ static __forceinline__ uint32_t lane() { return laneId() & 31u; }
    __device__ static __forceinline__ uint64_t vote(bool p) {
        const uint64_t m = ballot(p);
        return half() ? (m >> 32) : (m & 0xffffffffull);
    }
    __device__ static __forceinline__ uint32_t below(uint64_t m) {
        return uint32_t(__builtin_popcount(uint32_t(m) & ((1u << lane()) - 1u)));
    }
    __device__ static __forceinline__ uint32_t read(uint32_t v, uint32_t t) {
        return uint32_t(__shfl(int(v), int(half() | t), 64));
    }
    __device__ static __forceinline__ uint32_t uni(uint32_t v) { return v; }
    __device__ static __forceinline__ double maxD(double v) {
        for (int o = 16; o >= 1; o >>= 1) v = __builtin_fmax(v, __shfl_xor(v, o, 64));
        return v;
    }
};

__device__ __forceinline__ D4 loadSite(const CellCtx& c, uint32_t id) {
    if (id < c.n) return toD(c.sorted[id]);
    const uint32_t a = id - c.n;
    if (a < 6u * c.molecules) return c.scene[a / 6u].art[a % 6u];
    atomicOr(c.err, kErrCellSite);
    return D4{0, 0, 0, 0};
}

__device__ __forceinline__ D4 vertexPosition(const CellCtx& c, const D4& si, uint4 r) {
    D4 x0{0, 0, 0, 0}, x1{0, 0, 0, 0};
    quadrisector(si, loadSite(c, r.x), loadSite(c, r.y), loadSite(c, r.z), x0, x1);
    return selD4((r.w & 0xffu) != 0u, x1, x0);
}

// Distinct sites over the edges [0, E), in first-occurrence order.
template <uint32_t W = 64>
__device__ inline void buildNeighbours(const CellCtx& c, CellLds& L, WaveCell& w) {
    using G = Lanes<W>;
    const uint32_t lane = G::lane();
    uint32_t F = 0;
    for (uint32_t b = 0; b < w.E; b += W) {
        const uint32_t e = b + lane;
        const bool has = e < w.E;
        uint32_t x0 = kNone, x1 = kNone;
        if (has) {
            const uint4 ed = L.edg[e];
            x0 = ed.x;
            x1 = ed.y;
        }
        bool f0 = has, f1 = has && x1 != x0;
        const uint32_t lim = (w.E - b) < W ? w.E : b + W;
        // earlier edges, eight per step: independent LDS reads in flight
        uint32_t e2 = 0;
        for (; e2 + 8u <= lim; e2 += 8u) {
            uint4 y[8];
#pragma unroll
            for (uint32_t u = 0; u < 8u; u++) y[u] = L.edg[e2 + u];
#pragma unroll
            for (uint32_t u = 0; u < 8u; u++) {
                const bool before = e2 + u < e;
                if (before && (y[u].x == x0 || y[u].y == x0)) f0 = false;
                if (before && (y[u].x == x1 || y[u].y == x1)) f1 = false;
            }
        }
        for (; e2 < lim; e2++) {
            const uint4 y = L.edg[e2];
            if (e2 < e) {
                if (y.x == x0 || y.y == x0) f0 = false;
                if (y.x == x1 || y.y == x1) f1 = false;
            }
        }
        const uint64_t m0 = G::vote(f0), m1 = G::vote(f1);
        const uint32_t off = F + G::below(m0) + G::below(m1);
        if (f0) L.nb[off] = x0;
        if (f1) L.nb[off + (f0 ? 1u : 0u)] = x1;
        F += popc64(m0) + popc64(m1);
    }
    w.F = F;
    waveSync();
}

// Site of neighbour slot q (id = L.nb[q]).
__device__ __forceinline__ D4 nbSite(const CellCtx& c, const CellLds& L, uint32_t q, uint32_t id) {
    (void)L;
    (void)q;
    return loadSite(c, id);
}

// Candidate sphere is clear of every neighbour except skip1/skip2
// (topology_update.inl:578-592).  Per lane, serial over the neighbour set,
// U neighbours per step: their site loads are in flight together (callers
// with many live registers take fewer).
template <uint32_t U = 4, uint32_t W = 64>
__device__ __forceinline__ bool clearOfNeighbours(const CellCtx& c, const CellLds& L, uint32_t F, const D4& cand,
                                                  uint32_t skip1, uint32_t skip2) {
    using G = Lanes<W>;
    uint32_t q = 0;
    if constexpr (U > 1u) {
        for (; q + U <= F; q += U) {
            uint32_t id[U];
            D4 s[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) id[u] = G::uni(L.nb[q + u]);
#pragma unroll
            for (uint32_t u = 0; u < U; u++) s[u] = nbSite(c, L, q + u, id[u]);
            bool hit = false;
#pragma unroll
            for (uint32_t u = 0; u < U; u++) hit |= id[u] != skip1 && id[u] != skip2 && intersectD(s[u], cand);
            if (hit) return false;
        }
    }
    for (; q < F; q++) {
        const uint32_t id = G::uni(L.nb[q]);
        if (id == skip1 || id == skip2) continue;
        if (intersectD(nbSite(c, L, q, id), cand)) return false;
    }
    return true;
}

// Both candidate spheres of an edge against the neighbour set in one pass:
// the bits of `mask` (bit 0: x0, bit 1: x1) whose candidate is clear of
// every neighbour except skip1/skip2.  Equal to two clearOfNeighbours calls;
// each neighbour's site is loaded once for both.
template <uint32_t U = 4, uint32_t W = 64>
__device__ __forceinline__ uint32_t clearPairOfNeighbours(const CellCtx& c, const CellLds& L, uint32_t F,
                                                          const D4& x0, const D4& x1, uint32_t mask,
                                                          uint32_t skip1, uint32_t skip2) {
    using G = Lanes<W>;
    uint32_t ok = mask;
    uint32_t q = 0;
    if constexpr (U > 1u) {
        for (; q + U <= F && ok; q += U) {
            uint32_t id[U];
            D4 s[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) id[u] = G::uni(L.nb[q + u]);
#pragma unroll
            for (uint32_t u = 0; u < U; u++) s[u] = nbSite(c, L, q + u, id[u]);
            bool h0 = false, h1 = false;
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const bool use = id[u] != skip1 && id[u] != skip2;
                h0 |= use && (ok & 1u) && intersectD(s[u], x0);
                h1 |= use && (ok & 2u) && intersectD(s[u], x1);
            }
            if (h0) ok &= ~1u;
            if (h1) ok &= ~2u;
        }
    }
    for (; q < F && ok; q++) {
        const uint32_t id = G::uni(L.nb[q]);
        if (id == skip1 || id == skip2) continue;
        const D4 s = nbSite(c, L, q, id);
        if ((ok & 1u) && intersectD(s, x0)) ok &= ~1u;
        if ((ok & 2u) && intersectD(s, x1)) ok &= ~2u;
    }
    return ok;
}

// Stable in-place removal of records flagged by `drop`, step by step.
template <uint32_t W = 64, typename Drop>
__device__ __forceinline__ uint32_t compact(uint4* rec, uint32_t count, Drop drop) {
    using G = Lanes<W>;
    const uint32_t lane = G::lane();
    uint32_t removed = 0;
    for (uint32_t b = 0; b < count; b += W) {
        const uint32_t idx = b + lane;
        const bool has = idx < count;
        uint4 r = {0, 0, 0, 0};
        if (has) r = rec[idx];
        const bool d = has && drop(idx, r);
        const uint64_t m = G::vote(d);
        waveSync();
        if (has && !d) rec[idx - removed - G::below(m)] = r;
        removed += popc64(m);
        waveSync();
    }
    return count - removed;
}

__device__ __forceinline__ void sort3(uint32_t& a, uint32_t& b, uint32_t& c) {
    uint32_t t;
    if (a > c) { t = a; a = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
}

// One incremental update of cell i by site j (topology_update.inl:49-555).
// Fuse: solve steps (1) and (2) together when they fit the group (the cell
// pass, whose cells are small); the validation kernels, whose cells rarely
// fit, keep the registers.
template <uint32_t W = 64, bool Fuse = false>
__device__ inline bool updateCell(const CellCtx& c, CellLds& L, WaveCell& w, uint32_t i, const D4& si, uint32_t j,
                           const D4& sj) {
    using G = Lanes<W>;
    const uint32_t lane = G::lane();
    if (w.V == 0 && w.E == 0) return false;
    {
        const double ri = si.w, rj = sj.w;
        const double bigger = __builtin_fmax(ri, rj), smaller = __builtin_fmin(ri, rj);
        if (ltD(len(D3{si.x - sj.x, si.y - sj.y, si.z - sj.z}), bigger - smaller)) {
            if (ltD(rj, ri)) return false;
            w.V = 0;
            w.E = 0;
            w.F = 0;
            return false;
        }
    }
    const uint32_t baseV = w.V, baseE = w.E;
    buildNeighbours<W>(c, L, w);

    // Steps (1) and (2) solve one quadrisector per record: each vertex's
    // position (1), each edge's candidate vertices with s_j (2).  When the
    // vertices and edges fit the group together, both run as one solve:
    // edge e on lane e, vertex v on lane W - baseV + v (the same arithmetic
    // per record, one pass instead of two).
    const bool fused = Fuse && baseV + baseE <= W;
    const bool vLane = fused && baseV != 0u && lane >= W - baseV;
    D4 fq0{0, 0, 0, 0}, fq1{0, 0, 0, 0};
    uint32_t fmask = 0;
    if (fused) {
        D4 a{0, 0, 0, 0}, b{0, 0, 0, 0}, d{0, 0, 0, 0};
        bool act = false;
        if (lane < baseE) {
            const uint4 ed = L.edg[lane];
            a = sj;
            b = loadSite(c, ed.x);
            d = loadSite(c, ed.y);
            act = true;
        } else if (vLane) {
            const uint4 r = L.vtx[lane - (W - baseV)];
            a = loadSite(c, r.x);
            b = loadSite(c, r.y);
            d = loadSite(c, r.z);
            act = true;
        }
        if (act) fmask = quadrisector(si, a, b, d, fq0, fq1);
    }

    // (1) vertices whose empty sphere intersects s_j
    uint32_t nInv = 0;
    if (fused) {
        const uint32_t v = lane - (W - baseV);
        bool inv = false;
        if (vLane) {
            const uint4 r = L.vtx[v];
            inv = intersectD(selD4((r.w & 0xffu) != 0u, fq1, fq0), sj);
            if (inv) L.vtx[v].w = (r.w & 0xffu) | (0xffu << 8);
        }
        const uint64_t m = G::vote(inv);
        if (inv) L.list[G::below(m)] = v;
        nInv = popc64(m);
    } else {
        for (uint32_t b = 0; b < baseV; b += W) {
            const uint32_t v = b + lane;
            bool inv = false;
            if (v < baseV) {
                const uint4 r = L.vtx[v];
                inv = intersectD(vertexPosition(c, si, r), sj);
                if (inv) L.vtx[v].w = (r.w & 0xffu) | (0xffu << 8);
            }
            const uint64_t m = G::vote(inv);
            if (inv) L.list[nInv + G::below(m)] = v;
            nInv += popc64(m);
        }
    }
    waveSync();
    if (nInv) {
        for (uint32_t b = 0; b < baseE; b += W) {
            const uint32_t e = b + lane;
            if (e < baseE) {
                const uint4 ed = L.edg[e];
                uint32_t dec = 0;
                for (uint32_t u = 0; u < nInv; u++) {
                    const uint4 vx = L.vtx[L.list[u]];
                    const bool e1 = ed.x == vx.x && ed.y == vx.y;
                    const bool e2 = ed.x == vx.x && ed.y == vx.z;
                    const bool e3 = ed.x == vx.y && ed.y == vx.z;
                    if (e1 || e2 || e3) dec++;
                }
                if (dec) L.edg[e].z = ed.z - dec;
            }
        }
        waveSync();
        w.V = compact<W>(L.vtx, w.V, [&](uint32_t idx, uint4 r) { return ((r.w >> 8) & 0xffu) == 0xffu && idx < baseV; });
    }

    // (2) candidate vertices on every edge (k, l)
    uint32_t NE = 0, nClosed = 0;
    for (uint32_t b = 0; b < baseE; b += W) {
        const uint32_t e = b + lane;
        const bool has = e < baseE;
        uint4 ed = {0, 0, 0, 0};
        if (has) ed = L.edg[e];
        const uint32_t k = ed.x, l = ed.y;
        D4 q0{0, 0, 0, 0}, q1{0, 0, 0, 0};
        uint32_t mask = 0;
        D4 sk{0, 0, 0, 0}, sl{0, 0, 0, 0};
        if (has) {
            sk = loadSite(c, k);
            sl = loadSite(c, l);
            if (fused) {
                q0 = fq0;
                q1 = fq1;
                mask = fmask;
            } else {
                mask = quadrisector(si, sj, sk, sl, q0, q1);
            }
        }
        uint32_t vnb = ed.z;
        if (has && mask == 0 && vnb == kNone) {
            D4 t = q0;
            trisectorAliased(si, sk, sl, t);
            if (intersectD(t, sj)) vnb = 0;
        }
        const uint32_t valid = clearPairOfNeighbours<4, W>(c, L, w.F, q0, q1, mask, k, l);
        const uint32_t nNew = uint32_t(__builtin_popcount(valid));
        if (has && mask != 0 && nNew > 0) {
            if (vnb == kNone) vnb = 0;
            vnb += nNew;
        }
        if (has && vnb != ed.z) L.edg[e].z = vnb;

        const bool closedCand = has && mask == 0;
        const uint64_t mc = G::vote(closedCand);
        if (closedCand) L.list[nClosed + G::below(mc)] = e;
        nClosed += popc64(mc);

        const uint64_t b0 = G::vote(nNew >= 1), b1 = G::vote(nNew >= 2);
        const uint32_t off = G::below(b0) + G::below(b1);
        uint32_t jj = j, kk = k, ll = l;
        sort3(jj, kk, ll);
        uint32_t slot = w.V + off;
        uint32_t eslot = NE + 2 * off;
        const uint32_t two = nNew == 2 ? (1u << 30) : 0u;
        for (uint32_t nv = 0; nv < 2; nv++) {
            if (!((valid >> nv) & 1u)) continue;
            if (slot < kVertexCap) L.vtx[slot] = make_uint4(jj, kk, ll, nv);
            slot++;
            if (eslot < kNewEdgeCap) L.ne[eslot] = k | two;
            eslot++;
            if (eslot < kNewEdgeCap) L.ne[eslot] = l | two;
            eslot++;
        }
        const uint32_t tot = popc64(b0) + popc64(b1);
        w.V += tot;
        NE += 2 * tot;
    }
    if (w.V > kVertexCap) {
        w.V = kVertexCap;
        w.flags |= kFlagVertexOverflow;
    }
    if (NE > kNewEdgeCap) {
        NE = kNewEdgeCap;
        w.flags |= kFlagNewEdgeOverflow;
    }
    waveSync();

    // (3) closed elliptic trisectors (topology_update.inl:325-433).  The
    // candidates (edge order, endpoint order) are tested in parallel; an id
    // already among the new-edge entries, or accepted by a lower candidate,
    // is skipped, which yields the reference's sequential result.
    if (nClosed) {
        const uint32_t cands = 2 * nClosed;
        for (uint32_t qb = 0; qb < cands; qb += W) {
            const uint32_t q = qb + lane;
            const bool act = q < cands;
            uint32_t id = kNone;
            if (act) {
                const uint4 ed = L.edg[L.list[q >> 1]];
                id = (q & 1u) ? ed.y : ed.x;
            }
            bool ok = act;
            for (uint32_t a = 0; a < NE && ok; a++)
                if ((L.ne[a] & 0x3fffffffu) == id) ok = false;
            if (ok) {
                D4 x0{0, 0, 0, 0}, x1{0, 0, 0, 0};
                const uint32_t cnt = trisector(si, sj, loadSite(c, id), x0, x1);
                ok = cnt == 2 && clearOfNeighbours<4, W>(c, L, w.F, x0, id, i);
            }
            // keep the first accepted lane per id
            uint64_t m = G::vote(ok);
            uint64_t mm = m;
            while (mm) {
                const uint32_t t = uint32_t(__builtin_ctzll(mm));
                mm &= mm - 1;
                const uint32_t idt = G::read(id, t);
                if (ok && lane > t && id == idt) ok = false;
            }
            m = G::vote(ok);
            const uint32_t pos = NE + G::below(m);
            if (ok) {
                if (pos < kNewEdgeCap) L.ne[pos] = id | (1u << 31);
            }
            NE += popc64(m);
            if (NE > kNewEdgeCap) {
                NE = kNewEdgeCap;
                w.flags |= kFlagNewEdgeOverflow;
            }
            waveSync();
        }
    }

    // (4) old edges whose vertex count dropped to zero leave the cell
    w.E = compact<W>(L.edg, w.E, [&](uint32_t idx, uint4 r) { return r.z == 0u && idx < baseE; });

    // (5) one new edge (j, k) per distinct k, ascending, vertexNb = run length
    //     (or kNone for a closed trisector)
    // first-occurrence flags go to bit 30 of the entry
    for (uint32_t ab = 0; ab < NE; ab += W) {
        const uint32_t a = ab + lane;
        bool first = a < NE;
        uint32_t key = 0;
        if (first) {
            key = L.ne[a] & 0x3fffffffu;
            for (uint32_t a2 = 0; a2 < a; a2++)
                if ((L.ne[a2] & 0x3fffffffu) == key) {
                    first = false;
                    break;
                }
        }
        waveSync();
        if (a < NE) L.list[a] = first ? 1u : 0u;
        waveSync();
    }
    uint32_t D = 0;
    for (uint32_t ab = 0; ab < NE; ab += W) {
        const uint32_t a = ab + lane;
        const bool first = a < NE && L.list[a];
        if (first) {
            const uint32_t key = L.ne[a] & 0x3fffffffu;
            uint32_t count = 0, rank = 0;
            bool closed = false;
            for (uint32_t a2 = 0; a2 < NE; a2++) {
                const uint32_t en = L.ne[a2];
                const uint32_t k2 = en & 0x3fffffffu;
                if (k2 == key) {
                    count++;
                    closed = (en >> 31) != 0u;
                }
                if (k2 < key && L.list[a2]) rank++;
            }
            const uint32_t slot = w.E + rank;
            if (slot < kEdgeCap)
                L.edg[slot] = make_uint4(j < key ? j : key, j < key ? key : j, closed ? kNone : count, 0u);
        }
        D += popc64(G::vote(first));
    }
    w.E += D;
    if (w.E > kEdgeCap) {
        w.E = kEdgeCap;
        w.flags |= kFlagEdgeOverflow;
    }
    waveSync();
    return D != 0;
}

// cellRadius (topology_update.inl:595-653)
template <uint32_t W = 64>
__device__ inline double cellRadius(const CellCtx& c, const CellLds& L, const WaveCell& w, const D4& si) {
    using G = Lanes<W>;
    const uint32_t lane = G::lane();
    double maxR = 0.0;
    for (uint32_t b = 0; b < w.V; b += W) {
        const uint32_t v = b + lane;
        if (v < w.V) maxR = __builtin_fmax(maxR, vertexPosition(c, si, L.vtx[v]).w);
    }
    for (uint32_t b = 0; b < w.E; b += W) {
        const uint32_t e = b + lane;
        if (e < w.E) {
            const uint4 ed = L.edg[e];
            D4 x0{0, 0, 0, 0}, x1{0, 0, 0, 0};
            if (trisector(si, loadSite(c, ed.x), loadSite(c, ed.y), x0, x1) == 2)
                maxR = __builtin_fmax(maxR, __builtin_fmax(x0.w, x1.w));
        }
    }
    return G::maxD(maxR);
}

// The initial cell: octahedron of the 6 artificial sites (algorithm.inl:1472-1499)
// of the cell's molecule, whose ids start at n.
template <uint32_t W = 64>
__device__ __forceinline__ void initCell(CellLds& L, WaveCell& w, uint32_t n) {
    using G = Lanes<W>;
    const uint32_t lane = G::lane();
    const uint8_t av[8][3] = {{0, 2, 4}, {0, 2, 5}, {0, 3, 4}, {0, 3, 5}, {1, 2, 4}, {1, 2, 5}, {1, 3, 4}, {1, 3, 5}};
    const uint8_t ae[12][2] = {{0, 2}, {2, 4}, {1, 2}, {2, 5}, {1, 4}, {1, 5},
                               {0, 5}, {0, 4}, {1, 3}, {3, 5}, {0, 3}, {3, 4}};
    if (lane < 8) L.vtx[lane] = make_uint4(n + av[lane][0], n + av[lane][1], n + av[lane][2], 0u);
    if (lane < 12) L.edg[lane] = make_uint4(n + ae[lane][0], n + ae[lane][1], 2u, 0u);
    w.V = 8;
    w.E = 12;
    w.F = 0;
    w.flags = 0;
    waveSync();
}

// j is an endpoint of a current edge (wave-wide; algorithm.inl:368-376)
template <uint32_t W = 64>
__device__ __forceinline__ bool knownSite(const CellLds& L, const WaveCell& w, uint32_t j) {
    using G = Lanes<W>;
    const uint32_t lane = G::lane();
    bool hit = false;
    for (uint32_t b = 0; b < w.E; b += W) {
        const uint32_t e = b + lane;
        if (e < w.E) {
            const uint4 ed = L.edg[e];
            hit = hit || ed.x == j || ed.y == j;
        }
    }
    return G::vote(hit) != 0;
}

// The stored records of cell i (its vertices too when withVertices).
template <uint32_t W = 64>
__device__ __forceinline__ void loadCell(const CellCtx& c, CellLds& L, WaveCell& w, uint32_t i, bool withVertices) {
    using G = Lanes<W>;
    const uint32_t lane = G::lane();
    const uint32_t E = c.enb[i];
    const uint32_t V = withVertices ? c.vnb[i] : 0u;
    const uint4* ge = c.edg + size_t(i) * kEdgeCap;
    const uint4* gv = c.vtx + size_t(i) * kVertexCap;
    for (uint32_t e = lane; e < E; e += W) L.edg[e] = ge[e];
    for (uint32_t v = lane; v < V; v += W) L.vtx[v] = gv[v];
    w.V = V;
    w.E = E;
    w.F = 0;
    w.flags = 0;
    waveSync();
}

template <uint32_t W = 64>
__device__ __forceinline__ void storeCell(const CellCtx& c, const CellLds& L, const WaveCell& w, uint32_t i) {
    using G = Lanes<W>;
    const uint32_t lane = G::lane();
    uint4* gv = c.vtx + size_t(i) * kVertexCap;
    uint4* ge = c.edg + size_t(i) * kEdgeCap;
    for (uint32_t v = lane; v < w.V; v += W) gv[v] = L.vtx[v];
    for (uint32_t e = lane; e < w.E; e += W) ge[e] = L.edg[e];
    if (lane == 0) {
        c.vnb[i] = w.V;
        c.enb[i] = w.E;
        if (w.flags) c.flags[i] |= w.flags;
    }
}

}  // namespace apo_hip
