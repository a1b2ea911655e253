// edge_bound.h -- cover of the empty spheres of one cell edge.
//
// No counterpart in the reference.  The edge contributor search of
// apo/include/apo/gpu/algorithm.inl:572-793 bounds a closed (elliptic)
// trisector by the sphere of its two extreme points and a hyperbolic one by
// tangent planes; next to a large site that sphere can span the whole scene,
// and the search then tests every leaf.  The cover built here prunes the
// same traversal (same visiting order, fewer subtrees) without changing any
// search result: a subtree is skipped only when no site in it can touch an
// empty sphere of the edge.
//
// The centres of the spheres tangent to s1, s2, s3 with radius w are
// c(w) = p1 + u(w) +- h(w) n, where n is the unit normal of the centres'
// plane, u(w) = P + Q w lies in that plane and h(w)^2 = A w^2 + B w + Cc.
// A neighbour s becomes tangent where
//     |c(w) - p_s| = r_s + w  <=>  L0 + L1 w = +-2 sigma h(w),
// sigma = n.(p_s - p1); squared, a quadratic in w whose roots are all the
// points where the gap of s changes sign.  Between consecutive roots (of
// all neighbours, plus the extremum of h^2) every gap keeps its sign, so one
// midpoint test per interval and sign of h tells whether the interval can
// hold an accepted sphere.  Runs of such intervals are covered by balls: on
// [wa, wb] (h monotone) every sphere lies in the ball centred at c(wm),
// wm = (wa + wb) / 2, of radius |Q| (wb - wa) / 2 + max |h(w) - h(wm)| + wb.
// A contributor j touches an accepted sphere, so |p_j - centre| - r_j is at
// most that radius for one ball.
#pragma once

#include "apo_device.h"
#include "apo_internal.h"
#include "cell_update.h"

namespace apo_hip {

constexpr int kCoverBalls = 8;       // balls per cover
constexpr uint32_t kCoverRoots = 512;
constexpr int kCoverRuns = kCoverBalls + 1;

// Buffers of the cover construction in the wave's device region (covers
// with more than kCoverLdsRoots roots).
struct CoverLds {
    double roots[kCoverRoots];   // tangency roots, then the keep flags (bytes)
    double sorted[kCoverRoots];  // roots in ascending order
};
// Where a cover's roots live: in LDS (the cell's update scratch list, idle
// between the neighbour rebuild and the searches) when 2F + 1 of them fit,
// else in the wave's device region.  The capacity only decides where they
// live: a cover fails when it exceeds kCoverRoots either way.
constexpr uint32_t kCoverLdsRoots = uint32_t(sizeof(uint32_t) * kListCap / (2 * sizeof(double)));
struct CoverBuf {
    double* roots;
    double* sorted;
};
struct CoverOut {
    double runA[kCoverRuns], runB[kCoverRuns], runS[kCoverRuns];
    F4 ball[kCoverBalls];
};

struct TriCurve {
    D3 P, Q, n;  // P, Q relative to p1
    double A, B, Cc;
};

__device__ __forceinline__ bool triCurve(const D4& s1, const D4& s2, const D4& s3, TriCurve& t) {
    const D3 q2 = sub(xyz(s2), xyz(s1)), q3 = sub(xyz(s3), xyz(s1));
    const D3 nc = cross(q2, q3);
    const double nl = len(nc);
    if (!(nl > 1e-12 * len(q2) * len(q3))) return false;  // collinear centres
    t.n = mul(nc, 1.0 / nl);
    // rows: -q2 . u = b0x + b1x w, -q3 . u = b0y + b1y w, n . u = 0
    const D3 r1{-q2.x, -q2.y, -q2.z}, r2{-q3.x, -q3.y, -q3.z}, r3 = t.n;
    const double b0x = .5 * (-dot(q2, q2) - s1.w * s1.w + s2.w * s2.w);
    const double b0y = .5 * (-dot(q3, q3) - s1.w * s1.w + s3.w * s3.w);
    const double b1x = -(s1.w - s2.w), b1y = -(s1.w - s3.w);
    const D3 c23 = cross(r2, r3), c31 = cross(r3, r1);
    const double det = dot(r1, c23);
    if (!(__builtin_fabs(det) > 0.0)) return false;
    const double inv = 1.0 / det;
    t.P = mul(add(mul(c23, b0x), mul(c31, b0y)), inv);
    t.Q = mul(add(mul(c23, b1x), mul(c31, b1y)), inv);
    t.A = 1.0 - dot(t.Q, t.Q);
    t.B = 2.0 * (s1.w - dot(t.P, t.Q));
    t.Cc = s1.w * s1.w - dot(t.P, t.P);
    return __builtin_isfinite(t.A) && __builtin_isfinite(t.B) && __builtin_isfinite(t.Cc) &&
           __builtin_isfinite(dot(t.P, t.P) + dot(t.Q, t.Q));
}

// Real roots of a w^2 + b w + c; a slightly negative discriminant (rounding)
// counts as a double root.  Returns -1 when every w is a root.
__device__ __forceinline__ int realRoots(double a, double b, double c, double& r0, double& r1) {
    const double sa = __builtin_fabs(a), sb = __builtin_fabs(b), sc = __builtin_fabs(c);
    const double scale = fmax(sa, fmax(sb, sc));
    if (!(scale > 0.0)) return -1;
    if (sa <= 1e-13 * scale) {
        if (sb <= 1e-13 * scale) return -1;
        r0 = -c / b;
        r1 = r0;
        return 1;
    }
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) {
        if (disc < -1e-10 * (b * b + __builtin_fabs(4.0 * a * c))) return 0;
        disc = 0.0;
    }
    const double q = -.5 * (b + __builtin_copysign(__builtin_sqrt(disc), b));
    r0 = q / a;
    r1 = (q != 0.0) ? c / q : r0;
    return 2;
}

__device__ __forceinline__ D4 curvePoint(const TriCurve& t, const D3& p1, double w, double sgn) {
    const double h2 = t.A * w * w + t.B * w + t.Cc;
    const double h = __builtin_sqrt(fmax(h2, 0.0));
    const D3 c = add(p1, add(add(t.P, mul(t.Q, w)), mul(t.n, sgn * h)));
    return {c.x, c.y, c.z, w};
}

__device__ __forceinline__ double curveH(const TriCurve& t, double w) {
    return __builtin_sqrt(fmax(t.A * w * w + t.B * w + t.Cc, 0.0));
}

// Gap below which a neighbour clearly hits the sphere at x: more than
// margin + tol inside it.
__device__ __forceinline__ double clearLimit(const D4& x, double margin, double tolBase) {
    return -(margin + 1e-9 + 1e-7 * (__builtin_fabs(x.w) + tolBase));
}

// Some neighbour other than k, l has a gap below lim to the sphere at x, a
// lane's own point and limit: the midpoint test of one interval, one
// interval per lane.  Each lane walks the neighbour set.  With F <= 64
// neighbour q's site is lane q's preloaded `site`, read across the wave (a
// uniform value: one scalar copy per neighbour for all lanes); beyond that
// the sites are loaded.  The walk ends when every lane has found a hit.
__device__ __forceinline__ bool laneClearlyHit(const CellCtx& c, const CellLds& L, uint32_t F, const D4& x,
                                               double lim, uint32_t k, uint32_t l, const D4& site) {
    bool hit = false;
    if (F <= 64u) {
        for (uint32_t q = 0; q < F; q++) {
            const uint32_t s = __builtin_amdgcn_readfirstlane(L.nb[q]);
            if (s == k || s == l) continue;
            const D4 sq = readLaneD4(site, q);
            hit = hit || sphereDistD(sq, x) < lim;
            if (ballot(!hit) == 0) break;
        }
        return hit;
    }
    for (uint32_t q = 0; q < F; q++) {
        const uint32_t s = __builtin_amdgcn_readfirstlane(L.nb[q]);
        if (s == k || s == l) continue;
        hit = hit || sphereDistD(nbSite(c, L, q, s), x) < lim;
        if (ballot(!hit) == 0) break;
    }
    return hit;
}

// Wave-cooperative cover of edge (k,l) of the cell in L; si, k, l are
// uniform.  Returns the number of balls left in O.ball (uniform), or 0 when
// no cover can be certified (the search stays unpruned).
__device__ __forceinline__ int waveEdgeCover(const CellCtx& c, const CellLds& L, CoverBuf B, CoverOut& O,
                                             uint32_t F, const D4& si, uint32_t k, uint32_t l) {
    const uint32_t lane = laneId();
    TriCurve t;
    if (!triCurve(si, loadSite(c, k), loadSite(c, l), t)) return 0;
    const D3 p1 = xyz(si);
    const double r1 = si.w;
    const double tolBase = r1 + len(t.P);
    auto point = [&](double w, double sgn) -> D4 {
        const D3 cc = add(p1, add(add(t.P, mul(t.Q, w)), mul(t.n, sgn * curveH(t, w))));
        return {cc.x, cc.y, cc.z, w};
    };
    // pieces of the w axis where the trisector is real; accepted spheres
    // have w = |p1 - c| - r1 >= -r1
    double tr0 = 0.0, tr1 = 0.0;
    const int nt = realRoots(t.A, t.B, t.Cc, tr0, tr1);
    if (nt < 0 || nt == 1) return 0;
    if (nt == 2 && !(__builtin_isfinite(tr0) && __builtin_isfinite(tr1))) return 0;
    if (nt == 2 && tr0 > tr1) {
        const double x = tr0;
        tr0 = tr1;
        tr1 = x;
    }
    const double inf = __builtin_inf();
    double pa0 = 0.0, pb0 = 0.0, pa1 = 0.0, pb1 = 0.0;
    int np = 0;
    if (t.A < 0.0) {
        if (nt != 2) return 0;
        pa0 = fmax(tr0, -r1);
        pb0 = tr1;
        if (!(pb0 > pa0)) return 0;
        np = 1;
    } else if (nt == 2) {
        if (tr0 > -r1) {
            pa0 = -r1;
            pb0 = tr0;
            pa1 = fmax(tr1, -r1);
            pb1 = inf;
            np = 2;
        } else {
            pa0 = fmax(tr1, -r1);
            pb0 = inf;
            np = 1;
        }
    } else {
        pa0 = -r1;
        pb0 = inf;
        np = 1;
    }

    // tangency roots of every neighbour, then the extremum of h^2
    uint32_t total = 0;
    bool fail = false;
    for (uint32_t b = 0; b < F; b += 64) {
        const uint32_t q = b + lane;
        uint32_t cnt = 0;
        double ra = 0.0, rb = 0.0;
        if (q < F) {
            const uint32_t s = L.nb[q];
            if (s != k && s != l) {
                const D4 ss = nbSite(c, L, q, s);
                const D3 qs = sub(xyz(ss), p1);
                const double sigma = dot(t.n, qs);
                const double L0 = dot(qs, qs) - 2.0 * dot(t.P, qs) + r1 * r1 - ss.w * ss.w;
                const double L1 = -2.0 * dot(t.Q, qs) + 2.0 * (r1 - ss.w);
                const double s4 = 4.0 * sigma * sigma;
                const int nr = realRoots(L1 * L1 - s4 * t.A, 2.0 * L0 * L1 - s4 * t.B, L0 * L0 - s4 * t.Cc, ra, rb);
                if (nr < 0 || !__builtin_isfinite(ra) || !__builtin_isfinite(rb)) fail = true;
                else cnt = uint32_t(nr);
            }
        }
        const uint32_t incl = waveInclusiveSum(cnt);
        const uint32_t off = total + incl - cnt;
        if (cnt >= 1u && off < kCoverRoots) B.roots[off] = ra;
        if (cnt >= 2u && off + 1u < kCoverRoots) B.roots[off + 1u] = rb;
        total += readLane(incl, 63);
    }
    const double wv = -t.B / (2.0 * t.A);
    if (lane == 0 && total < kCoverRoots) B.roots[total] = wv;
    total += 1u;
    if (ballot(fail || !__builtin_isfinite(wv)) != 0 || total > kCoverRoots) return 0;
    waveSync();
    // rank sort (ties by position)
    for (uint32_t b = 0; b < total; b += 64) {
        const uint32_t q = b + lane;
        if (q < total) {
            const double x = B.roots[q];
            uint32_t rank = 0;
            for (uint32_t r = 0; r < total; r++) {
                const double y = B.roots[r];
                rank += (y < x || (y == x && r < q)) ? 1u : 0u;
            }
            B.sorted[rank] = x;
        }
    }
    waveSync();

    // intervals of each piece: keep flags, then runs of kept intervals.  A
    // lane evaluates one interval's test point; the wave then tests each
    // point against all neighbours at once.
    uint8_t* keep = reinterpret_cast<uint8_t*>(B.roots);
    const double qn = len(t.Q);
    D4 nbS{0, 0, 0, 0};
    if (F <= 64u && lane < F) {
        nbS = nbSite(c, L, lane, L.nb[lane]);
    }
    uint32_t nStart = 0, nEnd = 0;
    double wRun = -DBL_MAX;
    for (int p = 0; p < np; p++) {
        const double pa = p == 0 ? pa0 : pa1, pb = p == 0 ? pb0 : pb1;
        uint32_t lo = 0, hi = 0;
        for (uint32_t b = 0; b < total; b += 64) {
            const uint32_t q = b + lane;
            const bool has = q < total;
            const double x = has ? B.sorted[q] : 0.0;
            lo += popc64(ballot(has && x <= pa));
            hi += popc64(ballot(has && x < pb));
        }
        if (hi < lo) hi = lo;
        const uint32_t M = hi - lo + 1u;  // intervals of this piece
        auto bp = [&](uint32_t m) { return m == 0 ? pa : (m == M ? pb : B.sorted[lo + m - 1u]); };
        for (uint32_t b = 0; b < 2u * M; b += 64) {
            const uint32_t e = b + lane;
            const bool act = e < 2u * M;
            D4 x{0, 0, 0, 0};
            double lim = 0.0;
            bool unbounded = false;
            if (act) {
                const uint32_t h = e >= M ? 1u : 0u, m = e - h * M;
                const double sgn = h ? -1.0 : 1.0;
                const double u = bp(m), v = bp(m + 1u);
                if (v == inf) {
                    // must be clearly hit: the run may reach infinity
                    const double ws = u + .5 * fmax(u + r1, 0.0) + 1e-6 * (1.0 + r1);
                    x = point(ws, sgn);
                    lim = clearLimit(x, 0.0, tolBase);
                    unbounded = true;
                } else {
                    // a narrow interval must be cleared for all its points: the
                    // spheres on [u, v] move by at most the displacement bound
                    const double wm = .5 * (u + v), hm = curveH(t, wm);
                    double margin = 0.0;
                    if (v - u <= 1e-6 * (1.0 + __builtin_fabs(u) + r1))
                        margin = qn * (v - u) * .5 +
                                 fmax(__builtin_fabs(curveH(t, u) - hm), __builtin_fabs(curveH(t, v) - hm)) +
                                 (v - u) * .5;
                    x = point(wm, sgn);
                    lim = clearLimit(x, margin, tolBase);
                }
            }
            bool kp = false;
            if (ballot(act) != 0) {
                const bool hit = act && laneClearlyHit(c, L, F, x, lim, k, l, nbS);
                if (act) {
                    if (unbounded) {
                        if (!hit) fail = true;
                    } else {
                        kp = !hit;
                    }
                }
            }
            if (act) keep[e] = kp ? 1 : 0;
        }
        if (ballot(fail) != 0) return 0;
        waveSync();
        for (uint32_t b = 0; b < 2u * M; b += 64) {
            const uint32_t e = b + lane;
            bool st = false, en = false;
            double ua = 0.0, vb = 0.0, sgn = 1.0;
            if (e < 2u * M) {
                const uint32_t h = e >= M ? 1u : 0u, m = e - h * M;
                sgn = h ? -1.0 : 1.0;
                const bool kp = keep[e] != 0;
                ua = bp(m);
                vb = bp(m + 1u);
                const bool joinPrev = m > 0 && keep[e - 1u] != 0 && ua != wv;
                const bool joinNext = m + 1u < M && keep[e + 1u] != 0 && vb != wv;
                st = kp && !joinPrev;
                en = kp && !joinNext;
            }
            const uint64_t ms = ballot(st), me = ballot(en);
            if (st) {
                const uint32_t idx = nStart + countBelow(ms);
                if (idx < uint32_t(kCoverRuns)) {
                    O.runA[idx] = ua;
                    O.runS[idx] = sgn;
                }
            }
            if (en) {
                const uint32_t idx = nEnd + countBelow(me);
                if (idx < uint32_t(kCoverRuns)) O.runB[idx] = vb;
                wRun = fmax(wRun, vb);
            }
            nStart += popc64(ms);
            nEnd += popc64(me);
        }
        waveSync();
    }
    const double W = waveMaxD(wRun);
    if (nStart == 0 || nStart != nEnd) return 0;
    if (nStart > uint32_t(kCoverBalls)) {
        // one ball around s_i reaching the largest accepted radius
        const double m1 = fmax(__builtin_fabs(si.x), fmax(__builtin_fabs(si.y), __builtin_fabs(si.z)));
        if (lane == 0)
            O.ball[0] = {float(si.x), float(si.y), float(si.z),
                         float((r1 + 2.0 * W) * (1.0 + 1e-4) + 1e-6 * (1.0 + m1))};
        waveSync();
        return 1;
    }
    const uint32_t split = uint32_t(kCoverBalls) / nStart, nb = nStart * split;
    if (lane < nb) {
        const uint32_t run = lane / split, m = lane - run * split;
        const double ra = O.runA[run], rb = O.runB[run], sgn = O.runS[run];
        const double wa = ra + (rb - ra) * double(m) / double(split);
        const double wb = (m + 1u == split) ? rb : ra + (rb - ra) * double(m + 1u) / double(split);
        const double wm = .5 * (wa + wb);
        const D4 cm = point(wm, sgn);
        const double hm = curveH(t, wm);
        const double rad =
            qn * (wb - wa) * .5 + fmax(__builtin_fabs(curveH(t, wa) - hm), __builtin_fabs(curveH(t, wb) - hm)) + wb;
        const double mc = fmax(__builtin_fabs(cm.x), fmax(__builtin_fabs(cm.y), __builtin_fabs(cm.z)));
        O.ball[lane] = {float(cm.x), float(cm.y), float(cm.z),
                        float(rad * (1.0 + 1e-4) + 1e-6 * (1.0 + mc + __builtin_fabs(rad)))};
    }
    waveSync();
    return int(nb);
}

}  // namespace apo_hip
