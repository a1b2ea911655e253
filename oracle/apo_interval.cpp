// apo_interval.cpp -- CPU restatement of the reference's interval precision
// validator (app/precision/precision.cpp:14-163).  TEST INFRASTRUCTURE ONLY
// (part of the oracle library; never linked into the product).
//
// The reference uses boost::numeric::interval<double> with its default
// policies: every bound is computed with the FPU rounding mode switched to
// -inf or +inf.  Boost is not installed here, so this file restates those
// semantics directly: each operation runs under fesetround(FE_DOWNWARD) for
// the lower bound and fesetround(FE_UPWARD) for the upper bound (compiled
// with -frounding-math so the compiler keeps every operation where it is).
// The HIP kernel reaches the same bounds another way (error-free
// transformations), so agreement between the two is a check of both.
//
//   weightedDistance     precision.cpp:14-24   sqrt(sum (s_i - p_i)^2) - r
//   getBadQuadrisectorNb precision.cpp:26-100
//   getBadTrisectorNb    precision.cpp:102-161
// A point is "bad" as in the reference (another site's interval is
// certainly below the largest interval of its own sites).  "Uncertain"
// (not bad, but another site's interval reaches below that largest one)
// is this build's near-degenerate count.
#include <cfenv>
#include <cfloat>
#include <cmath>
#include <cstdint>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "apo_oracle.h"

namespace {

struct Iv {
    double lo, hi;
};

// one operation under an explicit rounding mode; volatile keeps the
// operation inside the mode switch
double addRounded(double a, double b, int mode) {
    std::fesetround(mode);
    volatile double x = a, y = b;
    volatile double r = x + y;
    std::fesetround(FE_TONEAREST);
    return r;
}
double subRounded(double a, double b, int mode) {
    std::fesetround(mode);
    volatile double x = a, y = b;
    volatile double r = x - y;
    std::fesetround(FE_TONEAREST);
    return r;
}
double mulRounded(double a, double b, int mode) {
    std::fesetround(mode);
    volatile double x = a, y = b;
    volatile double r = x * y;
    std::fesetround(FE_TONEAREST);
    return r;
}
double sqrtRounded(double a, int mode) {
    std::fesetround(mode);
    volatile double x = a;
    volatile double r = std::sqrt(x);
    std::fesetround(FE_TONEAREST);
    return r;
}

// Boost's interval product: the lowest and highest of the four endpoint
// products, each rounded outward
Iv mul(Iv a, Iv b) {
    const double c[4][2] = {{a.lo, b.lo}, {a.lo, b.hi}, {a.hi, b.lo}, {a.hi, b.hi}};
    Iv r{DBL_MAX, -DBL_MAX};
    for (const auto& p : c) {
        r.lo = std::fmin(r.lo, mulRounded(p[0], p[1], FE_DOWNWARD));
        r.hi = std::fmax(r.hi, mulRounded(p[0], p[1], FE_UPWARD));
    }
    return r;
}

// weightedDistance (precision.cpp:14-24) with point intervals for the site
// and the point
Iv distance(const double s[4], const double p[3]) {
    Iv res{0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        const Iv diff{subRounded(s[i], p[i], FE_DOWNWARD), subRounded(s[i], p[i], FE_UPWARD)};
        const Iv sq = mul(diff, diff);
        res = {addRounded(res.lo, sq.lo, FE_DOWNWARD), addRounded(res.hi, sq.hi, FE_UPWARD)};
    }
    const Iv root{res.lo <= 0.0 ? 0.0 : sqrtRounded(res.lo, FE_DOWNWARD), sqrtRounded(res.hi, FE_UPWARD)};
    return {subRounded(root.lo, s[3], FE_DOWNWARD), subRounded(root.hi, s[3], FE_UPWARD)};
}

}  // namespace

extern "C" {

void orc_round_dir(uint32_t op, double a, double b, double* lohi) {
    switch (op) {
        case 0:
            lohi[0] = addRounded(a, b, FE_DOWNWARD);
            lohi[1] = addRounded(a, b, FE_UPWARD);
            break;
        case 1:
            lohi[0] = mulRounded(a, b, FE_DOWNWARD);
            lohi[1] = mulRounded(a, b, FE_UPWARD);
            break;
        default:
            lohi[0] = sqrtRounded(a, FE_DOWNWARD);
            lohi[1] = sqrtRounded(a, FE_UPWARD);
            break;
    }
}

void orc_interval_distance(const double* site4, const double* p3, double* lohi) {
    const Iv d = distance(site4, p3);
    lohi[0] = d.lo;
    lohi[1] = d.hi;
}

// getBadQuadrisectorNb / getBadTrisectorNb by a scan of every site.  xyzw:
// count points (w unused), ids: count x 4 original ids (arity 3: ids[3]
// unused).  out2[0] = bad, out2[1] = uncertain.
void orc_precision_counts(const float* xyzr, uint32_t n, const double* xyzw, const uint32_t* ids, uint64_t count,
                          uint32_t arity, uint32_t threads, uint64_t* out2) {
#ifdef _OPENMP
    if (threads) omp_set_num_threads(int(threads));
#else
    (void)threads;
#endif
    uint64_t bad = 0, uncertain = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad, uncertain)
    for (int64_t q = 0; q < int64_t(count); q++) {
        const double p[3] = {xyzw[4 * q], xyzw[4 * q + 1], xyzw[4 * q + 2]};
        uint32_t own[4] = {ids[4 * q], ids[4 * q + 1], ids[4 * q + 2], ids[4 * q + 3]};
        if (arity == 3) own[3] = own[2];
        auto site = [&](uint32_t s, double out[4]) {
            for (int c = 0; c < 4; c++) out[c] = double(xyzr[4 * size_t(s) + c]);
        };
        Iv qmax{-DBL_MAX, -DBL_MAX};
        for (uint32_t t = 0; t < arity; t++) {
            double s[4];
            site(own[t], s);
            const Iv d = distance(s, p);
            qmax = {std::fmax(qmax.lo, d.lo), std::fmax(qmax.hi, d.hi)};
        }
        // A site whose round-to-nearest distance exceeds qmax.hi by far more
        // than its rounding error cannot reach below qmax.hi with intervals
        // either: only the others are evaluated with switched rounding.
        const double scale = std::fabs(p[0]) + std::fabs(p[1]) + std::fabs(p[2]) + std::fabs(qmax.hi) + 1.0;
        bool isBad = false, isUncertain = false;
        for (uint32_t s = 0; s < n && !isBad; s++) {
            if (s == own[0] || s == own[1] || s == own[2] || s == own[3]) continue;
            double si[4];
            site(s, si);
            const double dx = si[0] - p[0], dy = si[1] - p[1], dz = si[2] - p[2];
            const double near = std::sqrt(dx * dx + dy * dy + dz * dz) - si[3];
            if (near > qmax.hi + 1e-9 * (scale + std::fabs(si[0]) + std::fabs(si[1]) + std::fabs(si[2]) + si[3]))
                continue;
            const Iv d = distance(si, p);
            if (d.hi < qmax.lo) isBad = true;              // compare::certain <
            else if (d.lo < qmax.hi) isUncertain = true;   // possibly below
        }
        if (isBad) bad++;
        else if (isUncertain) uncertain++;
    }
    out2[0] = bad;
    out2[1] = uncertain;
}

}  // extern "C"
