// apo_oracle.cpp -- CPU restatement of apo's construction (test infrastructure).
//
// TEST INFRASTRUCTURE ONLY: this file is the parity oracle and the reported
// CPU baseline.  It is never linked into the product library.
//
// Numerics: IEEE FP32/FP64, round-to-nearest, no FMA contraction (built with
// -ffp-contract=off).  Every expression keeps the operand order of the
// reference expression it restates, so the HIP kernels (which keep the same
// order) produce the same bits.
//
// Deviations from the reference, all deliberate and shared with the HIP path:
//  * warp order: steps are 64 lanes wide and ballot masks are visited
//    lowest lane first (the reference visits 32-lane steps highest lane first,
//    apo/include/apo/gpu/utils.cuh:17-32);
//  * normalize() uses v * (1/sqrt(v.v)) instead of CUDA's rsqrt
//    (apo/include/apo/gpu/math.cuh:311-315);
//  * temporaries the reference leaves uninitialised (quadrisector outputs on
//    an early return) start at zero;
//  * capacities are checked on the final list sizes and overflowing cells are
//    flagged and truncated instead of corrupting neighbouring cells;
//  * edge and bisector validation restarts are capped like vertex
//    validation's (the reference has no cap there, algorithm.inl:835,1211).

#include "apo_oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Small vector types
// ---------------------------------------------------------------------------
struct F3 { float x, y, z; };
struct F4 { float x, y, z, w; };
struct D2 { double x, y; };
struct D3 { double x, y, z; };
struct D4 { double x, y, z, w; };

constexpr uint32_t kNone     = 0xffffffffu;
constexpr uint8_t  kBuried   = 0xff;
constexpr uint8_t  kCertified = 1;

// Flags per cell (shared meaning with the HIP library).
constexpr uint32_t kFlagVertexOverflow   = 1u;
constexpr uint32_t kFlagEdgeOverflow     = 2u;
constexpr uint32_t kFlagNewEdgeOverflow  = 4u;
constexpr uint32_t kFlagBisectorOverflow = 8u;
constexpr uint32_t kFlagRestartCap       = 16u;
constexpr uint32_t kFlagStackOverflow    = 32u;

constexpr int kWave = 64;

// eps predicates, apo/include/apo/core/math.inl:245-276 (eps = 1e-8)
inline bool ltD(double a, double b) { return (a - b) < 1e-8; }
inline bool gtD(double a, double b) { return (a - b) > -1e-8; }
inline bool ltF(float a, float b) { return (a - b) < 1e-8f; }
inline double signD(double a) { return ltD(a, 0.0) ? -1.0 : 1.0; }
inline float signF(float a) { return ltF(a, 0.f) ? -1.f : 1.f; }
inline bool isCloseToZeroD(double v) { return std::fabs(v) < 1e-8; }
// apo::isClose defaults atol 1e-5, rtol 1e-8 (core/math.hpp:65)
inline bool isCloseD(double a, double ref) { return std::fabs(a - ref) <= (1e-5 + 1e-8 * std::fabs(ref)); }

inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(D3 v) { return std::sqrt(dot(v, v)); }
inline D3 normalize(D3 v) {
    const double inv = 1.0 / std::sqrt(dot(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}
inline D3 xyz(D4 a) { return {a.x, a.y, a.z}; }

inline F3 subf(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 addf(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 mulf(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float lenf(F3 v) { return std::sqrt(dotf(v, v)); }
inline F3 normalizef(F3 v) {
    const float inv = 1.0f / std::sqrt(dotf(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}
inline F3 xyzf(F4 a) { return {a.x, a.y, a.z}; }
inline F4 toF(D4 a) { return {float(a.x), float(a.y), float(a.z), float(a.w)}; }
inline F3 toF3(D3 a) { return {float(a.x), float(a.y), float(a.z)}; }

// sphereDistance, math.cuh:368-378
inline float sphereDistF(F4 a, F4 b) { return lenf(subf(xyzf(a), xyzf(b))) - (a.w + b.w); }
inline double sphereDistD(D4 a, D4 b) {
    const D3 d{a.x - b.x, a.y - b.y, a.z - b.z};
    return len(d) - (a.w + b.w);
}
// intersect, math.cuh:380-388
inline bool intersectF(F4 a, F4 b) { return ltF(sphereDistF(a, b), 0.f); }
inline bool intersectD(D4 a, D4 b) { return ltD(sphereDistD(a, b), 0.0); }

// ---------------------------------------------------------------------------
// Aabbf (math.cuh:339-366, math.inl:5-101), FP32
// ---------------------------------------------------------------------------
struct Box { F3 mn, mx; };

// The bisector search's conservative subtree test (validation.hip,
// boxNearAxis): can the box hold the centre of a site whose distance to the
// axis through o along u, plus its radius (>= 0), is below cylR?
inline bool boxNearAxis(const Box& b, F3 o, F3 u, float cylR) {
    const F3 c{(b.mn.x + b.mx.x) * .5f, (b.mn.y + b.mx.y) * .5f, (b.mn.z + b.mx.z) * .5f};
    const F3 hv{(b.mx.x - b.mn.x) * .5f, (b.mx.y - b.mn.y) * .5f, (b.mx.z - b.mn.z) * .5f};
    const float h = lenf(hv);
    const F3 w = subf(c, o);
    const F3 x{w.y * u.z - w.z * u.y, w.z * u.x - w.x * u.z, w.x * u.y - w.y * u.x};
    const float lmax = lenf(w) + h;
    return lenf(x) - h < cylR * 1.000001f + 2e-3f * lmax;
}

inline F3 closestPoint(const Box& b, F4 s) {
    return {std::fmin(std::fmax(s.x, b.mn.x), b.mx.x), std::fmin(std::fmax(s.y, b.mn.y), b.mx.y),
            std::fmin(std::fmax(s.z, b.mn.z), b.mx.z)};
}
inline float boxDistance(const Box& b, F4 s) { return lenf(subf(xyzf(s), closestPoint(b, s))) - s.w; }
inline float boxFarthestDistance(const Box& b, F4 s) {
    const F3 centroid{(b.mn.x + b.mx.x) * .5f, (b.mn.y + b.mx.y) * .5f, (b.mn.z + b.mx.z) * .5f};
    const F3 radius{std::fabs(b.mn.x - b.mx.x) * .5f, std::fabs(b.mn.y - b.mx.y) * .5f,
                    std::fabs(b.mn.z - b.mx.z) * .5f};
    const F3 c = subf(xyzf(s), centroid);
    const F3 fp{radius.x * signF(-c.x), radius.y * signF(-c.y), radius.z * signF(-c.z)};
    return lenf(subf(c, fp)) - s.w;
}

// ---------------------------------------------------------------------------
// FP64 geometry, math.cuh:390-780
// ---------------------------------------------------------------------------
// Inverse of the 3x3 matrix whose columns are c1,c2,c3 (cofactors / det);
// outputs the columns of the inverse.
inline void inverse3(D3 c1, D3 c2, D3 c3, D3& i1, D3& i2, D3& i3) {
    const double invD = 1.0 / dot(c1, cross(c2, c3));
    const D3 r1 = cross(c2, c3);
    const D3 r2 = cross(c3, c1);
    const D3 r3 = cross(c1, c2);
    i1 = {invD * r1.x, invD * r2.x, invD * r3.x};
    i2 = {invD * r1.y, invD * r2.y, invD * r3.y};
    i3 = {invD * r1.z, invD * r2.z, invD * r3.z};
}
// (columns i1,i2,i3) * b
inline D3 matVec(D3 i1, D3 i2, D3 i3, D3 b) {
    return {dot(D3{i1.x, i2.x, i3.x}, b), dot(D3{i1.y, i2.y, i3.y}, b), dot(D3{i1.z, i2.z, i3.z}, b)};
}

// Apollonius vertex of 4 spheres (math.cuh:703-780).  Returns a 2-bit mask.
uint32_t quadrisector(const D4& s1, const D4& s2, const D4& s3, const D4& s4, D4& x1, D4& x2) {
    const D3 p1 = xyz(s1), p2 = xyz(s2), p3 = xyz(s3), p4 = xyz(s4);
    const double r1 = s1.w, r2 = s2.w, r3 = s3.w, r4 = s4.w;
    const D4 nh1{s1.x - s2.x, s1.y - s2.y, s1.z - s2.z, s1.w - s2.w};
    const D4 nh2{s1.x - s3.x, s1.y - s3.y, s1.z - s3.z, s1.w - s3.w};
    const D4 nh3{s1.x - s4.x, s1.y - s4.y, s1.z - s4.z, s1.w - s4.w};
    const D3 mC{.5 * (-dot(p1, p1) + dot(p2, p2) + r1 * r1 - r2 * r2),
                .5 * (-dot(p1, p1) + dot(p3, p3) + r1 * r1 - r3 * r3),
                .5 * (-dot(p1, p1) + dot(p4, p4) + r1 * r1 - r4 * r4)};
    D3 i1, i2, i3;
    inverse3(D3{nh1.x, nh2.x, nh3.x}, D3{nh1.y, nh2.y, nh3.y}, D3{nh1.z, nh2.z, nh3.z}, i1, i2, i3);
    const D3 t = matVec(i1, i2, i3, D3{nh1.w, nh2.w, nh3.w});
    const D3 mN{0.0 - t.x, 0.0 - t.y, 0.0 - t.z};
    const D3 mB = matVec(i1, i2, i3, mC);

    const double e1 = -mN.x * r1 - mB.x - p1.x;
    const double e2 = -mN.y * r1 - mB.y - p1.y;
    const double e3 = -mN.z * r1 - mB.z - p1.z;
    const double a = mN.x * mN.x + mN.y * mN.y + mN.z * mN.z - 1.0;
    const double b = 2.0 * (mN.x * e1 + mN.y * e2 + mN.z * e3);
    const double c = e1 * e1 + e2 * e2 + e3 * e3;
    const double delta = b * b - 4.0 * a * c;
    if (ltD(delta, 0.0)) return 0;

    const double q = -.5 * (b - signD(b) * std::sqrt(delta));
    const double xd1 = (c / q) - r1;
    const D3 x1n = sub(mul(mN, xd1), mB);
    x1 = {x1n.x, x1n.y, x1n.z, xd1};
    uint32_t mask = gtD(xd1 + r1, 0.0) ? 1u : 0u;
    if (gtD(xd1 + r1, 0.0)) x1.w = len(sub(p1, x1n)) - s1.w;
    if (!isCloseToZeroD(delta)) {
        const double xd2 = (q / a) - r1;
        const D3 x2n = sub(mul(mN, xd2), mB);
        x2 = {x2n.x, x2n.y, x2n.z, xd2};
        if (gtD(xd2 + r1, 0.0)) x2.w = len(sub(p1, x2n)) - s1.w;
        mask |= (gtD(xd2 + r1, 0.0) ? 1u : 0u) << 1;
    }
    return mask;
}

// Extreme points of the trisector of 3 spheres (math.cuh:636-701).
// x1 and x2 may alias (the reference calls it that way, topology_update.inl:188).
uint32_t trisector(const D4& s1, const D4& s2, const D4& s3, D4* x1, D4* x2) {
    const D3 p1 = xyz(s1), p2 = xyz(s2), p3 = xyz(s3);
    const double r1 = s1.w, r2 = s2.w, r3 = s3.w;
    const D4 nh1{s1.x - s2.x, s1.y - s2.y, s1.z - s2.z, s1.w - s2.w};
    const D4 nh2{s1.x - s3.x, s1.y - s3.y, s1.z - s3.z, s1.w - s3.w};
    const D3 nIjk = normalize(cross(sub(p2, p1), sub(p3, p1)));
    const double w = -dot(nIjk, p1);
    const D3 mC{.5 * (-dot(p1, p1) + dot(p2, p2) + r1 * r1 - r2 * r2),
                .5 * (-dot(p1, p1) + dot(p3, p3) + r1 * r1 - r3 * r3), w};
    D3 i1, i2, i3;
    inverse3(D3{nh1.x, nh2.x, nIjk.x}, D3{nh1.y, nh2.y, nIjk.y}, D3{nh1.z, nh2.z, nIjk.z}, i1, i2, i3);
    const D3 t = matVec(i1, i2, i3, D3{nh1.w, nh2.w, 0.0});
    const D3 mN{-t.x, -t.y, -t.z};
    const D3 mB = matVec(i1, i2, i3, mC);

    const double e1 = -mN.x * r1 - mB.x - p1.x;
    const double e2 = -mN.y * r1 - mB.y - p1.y;
    const double e3 = -mN.z * r1 - mB.z - p1.z;
    const double a = mN.x * mN.x + mN.y * mN.y + mN.z * mN.z - 1.0;
    const double b = 2.0 * (mN.x * e1 + mN.y * e2 + mN.z * e3);
    const double c = e1 * e1 + e2 * e2 + e3 * e3;
    const double delta = b * b - 4.0 * a * c;
    uint32_t count = 0;
    const double q = -.5 * (b - signD(b) * std::sqrt(delta));
    const double xd1 = (c / q) - r1;
    const D3 x1n = sub(mul(mN, xd1), mB);
    if (gtD(xd1 + r1, 0.0)) {
        *x1 = {x1n.x, x1n.y, x1n.z, xd1};
        count++;
    }
    const double xd2 = (q / a) - r1;
    const D3 x2n = sub(mul(mN, xd2), mB);
    if (gtD(xd2 + r1, 0.0)) {
        *x2 = {x2n.x, x2n.y, x2n.z, xd2};
        if (count == 0) *x1 = *x2;
        count++;
    }
    return count;
}

// 2x2 inverse with an explicit fma (math.cuh:390-396)
inline void inverse2(D2 a1, D2 a2, D2& c1, D2& c2) {
    const double invDet = 1.0 / std::fma(a1.x, a2.y, -a2.x * a1.y);
    c1 = {invDet * a2.y, -invDet * a1.y};
    c2 = {-invDet * a2.x, invDet * a1.x};
}
inline void matMul2(D2 a1, D2 a2, D2 b1, D2 b2, D2& c1, D2& c2) {
    c1 = {a1.x * b1.x + a2.x * b1.y, a1.y * b1.x + a2.y * b1.y};
    c2 = {a1.x * b2.x + a2.x * b2.y, a1.y * b2.x + a2.y * b2.y};
}

// Symmetric 2x2 eigendecomposition (math.cuh:428-503; LAPACK dlaev2 steps).
void eigen2(D2 m1, D2 m2, D2& e, D2& v) {
    const double a = m1.x, b = m2.x, c = m2.y;
    const double sm = a + c, df = a - c, adf = std::fabs(df), tb = b + b, ab = std::fabs(tb);
    double acmx = a, acmn = c;
    if (std::fabs(a) <= std::fabs(c)) std::swap(acmx, acmn);
    double aa = ab, bb = adf;
    if (adf < ab) std::swap(aa, bb);
    const double temp = aa / bb;
    const double rt = bb * std::sqrt(1.0 + temp * temp);
    const double sgn1 = std::copysign(1.0, sm);
    const double rt1 = .5 * (sm + sgn1 * rt);
    const double rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
    const double sgn2 = std::copysign(1.0, df);
    const double cs = df + sgn2 * rt;
    const double acs = std::fabs(cs);
    aa = tb;
    bb = cs;
    if (ltD(acs, ab)) std::swap(aa, bb);
    const double t = -aa / bb;
    double sn1 = 1.0 / std::sqrt(1.0 + t * t);
    double cs1 = t * sn1;
    if (ltD(acs, ab)) std::swap(sn1, cs1);
    if (ltD(acs, ab) && isCloseToZeroD(ab)) {
        cs1 = 1.0;
        sn1 = 0.0;
    }
    if (isCloseD(sgn1, sgn2)) {
        const double tn = cs1;
        cs1 = -sn1;
        sn1 = tn;
    }
    e = {rt2, rt1};
    v = {-sn1, cs1};
}

// Common tangent planes of 3 spheres (math.cuh:577-634).
void tangentPlanes(const D4& s1, const D4& s2, const D4& s3, D3& n1, D3& n2) {
    const D3 p1 = xyz(s1), p2 = xyz(s2), p3 = xyz(s3);
    const D4 nh1{s1.x - s2.x, s1.y - s2.y, s1.z - s2.z, s1.w - s2.w};
    const D4 nh2{s1.x - s3.x, s1.y - s3.y, s1.z - s3.z, s1.w - s3.w};
    const D2 aC1{nh1.x, nh2.x}, aC2{nh1.y, nh2.y}, bC1{nh1.z, nh2.z}, bC2{nh1.w, nh2.w};
    D2 iA1, iA2;
    inverse2(aC1, aC2, iA1, iA2);
    D2 cC1, cC2;
    matMul2(iA1, iA2, bC1, bC2, cC1, cC2);
    const D2 eC1{1.0 + cC1.x * cC1.x + cC1.y * cC1.y, 0.0 + cC1.x * cC2.x + cC1.y * cC2.y};
    const D2 eC2{0.0 + cC1.x * cC2.x + cC1.y * cC2.y, -1.0 + cC2.x * cC2.x + cC2.y * cC2.y};
    D2 ev, evec;
    eigen2(eC1, eC2, ev, evec);
    const double alpha0 = std::sqrt(std::fabs(ev.x));
    const double alpha1 = std::sqrt(std::fabs(ev.y));
    auto dot2 = [](D2 a, D2 b) { return a.x * b.x + a.y * b.y; };
    D2 z{1.0, alpha0 / alpha1};
    D2 y34{dot2(D2{evec.x, -evec.y}, z), dot2(D2{evec.y, evec.x}, z)};
    double x1 = -dot2(D2{cC1.x, cC2.x}, y34);
    double x2 = -dot2(D2{cC1.y, cC2.y}, y34);
    n1 = normalize(D3{x1, x2, y34.x});
    z = {1.0, -alpha0 / alpha1};
    y34 = {dot2(D2{evec.x, -evec.y}, z), dot2(D2{evec.y, evec.x}, z)};
    x1 = -dot2(D2{cC1.x, cC2.x}, y34);
    x2 = -dot2(D2{cC1.y, cC2.y}, y34);
    n2 = normalize(D3{x1, x2, y34.x});
    const D3 p = add(p1, mul(n1, s1.w));
    if (gtD(std::fabs(dot(n1, sub(p, p1)) - s1.w), 1e-4) || gtD(std::fabs(dot(n1, sub(p, p2)) - s2.w), 1e-4) ||
        gtD(std::fabs(dot(n1, sub(p, p3)) - s3.w), 1e-4)) {
        n1 = mul(n1, -1.0);
        n2 = mul(n2, -1.0);
    }
}

// 4x4 determinant by cofactors along the first column (math.cuh:519-531)
double det4(D4 c1, D4 c2, D4 c3, D4 c4) {
    const double d1 = dot(D3{c2.y, c2.z, c2.w}, cross(D3{c3.y, c3.z, c3.w}, D3{c4.y, c4.z, c4.w}));
    const double d2 = dot(D3{c2.x, c2.z, c2.w}, cross(D3{c3.x, c3.z, c3.w}, D3{c4.x, c4.z, c4.w}));
    const double d3 = dot(D3{c2.x, c2.y, c2.w}, cross(D3{c3.x, c3.y, c3.w}, D3{c4.x, c4.y, c4.w}));
    const double d4 = dot(D3{c2.x, c2.y, c2.z}, cross(D3{c3.x, c3.y, c3.z}, D3{c4.x, c4.y, c4.z}));
    return c1.x * d1 - c1.y * d2 + c1.z * d3 - c1.w * d4;
}

// Sphere through 4 points (math.cuh:533-575)
D4 sphereFromPoints(D3 p1, D3 p2, D3 p3, D3 p4) {
    const double t1 = -dot(p1, p1), t2 = -dot(p2, p2), t3 = -dot(p3, p3), t4 = -dot(p4, p4);
    const D4 X{p1.x, p2.x, p3.x, p4.x}, Y{p1.y, p2.y, p3.y, p4.y}, Z{p1.z, p2.z, p3.z, p4.z}, O{1, 1, 1, 1};
    const D4 T{t1, t2, t3, t4};
    const double t = det4(X, Y, Z, O);
    const double d = det4(T, Y, Z, O) / t;
    const double e = det4(X, T, Z, O) / t;
    const double f = det4(X, Y, T, O) / t;
    const double g = det4(X, Y, Z, T) / t;
    return {-d * .5, -e * .5, -f * .5, .5 * std::sqrt(d * d + e * e + f * f - 4. * g)};
}

// ---------------------------------------------------------------------------
// Morton codes (lbvh.cu:11-67), restated as a bit-by-bit interleave
// ---------------------------------------------------------------------------
uint64_t interleave10(uint32_t x, uint32_t y, uint32_t z) {
    uint64_t code = 0;
    for (int b = 0; b < 10; b++) {
        code |= uint64_t((x >> b) & 1u) << (3 * b);
        code |= uint64_t((y >> b) & 1u) << (3 * b + 1);
        code |= uint64_t((z >> b) & 1u) << (3 * b + 2);
    }
    return code;
}
uint64_t morton64(uint32_t x, uint32_t y, uint32_t z) {
    const uint64_t lo = interleave10(x & 1023u, y & 1023u, z & 1023u);
    const uint64_t hi = interleave10(x >> 10, y >> 10, z >> 10);
    return (hi << 30) | lo;
}
// truncating float -> u32 of a value known to lie in [0, 2^20]; NaN and
// negatives map to 0.
inline uint32_t quantize(float v) {
    if (!(v >= 0.f)) return 0u;
    if (v >= 4294967040.f) return 0xffffffffu;
    return uint32_t(v);
}

struct SceneBox {
    double mn[3], mx[3];
    double rmin, rmax;
    float fmn[3], fmx[3];
};

SceneBox sceneBox(const float* s, uint32_t n) {
    SceneBox b;
    for (int a = 0; a < 3; a++) {
        b.mn[a] = DBL_MAX;
        b.mx[a] = -DBL_MAX;
    }
    b.rmin = DBL_MAX;
    b.rmax = -DBL_MAX;
    for (uint32_t i = 0; i < n; i++) {
        const double r = s[4 * i + 3];
        for (int a = 0; a < 3; a++) {
            const double c = s[4 * i + a];
            b.mn[a] = std::fmin(b.mn[a], c - r);
            b.mx[a] = std::fmax(b.mx[a], c + r);
        }
        b.rmax = std::fmax(b.rmax, r);
        b.rmin = std::fmin(b.rmin, r);
    }
    for (int a = 0; a < 3; a++) {
        b.fmn[a] = float(b.mn[a]);
        b.fmx[a] = float(b.mx[a]);
    }
    return b;
}

uint64_t mortonOf(const float* p, const float* fmn, const float* fmx) {
    uint32_t u[3];
    for (int a = 0; a < 3; a++) {
        const float scale = 1.f / (fmx[a] - fmn[a]);
        const float q = (p[a] - fmn[a]) * scale;
        u[a] = quantize(q * 1048575.0f);
    }
    return morton64(u[0], u[1], u[2]);
}

// ---------------------------------------------------------------------------
// LBVH (lbvh.cu:88-323)
// ---------------------------------------------------------------------------
struct Lbvh {
    uint32_t n = 0;
    std::vector<uint64_t> codes;   // sorted
    std::vector<uint32_t> indices; // sorted position -> original index
    std::vector<int32_t> parent, left, right;
    std::vector<float> aabb;       // 6 floats per node
    std::vector<uint32_t> sizes;
    std::vector<float> maxr;

    Box box(uint32_t node) const {
        const float* a = &aabb[6 * size_t(node)];
        return {{a[0], a[1], a[2]}, {a[3], a[4], a[5]}};
    }
};

inline int clz64(uint64_t v) { return v == 0 ? 64 : __builtin_clzll(v); }

// Karras 2012 delta (lbvh.cu:144-155)
inline int delta(const std::vector<uint64_t>& c, int i, int j, int n) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = c[i], b = c[j];
    if (a != b) return clz64(a ^ b);
    return clz64(uint64_t(int64_t(i ^ j))) + 64;
}

void buildLbvh(const float* xyzr, uint32_t n, const SceneBox& sb, Lbvh& b) {
    b.n = n;
    const uint32_t nodes = 2 * n - 1;
    std::vector<uint64_t> codes(n);
    for (uint32_t i = 0; i < n; i++) codes[i] = mortonOf(&xyzr[4 * i], sb.fmn, sb.fmx);
    b.indices.resize(n);
    for (uint32_t i = 0; i < n; i++) b.indices[i] = i;
    std::stable_sort(b.indices.begin(), b.indices.end(),
                     [&](uint32_t x, uint32_t y) { return codes[x] < codes[y]; });
    b.codes.resize(n);
    for (uint32_t i = 0; i < n; i++) b.codes[i] = codes[b.indices[i]];

    b.parent.assign(nodes, -1);
    b.left.assign(nodes, 0);
    b.right.assign(nodes, 0);
    b.aabb.assign(6 * size_t(nodes), 0.f);
    b.sizes.assign(nodes, 0);
    b.maxr.assign(nodes, 0.f);

    // leaves (setLeaves): box in FP64 then rounded to FP32
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t node = n - 1 + i;
        b.left[node] = int32_t(i);
        b.right[node] = int32_t(i) + 1;
        const float* s = &xyzr[4 * size_t(b.indices[i])];
        const double x = s[0], y = s[1], z = s[2], r = s[3];
        float* a = &b.aabb[6 * size_t(node)];
        a[0] = float(x - r);
        a[1] = float(y - r);
        a[2] = float(z - r);
        a[3] = float(x + r);
        a[4] = float(y + r);
        a[5] = float(z + r);
        b.sizes[node] = 1;
        b.maxr[node] = s[3];
    }
    // internal nodes (construct)
    const int ni = int(n);
    for (int i = 0; i < ni - 1; i++) {
        const int dd = delta(b.codes, i, i + 1, ni) - delta(b.codes, i, i - 1, ni);
        const int d = (dd > 0) - (dd < 0);
        const int deltaMin = delta(b.codes, i, i - d, ni);
        int lmax = 2;
        while (delta(b.codes, i, i + lmax * d, ni) > deltaMin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(b.codes, i, i + (l + t) * d, ni) > deltaMin) l += t;
        const int j = i + l * d;
        const int deltaNode = delta(b.codes, i, j, ni);
        int s = 0, k = 2, t;
        do {
            t = (l + k - 1) / k;
            k <<= 1;
            if (delta(b.codes, i, i + (s + t) * d, ni) > deltaNode) s += t;
        } while (t > 1);
        const int gamma = i + s * d + std::min(d, 0);
        int lc = gamma, rc = gamma + 1;
        if (std::min(i, j) == gamma) lc += ni - 1;
        if (std::max(i, j) == gamma + 1) rc += ni - 1;
        b.left[i] = lc;
        b.right[i] = rc;
        b.parent[lc] = i;
        b.parent[rc] = i;
    }
    if (n > 1) b.parent[0] = -1;
    // refit, bottom-up by recursion order: children before parents
    std::vector<uint32_t> order;
    order.reserve(n);
    if (n > 1) {
        std::vector<uint32_t> st{0};
        while (!st.empty()) {
            const uint32_t v = st.back();
            st.pop_back();
            order.push_back(v);
            if (int(v) < ni - 1) {
                if (b.left[v] < ni - 1) st.push_back(uint32_t(b.left[v]));
                if (b.right[v] < ni - 1) st.push_back(uint32_t(b.right[v]));
            }
        }
    }
    for (auto it = order.rbegin(); it != order.rend(); ++it) {
        const uint32_t v = *it;
        const float* la = &b.aabb[6 * size_t(b.left[v])];
        const float* ra = &b.aabb[6 * size_t(b.right[v])];
        float* a = &b.aabb[6 * size_t(v)];
        for (int c = 0; c < 3; c++) {
            a[c] = std::fmin(std::fmin(std::fmin(FLT_MAX, la[c]), la[3 + c]), std::fmin(ra[c], ra[3 + c]));
            a[3 + c] = std::fmax(std::fmax(std::fmax(-FLT_MAX, la[c]), la[3 + c]), std::fmax(ra[c], ra[3 + c]));
        }
        b.sizes[v] = b.sizes[b.left[v]] + b.sizes[b.right[v]];
        b.maxr[v] = std::fmax(b.maxr[b.left[v]], b.maxr[b.right[v]]);
    }
}

// ---------------------------------------------------------------------------
// heap (heap.hpp:11-43)
// ---------------------------------------------------------------------------
void heapify(uint32_t* keys, float* vals, uint32_t node, uint32_t size) {
    uint32_t j = node;
    while (true) {
        const uint32_t l = 2 * j + 1, r = 2 * j + 2;
        uint32_t largest = j;
        if (l < size && vals[l] > vals[largest]) largest = l;
        if (r < size && vals[r] > vals[largest]) largest = r;
        if (largest == j) return;
        std::swap(vals[j], vals[largest]);
        std::swap(keys[j], keys[largest]);
        j = largest;
    }
}
void heapsort(uint32_t* keys, float* vals, int size) {
    while (size) {
        std::swap(vals[0], vals[size - 1]);
        std::swap(keys[0], keys[size - 1]);
        heapify(keys, vals, 0, uint32_t(--size));
    }
}

// kNN traversal (lbvh.inl:8-105). Returns max stack depth.
int knnQuery(const Lbvh& b, const std::vector<F4>& sf, uint32_t i, uint32_t K, float startingDistance,
             uint32_t* keys, float* vals, float top = FLT_MAX) {
    const uint32_t n = b.n;
    const F4 s = sf[i];
    // top < FLT_MAX: a range query, only gaps below top are listed
    for (uint32_t k = 0; k < K; k++) {
        keys[k] = kNone;
        vals[k] = top;
    }
    struct Entry { uint32_t node; float d; };
    std::vector<Entry> stack;
    stack.reserve(64);
    int maxDepth = 0;
    uint32_t node = 0;
    while (true) {
        while (true) {
            if (node >= n - 1) break;
            const uint32_t L = uint32_t(b.left[node]), R = uint32_t(b.right[node]);
            const Box lb = b.box(L), rb = b.box(R);
            const float dL = boxDistance(lb, s), dLmax = boxFarthestDistance(lb, s);
            const float dR = boxDistance(rb, s), dRmax = boxFarthestDistance(rb, s);
            const bool leftClosest = dL < dR;
            const float minFar = leftClosest ? dR : dL;
            const float maxFar = leftClosest ? dRmax : dLmax;
            const uint32_t farChild = leftClosest ? R : L;
            const float minClose = leftClosest ? dL : dR;
            const float maxClose = leftClosest ? dLmax : dRmax;
            const uint32_t closeChild = leftClosest ? L : R;
            const bool exploreFar = minFar < vals[0] && maxFar > startingDistance;
            const bool exploreClose = minClose < vals[0] && maxClose > startingDistance;
            if (!exploreFar && !exploreClose) break;
            node = exploreClose ? closeChild : farChild;
            if (exploreClose && exploreFar) {
                stack.push_back({farChild, minFar});
                maxDepth = std::max(maxDepth, int(stack.size()));
            }
        }
        if (node >= n - 1 && node < 2 * n - 1) {
            const uint32_t j = node - (n - 1);
            const float d = sphereDistF(s, sf[j]);
            if (j != i && d < vals[0] && d > startingDistance) {
                vals[0] = d;
                keys[0] = j;
                heapify(keys, vals, 0, K);
            }
        }
        bool found = false;
        while (!stack.empty()) {
            const Entry e = stack.back();
            stack.pop_back();
            if (e.d > vals[0]) continue;
            node = e.node;
            found = true;
            break;
        }
        if (!found) break;
    }
    heapsort(keys, vals, int(K));
    return maxDepth;
}

// ---------------------------------------------------------------------------
// Cell state (topology_update.cuh:20-97)
// ---------------------------------------------------------------------------
struct Vtx { uint32_t j, k, l; uint8_t type, status; };
struct Edg { uint32_t j, k, vertexNb; uint8_t status; };
struct Cell {
    std::vector<Vtx> V;
    std::vector<Edg> E;
};

struct NewEdge { uint32_t k; bool two; bool closed; };

struct Context {
    uint32_t n = 0;
    orc_config cfg{};

    std::vector<D4> S;   // sorted sites + 6 artificial, FP64
    std::vector<F4> Sf;  // float copies (toFloat)
    const Lbvh* bvh = nullptr;
    std::vector<uint8_t> status;
    std::vector<uint8_t> buried;  // snapshot of status == Buried after the kNN pass
    std::vector<float> knownRadii;
    std::vector<Cell> cells;
    std::vector<uint32_t> flags;
    std::vector<uint32_t> siteK;  // per-site neighbour count of the proposal path (empty: all K)
    std::vector<float> range;     // per-site search radius of the range path (empty: none)
    std::vector<uint32_t> want;   // completion mode: per cell, kNone, a wanted site, or a rebuild code
};

// Completion mode: a round's stuck-record list and the codes of want[]
// besides site ids (the HIP kernels use the same).
constexpr uint32_t kStuckCap = 16384;
constexpr uint32_t kRebuildSites = 528;        // LDS list of a rebuild (2 x the edge capacity)
constexpr uint32_t kWantRebuilt = 0xfffffffdu;  // rebuilt in this build: not again
constexpr uint32_t kWantRebuild = 0xfffffffeu;  // stuck: rebuild in this round

// Upper gap bound of a range query of radius b around a site of radius r:
// raised to the float just above -r, since a box holding the site's centre
// has box gap -r while the sites whose spheres hold that centre have gaps
// below -r (a lower bound there would prune every box around the centre).
inline float rangeTop(float b, float r) {
    const float negR = -r;
    uint32_t bits;
    std::memcpy(&bits, &negR, 4);
    float above;
    if (negR < 0.f) {
        bits -= 1u;
        std::memcpy(&above, &bits, 4);
    } else {
        const uint32_t one = 1u;
        std::memcpy(&above, &one, 4);
    }
    return b > above ? b : above;
}

// Known radius of a range query of bound b that listed fewer sites than its
// capacity: every site whose FP32 gap is below rangeTop(b, r_i) was listed,
// so every site whose exact gap is below it, less a margin of 2^-18 of the
// magnitudes involved (the FP32 rounding of a gap), is known.
inline double rangeKnownRadius(float b, const D4& si) {
    const double top = double(rangeTop(b, float(si.w)));
    const double scale = std::fabs(top) + std::fabs(si.x) + std::fabs(si.y) + std::fabs(si.z) + std::fabs(si.w);
    return top - scale * 0x1p-18;
}

// Steps of the validation loops: 64 records, contributors applied lowest
// lane first (the wave64 order the HIP kernels use), or with lanes = 32 the
// reference's own order, 32-lane warps visited highest lane first (forEach,
// apo/include/apo/gpu/utils.cuh:17-32).
inline int waveWidth(const Context& C) { return C.cfg.lanes == 32 ? 32 : kWave; }
inline int laneOrder(const Context& C, int t, int count) { return C.cfg.lanes == 32 ? count - 1 - t : t; }

// Neighbours asked for by each class of the learned proposal (mirrors the
// product's mapping; class 7, not certified within 64, runs `fallback`).
inline uint32_t proposalK(uint32_t c, uint32_t fallback) {
    static const uint32_t k[7] = {20, 24, 28, 32, 40, 48, 64};
    return c < 7 ? k[c] : fallback;
}

inline D4 vertexPosition(const Context& C, const D4& si, const Vtx& v) {
    D4 x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    quadrisector(si, C.S[v.j], C.S[v.k], C.S[v.l], x[0], x[1]);
    return x[v.type];
}

// Default validation predicate (topology_update.inl:578-592): no site on a
// cell edge, other than skip1/skip2, intersects the sphere.
inline bool edgeSitesClear(const Context& C, const Cell& cell, const D4& sphere, uint32_t skip1, uint32_t skip2) {
    for (const Edg& e : cell.E) {
        if (e.j != skip1 && e.j != skip2 && intersectD(C.S[e.j], sphere)) return false;
        if (e.k != skip1 && e.k != skip2 && intersectD(C.S[e.k], sphere)) return false;
    }
    return true;
}

inline bool known(const Cell& cell, uint32_t j) {
    for (const Edg& e : cell.E)
        if (e.j == j || e.k == j) return true;
    return false;
}

// updateCell (topology_update.inl:49-555).  `V` is the vertex list the
// update sees (the full cell list, or the list of vertices created by the
// current validation kernel, which starts empty in edge/bisector validation).
bool updateCell(const Context& C, uint32_t i, const D4& si, std::vector<Vtx>& V, Cell& cell, uint32_t j,
                const D4& sj, uint32_t& flags) {
    std::vector<Edg>& E = cell.E;
    if (V.empty() && E.empty()) return false;
    const double ri = si.w, rj = sj.w;
    const double bigger = std::fmax(ri, rj), smaller = std::fmin(ri, rj);
    if (ltD(len(D3{si.x - sj.x, si.y - sj.y, si.z - sj.z}), bigger - smaller)) {
        if (ltD(rj, ri)) return false;
        V.clear();
        E.clear();
        return false;
    }
    const size_t baseV = V.size(), baseE = E.size();

    // (1) vertices whose empty sphere intersects s_j are invalidated
    for (size_t v = 0; v < baseV; v++) {
        const D4 pos = vertexPosition(C, si, V[v]);
        if (!intersectD(pos, sj)) continue;
        V[v].status = 0xff;
        const Vtx vx = V[v];
        for (size_t e = 0; e < baseE; e++) {
            Edg& ed = E[e];
            const bool e1 = ed.j == vx.j && ed.k == vx.k;
            const bool e2 = ed.j == vx.j && ed.k == vx.l;
            const bool e3 = ed.j == vx.k && ed.k == vx.l;
            if (e1 || e2 || e3) ed.vertexNb--;
        }
    }

    // (2) candidate vertices on every edge (k,l): quadrisector(i, j, k, l)
    std::vector<Vtx> newV;
    std::vector<NewEdge> newE;
    std::vector<uint8_t> closedCand(baseE, 0);
    const uint32_t neCap = C.cfg.new_edge_capacity;
    bool neOverflow = false;
    auto pushNE = [&](NewEdge ne) {
        if (newE.size() < neCap) newE.push_back(ne);
        else neOverflow = true;
    };
    for (size_t e = 0; e < baseE; e++) {
        const Edg ed = E[e];
        const uint32_t k = ed.j, l = ed.k;
        D4 q[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        const uint32_t mask = quadrisector(si, sj, C.S[k], C.S[l], q[0], q[1]);
        if (mask == 0 && ed.vertexNb == kNone) {
            // closed edge fully cut by the bisector of j (topology_update.inl:186-191)
            trisector(si, C.S[k], C.S[l], &q[0], &q[0]);
            if (intersectD(q[0], sj)) E[e].vertexNb = 0;
        }
        uint32_t valid = mask;
        for (uint32_t nv = 0; nv < 2; nv++)
            if ((mask >> nv) & 1u)
                if (!edgeSitesClear(C, cell, q[nv], k, l)) valid &= ~(1u << nv);
        const uint32_t nNew = uint32_t(__builtin_popcount(valid));
        if (mask != 0 && nNew > 0) {
            if (E[e].vertexNb == kNone) E[e].vertexNb = 0;
            E[e].vertexNb += nNew;
        }
        uint32_t jj = j, kk = k, ll = l;
        if (jj > ll) std::swap(jj, ll);
        if (jj > kk) std::swap(jj, kk);
        if (kk > ll) std::swap(kk, ll);
        for (uint32_t nv = 0; nv < 2; nv++)
            if ((valid >> nv) & 1u) newV.push_back({jj, kk, ll, uint8_t(nv), 0});
        for (uint32_t nv = 0; nv < 2; nv++)
            if ((valid >> nv) & 1u) {
                pushNE({k, nNew == 2, false});
                pushNE({l, nNew == 2, false});
            }
        closedCand[e] = mask == 0;
    }

    // (3) closed elliptic trisectors (topology_update.inl:325-433)
    for (size_t e = 0; e < baseE; e++) {
        if (!closedCand[e]) continue;
        const uint32_t ids[2] = {E[e].j, E[e].k};
        for (int s = 0; s < 2; s++) {
            const uint32_t id = ids[s];
            bool share = false;
            for (const NewEdge& ne : newE)
                if (ne.k == id) { share = true; break; }
            if (share) continue;
            D4 x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            if (trisector(si, sj, C.S[id], &x[0], &x[1]) != 2) continue;
            if (!edgeSitesClear(C, cell, x[0], id, i)) continue;
            pushNE({id, false, true});
        }
    }
    if (neOverflow) flags |= kFlagNewEdgeOverflow;

    // (4) vertex list: surviving old vertices then new ones
    {
        std::vector<Vtx> out;
        out.reserve(baseV + newV.size());
        for (size_t v = 0; v < V.size(); v++)
            if (!(V[v].status == 0xff && v < baseV)) out.push_back(V[v]);
        for (const Vtx& v : newV) out.push_back(v);
        if (out.size() > C.cfg.vertex_capacity) {
            flags |= kFlagVertexOverflow;
            out.resize(C.cfg.vertex_capacity);
        }
        V.swap(out);
    }
    // (5) new edges: one per distinct k, ascending, vertexNb = run length
    std::vector<uint32_t> keys;
    keys.reserve(newE.size());
    for (const NewEdge& ne : newE) keys.push_back(ne.k);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<Edg> outE;
    outE.reserve(baseE + keys.size());
    for (size_t e = 0; e < E.size(); e++)
        if (!(E[e].vertexNb == 0 && e < baseE)) outE.push_back(E[e]);
    for (uint32_t key : keys) {
        uint32_t count = 0;
        bool closed = false;
        for (const NewEdge& ne : newE)
            if (ne.k == key) {
                count++;
                closed = ne.closed;
            }
        outE.push_back({std::min(j, key), std::max(j, key), closed ? kNone : count, 0});
    }
    if (outE.size() > C.cfg.edge_capacity) {
        flags |= kFlagEdgeOverflow;
        outE.resize(C.cfg.edge_capacity);
    }
    E.swap(outE);
    return !keys.empty();
}

// cellRadius (topology_update.inl:595-653)
double cellRadius(const Context& C, const D4& si, const Cell& cell) {
    double maxR = 0.0;
    for (const Vtx& v : cell.V) maxR = std::fmax(maxR, vertexPosition(C, si, v).w);
    for (const Edg& e : cell.E) {
        D4 x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        if (trisector(si, C.S[e.j], C.S[e.k], &x[0], &x[1]) == 2) maxR = std::fmax(maxR, std::fmax(x[0].w, x[1].w));
    }
    return maxR;
}

void initCell(const Context& C, Cell& cell) {
    const uint32_t n = C.n;
    // artificial vertices / edges (algorithm.inl:1472-1499)
    static const uint32_t av[8][3] = {{0, 2, 4}, {0, 2, 5}, {0, 3, 4}, {0, 3, 5},
                                      {1, 2, 4}, {1, 2, 5}, {1, 3, 4}, {1, 3, 5}};
    static const uint32_t ae[12][2] = {{0, 2}, {2, 4}, {1, 2}, {2, 5}, {1, 4}, {1, 5},
                                       {0, 5}, {0, 4}, {1, 3}, {3, 5}, {0, 3}, {3, 4}};
    cell.V.clear();
    cell.E.clear();
    for (auto& t : av) cell.V.push_back({n + t[0], n + t[1], n + t[2], 0, 0});
    for (auto& t : ae) cell.E.push_back({n + t[0], n + t[1], 2, 0});
}

// ---------------------------------------------------------------------------
// BVH searches of the validation kernels
// ---------------------------------------------------------------------------
// Nearest contributing site to a vertex sphere (algorithm.inl:305-414)
uint32_t vertexSearch(const Context& C, const Cell& cell, uint32_t i, const Vtx& vx, F4 cur, uint32_t& flags) {
    const Lbvh& b = *C.bvh;
    const uint32_t n = b.n;
    struct Entry { uint32_t node; float d; };
    std::vector<Entry> stack;
    uint32_t node = 0;
    float nd = FLT_MAX;
    uint32_t nb = kNone;
    while (true) {
        while (true) {
            if (node >= n - 1) break;
            const uint32_t L = uint32_t(b.left[node]), R = uint32_t(b.right[node]);
            const float dL = boxDistance(b.box(L), cur), dR = boxDistance(b.box(R), cur);
            const bool lc = dL < dR;
            const float minFar = lc ? dR : dL;
            const uint32_t farChild = lc ? R : L;
            const float minClose = lc ? dL : dR;
            const uint32_t closeChild = lc ? L : R;
            const bool ef = minFar < nd, ec = minClose < nd;
            if (!ef && !ec) break;
            node = ec ? closeChild : farChild;
            if (ec && ef) {
                stack.push_back({farChild, minFar});
                if (stack.size() > 32) flags |= kFlagStackOverflow;
            }
        }
        if (node >= n - 1 && node < 2 * n - 1) {
            const uint32_t j = node - (n - 1);
            if (j != i && j != vx.j && j != vx.k && j != vx.l && !C.buried[j] && !known(cell, j)) {
                const float d = sphereDistF(C.Sf[j], cur);
                if (d < nd) {
                    nd = d;
                    nb = j;
                }
            }
        }
        bool found = false;
        while (!stack.empty()) {
            const Entry e = stack.back();
            stack.pop_back();
            if (e.d > nd) continue;
            node = e.node;
            found = true;
            break;
        }
        if (!found) break;
    }
    return nb;
}

struct EdgeQuery {
    F4 bs;
    bool hyperbola;
    F3 n1, n2, pp1, pp2;
    int cover = 0;  // balls in `ball` (0: no pruning)
    F4 ball[8];     // every contributor j has |p_j - c| - r_j <= r for one ball
};
constexpr int kCoverBalls = 8;
// edge searches whose sphere exceeds this multiple of the known radius get a cover
constexpr float kCoverRatio = 2.0f;
struct EdgeCover {
    int count;
    F4 ball[kCoverBalls];
};

// ---------------------------------------------------------------------------
// Cover of the empty spheres of a cell edge.  No counterpart in the
// reference: an exact pruning aid for the edge contributor search, which
// must leave every search result unchanged (the traversal keeps its order
// and only skips subtrees that hold no possible contributor).
//
// The centres of the spheres tangent to s1, s2, s3 with radius w are
// c(w) = p1 + u(w) +- h(w) n, where n is the unit normal of the centres'
// plane, u(w) = P + Q w lies in that plane and h(w)^2 = A w^2 + B w + Cc
// (u is linear in w because subtracting the tangency equations pairwise
// leaves two linear equations).  Each sign of h traces a piece of the
// trisector parametrised by w.  A neighbour s becomes tangent where
//     |c(w) - p_s| = r_s + w  <=>  L0 + L1 w = +-2 sigma h(w),
// sigma = n.(p_s - p1); squared, a quadratic in w whose roots are all the
// points where the gap of s changes sign.
// ---------------------------------------------------------------------------
struct TriCurve {
    D3 P, Q, n;  // P, Q relative to p1
    double A, B, Cc;
};

bool triCurve(const D4& s1, const D4& s2, const D4& s3, TriCurve& t) {
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
    if (!(std::fabs(det) > 0.0)) return false;
    const double inv = 1.0 / det;
    t.P = mul(add(mul(c23, b0x), mul(c31, b0y)), inv);
    t.Q = mul(add(mul(c23, b1x), mul(c31, b1y)), inv);
    t.A = 1.0 - dot(t.Q, t.Q);
    t.B = 2.0 * (s1.w - dot(t.P, t.Q));
    t.Cc = s1.w * s1.w - dot(t.P, t.P);
    return std::isfinite(t.A) && std::isfinite(t.B) && std::isfinite(t.Cc) && std::isfinite(dot(t.P, t.P) + dot(t.Q, t.Q));
}

// Real roots of a w^2 + b w + c, counting a slightly negative discriminant
// (rounding) as a double root.  Returns -1 when every w is a root.
int realRoots(double a, double b, double c, double r[2]) {
    const double sa = std::fabs(a), sb = std::fabs(b), sc = std::fabs(c);
    const double scale = std::fmax(sa, std::fmax(sb, sc));
    if (!(scale > 0.0)) return -1;
    if (sa <= 1e-13 * scale) {
        if (sb <= 1e-13 * scale) return -1;
        r[0] = -c / b;
        return 1;
    }
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) {
        if (disc < -1e-10 * (b * b + std::fabs(4.0 * a * c))) return 0;
        disc = 0.0;
    }
    const double q = -.5 * (b + std::copysign(std::sqrt(disc), b));
    r[0] = q / a;
    r[1] = (q != 0.0) ? c / q : r[0];
    return 2;
}

template <class Nb>
bool clearOfAll(const Context& C, const Nb& N, const D4& x, uint32_t skip, double tol) {
    for (uint32_t s : N)
        if (s != skip && sphereDistD(C.S[s], x) < -tol) return false;
    return true;
}

// Cover of the empty spheres of edge (k,l): balls such that every sphere the
// contributor test can accept lies inside one of them.  The neighbours'
// tangency roots split each piece of the trisector into intervals on which
// no neighbour changes side, so one midpoint test per interval and half tells
// whether the interval can hold an accepted sphere.  An interval [wa, wb] of
// one half (h monotone on it) lies in the ball centred at c(wm), wm the
// midpoint, of radius |Q| (wb - wa) / 2 + max |h(w) - h(wm)| + wb.  A
// contributor j touches an accepted sphere, so |p_j - centre| - r_j <= radius
// for one ball.  false: no cover could be certified (search unpruned).
bool edgeCover(const Context& C, const Cell& cell, const D4& si, uint32_t k, uint32_t l, EdgeCover& cov) {
    cov.count = 0;
    TriCurve t;
    if (!triCurve(si, C.S[k], C.S[l], t)) return false;
    std::vector<uint32_t> N;
    N.reserve(2 * cell.E.size());
    for (const Edg& e : cell.E) {
        for (uint32_t s : {e.j, e.k}) {
            if (s == k || s == l) continue;
            bool dup = false;
            for (uint32_t m : N) dup = dup || m == s;
            if (!dup) N.push_back(s);
        }
    }
    const D3 p1 = xyz(si);
    const double r1 = si.w;
    const double tolBase = r1 + len(t.P);
    auto hAt = [&](double w) { return std::sqrt(std::fmax(t.A * w * w + t.B * w + t.Cc, 0.0)); };
    auto point = [&](double w, double sgn) -> D4 {
        const D3 c = add(p1, add(add(t.P, mul(t.Q, w)), mul(t.n, sgn * hAt(w))));
        return {c.x, c.y, c.z, w};
    };
    auto tolAt = [&](double w) { return 1e-9 + 1e-7 * (std::fabs(w) + tolBase); };
    // some neighbour intersects the sphere at x by more than margin + tol
    auto clearlyInvalid = [&](const D4& x, double margin) {
        for (uint32_t s : N)
            if (sphereDistD(C.S[s], x) < -(margin + tolAt(x.w))) return true;
        return false;
    };

    // pieces of the w axis where the trisector is real (h^2 >= 0); accepted
    // spheres have w = |p1 - c| - r1 >= -r1
    double tr[2];
    const int nt = realRoots(t.A, t.B, t.Cc, tr);
    if (nt < 0 || nt == 1) return false;
    if (nt == 2 && !(std::isfinite(tr[0]) && std::isfinite(tr[1]))) return false;
    if (nt == 2 && tr[0] > tr[1]) std::swap(tr[0], tr[1]);
    const double inf = std::numeric_limits<double>::infinity();
    struct Span { double a, b; };
    std::vector<Span> pieces;
    if (t.A < 0.0) {
        if (nt != 2) return false;
        const double a = std::fmax(tr[0], -r1);
        if (!(tr[1] > a)) return false;
        pieces.push_back({a, tr[1]});
    } else {
        if (nt == 2) {
            if (tr[0] > -r1) pieces.push_back({-r1, tr[0]});
            pieces.push_back({std::fmax(tr[1], -r1), inf});
        } else {
            pieces.push_back({-r1, inf});
        }
    }
    // breakpoints: tangency roots of every neighbour and the extremum of h^2
    std::vector<double> roots;
    for (uint32_t s : N) {
        const D4 ss = C.S[s];
        const D3 qs = sub(xyz(ss), p1);
        const double sigma = dot(t.n, qs);
        const double L0 = dot(qs, qs) - 2.0 * dot(t.P, qs) + r1 * r1 - ss.w * ss.w;
        const double L1 = -2.0 * dot(t.Q, qs) + 2.0 * (r1 - ss.w);
        const double s4 = 4.0 * sigma * sigma;
        double r[2];
        const int nr = realRoots(L1 * L1 - s4 * t.A, 2.0 * L0 * L1 - s4 * t.B, L0 * L0 - s4 * t.Cc, r);
        if (nr < 0) return false;
        for (int m = 0; m < nr; m++) {
            if (!std::isfinite(r[m])) return false;
            roots.push_back(r[m]);
        }
    }
    roots.push_back(-t.B / (2.0 * t.A));
    if (!std::isfinite(roots.back())) return false;
    std::sort(roots.begin(), roots.end());

    // possibly-valid runs of intervals, split at the extremum of h^2
    struct Run { double a, b, sgn; };
    std::vector<Run> runs;
    const double wv = -t.B / (2.0 * t.A);
    for (const Span& pc : pieces) {
        std::vector<double> bp{pc.a};
        for (double r : roots)
            if (r > pc.a && r < pc.b) bp.push_back(r);
        bp.push_back(pc.b);
        for (double sgn : {1.0, -1.0}) {
            bool open = false;
            for (size_t m = 0; m + 1 < bp.size(); m++) {
                const double u = bp[m], v = bp[m + 1];
                bool keep;
                if (v == inf) {
                    const double ws = u + .5 * std::fmax(u + r1, 0.0) + 1e-6 * (1.0 + r1);
                    if (!clearlyInvalid(point(ws, sgn), 0.0)) return false;  // may run to infinity
                    keep = false;
                } else {
                    // A wide interval has one sign per neighbour; on a narrow one
                    // the computed roots may be off by more than its width, so
                    // the midpoint must clear every point of it: the spheres
                    // on [u, v] move by at most the displacement bound plus the
                    // change of w.
                    const double wm = .5 * (u + v), hm = hAt(wm);
                    double margin = 0.0;
                    if (v - u <= 1e-6 * (1.0 + std::fabs(u) + r1))
                        margin = len(t.Q) * (v - u) * .5 + std::fmax(std::fabs(hAt(u) - hm), std::fabs(hAt(v) - hm)) +
                                 (v - u) * .5;
                    keep = !clearlyInvalid(point(wm, sgn), margin);
                }
                if (keep && open && u != wv) {
                    runs.back().b = v;
                } else if (keep) {
                    runs.push_back({u, v, sgn});
                }
                open = keep;
            }
        }
    }
    if (runs.empty()) return false;
    const double m1 = std::fmax(std::fabs(si.x), std::fmax(std::fabs(si.y), std::fabs(si.z)));
    if (int(runs.size()) > kCoverBalls) {
        // one ball around s_i reaching the largest accepted radius
        double W = -DBL_MAX;
        for (const Run& r : runs) W = std::fmax(W, r.b);
        cov.ball[0] = {float(si.x), float(si.y), float(si.z),
                       float((r1 + 2.0 * W) * (1.0 + 1e-4) + 1e-6 * (1.0 + m1))};
        cov.count = 1;
        return true;
    }
    const int split = kCoverBalls / int(runs.size());
    const double qn = len(t.Q);
    for (const Run& r : runs) {
        for (int m = 0; m < split; m++) {
            const double wa = r.a + (r.b - r.a) * double(m) / double(split);
            const double wb = (m + 1 == split) ? r.b : r.a + (r.b - r.a) * double(m + 1) / double(split);
            const double wm = .5 * (wa + wb);
            const D4 cm = point(wm, r.sgn);
            const double hm = hAt(wm);
            const double rad = qn * (wb - wa) * .5 + std::fmax(std::fabs(hAt(wa) - hm), std::fabs(hAt(wb) - hm)) + wb;
            const double mc = std::fmax(std::fabs(cm.x), std::fmax(std::fabs(cm.y), std::fabs(cm.z)));
            cov.ball[cov.count++] = {float(cm.x), float(cm.y), float(cm.z),
                                     float(rad * (1.0 + 1e-4) + 1e-6 * (1.0 + mc + std::fabs(rad)))};
        }
    }
    return true;
}

// Completion mode: can some point of box b lie strictly inside the
// half-space n.(x - p) < 0?  The box's lowest point along n is the corner
// taking the lower bound where n is positive, compared with a margin above
// the FP32 rounding of the three-term dot product, so no box that holds a
// site the leaf test accepts is skipped.  (The reference tests the box's
// point closest to the search sphere's centre, algorithm.inl:660-670, which
// can lie outside the half-space while part of the box lies inside.)
inline bool boxReachesHalfSpace(const Box& b, F3 n, F3 p) {
    const float dx = (n.x > 0.f ? b.mn.x : b.mx.x) - p.x;
    const float dy = (n.y > 0.f ? b.mn.y : b.mx.y) - p.y;
    const float dz = (n.z > 0.f ? b.mn.z : b.mx.z) - p.z;
    const float ex = n.x * dx, ey = n.y * dy, ez = n.z * dz;
    const float margin = 0x1p-20f * (std::fabs(ex) + std::fabs(ey) + std::fabs(ez));
    return ex + ey + ez < margin;
}

// Contributor search around an edge (algorithm.inl:616-793)
uint32_t edgeSearch(const Context& C, const Cell& cell, uint32_t i, const D4& si, uint32_t k, uint32_t l,
                    const EdgeQuery& q, uint32_t& flags) {
    const Lbvh& b = *C.bvh;
    const uint32_t n = b.n;
    std::vector<uint32_t> stack;
    uint32_t node = 0;
    while (true) {
        while (true) {
            if (node >= n - 1) break;
            const uint32_t L = uint32_t(b.left[node]), R = uint32_t(b.right[node]);
            const Box lb = b.box(L), rb = b.box(R);
            const F3 cpL = closestPoint(lb, q.bs), cpR = closestPoint(rb, q.bs);
            const float dL = boxDistance(lb, q.bs), dR = boxDistance(rb, q.bs);
            const bool lc = dL < dR;
            const float minFar = lc ? dR : dL;
            const F3 farPoint = lc ? cpR : cpL;
            const uint32_t farChild = lc ? R : L;
            const float minClose = lc ? dL : dR;
            const F3 closePoint = lc ? cpL : cpR;
            const uint32_t closeChild = lc ? L : R;
            bool ef = minFar < 0.f;
            bool ec = minClose < 0.f;
            if (q.hyperbola && C.cfg.complete) {
                const Box& fb = lc ? rb : lb;
                const Box& cb = lc ? lb : rb;
                ef = ef && boxReachesHalfSpace(fb, q.n1, q.pp1) && boxReachesHalfSpace(fb, q.n2, q.pp2);
                ec = ec && boxReachesHalfSpace(cb, q.n1, q.pp1) && boxReachesHalfSpace(cb, q.n2, q.pp2);
            } else if (q.hyperbola) {
                ef = ef && dotf(q.n1, subf(farPoint, q.pp1)) < 0.f && dotf(q.n2, subf(farPoint, q.pp2)) < 0.f;
                ec = ec && dotf(q.n1, subf(closePoint, q.pp1)) < 0.f && dotf(q.n2, subf(closePoint, q.pp2)) < 0.f;
            }
            if (q.cover) {
                const Box fb = lc ? rb : lb, cb = lc ? lb : rb;
                const float fr = b.maxr[farChild], cr = b.maxr[closeChild];
                bool tf = false, tc = false;
                for (int m = 0; m < q.cover; m++) {
                    tf = tf || boxDistance(fb, q.ball[m]) <= fr;
                    tc = tc || boxDistance(cb, q.ball[m]) <= cr;
                }
                ef = ef && tf;
                ec = ec && tc;
            }
            if (!ef && !ec) break;
            node = ec ? closeChild : farChild;
            if (ec && ef) {
                stack.push_back(farChild);
                if (stack.size() > 32) flags |= kFlagStackOverflow;
            }
        }
        if (node >= n - 1 && node < 2 * n - 1) {
            const uint32_t j = node - (n - 1);
            bool cc = j != i && j != k && j != l && !C.buried[j];
            if (cc) {
                const F4 fsj = C.Sf[j];
                cc = intersectF(fsj, q.bs);
                if (q.hyperbola) {
                    const F3 cj = xyzf(fsj);
                    cc = cc && dotf(q.n1, subf(cj, q.pp1)) < -fsj.w && dotf(q.n2, subf(cj, q.pp2)) < -fsj.w;
                }
            }
            if (cc && known(cell, j)) cc = false;
            if (cc) {
                D4 x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
                uint32_t mask = quadrisector(si, C.S[j], C.S[k], C.S[l], x[0], x[1]);
                if (mask != 3) mask = 0;
                uint32_t valid = mask;
                for (uint32_t nv = 0; nv < 2; nv++)
                    if ((mask >> nv) & 1u)
                        if (!edgeSitesClear(C, cell, x[nv], k, l)) valid &= ~(1u << nv);
                if (valid) return j;
            }
        }
        if (stack.empty()) break;
        node = stack.back();
        stack.pop_back();
    }
    return kNone;
}

// Contributor search along a bisector (algorithm.inl:1019-1176)
uint32_t bisectorSearch(const Context& C, const Cell& cell, uint32_t i, const D4& si, uint32_t bk, F3 fci, F3 fck,
                        F4 bs, float cylR, bool axisPrune, uint32_t& flags) {
    const Lbvh& b = *C.bvh;
    const uint32_t n = b.n;
    const D4 sk = C.S[bk];
    std::vector<uint32_t> stack;
    uint32_t node = 0;
    while (true) {
        while (true) {
            if (node >= n - 1) break;
            const uint32_t L = uint32_t(b.left[node]), R = uint32_t(b.right[node]);
            const float dL = boxDistance(b.box(L), bs), dR = boxDistance(b.box(R), bs);
            const bool lc = dL < dR;
            const float minFar = lc ? dR : dL;
            const uint32_t farChild = lc ? R : L;
            const float minClose = lc ? dL : dR;
            const uint32_t closeChild = lc ? L : R;
            bool ef = minFar < 0.f, ec = minClose < 0.f;
            // subtrees that cannot hold a cylinder candidate (no candidate
            // is lost, the rest keep their order: the same result)
            if (axisPrune && (ef || ec)) {
                const F3 u = normalizef(subf(fck, fci));
                if (ef) ef = boxNearAxis(b.box(farChild), fci, u, cylR);
                if (ec) ec = boxNearAxis(b.box(closeChild), fci, u, cylR);
            }
            if (!ef && !ec) break;
            node = ec ? closeChild : farChild;
            if (ec && ef) {
                stack.push_back(farChild);
                if (stack.size() > 32) flags |= kFlagStackOverflow;
            }
        }
        if (node >= n - 1 && node < 2 * n - 1) {
            const uint32_t j = node - (n - 1);
            bool cc = j != i && j != bk;
            if (cc) {
                const F4 fsj = C.Sf[j];
                const F3 fcj = xyzf(fsj);
                cc = bs.w >= fsj.w && (lenf(subf(xyzf(bs), fcj)) <= (bs.w - fsj.w));
                if (cc) {
                    const F3 itoj = subf(fcj, fci);
                    const float v = dotf(itoj, normalizef(subf(fck, fci)));
                    const float d = std::sqrt(dotf(itoj, itoj) - v * v);
                    cc = cc && (d + fsj.w) < cylR;
                }
            }
            if (cc && known(cell, j)) cc = false;
            if (cc) {
                D4 x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
                const uint32_t cnt = trisector(si, C.S[j], sk, &x[0], &x[1]);
                cc = cnt == 2 && edgeSitesClear(C, cell, x[1], bk, bk);
                if (cc) return j;
            }
        }
        if (stack.empty()) break;
        node = stack.back();
        stack.pop_back();
    }
    return kNone;
}

// ---------------------------------------------------------------------------
// Pipeline stages (algorithm.inl:101-1233)
// ---------------------------------------------------------------------------
// getCells.  The first pass starts every cell from the octahedron (the
// state initializeDataStructure leaves, topology_update.inl:5-47); later
// passes continue the stored cells that are still in validation.
void stageCells(Context& C, const std::vector<uint32_t>& knn, uint32_t K, bool resume) {
    const int64_t n = C.n;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < n; ii++) {
        const uint32_t i = uint32_t(ii);
        Cell& cell = C.cells[i];
        if (!resume) initCell(C, cell);
        if (C.status[i] == kCertified || C.status[i] == kBuried) continue;
        const D4 si = C.S[i];
        double knownRadius = -DBL_MAX;
        uint32_t flags = 0;
        const uint32_t use = C.siteK.empty() ? K : std::min(K, C.siteK[i]);
        uint32_t count = 0;
        while (count < K && knn[size_t(i) * K + count] != kNone) count++;
        for (uint32_t nId = 0; nId < use; nId++) {
            const uint32_t j = knn[size_t(i) * K + nId];
            if (j == kNone) break;
            if (C.buried[j]) continue;
            const D4 sj = C.S[j];
            updateCell(C, i, si, cell.V, cell, j, sj, flags);
            knownRadius = std::fmax(knownRadius, sphereDistD(si, sj));
        }
        if (!C.range.empty() && count < K) knownRadius = std::fmax(knownRadius, rangeKnownRadius(C.range[i], si));
        const double maxR = cellRadius(C, si, cell);
        if (gtD(knownRadius, 2.0 * maxR)) C.status[i] = kCertified;
        C.knownRadii[i] = float(knownRadius);
        C.flags[i] |= flags;
    }
}

void stageVertexValidation(Context& C, const std::vector<uint32_t>& tasks) {
    const int64_t T = int64_t(tasks.size());
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < T; t++) {
        const uint32_t i = tasks[size_t(t)];
        if (C.status[i] == kCertified || C.status[i] == kBuried) continue;
        Cell& cell = C.cells[i];
        const float kr = C.knownRadii[i];
        const D4 si = C.S[i];
        uint32_t flags = 0;
        uint32_t vs = 0, restarts = 0;
        const int W = waveWidth(C);
        while (vs < uint32_t(cell.V.size()) / uint32_t(W) + 1 && restarts < C.cfg.max_restarts) {
            const uint32_t base = vs * uint32_t(W);
            vs++;
            uint32_t invalid[kWave];
            int nInvalid = 0;
            for (int lane = 0; lane < W; lane++) {
                const uint32_t vid = base + uint32_t(lane);
                if (vid >= cell.V.size() || cell.V[vid].status != 0) continue;
                const D4 cur = vertexPosition(C, si, cell.V[vid]);
                if (2.0 * cur.w < double(kr)) {
                    cell.V[vid].status = 1;
                    continue;
                }
                const uint32_t nb = vertexSearch(C, cell, i, cell.V[vid], toF(cur), flags);
                cell.V[vid].status = 1;
                if (nb != kNone && intersectD(cur, C.S[nb])) invalid[nInvalid++] = nb;
            }
            for (int t3 = 0; t3 < nInvalid; t3++) {
                const uint32_t j = invalid[laneOrder(C, t3, nInvalid)];
                if (known(cell, j)) continue;
                updateCell(C, i, si, cell.V, cell, j, C.S[j], flags);
            }
            if (nInvalid) {
                vs = 0;
                restarts++;
            }
        }
        if (restarts >= C.cfg.max_restarts) flags |= kFlagRestartCap;
        C.flags[i] |= flags;
    }
}

void stageEdgeValidation(Context& C, const std::vector<uint32_t>& tasks) {
    const int64_t T = int64_t(tasks.size());
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < T; t++) {
        const uint32_t i = tasks[size_t(t)];
        if (C.status[i] == kCertified || C.status[i] == kBuried) continue;
        Cell& cell = C.cells[i];
        const float kr = C.knownRadii[i];
        const D4 si = C.S[i];
        const F4 fsi = toF(si);
        uint32_t flags = 0;
        for (Edg& e : cell.E)
            if (e.status != 1) e.status = 2;
        // the reference updates a vertex list that starts empty
        // (algorithm.inl:532); the completion mode the whole list
        std::vector<Vtx> Vown;
        std::vector<Vtx>& Vloc = C.cfg.complete ? cell.V : Vown;
        bool didChange = false;
        const int W = waveWidth(C);
        const uint32_t stepNb = uint32_t(cell.E.size()) / uint32_t(W) + 1;
        uint32_t es = 0, restarts = 0;
        while (es < stepNb) {
            uint32_t found[kWave];
            int nFound = 0;
            for (int lane = 0; lane < W; lane++) {
                const uint32_t eid = es * uint32_t(W) + uint32_t(lane);
                const bool inRange = eid < cell.E.size();
                bool needSearch = inRange && cell.E[eid].status == 2;
                uint32_t k = kNone, l = kNone;
                if (inRange) {
                    k = cell.E[eid].j;
                    l = cell.E[eid].k;
                }
                EdgeQuery q{};
                if (needSearch) {
                    const D4 sk = C.S[k], sl = C.S[l];
                    D4 ev[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
                    const uint32_t type = trisector(si, sk, sl, &ev[0], &ev[1]);
                    q.hyperbola = type == 1;
                    if (type == 2) {
                        const F4 v1 = toF(ev[0]), v2 = toF(ev[1]);
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
                    cell.E[eid].status = 1;
                    needSearch = false;
                }
                if (!needSearch) continue;
                if (C.cfg.edge_bound == 2 || (C.cfg.edge_bound == 1 && q.bs.w > kCoverRatio * kr)) {
                    EdgeCover cov;
                    if (edgeCover(C, cell, si, k, l, cov)) {
                        q.cover = cov.count;
                        for (int m = 0; m < cov.count; m++) q.ball[m] = cov.ball[m];
                    }
                }
                const uint32_t c = edgeSearch(C, cell, i, si, k, l, q, flags);
                if (c == kNone) cell.E[eid].status = 1;
                else found[nFound++] = c;
            }
            bool changed = false;
            for (int t3 = 0; t3 < nFound; t3++) {
                const uint32_t j = found[laneOrder(C, t3, nFound)];
                if (known(cell, j)) continue;
                const bool ch = updateCell(C, i, si, Vloc, cell, j, C.S[j], flags);
                didChange = didChange || ch;
                changed = changed || ch;
            }
            if (changed && restarts < C.cfg.max_restarts) {
                es = 1;
                restarts++;
            } else {
                if (changed) flags |= kFlagRestartCap;
                es++;
            }
        }
        C.status[i] = didChange ? 0 : 1;
        for (const Vtx& v : Vown) cell.V.push_back(v);
        if (cell.V.size() > C.cfg.vertex_capacity) {
            flags |= kFlagVertexOverflow;
            cell.V.resize(C.cfg.vertex_capacity);
        }
        C.flags[i] |= flags;
    }
}

void stageBisectorValidation(Context& C, uint64_t& changedCells) {
    const int64_t n = C.n;
    std::atomic<uint64_t> changedCount{0};
    // the subtree test assumes no negative radius (the scene's minimum)
    double rmin = DBL_MAX;
    for (int64_t j = 0; j < n; j++) rmin = std::fmin(rmin, C.S[j].w);
    const bool axisPrune = C.cfg.bisector_bound != 0 && rmin >= 0.0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < n; ii++) {
        const uint32_t i = uint32_t(ii);
        if (C.status[i] == kBuried) continue;
        Cell& cell = C.cells[i];
        const float kr = C.knownRadii[i];
        const D4 si = C.S[i];
        uint32_t flags = 0;
        std::vector<uint32_t> bis;
        for (const Edg& e : cell.E) {
            if (e.j < C.n && e.j > i) bis.push_back(e.j);
            if (e.k < C.n && e.k > i) bis.push_back(e.k);
        }
        std::sort(bis.begin(), bis.end());
        bis.erase(std::unique(bis.begin(), bis.end()), bis.end());
        if (bis.size() > C.cfg.bisector_capacity) {
            flags |= kFlagBisectorOverflow;
            bis.resize(C.cfg.bisector_capacity);
        }
        const F3 fci{float(si.x), float(si.y), float(si.z)};
        const float fri = float(si.w);
        std::vector<Vtx> Vown;
        std::vector<Vtx>& Vloc = C.cfg.complete ? cell.V : Vown;
        bool didChange = false;
        const int W = waveWidth(C);
        const uint32_t stepNb = uint32_t(bis.size()) / uint32_t(W) + 1;
        uint32_t bsi = 0, restarts = 0;
        while (bsi < stepNb) {
            uint32_t found[kWave];
            int nFound = 0;
            for (int lane = 0; lane < W; lane++) {
                const uint32_t bid = bsi * uint32_t(W) + uint32_t(lane);
                if (bid >= bis.size() || bis[bid] == kNone) continue;
                const uint32_t bk = bis[bid];
                const D4 sk = C.S[bk];
                const F3 fck{float(sk.x), float(sk.y), float(sk.z)};
                const float frk = float(sk.w);
                const float dist = lenf(subf(fci, fck)) + fri + frk;
                if (ltF(dist, kr)) {
                    bis[bid] = kNone;
                    continue;
                }
                const F3 center = mulf(addf(fci, fck), .5f);
                const float radius = lenf(subf(center, fci)) + std::fmax(fri, frk);
                const F4 bsph{center.x, center.y, center.z, radius};
                const float cylR = std::fmax(fri, frk);
                const uint32_t c = bisectorSearch(C, cell, i, si, bk, fci, fck, bsph, cylR, axisPrune, flags);
                if (c == kNone) bis[bid] = kNone;
                else found[nFound++] = c;
            }
            bool changed = false;
            for (int t3 = 0; t3 < nFound; t3++) {
                const uint32_t j = found[laneOrder(C, t3, nFound)];
                if (known(cell, j)) continue;
                const bool ch = updateCell(C, i, si, Vloc, cell, j, C.S[j], flags);
                didChange = didChange || ch;
                changed = changed || ch;
            }
            if (changed && restarts < C.cfg.max_restarts) {
                bsi = 1;
                restarts++;
            } else {
                if (changed) flags |= kFlagRestartCap;
                bsi++;
            }
        }
        C.status[i] = didChange ? 0 : 1;
        if (didChange) changedCount++;
        for (const Vtx& v : Vown) cell.V.push_back(v);
        if (cell.V.size() > C.cfg.vertex_capacity) {
            flags |= kFlagVertexOverflow;
            cell.V.resize(C.cfg.vertex_capacity);
        }
        C.flags[i] |= flags;
    }
    changedCells = changedCount.load();
}

// The vertex + edge validation loop of AlgorithmGPU::build (algorithm.inl:
// 1551-1593): rounds over the cells in validation (status 0) while there are
// any, at most max_iterations.  Adds the round task counts to *tasks;
// returns the rounds run.
uint32_t validationRounds(Context& C, uint64_t& tasksTotal) {
    const uint32_t n = C.n;
    uint32_t it = 0;
    for (; it < C.cfg.max_iterations; it++) {
        std::vector<uint32_t> tasks;
        for (uint32_t i = 0; i < n; i++)
            if (C.status[i] == 0) tasks.push_back(i);
        if (tasks.empty()) break;
        tasksTotal += tasks.size();
        stageVertexValidation(C, tasks);
        stageEdgeValidation(C, tasks);
    }
    return it;
}

// Records of cell c whose three sites are {a, b, d} (ascending).
inline uint32_t recordCount(const Cell& c, uint32_t a, uint32_t b, uint32_t d) {
    uint32_t m = 0;
    for (const Vtx& v : c.V) m += v.j == a && v.k == b && v.l == d;
    return m;
}

// One consistency round of the completion mode (no counterpart in the
// reference).  A vertex of the diagram is tangent to 4 sites and belongs to
// the 4 cells; a record (i; j, k, l) of cell i asks cells j, k and l for as
// many records of the same quadruple.  A cell holding fewer wants the
// quadruple sites it does not know; each cell is then updated with the
// smallest site it wants (one per round, so the result does not depend on
// the order of the requests) and goes back to validation when it changes.
// A cell that knows all three sites yet holds fewer records is stuck (the
// order in which its sites were inserted lost an edge that no update
// creates again): the record goes to the round's stuck list, and a stuck
// cell with no wanted site is rebuilt from its octahedron, first with the
// sites of its stuck records (ascending), then with its other real,
// unburied neighbours nearest first (FP32 gap, ties by index).  A cell is
// rebuilt at most once per build; a round whose stuck list exceeds
// kStuckCap rebuilds nothing, nor does a cell whose stuck records hold more
// than kRebuildSites sites (with repeats).  Buried cells, artificial sites
// and cells with a capacity flag take no part.  *stuckCount receives the
// round's stuck records, *rebuilt is increased by the cells rebuilt.
// Returns the cells the round changed.
uint64_t consistencyRound(Context& C, uint64_t& stuckCount, uint64_t& rebuilt) {
    const uint32_t n = C.n;
    auto part = [&](uint32_t m) { return C.status[m] != kBuried && C.flags[m] == 0u; };
#ifdef _OPENMP
    const int threads = omp_get_max_threads();
#else
    const int threads = 1;
#endif
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> req(threads);
    std::vector<std::vector<std::array<uint32_t, 4>>> stuckT(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < int64_t(n); ii++) {
        const uint32_t i = uint32_t(ii);
        if (!part(i)) continue;
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        const Cell& ci = C.cells[i];
        for (const Vtx& v : ci.V) {
            if (v.l >= n) continue;  // j < k < l: an artificial site is last
            const uint32_t own = recordCount(ci, v.j, v.k, v.l);
            const uint32_t q[4] = {i, v.j, v.k, v.l};
            for (int t = 1; t < 4; t++) {
                const uint32_t m = q[t];
                if (!part(m)) continue;
                uint32_t o[3], c = 0;
                for (int u = 0; u < 4; u++)
                    if (u != t) o[c++] = q[u];
                std::sort(o, o + 3);
                if (recordCount(C.cells[m], o[0], o[1], o[2]) >= own) continue;
                bool any = false;
                for (int u = 0; u < 3; u++)
                    if (!known(C.cells[m], o[u])) {
                        req[tid].push_back({m, o[u]});
                        any = true;
                    }
                if (!any) {
                    stuckT[tid].push_back({m, o[0], o[1], o[2]});
                    req[tid].push_back({m, kWantRebuild});
                }
            }
        }
    }
    std::vector<std::array<uint32_t, 4>> stuck;
    for (int t = 0; t < threads; t++) {
        for (const auto& r : req[t]) C.want[r.first] = std::min(C.want[r.first], r.second);
        stuck.insert(stuck.end(), stuckT[t].begin(), stuckT[t].end());
    }
    std::sort(stuck.begin(), stuck.end());
    stuckCount = stuck.size();
    const bool rebuildOk = stuck.size() <= kStuckCap;
    std::atomic<uint64_t> updated{0}, nRebuilt{0};
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t mm = 0; mm < int64_t(n); mm++) {
        const uint32_t m = uint32_t(mm);
        const uint32_t j = C.want[m];
        if (j == kNone || j == kWantRebuilt) continue;
        Cell& cell = C.cells[m];
        uint32_t flags = 0;
        if (j == kWantRebuild) {
            // the stuck records of m are contiguous in the sorted list
            const auto lo = std::lower_bound(stuck.begin(), stuck.end(), std::array<uint32_t, 4>{m, 0, 0, 0});
            std::vector<uint32_t> first;
            for (auto it = lo; it != stuck.end() && (*it)[0] == m; ++it)
                for (int u = 1; u < 4; u++) first.push_back((*it)[u]);
            if (!rebuildOk || first.size() > kRebuildSites) {
                C.want[m] = kNone;
                continue;
            }
            std::sort(first.begin(), first.end());
            first.erase(std::unique(first.begin(), first.end()), first.end());
            std::vector<std::pair<float, uint32_t>> rest;
            for (const Edg& e : cell.E)
                for (uint32_t x : {e.j, e.k})
                    if (x < n && C.status[x] != kBuried && !std::binary_search(first.begin(), first.end(), x))
                        rest.push_back({sphereDistF(C.Sf[m], C.Sf[x]), x});
            std::sort(rest.begin(), rest.end());
            rest.erase(std::unique(rest.begin(), rest.end()), rest.end());
            initCell(C, cell);
            for (uint32_t x : first) updateCell(C, m, C.S[m], cell.V, cell, x, C.S[x], flags);
            for (const auto& pr : rest) updateCell(C, m, C.S[m], cell.V, cell, pr.second, C.S[pr.second], flags);
            C.flags[m] |= flags;
            C.status[m] = 0;
            C.want[m] = kWantRebuilt;
            updated++;
            nRebuilt++;
            continue;
        }
        const bool changed = updateCell(C, m, C.S[m], cell.V, cell, j, C.S[j], flags);
        C.flags[m] |= flags;
        C.want[m] = kNone;
        if (changed) {
            C.status[m] = 0;
            updated++;
        }
    }
    rebuilt += nRebuilt.load();
    return updated.load();
}

// pcg hash of the reference (apo/src/core/utils.cpp:6-12)
inline uint32_t pcgWord(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

void artificialSites(const SceneBox& sb, D4 out[6]) {
    const D3 bbMin{sb.mn[0], sb.mn[1], sb.mn[2]}, bbMax{sb.mx[0], sb.mx[1], sb.mx[2]};
    const D3 cc{(bbMin.x + bbMax.x) / 2.0, (bbMin.y + bbMax.y) / 2.0, (bbMin.z + bbMax.z) / 2.0};
    const double aabbRadius = len(sub(cc, bbMax));
    const double radius = aabbRadius * (std::sqrt(3.0) + 1e-1);
    const double w = sb.rmin;
    const D4 base{cc.x, cc.y, cc.z, 0.0};
    const D4 offs[6] = {{radius, 0, 0, w}, {-radius, 0, 0, w}, {0, radius, 0, w},
                        {0, -radius, 0, w}, {0, 0, radius, w}, {0, 0, -radius, w}};
    for (int s = 0; s < 6; s++) {
        double v[4] = {base.x + offs[s].x, base.y + offs[s].y, base.z + offs[s].z, base.w + offs[s].w};
        // joggle(range = 0.1), utils.cpp:15-25
        for (int c = 0; c < 4; c++) v[c] += std::ldexp(double(pcgWord(uint32_t(s * 4 + c))), -32) * (2.0 * 1e-1) + (-1e-1);
        out[s] = {v[0], v[1], v[2], v[3]};
    }
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
struct orc_diagram {
    Context C;
    Lbvh bvh;
    std::vector<uint32_t> knn;
    std::vector<double> outXyzw;
    std::vector<uint32_t> outIds;
    std::vector<double> edgeXyzw;
    std::vector<uint32_t> edgeIds;
    uint64_t stats[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

extern "C" {

double orc_pcg_unit(uint32_t v) { return double(pcgWord(v)) / double(std::numeric_limits<uint32_t>::max()); }
double orc_pcg_ldexp(uint32_t v) { return std::ldexp(double(pcgWord(v)), -32); }

uint32_t orc_quadrisector(const double* s1, const double* s2, const double* s3, const double* s4, double* x1,
                          double* x2) {
    D4 a{s1[0], s1[1], s1[2], s1[3]}, b{s2[0], s2[1], s2[2], s2[3]}, c{s3[0], s3[1], s3[2], s3[3]},
        d{s4[0], s4[1], s4[2], s4[3]};
    D4 o1{0, 0, 0, 0}, o2{0, 0, 0, 0};
    const uint32_t m = quadrisector(a, b, c, d, o1, o2);
    std::memcpy(x1, &o1, sizeof(o1));
    std::memcpy(x2, &o2, sizeof(o2));
    return m;
}

uint32_t orc_trisector(const double* s1, const double* s2, const double* s3, double* x1, double* x2) {
    D4 a{s1[0], s1[1], s1[2], s1[3]}, b{s2[0], s2[1], s2[2], s2[3]}, c{s3[0], s3[1], s3[2], s3[3]};
    D4 o1{0, 0, 0, 0}, o2{0, 0, 0, 0};
    const uint32_t m = trisector(a, b, c, &o1, &o2);
    std::memcpy(x1, &o1, sizeof(o1));
    std::memcpy(x2, &o2, sizeof(o2));
    return m;
}

void orc_tangent_planes(const double* s1, const double* s2, const double* s3, double* n1, double* n2) {
    D4 a{s1[0], s1[1], s1[2], s1[3]}, b{s2[0], s2[1], s2[2], s2[3]}, c{s3[0], s3[1], s3[2], s3[3]};
    D3 o1, o2;
    tangentPlanes(a, b, c, o1, o2);
    std::memcpy(n1, &o1, sizeof(o1));
    std::memcpy(n2, &o2, sizeof(o2));
}

void orc_sphere_from_points(const double* p1, const double* p2, const double* p3, const double* p4, double* out) {
    const D4 s = sphereFromPoints(D3{p1[0], p1[1], p1[2]}, D3{p2[0], p2[1], p2[2]}, D3{p3[0], p3[1], p3[2]},
                                  D3{p4[0], p4[1], p4[2]});
    std::memcpy(out, &s, sizeof(s));
}

void orc_scene_box(const float* xyzr, uint32_t n, float* box6, double* minmax_radius) {
    const SceneBox sb = sceneBox(xyzr, n);
    for (int a = 0; a < 3; a++) {
        box6[a] = sb.fmn[a];
        box6[3 + a] = sb.fmx[a];
    }
    minmax_radius[0] = sb.rmin;
    minmax_radius[1] = sb.rmax;
}

uint64_t orc_morton64(uint32_t x, uint32_t y, uint32_t z) { return morton64(x, y, z); }

void orc_artificial_sites(const float* xyzr, uint32_t n, double* out24) {
    const SceneBox sb = sceneBox(xyzr, n);
    D4 art[6];
    artificialSites(sb, art);
    std::memcpy(out24, art, sizeof(art));
}

void orc_morton_codes(const float* xyzr, uint32_t n, const float* box6, uint64_t* codes) {
    for (uint32_t i = 0; i < n; i++) codes[i] = mortonOf(&xyzr[4 * size_t(i)], box6, box6 + 3);
}

int orc_lbvh_build(const float* xyzr, uint32_t n, uint64_t* sorted_codes, uint32_t* indices, int32_t* parent,
                   int32_t* left, int32_t* right, float* aabb, uint32_t* sizes, float* max_radius) {
    if (n < 2) return -1;
    const SceneBox sb = sceneBox(xyzr, n);
    Lbvh b;
    buildLbvh(xyzr, n, sb, b);
    const size_t nodes = 2 * size_t(n) - 1;
    std::memcpy(sorted_codes, b.codes.data(), n * sizeof(uint64_t));
    std::memcpy(indices, b.indices.data(), n * sizeof(uint32_t));
    std::memcpy(parent, b.parent.data(), nodes * sizeof(int32_t));
    std::memcpy(left, b.left.data(), nodes * sizeof(int32_t));
    std::memcpy(right, b.right.data(), nodes * sizeof(int32_t));
    std::memcpy(aabb, b.aabb.data(), nodes * 6 * sizeof(float));
    std::memcpy(sizes, b.sizes.data(), nodes * sizeof(uint32_t));
    std::memcpy(max_radius, b.maxr.data(), nodes * sizeof(float));
    return 0;
}

int orc_knn(const float* sorted_xyzr, uint32_t n, const int32_t* left, const int32_t* right, const float* aabb,
            uint32_t k, uint32_t* knn, float* knn_dist) {
    Lbvh b;
    b.n = n;
    const size_t nodes = 2 * size_t(n) - 1;
    b.left.assign(left, left + nodes);
    b.right.assign(right, right + nodes);
    b.aabb.assign(aabb, aabb + nodes * 6);
    std::vector<F4> sf(n);
    for (uint32_t i = 0; i < n; i++)
        sf[i] = {sorted_xyzr[4 * i], sorted_xyzr[4 * i + 1], sorted_xyzr[4 * i + 2], sorted_xyzr[4 * i + 3]};
    int maxDepth = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(max : maxDepth)
    for (int64_t ii = 0; ii < int64_t(n); ii++) {
        const uint32_t i = uint32_t(ii);
        const int dpt = knnQuery(b, sf, i, k, -FLT_MAX, &knn[size_t(i) * k], &knn_dist[size_t(i) * k]);
        maxDepth = std::max(maxDepth, dpt);
    }
    return maxDepth;
}

void orc_knn_brute(const float* s, uint32_t n, uint32_t k, uint32_t* knn, float* knn_dist) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < int64_t(n); ii++) {
        const uint32_t i = uint32_t(ii);
        const F4 si{s[4 * i], s[4 * i + 1], s[4 * i + 2], s[4 * i + 3]};
        std::vector<std::pair<float, uint32_t>> all;
        all.reserve(n);
        for (uint32_t j = 0; j < n; j++) {
            if (j == i) continue;
            const F4 sj{s[4 * j], s[4 * j + 1], s[4 * j + 2], s[4 * j + 3]};
            all.push_back({sphereDistF(si, sj), j});
        }
        std::sort(all.begin(), all.end());
        for (uint32_t t = 0; t < k; t++) {
            knn[size_t(i) * k + t] = t < all.size() ? all[t].second : kNone;
            knn_dist[size_t(i) * k + t] = t < all.size() ? all[t].first : FLT_MAX;
        }
    }
}

void orc_default_config(orc_config* cfg) {
    cfg->k = 16;
    cfg->vertex_capacity = 152;
    cfg->edge_capacity = 264;
    cfg->new_edge_capacity = 128;
    cfg->bisector_capacity = 512;
    cfg->max_iterations = 10;
    cfg->max_restarts = 100;
    cfg->threads = 0;
    cfg->stage_mask = 7;
    cfg->edge_bound = 0;
    cfg->fallback_k = 16;
    cfg->knn_passes = 1;
    cfg->complete = 0;
    cfg->lanes = 64;
    cfg->bisector_bound = 1;
}

static orc_diagram* buildDiagram(const float* xyzr, uint32_t n, const orc_config* cfgIn, const uint8_t* classes,
                                 const float* radii = nullptr) {
    if (n < 2) return nullptr;
    orc_config cfg;
    if (cfgIn) cfg = *cfgIn;
    else orc_default_config(&cfg);
#ifdef _OPENMP
    if (cfg.threads) omp_set_num_threads(int(cfg.threads));
#endif
    orc_diagram* d = new orc_diagram();
    Context& C = d->C;
    C.n = n;
    C.cfg = cfg;
    const SceneBox sb = sceneBox(xyzr, n);
    buildLbvh(xyzr, n, sb, d->bvh);
    C.bvh = &d->bvh;
    C.S.resize(size_t(n) + 6);
    C.Sf.resize(size_t(n) + 6);
    for (uint32_t i = 0; i < n; i++) {
        const float* s = &xyzr[4 * size_t(d->bvh.indices[i])];
        C.S[i] = {s[0], s[1], s[2], s[3]};
        C.Sf[i] = {s[0], s[1], s[2], s[3]};
    }
    D4 art[6];
    artificialSites(sb, art);
    for (int s = 0; s < 6; s++) {
        C.S[n + s] = art[s];
        C.Sf[n + s] = toF(art[s]);
    }
    C.status.assign(n, 0);
    C.knownRadii.assign(n, -FLT_MAX);
    C.cells.resize(n);
    C.flags.assign(n, 0);
    if (classes) {
        C.siteK.resize(n);
        for (uint32_t i = 0; i < n; i++) C.siteK[i] = proposalK(classes[d->bvh.indices[i]], cfg.fallback_k);
        cfg.k = 64;
        C.cfg.k = 64;
    }
    if (radii) {
        C.range.resize(n);
        for (uint32_t i = 0; i < n; i++) C.range[i] = radii[d->bvh.indices[i]];
        cfg.k = 64;
        C.cfg.k = 64;
    }
    const uint32_t K = cfg.k;
    // the learned paths run one pass of their own lists
    const uint32_t passes = (classes || radii) ? 1u : cfg.knn_passes;

    // getKnns + getCells, knnValidationStep times (algorithm.inl:32-206,
    // 1535-1549).  The lists start as 0xffffffff (the memset of :1526).
    d->knn.assign(size_t(n) * K, kNone);
    C.buried.assign(n, 0);
    std::vector<float> dist(size_t(n) * K);
    for (uint32_t pass = 0; pass < passes; pass++) {
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t ii = 0; ii < int64_t(n); ii++) {
            const uint32_t i = uint32_t(ii);
            uint32_t* keys = &d->knn[size_t(i) * K];
            float startingDistance = -FLT_MAX;
            if (pass > 0) {
                // continue after the previous pass's last neighbour, or end the
                // list when that pass did not fill it (algorithm.inl:46-60)
                const uint32_t last = keys[K - 1];
                if (last >= n) {
                    for (uint32_t kk = 0; kk < K; kk++) keys[kk] = kNone;
                    continue;
                }
                startingDistance = sphereDistF(C.Sf[last], C.Sf[i]);
            }
            const float top = C.range.empty() ? FLT_MAX : rangeTop(C.range[i], C.Sf[i].w);
            const int depth = knnQuery(d->bvh, C.Sf, i, K, startingDistance, keys, &dist[size_t(i) * K], top);
            if (depth > 32) C.flags[i] |= kFlagStackOverflow;
            if (pass > 0 || keys[0] == kNone) continue;
            // buried test against the nearest site, first pass only (:85-98)
            const F4 si = C.Sf[i];
            const D4 closest = C.S[keys[0]];
            const double ri = si.w, rj = closest.w;
            const double bigger = std::fmax(ri, rj), smaller = std::fmin(ri, rj);
            const D3 dc{double(si.x) - closest.x, double(si.y) - closest.y, double(si.z) - closest.z};
            if (ltD(len(dc), bigger - smaller)) {
                if (ltD(rj, ri)) continue;
                C.status[i] = kBuried;
            }
        }
        if (pass == 0) {
            uint64_t buried = 0;
            for (uint32_t i = 0; i < n; i++) {
                C.buried[i] = C.status[i] == kBuried;
                buried += C.buried[i];
            }
            d->stats[5] = buried;
        }
        if (cfg.stage_mask & 1u) stageCells(C, d->knn, K, pass > 0);
        else if (pass == 0)
            for (uint32_t i = 0; i < n; i++) initCell(C, C.cells[i]);
    }
    if (passes == 0)
        for (uint32_t i = 0; i < n; i++) initCell(C, C.cells[i]);

    uint64_t uncertified = 0;
    for (uint32_t i = 0; i < n; i++) uncertified += C.status[i] == 0;
    d->stats[0] = uncertified;

    if (cfg.stage_mask & 2u) d->stats[2] = validationRounds(C, d->stats[1]);
    if (cfg.stage_mask & 4u) {
        uint64_t changed = 0;
        stageBisectorValidation(C, changed);
        d->stats[6] = changed;
    }
    if (cfg.complete && (cfg.stage_mask & 2u)) {
        // the cells the bisector stage changed (status 0) are validated again,
        // then the consistency rounds
        validationRounds(C, d->stats[10]);
        // every round's validation rounds run (over the cells still in
        // validation); a round after one that updated no cell checks nothing
        uint64_t updated = 1;
        C.want.assign(n, kNone);
        for (uint32_t round = 0; round < cfg.complete; round++) {
            if (updated) {
                d->stats[7]++;
                updated = consistencyRound(C, d->stats[9], d->stats[11]);
                d->stats[8] += updated;
            }
            validationRounds(C, d->stats[10]);
        }
    }
    uint64_t overflowed = 0;
    for (uint32_t i = 0; i < n; i++) overflowed += C.flags[i] != 0;
    d->stats[3] = overflowed;

    // toHost: getWriteableVertices + exclusive scan + saveVertices (algorithm.inl:1235-1307)
    for (uint32_t i = 0; i < n; i++) {
        if (C.status[i] == kBuried) continue;
        const D4 si = C.S[i];
        for (const Vtx& v : C.cells[i].V) {
            if (i > v.j || v.j >= n || v.k >= n || v.l >= n) continue;
            D4 x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            quadrisector(si, C.S[v.j], C.S[v.k], C.S[v.l], x[0], x[1]);
            const D4 p = x[v.type];
            d->outXyzw.insert(d->outXyzw.end(), {p.x, p.y, p.z, p.w});
            d->outIds.insert(d->outIds.end(), {d->bvh.indices[i], d->bvh.indices[v.j], d->bvh.indices[v.k],
                                               d->bvh.indices[v.l]});
        }
    }
    d->stats[4] = d->outIds.size() / 4;

    // toHostFull's closed edges: getWriteableClosedEdges + saveClosedEdges
    // (algorithm.inl:1309-1383).  Both extreme points start at zero where the
    // reference leaves them uninitialised.
    for (uint32_t i = 0; i < n; i++) {
        if (C.status[i] == kBuried) continue;
        const D4 si = C.S[i];
        for (const Edg& e : C.cells[i].E) {
            if (e.vertexNb != kNone || i > e.j || e.j >= n || e.k >= n) continue;
            D4 x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
            trisector(si, C.S[e.j], C.S[e.k], &x[0], &x[1]);
            if (x[1].w < x[0].w) x[0] = x[1];
            d->edgeXyzw.insert(d->edgeXyzw.end(), {x[0].x, x[0].y, x[0].z, x[0].w});
            d->edgeIds.insert(d->edgeIds.end(), {d->bvh.indices[i], d->bvh.indices[e.j], d->bvh.indices[e.k], 0u});
        }
    }
    return d;
}

orc_diagram* orc_build(const float* xyzr, uint32_t n, const orc_config* cfgIn) {
    return buildDiagram(xyzr, n, cfgIn, nullptr);
}

// The learned-proposal path with given classes (input order): kNN with 64
// neighbours, each cell runs its class's neighbour count, then the plain
// certification and validation.
orc_diagram* orc_build_proposal(const float* xyzr, uint32_t n, const orc_config* cfgIn, const uint8_t* classes) {
    return buildDiagram(xyzr, n, cfgIn, classes);
}

// The range path with given search radii (input order): 64-long lists of the
// sites whose gap is below the radius, each cell updated with its whole list,
// and the radius as the known radius when the list is not full.
orc_diagram* orc_build_radii(const float* xyzr, uint32_t n, const orc_config* cfgIn, const float* radii) {
    return buildDiagram(xyzr, n, cfgIn, nullptr, radii);
}

void orc_free(orc_diagram* d) { delete d; }

uint64_t orc_vertex_count(const orc_diagram* d) { return d->outIds.size() / 4; }

void orc_get_vertices(const orc_diagram* d, double* xyzw, uint32_t* ids) {
    std::memcpy(xyzw, d->outXyzw.data(), d->outXyzw.size() * sizeof(double));
    std::memcpy(ids, d->outIds.data(), d->outIds.size() * sizeof(uint32_t));
}

uint64_t orc_closed_edge_count(const orc_diagram* d) { return d->edgeIds.size() / 4; }

void orc_get_closed_edges(const orc_diagram* d, double* xyzw, uint32_t* ids) {
    std::memcpy(xyzw, d->edgeXyzw.data(), d->edgeXyzw.size() * sizeof(double));
    std::memcpy(ids, d->edgeIds.data(), d->edgeIds.size() * sizeof(uint32_t));
}

void orc_get_sorted_sites(const orc_diagram* d, float* xyzr, uint32_t* indices) {
    const uint32_t n = d->C.n;
    for (uint32_t i = 0; i < n; i++) {
        const F4 s = d->C.Sf[i];
        xyzr[4 * i] = s.x;
        xyzr[4 * i + 1] = s.y;
        xyzr[4 * i + 2] = s.z;
        xyzr[4 * i + 3] = s.w;
        indices[i] = d->bvh.indices[i];
    }
}

void orc_get_cell_counts(const orc_diagram* d, uint32_t* vertex_nb, uint32_t* edge_nb, uint8_t* status,
                         float* known_radius, uint32_t* flags) {
    const uint32_t n = d->C.n;
    for (uint32_t i = 0; i < n; i++) {
        vertex_nb[i] = uint32_t(d->C.cells[i].V.size());
        edge_nb[i] = uint32_t(d->C.cells[i].E.size());
        status[i] = d->C.status[i];
        known_radius[i] = d->C.knownRadii[i];
        flags[i] = d->C.flags[i];
    }
}

static void packCell(const Cell& c, uint32_t* vertices, uint32_t* edges) {
    for (size_t v = 0; v < c.V.size(); v++) {
        vertices[4 * v] = c.V[v].j;
        vertices[4 * v + 1] = c.V[v].k;
        vertices[4 * v + 2] = c.V[v].l;
        vertices[4 * v + 3] = uint32_t(c.V[v].type) | (uint32_t(c.V[v].status) << 8);
    }
    for (size_t e = 0; e < c.E.size(); e++) {
        edges[4 * e] = c.E[e].j;
        edges[4 * e + 1] = c.E[e].k;
        edges[4 * e + 2] = c.E[e].vertexNb;
        edges[4 * e + 3] = c.E[e].status;
    }
}

void orc_get_cell(const orc_diagram* d, uint32_t i, uint32_t* vertices, uint32_t* edges) {
    packCell(d->C.cells[i], vertices, edges);
}

void orc_get_all_cells(const orc_diagram* d, uint32_t* vertices, uint32_t* edges) {
    size_t vo = 0, eo = 0;
    for (const Cell& c : d->C.cells) {
        packCell(c, vertices + 4 * vo, edges + 4 * eo);
        vo += c.V.size();
        eo += c.E.size();
    }
}

// toTopology (algorithm.inl:1732-1821): per cell in sorted order, the std::set
// of bisectors {i, j}, trisectors {i, j, k}, quadrisectors {i, j, k, l} with i
// below the others and no artificial site, emitted in set order as original
// ids.  With null outputs only the counts are produced.
void orc_topology(const orc_diagram* d, uint64_t* counts3, uint32_t* bis, uint32_t* tri, uint32_t* quad) {
    const Context& C = d->C;
    const uint32_t n = C.n;
    const std::vector<uint32_t>& idx = d->bvh.indices;
    uint64_t nb = 0, nt = 0, nq = 0;
    for (uint32_t i = 0; i < n; i++) {
        const Cell& cell = C.cells[i];
        std::set<std::array<uint32_t, 2>> bs;
        std::set<std::array<uint32_t, 3>> ts;
        std::set<std::array<uint32_t, 4>> qs;
        for (const Edg& e : cell.E) {
            if (i < e.j && e.j < n) bs.insert({i, e.j});
            if (i < e.k && e.k < n) bs.insert({i, e.k});
            if (i < e.j && i < e.k && e.j < n && e.k < n) ts.insert({i, e.j, e.k});
        }
        for (const Vtx& v : cell.V)
            if (i < v.j && i < v.k && i < v.l && v.j < n && v.k < n && v.l < n) qs.insert({i, v.j, v.k, v.l});
        for (const auto& b : bs) {
            if (bis) {
                bis[2 * nb] = idx[b[0]];
                bis[2 * nb + 1] = idx[b[1]];
            }
            nb++;
        }
        for (const auto& t : ts) {
            if (tri)
                for (int a = 0; a < 3; a++) tri[3 * nt + a] = idx[t[a]];
            nt++;
        }
        for (const auto& q : qs) {
            if (quad)
                for (int a = 0; a < 4; a++) quad[4 * nq + a] = idx[q[a]];
            nq++;
        }
    }
    counts3[0] = nb;
    counts3[1] = nt;
    counts3[2] = nq;
}

void orc_get_knns(const orc_diagram* d, uint32_t* knn) { std::memcpy(knn, d->knn.data(), d->knn.size() * 4); }

int orc_box_reaches_half_space(const float* box6, const float* n3, const float* p3) {
    const Box b{{box6[0], box6[1], box6[2]}, {box6[3], box6[4], box6[5]}};
    return boxReachesHalfSpace(b, F3{n3[0], n3[1], n3[2]}, F3{p3[0], p3[1], p3[2]}) ? 1 : 0;
}

void orc_get_stats(const orc_diagram* d, uint64_t* stats12) { std::memcpy(stats12, d->stats, sizeof(d->stats)); }

void orc_certified_k(const float* xyzr, uint32_t n, uint32_t kmax, uint32_t* out_k, uint32_t threads) {
#ifdef _OPENMP
    if (threads) omp_set_num_threads(int(threads));
#else
    (void)threads;
#endif
    orc_config cfg;
    orc_default_config(&cfg);
    cfg.k = kmax;
    orc_diagram* d = new orc_diagram();
    Context& C = d->C;
    C.n = n;
    C.cfg = cfg;
    const SceneBox sb = sceneBox(xyzr, n);
    buildLbvh(xyzr, n, sb, d->bvh);
    C.bvh = &d->bvh;
    C.S.resize(size_t(n) + 6);
    C.Sf.resize(size_t(n) + 6);
    for (uint32_t i = 0; i < n; i++) {
        const float* s = &xyzr[4 * size_t(d->bvh.indices[i])];
        C.S[i] = {s[0], s[1], s[2], s[3]};
        C.Sf[i] = {s[0], s[1], s[2], s[3]};
    }
    D4 art[6];
    artificialSites(sb, art);
    for (int s = 0; s < 6; s++) {
        C.S[n + s] = art[s];
        C.Sf[n + s] = toF(art[s]);
    }
    C.status.assign(n, 0);
    C.knownRadii.assign(n, -FLT_MAX);
    C.cells.resize(n);
    C.flags.assign(n, 0);
    std::vector<uint32_t> knn(size_t(n) * kmax);
    std::vector<float> dist(size_t(n) * kmax);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < int64_t(n); ii++) {
        const uint32_t i = uint32_t(ii);
        uint32_t* keys = &knn[size_t(i) * kmax];
        knnQuery(d->bvh, C.Sf, i, kmax, -FLT_MAX, keys, &dist[size_t(i) * kmax]);
        if (keys[0] == kNone) continue;
        const F4 si = C.Sf[i];
        const D4 closest = C.S[keys[0]];
        const double ri = si.w, rj = closest.w;
        const D3 dc{double(si.x) - closest.x, double(si.y) - closest.y, double(si.z) - closest.z};
        if (ltD(len(dc), std::fmax(ri, rj) - std::fmin(ri, rj)) && !ltD(rj, ri)) C.status[i] = kBuried;
    }
    C.buried.assign(n, 0);
    for (uint32_t i = 0; i < n; i++) C.buried[i] = C.status[i] == kBuried;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < int64_t(n); ii++) {
        const uint32_t i = uint32_t(ii);
        const uint32_t orig = d->bvh.indices[i];
        out_k[orig] = kNone;
        if (C.status[i] == kBuried) {
            out_k[orig] = 0;
            continue;
        }
        Cell cell;
        initCell(C, cell);
        const D4 si = C.S[i];
        double knownRadius = -DBL_MAX;
        uint32_t flags = 0;
        for (uint32_t nId = 0; nId < kmax; nId++) {
            const uint32_t j = knn[size_t(i) * kmax + nId];
            if (j == kNone) break;
            if (C.buried[j]) continue;
            updateCell(C, i, si, cell.V, cell, j, C.S[j], flags);
            knownRadius = std::fmax(knownRadius, sphereDistD(si, C.S[j]));
            if (gtD(knownRadius, 2.0 * cellRadius(C, si, cell))) {
                out_k[orig] = nId + 1;
                break;
            }
        }
    }
    delete d;
}

uint64_t orc_empty_sphere_violations(const float* xyzr, uint32_t n, const double* xyzw, const uint32_t* ids,
                                     uint64_t v, double tol, uint32_t threads) {
#ifdef _OPENMP
    if (threads) omp_set_num_threads(int(threads));
#else
    (void)threads;
#endif
    uint64_t bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
    for (int64_t q = 0; q < int64_t(v); q++) {
        const double px = xyzw[4 * q], py = xyzw[4 * q + 1], pz = xyzw[4 * q + 2];
        auto wd = [&](uint32_t s) {
            const double dx = double(xyzr[4 * s]) - px, dy = double(xyzr[4 * s + 1]) - py,
                         dz = double(xyzr[4 * s + 2]) - pz;
            return std::sqrt(dx * dx + dy * dy + dz * dz) - double(xyzr[4 * s + 3]);
        };
        double wmax = -DBL_MAX;
        for (int t = 0; t < 4; t++) wmax = std::fmax(wmax, wd(ids[4 * q + t]));
        for (uint32_t s = 0; s < n; s++) {
            if (s == ids[4 * q] || s == ids[4 * q + 1] || s == ids[4 * q + 2] || s == ids[4 * q + 3]) continue;
            if (wd(s) < wmax - tol) {
                bad++;
                break;
            }
        }
    }
    return bad;
}

uint64_t orc_bruteforce_vertices(const float* xyzr, uint32_t n, uint32_t* ids, uint64_t cap, double tol) {
    const SceneBox sb = sceneBox(xyzr, n);
    D4 art[6];
    artificialSites(sb, art);
    std::vector<D4> all(size_t(n) + 6);
    for (uint32_t i = 0; i < n; i++) all[i] = {xyzr[4 * i], xyzr[4 * i + 1], xyzr[4 * i + 2], xyzr[4 * i + 3]};
    for (int s = 0; s < 6; s++) all[n + s] = art[s];
    std::vector<std::array<uint32_t, 4>> found;
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t b = a + 1; b < n; b++)
            for (uint32_t c = b + 1; c < n; c++)
                for (uint32_t e = c + 1; e < n; e++) {
                    D4 x[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
                    const uint32_t mask = quadrisector(all[a], all[b], all[c], all[e], x[0], x[1]);
                    for (int r = 0; r < 2; r++) {
                        if (!((mask >> r) & 1u)) continue;
                        const D4 p = x[r];
                        bool empty = true;
                        for (size_t s = 0; s < all.size() && empty; s++) {
                            if (s == a || s == b || s == c || s == e) continue;
                            const double dx = all[s].x - p.x, dy = all[s].y - p.y, dz = all[s].z - p.z;
                            if (std::sqrt(dx * dx + dy * dy + dz * dz) - all[s].w < p.w - tol) empty = false;
                        }
                        if (empty) found.push_back({a, b, c, e});
                    }
                }
    std::sort(found.begin(), found.end());
    for (size_t q = 0; q < found.size() && q < cap; q++)
        for (int t = 0; t < 4; t++) ids[4 * q + t] = found[q][t];
    return found.size();
}

}  // extern "C"
