// apo_device.h -- device-side types, wave64 helpers and FP32/FP64 geometry
// for the gfx950 Apollonius-diagram kernels.
//
// All HIP code is compiled with -ffp-contract=off: every expression below
// keeps the operand order of the apo expression it restates (file:line
// given per function), so results are bit-identical to an IEEE host
// evaluation of the same expression.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace apo_hip {

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint8_t kStatusInValidation = 0;
constexpr uint8_t kStatusCertified = 1;
constexpr uint8_t kStatusBuried = 0xff;

constexpr uint32_t kFlagVertexOverflow = 1u;
constexpr uint32_t kFlagEdgeOverflow = 2u;
constexpr uint32_t kFlagNewEdgeOverflow = 4u;
constexpr uint32_t kFlagBisectorOverflow = 8u;
constexpr uint32_t kFlagRestartCap = 16u;
constexpr uint32_t kFlagStackOverflow = 32u;

constexpr int kWave = 64;

// Device guard bits (a guard turns an inconsistent intermediate into an
// error report instead of an out-of-bounds access).
constexpr uint32_t kErrSortTile = 1u;
constexpr uint32_t kErrSortScatter = 2u;
constexpr uint32_t kErrSortSpin = 4u;
constexpr uint32_t kErrGatherIndex = 8u;
constexpr uint32_t kErrRefitParent = 16u;
constexpr uint32_t kErrRefitChild = 32u;
constexpr uint32_t kErrKnnChild = 64u;
constexpr uint32_t kErrCellSite = 128u;
constexpr uint32_t kErrCellNeighbour = 256u;
constexpr uint32_t kErrValidation = 512u;

// ---------------------------------------------------------------------------
// vector types with explicit, ordered arithmetic
// ---------------------------------------------------------------------------
struct F3 { float x, y, z; };
struct __attribute__((aligned(16))) F4 { float x, y, z, w; };
struct D2 { double x, y; };
struct D3 { double x, y, z; };
struct __attribute__((aligned(16))) D4 { double x, y, z, w; };

__device__ __forceinline__ D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double len(D3 v) { return __builtin_sqrt(dot(v, v)); }
__device__ __forceinline__ D3 normalize(D3 v) {
    const double inv = 1.0 / __builtin_sqrt(dot(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}
__device__ __forceinline__ D3 xyz(D4 a) { return {a.x, a.y, a.z}; }

__device__ __forceinline__ F3 subf(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 addf(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 mulf(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float lenf(F3 v) { return __builtin_sqrtf(dotf(v, v)); }
__device__ __forceinline__ F3 normalizef(F3 v) {
    const float inv = 1.0f / __builtin_sqrtf(dotf(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}
__device__ __forceinline__ F3 xyzf(F4 a) { return {a.x, a.y, a.z}; }
// Component-wise selects: a conditional on aggregates would select between
// addresses and push both operands to private memory.
__device__ __forceinline__ F3 selF3(bool c, F3 a, F3 b) { return {c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
__device__ __forceinline__ F4 selF4(bool c, F4 a, F4 b) {
    return {c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w};
}
__device__ __forceinline__ D4 selD4(bool c, D4 a, D4 b) {
    return {c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w};
}
__device__ __forceinline__ F4 toF(D4 a) { return {float(a.x), float(a.y), float(a.z), float(a.w)}; }
__device__ __forceinline__ F3 toF3(D3 a) { return {float(a.x), float(a.y), float(a.z)}; }
__device__ __forceinline__ D4 toD(F4 a) { return {double(a.x), double(a.y), double(a.z), double(a.w)}; }

// eps predicates (apo/include/apo/core/math.inl:245-276), eps = 1e-8
__device__ __forceinline__ bool ltD(double a, double b) { return (a - b) < 1e-8; }
__device__ __forceinline__ bool gtD(double a, double b) { return (a - b) > -1e-8; }
__device__ __forceinline__ bool ltF(float a, float b) { return (a - b) < 1e-8f; }
__device__ __forceinline__ double signD(double a) { return ltD(a, 0.0) ? -1.0 : 1.0; }
__device__ __forceinline__ float signF(float a) { return ltF(a, 0.f) ? -1.f : 1.f; }
__device__ __forceinline__ bool isCloseToZeroD(double v) { return __builtin_fabs(v) < 1e-8; }
__device__ __forceinline__ bool isCloseD(double a, double ref) {
    return __builtin_fabs(a - ref) <= (1e-5 + 1e-8 * __builtin_fabs(ref));
}

// sphereDistance / intersect (apo/include/apo/gpu/math.cuh:368-388)
__device__ __forceinline__ float sphereDistF(F4 a, F4 b) { return lenf(subf(xyzf(a), xyzf(b))) - (a.w + b.w); }
__device__ __forceinline__ double sphereDistD(D4 a, D4 b) {
    const D3 d{a.x - b.x, a.y - b.y, a.z - b.z};
    return len(d) - (a.w + b.w);
}
// intersect (math.cuh:380-388): the rounded value sqrt(d2) - s, with d2 the
// squared centre distance and s = ra + rb, below the tolerance T (1e-8).
// Away from the boundary d2 = (s + T)^2 the square root decides nothing:
// with s >= 0, d2 below (s + T)^2 (1 - g) puts the rooted value more than
// g/4 (s + T) under T, more than its two roundings (square root, then
// subtraction: one unit each, relative to at most s + T) can move it, and
// d2 above (s + T)^2 (1 + g) puts the root above s + T, so the rounded
// difference is at least T.  g = 2^-40 in FP64 and 2^-12 in FP32 dwarf those
// units (2^-53, 2^-24) and the rounding of the threshold itself.  Only a
// distance inside the band (or s < 0, or an operand out of range) takes the
// rooted form, so every decision is the rooted form's bit for bit.
__device__ __forceinline__ bool intersectF(F4 a, F4 b) {
    const F3 d = subf(xyzf(a), xyzf(b));
    const float d2 = dotf(d, d);
    const float s = a.w + b.w;
    const float t = s + 1e-8f;
    const float t2 = t * t;
    if (s >= 0.f && t2 < 1e30f) {
        if (d2 < t2 * (1.f - 0x1p-12f)) return true;
        if (d2 > t2 * (1.f + 0x1p-12f)) return false;
    }
    return ltF(__builtin_sqrtf(d2) - s, 0.f);
}
__device__ __forceinline__ bool intersectD(D4 a, D4 b) {
    const D3 d{a.x - b.x, a.y - b.y, a.z - b.z};
    const double d2 = dot(d, d);
    const double s = a.w + b.w;
    const double t = s + 1e-8;
    const double t2 = t * t;
    if (s >= 0.0 && t2 < 1e300) {
        if (d2 < t2 * (1.0 - 0x1p-40)) return true;
        if (d2 > t2 * (1.0 + 0x1p-40)) return false;
    }
    return ltD(__builtin_sqrt(d2) - s, 0.0);
}

// ---------------------------------------------------------------------------
// boxes (apo/include/apo/gpu/math.inl:79-101), FP32
// ---------------------------------------------------------------------------
struct Box { F3 mn, mx; };
__device__ __forceinline__ Box selBox(bool c, const Box& a, const Box& b) { return {selF3(c, a.mn, b.mn), selF3(c, a.mx, b.mx)}; }

// Can the box hold the centre of a site that the bisector search's
// cylinder test (distance to the axis through o along unit u, plus the
// site's radius, below cylR) accepts?  A conservative bound for radii >= 0:
// the distance of the box's bounding sphere to the axis (a cross product, no
// cancellation) against cylR plus 2e-3 of the farthest box point's distance
// to o, twice what the FP32 rounding of the leaf test's distance can lose
// (its a - v^2 cancels: at most sqrt(15 x 2^-24) of that distance).
__device__ __forceinline__ bool boxNearAxis(const Box& b, F3 o, F3 u, float cylR) {
    const F3 c{(b.mn.x + b.mx.x) * .5f, (b.mn.y + b.mx.y) * .5f, (b.mn.z + b.mx.z) * .5f};
    const F3 hv{(b.mx.x - b.mn.x) * .5f, (b.mx.y - b.mn.y) * .5f, (b.mx.z - b.mn.z) * .5f};
    const float h = lenf(hv);
    const F3 w = subf(c, o);
    const F3 x{w.y * u.z - w.z * u.y, w.z * u.x - w.x * u.z, w.x * u.y - w.y * u.x};
    const float lmax = lenf(w) + h;
    return lenf(x) - h < cylR * 1.000001f + 2e-3f * lmax;
}

__device__ __forceinline__ F3 closestPoint(const Box& b, F4 s) {
    return {__builtin_fminf(__builtin_fmaxf(s.x, b.mn.x), b.mx.x),
            __builtin_fminf(__builtin_fmaxf(s.y, b.mn.y), b.mx.y),
            __builtin_fminf(__builtin_fmaxf(s.z, b.mn.z), b.mx.z)};
}
__device__ __forceinline__ float boxDistance(const Box& b, F4 s) {
    return lenf(subf(xyzf(s), closestPoint(b, s))) - s.w;
}
// Can box b hold a site of radius up to rmax that reaches into the ball?
// The edge cover's subtree test, boxDistance(b, ball) <= rmax, without the
// square root: the squared distance against (ball radius + rmax)^2 with a
// relative slack of 2^-16.  The rooted form's FP32 roundings (square root,
// subtraction) move its decision by less than 2^-22 of the squared
// distance, so every box it accepts is accepted here: a conservative
// superset, which is all a pruning test must be.
__device__ __forceinline__ bool boxNearBall(const Box& b, F4 ball, float rmax) {
    const F3 d = subf(xyzf(ball), closestPoint(b, ball));
    const float s = ball.w + rmax;
    return s >= 0.f && dotf(d, d) <= s * s * (1.f + 0x1p-16f);
}
__device__ __forceinline__ float boxFarthestDistance(const Box& b, F4 s) {
    const F3 centroid{(b.mn.x + b.mx.x) * .5f, (b.mn.y + b.mx.y) * .5f, (b.mn.z + b.mx.z) * .5f};
    const F3 radius{__builtin_fabsf(b.mn.x - b.mx.x) * .5f, __builtin_fabsf(b.mn.y - b.mx.y) * .5f,
                    __builtin_fabsf(b.mn.z - b.mx.z) * .5f};
    const F3 c = subf(xyzf(s), centroid);
    const F3 fp{radius.x * signF(-c.x), radius.y * signF(-c.y), radius.z * signF(-c.z)};
    return lenf(subf(c, fp)) - s.w;
}

// Can some point of box b lie strictly inside the half-space n.(x - p) < 0?
// The box's lowest point along n is the corner taking the lower bound where
// n is positive; the margin covers the FP32 rounding of the three-term dot
// product (8 units in the last place of the sum of magnitudes, 2^-20), so no
// box holding a site inside the half-space is rejected.  The
// completion mode's hyperbolic edge searches prune with it.
__device__ __forceinline__ bool boxReachesHalfSpace(const Box& b, F3 n, F3 p) {
    const float dx = (n.x > 0.f ? b.mn.x : b.mx.x) - p.x;
    const float dy = (n.y > 0.f ? b.mn.y : b.mx.y) - p.y;
    const float dz = (n.z > 0.f ? b.mn.z : b.mx.z) - p.z;
    const float ex = n.x * dx, ey = n.y * dy, ez = n.z * dz;
    const float margin = 0x1p-20f * (__builtin_fabsf(ex) + __builtin_fabsf(ey) + __builtin_fabsf(ez));
    return ex + ey + ez < margin;
}

// Upper gap bound of a range query of radius b around a site of radius r.
// A subtree box holding the site's centre has box gap -r, and the gap of a
// site whose sphere holds that centre is below -r, so the box gap bounds
// such sites' gaps from above, not below: a bound at or below -r would
// prune every box around the centre.  The bound is therefore raised to the
// float just above -r; every site whose gap is below it is then in reach
// (box gaps bound the gaps of the other sites from below), the site that
// buries this one (gap < -2 r) included.
__device__ __host__ __forceinline__ float rangeTop(float b, float r) {
    const float negR = -r;
    const uint32_t bits = __builtin_bit_cast(uint32_t, negR);
    const float above = negR < 0.f ? __builtin_bit_cast(float, bits - 1u) : __builtin_bit_cast(float, 1u);
    return b > above ? b : above;
}

// ---------------------------------------------------------------------------
// LBVH node: both child boxes live in the parent so one 64-byte line per
// visited node carries everything the traversal needs.
//   a = (lmin.x, lmin.y, lmin.z, lmax.x)
//   b = (lmax.y, lmax.z, rmin.x, rmin.y)
//   c = (rmin.z, rmax.x, rmax.y, rmax.z)
//   d = (left, right, bits(lmaxr), bits(rmaxr))
// Node numbering is the reference's: internal [0, n-1), leaves [n-1, 2n-1).
// ---------------------------------------------------------------------------
struct __attribute__((aligned(64))) Node {
    float4 a, b, c;
    int4 d;
};

__device__ __forceinline__ void nodeBoxes(const Node& nd, Box& l, Box& r) {
    l.mn = {nd.a.x, nd.a.y, nd.a.z};
    l.mx = {nd.a.w, nd.b.x, nd.b.y};
    r.mn = {nd.b.z, nd.b.w, nd.c.x};
    r.mx = {nd.c.y, nd.c.z, nd.c.w};
}

// ---------------------------------------------------------------------------
// FP64 geometry (apo/include/apo/gpu/math.cuh:390-780)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void inverse3(D3 c1, D3 c2, D3 c3, D3& i1, D3& i2, D3& i3) {
    const double invD = 1.0 / dot(c1, cross(c2, c3));
    const D3 r1 = cross(c2, c3);
    const D3 r2 = cross(c3, c1);
    const D3 r3 = cross(c1, c2);
    i1 = {invD * r1.x, invD * r2.x, invD * r3.x};
    i2 = {invD * r1.y, invD * r2.y, invD * r3.y};
    i3 = {invD * r1.z, invD * r2.z, invD * r3.z};
}
__device__ __forceinline__ D3 matVec(D3 i1, D3 i2, D3 i3, D3 b) {
    return {dot(D3{i1.x, i2.x, i3.x}, b), dot(D3{i1.y, i2.y, i3.y}, b), dot(D3{i1.z, i2.z, i3.z}, b)};
}

// Apollonius vertices of 4 spheres (math.cuh:703-780); 2-bit validity mask.
__device__ __forceinline__ uint32_t quadrisector(D4 s1, D4 s2, D4 s3, D4 s4, D4& x1, D4& x2) {
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
    const double q = -.5 * (b - signD(b) * __builtin_sqrt(delta));
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

// Extreme points of the trisector of 3 spheres (math.cuh:636-701).  `alias`
// reproduces the reference call with x1 and x2 bound to the same variable
// (topology_update.inl:188): the result is then written to x1 only.
template <bool alias>
__device__ __forceinline__ uint32_t trisectorImpl(D4 s1, D4 s2, D4 s3, D4& x1, D4& x2) {
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
    const double q = -.5 * (b - signD(b) * __builtin_sqrt(delta));
    const double xd1 = (c / q) - r1;
    const D3 x1n = sub(mul(mN, xd1), mB);
    if (gtD(xd1 + r1, 0.0)) {
        x1 = {x1n.x, x1n.y, x1n.z, xd1};
        count++;
    }
    const double xd2 = (q / a) - r1;
    const D3 x2n = sub(mul(mN, xd2), mB);
    if (gtD(xd2 + r1, 0.0)) {
        if (alias) x1 = {x2n.x, x2n.y, x2n.z, xd2};
        else {
            x2 = {x2n.x, x2n.y, x2n.z, xd2};
            if (count == 0) x1 = x2;
        }
        count++;
    }
    return count;
}
__device__ __forceinline__ uint32_t trisector(D4 s1, D4 s2, D4 s3, D4& x1, D4& x2) {
    return trisectorImpl<false>(s1, s2, s3, x1, x2);
}
__device__ __forceinline__ uint32_t trisectorAliased(D4 s1, D4 s2, D4 s3, D4& x) {
    D4 unused;
    return trisectorImpl<true>(s1, s2, s3, x, unused);
}

__device__ __forceinline__ void inverse2(D2 a1, D2 a2, D2& c1, D2& c2) {
    const double invDet = 1.0 / __builtin_fma(a1.x, a2.y, -a2.x * a1.y);
    c1 = {invDet * a2.y, -invDet * a1.y};
    c2 = {-invDet * a2.x, invDet * a1.x};
}
__device__ __forceinline__ void matMul2(D2 a1, D2 a2, D2 b1, D2 b2, D2& c1, D2& c2) {
    c1 = {a1.x * b1.x + a2.x * b1.y, a1.y * b1.x + a2.y * b1.y};
    c2 = {a1.x * b2.x + a2.x * b2.y, a1.y * b2.x + a2.y * b2.y};
}
__device__ __forceinline__ void swapD(double& a, double& b) {
    const double t = a;
    a = b;
    b = t;
}

// Symmetric 2x2 eigendecomposition (math.cuh:428-503)
__device__ __forceinline__ void eigen2(D2 m1, D2 m2, D2& e, D2& v) {
    const double a = m1.x, b = m2.x, c = m2.y;
    const double sm = a + c, df = a - c, adf = __builtin_fabs(df), tb = b + b, ab = __builtin_fabs(tb);
    double acmx = a, acmn = c;
    if (__builtin_fabs(a) <= __builtin_fabs(c)) swapD(acmx, acmn);
    double aa = ab, bb = adf;
    if (adf < ab) swapD(aa, bb);
    const double temp = aa / bb;
    const double rt = bb * __builtin_sqrt(1.0 + temp * temp);
    const double sgn1 = __builtin_copysign(1.0, sm);
    const double rt1 = .5 * (sm + sgn1 * rt);
    const double rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
    const double sgn2 = __builtin_copysign(1.0, df);
    const double cs = df + sgn2 * rt;
    const double acs = __builtin_fabs(cs);
    aa = tb;
    bb = cs;
    if (ltD(acs, ab)) swapD(aa, bb);
    const double t = -aa / bb;
    double sn1 = 1.0 / __builtin_sqrt(1.0 + t * t);
    double cs1 = t * sn1;
    if (ltD(acs, ab)) swapD(sn1, cs1);
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

__device__ __forceinline__ double dot2(D2 a, D2 b) { return a.x * b.x + a.y * b.y; }

// Common tangent planes of 3 spheres (math.cuh:577-634)
__device__ __forceinline__ void tangentPlanes(D4 s1, D4 s2, D4 s3, D3& n1, D3& n2) {
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
    const double alpha0 = __builtin_sqrt(__builtin_fabs(ev.x));
    const double alpha1 = __builtin_sqrt(__builtin_fabs(ev.y));
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
    if (gtD(__builtin_fabs(dot(n1, sub(p, p1)) - s1.w), 1e-4) || gtD(__builtin_fabs(dot(n1, sub(p, p2)) - s2.w), 1e-4) ||
        gtD(__builtin_fabs(dot(n1, sub(p, p3)) - s3.w), 1e-4)) {
        n1 = mul(n1, -1.0);
        n2 = mul(n2, -1.0);
    }
}

__device__ __forceinline__ double det4(D4 c1, D4 c2, D4 c3, D4 c4) {
    const double d1 = dot(D3{c2.y, c2.z, c2.w}, cross(D3{c3.y, c3.z, c3.w}, D3{c4.y, c4.z, c4.w}));
    const double d2 = dot(D3{c2.x, c2.z, c2.w}, cross(D3{c3.x, c3.z, c3.w}, D3{c4.x, c4.z, c4.w}));
    const double d3 = dot(D3{c2.x, c2.y, c2.w}, cross(D3{c3.x, c3.y, c3.w}, D3{c4.x, c4.y, c4.w}));
    const double d4 = dot(D3{c2.x, c2.y, c2.z}, cross(D3{c3.x, c3.y, c3.z}, D3{c4.x, c4.y, c4.z}));
    return c1.x * d1 - c1.y * d2 + c1.z * d3 - c1.w * d4;
}

// Sphere through 4 points (math.cuh:533-575)
__device__ __forceinline__ D4 sphereFromPoints(D3 p1, D3 p2, D3 p3, D3 p4) {
    const double t1 = -dot(p1, p1), t2 = -dot(p2, p2), t3 = -dot(p3, p3), t4 = -dot(p4, p4);
    const D4 X{p1.x, p2.x, p3.x, p4.x}, Y{p1.y, p2.y, p3.y, p4.y}, Z{p1.z, p2.z, p3.z, p4.z}, O{1, 1, 1, 1};
    const D4 T{t1, t2, t3, t4};
    const double t = det4(X, Y, Z, O);
    const double d = det4(T, Y, Z, O) / t;
    const double e = det4(X, T, Z, O) / t;
    const double f = det4(X, Y, T, O) / t;
    const double g = det4(X, Y, Z, T) / t;
    return {-d * .5, -e * .5, -f * .5, .5 * __builtin_sqrt(d * d + e * e + f * f - 4. * g)};
}

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t laneId() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// number of set bits of `mask` in lanes below the calling lane
__device__ __forceinline__ uint32_t countBelow(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}
__device__ __forceinline__ uint32_t popc64(uint64_t m) { return uint32_t(__builtin_popcountll(m)); }
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint32_t readFirst(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t readLane(uint32_t v, uint32_t lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ double readLaneD(double v, uint32_t lane) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(b), lane);
    const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(b >> 32), lane);
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ D4 readLaneD4(D4 v, uint32_t lane) {
    return {readLaneD(v.x, lane), readLaneD(v.y, lane), readLaneD(v.z, lane), readLaneD(v.w, lane)};
}
__device__ __forceinline__ double waveMaxD(double v) {
    for (int o = 32; o >= 1; o >>= 1) v = __builtin_fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t waveSumU(uint32_t v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// inclusive prefix sum of small per-lane values
__device__ __forceinline__ uint32_t waveInclusiveSum(uint32_t v) {
    const uint32_t lane = laneId();
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t other = __shfl_up(v, o, 64);
        if (lane >= uint32_t(o)) v += other;
    }
    return v;
}

}  // namespace apo_hip
