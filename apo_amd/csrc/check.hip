// check.hip -- FP64 empty-sphere check of Apollonius vertices through the LBVH.
//
// Same criterion as the reference's precision validator
// (app/precision/precision.cpp:26-100): with w = the largest weighted
// distance |p - c| - r from the vertex centre p to its 4 tangent sites, the
// vertex is violated when some other site is closer than w - tol.  The
// reference scans every site (O(V*N)); here each vertex walks the LBVH and
// only tests sites whose FP32 leaf box can come within the bound.
#include "apo_device.h"
#include "apo_internal.h"

namespace apo_hip {

__device__ __forceinline__ double weightedDistance(const D4& s, const D3& p) {
    const double dx = s.x - p.x, dy = s.y - p.y, dz = s.z - p.z;
    return __builtin_sqrt(dx * dx + dy * dy + dz * dz) - s.w;
}

// Vertices of a batch are checked against the sites of their own molecule
// (the molecule of their first tangent site), through that molecule's tree.
__global__ __launch_bounds__(256) void k_check_empty(CellCtx c, const uint32_t* __restrict__ indices,
                                                     const uint32_t* __restrict__ rank, const D4* __restrict__ xyzw,
                                                     const uint4* __restrict__ ids, uint64_t count, double tol,
                                                     unsigned long long* violations) {
    const F4* __restrict__ sorted = c.sorted;
    const Node* __restrict__ nodes = c.nodes;
    const uint32_t n = c.n;
    uint32_t* err = c.err;
    const uint64_t q = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const D4 v = xyzw[q];
    const uint4 id = ids[q];
    const D3 p{v.x, v.y, v.z};
    const uint32_t o[4] = {id.x, id.y, id.z, id.w};
    double wmax = -DBL_MAX;
    for (int t = 0; t < 4; t++) {
        if (o[t] >= n) {
            atomicOr(err, kErrCellSite);
            return;
        }
        wmax = __builtin_fmax(wmax, weightedDistance(toD(sorted[rank[o[t]]]), p));
    }
    const double bound = wmax - tol;
    // A site s below a node violates only if |p - c_s| - r_s < bound; with d
    // the distance from p to the node's extent box and R its largest radius,
    // |p - c_s| - r_s >= d - R.  FP32 rounding is covered by `slack`.
    const float slack = float(1e-6 * (__builtin_fabs(p.x) + __builtin_fabs(p.y) + __builtin_fabs(p.z) +
                                      __builtin_fabs(wmax)) + 1e-6);
    const float fbound = float(bound);
    const F4 pt{float(p.x), float(p.y), float(p.z), 0.f};
    uint32_t stack[64];
    uint32_t sp = 0, node = rootNode(c, cellMolecule(c, rank[o[0]]));
    const uint32_t internal = c.internal, total = nodeCount(c);
    bool bad = false;
    while (!bad) {
        if (node < internal) {
            const Node nn = nodes[node];
            Box lb, rb;
            nodeBoxes(nn, lb, rb);
            const uint32_t L = uint32_t(nn.d.x), R = uint32_t(nn.d.y);
            const float lmax = __builtin_bit_cast(float, nn.d.z), rmax = __builtin_bit_cast(float, nn.d.w);
            const bool goL = boxDistance(lb, pt) - lmax < fbound + slack;
            const bool goR = boxDistance(rb, pt) - rmax < fbound + slack;
            if (goL && goR) {
                if (sp < 64) stack[sp++] = R;
                else atomicOr(err, kErrKnnChild);
                node = L;
                continue;
            }
            if (goL || goR) {
                node = goL ? L : R;
                continue;
            }
        } else if (node < total) {
            const uint32_t j = node - internal;
            const uint32_t orig = indices[j];
            if (orig != id.x && orig != id.y && orig != id.z && orig != id.w)
                bad = weightedDistance(toD(sorted[j]), p) < bound;
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    if (bad) atomicAdd(violations, 1ull);
}

// ---------------------------------------------------------------------------
// Interval validator of the reference's precision app
// (app/precision/precision.cpp:14-163, Boost.Interval over double).  Every
// operation rounds outward: the FP64 result of round-to-nearest is corrected
// by one ulp towards -inf (lower bound) or +inf (upper bound) when the exact
// error term, obtained by an error-free transformation (TwoSum for +/-, an
// FMA for * and for the square-root residual), shows that rounding went the
// wrong way.  That is exactly directed rounding, so the bounds are those
// Boost computes with the rounding mode switched.
// ---------------------------------------------------------------------------
struct Ival { double lo, hi; };

__global__ void k_rank(const uint32_t* __restrict__ indices, uint32_t n, uint32_t* rank);

__device__ __forceinline__ double nextDown(double x) {
    if (x == 0.0) return -__builtin_bit_cast(double, uint64_t(1));
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    return __builtin_bit_cast(double, x > 0.0 ? b - 1u : b + 1u);
}
__device__ __forceinline__ double nextUp(double x) { return -nextDown(-x); }
// exact error of s = RN(a + b) (Knuth's TwoSum)
__device__ __forceinline__ double sumError(double a, double b, double s) {
    const double bb = s - a;
    return (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ double addDown(double a, double b) {
    const double s = a + b;
    return sumError(a, b, s) < 0.0 ? nextDown(s) : s;
}
__device__ __forceinline__ double addUp(double a, double b) {
    const double s = a + b;
    return sumError(a, b, s) > 0.0 ? nextUp(s) : s;
}
__device__ __forceinline__ double mulDown(double a, double b) {
    const double p = a * b;
    return __builtin_fma(a, b, -p) < 0.0 ? nextDown(p) : p;
}
__device__ __forceinline__ double mulUp(double a, double b) {
    const double p = a * b;
    return __builtin_fma(a, b, -p) > 0.0 ? nextUp(p) : p;
}
__device__ __forceinline__ double sqrtDown(double x) {
    const double s = __builtin_sqrt(x);
    return __builtin_fma(-s, s, x) < 0.0 ? nextDown(s) : s;
}
__device__ __forceinline__ double sqrtUp(double x) {
    const double s = __builtin_sqrt(x);
    return __builtin_fma(-s, s, x) > 0.0 ? nextUp(s) : s;
}
__device__ __forceinline__ Ival imul(Ival a, Ival b) {
    const double l = __builtin_fmin(__builtin_fmin(mulDown(a.lo, b.lo), mulDown(a.lo, b.hi)),
                                    __builtin_fmin(mulDown(a.hi, b.lo), mulDown(a.hi, b.hi)));
    const double h = __builtin_fmax(__builtin_fmax(mulUp(a.lo, b.lo), mulUp(a.lo, b.hi)),
                                    __builtin_fmax(mulUp(a.hi, b.lo), mulUp(a.hi, b.hi)));
    return {l, h};
}
// weightedDistance (precision.cpp:14-24): sqrt(sum (s_i - p_i)^2) - r
__device__ __forceinline__ Ival intervalDistance(const D4& s, const D3& p) {
    const double sc[3] = {s.x, s.y, s.z}, pc[3] = {p.x, p.y, p.z};
    Ival res{0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        const Ival diff{addDown(sc[i], -pc[i]), addUp(sc[i], -pc[i])};
        const Ival sq = imul(diff, diff);
        res = {addDown(res.lo, sq.lo), addUp(res.hi, sq.hi)};
    }
    // the sum's upper bound is never negative here (a sum of squares)
    const Ival root{res.lo <= 0.0 ? 0.0 : sqrtDown(res.lo), sqrtUp(res.hi)};
    return {addDown(root.lo, -s.w), addUp(root.hi, -s.w)};
}

// One point per thread: bad when another site's distance interval lies
// entirely below the largest of the point's own sites (precision.cpp:71-89,
// compare::certain), uncertain when it is not bad and some other site's
// interval reaches below that largest one (the configuration is
// near-degenerate: FP64 intervals cannot decide it).  arity 4: Apollonius
// vertices (getBadQuadrisectorNb); arity 3: closed trisector extreme points
// (getBadTrisectorNb, ids.w ignored).
__global__ __launch_bounds__(256) void k_precision(CellCtx c, const uint32_t* __restrict__ indices,
                                                   const uint32_t* __restrict__ rank, const D4* __restrict__ xyzw,
                                                   const uint4* __restrict__ ids, uint64_t count, uint32_t arity,
                                                   unsigned long long* counts2) {
    const F4* __restrict__ sorted = c.sorted;
    const Node* __restrict__ nodes = c.nodes;
    const uint32_t n = c.n;
    uint32_t* err = c.err;
    const uint64_t q = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const D4 v = xyzw[q];
    const uint4 id4 = ids[q];
    const D3 p{v.x, v.y, v.z};
    const uint32_t o[4] = {id4.x, id4.y, id4.z, arity == 4u ? id4.w : id4.z};
    Ival qmax{-DBL_MAX, -DBL_MAX};
    for (uint32_t t = 0; t < arity; t++) {
        if (o[t] >= n) {
            atomicOr(err, kErrCellSite);
            return;
        }
        const Ival d = intervalDistance(toD(sorted[rank[o[t]]]), p);
        qmax = {__builtin_fmax(qmax.lo, d.lo), __builtin_fmax(qmax.hi, d.hi)};
    }
    // A site s below a node can reach below qmax.hi only if
    // |p - c_s| - r_s < qmax.hi; |p - c_s| - r_s >= d - R with d the distance
    // from p to the node's box and R its largest radius; `slack` covers the
    // FP32 evaluation of d - R.
    const float slack = float(1e-6 * (__builtin_fabs(p.x) + __builtin_fabs(p.y) + __builtin_fabs(p.z) +
                                      __builtin_fabs(qmax.hi)) + 1e-6);
    const float fbound = float(qmax.hi);
    const F4 pt{float(p.x), float(p.y), float(p.z), 0.f};
    uint32_t stack[64];
    uint32_t sp = 0, node = rootNode(c, cellMolecule(c, rank[o[0]]));
    const uint32_t internal = c.internal, total = nodeCount(c);
    bool bad = false, uncertain = false;
    while (!bad) {
        if (node < internal) {
            const Node nn = nodes[node];
            Box lb, rb;
            nodeBoxes(nn, lb, rb);
            const uint32_t L = uint32_t(nn.d.x), R = uint32_t(nn.d.y);
            const float lmax = __builtin_bit_cast(float, nn.d.z), rmax = __builtin_bit_cast(float, nn.d.w);
            const bool goL = boxDistance(lb, pt) - lmax < fbound + slack;
            const bool goR = boxDistance(rb, pt) - rmax < fbound + slack;
            if (goL && goR) {
                if (sp < 64) stack[sp++] = R;
                else atomicOr(err, kErrKnnChild);
                node = L;
                continue;
            }
            if (goL || goR) {
                node = goL ? L : R;
                continue;
            }
        } else if (node < total) {
            const uint32_t j = node - internal;
            const uint32_t orig = indices[j];
            if (orig != o[0] && orig != o[1] && orig != o[2] && orig != o[3]) {
                const Ival d = intervalDistance(toD(sorted[j]), p);
                bad = d.hi < qmax.lo;
                uncertain = uncertain || d.lo < qmax.hi;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    if (bad) atomicAdd(&counts2[0], 1ull);
    else if (uncertain) atomicAdd(&counts2[1], 1ull);
}

hipError_t precisionLaunch(const CellCtx& c, const uint32_t* indices, uint32_t* rank, const D4* xyzw,
                           const uint4* ids, uint64_t count, uint32_t arity, unsigned long long* counts2,
                           hipStream_t st) {
    if (arity != 3u && arity != 4u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rank, dim3((c.n + 255) / 256), dim3(256), 0, st, indices, c.n, rank);
    if (count == 0) return hipGetLastError();
    const uint64_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(k_precision, dim3(uint32_t(blocks)), dim3(256), 0, st, c, indices, rank, xyzw, ids, count,
                       arity, counts2);
    return hipGetLastError();
}

__global__ void k_rank(const uint32_t* __restrict__ indices, uint32_t n, uint32_t* rank) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) rank[indices[j]] = j;
}

hipError_t rankLaunch(const uint32_t* indices, uint32_t n, uint32_t* rank, hipStream_t st) {
    hipLaunchKernelGGL(k_rank, dim3((n + 255) / 256), dim3(256), 0, st, indices, n, rank);
    return hipGetLastError();
}

hipError_t checkEmptyLaunch(const CellCtx& c, const uint32_t* indices, uint32_t* rank, const D4* xyzw,
                            const uint4* ids, uint64_t count, double tol, unsigned long long* violations,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_rank, dim3((c.n + 255) / 256), dim3(256), 0, st, indices, c.n, rank);
    if (count == 0) return hipGetLastError();
    const uint64_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(k_check_empty, dim3(uint32_t(blocks)), dim3(256), 0, st, c, indices, rank, xyzw, ids, count,
                       tol, violations);
    return hipGetLastError();
}

}  // namespace apo_hip
