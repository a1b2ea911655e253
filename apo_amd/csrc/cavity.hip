// cavity.hip -- cavity detection on the built diagram (gfx950).
//
// Reference: app/cavity/utils.cu.  The reference's cavity app takes the
// diagram's vertices (their empty spheres are the candidate cavity balls),
// measures how enclosed each one is by casting rays from its centre
// (getHitRatios, :17-65, traced with OptiX sphere primitives by the
// ray-generation program of app/cavity/shaders: 64 directions, tmin 0.01,
// tmax 1e16, a ray counts when it hits any site) and keeps the vertices whose
// hit ratio and radius pass two thresholds (filterByHitRatios /
// filterByRadius / filterVertices, :244-264, :527-558).
//
// Here there is no ray-tracing hardware: each vertex is one wave64 and each
// lane traces one direction through the LBVH of the build (any hit ends the
// lane's walk; its stack lives in LDS as [depth][lane]).  Boxes are tested
// with the slab method, widened by a margin far above FP32 rounding, so a
// subtree that holds a sphere the ray hits is never skipped; the sphere test
// itself is the FP32 quadratic
//   b = (o - c).d, q = |o - c|^2 - r^2, s = sqrt(b^2 - q), t = -b -/+ s,
// a hit when either root lies in [0.01, 1e16].
#include "apo_device.h"
#include "apo_internal.h"

namespace apo_hip {

constexpr uint32_t kCavityStack = 64;
constexpr float kRayTMin = 0.01f;
constexpr float kRayTMax = 1e16f;

__device__ __forceinline__ bool raySphere(F3 o, F3 d, F4 s) {
    const F3 oc{o.x - s.x, o.y - s.y, o.z - s.z};
    const float b = oc.x * d.x + oc.y * d.y + oc.z * d.z;
    const float q = (oc.x * oc.x + oc.y * oc.y + oc.z * oc.z) - s.w * s.w;
    const float disc = b * b - q;
    if (disc < 0.f) return false;
    const float r = __builtin_sqrtf(disc);
    const float t1 = -b - r, t2 = -b + r;
    return (t1 >= kRayTMin && t1 <= kRayTMax) || (t2 >= kRayTMin && t2 <= kRayTMax);
}

// Slab test of the ray against a box widened by a margin: false only when
// no point of the box can be hit for t >= 0.
__device__ __forceinline__ bool rayBox(F3 o, F3 d, const Box& b) {
    const float m = 1e-5f * (1.f + __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(b.mn.x), __builtin_fabsf(b.mx.x)),
                                                   __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(b.mn.y),
                                                                                   __builtin_fabsf(b.mx.y)),
                                                                   __builtin_fmaxf(__builtin_fabsf(b.mn.z),
                                                                                   __builtin_fabsf(b.mx.z)))));
    float tn = 0.f, tf = kRayTMax;
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    const float lo[3] = {b.mn.x - m, b.mn.y - m, b.mn.z - m}, hi[3] = {b.mx.x + m, b.mx.y + m, b.mx.z + m};
    for (int a = 0; a < 3; a++) {
        if (dd[a] == 0.f) {
            if (oo[a] < lo[a] || oo[a] > hi[a]) return false;
            continue;
        }
        const float inv = 1.f / dd[a];
        const float t1 = (lo[a] - oo[a]) * inv, t2 = (hi[a] - oo[a]) * inv;
        tn = __builtin_fmaxf(tn, __builtin_fminf(t1, t2));
        tf = __builtin_fminf(tf, __builtin_fmaxf(t1, t2));
    }
    // widen the interval test by a relative margin too (FP32 slab rounding)
    return tn <= tf * (1.f + 1e-5f) + 1e-5f;
}

// One wave per vertex, one lane per direction (64 per pass).  The vertex is
// traced against its own molecule (the molecule of its first site).
__global__ __launch_bounds__(64) void k_hit_ratios(CellCtx c, const uint32_t* __restrict__ rank,
                                                   const D4* __restrict__ xyzw, const uint4* __restrict__ ids,
                                                   uint64_t count, const F3* __restrict__ dirs, uint32_t samples,
                                                   float* ratios) {
    __shared__ uint32_t stk[kCavityStack][64];
    const uint32_t lane = laneId();
    const F4* __restrict__ sorted = c.sorted;
    const Node* __restrict__ nodes = c.nodes;
    const uint32_t internal = c.internal, total = nodeCount(c);
    for (uint64_t v = blockIdx.x; v < count; v += gridDim.x) {
        const D4 x = xyzw[v];
        const F3 o{float(x.x), float(x.y), float(x.z)};
        const uint32_t first = ids[v].x;
        if (first >= c.n) {
            if (lane == 0) atomicOr(c.err, kErrCellSite);
            continue;
        }
        const uint32_t root = rootNode(c, cellMolecule(c, rank[first]));
        uint32_t hits = 0;
        for (uint32_t s0 = 0; s0 < samples; s0 += 64u) {
            const uint32_t s = s0 + lane;
            bool hit = false;
            if (s < samples) {
                const F3 d = dirs[s];
                uint32_t sp = 0, node = root;
                for (;;) {
                    bool next = false;
                    if (node < internal) {
                        const Node nn = nodes[node];
                        Box lb, rb;
                        nodeBoxes(nn, lb, rb);
                        const bool goL = rayBox(o, d, lb), goR = rayBox(o, d, rb);
                        if (goL || goR) {
                            if (goL && goR) {
                                if (sp < kCavityStack) stk[sp++][lane] = uint32_t(nn.d.y);
                                else atomicOr(c.err, kErrKnnChild);
                            }
                            node = uint32_t(goL ? nn.d.x : nn.d.y);
                            next = true;
                        }
                    } else if (node < total) {
                        hit = raySphere(o, d, sorted[node - internal]);
                        if (hit) break;
                    }
                    if (next) continue;
                    if (sp == 0) break;
                    node = stk[--sp][lane];
                }
            }
            hits += popc64(ballot(hit));
        }
        if (lane == 0) ratios[v] = float(hits) / float(samples);
    }
}

// Filter (filterByHitRatios and filterByRadius) and order-preserving
// compaction (filterVertices' copy_if): block counts, a scan, the writes.
__device__ __forceinline__ bool keepVertex(const D4& x, float ratio, float hitThreshold, float radiusThreshold) {
    return ratio > hitThreshold && float(x.w) > radiusThreshold;
}

__global__ __launch_bounds__(256) void k_cavity_count(const D4* __restrict__ xyzw, const float* __restrict__ ratios,
                                                      uint64_t count, float hitThreshold, float radiusThreshold,
                                                      uint32_t* blockSums) {
    const uint64_t v = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    const bool keep = v < count && keepVertex(xyzw[v], ratios[v], hitThreshold, radiusThreshold);
    __shared__ uint32_t s[4];
    const uint32_t w = popc64(ballot(keep));
    if (laneId() == 0) s[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) blockSums[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ __launch_bounds__(256) void k_cavity_write(const D4* __restrict__ xyzw, const float* __restrict__ ratios,
                                                      uint64_t count, float hitThreshold, float radiusThreshold,
                                                      const uint32_t* __restrict__ blockOffsets, float4* out,
                                                      uint64_t capacity) {
    const uint64_t v = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    D4 x{0.0, 0.0, 0.0, 0.0};
    if (v < count) x = xyzw[v];
    const bool keep = v < count && keepVertex(x, ratios[v], hitThreshold, radiusThreshold);
    const uint64_t m = ballot(keep);
    __shared__ uint32_t s[4];
    if (laneId() == 0) s[threadIdx.x >> 6] = popc64(m);
    __syncthreads();
    uint32_t off = blockOffsets[blockIdx.x];
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) off += s[w];
    off += countBelow(m);
    if (keep && off < capacity) out[off] = make_float4(float(x.x), float(x.y), float(x.z), float(x.w));
}

hipError_t hitRatiosLaunch(const CellCtx& c, const uint32_t* indices, uint32_t* rank, const D4* xyzw,
                           const uint4* ids, uint64_t count, const F3* dirs, uint32_t samples, float* ratios,
                           hipStream_t st) {
    if (samples == 0) return hipErrorInvalidValue;
    const hipError_t e = rankLaunch(indices, c.n, rank, st);
    if (e != hipSuccess || count == 0) return e;
    const uint32_t grid = count < (1u << 20) ? uint32_t(count) : (1u << 20);
    hipLaunchKernelGGL(k_hit_ratios, dim3(grid), dim3(64), 0, st, c, rank, xyzw, ids, count, dirs, samples, ratios);
    return hipGetLastError();
}

hipError_t cavityCountLaunch(const D4* xyzw, const float* ratios, uint64_t count, float hitThreshold,
                             float radiusThreshold, uint32_t* blockSums, uint32_t* total, hipStream_t st) {
    const uint32_t blocks = uint32_t((count + 255u) / 256u);
    hipLaunchKernelGGL(k_cavity_count, dim3(blocks), dim3(256), 0, st, xyzw, ratios, count, hitThreshold,
                       radiusThreshold, blockSums);
    return scanBlockSumsLaunch(blockSums, blocks, total, st);
}

hipError_t cavityWriteLaunch(const D4* xyzw, const float* ratios, uint64_t count, float hitThreshold,
                             float radiusThreshold, const uint32_t* blockOffsets, float4* out, uint64_t capacity,
                             hipStream_t st) {
    const uint32_t blocks = uint32_t((count + 255u) / 256u);
    hipLaunchKernelGGL(k_cavity_write, dim3(blocks), dim3(256), 0, st, xyzw, ratios, count, hitThreshold,
                       radiusThreshold, blockOffsets, out, capacity);
    return hipGetLastError();
}

}  // namespace apo_hip
