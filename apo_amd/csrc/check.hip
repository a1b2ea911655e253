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

__global__ void k_rank(const uint32_t* __restrict__ indices, uint32_t n, uint32_t* rank) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) rank[indices[j]] = j;
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
