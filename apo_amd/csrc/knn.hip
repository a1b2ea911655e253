// knn.hip -- k nearest sites by sphere gap over the LBVH, one lane per site.
//
// Reference: apo/include/apo/gpu/lbvh.inl:8-105 (closest-child-first
// stackful traversal), apo/include/apo/gpu/heap.hpp:11-43 (bounded max-heap,
// heapsort) and apo/include/apo/gpu/algorithm.inl:32-99 (getKnns with the
// buried test on the first pass).  The traversal order, the strict
// acceptance tests and the heap moves are those of the reference, so the
// neighbour lists (including the order of equal gaps) are identical.
//
// Layout: the heap (keys, gaps) and the traversal stack (node, gap bits) of
// every lane sit in LDS as [slot][lane] arrays, so any mix of per-lane slot
// indices is bank-conflict free.  The result is written through LDS so the
// global stores are coalesced.
#include "apo_device.h"
#include "apo_internal.h"

namespace apo_hip {

// One pass of getKnns for 64 sites.  KT bounds the list length k in LDS;
// pass > 0 continues each list after the previous pass's last neighbour
// (algorithm.inl:43-60): the lists are read and rewritten in place.
// bound (sorted order, may be null; first pass only): the list of site i
// holds only sites whose gap is below rangeTop(bound[i], r_i) (the heap
// starts full of that gap instead of FLT_MAX, so the traversal also prunes
// beyond it).
template <uint32_t KT>
__global__ __launch_bounds__(64) void k_knn(CellCtx c, uint32_t* __restrict__ knn, uint32_t K, uint32_t pass,
                                            const float* __restrict__ bound) {
    __shared__ uint32_t hKey[KT][64];
    __shared__ float hVal[KT][64];
    // stacked entries: parent node << 1 | the stacked child's side; its gap
    // is recomputed from the parent's copy of the child box when popped
    // (the FP32 value the reference stacks), so the stack takes half the LDS
    __shared__ uint32_t stNode[kStackDepth][64];

    const uint32_t lane = laneId();
    const uint32_t base = blockIdx.x * 64u;
    const uint32_t i = base + lane;
    const uint32_t n = c.n;
    bool active = i < n;

    float top = FLT_MAX;
    if (bound && active) top = rangeTop(bound[i], c.sorted[i].w);
    for (uint32_t k = 0; k < K; k++) {
        hKey[k][lane] = kNone;
        hVal[k][lane] = top;
    }
    uint32_t overflow = 0;
    float startingDistance = -FLT_MAX;
    if (active && pass > 0) {
        // a list the previous pass did not fill ends here (0xffffffff x K)
        const uint32_t last = knn[size_t(i) * K + K - 1u];
        if (last < n) startingDistance = sphereDistF(c.sorted[last], c.sorted[i]);
        else active = false;
    }
    if (active) {
        const F4 s = c.sorted[i];
        uint32_t sp = 0;
        uint32_t node = rootNode(c, cellMolecule(c, i));
        const uint32_t internal = c.internal, nodes = nodeCount(c);
        while (true) {
            while (node < internal) {
                const Node nd = c.nodes[node];
                Box lb, rb;
                nodeBoxes(nd, lb, rb);
                const uint32_t L = uint32_t(nd.d.x), R = uint32_t(nd.d.y);
                if (L >= nodes || R >= nodes) {
                    atomicOr(c.err, kErrKnnChild);
                    node = nodes;  // neither internal nor leaf: ends this descent
                    break;
                }
                const float dL = boxDistance(lb, s), dLmax = boxFarthestDistance(lb, s);
                const float dR = boxDistance(rb, s), dRmax = boxFarthestDistance(rb, s);
                const bool leftClosest = dL < dR;
                const float minFar = leftClosest ? dR : dL;
                const float maxFar = leftClosest ? dRmax : dLmax;
                const uint32_t farChild = leftClosest ? R : L;
                const float minClose = leftClosest ? dL : dR;
                const float maxClose = leftClosest ? dLmax : dRmax;
                const uint32_t closeChild = leftClosest ? L : R;
                const float root = hVal[0][lane];
                const bool exploreFar = minFar < root && maxFar > startingDistance;
                const bool exploreClose = minClose < root && maxClose > startingDistance;
                if (!exploreFar && !exploreClose) break;
                const uint32_t here = node;
                node = exploreClose ? closeChild : farChild;
                if (exploreClose && exploreFar) {
                    if (sp < kStackDepth) {
                        stNode[sp][lane] = (here << 1) | (leftClosest ? 1u : 0u);
                        sp++;
                    } else {
                        overflow = 1;
                    }
                }
            }
            if (node >= internal && node < nodes) {
                const uint32_t j = node - internal;
                const float d = sphereDistF(s, c.sorted[j]);
                if (j != i && d < hVal[0][lane] && d > startingDistance) {
                    // replace the root and sift down (heap.hpp:11-31)
                    hVal[0][lane] = d;
                    hKey[0][lane] = j;
                    uint32_t p = 0;
                    while (true) {
                        const uint32_t l = 2 * p + 1, r = 2 * p + 2;
                        uint32_t largest = p;
                        float vl = hVal[p][lane];
                        if (l < K && hVal[l][lane] > vl) {
                            largest = l;
                            vl = hVal[l][lane];
                        }
                        if (r < K && hVal[r][lane] > vl) largest = r;
                        if (largest == p) break;
                        const float tv = hVal[p][lane];
                        hVal[p][lane] = hVal[largest][lane];
                        hVal[largest][lane] = tv;
                        const uint32_t tk = hKey[p][lane];
                        hKey[p][lane] = hKey[largest][lane];
                        hKey[largest][lane] = tk;
                        p = largest;
                    }
                }
            }
            bool found = false;
            while (sp > 0) {
                sp--;
                const uint32_t e = stNode[sp][lane];
                const Node pn = c.nodes[e >> 1];
                Box lb, rb;
                nodeBoxes(pn, lb, rb);
                const bool right = (e & 1u) != 0u;
                if (boxDistance(selBox(right, rb, lb), s) > hVal[0][lane]) continue;
                node = uint32_t(right ? pn.d.y : pn.d.x);
                found = true;
                break;
            }
            if (!found) break;
        }
        // heapsort into ascending gaps (heap.hpp:33-43)
        for (uint32_t size = K; size > 0;) {
            const float tv = hVal[0][lane];
            hVal[0][lane] = hVal[size - 1][lane];
            hVal[size - 1][lane] = tv;
            const uint32_t tk = hKey[0][lane];
            hKey[0][lane] = hKey[size - 1][lane];
            hKey[size - 1][lane] = tk;
            --size;
            uint32_t p = 0;
            while (true) {
                const uint32_t l = 2 * p + 1, r = 2 * p + 2;
                uint32_t largest = p;
                if (l < size && hVal[l][lane] > hVal[largest][lane]) largest = l;
                if (r < size && hVal[r][lane] > hVal[largest][lane]) largest = r;
                if (largest == p) break;
                const float t2 = hVal[p][lane];
                hVal[p][lane] = hVal[largest][lane];
                hVal[largest][lane] = t2;
                const uint32_t k2 = hKey[p][lane];
                hKey[p][lane] = hKey[largest][lane];
                hKey[largest][lane] = k2;
                p = largest;
            }
        }
        if (pass > 0) {
            if (overflow) c.flags[i] |= kFlagStackOverflow;
        } else {
            // buried test against the nearest site, first pass only (algorithm.inl:85-98)
            const uint32_t j0 = hKey[0][lane];
            uint8_t status = kStatusInValidation;
            if (j0 != kNone) {
                const F4 cl = c.sorted[j0];
                const double ri = s.w, rj = cl.w;
                const double bigger = __builtin_fmax(ri, rj), smaller = __builtin_fmin(ri, rj);
                const D3 dc{double(s.x) - double(cl.x), double(s.y) - double(cl.y), double(s.z) - double(cl.z)};
                if (ltD(len(dc), bigger - smaller) && !ltD(rj, ri)) status = kStatusBuried;
            }
            // the kNN pass is the first writer of status and flags
            c.status[i] = status;
            c.flags[i] = overflow ? kFlagStackOverflow : 0u;
        }
    }
    waveSync();
    // coalesced write of the 64 x K block
    const uint32_t rows = (n - base) < 64u ? (n - base) : 64u;
    for (uint32_t t = lane; t < rows * K; t += 64u) {
        const uint32_t row = t / K, col = t % K;
        knn[size_t(base + row) * K + col] = hKey[col][row];
    }
}

hipError_t knnLaunch(const CellCtx& c, uint32_t k, uint32_t* knn, hipStream_t st, uint32_t pass,
                     const float* bound) {
    if (bound && pass) return hipErrorInvalidValue;
    const uint32_t blocks = (c.n + 63u) / 64u;
    if (k == 0 || k > 64) return hipErrorInvalidValue;
    // the smallest LDS list that holds k entries
    if (k <= 8) hipLaunchKernelGGL(k_knn<8>, dim3(blocks), dim3(64), 0, st, c, knn, k, pass, bound);
    else if (k <= 16) hipLaunchKernelGGL(k_knn<16>, dim3(blocks), dim3(64), 0, st, c, knn, k, pass, bound);
    else if (k <= 32) hipLaunchKernelGGL(k_knn<32>, dim3(blocks), dim3(64), 0, st, c, knn, k, pass, bound);
    else hipLaunchKernelGGL(k_knn<64>, dim3(blocks), dim3(64), 0, st, c, knn, k, pass, bound);
    return hipGetLastError();
}

}  // namespace apo_hip
