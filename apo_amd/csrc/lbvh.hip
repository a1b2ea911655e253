// lbvh.hip -- LBVH build for gfx950: scene box, Morton codes, onesweep LSD
// radix sort, Karras hierarchy, atomic bottom-up refit (AABB + max radius).
//
// Reference: apo/src/gpu/lbvh.cu:11-323 (Morton, setLeaves, construct,
// refit) and apo/src/gpu/algorithm.cpp:5-25 (scene box).  The sort replaces
// thrust::sort_by_key (lbvh.cu:294-295) with a stable hand-written onesweep.
#include "apo_device.h"
#include "apo_internal.h"

namespace apo_hip {

// ---------------------------------------------------------------------------
// Scene box: min/max of sphere extents in FP64 (algorithm.cpp:5-25), then
// the float box, Morton scale and the 6 artificial sites (algorithm.inl:1437-1464)
// computed by the last block to finish.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t orderedBits(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double fromOrderedBits(uint64_t o) {
    const uint64_t b = (o >> 63) ? (o & 0x7fffffffffffffffull) : ~o;
    return __builtin_bit_cast(double, b);
}

__device__ __forceinline__ uint32_t pcgWord(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Scene record from the 8 reduced values (min x-r, y-r, z-r, r; max x+r,
// y+r, z+r, r): FP32 Morton frame and the 6 artificial sites.
__device__ Scene makeScene(const double* r) {
    Scene sc;
    for (int a = 0; a < 3; a++) {
        sc.mn[a] = r[a];
        sc.mx[a] = r[4 + a];
        sc.fmn[a] = float(r[a]);
        sc.fmx[a] = float(r[4 + a]);
        sc.scale[a] = 1.f / (sc.fmx[a] - sc.fmn[a]);
    }
    sc.rmin = r[3];
    sc.rmax = r[7];
    // artificial sites: centroid +- R * axis, w = min radius, joggled by 0.1
    const D3 cc{(sc.mn[0] + sc.mx[0]) / 2.0, (sc.mn[1] + sc.mx[1]) / 2.0, (sc.mn[2] + sc.mx[2]) / 2.0};
    const double aabbRadius = len(sub(cc, D3{sc.mx[0], sc.mx[1], sc.mx[2]}));
    const double radius = aabbRadius * (__builtin_sqrt(3.0) + 1e-1);
    const double w = sc.rmin;
    const double offs[6][4] = {{radius, 0, 0, w}, {-radius, 0, 0, w}, {0, radius, 0, w},
                               {0, -radius, 0, w}, {0, 0, radius, w}, {0, 0, -radius, w}};
    for (int s = 0; s < 6; s++) {
        double v[4] = {cc.x + offs[s][0], cc.y + offs[s][1], cc.z + offs[s][2], 0.0 + offs[s][3]};
        for (int c = 0; c < 4; c++)
            v[c] += __builtin_ldexp(double(pcgWord(uint32_t(s * 4 + c))), -32) * (2.0 * 1e-1) + (-1e-1);
        sc.art[s] = {v[0], v[1], v[2], v[3]};
    }
    return sc;
}

// Sphere extent of one site folded into the running min / max.
__device__ __forceinline__ void accumulateExtent(const F4 s, double* mn, double* mx) {
    const double r = s.w;
    const double c[3] = {s.x, s.y, s.z};
    for (int a = 0; a < 3; a++) {
        mn[a] = __builtin_fmin(mn[a], c[a] - r);
        mx[a] = __builtin_fmax(mx[a], c[a] + r);
    }
    mn[3] = __builtin_fmin(mn[3], r);
    mx[3] = __builtin_fmax(mx[3], r);
}

// Block-wide min / max of a 256-thread block; thread 0 receives r[8].
__device__ __forceinline__ void blockExtent(double* mn, double* mx, double* r) {
    for (int a = 0; a < 4; a++) {
        for (int o = 32; o >= 1; o >>= 1) {
            mn[a] = __builtin_fmin(mn[a], __shfl_xor(mn[a], o, 64));
            mx[a] = __builtin_fmax(mx[a], __shfl_xor(mx[a], o, 64));
        }
    }
    __shared__ double sred[4][8];
    const uint32_t wave = threadIdx.x / 64, lane = laneId();
    if (lane == 0)
        for (int a = 0; a < 4; a++) {
            sred[wave][a] = mn[a];
            sred[wave][4 + a] = mx[a];
        }
    __syncthreads();
    for (int a = 0; a < 8; a++) r[a] = sred[0][a];
    for (uint32_t w = 1; w < 4; w++)
        for (int a = 0; a < 4; a++) {
            r[a] = __builtin_fmin(r[a], sred[w][a]);
            r[4 + a] = __builtin_fmax(r[4 + a], sred[w][4 + a]);
        }
}

__global__ __launch_bounds__(256) void k_scene_box(const F4* __restrict__ sites, uint32_t n, Scene* scene,
                                                   double* partials, uint32_t* ticket, uint32_t* zero,
                                                   uint32_t zeroWords) {
    // 8 reduced values: min x-r, y-r, z-r, r ; max x+r, y+r, z+r, r
    double mn[4] = {DBL_MAX, DBL_MAX, DBL_MAX, DBL_MAX};
    double mx[4] = {-DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        accumulateExtent(sites[i], mn, mx);
    }
    // clear the scratch words that later kernels accumulate into
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < zeroWords; w += stride) zero[w] = 0u;

    double r[8];
    blockExtent(mn, mx, r);
    if (threadIdx.x == 0)
        for (int a = 0; a < 8; a++) partials[blockIdx.x * 8 + a] = r[a];
}

// Reduction of the block partials (one block of kBoxBlocks threads) and the
// scene record: FP32 Morton frame and the 6 artificial sites.
__global__ __launch_bounds__(kBoxBlocks) void k_scene_final(const double* __restrict__ partials, Scene* scene) {
    double v[8];
    for (int a = 0; a < 8; a++) v[a] = partials[threadIdx.x * 8 + a];
    for (int a = 0; a < 4; a++) {
        for (int o = 32; o >= 1; o >>= 1) {
            v[a] = __builtin_fmin(v[a], __shfl_xor(v[a], o, 64));
            v[4 + a] = __builtin_fmax(v[4 + a], __shfl_xor(v[4 + a], o, 64));
        }
    }
    __shared__ double sred[kBoxBlocks / 64][8];
    const uint32_t wave = threadIdx.x / 64, lane = laneId();
    if (lane == 0)
        for (int a = 0; a < 8; a++) sred[wave][a] = v[a];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double r[8];
    for (int a = 0; a < 8; a++) r[a] = sred[0][a];
    for (uint32_t w = 1; w < kBoxBlocks / 64; w++)
        for (int a = 0; a < 4; a++) {
            r[a] = __builtin_fmin(r[a], sred[w][a]);
            r[4 + a] = __builtin_fmax(r[4 + a], sred[w][4 + a]);
        }
    *scene = makeScene(r);
}

// Scene records of a batch: block m reduces molecule m's sites and writes
// scene[m]; the blocks also clear the scratch words later kernels
// accumulate into.
__global__ __launch_bounds__(256) void k_scene_segments(const F4* __restrict__ sites, const uint32_t* __restrict__ offsets,
                                                        Scene* scene, uint32_t* zero, uint32_t zeroWords) {
    const uint32_t m = blockIdx.x;
    const uint32_t lo = offsets[m], hi = offsets[m + 1];
    double mn[4] = {DBL_MAX, DBL_MAX, DBL_MAX, DBL_MAX};
    double mx[4] = {-DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        accumulateExtent(sites[i], mn, mx);
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < zeroWords; w += stride) zero[w] = 0u;
    double r[8];
    blockExtent(mn, mx, r);
    if (threadIdx.x == 0) scene[m] = makeScene(r);
}

// ---------------------------------------------------------------------------
// Morton codes (lbvh.cu:11-67): 20 bits per axis, 60-bit code, with the 8
// digit histograms of the radix sort accumulated on the fly.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t expandBits10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t quantize(float v) {
    if (!(v >= 0.f)) return 0u;
    if (v >= 4294967040.f) return 0xffffffffu;
    return uint32_t(v);
}
__device__ __forceinline__ uint64_t mortonCode(F4 s, const float* fmn, const float* scale) {
    uint32_t u[3];
    const float c[3] = {s.x, s.y, s.z};
    for (int a = 0; a < 3; a++) {
        const float q = (c[a] - fmn[a]) * scale[a];
        u[a] = quantize(q * 1048575.0f);
    }
    const uint64_t lo = expandBits10(u[0]) | (expandBits10(u[1]) << 1) | (expandBits10(u[2]) << 2);
    const uint64_t hi = expandBits10(u[0] >> 10) | (expandBits10(u[1] >> 10) << 1) | (expandBits10(u[2] >> 10) << 2);
    return (hi << 30) | lo;
}

// Sort digit of pass p: byte p of the key for p < 8; the molecule index's
// bits 4-11 and 12-19 for passes 8 and 9.
__device__ __forceinline__ uint32_t moleculeDigit(uint32_t m, int p) { return (m >> (4 + 8 * (p - 8))) & 255u; }

// One molecule: codes in the molecule's frame.  A batch: site i is coded in
// the frame of its molecule m, and m's low 4 bits go to key bits 60-63.
__global__ __launch_bounds__(256) void k_morton(const F4* __restrict__ sites, uint32_t n, Segments seg, int passes,
                                                const Scene* __restrict__ scene, uint64_t* __restrict__ codes,
                                                uint32_t* __restrict__ indices, uint32_t* __restrict__ hist) {
    __shared__ uint32_t sh[kMaxRadixPasses][256];
    for (uint32_t t = threadIdx.x; t < kMaxRadixPasses * 256; t += blockDim.x) (&sh[0][0])[t] = 0u;
    __syncthreads();
    float fmn[3], scale[3];
    for (int a = 0; a < 3; a++) {
        fmn[a] = scene[0].fmn[a];
        scale[a] = scene[0].scale[a];
    }
    const bool batch = seg.molecules > 1u;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t m = 0;
        if (batch) {
            m = moleculeOf(seg.offsets, seg.molecules, i);
            for (int a = 0; a < 3; a++) {
                fmn[a] = scene[m].fmn[a];
                scale[a] = scene[m].scale[a];
            }
        }
        const uint64_t code = mortonCode(sites[i], fmn, scale) | (uint64_t(m & 15u) << 60);
        codes[i] = code;
        indices[i] = i;
        for (int p = 0; p < kRadixPasses; p++) atomicAdd(&sh[p][(code >> (8 * p)) & 255u], 1u);
        for (int p = kRadixPasses; p < passes; p++) atomicAdd(&sh[p][moleculeDigit(m, p)], 1u);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < uint32_t(passes) * 256u; t += blockDim.x) {
        const uint32_t v = (&sh[0][0])[t];
        if (v) atomicAdd(&hist[t], v);
    }
}

// ---------------------------------------------------------------------------
// Onesweep LSD radix sort pass: 8-bit digit, 4096 keys per tile (4 waves x
// 16 rounds x 64 lanes).  Ranking inside a wave uses 8 ballots to find the
// lanes holding the same digit, so equal digits keep input order (stable).
// Tiles take an index from an atomic counter and chain their per-digit
// prefixes with decoupled look-back.
// ---------------------------------------------------------------------------
constexpr uint32_t kTileKeys = 4096;
constexpr uint32_t kFlagAgg = 1u << 30, kFlagInc = 2u << 30, kValueMask = (1u << 30) - 1u;

// Digit of a key / value pair in pass p: a key byte, or (kMol) a digit of
// the molecule of the value (an original site index).
template <bool kMol>
__device__ __forceinline__ uint32_t sortDigit(uint64_t key, uint32_t val, int p, const Segments& seg) {
    if (kMol) return moleculeDigit(moleculeOf(seg.offsets, seg.molecules, val), p);
    return uint32_t(key >> (8 * p)) & 255u;
}

template <bool kMol>
__global__ __launch_bounds__(256) void k_onesweep(const uint64_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                                  uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut,
                                                  uint32_t n, int pass, Segments seg,
                                                  const uint32_t* __restrict__ digitHist, uint32_t* lookback,
                                                  uint32_t* tileCounter, uint32_t numTiles, uint32_t* err) {
    __shared__ uint64_t sKeys[kTileKeys];
    __shared__ uint32_t sVals[kTileKeys];
    __shared__ uint32_t sWave[4][256];
    __shared__ uint32_t sTileStart[256];
    __shared__ uint32_t sGlobal[256];
    __shared__ uint32_t sScan[4];
    __shared__ uint32_t sScan2[4];
    __shared__ uint32_t sTile;

    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = laneId();
    if (tid == 0) sTile = atomicAdd(tileCounter, 1u);
    sWave[0][tid] = sWave[1][tid] = sWave[2][tid] = sWave[3][tid] = 0u;
    __syncthreads();
    const uint32_t tile = sTile;
    if (tile >= numTiles) {
        // more blocks than tiles cannot happen with grid == numTiles; report it
        if (tid == 0) atomicOr(err, kErrSortTile);
        return;
    }
    const uint32_t base = tile * kTileKeys + wave * 1024u;

    uint64_t key[16];
    uint32_t val[16];
    uint32_t rank[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t idx = base + uint32_t(r) * 64u + lane;
        const bool valid = idx < n;
        key[r] = valid ? keysIn[idx] : 0ull;
        val[r] = valid ? valsIn[idx] : 0u;
        const uint32_t d = sortDigit<kMol>(key[r], val[r], pass, seg);
        uint64_t peers = ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bb = ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t below = countBelow(peers);
        const uint32_t prev = sWave[wave][d];
        waveSync();
        if (valid && below == 0) sWave[wave][d] = prev + popc64(peers);
        waveSync();
        rank[r] = prev + below;
    }
    __syncthreads();

    // per-digit tile totals, per-wave exclusive offsets
    const uint32_t d = tid;
    const uint32_t c0 = sWave[0][d], c1 = sWave[1][d], c2 = sWave[2][d], c3 = sWave[3][d];
    const uint32_t tileCount = c0 + c1 + c2 + c3;
    // publish the tile aggregate, then look back for the exclusive prefix
    uint32_t* lb = lookback + size_t(tile) * 256u + d;
    uint32_t excl = 0;
    if (tile == 0) {
        __hip_atomic_store(lb, kFlagInc | tileCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(lb, kFlagAgg | tileCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int t = int(tile) - 1;
        uint32_t spins = 0;
        while (t >= 0) {
            const uint32_t v = __hip_atomic_load(lookback + size_t(t) * 256u + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t flag = v & ~kValueMask;
            if (flag == 0u) {
                if (++spins > (1u << 22)) {
                    atomicOr(err, kErrSortSpin);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += v & kValueMask;
            if (flag == kFlagInc) break;
            t--;
        }
        __hip_atomic_store(lb, kFlagInc | (excl + tileCount), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // exclusive scans over the 256 digits: global digit starts and tile digit starts
    const uint32_t g = digitHist[d];
    uint32_t gInc = waveInclusiveSum(g);
    uint32_t tInc = waveInclusiveSum(tileCount);
    if (lane == 63) {
        sScan[wave] = gInc;
        sScan2[wave] = tInc;
    }
    sWave[1][d] = c0;
    sWave[2][d] = c0 + c1;
    sWave[3][d] = c0 + c1 + c2;
    sWave[0][d] = 0u;
    __syncthreads();
    uint32_t gOff = 0, tOff = 0;
    for (uint32_t w = 0; w < wave; w++) {
        gOff += sScan[w];
        tOff += sScan2[w];
    }
    sGlobal[d] = gOff + gInc - g + excl;
    sTileStart[d] = tOff + tInc - tileCount;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t idx = base + uint32_t(r) * 64u + lane;
        if (idx < n) {
            const uint32_t dg = sortDigit<kMol>(key[r], val[r], pass, seg);
            const uint32_t pos = sTileStart[dg] + sWave[wave][dg] + rank[r];
            if (pos < kTileKeys) {
                sKeys[pos] = key[r];
                sVals[pos] = val[r];
            } else {
                atomicOr(err, kErrSortScatter);
            }
        }
    }
    __syncthreads();
    const uint32_t tileBase = tile * kTileKeys;
    const uint32_t tileValid = (n - tileBase) < kTileKeys ? (n - tileBase) : kTileKeys;
    for (uint32_t p = tid; p < tileValid; p += 256u) {
        const uint64_t k = sKeys[p];
        const uint32_t v = sVals[p];
        const uint32_t dg = sortDigit<kMol>(k, v, pass, seg);
        const uint32_t out = sGlobal[dg] + (p - sTileStart[dg]);
        if (out < n) {
            keysOut[out] = k;
            valsOut[out] = v;
        } else {
            atomicOr(err, kErrSortScatter);
        }
    }
}

// ---------------------------------------------------------------------------
// Karras 2012 hierarchy (lbvh.cu:144-223) + gather of the sorted sites
// (setLeaves, lbvh.cu:88-114).  Thread i handles internal node i and leaf i.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int deltaK(const uint64_t* __restrict__ c, int i, int j, int n) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = c[i], b = c[j];
    if (a != b) return __builtin_clzll(a ^ b);
    return __builtin_clzll(uint64_t(int64_t(i ^ j))) + 64;
}

__global__ __launch_bounds__(256) void k_karras(const uint64_t* __restrict__ codes, const uint32_t* __restrict__ indices,
                                                const F4* __restrict__ sitesIn, uint32_t n, Segments seg,
                                                F4* __restrict__ sorted, Node* __restrict__ nodes,
                                                int32_t* __restrict__ parent, uint32_t* __restrict__ refitCounter,
                                                uint32_t* __restrict__ sizes, uint32_t* __restrict__ range,
                                                uint32_t* __restrict__ crossList, uint32_t* crossCount, uint32_t* err) {
    const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= n) return;
    const uint32_t src = indices[gi];
    if (src < n) {
        sorted[gi] = sitesIn[src];
    } else {
        sorted[gi] = F4{0.f, 0.f, 0.f, 0.f};
        atomicOr(err, kErrGatherIndex);
    }
    const uint32_t internal = n - seg.molecules;
    sizes[internal + gi] = 1u;
    // the molecule's segment: Karras runs on local indices, so each
    // molecule's tree is the one a build of that molecule alone emits
    uint32_t m = 0, base = 0, count = n;
    if (seg.molecules > 1u) {
        m = moleculeOf(seg.offsets, seg.molecules, gi);
        base = seg.offsets[m];
        count = seg.offsets[m + 1] - base;
    }
    const int i = int(gi - base);
    const int ni = int(count);
    if (i >= ni - 1) return;
    const uint64_t* c = codes + base;
    const uint32_t nodeBase = base - m;  // the molecule's first internal node (its root)
    const uint32_t g = nodeBase + uint32_t(i);
    refitCounter[g] = 0u;
    const int dd = deltaK(c, i, i + 1, ni) - deltaK(c, i, i - 1, ni);
    const int d = (dd > 0) - (dd < 0);
    const int deltaMin = deltaK(c, i, i - d, ni);
    int lmax = 2;
    while (deltaK(c, i, i + lmax * d, ni) > deltaMin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (deltaK(c, i, i + (l + t) * d, ni) > deltaMin) l += t;
    const int j = i + l * d;
    const int deltaNode = deltaK(c, i, j, ni);
    int s = 0, k = 2, t;
    do {
        t = (l + k - 1) / k;
        k <<= 1;
        if (deltaK(c, i, i + (s + t) * d, ni) > deltaNode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    // a child whose range is one leaf is that leaf, else the internal node
    // with the same local index
    const uint32_t lc = (i < j ? i : j) == gamma ? internal + base + uint32_t(gamma) : nodeBase + uint32_t(gamma);
    const uint32_t rc = (i > j ? i : j) == gamma + 1 ? internal + base + uint32_t(gamma) + 1u
                                                     : nodeBase + uint32_t(gamma) + 1u;
    nodes[g].d.x = int32_t(lc);
    nodes[g].d.y = int32_t(rc);
    parent[lc] = int32_t(g);
    parent[rc] = int32_t(g);
    if (i == 0) parent[g] = -1;
    // leaf range (a node's range starts or ends at its own local index)
    const uint32_t first = base + uint32_t(i < j ? i : j), last = base + uint32_t(i > j ? i : j);
    range[2 * g] = first;
    range[2 * g + 1] = last;
    sizes[g] = last - first + 1u;
    if (first / kRefitLeaves != last / kRefitLeaves) crossList[atomicAdd(crossCount, 1u)] = g;
}

// ---------------------------------------------------------------------------
// Bottom-up refit (lbvh.cu:225-272): one thread per leaf climbs; the second
// thread to reach a node unions both child boxes.  Child boxes and max radii
// are written into the parent's 64-byte record.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void writeSlot(Node* nd, bool left, const float* mn, const float* mx, float maxr) {
    float* f = reinterpret_cast<float*>(nd);
    if (left) {
        f[0] = mn[0]; f[1] = mn[1]; f[2] = mn[2];
        f[3] = mx[0]; f[4] = mx[1]; f[5] = mx[2];
        f[14] = maxr;
    } else {
        f[6] = mn[0]; f[7] = mn[1]; f[8] = mn[2];
        f[9] = mx[0]; f[10] = mx[1]; f[11] = mx[2];
        f[15] = maxr;
    }
}

__device__ __forceinline__ void leafBox(const F4& s, float* mn, float* mx) {
    const double x = s.x, y = s.y, z = s.z, r = s.w;
    mn[0] = float(x - r);
    mn[1] = float(y - r);
    mn[2] = float(z - r);
    mx[0] = float(x + r);
    mx[1] = float(y + r);
    mx[2] = float(z + r);
}

// Union of both child slots of a node record.
__device__ __forceinline__ void nodeUnion(const Node* P, float* mn, float* mx, float& maxr) {
    const float* f = reinterpret_cast<const float*>(P);
    for (int c = 0; c < 3; c++) {
        mn[c] = __builtin_fminf(f[c], f[6 + c]);
        mx[c] = __builtin_fmaxf(f[3 + c], f[9 + c]);
    }
    maxr = __builtin_fmaxf(f[14], f[15]);
}

__device__ __forceinline__ bool crossesBlocks(const uint32_t* range, uint32_t node) {
    return range[2 * node] / kRefitLeaves != range[2 * node + 1] / kRefitLeaves;
}

// Phase 1: one block per kRefitLeaves consecutive leaves climbs through the
// nodes whose leaves all lie in its range.  The first child to arrive at a
// node writes its slot and stops; the second unions both slots and climbs
// on.  Arrivals are counted in LDS, at the node's split (the last leaf of
// its left child), which is unique among the block's nodes; slots travel
// through the node records with workgroup-scope ordering only (no
// device-wide fence).  A node whose leaves span two blocks only receives its
// child's slot here.  Trees of a batch never share a node, so each climb
// ends at its molecule's root.
__global__ __launch_bounds__(512) void k_refit_local(const F4* __restrict__ sorted, uint32_t n, uint32_t internal,
                                                    Node* nodes, const int32_t* __restrict__ parent,
                                                    const uint32_t* __restrict__ range, float* rootBox, uint32_t* err) {
    __shared__ uint32_t arrivals[kRefitLeaves];
    const uint32_t lo = blockIdx.x * kRefitLeaves;
    const uint32_t hi = (n - lo) < kRefitLeaves ? n : lo + kRefitLeaves;
    const bool single = internal + 1u == n;
    for (uint32_t t = threadIdx.x; t < kRefitLeaves; t += blockDim.x) arrivals[t] = 0u;
    __syncthreads();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        float mn[3], mx[3];
        leafBox(sorted[i], mn, mx);
        float maxr = sorted[i].w;
        uint32_t node = internal + i, first = i, last = i;
        for (uint32_t climb = 0; climb < n; climb++) {
            const int p = parent[node];
            if (p < 0) {
                if (node >= internal || (single && node != 0u)) {
                    atomicOr(err, kErrRefitParent);
                } else if (single) {
                    for (int c = 0; c < 3; c++) {
                        rootBox[c] = mn[c];
                        rootBox[3 + c] = mx[c];
                    }
                    rootBox[6] = maxr;
                }
                break;
            }
            if (uint32_t(p) >= internal) {
                atomicOr(err, kErrRefitParent);
                break;
            }
            Node* P = nodes + p;
            const bool left = uint32_t(P->d.x) == node;
            writeSlot(P, left, mn, mx, maxr);
            const uint32_t pf = range[2 * p], pl = range[2 * p + 1];
            if (pf < lo || pl >= hi) break;  // finished by the second phase
            const uint32_t split = left ? last : first - 1u;
            if (split < pf || split >= pl) {
                atomicOr(err, kErrRefitParent);
                break;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            const uint32_t old = atomicAdd(&arrivals[split - lo], 1u);
            if (old == 0u) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            nodeUnion(P, mn, mx, maxr);
            node = uint32_t(p);
            first = pf;
            last = pl;
        }
    }
}

// Phase 2: one block finishes the nodes whose leaves span two phase-1
// blocks.  Climbs start at those with no such child; a node continues once
// all its spanning children have arrived.  One block, so workgroup-scope
// ordering suffices.
__global__ __launch_bounds__(1024) void k_refit_top(uint32_t n, uint32_t internal, Node* nodes,
                                                   const int32_t* __restrict__ parent,
                                                   const uint32_t* __restrict__ range,
                                                   const uint32_t* __restrict__ crossList,
                                                   const uint32_t* __restrict__ crossCount, uint32_t* refitCounter,
                                                   float* rootBox, uint32_t* err) {
    const bool single = internal + 1u == n;
    const uint32_t count = *crossCount < internal ? *crossCount : internal;
    auto spanningChildren = [&](uint32_t q) {
        const uint32_t l = uint32_t(nodes[q].d.x), r = uint32_t(nodes[q].d.y);
        return uint32_t(l < internal && crossesBlocks(range, l)) + uint32_t(r < internal && crossesBlocks(range, r));
    };
    for (uint32_t u = threadIdx.x; u < count; u += blockDim.x) {
        uint32_t node = crossList[u];
        if (node >= internal) {
            atomicOr(err, kErrRefitParent);
            continue;
        }
        if (spanningChildren(node) != 0u) continue;
        for (uint32_t climb = 0; climb < n; climb++) {
            float mn[3], mx[3], maxr;
            nodeUnion(nodes + node, mn, mx, maxr);
            const int p = parent[node];
            if (p < 0) {
                if (single) {
                    if (node != 0u) atomicOr(err, kErrRefitParent);
                    for (int c = 0; c < 3; c++) {
                        rootBox[c] = mn[c];
                        rootBox[3 + c] = mx[c];
                    }
                    rootBox[6] = maxr;
                }
                break;
            }
            if (uint32_t(p) >= internal) {
                atomicOr(err, kErrRefitParent);
                break;
            }
            Node* P = nodes + p;
            writeSlot(P, P->d.x == int(node), mn, mx, maxr);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            const uint32_t old = atomicAdd(&refitCounter[p], 1u);
            if (old + 1u < spanningChildren(uint32_t(p))) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            node = uint32_t(p);
        }
    }
}

__global__ __launch_bounds__(256) void k_copy_sorted(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                     uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        kout[i] = kin[i];
        vout[i] = vin[i];
    }
}

// Depth of the deepest leaf (root depth 0).  A depth-first traversal keeps
// at most one stacked node per level of its current path, so no stack of
// this depth or more can overflow.
__global__ __launch_bounds__(256) void k_tree_height(const int32_t* __restrict__ parent, uint32_t n,
                                                     uint32_t internal, uint32_t* height) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t d = 0;
    if (i < n)
        for (int32_t p = parent[internal + i]; p >= 0 && d < n; p = parent[p]) d++;
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t other = __shfl_xor(d, o, 64);
        d = other > d ? other : d;
    }
    if ((threadIdx.x & 63u) == 0u) atomicMax(height, d);
}

// ---------------------------------------------------------------------------
// host-side launcher
// ---------------------------------------------------------------------------
hipError_t treeHeightLaunch(const int32_t* parent, uint32_t n, uint32_t internal, uint32_t* height, hipStream_t st) {
    hipLaunchKernelGGL(k_tree_height, dim3((n + 255u) / 256u), dim3(256), 0, st, parent, n, internal, height);
    return hipGetLastError();
}

size_t sortScratchWords(uint32_t n) {
    const size_t tiles = (n + kTileKeys - 1) / kTileKeys;
    return size_t(kMaxRadixPasses) * 256u + kMaxRadixPasses + size_t(kMaxRadixPasses) * tiles * 256u + 1u;
}

hipError_t lbvhBuild(const LbvhBuffers& b, const F4* sitesIn, uint32_t n, const Segments& seg, hipStream_t st) {
    const uint32_t tiles = (n + kTileKeys - 1) / kTileKeys;
    const int passes = seg.molecules > 1u ? radixPassesFor(seg.molecules) : kRadixPasses;
    // sort counters and look-back, then the spanning-node counter
    const uint32_t zeroWords = uint32_t(passes) * 256u + uint32_t(passes) + uint32_t(passes) * tiles * 256u + 1u;
    if (seg.molecules > 1u) {
        hipLaunchKernelGGL(k_scene_segments, dim3(seg.molecules), dim3(256), 0, st, sitesIn, seg.offsets, b.scene,
                           b.sortScratch, zeroWords);
    } else {
        hipLaunchKernelGGL(k_scene_box, dim3(kBoxBlocks), dim3(256), 0, st, sitesIn, n, b.scene, b.boxPartials,
                           b.boxTicket, b.sortScratch, zeroWords);
        hipLaunchKernelGGL(k_scene_final, dim3(1), dim3(kBoxBlocks), 0, st, b.boxPartials, b.scene);
    }
    uint32_t* hist = b.sortScratch;
    uint32_t* tileCounters = hist + passes * 256u;
    uint32_t* lookback = tileCounters + passes;
    const uint32_t mortonBlocks = (n + 255u) / 256u < 2048u ? (n + 255u) / 256u : 2048u;
    hipLaunchKernelGGL(k_morton, dim3(mortonBlocks), dim3(256), 0, st, sitesIn, n, seg, passes, b.scene, b.codesA,
                       b.idxA, hist);
    uint64_t* kin = b.codesA;
    uint32_t* vin = b.idxA;
    uint64_t* kout = b.codesB;
    uint32_t* vout = b.idxB;
    for (int p = 0; p < passes; p++) {
        if (p < kRadixPasses)
            hipLaunchKernelGGL(k_onesweep<false>, dim3(tiles), dim3(256), 0, st, kin, vin, kout, vout, n, p, seg,
                               hist + p * 256, lookback + size_t(p) * tiles * 256u, tileCounters + p, tiles, b.err);
        else
            hipLaunchKernelGGL(k_onesweep<true>, dim3(tiles), dim3(256), 0, st, kin, vin, kout, vout, n, p, seg,
                               hist + p * 256, lookback + size_t(p) * tiles * 256u, tileCounters + p, tiles, b.err);
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    if (passes & 1) {
        // an odd number of passes leaves the sorted data in B: move it to A
        hipLaunchKernelGGL(k_copy_sorted, dim3((n + 255u) / 256u), dim3(256), 0, st, kin, vin, b.codesA, b.idxA, n);
    }
    const uint32_t blocks = (n + 255u) / 256u;
    const uint32_t internal = n - seg.molecules;
    uint32_t* crossCount = lookback + size_t(passes) * tiles * 256u;
    hipLaunchKernelGGL(k_karras, dim3(blocks), dim3(256), 0, st, b.codesA, b.idxA, sitesIn, n, seg, b.sorted, b.nodes,
                       b.parent, b.refitCounter, b.sizes, b.range, b.crossList, crossCount, b.err);
    const uint32_t refitBlocks = (n + kRefitLeaves - 1u) / kRefitLeaves;
    hipLaunchKernelGGL(k_refit_local, dim3(refitBlocks), dim3(512), 0, st, b.sorted, n, internal, b.nodes, b.parent,
                       b.range, b.rootBox, b.err);
    hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, st, n, internal, b.nodes, b.parent, b.range, b.crossList,
                       crossCount, b.refitCounter, b.rootBox, b.err);
    return hipGetLastError();
}

}  // namespace apo_hip
