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

constexpr uint32_t kStreamUnroll = 8;  // site loads in flight per thread in the streaming kernels

__global__ __launch_bounds__(256) void k_scene_box(const F4* __restrict__ sites, uint32_t n, Scene* scene,
                                                   double* partials, uint32_t* ticket, uint32_t* zero,
                                                   uint32_t zeroWords) {
    // 8 reduced values: min x-r, y-r, z-r, r ; max x+r, y+r, z+r, r
    double mn[4] = {DBL_MAX, DBL_MAX, DBL_MAX, DBL_MAX};
    double mx[4] = {-DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    const uint32_t stride = gridDim.x * blockDim.x;
    // kStreamUnroll loads in flight per thread before their use
    for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += kStreamUnroll * stride) {
        F4 sv[kStreamUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kStreamUnroll; u++)
            if (i0 + u * stride < n) sv[u] = sites[i0 + u * stride];
#pragma unroll
        for (uint32_t u = 0; u < kStreamUnroll; u++)
            if (i0 + u * stride < n) accumulateExtent(sv[u], mn, mx);
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
                                                uint32_t* __restrict__ indices, uint32_t* __restrict__ hist,
                                                RefitStarts rs) {
    __shared__ uint32_t sh[kMaxRadixPasses][256];
    // the refit's start lists begin empty (the hierarchy kernel fills them)
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < refitStartWords(n); w += gridDim.x * blockDim.x)
        rs.list[w] = ~0u;
    for (uint32_t t = threadIdx.x; t < kMaxRadixPasses * 256; t += blockDim.x) (&sh[0][0])[t] = 0u;
    __syncthreads();
    float fmn[3], scale[3];
    for (int a = 0; a < 3; a++) {
        fmn[a] = scene[0].fmn[a];
        scale[a] = scene[0].scale[a];
    }
    const bool batch = seg.molecules > 1u;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += kStreamUnroll * stride) {
        F4 sv[kStreamUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kStreamUnroll; u++)
            if (i0 + u * stride < n) sv[u] = sites[i0 + u * stride];
#pragma unroll
        for (uint32_t u = 0; u < kStreamUnroll; u++) {
            const uint32_t i = i0 + u * stride;
            if (i >= n) break;
            uint32_t m = 0;
            if (batch) {
                m = moleculeOf(seg.offsets, seg.molecules, i);
                for (int a = 0; a < 3; a++) {
                    fmn[a] = scene[m].fmn[a];
                    scale[a] = scene[m].scale[a];
                }
            }
            const uint64_t code = mortonCode(sv[u], fmn, scale) | (uint64_t(m & 15u) << 60);
            codes[i] = code;
            indices[i] = i;
            for (int p = 0; p < kRadixPasses; p++) atomicAdd(&sh[p][(code >> (8 * p)) & 255u], 1u);
            for (int p = kRadixPasses; p < passes; p++) atomicAdd(&sh[p][moleculeDigit(m, p)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < uint32_t(passes) * 256u; t += blockDim.x) {
        const uint32_t v = (&sh[0][0])[t];
        if (v) atomicAdd(&hist[t], v);
    }
}

// ---------------------------------------------------------------------------
// Onesweep LSD radix sort pass: 8-bit digit, 4096 keys per tile (16 waves x
// 4 rounds x 64 lanes).  Ranking inside a wave uses 8 ballots to find the
// lanes holding the same digit, so equal digits keep input order (stable).
// Tiles take an index from an atomic counter and chain their per-digit
// prefixes with decoupled look-back.
// ---------------------------------------------------------------------------
constexpr uint32_t kTileKeys = 4096;
constexpr uint32_t kSortWaves = 16;
constexpr uint32_t kSortThreads = 64u * kSortWaves;
constexpr uint32_t kSortRounds = kTileKeys / kSortThreads;
constexpr uint32_t kFlagAgg = 1u << 30, kFlagInc = 2u << 30, kValueMask = (1u << 30) - 1u;
constexpr uint32_t kLookBack = 8;  // predecessor tiles read per look-back step

// Digit of a key / value pair in pass p: a key byte, or (kMol) a digit of
// the molecule of the value (an original site index).
template <bool kMol>
__device__ __forceinline__ uint32_t sortDigit(uint64_t key, uint32_t val, int p, const Segments& seg) {
    if (kMol) return moleculeDigit(moleculeOf(seg.offsets, seg.molecules, val), p);
    return uint32_t(key >> (8 * p)) & 255u;
}

template <bool kMol>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(const uint64_t* __restrict__ keysIn,
                                                           const uint32_t* __restrict__ valsIn,
                                                           uint64_t* __restrict__ keysOut,
                                                           uint32_t* __restrict__ valsOut, uint32_t n, int pass,
                                                           Segments seg, const uint32_t* __restrict__ digitHist,
                                                           uint32_t* lookback, uint32_t* tileCounter,
                                                           uint32_t numTiles, uint32_t* err) {
    __shared__ uint64_t sKeys[kTileKeys];
    __shared__ uint32_t sVals[kTileKeys];
    __shared__ uint32_t sWave[kSortWaves][256];  // per-wave digit counts, then per-wave digit offsets
    __shared__ uint32_t sTileStart[256];
    __shared__ uint32_t sGlobal[256];
    __shared__ uint32_t sScan[4];
    __shared__ uint32_t sScan2[4];
    __shared__ uint32_t sTile;

    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = laneId();
    if (tid == 0) sTile = atomicAdd(tileCounter, 1u);
    for (uint32_t t = tid; t < kSortWaves * 256u; t += kSortThreads) (&sWave[0][0])[t] = 0u;
    __syncthreads();
    const uint32_t tile = sTile;
    if (tile >= numTiles) {
        // more blocks than tiles cannot happen with grid == numTiles; report it
        if (tid == 0) atomicOr(err, kErrSortTile);
        return;
    }
    const uint32_t base = tile * kTileKeys + wave * (kSortRounds * 64u);

    uint64_t key[kSortRounds];
    uint32_t val[kSortRounds];
    uint32_t rank[kSortRounds];
#pragma unroll
    for (uint32_t r = 0; r < kSortRounds; r++) {
        const uint32_t idx = base + r * 64u + lane;
        const bool valid = idx < n;
        key[r] = valid ? keysIn[idx] : 0ull;
        val[r] = valid ? valsIn[idx] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < kSortRounds; r++) {
        const bool valid = base + r * 64u + lane < n;
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

    // digit threads: tile totals, per-wave exclusive offsets, look-back, scans
    uint32_t tileCount = 0, excl = 0, g = 0, gInc = 0, tInc = 0;
    const uint32_t d = tid;
    if (tid < 256u) {
        for (uint32_t w = 0; w < kSortWaves; w++) {
            const uint32_t c = sWave[w][d];
            sWave[w][d] = tileCount;
            tileCount += c;
        }
        // publish the tile aggregate, then look back for the exclusive prefix
        uint32_t* lb = lookback + size_t(tile) * 256u + d;
        if (tile == 0) {
            __hip_atomic_store(lb, kFlagInc | tileCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(lb, kFlagAgg | tileCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // kLookBack predecessors are read at once; their values are taken
            // in order up to the first inclusive one, or up to the first not
            // yet published, where the next read starts
            int t = int(tile) - 1;
            uint32_t spins = 0;
            while (t >= 0) {
                uint32_t v[kLookBack];
#pragma unroll
                for (int j = 0; j < int(kLookBack); j++)
                    v[j] = t - j >= 0 ? __hip_atomic_load(lookback + size_t(t - j) * 256u + d, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : kFlagInc;
                int used = 0;
                bool done = false;
#pragma unroll
                for (int j = 0; j < int(kLookBack); j++) {
                    const uint32_t flag = v[j] & ~kValueMask;
                    if (flag == 0u) break;
                    excl += v[j] & kValueMask;
                    used = j + 1;
                    if (flag == kFlagInc) {
                        done = true;
                        break;
                    }
                }
                if (done) break;
                t -= used;
                if (used < int(kLookBack)) {
                    if (++spins > (1u << 22)) {
                        atomicOr(err, kErrSortSpin);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __hip_atomic_store(lb, kFlagInc | (excl + tileCount), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // exclusive scans over the 256 digits: global digit starts and tile digit starts
        g = digitHist[d];
        gInc = waveInclusiveSum(g);
        tInc = waveInclusiveSum(tileCount);
        if (lane == 63) {
            sScan[wave] = gInc;
            sScan2[wave] = tInc;
        }
    }
    __syncthreads();
    if (tid < 256u) {
        uint32_t gOff = 0, tOff = 0;
        for (uint32_t w = 0; w < wave; w++) {
            gOff += sScan[w];
            tOff += sScan2[w];
        }
        sGlobal[d] = gOff + gInc - g + excl;
        sTileStart[d] = tOff + tInc - tileCount;
    }
    __syncthreads();

#pragma unroll
    for (uint32_t r = 0; r < kSortRounds; r++) {
        const uint32_t idx = base + r * 64u + lane;
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
    for (uint32_t p = tid; p < tileValid; p += kSortThreads) {
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
// The same pass over key digits with tiles of 8192 keys (8 waves x 16
// rounds x 64 lanes) for large sorts: fewer tiles, so a shorter look-back
// chain and longer runs per digit in the scatter.  The keys are staged in
// LDS by output position and written; each slot then takes its output index
// (low word) and value (high word), and the values are written: one 64 KB
// staging buffer for both, so two blocks fit a CU.  Same ranking, so the
// same stable order.  Below kWideMinKeys keys the 4096-key tiles fill the
// chip better (1M keys: 0.35 against 0.37 ms for the LBVH with 6144-key
// tiles); from 2M keys the wide ones win (4M: 0.81 -> 0.78 ms; 16M: 2.70 ->
// 2.62 ms against 6144-key tiles from 4M keys).
// ---------------------------------------------------------------------------
constexpr uint32_t kWideWaves = 8;
constexpr uint32_t kWideRounds = 16;
constexpr uint32_t kWideThreads = 64u * kWideWaves;
constexpr uint32_t kWideKeys = kWideThreads * kWideRounds;
constexpr uint32_t kWideMinKeys = 1u << 21;

__global__ __launch_bounds__(kWideThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_onesweep_wide(const uint64_t* __restrict__ keysIn,
                                                               const uint32_t* __restrict__ valsIn,
                                                               uint64_t* __restrict__ keysOut,
                                                               uint32_t* __restrict__ valsOut, uint32_t n, int pass,
                                                               const uint32_t* __restrict__ digitHist,
                                                               uint32_t* lookback, uint32_t* tileCounter,
                                                               uint32_t numTiles, uint32_t* err) {
    __shared__ uint64_t sSlot[kWideKeys];
    __shared__ uint32_t sWave[kWideWaves][256];
    __shared__ uint32_t sTileStart[256];
    __shared__ uint32_t sGlobal[256];
    __shared__ uint32_t sScan[4];
    __shared__ uint32_t sScan2[4];
    __shared__ uint32_t sTile;

    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = laneId();
    if (tid == 0) sTile = atomicAdd(tileCounter, 1u);
    for (uint32_t t = tid; t < kWideWaves * 256u; t += kWideThreads) (&sWave[0][0])[t] = 0u;
    __syncthreads();
    const uint32_t tile = sTile;
    if (tile >= numTiles) {
        if (tid == 0) atomicOr(err, kErrSortTile);
        return;
    }
    const uint32_t base = tile * kWideKeys + wave * (kWideRounds * 64u);
    const int shift = 8 * pass;

    uint64_t key[kWideRounds];
    uint32_t val[kWideRounds];
    uint32_t pos[kWideRounds];
#pragma unroll
    for (uint32_t r = 0; r < kWideRounds; r++) {
        const uint32_t idx = base + r * 64u + lane;
        const bool valid = idx < n;
        key[r] = valid ? keysIn[idx] : 0ull;
        val[r] = valid ? valsIn[idx] : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < kWideRounds; r++) {
        const bool valid = base + r * 64u + lane < n;
        const uint32_t d = uint32_t(key[r] >> shift) & 255u;
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
        pos[r] = prev + below;
    }
    __syncthreads();

    uint32_t tileCount = 0, excl = 0, g = 0, gInc = 0, tInc = 0;
    const uint32_t d = tid;
    if (tid < 256u) {
        for (uint32_t w = 0; w < kWideWaves; w++) {
            const uint32_t c = sWave[w][d];
            sWave[w][d] = tileCount;
            tileCount += c;
        }
        uint32_t* lb = lookback + size_t(tile) * 256u + d;
        if (tile == 0) {
            __hip_atomic_store(lb, kFlagInc | tileCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(lb, kFlagAgg | tileCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int t = int(tile) - 1;
            uint32_t spins = 0;
            while (t >= 0) {
                uint32_t v[kLookBack];
#pragma unroll
                for (int j = 0; j < int(kLookBack); j++)
                    v[j] = t - j >= 0 ? __hip_atomic_load(lookback + size_t(t - j) * 256u + d, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : kFlagInc;
                int used = 0;
                bool done = false;
#pragma unroll
                for (int j = 0; j < int(kLookBack); j++) {
                    const uint32_t flag = v[j] & ~kValueMask;
                    if (flag == 0u) break;
                    excl += v[j] & kValueMask;
                    used = j + 1;
                    if (flag == kFlagInc) {
                        done = true;
                        break;
                    }
                }
                if (done) break;
                t -= used;
                if (used < int(kLookBack)) {
                    if (++spins > (1u << 22)) {
                        atomicOr(err, kErrSortSpin);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __hip_atomic_store(lb, kFlagInc | (excl + tileCount), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        g = digitHist[d];
        gInc = waveInclusiveSum(g);
        tInc = waveInclusiveSum(tileCount);
        if (lane == 63) {
            sScan[wave] = gInc;
            sScan2[wave] = tInc;
        }
    }
    __syncthreads();
    if (tid < 256u) {
        uint32_t gOff = 0, tOff = 0;
        for (uint32_t w = 0; w < wave; w++) {
            gOff += sScan[w];
            tOff += sScan2[w];
        }
        sGlobal[d] = gOff + gInc - g + excl;
        sTileStart[d] = tOff + tInc - tileCount;
    }
    __syncthreads();

    // keys by output position
#pragma unroll
    for (uint32_t r = 0; r < kWideRounds; r++) {
        const uint32_t idx = base + r * 64u + lane;
        uint32_t p = kNone;
        if (idx < n) {
            const uint32_t dg = uint32_t(key[r] >> shift) & 255u;
            p = sTileStart[dg] + sWave[wave][dg] + pos[r];
            if (p < kWideKeys) {
                sSlot[p] = key[r];
            } else {
                atomicOr(err, kErrSortScatter);
                p = kNone;
            }
        }
        pos[r] = p;
    }
    __syncthreads();
    const uint32_t tileBase = tile * kWideKeys;
    const uint32_t tileValid = (n - tileBase) < kWideKeys ? (n - tileBase) : kWideKeys;
    uint32_t* slot32 = reinterpret_cast<uint32_t*>(sSlot);
    for (uint32_t p = tid; p < tileValid; p += kWideThreads) {
        const uint64_t k = sSlot[p];
        const uint32_t dg = uint32_t(k >> shift) & 255u;
        const uint32_t out = sGlobal[dg] + (p - sTileStart[dg]);
        if (out < n) keysOut[out] = k;
        else atomicOr(err, kErrSortScatter);
        slot32[2u * p] = out;  // the slot's key is read: its low word takes the output index
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kWideRounds; r++)
        if (pos[r] != kNone) slot32[2u * pos[r] + 1u] = val[r];
    __syncthreads();
    for (uint32_t p = tid; p < tileValid; p += kWideThreads) {
        const uint32_t out = slot32[2u * p];
        if (out < n) valsOut[out] = slot32[2u * p + 1u];
    }
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

// The refit runs in phases over nested aligned leaf blocks: phase 0 over
// blocks of 2^kRefitShift leaves, phase p >= 1 over blocks of
// 2^(kRefitShift + kRefitFanShift * p).  A node belongs to the first phase
// whose blocks contain its whole leaf range, so each phase's climbs stay
// inside one block and need only workgroup-scope ordering; later phases see
// earlier slots through the kernel boundary.
__device__ __forceinline__ uint32_t refitShift(uint32_t phase) { return kRefitShift + kRefitFanShift * phase; }

// Level of a leaf range: the last phase whose blocks it crosses (-1: none).
__device__ __forceinline__ int rangeLevel(uint32_t first, uint32_t last) {
    const uint32_t x = first ^ last;
    if ((x >> kRefitShift) == 0u) return -1;
    return int((31u - uint32_t(__builtin_clz(x)) - kRefitShift) / kRefitFanShift);
}

// A node whose children both finish in earlier phases starts a climb in
// its own phase.  Its split sits on a block boundary of its level, which
// indexes its slot in that level's start list; the writer of its left slot
// (the left child's finisher) records it when the right child is also
// finished before the node's phase.
__device__ __forceinline__ void recordStart(const RefitStarts& rs, uint32_t p, uint32_t pf, uint32_t pl,
                                            uint32_t rightFirst, uint32_t* err) {
    const int level = rangeLevel(pf, pl);
    if (level < 0 || level >= int(kRefitLevels)) {
        atomicOr(err, kErrRefitParent);
        return;
    }
    if (((rightFirst ^ pl) >> refitShift(uint32_t(level))) != 0u) return;  // right child finishes in the same phase
    const uint32_t slot = rightFirst >> refitShift(uint32_t(level));
    if (slot >= refitSlots(rs.n, uint32_t(level))) {
        atomicOr(err, kErrRefitParent);
        return;
    }
    rs.list[refitSlotOffset(rs.n, uint32_t(level)) + slot] = p;
}

// Phase p >= 1: one wave per block of 2^refitShift(p) leaves.  Lane t
// starts at the node recorded for the block's t-th boundary of level p-1
// (if any) and climbs while the parent stays inside the block.  A parent
// with a second child of level p-1 is continued by the later of the two
// (arrival counter in refitCounter, zeroed by k_hierarchy); a parent whose
// other child finished in an earlier phase is continued at once.
__global__ __launch_bounds__(64) void k_refit_phase(uint32_t n, uint32_t internal, uint32_t phase, Node* nodes,
                                                    const int32_t* __restrict__ parent,
                                                    const uint32_t* __restrict__ range, RefitStarts rs,
                                                    uint32_t* refitCounter, float* rootBox, uint32_t* err) {
    const uint32_t level = phase - 1u;
    const uint32_t slot = blockIdx.x * 64u + threadIdx.x;
    if (slot >= refitSlots(n, level)) return;
    uint32_t node = rs.list[refitSlotOffset(n, level) + slot];
    if (node == ~0u) return;
    if (node >= internal) {
        atomicOr(err, kErrRefitParent);
        return;
    }
    const bool single = internal + 1u == n;
    const uint32_t shift = refitShift(phase), inner = refitShift(level);
    uint32_t first = range[2 * node], last = range[2 * node + 1];
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
        const uint32_t pf = range[2 * p], pl = range[2 * p + 1];
        const bool left = pf == first;
        if (left ? pl <= last : (pl != last || pf >= first)) {
            atomicOr(err, kErrRefitParent);
            break;
        }
        writeSlot(nodes + p, left, mn, mx, maxr);
        if (((pf ^ pl) >> shift) != 0u) {  // finished by a later phase
            if (left) recordStart(rs, uint32_t(p), pf, pl, last + 1u, err);
            break;
        }
        const uint32_t sf = left ? last + 1u : pf, sl = left ? pl : first - 1u;
        if (((sf ^ sl) >> inner) != 0u) {  // the sibling climbs in this phase too
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            const uint32_t old = atomicAdd(&refitCounter[p], 1u);
            if (old == 0u) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        node = uint32_t(p);
        first = pf;
        last = pl;
    }
}

// ---------------------------------------------------------------------------
// Hierarchy and the first refit phase in one kernel: block b takes leaves
// [512 b, 512 b + 512) and the internal nodes whose index lies in the same
// window (a node's index is one end of its leaf range, minus the molecule
// index).  Each thread first runs Karras 2012 for its node (lbvh.cu:144-223)
// and gathers its leaf's site (setLeaves, lbvh.cu:88-114), then climbs from
// its leaf as in the refit (lbvh.cu:225-272) through the nodes whose leaf
// ranges lie inside the block.  Those nodes' parents, ranges and children
// are known in the block (LDS), and the second arrival at a node writes its
// whole 64-byte record (child boxes, child ids, max radii) once.  A node
// whose parent was computed by another block is appended to `pending`
// (k_refit_pending writes its slot after the kernel boundary); a node whose
// parent was computed here but whose range leaves the block gets its slot
// written now.  Child ids of the nodes whose ranges leave the block are
// written by the Karras step.
// ---------------------------------------------------------------------------
constexpr uint32_t kHierHalo = 256;  // codes staged in LDS on each side of the block's own
constexpr int32_t kParentUnknown = -2;

__global__ __launch_bounds__(kRefitLeaves) void k_hierarchy(const uint64_t* __restrict__ codes,
                                                            const uint32_t* __restrict__ indices,
                                                            const F4* __restrict__ sitesIn, uint32_t n, Segments seg,
                                                            F4* __restrict__ sorted, Node* __restrict__ nodes,
                                                            int32_t* __restrict__ parent,
                                                            uint32_t* __restrict__ refitCounter,
                                                            uint32_t* __restrict__ range,
                                                            RefitStarts rs, uint32_t* __restrict__ pending,
                                                            uint32_t* pendingCount, float* rootBox, uint32_t* err) {
    // the staged codes and the climb's exchange slots share one buffer
    __shared__ uint64_t pool[7 * kRefitLeaves];
    __shared__ int32_t wParent[kRefitLeaves];      // parent of window node w, kParentUnknown if not computed here
    __shared__ int32_t wLeafParent[kRefitLeaves];  // parent of the block's leaf t
    __shared__ uint2 wRange[kRefitLeaves];
    __shared__ uint32_t arrivals[kRefitLeaves];
    __shared__ uint32_t pendLocal, pendBase;
    uint64_t* wc = pool;
    float(*xch)[7][kRefitLeaves] = reinterpret_cast<float(*)[7][kRefitLeaves]>(pool);  // [side][min, max, r][split - lo]

    const uint32_t tid = threadIdx.x;
    const uint32_t lo = blockIdx.x * kRefitLeaves;
    const uint32_t hi = (n - lo) < kRefitLeaves ? n : lo + kRefitLeaves;
    const uint32_t internal = n - seg.molecules;
    uint32_t wb = lo, we = hi;
    if (seg.molecules > 1u) {
        wb = lo - moleculeOf(seg.offsets, seg.molecules, lo);
        we = hi - moleculeOf(seg.offsets, seg.molecules, hi - 1u);
    }
    we = we < internal ? we : internal;
    const uint32_t wn = we > wb ? we - wb : 0u;
    const uint32_t wlo = lo > kHierHalo ? lo - kHierHalo : 0u;
    const uint32_t whi = n - hi > kHierHalo ? hi + kHierHalo : n;
    for (uint32_t t = tid; t < whi - wlo; t += kRefitLeaves) wc[t] = codes[wlo + t];
    wParent[tid] = kParentUnknown;
    wLeafParent[tid] = kParentUnknown;
    arrivals[tid] = 0u;
    if (tid == 0) pendLocal = 0u;
    __syncthreads();

    const uint32_t gi = lo + tid;
    const bool live = gi < hi;
    F4 site{0.f, 0.f, 0.f, 0.f};
    uint32_t m = 0, base = 0, count = n;
    if (live) {
        const uint32_t src = indices[gi];
        if (src < n) site = sitesIn[src];
        else atomicOr(err, kErrGatherIndex);
        sorted[gi] = site;
        if (seg.molecules > 1u) {
            m = moleculeOf(seg.offsets, seg.molecules, gi);
            base = seg.offsets[m];
            count = seg.offsets[m + 1] - base;
        }
    }
    const int i = int(gi - base);
    const int ni = int(count);
    const uint32_t nodeBase = base - m;  // the molecule's first internal node (its root)
    if (live && i < ni - 1) {
        const uint64_t ci = wc[gi - wlo];
        // delta(i, j): common prefix of the codes, or 64 + that of the indices
        // when the codes are equal; -1 outside the molecule
        auto delta = [&](int j) -> int {
            if (j < 0 || j >= ni) return -1;
            const uint32_t gp = base + uint32_t(j);
            const uint64_t cj = gp - wlo < whi - wlo ? wc[gp - wlo] : codes[gp];
            if (ci != cj) return __builtin_clzll(ci ^ cj);
            return __builtin_clzll(uint64_t(int64_t(i ^ j))) + 64;
        };
        const uint32_t g = nodeBase + uint32_t(i);
        const int dd = delta(i + 1) - delta(i - 1);
        const int d = (dd > 0) - (dd < 0);
        const int deltaMin = delta(i - d);
        int lmax = 2;
        while (delta(i + lmax * d) > deltaMin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(i + (l + t) * d) > deltaMin) l += t;
        const int j = i + l * d;
        const int deltaNode = delta(j);
        int s = 0, k = 2, t;
        do {
            t = (l + k - 1) / k;
            k <<= 1;
            if (delta(i + (s + t) * d) > deltaNode) s += t;
        } while (t > 1);
        const int gamma = i + s * d + (d < 0 ? d : 0);
        // a child whose range is one leaf is that leaf, else the internal node
        // with the same local index
        const bool leftLeaf = (i < j ? i : j) == gamma, rightLeaf = (i > j ? i : j) == gamma + 1;
        const uint32_t lc = leftLeaf ? internal + base + uint32_t(gamma) : nodeBase + uint32_t(gamma);
        const uint32_t rc = rightLeaf ? internal + base + uint32_t(gamma) + 1u : nodeBase + uint32_t(gamma) + 1u;
        parent[lc] = int32_t(g);
        parent[rc] = int32_t(g);
        if (i == 0) {
            parent[g] = -1;
            wParent[g - wb] = -1;
        }
        const uint32_t first = base + uint32_t(i < j ? i : j), last = base + uint32_t(i > j ? i : j);
        reinterpret_cast<uint2*>(range)[g] = make_uint2(first, last);
        wRange[g - wb] = make_uint2(first, last);
        // the children's parent, for the climbs that reach them in this block
        const uint32_t ls = base + uint32_t(gamma);  // global index of the split leaf
        if (leftLeaf) {
            if (ls - lo < kRefitLeaves) wLeafParent[ls - lo] = int32_t(g);
        } else if (lc - wb < wn) {
            wParent[lc - wb] = int32_t(g);
        }
        if (rightLeaf) {
            if (ls + 1u - lo < kRefitLeaves) wLeafParent[ls + 1u - lo] = int32_t(g);
        } else if (rc - wb < wn) {
            wParent[rc - wb] = int32_t(g);
        }
        // a node finished in a later phase gets its child ids and a cleared
        // arrival counter now; the others their whole record at the end of
        // their climb
        if (first < lo || last >= hi) {
            reinterpret_cast<int2*>(&nodes[g].d)[0] = make_int2(int32_t(lc), int32_t(rc));
            refitCounter[g] = 0u;
        }
    }
    __syncthreads();

    uint32_t pend = ~0u, pendAt = 0u;
    if (live) {
        const bool single = internal + 1u == n;
        float mn[3], mx[3];
        leafBox(site, mn, mx);
        float maxr = site.w;
        uint32_t node = internal + gi, first = gi, last = gi;
        int p = wLeafParent[tid];
        for (uint32_t climb = 0; climb < n; climb++) {
            if (p == kParentUnknown) {  // parent computed by another block: its slot is written later
                pend = node;
                pendAt = atomicAdd(&pendLocal, 1u);
                break;
            }
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
            const uint32_t w = uint32_t(p) - wb;
            if (uint32_t(p) >= internal || w >= wn) {
                atomicOr(err, kErrRefitParent);
                break;
            }
            const uint2 r = wRange[w];
            const uint32_t pf = r.x, pl = r.y;
            const bool left = pf == first;
            if (left ? pl <= last : (pl != last || pf >= first)) {
                atomicOr(err, kErrRefitParent);
                break;
            }
            if (pf < lo || pl >= hi) {  // finished by a later phase
                writeSlot(nodes + p, left, mn, mx, maxr);
                if (left) recordStart(rs, uint32_t(p), pf, pl, last + 1u, err);
                break;
            }
            const uint32_t sg = left ? last : first - 1u;  // split leaf (global)
            const uint32_t split = sg - lo;
            const uint32_t side = left ? 0u : 1u;
            for (int c = 0; c < 3; c++) {
                xch[side][c][split] = mn[c];
                xch[side][3 + c][split] = mx[c];
            }
            xch[side][6][split] = maxr;
            // (LDS-only ordering: the node records written here are read by
            // later kernels, so no wait for global stores)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            const uint32_t old = atomicAdd(&arrivals[split], 1u);
            if (old == 0u) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            float lmn[3], lmx[3], rmn[3], rmx[3];
            for (int c = 0; c < 3; c++) {
                lmn[c] = xch[0][c][split];
                lmx[c] = xch[0][3 + c][split];
                rmn[c] = xch[1][c][split];
                rmx[c] = xch[1][3 + c][split];
            }
            const float lr = xch[0][6][split], rr = xch[1][6][split];
            // the Karras children of a node split after leaf sg
            const uint32_t lc = pf == sg ? internal + sg : sg - m;
            const uint32_t rc = pl == sg + 1u ? internal + sg + 1u : sg + 1u - m;
            Node* P = nodes + p;
            P->a = make_float4(lmn[0], lmn[1], lmn[2], lmx[0]);
            P->b = make_float4(lmx[1], lmx[2], rmn[0], rmn[1]);
            P->c = make_float4(rmn[2], rmx[0], rmx[1], rmx[2]);
            P->d = make_int4(int32_t(lc), int32_t(rc), __float_as_int(lr), __float_as_int(rr));
            for (int c = 0; c < 3; c++) {
                mn[c] = __builtin_fminf(lmn[c], rmn[c]);
                mx[c] = __builtin_fmaxf(lmx[c], rmx[c]);
            }
            maxr = __builtin_fmaxf(lr, rr);
            node = uint32_t(p);
            first = pf;
            last = pl;
            p = wParent[w];
        }
    }
    __syncthreads();
    if (tid == 0 && pendLocal) pendBase = atomicAdd(pendingCount, pendLocal);
    __syncthreads();
    if (pend != ~0u) pending[pendBase + pendAt] = pend;
}

// The nodes k_hierarchy appended to `pending` (their parent was computed by
// another block, so its range leaves theirs): each writes its box into its
// side of the parent's record, and the left child records the parent as a
// climb start of the parent's phase when due.
__global__ __launch_bounds__(256) void k_refit_pending(const uint32_t* __restrict__ pending,
                                                       const uint32_t* __restrict__ pendingCount, uint32_t internal,
                                                       const F4* __restrict__ sorted, Node* nodes,
                                                       const int32_t* __restrict__ parent,
                                                       const uint32_t* __restrict__ range, RefitStarts rs,
                                                       uint32_t* err) {
    const uint32_t count = *pendingCount;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const uint32_t c = pending[k];
        float mn[3], mx[3], maxr;
        uint32_t first, last;
        if (c >= internal) {
            first = last = c - internal;
            const F4 s = sorted[first];
            leafBox(s, mn, mx);
            maxr = s.w;
        } else {
            nodeUnion(nodes + c, mn, mx, maxr);
            first = range[2 * c];
            last = range[2 * c + 1];
        }
        const int p = parent[c];
        if (p < 0 || uint32_t(p) >= internal) {
            atomicOr(err, kErrRefitParent);
            continue;
        }
        const uint32_t pf = range[2 * p], pl = range[2 * p + 1];
        const bool left = pf == first;
        if (left ? pl <= last : (pl != last || pf >= first)) {
            atomicOr(err, kErrRefitParent);
            continue;
        }
        writeSlot(nodes + p, left, mn, mx, maxr);
        if (left) recordStart(rs, uint32_t(p), pf, pl, last + 1u, err);
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
// this depth or more can overflow.  A leaf's depth is the number of
// internal nodes whose leaf range holds it: each node adds +1 at its first
// leaf and -1 after its last, and the largest prefix sum is the height.
// Block b takes the nodes and owns the mark positions [4096 b, 4096 b +
// 4096): marks that land there are summed in LDS and stored once; the rest
// (large nodes and block edges) go to `far` with global atomics.
constexpr uint32_t kDepthMarks = 4096;  // nodes and positions per block (1024 threads x 4)

__global__ __launch_bounds__(1024) void k_depth_marks(const uint32_t* __restrict__ range, uint32_t n,
                                                      uint32_t internal, int32_t* __restrict__ diff, int32_t* far) {
    __shared__ int32_t own[kDepthMarks];
    const uint32_t lo = blockIdx.x * kDepthMarks;
    for (uint32_t t = threadIdx.x; t < kDepthMarks; t += blockDim.x) own[t] = 0;
    __syncthreads();
    auto mark = [&](uint32_t x, int32_t v) {
        if (x - lo < kDepthMarks) atomicAdd(&own[x - lo], v);
        else atomicAdd(&far[x], v);
    };
    for (uint32_t q = 0; q < kDepthMarks / 1024u; q++) {
        const uint32_t g = lo + q * 1024u + threadIdx.x;
        if (g >= internal) break;
        const uint2 r = reinterpret_cast<const uint2*>(range)[g];
        mark(r.x, 1);
        if (r.y + 1u < n) mark(r.y + 1u, -1);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < kDepthMarks && lo + t < n; t += blockDim.x) diff[lo + t] = own[t];
}

constexpr uint32_t kDepthBlock = 4096;  // prefix-sum elements per block (1024 threads x 4)

// A run of marks as (sum, largest prefix sum); runs join in order.
__device__ __forceinline__ int2 joinRuns(int2 a, int2 b) {
    const int32_t m = a.x + b.y;
    return make_int2(a.x + b.x, a.y > m ? a.y : m);
}

// Ordered join over the block: each thread holds consecutive runs.
__device__ __forceinline__ int2 blockJoin(int2 r, int2* waveRuns) {
    for (int o = 1; o < 64; o <<= 1) {
        const int2 other = make_int2(__shfl_down(r.x, o, 64), __shfl_down(r.y, o, 64));
        if (int(laneId()) + o < 64) r = joinRuns(r, other);
    }
    const uint32_t waves = blockDim.x >> 6, wave = threadIdx.x >> 6;
    if (laneId() == 0) waveRuns[wave] = r;
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t w = 1; w < waves; w++) r = joinRuns(r, waveRuns[w]);
    return r;
}

// Per block of kDepthBlock marks: its sum and its largest prefix sum.
__global__ __launch_bounds__(1024) void k_depth_scan(const int32_t* __restrict__ diff, const int32_t* __restrict__ far,
                                                     uint32_t n, int2* blocks) {
    __shared__ int2 waveRuns[16];
    const uint32_t base = blockIdx.x * kDepthBlock + threadIdx.x * 4u;
    int2 r = make_int2(0, 0);
    for (uint32_t q = 0; q < 4u; q++) {
        r.x += base + q < n ? diff[base + q] + far[base + q] : 0;
        r.y = r.y > r.x ? r.y : r.x;
    }
    r = blockJoin(r, waveRuns);
    if (threadIdx.x == 0) blocks[blockIdx.x] = r;
}

// One block joins the blocks' runs in order: the largest prefix overall.
__global__ __launch_bounds__(1024) void k_depth_final(const int2* __restrict__ blocks, uint32_t count,
                                                      uint32_t* height) {
    __shared__ int2 waveRuns[16];
    const uint32_t per = (count + blockDim.x - 1u) / blockDim.x;
    const uint32_t b0 = threadIdx.x * per;
    int2 r = make_int2(0, 0);
    for (uint32_t b = b0; b < b0 + per && b < count; b++) r = joinRuns(r, blocks[b]);
    r = blockJoin(r, waveRuns);
    if (threadIdx.x == 0) *height = uint32_t(r.y > 0 ? r.y : 0);
}

// ---------------------------------------------------------------------------
// host-side launcher
// ---------------------------------------------------------------------------
hipError_t treeHeightLaunch(const uint32_t* range, uint32_t n, uint32_t internal, int32_t* diff, int32_t* far,
                            int2* blocks, uint32_t* height, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_words, dim3((n + 255u) / 256u), dim3(256), 0, st, reinterpret_cast<uint32_t*>(far), n,
                       0u);
    hipLaunchKernelGGL(k_depth_marks, dim3((n + kDepthMarks - 1u) / kDepthMarks), dim3(1024), 0, st, range, n,
                       internal, diff, far);
    const uint32_t nb = (n + kDepthBlock - 1u) / kDepthBlock;
    hipLaunchKernelGGL(k_depth_scan, dim3(nb), dim3(1024), 0, st, diff, far, n, blocks);
    hipLaunchKernelGGL(k_depth_final, dim3(1), dim3(1024), 0, st, blocks, nb, height);
    return hipGetLastError();
}

// Sort pass p of n keys runs the wide kernel: key digits of a large sort.
bool widePass(uint32_t n, int p) { return p < kRadixPasses && n >= kWideMinKeys; }
uint32_t passTiles(uint32_t n, int p) {
    return widePass(n, p) ? (n + kWideKeys - 1u) / kWideKeys : (n + kTileKeys - 1u) / kTileKeys;
}

size_t sortScratchWords(uint32_t n) {
    const size_t tiles = (n + kTileKeys - 1) / kTileKeys;  // the most any pass has
    // (+1: the hierarchy's pending count)
    return size_t(kMaxRadixPasses) * 256u + kMaxRadixPasses + size_t(kMaxRadixPasses) * tiles * 256u + 1u;
}

hipError_t lbvhBuild(const LbvhBuffers& b, const F4* sitesIn, uint32_t n, const Segments& seg, hipStream_t st) {
    const int passes = seg.molecules > 1u ? radixPassesFor(seg.molecules) : kRadixPasses;
    // sort counters and look-back (each pass's tiles x 256 words)
    size_t lbWords = 0;
    for (int p = 0; p < passes; p++) lbWords += size_t(passTiles(n, p)) * 256u;
    const uint32_t zeroWords = uint32_t(passes) * 256u + uint32_t(passes) + uint32_t(lbWords) + 1u;
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
    uint32_t* pendingCount = lookback + lbWords;
    // (few blocks: each flushes passes x 256 histogram counts with global atomics)
    const uint32_t mortonBlocks = (n + 255u) / 256u < 512u ? (n + 255u) / 256u : 512u;
    hipLaunchKernelGGL(k_morton, dim3(mortonBlocks), dim3(256), 0, st, sitesIn, n, seg, passes, b.scene, b.codesA,
                       b.idxA, hist, RefitStarts{b.startLists, n});
    uint64_t* kin = b.codesA;
    uint32_t* vin = b.idxA;
    uint64_t* kout = b.codesB;
    uint32_t* vout = b.idxB;
    size_t lbOff = 0;
    for (int p = 0; p < passes; p++) {
        const uint32_t tiles = passTiles(n, p);
        if (widePass(n, p))
            hipLaunchKernelGGL(k_onesweep_wide, dim3(tiles), dim3(kWideThreads), 0, st, kin, vin, kout, vout, n, p,
                               hist + p * 256, lookback + lbOff, tileCounters + p, tiles, b.err);
        else if (p < kRadixPasses)
            hipLaunchKernelGGL(k_onesweep<false>, dim3(tiles), dim3(kSortThreads), 0, st, kin, vin, kout, vout, n, p, seg,
                               hist + p * 256, lookback + lbOff, tileCounters + p, tiles, b.err);
        else
            hipLaunchKernelGGL(k_onesweep<true>, dim3(tiles), dim3(kSortThreads), 0, st, kin, vin, kout, vout, n, p, seg,
                               hist + p * 256, lookback + lbOff, tileCounters + p, tiles, b.err);
        lbOff += size_t(tiles) * 256u;
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    if (passes & 1) {
        // an odd number of passes leaves the sorted data in B: move it to A
        hipLaunchKernelGGL(k_copy_sorted, dim3((n + 255u) / 256u), dim3(256), 0, st, kin, vin, b.codesA, b.idxA, n);
    }
    const uint32_t internal = n - seg.molecules;
    const RefitStarts rs{b.startLists, n};
    // the second key buffer is free once the sort is done: it holds the
    // hierarchy's pending nodes (at most one per node and leaf)
    uint32_t* pending = reinterpret_cast<uint32_t*>(b.codesB);
    const uint32_t refitBlocks = (n + kRefitLeaves - 1u) / kRefitLeaves;
    hipLaunchKernelGGL(k_hierarchy, dim3(refitBlocks), dim3(kRefitLeaves), 0, st, b.codesA, b.idxA, sitesIn, n, seg,
                       b.sorted, b.nodes, b.parent, b.refitCounter, b.range, rs, pending, pendingCount,
                       b.rootBox, b.err);
    if (refitBlocks > 1u) {
        const uint32_t pendingBlocks = refitBlocks < 1024u ? refitBlocks : 1024u;
        hipLaunchKernelGGL(k_refit_pending, dim3(pendingBlocks), dim3(256), 0, st, pending, pendingCount, internal,
                           b.sorted, b.nodes, b.parent, b.range, rs, b.err);
    }
    for (uint32_t phase = 1; phase <= refitPhases(n); phase++)
        hipLaunchKernelGGL(k_refit_phase, dim3((refitSlots(n, phase - 1u) + 63u) / 64u), dim3(64), 0, st, n, internal,
                           phase, b.nodes, b.parent, b.range, rs, b.refitCounter, b.rootBox, b.err);
    return hipGetLastError();
}

}  // namespace apo_hip
