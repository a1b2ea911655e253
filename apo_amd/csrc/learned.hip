// learned.hip -- the learned neighbour-count proposal on gfx950.
//
// A small MLP (52 -> 64 -> 64 -> 8, ReLU) predicts for every site how many
// nearest neighbours its cell needs before the security-radius test
// (algorithm.inl:197-201) certifies it.  The cell pass then runs exactly
// that many updates and judges the cell with the plain test; any cell the
// test does not certify goes through the plain validation kernels, so the
// proposal can only change how much work reaches validation.
//
// Labels are the reference's own (app/benchmarks/
// benchmark_contributing_nearest.cu:75-113): the number of neighbours, in
// kNN order, after which the test first holds, checking after every update.
//
// Inference runs on the matrix cores: BF16 weights and activations, FP32
// accumulation (v_mfma_f32_16x16x16_bf16), one wave per 16 sites.  Training
// runs in FP32 on the vector units with BF16-rounded weights in the forward
// pass (the weights inference will see), Adam on FP32 master weights.
#include "apo_device.h"
#include "apo_internal.h"
#include "cell_update.h"

namespace apo_hip {

// ---------------------------------------------------------------------------
// Cell pass with a per-site neighbour count (mode kProposal) or with the
// test after every update (mode kLabels, getCellsContribution).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_cells_k(CellCtx c, const uint32_t* __restrict__ knn, uint32_t K,
                                                const uint8_t* __restrict__ cls, uint8_t* labels, uint32_t mode,
                                                uint32_t fallbackK) {
    __shared__ CellLds L;
    const uint32_t lane = laneId();
    const uint32_t n = c.n;
    const uint32_t i = blockIdx.x;
    WaveCell w;
    initCell(L, w, artificialBase(c, cellMolecule(c, i)));
    const uint8_t st = c.status[i];
    const bool active = !(st == kStatusCertified || st == kStatusBuried);
    float kr = -FLT_MAX;
    bool certified = false;
    uint32_t label = st == kStatusBuried ? kLabelBuried : 0u;
    if (active) {
        const D4 si = loadSite(c, i);
        uint32_t use = K;
        if (mode == kProposal) {
            const uint32_t q = cls[i];
            use = q < kProposalClasses ? proposalK(q, fallbackK) : K;
            if (use > K) use = K;
        }
        uint32_t myJ = kNone;
        bool myBuried = false;
        if (lane < K) {
            myJ = knn[size_t(i) * K + lane];
            if (myJ != kNone && myJ >= n) {
                atomicOr(c.err, kErrCellNeighbour);
                myJ = kNone;
            }
            if (myJ != kNone) myBuried = c.status[myJ] == kStatusBuried;
        }
        const uint64_t buriedMask = ballot(myBuried);
        const uint64_t noneMask = ballot(myJ == kNone);
        const uint32_t count = noneMask ? uint32_t(__builtin_ctzll(noneMask)) : 64u;
        const uint32_t limit = count < use ? count : use;
        double knownRadius = -DBL_MAX;
        for (uint32_t t = 0; t < limit; t++) {
            if ((buriedMask >> t) & 1u) continue;
            const uint32_t j = readLane(myJ, t);
            const D4 sj = loadSite(c, j);
            updateCell(c, L, w, i, si, j, sj);
            knownRadius = __builtin_fmax(knownRadius, sphereDistD(si, sj));
            if (mode == kLabels) {
                // the test's inputs are wave-uniform; read it from one lane so
                // the loop exit is a scalar branch taken by the whole wave
                const bool cert = gtD(knownRadius, 2.0 * cellRadius(c, L, w, si));
                if (__builtin_amdgcn_readfirstlane(cert ? 1u : 0u) != 0u) {
                    certified = true;
                    label = t + 1u;
                    break;
                }
            }
        }
        if (mode == kProposal) certified = gtD(knownRadius, 2.0 * cellRadius(c, L, w, si));
        kr = float(knownRadius);
    }
    storeCell(c, L, w, i);
    if (lane == 0) {
        if (certified) c.status[i] = kStatusCertified;
        c.knownRadii[i] = kr;
        if (mode == kLabels) labels[i] = uint8_t(label);
    }
}

// ---------------------------------------------------------------------------
// Features: 52 floats per site from the site and its 16 nearest neighbours,
// lengths scaled by d16, the centre distance of the 16th neighbour:
//   [0]            r_i / d16
//   [1 + 3t ...]   for t < 16: centre distance, radius, sphere gap of the
//                  t-th neighbour, each / d16 (zeros for a missing one)
//   [49]           fraction of the 16 with a larger radius than r_i
//   [50]           log2(r_i / mean neighbour radius) / 8, clamped to [-1, 1]
//   [51]           (largest - smallest gap) / d16
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_features(CellCtx c, const uint32_t* __restrict__ knn, uint32_t K,
                                                  float* __restrict__ feat, float* __restrict__ d16Out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const F4 s = c.sorted[i];
    float* f = feat + size_t(i) * kFeatures;
    const uint32_t* nb = knn + size_t(i) * K;
    float d16 = 0.f;
    for (int t = 15; t >= 0; t--) {
        const uint32_t j = nb[t];
        if (j < c.n) {
            const F4 o = c.sorted[j];
            d16 = lenf(subf(xyzf(o), xyzf(s)));
            break;
        }
    }
    if (d16Out) d16Out[i] = d16;
    const float inv = d16 > 0.f ? 1.f / d16 : 0.f;
    f[0] = s.w * inv;
    float larger = 0.f, rsum = 0.f, gmin = FLT_MAX, gmax = -FLT_MAX;
    int have = 0;
    for (int t = 0; t < 16; t++) {
        const uint32_t j = nb[t];
        float dd = 0.f, rr = 0.f, gg = 0.f;
        if (j < c.n) {
            const F4 o = c.sorted[j];
            const float d = lenf(subf(xyzf(o), xyzf(s)));
            const float g = d - s.w - o.w;
            dd = d * inv;
            rr = o.w * inv;
            gg = g * inv;
            larger += o.w > s.w ? 1.f : 0.f;
            rsum += o.w;
            gmin = fminf(gmin, gg);
            gmax = fmaxf(gmax, gg);
            have++;
        }
        f[1 + 3 * t] = dd;
        f[2 + 3 * t] = rr;
        f[3 + 3 * t] = gg;
    }
    f[49] = have ? larger / float(have) : 0.f;
    float lr = 0.f;
    if (have && rsum > 0.f && s.w > 0.f) lr = __builtin_log2f(s.w / (rsum / float(have))) * 0.125f;
    f[50] = fminf(1.f, fmaxf(-1.f, lr));
    f[51] = have ? gmax - gmin : 0.f;
}

// ---------------------------------------------------------------------------
// Inference on the matrix cores.  Weights are transposed BF16 [out][in] with
// in = 64 (layer 1 zero-padded from 52) and out = 64, 64, 16 (layer 3
// zero-padded from 9: 8 class logits, then the radius head).  v_mfma_f32_16x16x16_bf16 lane layout: A[row = l % 16]
// [k = 4 (l / 16) + 0..3], B[k = 4 (l / 16) + 0..3][col = l % 16],
// D[row = 4 (l / 16) + 0..3][col = l % 16].
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t toBf16(float x) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    return uint16_t((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float fromBf16(uint16_t h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

// out[16][N] = act(in[16][64] * W^T + b), in/out bf16 in LDS, N = 16 * blocks
template <uint32_t Blocks, bool Relu>
__device__ __forceinline__ void mlpLayer(const uint16_t (*in)[kHidden], const uint16_t* __restrict__ wt,
                                         const float* __restrict__ bias, uint16_t (*out)[kHidden], float (*logits)[16]) {
    const uint32_t lane = laneId();
    const uint32_t row = lane & 15u, kq = (lane >> 4) * 4u;
#pragma unroll
    for (uint32_t cb = 0; cb < Blocks; cb++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const uint32_t col = cb * 16u + row;
#pragma unroll
        for (uint32_t s = 0; s < kHidden / 16u; s++) {
            bf16x4 a, b;
            const uint32_t k0 = s * 16u + kq;
            for (int e = 0; e < 4; e++) {
                a[e] = short(in[row][k0 + e]);
                b[e] = short(wt[col * kHidden + k0 + e]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        const float bb = bias[col];
        for (int e = 0; e < 4; e++) {
            float v = acc[e] + bb;
            const uint32_t r = kq + e;
            if (Relu) {
                v = fmaxf(v, 0.f);
                out[r][col] = toBf16(v);
            } else {
                logits[r][col] = v;
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_mlp_infer(const float* __restrict__ feat, uint32_t n, const uint16_t* w1,
                                                  const float* b1, const uint16_t* w2, const float* b2,
                                                  const uint16_t* w3, const float* b3, uint8_t* cls,
                                                  float* logitsOut, float cost, uint32_t fallbackK,
                                                  uint32_t maxK, const float* __restrict__ d16,
                                                  float* __restrict__ radii) {
    __shared__ uint16_t X[16][kHidden];
    __shared__ uint16_t H[16][kHidden];
    __shared__ float Lg[16][16];
    const uint32_t lane = laneId();
    const uint32_t tiles = (n + 15u) / 16u;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t base = tile * 16u;
        for (uint32_t idx = lane; idx < 16u * kHidden; idx += 64u) {
            const uint32_t r = idx / kHidden, col = idx % kHidden;
            const uint32_t site = base + r;
            const float v = (col < kFeatures && site < n) ? feat[size_t(site) * kFeatures + col] : 0.f;
            X[r][col] = toBf16(v);
        }
        waveSync();
        mlpLayer<kHidden / 16u, true>(X, w1, b1, H, Lg);
        waveSync();
        mlpLayer<kHidden / 16u, true>(H, w2, b2, X, Lg);
        waveSync();
        mlpLayer<1, false>(X, w3, b3, H, Lg);
        waveSync();
        if (lane < 16u && base + lane < n) {
            uint32_t best = 0;
            float bv = Lg[lane][0];
            for (uint32_t q = 1; q < kProposalClasses; q++)
                if (Lg[lane][q] > bv) {
                    bv = Lg[lane][q];
                    best = q;
                }
            // a class asking for more than maxK neighbours predicts a cell
            // that maxK neighbours do not certify: the last class
            if (proposalK(best) > maxK) best = kProposalClasses - 1;
            if (cost > 0.f) {
                // expected work: k_c updates, plus a validation pass when the
                // cell is not certified at k_c (classes above c); the last
                // class runs fallbackK neighbours and always validates; only
                // classes of at most maxK neighbours are candidates
                float p[kProposalClasses], z = 0.f;
                for (uint32_t q = 0; q < kProposalClasses; q++) {
                    p[q] = __expf(Lg[lane][q] - bv);
                    z += p[q];
                }
                float bestCost = float(proposalK(kProposalClasses - 1, fallbackK)) + cost, cum = 0.f;
                best = kProposalClasses - 1;
                for (uint32_t q = 0; q + 1 < kProposalClasses && proposalK(q) <= maxK; q++) {
                    cum += p[q] / z;
                    const float e = float(proposalK(q)) + cost * (1.f - cum);
                    if (e < bestCost) {
                        bestCost = e;
                        best = q;
                    }
                }
            }
            cls[base + lane] = uint8_t(best);
            if (logitsOut)
                for (uint32_t q = 0; q < kProposalOutputs; q++)
                    logitsOut[size_t(base + lane) * kProposalOutputs + q] = Lg[lane][q];
            // the radius head: log2 of the search radius over d16
            if (radii) radii[base + lane] = d16[base + lane] * exp2f(Lg[lane][kRadiusOutput]);
        }
        waveSync();
    }
}

// ---------------------------------------------------------------------------
// Training on the matrix cores.  Parameters, flat: W1 [52][64], b1 [64],
// W2 [64][64], b2 [64], W3 [64][9], b3 [9] (weights [in][out]).  The loss
// of a row: softmax cross-entropy over the 8 class logits, plus
// kRadiusLossWeight x the squared error of the radius head against the
// row's target (rows without a target, NaN, add none).
//
// The forward pass is the inference pass above (BF16 weights and
// activations, FP32 accumulation on v_mfma_f32_16x16x16_bf16), so the
// network trained is the network that runs.  The backward pass runs in
// FP32 on v_mfma_f32_16x16x4_f32 (lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; D as above) against the BF16-rounded weights the
// forward pass used.  Each wave takes kTrainTiles tiles of 16 rows of the
// batch and writes its partial gradient (and loss) once; the Adam kernel
// sums the partials, so no atomics and the sum order is fixed.
// ---------------------------------------------------------------------------
constexpr uint32_t kTrainTiles = 8;
constexpr uint32_t kOffB1 = kFeatures * kHidden, kOffW2 = kOffB1 + kHidden, kOffB2 = kOffW2 + kHidden * kHidden,
                   kOffW3 = kOffB2 + kHidden, kOffB3 = kOffW3 + kHidden * kProposalOutputs;

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__global__ __launch_bounds__(64) void k_train_grad(const float* __restrict__ feat, const uint8_t* __restrict__ cls,
                                                   const float* __restrict__ radiusTarget,
                                                   const uint32_t* __restrict__ order, uint32_t rows,
                                                   float invBatch, const uint16_t* w1t, const float* b1,
                                                   const uint16_t* w2t, const float* b2, const uint16_t* w3t,
                                                   const float* b3, const float* __restrict__ wq,
                                                   float* __restrict__ partial, float* __restrict__ lossPartial) {
    __shared__ uint16_t X[16][kHidden];
    __shared__ uint16_t H1[16][kHidden];
    __shared__ uint16_t H2[16][kHidden];
    __shared__ float Z[16][16];    // logits, then dLoss/dlogits
    __shared__ float D2[16][kHidden];
    __shared__ float D1[16][kHidden];
    const uint32_t lane = laneId();
    const uint32_t lr16 = lane & 15u, lq = lane >> 4;
    const float* W2 = wq + kOffW2;
    const float* W3 = wq + kOffW3;
    f32x4 g1[4][4], g2[4][4], g3[4];
    for (int rb = 0; rb < 4; rb++) {
        g3[rb] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int cb = 0; cb < 4; cb++) g1[rb][cb] = g2[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
    const uint32_t first = blockIdx.x * kTrainTiles * 16u;
    for (uint32_t t = 0; t < kTrainTiles; t++) {
        const uint32_t base = first + t * 16u;
        if (base >= rows) break;
        for (uint32_t idx = lane; idx < 16u * kHidden; idx += 64u) {
            const uint32_t r = idx / kHidden, col = idx % kHidden;
            const uint32_t row = base + r;
            const float v = (col < kFeatures && row < rows) ? feat[size_t(order[row]) * kFeatures + col] : 0.f;
            X[r][col] = toBf16(v);
        }
        waveSync();
        mlpLayer<kHidden / 16u, true>(X, w1t, b1, H1, Z);
        waveSync();
        mlpLayer<kHidden / 16u, true>(H1, w2t, b2, H2, Z);
        waveSync();
        mlpLayer<1, false>(H2, w3t, b3, H1, Z);
        waveSync();
        // softmax cross-entropy of the batch mean: dL/dz = (p - onehot) / batch
        if (lane < 16u) {
            const uint32_t row = base + lane;
            float z[kProposalClasses];
            for (uint32_t q = 0; q < kProposalClasses; q++) z[q] = Z[lane][q];
            if (row < rows) {
                const uint32_t y = cls[order[row]];
                float m = z[0];
                for (uint32_t q = 1; q < kProposalClasses; q++) m = fmaxf(m, z[q]);
                float e[kProposalClasses], sum = 0.f;
                for (uint32_t q = 0; q < kProposalClasses; q++) {
                    e[q] = expf(z[q] - m);
                    sum += e[q];
                }
                for (uint32_t q = 0; q < kProposalClasses; q++)
                    Z[lane][q] = (e[q] / sum - (q == y ? 1.f : 0.f)) * invBatch;
                loss += (m + logf(sum)) - z[y < kProposalClasses ? y : 0u];
                // the radius head: w (z - t)^2, dL/dz = 2 w (z - t) / batch
                const float t = radiusTarget ? radiusTarget[order[row]] : __builtin_nanf("");
                const float zr = Z[lane][kRadiusOutput];
                const bool has = t == t;
                Z[lane][kRadiusOutput] = has ? 2.f * kRadiusLossWeight * (zr - t) * invBatch : 0.f;
                if (has) loss += kRadiusLossWeight * (zr - t) * (zr - t);
            } else {
                for (uint32_t q = 0; q < kProposalOutputs; q++) Z[lane][q] = 0.f;
            }
            for (uint32_t q = kProposalOutputs; q < 16u; q++) Z[lane][q] = 0.f;
        }
        waveSync();
        // dH2 = dZ W3^T through the ReLU
        for (uint32_t cb = 0; cb < 4u; cb++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            const uint32_t i = cb * 16u + lr16;
            for (uint32_t ks = 0; ks < (kProposalOutputs + 3u) / 4u; ks++) {
                const uint32_t o = ks * 4u + lq;
                acc = mfma4(Z[lr16][o], o < kProposalOutputs ? W3[i * kProposalOutputs + o] : 0.f, acc);
            }
            for (int e = 0; e < 4; e++) {
                const uint32_t r = lq * 4u + e;
                D2[r][i] = fromBf16(H2[r][i]) > 0.f ? acc[e] : 0.f;
            }
        }
        waveSync();
        // dH1 = dH2 W2^T through the ReLU
        for (uint32_t cb = 0; cb < 4u; cb++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            const uint32_t i = cb * 16u + lr16;
            for (uint32_t ks = 0; ks < kHidden / 4u; ks++) {
                const uint32_t j = ks * 4u + lq;
                acc = mfma4(D2[lr16][j], W2[i * kHidden + j], acc);
            }
            for (int e = 0; e < 4; e++) {
                const uint32_t r = lq * 4u + e;
                D1[r][i] = fromBf16(H1[r][i]) > 0.f ? acc[e] : 0.f;
            }
        }
        waveSync();
        // weight gradients: G[in][out] += sum over rows of act[row][in] * d[row][out]
        for (uint32_t ks = 0; ks < 4u; ks++) {
            const uint32_t r = ks * 4u + lq;
            for (uint32_t rb = 0; rb < 4u; rb++) {
                const uint32_t in = rb * 16u + lr16;
                const float a1 = fromBf16(X[r][in]), a2 = fromBf16(H1[r][in]), a3 = fromBf16(H2[r][in]);
                g3[rb] = mfma4(a3, Z[r][lr16], g3[rb]);
                for (uint32_t cb = 0; cb < 4u; cb++) {
                    const uint32_t out = cb * 16u + lr16;
                    g2[rb][cb] = mfma4(a2, D2[r][out], g2[rb][cb]);
                    g1[rb][cb] = mfma4(a1, D1[r][out], g1[rb][cb]);
                }
            }
        }
        for (uint32_t r = 0; r < 16u; r++) {
            gb1 += D1[r][lane];
            gb2 += D2[r][lane];
            if (lane < kProposalOutputs) gb3 += Z[r][lane];
        }
        waveSync();
    }
    float* p = partial + size_t(blockIdx.x) * kProposalParams;
    for (uint32_t rb = 0; rb < 4u; rb++)
        for (int e = 0; e < 4; e++) {
            const uint32_t in = rb * 16u + lq * 4u + e;
            for (uint32_t cb = 0; cb < 4u; cb++) {
                const uint32_t out = cb * 16u + lr16;
                if (in < kFeatures) p[in * kHidden + out] = g1[rb][cb][e];
                p[kOffW2 + in * kHidden + out] = g2[rb][cb][e];
            }
            if (lr16 < kProposalOutputs) p[kOffW3 + in * kProposalOutputs + lr16] = g3[rb][e];
        }
    p[kOffB1 + lane] = gb1;
    p[kOffB2 + lane] = gb2;
    if (lane < kProposalOutputs) p[kOffB3 + lane] = gb3;
    float tot = loss;
    for (int o = 32; o >= 1; o >>= 1) tot += __shfl_xor(tot, o, 64);
    if (lane == 0) lossPartial[blockIdx.x] = tot;
}

// Adam on the FP32 master weights with the gradient summed over the waves'
// partials, the BF16-rounded copy the backward pass reads, and the epoch's
// summed loss
__global__ void k_adam(float* w, const float* __restrict__ partial, uint32_t waves, float* m, float* v, float* wq,
                       float lr, float b1t, float b2t, const float* __restrict__ lossPartial, float* epochLoss) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx == 0) {
        float l = 0.f;
        for (uint32_t q = 0; q < waves; q++) l += lossPartial[q];
        *epochLoss += l;
    }
    if (idx >= kProposalParams) return;
    float gi = 0.f;
    for (uint32_t q = 0; q < waves; q++) gi += partial[size_t(q) * kProposalParams + idx];
    const float mi = 0.9f * m[idx] + 0.1f * gi;
    const float vi = 0.999f * v[idx] + 0.001f * gi * gi;
    m[idx] = mi;
    v[idx] = vi;
    const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
    const float wi = w[idx] - lr * mh / (__builtin_sqrtf(vh) + 1e-8f);
    w[idx] = wi;
    wq[idx] = fromBf16(toBf16(wi));
}

// FP32 master parameters -> the inference layout (transposed, padded BF16)
__global__ void k_pack_weights(const float* p, uint16_t* w1, float* b1, uint16_t* w2, float* b2, uint16_t* w3,
                               float* b3) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const float* W1 = p;
    const float* B1 = W1 + kFeatures * kHidden;
    const float* W2 = B1 + kHidden;
    const float* B2 = W2 + kHidden * kHidden;
    const float* W3 = B2 + kHidden;
    const float* B3 = W3 + kHidden * kProposalOutputs;
    if (idx < kHidden * kHidden) {
        const uint32_t o = idx / kHidden, in = idx % kHidden;
        w1[idx] = toBf16(in < kFeatures ? W1[in * kHidden + o] : 0.f);
        w2[idx] = toBf16(W2[in * kHidden + o]);
    }
    if (idx < 16u * kHidden) {
        const uint32_t o = idx / kHidden, in = idx % kHidden;
        w3[idx] = toBf16(o < kProposalOutputs ? W3[in * kProposalOutputs + o] : 0.f);
    }
    if (idx < kHidden) {
        b1[idx] = B1[idx];
        b2[idx] = B2[idx];
    }
    if (idx < 16u) b3[idx] = idx < kProposalOutputs ? B3[idx] : 0.f;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t cellsKLaunch(const CellCtx& c, const uint32_t* knn, uint32_t K, const uint8_t* cls, uint8_t* labels,
                        uint32_t mode, uint32_t fallbackK, hipStream_t st) {
    if (K > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cells_k, dim3(c.n), dim3(64), 0, st, c, knn, K, cls, labels, mode, fallbackK);
    return hipGetLastError();
}

// Classes given by the caller in input order, moved to sorted order.
__global__ __launch_bounds__(256) void k_gather_classes(const uint32_t* __restrict__ indices,
                                                        const uint8_t* __restrict__ given, uint32_t n,
                                                        uint8_t* cls) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) cls[i] = given[indices[i]];
}

hipError_t gatherClassesLaunch(const uint32_t* indices, const uint8_t* given, uint32_t n, uint8_t* cls,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_gather_classes, dim3((n + 255u) / 256u), dim3(256), 0, st, indices, given, n, cls);
    return hipGetLastError();
}

// Per-site search radii given by the caller in input order, moved to sorted order.
__global__ __launch_bounds__(256) void k_gather_radii(const uint32_t* __restrict__ indices,
                                                      const float* __restrict__ given, uint32_t n, float* radii) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) radii[i] = given[indices[i]];
}

hipError_t gatherRadiiLaunch(const uint32_t* indices, const float* given, uint32_t n, float* radii, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_radii, dim3((n + 255u) / 256u), dim3(256), 0, st, indices, given, n, radii);
    return hipGetLastError();
}

hipError_t featuresLaunch(const CellCtx& c, const uint32_t* knn, uint32_t K, float* feat, hipStream_t st,
                          float* d16) {
    if (K < 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_features, dim3((c.n + 255u) / 256u), dim3(256), 0, st, c, knn, K, feat, d16);
    return hipGetLastError();
}

hipError_t mlpInferLaunch(const float* feat, uint32_t n, const ProposalWeights& w, uint8_t* cls, float* logits,
                          float cost, uint32_t fallbackK, uint32_t maxK, hipStream_t st, const float* d16,
                          float* radii) {
    if (radii && !d16) return hipErrorInvalidValue;
    const uint32_t tiles = (n + 15u) / 16u;
    const uint32_t grid = tiles < 8192u ? tiles : 8192u;
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(64), 0, st, feat, n, w.w1, w.b1, w.w2, w.b2, w.w3, w.b3, cls,
                       logits, cost, fallbackK, maxK, d16, radii);
    return hipGetLastError();
}

hipError_t packWeightsLaunch(const float* params, const ProposalWeights& w, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_weights, dim3((kHidden * kHidden + 255u) / 256u), dim3(256), 0, st, params, w.w1, w.b1,
                       w.w2, w.b2, w.w3, w.b3);
    return hipGetLastError();
}

uint32_t trainWaves(uint32_t rows) { return (rows + kTrainTiles * 16u - 1u) / (kTrainTiles * 16u); }

hipError_t trainGradLaunch(const TrainBuffers& t, const uint32_t* order, uint32_t rows, float invBatch,
                           hipStream_t st) {
    const uint32_t waves = trainWaves(rows);
    if (waves > t.maxWaves) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_train_grad, dim3(waves), dim3(64), 0, st, t.feat, t.cls, t.radius, order, rows, invBatch,
                       t.pw.w1,
                       t.pw.b1, t.pw.w2, t.pw.b2, t.pw.w3, t.pw.b3, t.wq, t.partial, t.lossPartial);
    return hipGetLastError();
}

// One step: the batch's gradient, Adam, and the forward weights of the next step.
hipError_t mlpTrainStep(const TrainBuffers& t, const uint32_t* order, uint32_t rows, float lr, uint32_t step,
                        hipStream_t st) {
    hipError_t e = trainGradLaunch(t, order, rows, 1.f / float(rows), st);
    if (e != hipSuccess) return e;
    const float b1t = __builtin_powf(0.9f, float(step + 1)), b2t = __builtin_powf(0.999f, float(step + 1));
    hipLaunchKernelGGL(k_adam, dim3((kProposalParams + 255u) / 256u), dim3(256), 0, st, t.w, t.partial,
                       trainWaves(rows), t.m, t.v, t.wq, lr, b1t, b2t, t.lossPartial, t.epochLoss);
    return packWeightsLaunch(t.w, t.pw, st);
}

}  // namespace apo_hip
