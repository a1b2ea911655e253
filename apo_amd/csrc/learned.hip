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
                                                  float* __restrict__ feat) {
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
// zero-padded from 8).  v_mfma_f32_16x16x16_bf16 lane layout: A[row = l % 16]
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
                                                  uint32_t maxK) {
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
                for (uint32_t q = 0; q < kProposalClasses; q++) logitsOut[size_t(base + lane) * kProposalClasses + q] = Lg[lane][q];
        }
        waveSync();
    }
}

// ---------------------------------------------------------------------------
// Training (FP32).  Parameters, flat: W1 [52][64], b1 [64], W2 [64][64],
// b2 [64], W3 [64][8], b3 [8] (weights [in][out]).
// ---------------------------------------------------------------------------
// C[M][N] (+)= sum_k A(m, k) B(k, n); A(m, k) = ta ? A[k * lda + m] : A[m * lda + k]
__global__ void k_gemm(const float* A, uint32_t lda, uint32_t ta, const float* B, uint32_t ldb, uint32_t tb,
                       float* C, uint32_t ldc, uint32_t M, uint32_t N, uint32_t K, float beta) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * N) return;
    const uint32_t m = idx / N, nn = idx % N;
    float acc = 0.f;
    for (uint32_t k = 0; k < K; k++) {
        const float a = ta ? A[size_t(k) * lda + m] : A[size_t(m) * lda + k];
        const float b = tb ? B[size_t(nn) * ldb + k] : B[size_t(k) * ldb + nn];
        acc += a * b;
    }
    float* c = C + size_t(m) * ldc + nn;
    *c = beta != 0.f ? beta * *c + acc : acc;
}

// rows of x[batch][cols] += bias, optional ReLU
__global__ void k_bias_act(float* x, const float* bias, uint32_t rows, uint32_t cols, uint32_t relu) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * cols) return;
    float v = x[idx] + bias[idx % cols];
    if (relu) v = fmaxf(v, 0.f);
    x[idx] = v;
}

// softmax cross-entropy: grad = (softmax - onehot) / batch, loss summed
__global__ void k_softmax_xent(const float* logits, const uint8_t* labels, uint32_t rows, float* grad, float* loss) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float* z = logits + size_t(r) * kProposalClasses;
    float m = z[0];
    for (uint32_t q = 1; q < kProposalClasses; q++) m = fmaxf(m, z[q]);
    float s = 0.f;
    for (uint32_t q = 0; q < kProposalClasses; q++) s += __expf(z[q] - m);
    const uint32_t y = labels[r];
    for (uint32_t q = 0; q < kProposalClasses; q++) {
        const float p = __expf(z[q] - m) / s;
        grad[size_t(r) * kProposalClasses + q] = (p - (q == y ? 1.f : 0.f)) / float(rows);
    }
    atomicAdd(loss, (m + __logf(s)) - z[y]);
}

// d *= (h > 0)
__global__ void k_relu_back(float* d, const float* h, uint32_t count) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count && !(h[idx] > 0.f)) d[idx] = 0.f;
}

// g[col] = sum over rows of d[row][col]
__global__ void k_col_sum(const float* d, uint32_t rows, uint32_t cols, float* g) {
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= cols) return;
    float s = 0.f;
    for (uint32_t r = 0; r < rows; r++) s += d[size_t(r) * cols + col];
    g[col] = s;
}

// gather batch rows of features and labels
__global__ void k_gather(const float* feat, const uint8_t* labels, const uint32_t* order, uint32_t rows, float* x,
                         uint8_t* y) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * kFeatures) return;
    const uint32_t r = idx / kFeatures, col = idx % kFeatures;
    const uint32_t src = order[r];
    x[idx] = feat[size_t(src) * kFeatures + col];
    if (col == 0) y[r] = labels[src];
}

// Adam step on the FP32 master weights, then the BF16-rounded copy the
// forward pass uses
__global__ void k_adam(float* w, const float* g, float* m, float* v, float* wq, uint32_t count, float lr, float b1t,
                       float b2t) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const float gi = g[idx];
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
    const float* B3 = W3 + kHidden * kProposalClasses;
    if (idx < kHidden * kHidden) {
        const uint32_t o = idx / kHidden, in = idx % kHidden;
        w1[idx] = toBf16(in < kFeatures ? W1[in * kHidden + o] : 0.f);
        w2[idx] = toBf16(W2[in * kHidden + o]);
    }
    if (idx < 16u * kHidden) {
        const uint32_t o = idx / kHidden, in = idx % kHidden;
        w3[idx] = toBf16(o < kProposalClasses ? W3[in * kProposalClasses + o] : 0.f);
    }
    if (idx < kHidden) {
        b1[idx] = B1[idx];
        b2[idx] = B2[idx];
    }
    if (idx < 16u) b3[idx] = idx < kProposalClasses ? B3[idx] : 0.f;
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

hipError_t featuresLaunch(const CellCtx& c, const uint32_t* knn, uint32_t K, float* feat, hipStream_t st) {
    if (K < 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_features, dim3((c.n + 255u) / 256u), dim3(256), 0, st, c, knn, K, feat);
    return hipGetLastError();
}

hipError_t mlpInferLaunch(const float* feat, uint32_t n, const ProposalWeights& w, uint8_t* cls, float* logits,
                          float cost, uint32_t fallbackK, uint32_t maxK, hipStream_t st) {
    const uint32_t tiles = (n + 15u) / 16u;
    const uint32_t grid = tiles < 8192u ? tiles : 8192u;
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(64), 0, st, feat, n, w.w1, w.b1, w.w2, w.b2, w.w3, w.b3, cls,
                       logits, cost, fallbackK, maxK);
    return hipGetLastError();
}

hipError_t packWeightsLaunch(const float* params, const ProposalWeights& w, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_weights, dim3((kHidden * kHidden + 255u) / 256u), dim3(256), 0, st, params, w.w1, w.b1,
                       w.w2, w.b2, w.w3, w.b3);
    return hipGetLastError();
}

static inline dim3 blocksFor(size_t count) { return dim3(uint32_t((count + 255u) / 256u)); }

// One training step on a gathered batch; returns the summed loss in *loss.
hipError_t mlpTrainStep(const TrainBuffers& t, uint32_t rows, float lr, uint32_t step, hipStream_t st) {
    const float* W1 = t.wq;
    const float* B1 = W1 + kFeatures * kHidden;
    const float* W2 = B1 + kHidden;
    const float* B2 = W2 + kHidden * kHidden;
    const float* W3 = B2 + kHidden;
    const float* B3 = W3 + kHidden * kProposalClasses;
    float* gW1 = t.grad;
    float* gB1 = gW1 + kFeatures * kHidden;
    float* gW2 = gB1 + kHidden;
    float* gB2 = gW2 + kHidden * kHidden;
    float* gW3 = gB2 + kHidden;
    float* gB3 = gW3 + kHidden * kProposalClasses;
    // forward
    hipLaunchKernelGGL(k_gemm, blocksFor(rows * kHidden), dim3(256), 0, st, t.x, kFeatures, 0u, W1, kHidden, 0u, t.h1,
                       kHidden, rows, kHidden, kFeatures, 0.f);
    hipLaunchKernelGGL(k_bias_act, blocksFor(rows * kHidden), dim3(256), 0, st, t.h1, B1, rows, kHidden, 1u);
    hipLaunchKernelGGL(k_gemm, blocksFor(rows * kHidden), dim3(256), 0, st, t.h1, kHidden, 0u, W2, kHidden, 0u, t.h2,
                       kHidden, rows, kHidden, kHidden, 0.f);
    hipLaunchKernelGGL(k_bias_act, blocksFor(rows * kHidden), dim3(256), 0, st, t.h2, B2, rows, kHidden, 1u);
    hipLaunchKernelGGL(k_gemm, blocksFor(rows * kProposalClasses), dim3(256), 0, st, t.h2, kHidden, 0u, W3,
                       kProposalClasses, 0u, t.z, kProposalClasses, rows, kProposalClasses, kHidden, 0.f);
    hipLaunchKernelGGL(k_bias_act, blocksFor(rows * kProposalClasses), dim3(256), 0, st, t.z, B3, rows,
                       kProposalClasses, 0u);
    hipLaunchKernelGGL(k_softmax_xent, blocksFor(rows), dim3(256), 0, st, t.z, t.y, rows, t.dz, t.loss);
    // backward
    hipLaunchKernelGGL(k_gemm, blocksFor(kHidden * kProposalClasses), dim3(256), 0, st, t.h2, kHidden, 1u, t.dz,
                       kProposalClasses, 0u, gW3, kProposalClasses, kHidden, kProposalClasses, rows, 0.f);
    hipLaunchKernelGGL(k_col_sum, blocksFor(kProposalClasses), dim3(256), 0, st, t.dz, rows, kProposalClasses, gB3);
    hipLaunchKernelGGL(k_gemm, blocksFor(rows * kHidden), dim3(256), 0, st, t.dz, kProposalClasses, 0u, W3,
                       kProposalClasses, 1u, t.d2, kHidden, rows, kHidden, kProposalClasses, 0.f);
    hipLaunchKernelGGL(k_relu_back, blocksFor(rows * kHidden), dim3(256), 0, st, t.d2, t.h2, rows * kHidden);
    hipLaunchKernelGGL(k_gemm, blocksFor(kHidden * kHidden), dim3(256), 0, st, t.h1, kHidden, 1u, t.d2, kHidden, 0u,
                       gW2, kHidden, kHidden, kHidden, rows, 0.f);
    hipLaunchKernelGGL(k_col_sum, blocksFor(kHidden), dim3(256), 0, st, t.d2, rows, kHidden, gB2);
    hipLaunchKernelGGL(k_gemm, blocksFor(rows * kHidden), dim3(256), 0, st, t.d2, kHidden, 0u, W2, kHidden, 1u, t.d1,
                       kHidden, rows, kHidden, kHidden, 0.f);
    hipLaunchKernelGGL(k_relu_back, blocksFor(rows * kHidden), dim3(256), 0, st, t.d1, t.h1, rows * kHidden);
    hipLaunchKernelGGL(k_gemm, blocksFor(kFeatures * kHidden), dim3(256), 0, st, t.x, kFeatures, 1u, t.d1, kHidden, 0u,
                       gW1, kHidden, kFeatures, kHidden, rows, 0.f);
    hipLaunchKernelGGL(k_col_sum, blocksFor(kHidden), dim3(256), 0, st, t.d1, rows, kHidden, gB1);
    // update
    const float b1t = __builtin_powf(0.9f, float(step + 1)), b2t = __builtin_powf(0.999f, float(step + 1));
    hipLaunchKernelGGL(k_adam, blocksFor(kProposalParams), dim3(256), 0, st, t.w, t.grad, t.m, t.v, t.wq,
                       kProposalParams, lr, b1t, b2t);
    return hipGetLastError();
}

hipError_t gatherLaunch(const float* feat, const uint8_t* labels, const uint32_t* order, uint32_t rows, float* x,
                        uint8_t* y, hipStream_t st) {
    hipLaunchKernelGGL(k_gather, blocksFor(size_t(rows) * kFeatures), dim3(256), 0, st, feat, labels, order, rows, x,
                       y);
    return hipGetLastError();
}

}  // namespace apo_hip
