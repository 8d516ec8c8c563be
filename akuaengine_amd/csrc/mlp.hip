// mlp.hip -- forward, backward and Adam kernels of the warm-start network on the gfx950 MFMA
// units (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation).
//
// Every product is C(16x16) += A(16xK) B(Kx16) with both operands stored K-contiguous in LDS:
// A as rows [m][k], B as its transpose [n][k].  For the 16x16x16 instruction lane l = 16 g + r
// supplies A[r][4g..4g+3] and B[4g..4g+3][r] and receives C[4g+i][r], i = 0..3.  A wave owns
// 16 samples at a time; activations are kept in LDS in both orientations ([sample][unit] for the
// next layer's A operand, [unit][sample] for the weight-gradient products, which contract over
// samples).
#include "mlp.h"

#include <algorithm>

#include "pbf_kernels.h"

namespace akua {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (inputs here are finite).
__device__ __forceinline__ uint16_t to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// LDS written by some lanes of a wave and read by others: complete the writes and keep the
// compiler from moving accesses across this point.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ f32x4 mma16(f32x4 c, const uint16_t* a, int lda, const uint16_t* bt,
                                       int ldb, int K) {
    const int l = __lane_id();
    const int r = l & 15, g = l >> 4;
    for (int kb = 0; kb < K; kb += 16) {
        const bf16x4 av = *reinterpret_cast<const bf16x4*>(a + r * lda + kb + 4 * g);
        const bf16x4 bv = *reinterpret_cast<const bf16x4*>(bt + r * ldb + kb + 4 * g);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, c, 0, 0, 0);
    }
    return c;
}

// BF16 weights in the orientations the products read, plus the FP32 vectors.
struct alignas(16) WeightsLds {
    uint16_t w1t[kMlpHid * kMlpIn];   // [o][f]
    uint16_t w2t[kMlpHid * kMlpHid];  // [o][i]
    uint16_t w3t[16 * kMlpHid];       // [o (padded to 16)][j]
    float b1[kMlpHid], b2[kMlpHid], b3[16];
    float mean[kMlpIn], inv[kMlpIn], ymul[kMlpOut];
};
struct alignas(16) WeightsBackLds {
    uint16_t w2[kMlpHid * kMlpHid];   // [i][o]
    uint16_t w3[kMlpHid * 16];        // [j][o (padded to 16)]
};

__device__ void load_weights(const float* __restrict__ P, WeightsLds& W, WeightsBackLds* B) {
    for (int t = threadIdx.x; t < kMlpIn * kMlpHid; t += blockDim.x) {
        const int f = t / kMlpHid, o = t % kMlpHid;
        W.w1t[o * kMlpIn + f] = to_bf16(P[kOffW1 + t]);
    }
    for (int t = threadIdx.x; t < kMlpHid * kMlpHid; t += blockDim.x) {
        const int i = t / kMlpHid, o = t % kMlpHid;
        const uint16_t w = to_bf16(P[kOffW2 + t]);
        W.w2t[o * kMlpHid + i] = w;
        if (B) B->w2[t] = w;
    }
    for (int t = threadIdx.x; t < 16 * kMlpHid; t += blockDim.x) {
        const int j = t / 16, o = t % 16;
        const uint16_t w = o < kMlpOut ? to_bf16(P[kOffW3 + j * kMlpOut + o]) : (uint16_t)0;
        W.w3t[o * kMlpHid + j] = w;
        if (B) B->w3[j * 16 + o] = w;
    }
    for (int t = threadIdx.x; t < kMlpHid; t += blockDim.x) {
        W.b1[t] = P[kOffB1 + t];
        W.b2[t] = P[kOffB2 + t];
    }
    for (int t = threadIdx.x; t < 16; t += blockDim.x) {
        W.b3[t] = t < kMlpOut ? P[kOffB3 + t] : 0.0f;
        W.mean[t] = P[kOffMean + t];
        W.inv[t] = P[kOffInv + t];
        if (t < kMlpOut) W.ymul[t] = P[kOffYmul + t];
    }
}

// Per-wave activations and gradients of training.
struct alignas(16) ActLds {
    uint16_t x[16 * 16], xt_[16 * 16];
    uint16_t h1[16 * kMlpHid], h1t_[kMlpHid * 16];
    uint16_t h2[16 * kMlpHid], h2t_[kMlpHid * 16];
    uint16_t g3[16 * 16], g3t[16 * 16];
    uint16_t g2[16 * kMlpHid], g2t[kMlpHid * 16];
    uint16_t g1t[kMlpHid * 16];
    float y[16 * 16];
    __device__ uint16_t* xt() { return xt_; }
    __device__ uint16_t* h1t() { return h1t_; }
    __device__ uint16_t* h2t() { return h2t_; }
};

// Normalised BF16 features of samples row0..row0+15 (rows >= n are zero) into A.x (and, for
// training, its transpose).  Training path only; inference keeps activations in registers.
template <bool kTrain, class Act>
__device__ __forceinline__ void load_features(const float* __restrict__ feat, uint32_t row0,
                                              uint32_t n, const WeightsLds& W, Act& A) {
    const int l = __lane_id();
    const int s = l & 15, f0 = 4 * (l >> 4);
    const uint32_t row = row0 + (uint32_t)s;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (row < n) v = *reinterpret_cast<const float4*>(feat + (size_t)row * kMlpIn + f0);
    const float fv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float xn = row < n ? (fv[k] - W.mean[f0 + k]) * W.inv[f0 + k] : 0.0f;
        const uint16_t b = to_bf16(xn);
        A.x[s * 16 + f0 + k] = b;
        if (kTrain) A.xt()[(f0 + k) * 16 + s] = b;
    }
}

// The forward pass of the wave's 16 samples (features already in A.x).  mask1/mask2 receive the
// ReLU activity bits of this lane's accumulator entries (bit 4*n0 + i).  A.y[s*16 + o] receives
// the normalised outputs (o < 4).
template <bool kTrain, class Act>
__device__ __forceinline__ void forward_tile(const WeightsLds& W, Act& A, uint32_t& mask1,
                                             uint32_t& mask2) {
    const int l = __lane_id();
    const int r = l & 15, g = l >> 4;
    mask1 = mask2 = 0u;
    wave_sync();
#pragma unroll
    for (int n0 = 0; n0 < 4; ++n0) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = mma16(c, A.x, 16, W.w1t + n0 * 16 * kMlpIn, kMlpIn, 16);
        const int j = n0 * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int s = 4 * g + i;
            const float z = c[i] + W.b1[j];
            const bool on = z > 0.0f;
            mask1 |= (on ? 1u : 0u) << (4 * n0 + i);
            const uint16_t b = to_bf16(on ? z : 0.0f);
            A.h1[s * kMlpHid + j] = b;
            if (kTrain) A.h1t()[j * 16 + s] = b;
        }
    }
    wave_sync();
#pragma unroll
    for (int n0 = 0; n0 < 4; ++n0) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = mma16(c, A.h1, kMlpHid, W.w2t + n0 * 16 * kMlpHid, kMlpHid, kMlpHid);
        const int j = n0 * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int s = 4 * g + i;
            const float z = c[i] + W.b2[j];
            const bool on = z > 0.0f;
            mask2 |= (on ? 1u : 0u) << (4 * n0 + i);
            const uint16_t b = to_bf16(on ? z : 0.0f);
            A.h2[s * kMlpHid + j] = b;
            if (kTrain) A.h2t()[j * 16 + s] = b;
        }
    }
    wave_sync();
    {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = mma16(c, A.h2, kMlpHid, W.w3t, kMlpHid, kMlpHid);
#pragma unroll
        for (int i = 0; i < 4; ++i) A.y[(4 * g + i) * 16 + r] = c[i] + W.b3[r];
    }
    wave_sync();
}

constexpr int kInferBlock = 256;
constexpr int kInferWaves = kInferBlock / 64;

__device__ __forceinline__ bf16x4 pack_bf16(float a, float b, float c, float d) {
    typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
    const bf16x2v lo = {(__bf16)a, (__bf16)b};  // v_cvt_pk_bf16_f32: round to nearest even
    const bf16x2v hi = {(__bf16)c, (__bf16)d};
    const uint32_t u0 = __builtin_bit_cast(uint32_t, lo), u1 = __builtin_bit_cast(uint32_t, hi);
    bf16x4 r;
    r[0] = (short)(u0 & 0xFFFFu);
    r[1] = (short)(u0 >> 16);
    r[2] = (short)(u1 & 0xFFFFu);
    r[3] = (short)(u1 >> 16);
    return r;
}

// Inference in the transposed orientation C[unit][sample] = W^T H^T: lane l = 16 g + r owns
// sample r; the accumulator of output tile m0 holds units 16 m0 + 4 g + i of that sample, which
// is exactly the B operand of the next layer's k-step m0, so activations stay in registers.  The
// weight fragments (A operands) are loaded from LDS into registers once per wave.
struct InferRegs {
    bf16x4 w1[4], w2[4][4], w3[4];
};

__device__ __forceinline__ void load_infer_regs(const WeightsLds& W, InferRegs& R) {
    const int l = __lane_id();
    const int r = l & 15, g = l >> 4;
#pragma unroll
    for (int m0 = 0; m0 < 4; ++m0) {
        R.w1[m0] = *reinterpret_cast<const bf16x4*>(W.w1t + (16 * m0 + r) * kMlpIn + 4 * g);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            R.w2[m0][kk] = *reinterpret_cast<const bf16x4*>(W.w2t + (16 * m0 + r) * kMlpHid + 16 * kk + 4 * g);
        R.w3[m0] = *reinterpret_cast<const bf16x4*>(W.w3t + r * kMlpHid + 16 * m0 + 4 * g);
    }
}

// Normalised outputs y[i] = output 4g + i of sample row0 + r (meaningful for g == 0).
__device__ __forceinline__ f32x4 infer_tile(const WeightsLds& W, const InferRegs& R,
                                            const float* __restrict__ feat, uint32_t row0,
                                            uint32_t n) {
    const int l = __lane_id();
    const int r = l & 15, g = l >> 4;
    const uint32_t row = row0 + (uint32_t)r;
    bf16x4 xb;
    if (row < n) {
        const float4 v = *reinterpret_cast<const float4*>(feat + (size_t)row * kMlpIn + 4 * g);
        xb = pack_bf16((v.x - W.mean[4 * g + 0]) * W.inv[4 * g + 0],
                       (v.y - W.mean[4 * g + 1]) * W.inv[4 * g + 1],
                       (v.z - W.mean[4 * g + 2]) * W.inv[4 * g + 2],
                       (v.w - W.mean[4 * g + 3]) * W.inv[4 * g + 3]);
    } else {
        xb = bf16x4{0, 0, 0, 0};
    }
    bf16x4 h1[4], h2[4];
#pragma unroll
    for (int m0 = 0; m0 < 4; ++m0) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w1[m0], xb, c, 0, 0, 0);
        const float* b = W.b1 + 16 * m0 + 4 * g;
        h1[m0] = pack_bf16(fmaxf(c[0] + b[0], 0.0f), fmaxf(c[1] + b[1], 0.0f),
                           fmaxf(c[2] + b[2], 0.0f), fmaxf(c[3] + b[3], 0.0f));
    }
#pragma unroll
    for (int m0 = 0; m0 < 4; ++m0) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w2[m0][kk], h1[kk], c, 0, 0, 0);
        const float* b = W.b2 + 16 * m0 + 4 * g;
        h2[m0] = pack_bf16(fmaxf(c[0] + b[0], 0.0f), fmaxf(c[1] + b[1], 0.0f),
                           fmaxf(c[2] + b[2], 0.0f), fmaxf(c[3] + b[3], 0.0f));
    }
    f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w3[kk], h2[kk], c, 0, 0, 0);
    const float* b = W.b3 + 4 * g;
    return f32x4{c[0] + b[0], c[1] + b[1], c[2] + b[2], c[3] + b[3]};
}

__global__ __launch_bounds__(kInferBlock) void k_mlp_eval(const float* __restrict__ P,
                                                          const float* __restrict__ feat,
                                                          uint32_t n, float* __restrict__ y) {
    __shared__ WeightsLds W;
    load_weights(P, W, nullptr);
    __syncthreads();
    InferRegs R;
    load_infer_regs(W, R);
    const int l = __lane_id();
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t wave = blockIdx.x * kInferWaves + threadIdx.x / 64;
    const uint32_t nwaves = gridDim.x * kInferWaves;
    for (uint32_t t = wave; t < tiles; t += nwaves) {
        const f32x4 o = infer_tile(W, R, feat, t * 16, n);
        const uint32_t row = t * 16 + (uint32_t)l;
        if (l < 16 && row < n)
            *reinterpret_cast<float4*>(y + (size_t)row * 4) =
                make_float4(o[0] * W.ymul[0], o[1] * W.ymul[1], o[2] * W.ymul[2], o[3] * W.ymul[3]);
    }
}

// The warm start of a learned step: x*_w = x*_pred + h * D for every particle slot of a scene,
// with D and the auxiliary lambda the network's outputs for the slot's features.
__global__ __launch_bounds__(kInferBlock) void k_mlp_apply(DeviceState d, const float* __restrict__ P,
                                                           float h, const float4* __restrict__ XsPred,
                                                           float4* __restrict__ XsWarm,
                                                           float* __restrict__ lamHat) {
    __shared__ WeightsLds W;
    load_weights(P, W, nullptr);
    __syncthreads();
    InferRegs R;
    load_infer_regs(W, R);
    const int l = __lane_id();
    const uint32_t n = d.nslots;
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t wave = blockIdx.x * kInferWaves + threadIdx.x / 64;
    const uint32_t nwaves = gridDim.x * kInferWaves;
    for (uint32_t t = wave; t < tiles; t += nwaves) {
        const f32x4 o = infer_tile(W, R, d.feat, t * 16, n);
        const uint32_t p = t * 16 + (uint32_t)l;
        if (l < 16 && p < n && p < d.scenes[d.chunkScene[p >> 8]].end) {
            const float4 x = XsPred[p];
            // the network was trained on targets within +-kMlpTargetClip normalised units: the
            // warm start does not extrapolate beyond them
            const float c = kMlpTargetClip;
            const float dx = fminf(fmaxf(o[0], -c), c) * W.ymul[0];
            const float dy = fminf(fmaxf(o[1], -c), c) * W.ymul[1];
            const float dz = fminf(fmaxf(o[2], -c), c) * W.ymul[2];
            XsWarm[p] = make_float4(x.x + h * dx, x.y + h * dy, x.z + h * dz, x.w);
            lamHat[p] = o[3] * W.ymul[3];
        }
    }
}

// Gradient partials: one row per wave (kMlpGradBlocks workgroups of kMlpGradWaves waves).
__global__ __launch_bounds__(64 * kMlpGradWaves) void k_mlp_grad(const float* __restrict__ P,
                                                                 const float* __restrict__ feat,
                                                                 const float* __restrict__ targ,
                                                                 uint32_t n, MlpLossWeights lw,
                                                                 float* __restrict__ part) {
    __shared__ WeightsLds W;
    __shared__ WeightsBackLds WB;
    __shared__ ActLds act[kMlpGradWaves];
    load_weights(P, W, &WB);
    __syncthreads();
    ActLds& A = act[threadIdx.x / 64];
    const int l = __lane_id();
    const int r = l & 15, g = l >> 4;
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t wave = blockIdx.x * kMlpGradWaves + threadIdx.x / 64;
    const uint32_t nwaves = gridDim.x * kMlpGradWaves;
    const float invN = 1.0f / (float)n;

    f32x4 acc1[4], acc2[4][4], acc3[4];
    float db1[4], db2[4], db3 = 0.0f, loss = 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        acc1[a] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        acc3[a] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        db1[a] = db2[a] = 0.0f;
#pragma unroll
        for (int b = 0; b < 4; ++b) acc2[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }

    for (uint32_t t = wave; t < tiles; t += nwaves) {
        load_features<true>(feat, t * 16, n, W, A);
        uint32_t m1, m2;
        forward_tile<true>(W, A, m1, m2);
        // output gradient of the mean weighted squared error, in normalised units
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int s = 4 * g + i;
            const uint32_t row = t * 16 + (uint32_t)s;
            float G = 0.0f;
            if (r < kMlpOut && row < n) {
                const float tn = fminf(fmaxf(targ[(size_t)row * kMlpOut + r] / W.ymul[r],
                                             -kMlpTargetClip), kMlpTargetClip);
                const float diff = A.y[s * 16 + r] - tn;
                loss += lw.w[r] * diff * diff;
                G = 2.0f * lw.w[r] * diff * invN;
            }
            db3 += G;
            const uint16_t b = to_bf16(G);
            A.g3[s * 16 + r] = b;
            A.g3t[r * 16 + s] = b;
        }
        wave_sync();
        // dH2 = G3 W3^T, masked by layer 2's activity
#pragma unroll
        for (int n0 = 0; n0 < 4; ++n0) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = mma16(c, A.g3, 16, WB.w3 + n0 * 16 * 16, 16, 16);
            const int j = n0 * 16 + r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int s = 4 * g + i;
                const float G = (m2 >> (4 * n0 + i)) & 1u ? c[i] : 0.0f;
                db2[n0] += G;
                const uint16_t b = to_bf16(G);
                A.g2[s * kMlpHid + j] = b;
                A.g2t[j * 16 + s] = b;
            }
        }
        wave_sync();
        // dH1 = G2 W2^T, masked by layer 1's activity
#pragma unroll
        for (int n0 = 0; n0 < 4; ++n0) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = mma16(c, A.g2, kMlpHid, WB.w2 + n0 * 16 * kMlpHid, kMlpHid, kMlpHid);
            const int j = n0 * 16 + r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int s = 4 * g + i;
                const float G = (m1 >> (4 * n0 + i)) & 1u ? c[i] : 0.0f;
                db1[n0] += G;
                A.g1t[j * 16 + s] = to_bf16(G);
            }
        }
        wave_sync();
        // weight gradients, contracted over the 16 samples
#pragma unroll
        for (int m0 = 0; m0 < 4; ++m0) acc3[m0] = mma16(acc3[m0], A.h2t() + m0 * 16 * 16, 16, A.g3t, 16, 16);
#pragma unroll
        for (int m0 = 0; m0 < 4; ++m0)
#pragma unroll
            for (int n0 = 0; n0 < 4; ++n0)
                acc2[m0][n0] = mma16(acc2[m0][n0], A.h1t() + m0 * 16 * 16, 16, A.g2t + n0 * 16 * 16, 16, 16);
#pragma unroll
        for (int n0 = 0; n0 < 4; ++n0) acc1[n0] = mma16(acc1[n0], A.xt(), 16, A.g1t + n0 * 16 * 16, 16, 16);
        wave_sync();
    }

    float* row = part + (size_t)wave * kMlpPartRow;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = 4 * g + i;
#pragma unroll
        for (int n0 = 0; n0 < 4; ++n0) row[kOffW1 + m * kMlpHid + n0 * 16 + r] = acc1[n0][i];
#pragma unroll
        for (int m0 = 0; m0 < 4; ++m0)
#pragma unroll
            for (int n0 = 0; n0 < 4; ++n0)
                row[kOffW2 + (m0 * 16 + m) * kMlpHid + n0 * 16 + r] = acc2[m0][n0][i];
        if (r < kMlpOut)
#pragma unroll
            for (int m0 = 0; m0 < 4; ++m0) row[kOffW3 + (m0 * 16 + m) * kMlpOut + r] = acc3[m0][i];
    }
    // bias gradients: sum the four lane groups of each column
#pragma unroll
    for (int n0 = 0; n0 < 4; ++n0) {
        float a = db1[n0], b = db2[n0];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        b += __shfl_xor(b, 16, 64);
        b += __shfl_xor(b, 32, 64);
        if (g == 0) {
            row[kOffB1 + n0 * 16 + r] = a;
            row[kOffB2 + n0 * 16 + r] = b;
        }
    }
    db3 += __shfl_xor(db3, 16, 64);
    db3 += __shfl_xor(db3, 32, 64);
    if (g == 0 && r < kMlpOut) row[kOffB3 + r] = db3;
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (l == 0) row[kMlpTrainable] = loss * invN;
}

__global__ void k_mlp_reduce(const float* __restrict__ part, float* __restrict__ grad,
                             float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpTrainable) return;
    float s = 0.0f;
    for (int w = 0; w < kMlpPartRows; ++w) s += part[(size_t)w * kMlpPartRow + p];
    if (p < kMlpTrainable) grad[p] = s;
    else *loss = s;
}

__global__ void k_mlp_adam(const float* __restrict__ part, float* __restrict__ P,
                           float* __restrict__ M, float* __restrict__ V, float lr, float b1,
                           float b2, float eps, float c1, float c2, float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpTrainable) return;
    float s = 0.0f;
    for (int w = 0; w < kMlpPartRows; ++w) s += part[(size_t)w * kMlpPartRow + p];
    if (p == kMlpTrainable) {
        *loss = s;
        return;
    }
    const float m = b1 * M[p] + (1.0f - b1) * s;
    const float v = b2 * V[p] + (1.0f - b2) * (s * s);
    M[p] = m;
    V[p] = v;
    P[p] = P[p] - lr * (m * c1) / (sqrtf(v * c2) + eps);
}

}  // namespace

// Workgroups of an inference launch: the waves loop over the 16-sample tiles, and each workgroup
// first stages the weights in LDS, so the grid is one resident set and no more: the kernels'
// 128 VGPRs leave room for 4 waves per SIMD, i.e. 4 workgroups of 4 waves per CU (1,024 on 256
// CUs measured 59 us per 1M-sample launch against 69 us with 2,048).
static uint32_t infer_blocks(uint32_t tiles) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    const uint32_t resident = 4u * (uint32_t)cus;
    return std::max<uint32_t>(1u, std::min<uint32_t>((tiles + kInferWaves - 1) / kInferWaves, resident));
}

void launch_mlp_eval(hipStream_t st, const float* params, const float* feat, uint32_t n, float* y) {
    if (!n) return;
    const uint32_t tiles = (n + 15) / 16;
    const uint32_t blocks = infer_blocks(tiles);
    k_mlp_eval<<<blocks, kInferBlock, 0, st>>>(params, feat, n, y);
}

void launch_mlp_apply(hipStream_t st, const DeviceState& d, const float* params, float h,
                      const float4* XsPred, float4* XsWarm, float* lamHat) {
    if (!d.nslots) return;
    const uint32_t tiles = (d.nslots + 15) / 16;
    const uint32_t blocks = infer_blocks(tiles);
    k_mlp_apply<<<blocks, kInferBlock, 0, st>>>(d, params, h, XsPred, XsWarm, lamHat);
}

void launch_mlp_grad(hipStream_t st, const float* params, const float* feat, const float* targ,
                     uint32_t n, MlpLossWeights lw, float* part) {
    k_mlp_grad<<<kMlpGradBlocks, 64 * kMlpGradWaves, 0, st>>>(params, feat, targ, n, lw, part);
}

void launch_mlp_reduce(hipStream_t st, const float* part, float* grad, float* loss) {
    k_mlp_reduce<<<(kMlpPartRow + 255) / 256, 256, 0, st>>>(part, grad, loss);
}

void launch_mlp_adam(hipStream_t st, const float* part, float* params, float* m, float* v, float lr,
                     float beta1, float beta2, float eps, float c1, float c2, float* loss) {
    k_mlp_adam<<<(kMlpPartRow + 255) / 256, 256, 0, st>>>(part, params, m, v, lr, beta1, beta2, eps,
                                                          c1, c2, loss);
}

}  // namespace akua
