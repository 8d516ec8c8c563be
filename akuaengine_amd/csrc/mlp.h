// mlp.h -- the per-particle warm-start network: 16 features -> 64 -> 64 (ReLU) -> 4 outputs,
// BF16 weights and activations on the MFMA units with FP32 accumulation, FP32 master weights.
//
// Parameter blob (floats), the layout every entry point and the network file share:
//   W1 [16][64] (input-major), b1 [64], W2 [64][64], b2 [64], W3 [64][4], b3 [4]   (trainable)
//   fmean [16], finv [16]: features are normalised as (f - fmean) * finv before the first layer
//   ymul [4]: outputs are trained on targets / ymul and returned multiplied by ymul
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace akua {

constexpr int kMlpIn = 16, kMlpHid = 64, kMlpOut = 4;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHid;    // 1024
constexpr int kOffW2 = kOffB1 + kMlpHid;             // 1088
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;   // 5184
constexpr int kOffW3 = kOffB2 + kMlpHid;             // 5248
constexpr int kOffB3 = kOffW3 + kMlpHid * kMlpOut;   // 5504
constexpr int kMlpTrainable = kOffB3 + kMlpOut;      // 5508
constexpr int kOffMean = kMlpTrainable;              // 5508
constexpr int kOffInv = kOffMean + kMlpIn;           // 5524
constexpr int kOffYmul = kOffInv + kMlpIn;           // 5540
constexpr int kMlpParams = kOffYmul + kMlpOut;       // 5544
constexpr int kMlpPartRow = kMlpTrainable + 1;       // gradient partial row: params, then loss

// Normalised targets are clipped to [-kMlpTargetClip, kMlpTargetClip] in the loss, so that the
// rare particles of a splash do not dominate a minibatch.
constexpr float kMlpTargetClip = 8.0f;

// Loss weights per output in normalised units (three position components, lambda).
struct MlpLossWeights {
    float w[4];
};

// Workgroups of the gradient kernel: each of its kMlpGradWaves waves writes one partial row.
constexpr int kMlpGradBlocks = 128;
constexpr int kMlpGradWaves = 2;
constexpr int kMlpPartRows = kMlpGradBlocks * kMlpGradWaves;

struct DeviceState;

// Outputs (denormalised) of n feature rows: y[k*4 + o].
void launch_mlp_eval(hipStream_t st, const float* params, const float* feat, uint32_t n, float* y);

// The learned step's warm start: XsWarm = XsPred + h * (network position outputs) for every
// particle slot, lamHat = the network's lambda output.  Features are read from d.feat.
void launch_mlp_apply(hipStream_t st, const DeviceState& d, const float* params, float h,
                      const float4* XsPred, float4* XsWarm, float* lamHat);

// Per-wave gradient partials of the mean loss over n samples (rows of feat/targ), for
// launch_mlp_reduce / launch_mlp_adam.  part: kMlpPartRows * kMlpPartRow floats.
void launch_mlp_grad(hipStream_t st, const float* params, const float* feat, const float* targ,
                     uint32_t n, MlpLossWeights lw, float* part);
// Sums the partial rows in a fixed order: grad[kMlpTrainable], loss -> *loss.
void launch_mlp_reduce(hipStream_t st, const float* part, float* grad, float* loss);
// Sums the partials and applies one Adam step to the trainable parameters in place.
// c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t).
void launch_mlp_adam(hipStream_t st, const float* part, float* params, float* m, float* v, float lr,
                     float beta1, float beta2, float eps, float c1, float c2, float* loss);

}  // namespace akua
