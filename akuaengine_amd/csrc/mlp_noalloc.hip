// mlp_noalloc.hip -- one Adam step of the warm-start network that allocates no memory
// (akua_mlp_adam_step_device): parameters, moments, samples and the loss are the caller's device
// buffers, and the gradient partial rows live in a file-scope device array of fixed size.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/akua.h"
#include "last_error.h"
#include "mlp.h"

namespace akua {

namespace {

// The gradient partial rows of k_mlp_grad: kMlpPartRows rows of kMlpPartRow floats, 5.64 MB,
// whatever the number of samples.  Every call rewrites every row before k_mlp_adam reads them,
// so nothing needs clearing between calls.
constexpr size_t kScratchFloats = (size_t)kMlpPartRows * kMlpPartRow;
static __device__ __align__(256) float scratch[kScratchFloats];

// Device address of `scratch` per device, taken once per device with that device current.  In
// host code the array's name is not a device address: only these pointers reach the kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
float* g_scratchPtr[kMaxDevices] = {};

hipError_t scratch_on_current_device(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[dev] = static_cast<float*>(p);
    }
    *out = g_scratchPtr[dev];
    return hipSuccess;
}

akua_status fail_with(akua_status code, const std::string& message) {
    set_last_error(message);
    return code;
}

}  // namespace

}  // namespace akua

extern "C" akua_status akua_mlp_adam_step_device(float* params, float* m, float* v, int64_t t,
                                                 float learningRate, const float* features,
                                                 const float* targets, int32_t n,
                                                 const float lossWeights[4], float* loss,
                                                 void* stream) {
    using namespace akua;
    if (!params || !m || !v || !features || !targets || !lossWeights || !loss)
        return fail_with(AKUA_ERR_INVALID_ARG, "NULL argument");
    if (n < 1) return fail_with(AKUA_ERR_INVALID_ARG, "n must be >= 1");
    if (t < 1) return fail_with(AKUA_ERR_INVALID_ARG, "t must be >= 1");
    if (!(learningRate > 0.0f) || !std::isfinite(learningRate))
        return fail_with(AKUA_ERR_INVALID_ARG, "learningRate must be > 0");
    for (int k = 0; k < 4; ++k)
        if (!(lossWeights[k] >= 0.0f) || !std::isfinite(lossWeights[k]))
            return fail_with(AKUA_ERR_INVALID_ARG, "loss weights must be >= 0");
    float* part = nullptr;
    hipError_t e = scratch_on_current_device(&part);
    if (e != hipSuccess)
        return fail_with(AKUA_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const MlpLossWeights lw{{lossWeights[0], lossWeights[1], lossWeights[2], lossWeights[3]}};
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)t)));
    launch_mlp_grad(st, params, features, targets, (uint32_t)n, lw, part);
    launch_mlp_adam(st, part, params, m, v, learningRate, 0.9f, 0.999f, 1e-8f, c1, c2, loss);
    e = hipGetLastError();
    if (e != hipSuccess) return fail_with(AKUA_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    set_last_error("");
    return AKUA_OK;
}
