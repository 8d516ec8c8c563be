// selftest.hip -- on-device check that the written-out division (and its shared reciprocal) and
// square root (pbf_math.h) return the same bits as the correctly rounded '/' and sqrtf on this GPU.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/akua.h"
#include "pbf_math.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ bool same_bits(float a, float b) {
    return (__float_as_uint(a) == __float_as_uint(b)) || (a != a && b != b);
}

// A float with a random sign, a random exponent in [lo, hi] and a random mantissa.
__device__ __forceinline__ float rand_float(uint64_t z, int lo, int hi) {
    const uint32_t m = (uint32_t)z & 0x7FFFFFu;
    const int e = lo + (int)((z >> 23) % (uint64_t)(hi - lo + 1));
    const uint32_t sign = (uint32_t)(z >> 63) << 31;
    return __uint_as_float(sign | ((uint32_t)(e + 127) << 23) | m);
}

__global__ void k_selftest(uint64_t n, uint64_t seed, unsigned long long* counts) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t z0 = mix64(seed + 0x9E3779B97F4A7C15ull * (i + 1));
    const uint64_t z1 = mix64(z0 ^ 0xD1B54A32D192ED03ull);
    const uint64_t z2 = mix64(z1 + 0x2545F4914F6CDD1Dull);
    unsigned long long bad[4] = {0, 0, 0, 0};
    // division: divisors over the whole normal range, numerators over the whole range with
    // zeros, denormals and infinities mixed in
    {
        float b = fabsf(rand_float(z0, -126, 127));
        float a = rand_float(z1, -149 + 23, 127);
        switch (z2 & 15) {
            case 0: a = 0.0f; break;
            case 1: a = -0.0f; break;
            case 2: a = __uint_as_float((uint32_t)z1 & 0x807FFFFFu); break;  // denormal
            case 3: a = rand_float(z1, -70, 70); b = fabsf(rand_float(z0, -20, 1)); break;
            case 4: a = rand_float(z1, -16, 16); b = fabsf(rand_float(z0, -17, -3)); break;
            default: break;
        }
        const akua::Recip R = akua::recip(b);
        if (!same_bits(akua::div_rn(a, R), a / b)) bad[0]++;
        // the shared reciprocal is the correctly rounded one over the divisors div_rn keeps
        const float c = fabsf(rand_float(z2, -40, 39));
        if (!same_bits(akua::recip(c).r, 1.0f / c)) bad[0]++;
    }
    // square root: x >= 2^-96 must match sqrtf; below, the result must stay under 2^-47
    {
        const float x = fabsf(rand_float(z2, -96, 127));
        if (!same_bits(akua::sqrt_rn_unscaled(x), sqrtf(x))) bad[1]++;
        const float y = __uint_as_float((uint32_t)(z0 >> 11) % 0x0F800000u);  // [0, 2^-96)
        if (!(akua::sqrt_rn_unscaled(y) < 0x1p-47f)) bad[2]++;
        const float w = (float)(z1 >> 40) * 0x1p-26f;  // squared lengths of the SPH range
        if (!same_bits(akua::sqrt_rn_unscaled(w), sqrtf(w))) bad[3]++;
    }
    for (int k = 0; k < 4; ++k)
        if (bad[k]) atomicAdd(&counts[k], bad[k]);
}

}  // namespace

extern "C" akua_status akua_selftest_arith(uint64_t n, uint64_t seed, uint64_t mismatches[4]) {
    if (!mismatches) return AKUA_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return AKUA_ERR_NO_DEVICE;
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, 4 * sizeof(unsigned long long)) != hipSuccess) return AKUA_ERR_OUT_OF_MEMORY;
    akua_status st = AKUA_OK;
    if (hipMemset(d, 0, 4 * sizeof(unsigned long long)) != hipSuccess) st = AKUA_ERR_HIP;
    const uint64_t chunk = 1ull << 26;
    for (uint64_t off = 0; st == AKUA_OK && off < n; off += chunk) {
        const uint64_t m = (n - off < chunk) ? n - off : chunk;
        k_selftest<<<(unsigned)((m + 255) / 256), 256>>>(m, seed + off, d);
        if (hipGetLastError() != hipSuccess) st = AKUA_ERR_HIP;
    }
    unsigned long long h[4] = {0, 0, 0, 0};
    if (st == AKUA_OK && hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
        st = AKUA_ERR_HIP;
    for (int k = 0; k < 4; ++k) mismatches[k] = h[k];
    (void)hipFree(d);
    return st;
}
