// pbf_math.h -- device arithmetic of the PBF step.
//
// Every file of the library is compiled with -ffp-contract=off and correctly rounded FP32
// division and square root, so each `a * b + c` below rounds twice, exactly like the CPU oracle.
// Operation order follows the reference kernels (file:line next to each helper): glm's dot is
// (x*x + y*y) + z*z, vec*scalar rounds per component, integer powers are left-to-right products.
#pragma once

#include <hip/hip_runtime.h>

#include "pbf_types.h"

namespace akua {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 make3(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }

// glm::dot for vec3 (glm/detail/func_geometric.inl:48-54)
__device__ __forceinline__ float dot3(f3 a, f3 b) {
    float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
    return (tx + ty) + tz;
}
__device__ __forceinline__ float length3(f3 a) { return sqrtf(dot3(a, a)); }

// MathUtilsCUDA::cross (MathUtilsCUDA.h:12-18)
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// device_poly6 (SmoothingKernelsCUDA.h:16-21): cp * (h^2 - r^2)^3, 0 beyond h.
__device__ __forceinline__ float poly6(float d2, const SphParams& P) {
    if (d2 > P.h2) return 0.0f;
    float t = P.h2 - d2;
    return P.cp * ((t * t) * t);
}

// device_gradient_spiky (SmoothingKernelsCUDA.h:23-28): cs * (h - |r|)^2 * r / |r|.
__device__ __forceinline__ f3 grad_spiky(f3 r, const SphParams& P) {
    float rm = length3(r);
    if (rm > P.h || rm < 1e-5f) return {0.0f, 0.0f, 0.0f};
    float t = P.h - rm;
    float s = P.cs * (t * t);
    return (s * r) / rm;
}

// powf(x, n) of the artificial pressure (ConstraintSolverCUDA.cu:124): an integer n is the n-fold
// left-to-right product; any other n calls powf.
__device__ __forceinline__ float corr_pow(float x, const SphParams& P) {
    int k = P.corrNInt;
    if (k < 0) return powf(x, P.corrN);
    if (k == 0) return 1.0f;
    float r = x;
    for (int i = 1; i < k; ++i) r = r * x;
    return r;
}

// float -> int with NaN -> 0 and saturation, spelled out so the result never depends on what
// the compiler assumes about an out-of-range conversion.
__device__ __forceinline__ int to_int_sat(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x < -2147483648.0f) return (int)0x80000000u;
    return (int)x;
}

// device_discretize_position (NeighbourSearchCUDA.cu:15-21)
__device__ __forceinline__ void discretize(f3 p, float cs, int& cx, int& cy, int& cz) {
    cx = to_int_sat(floorf(p.x / cs));
    cy = to_int_sat(floorf(p.y / cs));
    cz = to_int_sat(floorf(p.z / cs));
}

// device_get_hash (NeighbourSearchCUDA.cu:23-27), with the int multiply's wrap done in uint32
// and the modulo by the scene's table size done by multiply-high (exact for every input).
__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, const SceneDev& sc) {
    uint32_t a = (uint32_t)cx * 73856093u;
    uint32_t b = (uint32_t)cy * 19349663u;
    uint32_t c = (uint32_t)cz * 83492791u;
    return fast_mod(a ^ b ^ c, sc.tableSize, sc.modMagic, sc.modShift);
}

// kernel_predict_position (IntegrationCUDA.cu:27-36): v* = v + g dt; x* = x + v* dt.
__device__ __forceinline__ f3 predict_x(f3 x, f3 v, const StepParams& S) {
    f3 g{S.g[0], S.g[1], S.g[2]};
    f3 vs = v + g * S.dt;
    return x + vs * S.dt;
}

// First slot of a hash in the scene's open-addressing cell table.
__device__ __forceinline__ uint32_t table_slot(uint32_t hash, uint32_t log2cap) {
    return (hash * 2654435761u) >> (32u - log2cap);
}

}  // namespace akua
