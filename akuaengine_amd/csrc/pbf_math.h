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

// ---- correctly rounded division and square root, written out ------------------------------
// The gfx950 sequence for a correctly rounded a / b is
//   den = div_scale(b); num = div_scale(a); r = rcp(den); r = fma(fma(-den, r, 1), r, r);
//   q = num * r; t = fma(-den, q, num); q = fma(t, r, q); t = fma(-den, q, num);
//   q = div_fmas(t, r, q); q = div_fixup(q, b, a).
// For a positive divisor and a numerator of magnitude in [2^-40, 2^40] neither div_scale scales
// (exponents far from the limits, difference below 96) nor div_fixup changes anything, and
// div_fmas is a plain fma.  Two facts shorten the rest:
//  - the refined reciprocal r = fma(fma(-b, rcp(b), 1), rcp(b), rcp(b)) is RN(1/b), the correctly
//    rounded reciprocal, for every significand of b (checked exhaustively on the device; the
//    exponent only scales rcp's result in this range), and it is shared by every numerator over
//    one divisor;
//  - with y = RN(1/b), q = RN(a y) and the exact remainder t = a - b q (one fma), RN(q + t y) is
//    RN(a / b) (Markstein's theorem on the final correction of a division), so one correction
//    step instead of two gives the same quotient.
// Outside that range the '/' operator is used.  akua_selftest_arith checks the reciprocal and the
// quotient against '/' on the device.
struct Recip {
    float b, r;
};

__device__ __forceinline__ Recip recip(float b) {
    float r = __builtin_amdgcn_rcpf(b);
    const float e = fmaf(-b, r, 1.0f);
    r = fmaf(e, r, r);
    return {b, r};
}

// a / b by the shared reciprocal, valid (equal to '/') when in_div_range(a) and
// in_div_range(b).  The quotient takes a's sign bit: for b > 0 a nonzero quotient has it
// already, and a zero numerator gives a quotient of magnitude zero whose sign the fmas may
// lose (one bit insert instead of a compare and a select).
__device__ __forceinline__ float div_rn_unchecked(float a, const Recip& R) {
    float q = a * R.r;
    const float t = fmaf(-R.b, q, a);
    q = fmaf(t, R.r, q);
    return __uint_as_float((__float_as_uint(q) & 0x7FFFFFFFu) | (__float_as_uint(a) & 0x80000000u));
}

// Range tests on the bit patterns: |x| in [2^-40, 2^40] (0x2B800000 .. 0x53800000) or x == 0.
// For several numerators at once, (|bits| - 1) maps zero above every finite pattern, so one
// unsigned min and one max test them all.
constexpr uint32_t kDivLo = 0x2B800000u, kDivHi = 0x53800000u;

__device__ __forceinline__ uint32_t mag_bits(float x) { return __float_as_uint(x) & 0x7FFFFFFFu; }

// 2 (|bits| - 1) mod 2^32, one shift-and-add: the sign bit shifts out, and zero maps to
// 0xFFFFFFFE, above every finite pattern.  So low2(x) >= kDivLo2 exactly when |x| >= 2^-40 or
// x == 0, the lower half of in_div_range.
constexpr uint32_t kDivLo2 = 2u * kDivLo - 2u;
__device__ __forceinline__ uint32_t low2(float x) { return (__float_as_uint(x) << 1) - 2u; }

__device__ __forceinline__ bool in_div_range(float x) {
    const uint32_t m = mag_bits(x);
    return m - 1u >= kDivLo - 1u && m <= kDivHi;
}

// b > 0 and |b| in [2^-40, 2^40]
__device__ __forceinline__ bool divisor_ok(float b) {
    const uint32_t u = __float_as_uint(b);
    return u >= kDivLo && u <= kDivHi;
}

__device__ __forceinline__ float div_rn(float a, const Recip& R) {
    float q = div_rn_unchecked(a, R);
    if (!(in_div_range(a) && divisor_ok(R.b))) q = a / R.b;
    return q;
}

// (x, y, z) / b with one range check for the four operands.
__device__ __forceinline__ f3 div3_rn(f3 v, float b) {
    const Recip R = recip(b);
    f3 q{div_rn_unchecked(v.x, R), div_rn_unchecked(v.y, R), div_rn_unchecked(v.z, R)};
    const uint32_t mx = mag_bits(v.x), my = mag_bits(v.y), mz = mag_bits(v.z);
    const uint32_t lo = min(min(mx - 1u, my - 1u), mz - 1u);
    const uint32_t hi = max(max(mx, my), mz);
    const bool ok = divisor_ok(b) && lo >= kDivLo - 1u && hi <= kDivHi;
    if (!ok) q = {v.x / b, v.y / b, v.z / b};
    return q;
}

// a / b where b and an upper bound of |a| are known to lie in the exact-division range: only
// the lower bound of |a| (zero or >= 2^-40) is tested.
// The fallback is marked unlikely so that the compiler keeps it behind a branch that every wave
// skips, instead of computing the full division for every term and selecting.
__device__ __forceinline__ float div_rn_small(float a, const Recip& R) {
    float q = div_rn_unchecked(a, R);
    if (__builtin_expect(low2(a) < kDivLo2, 0)) q = a / R.b;
    return q;
}

// div_rn_small for a numerator that is +0 or positive: the one-correction quotient keeps the
// sign by itself there (+0 stays +0), so no sign fix is needed.
__device__ __forceinline__ float div_rn_small_nonneg(float a, const Recip& R) {
    float q = a * R.r;
    const float t = fmaf(-R.b, q, a);
    q = fmaf(t, R.r, q);
    if (__builtin_expect(low2(a) < kDivLo2, 0)) q = a / R.b;
    return q;
}

// (x, y, z) / b on the same terms as div_rn_small.
__device__ __forceinline__ f3 div3_rn_small(f3 v, float b) {
    const Recip R = recip(b);
    f3 q{div_rn_unchecked(v.x, R), div_rn_unchecked(v.y, R), div_rn_unchecked(v.z, R)};
    const uint32_t lo = min(min(low2(v.x), low2(v.y)), low2(v.z));
    if (lo < kDivLo2) q = {v.x / b, v.y / b, v.z / b};
    return q;
}

// Correctly rounded sqrt for x >= 2^-96 (no scaling needed there): the 1-ulp hardware root,
// corrected by the residuals of its two neighbours, as the compiler's own lowering does.  Below
// 2^-96 it returns a value below 2^-47 (not necessarily correctly rounded): callers use it only
// where every such value leads to the same branch.
__device__ __forceinline__ float sqrt_rn_unscaled(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float em = fmaf(-sm, s, x);
    const float ep = fmaf(-sp, s, x);
    s = (em <= 0.0f) ? sm : s;
    s = (ep > 0.0f) ? sp : s;
    return s;
}

// device_gradient_spiky (SmoothingKernelsCUDA.h:23-28): cs * (h - |r|)^2 * r / |r|.  |r| and
// the three quotients are correctly rounded (the reference's length and '/'): a squared length
// below 2^-96 gives |r| < 1e-5, the zero branch, whatever its last bits.  kFast: |r| lies in
// [1e-5, h] and |cs| h^3 bounds the numerators, both inside the exact-division range
// (SphParams::fast), so only the numerators' lower bound is tested.
// grad_spiky_nonzero: the same gradient into g, or false (g untouched) where it is zero, for
// callers whose sums a zero gradient leaves unchanged.
template <bool kFast = false>
__device__ __forceinline__ bool grad_spiky_nonzero(f3 r, const SphParams& P, f3& g) {
    const float d2 = dot3(r, r);
    const float rm = sqrt_rn_unscaled(d2);
    if (rm > P.h || rm < 1e-5f) return false;
    float t = P.h - rm;
    float s = P.cs * (t * t);
    g = kFast ? div3_rn_small(s * r, rm) : div3_rn(s * r, rm);
    return true;
}

template <bool kFast = false>
__device__ __forceinline__ f3 grad_spiky(f3 r, const SphParams& P) {
    f3 g{0.0f, 0.0f, 0.0f};
    grad_spiky_nonzero<kFast>(r, P, g);
    return g;
}

// powf(x, n) of the artificial pressure (ConstraintSolverCUDA.cu:124): an integer n is the n-fold
// left-to-right product; any other n calls powf.
__device__ __forceinline__ float corr_pow(float x, const SphParams& P) {
    int k = P.corrNInt;
    if (k == 4) return ((x * x) * x) * x;  // the usual exponent, without the loop
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

// Key of a cell of the non-parity search over h / 2 cells: the low 10 bits of each coordinate,
// packed (never kEmptyKey).  Cells less than 1024 apart along every axis get distinct keys, so no
// two cells of one 5x5x5 stencil share a bucket.
__device__ __forceinline__ uint32_t cell_key_fine(int cx, int cy, int cz) {
    return (((uint32_t)cx & 1023u) << 20) | (((uint32_t)cy & 1023u) << 10) | ((uint32_t)cz & 1023u);
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
