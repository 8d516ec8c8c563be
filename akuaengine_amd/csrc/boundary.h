// boundary.h -- boundary samples on the box walls (opt-in; no reference counterpart).
//
// The reference keeps particles in the box with a soft clamp and plans "virtual particles" in its
// place (ConstraintSolverCUDA.cu:132-135).  This extension follows Akinci et al. 2012 ("Versatile
// rigid-fluid coupling for incompressible SPH") in the PBF form of Bender's PositionBasedDynamics
// library: static samples b on the six walls, each with the volume weight
// psi_b = rho0 / sum_k W(x_b - x_k) over the samples k, add psi_b W(x*_i - x_b) to rho_i and
// psi_b grad W(x*_i - x_b) to the sum behind grad_i C_i; lambda_i times that gradient sum joins
// the sum behind dp_i (both sums are then scaled by 1/rho0).  The clamp and damping stay.
//
// Layout (BoundaryDev):
//   per axis a: n_a = max(2, int(ceil((hi_a - lo_a) / s)) + 1) points, spacing
//   d_a = (hi_a - lo_a) / (n_a - 1);
//   wall w = 0..5: normal axis a = w / 2, side w % 2 (lo, hi), tangential axes b < c;
//   layer l = 0..L-1 on the plane lo_a - l s (side 0) or hi_a + l s (side 1);
//   tangential index t along axis e runs over [-(L-1), n_e - 1 + (L-1)] when e > a (the wall
//   owns that edge) and over [1, n_e - 2] when e < a, at coordinate lo_e + t d_e;
//   sample index base_w + (l N_b + (t_b - t0_b)) N_c + (t_c - t0_c).
// A point x visits, per wall, the layers whose plane lies within h, the tangential indices from
// floor((x_e - lo_e - h) / d_e) - 1 to floor((x_e - lo_e + h) / d_e) + 1 (clamped to the wall),
// skipping rows and columns whose coordinate lies h or more away, and keeps the samples with
// |x - x_b|^2 < h^2; the visit runs in ascending sample index, which fixes every sum's order.
// The CPU oracle states the same definition independently (oracle/pbf_oracle.hpp).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "pbf_types.h"

namespace akua {

constexpr int kMaxBoundaryLayers = 4;
constexpr uint32_t kMaxBoundarySamples = 1u << 26;

__host__ __device__ inline long long floor_sat(float v) {
    if (!(v == v)) return 0;
    if (v >= 1e9f) return 1000000000ll;
    if (v <= -1e9f) return -1000000000ll;
    return (long long)floorf(v);
}

__host__ __device__ inline float wall_plane(const BoundaryDev& B, int w, int l) {
    const int a = w >> 1;
    return (w & 1) ? B.hi[a] + (float)l * B.s : B.lo[a] - (float)l * B.s;
}

// Calls f(index, x - x_b (three floats), |x - x_b|^2) for every sample that counts for (x, y, z),
// in ascending sample index.  The wall loop is unrolled so every axis index is a constant.
template <class F>
__host__ __device__ inline void boundary_visit(const BoundaryDev& B, float x, float y, float z,
                                               float h, float h2, F&& f) {
    const float xa[3] = {x, y, z};
#pragma unroll
    for (int w = 0; w < 6; ++w) {
        const int a = w >> 1;
        const int b = a == 0 ? 1 : 0;
        const int c = a == 2 ? 1 : 2;
        const int tb0 = B.t0[w][0], tnb = B.tn[w][0], tc0 = B.t0[w][1], tnc = B.tn[w][1];
        if (tnb == 0 || tnc == 0) continue;
        // the layers whose plane lies within h (most walls have none: skip their windows)
        uint32_t near = 0u;
        for (int l = 0; l < B.layers; ++l) {
            const float ra = xa[a] - wall_plane(B, w, l);
            if (ra * ra < h2) near |= 1u << l;
        }
        if (!near) continue;
        long long b0 = floor_sat((xa[b] - B.lo[b] - h) / B.d[b]) - 1;
        long long b1 = floor_sat((xa[b] - B.lo[b] + h) / B.d[b]) + 1;
        long long c0 = floor_sat((xa[c] - B.lo[c] - h) / B.d[c]) - 1;
        long long c1 = floor_sat((xa[c] - B.lo[c] + h) / B.d[c]) + 1;
        b0 = b0 < tb0 ? tb0 : b0;
        b1 = b1 > tb0 + tnb - 1 ? tb0 + tnb - 1 : b1;
        c0 = c0 < tc0 ? tc0 : c0;
        c1 = c1 > tc0 + tnc - 1 ? tc0 + tnc - 1 : c1;
        for (int l = 0; l < B.layers; ++l) {
            if (!((near >> l) & 1u)) continue;
            const float na = wall_plane(B, w, l);
            const float ra = xa[a] - na;
            for (int tb = (int)b0; tb <= (int)b1; ++tb) {
                const float cb = B.lo[b] + (float)tb * B.d[b];
                const float rb = xa[b] - cb;
                if (rb * rb >= h2) continue;
                for (int tc = (int)c0; tc <= (int)c1; ++tc) {
                    const float cc = B.lo[c] + (float)tc * B.d[c];
                    const float rc = xa[c] - cc;
                    if (rc * rc >= h2) continue;
                    float sep[3];
                    sep[a] = ra;
                    sep[b] = rb;
                    sep[c] = rc;
                    const float d2 = (sep[0] * sep[0] + sep[1] * sep[1]) + sep[2] * sep[2];
                    if (!(d2 < h2)) continue;
                    const uint32_t idx = B.base[w] + ((uint32_t)l * (uint32_t)tnb + (uint32_t)(tb - tb0)) *
                                                         (uint32_t)tnc +
                                         (uint32_t)(tc - tc0);
                    f(idx, sep[0], sep[1], sep[2], d2);
                }
            }
        }
    }
}

// True when some layer plane lies within h of the point along its normal: only then can a
// sample count for it.
__host__ __device__ inline bool boundary_near(const BoundaryDev& B, float x, float y, float z, float h2) {
    const float xa[3] = {x, y, z};
    bool near = false;
#pragma unroll
    for (int w = 0; w < 6; ++w)
        for (int l = 0; l < B.layers; ++l) {
            const float r = xa[w >> 1] - wall_plane(B, w, l);
            near = near || r * r < h2;
        }
    return near;
}

// Host side (boundary.cpp).
// Builds the layout for a box and computes psi; returns false (with a message) when the box or the
// spacing admits no boundary.  B.psi is left null: the caller points it at the device copy.
bool build_boundary(const float lo[3], const float hi[3], float s, int layers, const SphParams& P,
                    float restDensity, BoundaryDev& B, std::vector<float>& psi, std::string& err);
// Position of sample idx.
void boundary_position(const BoundaryDev& B, uint32_t idx, float out[3]);

}  // namespace akua
