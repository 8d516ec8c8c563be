// boundary.cpp -- host construction of the boundary samples (layout and psi; see boundary.h).
#include "boundary.h"

#include <algorithm>

namespace akua {

namespace {

// poly6 with the constants of SphParams, in the operation order of the device kernel.
float poly6_h(float d2, const SphParams& P) {
    if (d2 > P.h2) return 0.0f;
    const float t = P.h2 - d2;
    return P.cp * ((t * t) * t);
}

}  // namespace

bool build_boundary(const float lo[3], const float hi[3], float s, int layers, const SphParams& P,
                    float restDensity, BoundaryDev& B, std::vector<float>& psi, std::string& err) {
    B = BoundaryDev{};
    if (!(s > 0.0f) || !std::isfinite(s)) {
        err = "boundary spacing must be positive";
        return false;
    }
    if (layers < 1 || layers > kMaxBoundaryLayers) {
        err = "boundary layers must be 1..4";
        return false;
    }
    B.layers = layers;
    B.s = s;
    int n[3];
    for (int a = 0; a < 3; ++a) {
        B.lo[a] = lo[a];
        B.hi[a] = hi[a];
        const float E = hi[a] - lo[a];
        if (!(E > 0.0f) || !std::isfinite(E)) {
            err = "the box must have a positive extent on every axis for boundary samples";
            return false;
        }
        const float q = std::ceil(E / s);
        if (!(q < (float)(1 << 20))) {
            err = "boundary spacing too small for the box";
            return false;
        }
        n[a] = std::max(2, (int)q + 1);
        B.d[a] = E / (float)(n[a] - 1);
    }
    uint64_t total = 0;
    for (int w = 0; w < 6; ++w) {
        const int a = w >> 1;
        const int e2[2] = {a == 0 ? 1 : 0, a == 2 ? 1 : 2};
        for (int k = 0; k < 2; ++k) {
            const int e = e2[k];
            if (e > a) {
                B.t0[w][k] = -(layers - 1);
                B.tn[w][k] = n[e] + 2 * (layers - 1);
            } else {
                B.t0[w][k] = 1;
                B.tn[w][k] = std::max(0, n[e] - 2);
            }
        }
        B.base[w] = (uint32_t)total;
        total += (uint64_t)layers * (uint64_t)B.tn[w][0] * (uint64_t)B.tn[w][1];
        if (total > kMaxBoundarySamples) {
            err = "too many boundary samples (more than 2^26): use a larger spacing";
            return false;
        }
    }
    B.base[6] = (uint32_t)total;
    psi.assign((size_t)total, 0.0f);
    for (uint32_t k = 0; k < (uint32_t)total; ++k) {
        float x[3];
        boundary_position(B, k, x);
        float delta = 0.0f;
        boundary_visit(B, x[0], x[1], x[2], P.h, P.h2,
                       [&](uint32_t, float, float, float, float d2) { delta = delta + poly6_h(d2, P); });
        psi[k] = restDensity / delta;
    }
    B.on = 1;
    return true;
}

void boundary_position(const BoundaryDev& B, uint32_t idx, float out[3]) {
    int w = 0;
    while (w < 5 && idx >= B.base[w + 1]) ++w;
    const int a = w >> 1;
    const int b = a == 0 ? 1 : 0;
    const int c = a == 2 ? 1 : 2;
    uint32_t r = idx - B.base[w];
    const uint32_t nc = (uint32_t)B.tn[w][1], nb = (uint32_t)B.tn[w][0];
    const int tc = (int)(r % nc) + B.t0[w][1];
    r /= nc;
    const int tb = (int)(r % nb) + B.t0[w][0];
    const int l = (int)(r / nb);
    out[a] = wall_plane(B, w, l);
    out[b] = B.lo[b] + (float)tb * B.d[b];
    out[c] = B.lo[c] + (float)tc * B.d[c];
}

}  // namespace akua
