// aos.h -- the reference's 108-byte particle record and its device conversion kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace akua {

// Same layout as akua_particle (include/akua.h) and AkuaEngine::Particle (Particle.h:8-31).
struct AosParticle {
    float position[3];
    float velocity[3];
    float new_position[3];
    float new_velocity[3];
    float position_delta[3];
    float vorticity[3];
    float mass;
    float density;
    float lambda;
    uint32_t hash;
    float color[4];
    float size;
};
static_assert(sizeof(AosParticle) == 108, "particle record must be 108 bytes");

void launch_import_aos(hipStream_t st, const AosParticle* in, uint32_t n, float4* X, float4* V,
                       float4* VS, float4* DP, float* dens, float* lam, uint32_t* id, uint32_t* lid,
                       uint32_t* phys, float4* color, float* size);
void launch_export_aos(hipStream_t st, AosParticle* out, uint32_t n, const float4* X,
                       const float4* V, const float4* VS, const float4* DP, const float* dens,
                       const float* lam, const uint32_t* id, const uint32_t* lid, const float4* W,
                       const uint32_t* hash, const float4* color, const float* size);

}  // namespace akua
