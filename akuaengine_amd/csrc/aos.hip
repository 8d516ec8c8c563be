// aos.hip -- conversion between the reference's 108-byte particle records (Particle.h:8-31) and
// the solver's structure-of-arrays state, on the device (for akua_solver_step_device).
#include "aos.h"

namespace akua {

namespace {

// Records in buffer order become both the logical and the physical order (lid = phys = identity).
__global__ void k_import_aos(const AosParticle* __restrict__ in, uint32_t n, float4* X, float4* V,
                             float4* VS, float4* DP, float* dens, float* lam, uint32_t* id,
                             uint32_t* lid, uint32_t* phys, float4* color, float* size) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const AosParticle& p = in[i];
    X[i] = make_float4(p.position[0], p.position[1], p.position[2], p.mass);
    V[i] = make_float4(p.velocity[0], p.velocity[1], p.velocity[2], 0.0f);
    VS[i] = make_float4(p.new_velocity[0], p.new_velocity[1], p.new_velocity[2], 0.0f);
    DP[i] = make_float4(p.position_delta[0], p.position_delta[1], p.position_delta[2], 0.0f);
    dens[i] = p.density;
    lam[i] = p.lambda;
    id[i] = i;
    lid[i] = i;
    phys[i] = i;
    color[i] = make_float4(p.color[0], p.color[1], p.color[2], p.color[3]);
    size[i] = p.size;
}

// Physical slot i is written to its logical position, the order of the reference's buffer.
// new_velocity is the step's velocity before the wall damping, vorticity and XSPH
// (IntegrationCUDA.cu:45), position_delta the last iteration's delta (ConstraintSolverCUDA.cu:129).
__global__ void k_export_aos(AosParticle* __restrict__ out, uint32_t n, const float4* X,
                             const float4* V, const float4* VS, const float4* DP,
                             const float* dens, const float* lam, const uint32_t* id,
                             const uint32_t* lid, const float4* W, const uint32_t* hash,
                             const float4* color, const float* size) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    AosParticle p;
    float4 x = X[i], v = V[i], w = W[i], c = color[id[i]], vs = VS[i], dp = DP[i];
    p.position[0] = x.x; p.position[1] = x.y; p.position[2] = x.z;
    p.velocity[0] = v.x; p.velocity[1] = v.y; p.velocity[2] = v.z;
    p.new_position[0] = x.x; p.new_position[1] = x.y; p.new_position[2] = x.z;
    p.new_velocity[0] = vs.x; p.new_velocity[1] = vs.y; p.new_velocity[2] = vs.z;
    p.position_delta[0] = dp.x; p.position_delta[1] = dp.y; p.position_delta[2] = dp.z;
    p.vorticity[0] = w.x; p.vorticity[1] = w.y; p.vorticity[2] = w.z;
    p.mass = x.w;
    p.density = dens[i];
    p.lambda = lam[i];
    const uint32_t k = lid[i];
    p.hash = hash[k];
    p.color[0] = c.x; p.color[1] = c.y; p.color[2] = c.z; p.color[3] = c.w;
    p.size = size[id[i]];
    out[k] = p;
}

}  // namespace

void launch_import_aos(hipStream_t st, const AosParticle* in, uint32_t n, float4* X, float4* V,
                       float4* VS, float4* DP, float* dens, float* lam, uint32_t* id, uint32_t* lid,
                       uint32_t* phys, float4* color, float* size) {
    if (n)
        k_import_aos<<<(n + 255) / 256, 256, 0, st>>>(in, n, X, V, VS, DP, dens, lam, id, lid, phys,
                                                      color, size);
}

void launch_export_aos(hipStream_t st, AosParticle* out, uint32_t n, const float4* X,
                       const float4* V, const float4* VS, const float4* DP, const float* dens,
                       const float* lam, const uint32_t* id, const uint32_t* lid, const float4* W,
                       const uint32_t* hash, const float4* color, const float* size) {
    if (n)
        k_export_aos<<<(n + 255) / 256, 256, 0, st>>>(out, n, X, V, VS, DP, dens, lam, id, lid, W,
                                                      hash, color, size);
}

}  // namespace akua
