// pbf_kernels.h -- host launchers of the PBF step kernels (pbf_kernels.hip).
//
// Data layout in HBM (structure of arrays, slots in the current sorted order):
//   X    float4 (x, y, z, mass)   positions after the last step        16 B
//   V    float4 (vx, vy, vz, 0)   velocities after the last step       16 B
//   dens, lam  float              density and lambda of the last step   4 B each
//   id   uint32                   original particle index               4 B
// Per-step scratch: X0 (old positions, new order), XsA/XsB (x* ping-pong), W (vorticity and its
// length), Vc (velocity after vorticity confinement), the sort keys/values, the cell table, and
// the neighbour list.  The list is wave-interleaved: entry k of slot i is at
//   ((i / 64) * maxN + k) * 64 + i % 64
// so the 64 lanes of a wave read entry k of their 64 lists as one 256-byte line.
#pragma once

#include <hip/hip_runtime.h>

#include "pbf_types.h"

namespace akua {

struct DeviceState {
    uint32_t nslots;           // particle slots (all scenes, with alignment padding)
    const SceneDev* scenes;
    const uint32_t* chunkScene;  // scene of each 256-slot chunk
    float4* X;
    float4* V;
    float* dens;
    float* lam;
    uint32_t* id;
    float4* X0;
    float4* XsA;
    float4* XsB;
    float4* W;
    float4* Vc;
    float* Tdens;
    float* Tlam;
    uint32_t* Tid;
    uint32_t* keys;        // hashes in the pre-sort order
    const uint32_t* skeys; // sorted hashes
    const uint32_t* perm;  // sorted slot -> pre-sort slot
    uint32_t* tkey;        // cell table keys (all scenes)
    uint2* trange;         // cell table (start, end)
    uint32_t tableTotal;
    uint32_t* nbr;
    uint32_t* cnt;
    unsigned long long* nbrSum;  // running sum of neighbour counts (statistics)
};

constexpr int kBlock = 256;

void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t n, uint32_t v);
void launch_predict_hash(hipStream_t st, const DeviceState& d, const SphParams& P,
                         const StepParams& S);
void launch_gather(hipStream_t st, const DeviceState& d, const SphParams& P, const StepParams& S);
void launch_neighbours(hipStream_t st, const DeviceState& d, const SphParams& P);
void launch_density_lambda(hipStream_t st, const DeviceState& d, const SphParams& P,
                           const float4* Xs);
// Position correction with the wall clamp.  final == true also runs the velocity update and the
// wall damping and writes X, V and id (the last iteration); applyDelta == false skips the
// correction (a step with zero iterations).
void launch_delta_correct(hipStream_t st, const DeviceState& d, const SphParams& P,
                          const StepParams& S, const float4* XsIn, float4* XsOut, bool final,
                          bool applyDelta);
void launch_vorticity(hipStream_t st, const DeviceState& d, const SphParams& P);
void launch_confinement(hipStream_t st, const DeviceState& d, const SphParams& P,
                        const StepParams& S);
void launch_xsph(hipStream_t st, const DeviceState& d, const SphParams& P);
// Exact SPH density at the positions held in XsA (sorted order), per-scene error max and sum.
void launch_density_probe(hipStream_t st, const DeviceState& d, const SphParams& P, float* rho,
                          unsigned int* errMaxBits, double* errSum);

}  // namespace akua
