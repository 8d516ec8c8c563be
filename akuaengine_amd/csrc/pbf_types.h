// pbf_types.h -- structures shared by the host solver and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace akua {

// Per-scene constants kept in device memory.  A solver holds one scene (the PBFSolver API) or
// many independent scenes (batched stepping).  Scene s owns the particle slots [begin, end);
// begin is a multiple of kSlotAlign so that no 256-slot chunk straddles two scenes.
struct SceneDev {
    uint32_t begin;        // first particle slot
    uint32_t end;          // one past the last particle slot
    uint32_t tableSize;    // maxNeighbours * N_s: the spatial-hash modulus (PBFSolver.cpp:15)
    uint32_t tabOffset;    // first slot of this scene's cell table
    uint32_t tabLog2;      // log2 of this scene's cell-table capacity
    uint32_t pad[3];
    float bmin[4];         // box (xyz) for the wall clamp and the velocity damping
    float bmax[4];
};

// SPH and solver constants.  Every value is computed on the host with the same float expressions
// the reference evaluates inline (SmoothingKernelsCUDA.h:16-28, ConstraintSolverCUDA.cu:201,
// :123-124), so device code never re-derives a constant.
struct SphParams {
    float h;           // smoothRadius
    float h2;          // h * h
    float cp;          // 315 / (64 * 3.14 * h^9)
    float cs;          // -45 / (3.14 * h^6)
    float w0;          // poly6(0)
    float wdq;         // poly6(delta_q^2)
    float invRest;     // 1 / restDensity
    float relax;       // relaxation (epsilon)
    float negCorrK;    // -k of the artificial pressure
    float corrN;       // n of the artificial pressure
    int32_t corrNInt;  // n as an integer in [0, 16], or -1 when n is not such an integer
    float cellHash;    // cell size used to hash (smoothRadius, NeighbourSearchCUDA.cu:163)
    float cellSearch;  // cell size used by the stencil search (spatialHashCellSize)
    float vortEps;
    float visc;
    int32_t maxN;
};

// Per-step scalars passed by value to the kernels.
struct StepParams {
    float dt;
    float g[3];
};

constexpr uint32_t kSlotAlign = 256;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;  // never a hash: hashes are < tableSize <= 2^32 - 1

}  // namespace akua
