// pbf_types.h -- structures shared by the host solver and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

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
    uint32_t layoutBits;   // bits per axis of the super-cell Morton key (memory layout only)
    uint32_t modMagic;     // x % tableSize by multiply-high (fast_mod below)
    uint32_t modShift;
    int32_t layoutOrigin[4];  // super-cell coordinates mapped to Morton coordinate 0
    float bmin[4];         // box (xyz) for the wall clamp and the velocity damping
    float bmax[4];
};

// Boundary samples on the six box walls (opt-in; no reference counterpart).  Their layout,
// volume weights and the order in which a particle visits them are defined in boundary.h.
struct BoundaryDev {
    int32_t on;          // 0: no samples (the reference's clamp and damping only)
    int32_t layers;      // L, 1..4
    float s;             // layer spacing
    float lo[3], hi[3];  // the box the samples were built for
    float d[3];          // tangential spacing per axis
    int32_t t0[6][2];    // first tangential index per wall and tangential axis (b < c)
    int32_t tn[6][2];    // tangential index counts
    uint32_t base[7];    // first sample index of each wall; base[6] = number of samples
    const float* psi;    // volume weight of each sample (device memory)
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
    int32_t corrOn;    // LambdaCorrParams::enabled: 0 drops the artificial pressure
    float corrN;       // n of the artificial pressure
    int32_t corrNInt;  // n as an integer in [0, 16], or -1 when n is not such an integer
    float cellHash;    // cell size used to hash (smoothRadius, NeighbourSearchCUDA.cu:163)
    float cellSearch;  // cell size used by the stencil search (spatialHashCellSize)
    // 1: the reference's search (cells as above, the 27-cell stencil).  2: the non-parity search
    // over cells of h / 2 (cellHash = cellSearch = h / 2) and the 5x5x5 cells that come within h
    // of the particle; same neighbour set, another list order (results equal the reference's up
    // to the rounding of the reordered sums).
    int32_t searchCells;
    float vortEps;
    float visc;
    int32_t maxN;
    // 1 when every particle's mass is exactly 1 and the quotients of the spiky gradient and of
    // the artificial pressure need no upper-bound or divisor range test (their bounds, from h,
    // the kernel constants and delta_q, lie inside the exact-division range): the neighbour
    // sweeps then skip the multiplications by a neighbour's mass (identities) and those tests.
    int32_t fast;
    BoundaryDev bd;    // boundary samples (bd.on == 0: none)
};

// Per-step scalars passed by value to the kernels.
struct StepParams {
    float dt;
    float g[3];
};

// x % d for 32-bit unsigned x and a divisor d fixed per scene, by the round-up multiply-high
// method (Granlund & Montgomery 1994; Hacker's Delight, 2nd ed., 10-8): with l = ceil(log2 d),
// magic = floor(2^32 (2^l - d) / d) + 1, t = mulhi(magic, x), the quotient is
// (t + ((x - t) >> 1)) >> (l - 1) for every x (d >= 2); d == 1 uses shift 0 and magic 0 as a flag.
struct FastMod {
    uint32_t d, magic, shift;
};

inline FastMod make_fast_mod(uint32_t d) {
    FastMod f{d, 0u, 0u};
    if (d <= 1) return f;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.magic = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    f.shift = l - 1;
    return f;
}

__host__ __device__ inline uint32_t fast_mod(uint32_t x, uint32_t d, uint32_t magic,
                                             uint32_t shift) {
    if (d <= 1) return 0u;
    const uint32_t t = (uint32_t)(((uint64_t)magic * x) >> 32);
    const uint32_t q = (t + ((x - t) >> 1)) >> shift;
    return x - q * d;
}

constexpr uint32_t kSlotAlign = 256;
// Most slots of a solver: per-slot arrays of up to 16-byte elements are then addressed by 32-bit
// byte offsets in the kernels.
constexpr uint32_t kMaxSlots = 1u << 28;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;  // never a hash: hashes are < tableSize <= 2^32 - 1

}  // namespace akua
