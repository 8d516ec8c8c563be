// radix_sort.h -- stable LSD radix sort of (uint32 key, uint32 value) pairs on gfx950.
//
// Replaces the reference's thrust::sort of the 108-byte particle records by hash
// (NeighbourSearchCUDA.cu:166-170) with a sort of (hash, slot) pairs; the particle data is
// then gathered once by the resulting permutation.  LSD radix sort is stable, so equal hashes
// keep their current order, which is the order the CPU oracle fixes.
//
// Keys live in particle slots; the slots of each scene form one segment and the sort never
// moves a key across segments (the scan runs scene-major, so the scene acts as the most
// significant digit without costing a pass).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace akua {

constexpr int kSortBlock = 256;   // threads per tile (4 wave64)
constexpr int kSortItems = 16;    // keys per thread
constexpr int kSortTile = kSortBlock * kSortItems;

struct SortTile {
    uint32_t begin;       // first slot of the tile
    uint32_t end;         // one past the last slot
    uint32_t countBase;   // tiles of all earlier scenes; counts are [scene][digit][tile]
    uint32_t sceneTiles;  // number of tiles of the scene
    uint32_t localTile;   // index of this tile within its scene
    int32_t slotShift;    // slot of the scene's first particle minus its rank among all particles
    uint32_t pad[2];
};

// Tiles for scenes given as [begin, end) slot ranges.  Returns the number of counts per digit
// (the sum of the scenes' tile counts).
uint32_t build_sort_tiles(const std::vector<uint32_t>& sceneBegin,
                          const std::vector<uint32_t>& sceneEnd, std::vector<SortTile>& tiles);

struct SortPlan {
    int bits = 32;        // significant key bits
    int digitBits = 8;    // bits per pass (4..8)
    int passes() const { return (bits + digitBits - 1) / digitBits; }
};

// Sorts (keys, identity values) held in keysA; values are generated as the slot index in the
// first pass.  Uses keysB/valsB as the second buffer.  counts must hold
// (1 << digitBits) * numTiles words.  Returns 0 if the result is in (keysA, valsA), 1 if in
// (keysB, valsB).  limit: slots in the key buffers.  hook (optional) is called after each kernel launch with its name.
using LaunchHook = void (*)(void* ctx, const char* name, int phase);
int radix_sort_pairs(hipStream_t stream, const SortPlan& plan, uint32_t* keysA, uint32_t* valsA,
                     uint32_t* keysB, uint32_t* valsB, const SortTile* dTiles, uint32_t numTiles,
                     uint32_t countsPerDigit, uint32_t* dCounts, uint32_t limit, LaunchHook hook,
                     void* hookCtx);

}  // namespace akua
