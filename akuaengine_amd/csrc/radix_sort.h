// radix_sort.h -- stable LSD radix sort of (uint32 key, uint32 value) pairs on gfx950.
//
// Replaces the reference's thrust::sort of the 108-byte particle records by hash
// (NeighbourSearchCUDA.cu:166-170) with sorts of (key, slot) pairs; particle data is gathered
// once by the resulting permutation.  LSD radix sort is stable, so equal keys keep their input
// order, which is the order the CPU oracle fixes.
//
// Keys live in particle slots; the slots of each scene form one segment and the sort never
// moves a key across segments: offsets are computed per (scene, digit), so the scene acts as the
// most significant digit without costing a pass.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace akua {

constexpr int kSortBlock = 256;   // threads per tile (4 wave64)
constexpr int kSortItems = 8;     // keys per thread
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kMaxDigitBits = 8;
constexpr int kMaxPasses = 8;

struct SortTile {
    uint32_t begin;       // first slot of the tile
    uint32_t end;         // one past the last slot
    uint32_t countBase;   // tiles of all earlier scenes; counts are [scene][digit][tile]
    uint32_t sceneTiles;  // number of tiles of the scene
    uint32_t localTile;   // index of this tile within its scene
    uint32_t scene;
    uint32_t pad[2];
};

struct SortScene {
    uint32_t begin;       // first slot of the scene
    uint32_t countBase;   // tiles of all earlier scenes
    uint32_t tiles;       // tiles of this scene
    uint32_t pad;
};

// Tiles for scenes given as [begin, end) slot ranges.  Returns the total number of tiles.
uint32_t build_sort_tiles(const std::vector<uint32_t>& sceneBegin,
                          const std::vector<uint32_t>& sceneEnd, std::vector<SortTile>& tiles,
                          std::vector<SortScene>& scenes);

struct SortPlan {
    int bits = 32;        // significant key bits
    int digitBits = 8;    // bits per pass (4..8)
    int passes() const { return (bits + digitBits - 1) / digitBits; }
    // The fewest passes of at most 8 bits, with the bits spread evenly over them.
    static SortPlan balanced(int bits);
};

struct SortBuffers {
    uint32_t* keysA;      // input keys (clobbered)
    uint32_t* valsA;
    uint32_t* keysB;
    uint32_t* valsB;
    uint32_t* counts;     // (1 << kMaxDigitBits) * numTiles
    uint32_t* totals;     // kMaxPasses * (1 << kMaxDigitBits) * numScenes
    uint32_t* look;       // look-back words of sorts of at most kOneSweepMaxTiles tiles:
                          // 2 * (1 << kMaxDigitBits) * numTiles (NULL: such sorts use three
                          // launches per pass too)
    uint32_t* tickets;    // kMaxPasses tile tickets of those sorts
};
// Sorts of at most this many tiles use the single-pass look-back scatter.
constexpr uint32_t kOneSweepMaxTiles = 64;
// Largest slot count the look-back words can hold (30-bit counts).
constexpr uint32_t kSortMaxSlots = (1u << 30) - 1u;

// Sorts keysA with values equal to the slot index (generated in the first pass).  Returns 0 if
// the result is in (keysA, valsA), 1 if in (keysB, valsB).  limit: slots in the key buffers.
// hook (optional) is called around each kernel launch with its name (phase 0 before, 1 after).
using LaunchHook = void (*)(void* ctx, const char* name, int phase);
int radix_sort_pairs(hipStream_t stream, const SortPlan& plan, const SortBuffers& buf,
                     const SortTile* dTiles, uint32_t numTiles, const SortScene* dScenes,
                     uint32_t numScenes, uint32_t limit, LaunchHook hook, void* hookCtx);
// Two independent sorts over the same tiles and scenes (separate buffers, counts, totals and
// look-back words) with the same digit width, every pass of both in shared launches.  Returns
// false, launching nothing, if the digit widths differ; else sets each result as
// radix_sort_pairs returns it.
// scratchZeroed: the spans radix_sort_zero_spans names were zeroed since the last sort (by a
// launch of the caller's), so the sort skips its own clearing launch.
bool radix_sort_pairs2(hipStream_t stream, const SortPlan& planA, const SortBuffers& bufA,
                       const SortPlan& planB, const SortBuffers& bufB, const SortTile* dTiles,
                       uint32_t numTiles, const SortScene* dScenes, uint32_t numScenes, uint32_t limit,
                       LaunchHook hook, void* hookCtx, int& resultA, int& resultB,
                       bool scratchZeroed = false);

// Scratch that two sorts of at most kOneSweepMaxTiles tiles need zeroed before they start: per sort
// the totals of every pass, the first pass's look-back words and the tile tickets.  count is 0
// when the sorts take three launches per pass (nothing to zero).
struct SortZeroSpans {
    uint32_t* p[6];
    uint32_t n[6];
    uint32_t count;
};
SortZeroSpans radix_sort_zero_spans(const SortPlan& planA, const SortBuffers& bufA, const SortPlan& planB,
                                    const SortBuffers& bufB, uint32_t numTiles, uint32_t numScenes);

}  // namespace akua
