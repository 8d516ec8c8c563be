// radix_sort.hip -- stable LSD radix sort of (key, slot) pairs for gfx950.
//
// Per pass (digit of D bits, R = 2^D buckets), three kernels:
//   rs_hist    : per tile, the digit histogram (wave64 match-by-ballot, one LDS add per group).
//   rs_scan    : one workgroup turns the scene-major [scene][digit][tile] counts into exclusive
//                offsets; the scan order makes every scene's keys land inside its own segment.
//   rs_scatter : per tile, stable ranks by wave64 ballots (wave w owns a contiguous quarter of
//                the tile, in 16 rounds of 64 consecutive keys), a tile-local counting sort in
//                LDS, then coalesced runs written to the global offsets.
#include "radix_sort.h"

namespace akua {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Lanes of this wave (among `valid`) whose digit equals mine.
template <int D>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
    uint64_t peers = valid;
#pragma unroll
    for (int b = 0; b < D; ++b) {
        bool bit = (d >> b) & 1u;
        uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// Exclusive scan of one value per thread over a 256-thread block; also returns the total.
__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* ldsWave,
                                                             uint32_t& total) {
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(inc, off, 64);
        if (lane >= (uint32_t)off) inc += t;
    }
    if (lane == 63) ldsWave[wave] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += ldsWave[w];
    total = ldsWave[0] + ldsWave[1] + ldsWave[2] + ldsWave[3];
    __syncthreads();
    return base + inc - v;
}

template <int D>
__global__ __launch_bounds__(kSortBlock) void rs_hist(const uint32_t* __restrict__ keys,
                                                      const SortTile* __restrict__ tiles,
                                                      uint32_t shift, uint32_t* __restrict__ counts) {
    constexpr uint32_t R = 1u << D;
    __shared__ uint32_t hist[R];
    const SortTile t = tiles[blockIdx.x];
    for (uint32_t d = threadIdx.x; d < R; d += kSortBlock) hist[d] = 0;
    __syncthreads();
    const uint32_t lane = lane_id();
#pragma unroll 4
    for (int it = 0; it < kSortItems; ++it) {
        uint32_t idx = t.begin + it * kSortBlock + threadIdx.x;
        bool valid = idx < t.end;
        uint32_t d = valid ? (keys[idx] >> shift) & (R - 1) : 0u;
        uint64_t vmask = __ballot(valid);
        uint64_t peers = match_digit<D>(d, vmask);
        // the highest lane of each group adds the group's size
        if (valid && (peers >> lane) == 1ull) atomicAdd(&hist[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < R; d += kSortBlock)
        counts[t.countBase * R + d * t.sceneTiles + t.localTile] = hist[d];
}

// Exclusive scan of counts[0, n) in place, one 1024-thread workgroup, each thread a contiguous
// chunk.
__global__ __launch_bounds__(1024) void rs_scan(uint32_t* __restrict__ counts, uint32_t n) {
    __shared__ uint32_t waveSums[16];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (n + 1023) / 1024;
    const uint32_t b = min(n, tid * chunk), e = min(n, b + chunk);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += counts[i];
    // block exclusive scan of sum over 16 waves
    const uint32_t lane = lane_id(), wave = tid >> 6;
    uint32_t inc = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(inc, off, 64);
        if (lane >= (uint32_t)off) inc += t;
    }
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += waveSums[w];
    uint32_t run = base + inc - sum;
    for (uint32_t i = b; i < e; ++i) {
        uint32_t c = counts[i];
        counts[i] = run;
        run += c;
    }
}

template <int D, bool kIdentityVals>
__global__ __launch_bounds__(kSortBlock) void rs_scatter(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut,
    const SortTile* __restrict__ tiles, uint32_t shift, const uint32_t* __restrict__ offsets,
    uint32_t limit) {
    constexpr uint32_t R = 1u << D;
    constexpr int kWaves = kSortBlock / 64;
    constexpr int kRounds = kSortTile / kSortBlock;  // rounds of 64 keys per wave
    __shared__ uint32_t sKey[kSortTile];
    __shared__ uint32_t sVal[kSortTile];
    __shared__ uint32_t waveCnt[kWaves][R];
    __shared__ uint32_t tileStart[R];
    __shared__ uint32_t globalBase[R];
    __shared__ uint32_t scanTmp[4];

    const SortTile t = tiles[blockIdx.x];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t d = threadIdx.x; d < R; d += kSortBlock)
        for (int w = 0; w < kWaves; ++w) waveCnt[w][d] = 0;
    __syncthreads();

    // Phase 1: stable ranks.  Wave w owns slots [begin + w*64*kRounds, +64*kRounds) in order.
    uint32_t key[kRounds], val[kRounds], rankDigit[kRounds];
    const uint32_t waveBase = t.begin + wave * 64u * kRounds;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint32_t idx = waveBase + r * 64u + lane;
        bool valid = idx < t.end;
        key[r] = valid ? keysIn[idx] : 0u;
        val[r] = valid ? (kIdentityVals ? idx : valsIn[idx]) : 0u;
        uint32_t d = (key[r] >> shift) & (R - 1);
        uint64_t vmask = __ballot(valid);
        uint64_t peers = match_digit<D>(d, vmask);
        uint32_t below = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
        uint32_t before = valid ? waveCnt[wave][d] : 0u;
        // all lanes have read the counter before its group's highest lane advances it
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers >> lane) == 1ull) waveCnt[wave][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rankDigit[r] = valid ? ((before + below) << 8) | d : 0xFFFFFFFFu;
    }
    __syncthreads();

    // Phase 2: per digit, offsets of each wave and of the digit within the tile.
    {
        const uint32_t d = threadIdx.x;  // R <= 256 == kSortBlock
        uint32_t tot = 0;
        if (d < R) {
            for (int w = 0; w < kWaves; ++w) {
                uint32_t c = waveCnt[w][d];
                waveCnt[w][d] = tot;
                tot += c;
            }
        }
        uint32_t total;
        uint32_t start = block_exclusive_scan_256(d < R ? tot : 0u, scanTmp, total);
        if (d < R) {
            tileStart[d] = start;
            globalBase[d] = offsets[t.countBase * R + d * t.sceneTiles + t.localTile] + t.slotShift;
        }
    }
    __syncthreads();

    // Phase 3: tile-local counting sort into LDS.
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint32_t rd = rankDigit[r];
        if (rd != 0xFFFFFFFFu) {
            uint32_t d = rd & 0xFFu, rank = rd >> 8;
            uint32_t pos = tileStart[d] + waveCnt[wave][d] + rank;
            sKey[pos] = key[r];
            sVal[pos] = val[r];
        }
    }
    __syncthreads();

    // Phase 4: runs of equal digits go to consecutive global slots.
    const uint32_t n = t.end - t.begin;
    for (uint32_t p = threadIdx.x; p < n; p += kSortBlock) {
        uint32_t k = sKey[p];
        uint32_t d = (k >> shift) & (R - 1);
        uint32_t g = globalBase[d] + (p - tileStart[d]);
        if (g < limit) {  // always true for a consistent scan; guards the write regardless
            keysOut[g] = k;
            valsOut[g] = sVal[p];
        }
    }
}

template <int D>
void launch_pass(hipStream_t st, bool first, const uint32_t* kin, const uint32_t* vin,
                 uint32_t* kout, uint32_t* vout, const SortTile* tiles, uint32_t ntiles,
                 uint32_t shift, uint32_t countsPerDigit, uint32_t* counts, uint32_t limit,
                 LaunchHook hook, void* ctx) {
    if (hook) hook(ctx, "rs_hist", 0);
    rs_hist<D><<<ntiles, kSortBlock, 0, st>>>(kin, tiles, shift, counts);
    if (hook) hook(ctx, "rs_hist", 1);
    const uint32_t n = countsPerDigit * (1u << D);
    if (hook) hook(ctx, "rs_scan", 0);
    rs_scan<<<1, 1024, 0, st>>>(counts, n);
    if (hook) hook(ctx, "rs_scan", 1);
    if (hook) hook(ctx, "rs_scatter", 0);
    if (first)
        rs_scatter<D, true><<<ntiles, kSortBlock, 0, st>>>(kin, vin, kout, vout, tiles, shift, counts, limit);
    else
        rs_scatter<D, false><<<ntiles, kSortBlock, 0, st>>>(kin, vin, kout, vout, tiles, shift, counts, limit);
    if (hook) hook(ctx, "rs_scatter", 1);
}

}  // namespace

uint32_t build_sort_tiles(const std::vector<uint32_t>& sceneBegin,
                          const std::vector<uint32_t>& sceneEnd, std::vector<SortTile>& tiles) {
    tiles.clear();
    // First pass: tiles per scene, to know each scene's count base.
    std::vector<uint32_t> nt(sceneBegin.size());
    uint32_t total = 0;
    for (size_t s = 0; s < sceneBegin.size(); ++s) {
        uint32_t n = sceneEnd[s] - sceneBegin[s];
        nt[s] = (n + kSortTile - 1) / kSortTile;
        total += nt[s];
    }
    uint32_t tileBase = 0;  // tiles of all earlier scenes
    uint32_t rankBase = 0;  // particles of all earlier scenes
    for (size_t s = 0; s < sceneBegin.size(); ++s) {
        for (uint32_t k = 0; k < nt[s]; ++k) {
            SortTile t{};
            t.begin = sceneBegin[s] + k * kSortTile;
            t.end = std::min(sceneEnd[s], t.begin + (uint32_t)kSortTile);
            t.countBase = tileBase;
            t.sceneTiles = nt[s];
            t.localTile = k;
            t.slotShift = (int32_t)sceneBegin[s] - (int32_t)rankBase;
            tiles.push_back(t);
        }
        tileBase += nt[s];
        rankBase += sceneEnd[s] - sceneBegin[s];
    }
    return total;
}

int radix_sort_pairs(hipStream_t stream, const SortPlan& plan, uint32_t* keysA, uint32_t* valsA,
                     uint32_t* keysB, uint32_t* valsB, const SortTile* dTiles, uint32_t numTiles,
                     uint32_t countsPerDigit, uint32_t* dCounts, uint32_t limit, LaunchHook hook,
                     void* hookCtx) {
    const int P = plan.passes();
    uint32_t* kin = keysA;
    uint32_t* vin = valsA;
    uint32_t* kout = keysB;
    uint32_t* vout = valsB;
    for (int p = 0; p < P; ++p) {
        uint32_t shift = (uint32_t)(p * plan.digitBits);
        switch (plan.digitBits) {
            case 4: launch_pass<4>(stream, p == 0, kin, vin, kout, vout, dTiles, numTiles, shift, countsPerDigit, dCounts, limit, hook, hookCtx); break;
            case 5: launch_pass<5>(stream, p == 0, kin, vin, kout, vout, dTiles, numTiles, shift, countsPerDigit, dCounts, limit, hook, hookCtx); break;
            case 6: launch_pass<6>(stream, p == 0, kin, vin, kout, vout, dTiles, numTiles, shift, countsPerDigit, dCounts, limit, hook, hookCtx); break;
            case 7: launch_pass<7>(stream, p == 0, kin, vin, kout, vout, dTiles, numTiles, shift, countsPerDigit, dCounts, limit, hook, hookCtx); break;
            default: launch_pass<8>(stream, p == 0, kin, vin, kout, vout, dTiles, numTiles, shift, countsPerDigit, dCounts, limit, hook, hookCtx); break;
        }
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    return (P % 2 == 0) ? 0 : 1;
}

}  // namespace akua
