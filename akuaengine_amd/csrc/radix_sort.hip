// radix_sort.hip -- stable LSD radix sort of (key, slot) pairs for gfx950.
//
// Per sort, one fill of the per-(pass, scene, digit) totals; then per pass (digit of D bits,
// R = 2^D buckets) three kernels:
//   rs_hist    : per tile, the digit histogram (wave64 match-by-ballot, one LDS add per group),
//                stored as [scene][digit][tile] counts and added to the scene's digit totals.
//   rs_scan    : one wave per (scene, digit): the digit's base slot (scene begin + the totals of
//                the smaller digits) plus the exclusive scan of its tile counts.
//   rs_scatter : per tile, stable ranks by wave64 ballots (wave w owns a contiguous quarter of
//                the tile, in kSortItems rounds of 64 consecutive keys), a tile-local counting sort in
//                LDS, then coalesced runs written to the global offsets.
#include "radix_sort.h"

#include <algorithm>

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

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* ldsWave) {
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
    __syncthreads();
    return base + inc - v;
}

template <int D>
__global__ __launch_bounds__(kSortBlock) void rs_hist(const uint32_t* __restrict__ keys,
                                                      const SortTile* __restrict__ tiles,
                                                      uint32_t shift, uint32_t* __restrict__ counts) {
    constexpr uint32_t R = 1u << D;
    constexpr int kWaves = kSortBlock / 64;
    __shared__ uint32_t hist[kWaves][R];  // one histogram per wave: fewer atomic collisions
    const SortTile t = tiles[blockIdx.x];
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t d = threadIdx.x; d < R; d += kSortBlock)
        for (int w = 0; w < kWaves; ++w) hist[w][d] = 0;
    __syncthreads();
    uint32_t dg[kSortItems];
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const uint32_t idx = t.begin + it * kSortBlock + threadIdx.x;
        dg[it] = idx < t.end ? (keys[idx] >> shift) & (R - 1) : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int it = 0; it < kSortItems; ++it)
        if (dg[it] != 0xFFFFFFFFu) atomicAdd(&hist[wave][dg[it]], 1u);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < R; d += kSortBlock) {
        uint32_t c = 0;
        for (int w = 0; w < kWaves; ++w) c += hist[w][d];
        counts[t.countBase * R + d * t.sceneTiles + t.localTile] = c;
    }
}

// One wave per (scene, digit): offsets[scene][digit][tile] = the keys of this digit in earlier
// tiles of the scene, and totals[scene][digit] = the scene's keys with this digit.  The scatter
// adds the scene's first slot and the keys of smaller digits (no atomics anywhere).
template <int D>
__global__ __launch_bounds__(64) void rs_scan(uint32_t* __restrict__ counts,
                                              uint32_t* __restrict__ totals,
                                              const SortScene* __restrict__ scenes) {
    constexpr uint32_t R = 1u << D;
    constexpr uint32_t kPer = 8;  // consecutive tiles per lane: one memory round trip per 512
    const uint32_t s = blockIdx.x / R, d = blockIdx.x % R;
    const SortScene sc = scenes[s];
    const uint32_t lane = lane_id();
    uint32_t carry = 0;
    uint32_t* c = counts + (size_t)sc.countBase * R + (size_t)d * sc.tiles;
    for (uint32_t base = 0; base < sc.tiles; base += 64u * kPer) {
        const uint32_t t0 = base + lane * kPer;
        uint32_t v[kPer];
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i) v[i] = t0 + i < sc.tiles ? c[t0 + i] : 0u;
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i) sum += v[i];
        uint32_t inc = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(inc, off, 64);
            if (lane >= (uint32_t)off) inc += y;
        }
        uint32_t run = carry + inc - sum;
#pragma unroll
        for (uint32_t i = 0; i < kPer; ++i) {
            if (t0 + i < sc.tiles) c[t0 + i] = run;
            run += v[i];
        }
        carry += __shfl(inc, 63, 64);
    }
    if (lane == 0) totals[s * R + d] = carry;
}

template <int D, bool kIdentityVals>
__global__ __launch_bounds__(kSortBlock) void rs_scatter(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut,
    const SortTile* __restrict__ tiles, uint32_t shift, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ totals, uint32_t limit) {
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
        uint32_t start = block_exclusive_scan_256(d < R ? tot : 0u, scanTmp);
        // the scene's keys with a smaller digit
        const uint32_t below = block_exclusive_scan_256(d < R ? totals[t.scene * R + d] : 0u, scanTmp);
        if (d < R) {
            tileStart[d] = start;
            const uint32_t sceneBegin = t.begin - t.localTile * (uint32_t)kSortTile;
            globalBase[d] = sceneBegin + below + offsets[t.countBase * R + d * t.sceneTiles + t.localTile];
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
                 const SortScene* scenes, uint32_t nscenes, uint32_t shift, uint32_t* counts,
                 uint32_t* totals, uint32_t limit, LaunchHook hook, void* ctx) {
    if (hook) hook(ctx, "rs_hist", 0);
    rs_hist<D><<<ntiles, kSortBlock, 0, st>>>(kin, tiles, shift, counts);
    if (hook) hook(ctx, "rs_hist", 1);
    if (hook) hook(ctx, "rs_scan", 0);
    rs_scan<D><<<nscenes * (1u << D), 64, 0, st>>>(counts, totals, scenes);
    if (hook) hook(ctx, "rs_scan", 1);
    if (hook) hook(ctx, "rs_scatter", 0);
    if (first)
        rs_scatter<D, true><<<ntiles, kSortBlock, 0, st>>>(kin, vin, kout, vout, tiles, shift, counts,
                                                           totals, limit);
    else
        rs_scatter<D, false><<<ntiles, kSortBlock, 0, st>>>(kin, vin, kout, vout, tiles, shift, counts,
                                                            totals, limit);
    if (hook) hook(ctx, "rs_scatter", 1);
}

}  // namespace

SortPlan SortPlan::balanced(int bits) {
    SortPlan p;
    p.bits = bits < 1 ? 1 : bits;
    const int passes = (p.bits + kMaxDigitBits - 1) / kMaxDigitBits;
    p.digitBits = std::max(4, (p.bits + passes - 1) / passes);
    return p;
}

uint32_t build_sort_tiles(const std::vector<uint32_t>& sceneBegin,
                          const std::vector<uint32_t>& sceneEnd, std::vector<SortTile>& tiles,
                          std::vector<SortScene>& scenes) {
    tiles.clear();
    scenes.clear();
    uint32_t tileBase = 0;
    for (size_t s = 0; s < sceneBegin.size(); ++s) {
        const uint32_t n = sceneEnd[s] - sceneBegin[s];
        const uint32_t nt = (n + kSortTile - 1) / kSortTile;
        for (uint32_t k = 0; k < nt; ++k) {
            SortTile t{};
            t.begin = sceneBegin[s] + k * kSortTile;
            t.end = std::min(sceneEnd[s], t.begin + (uint32_t)kSortTile);
            t.countBase = tileBase;
            t.sceneTiles = nt;
            t.localTile = k;
            t.scene = (uint32_t)s;
            tiles.push_back(t);
        }
        scenes.push_back(SortScene{sceneBegin[s], tileBase, nt, 0u});
        tileBase += nt;
    }
    return tileBase;
}

int radix_sort_pairs(hipStream_t stream, const SortPlan& plan, const SortBuffers& buf,
                     const SortTile* dTiles, uint32_t numTiles, const SortScene* dScenes,
                     uint32_t numScenes, uint32_t limit, LaunchHook hook, void* hookCtx) {
    const int P = plan.passes();
    const uint32_t R = 1u << plan.digitBits;
    uint32_t* kin = buf.keysA;
    uint32_t* vin = buf.valsA;
    uint32_t* kout = buf.keysB;
    uint32_t* vout = buf.valsB;
    for (int p = 0; p < P; ++p) {
        const uint32_t shift = (uint32_t)(p * plan.digitBits);
        uint32_t* tot = buf.totals + (size_t)p * R * numScenes;
#define AKUA_PASS(DB)                                                                        \
    launch_pass<DB>(stream, p == 0, kin, vin, kout, vout, dTiles, numTiles, dScenes, numScenes, \
                    shift, buf.counts, tot, limit, hook, hookCtx)
        switch (plan.digitBits) {
            case 4: AKUA_PASS(4); break;
            case 5: AKUA_PASS(5); break;
            case 6: AKUA_PASS(6); break;
            case 7: AKUA_PASS(7); break;
            default: AKUA_PASS(8); break;
        }
#undef AKUA_PASS
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    return (P % 2 == 0) ? 0 : 1;
}

}  // namespace akua
