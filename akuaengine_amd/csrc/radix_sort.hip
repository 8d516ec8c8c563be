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
// Sorts of at most kOneSweepMaxTiles tiles (small scenes, where the three launches per pass cost
// more than the work) run instead: rs_init and rs_hist_all once per sort (the digit totals of every
// pass), then per pass one rs_scatter that finds each tile's offsets by decoupled look-back over
// the earlier tiles' published counts.  Both give the same stable order.
// Two sorts over the same tiles with the same digit width (radix_sort_pairs2) share every launch:
// each kernel's grid holds both sorts' blocks, the second sort's after the first's, and a pass of
// only one sort (the longer one's last) runs alone.
#include "radix_sort.h"

#include <algorithm>

namespace akua {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// One sort's buffers in one pass.
struct SortIO {
    const uint32_t* keysIn;
    const uint32_t* valsIn;
    uint32_t* keysOut;
    uint32_t* valsOut;
    uint32_t* counts;    // [scene][digit][tile] counts, then offsets (three-launch passes)
    uint32_t* totals;    // this pass's [scene][digit] totals
    uint32_t* look;      // look-back words of this pass, and of the next (cleared here)
    uint32_t* lookNext;
    uint32_t* ticket;    // this pass's tile ticket
};

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

// Blocks [0, ntiles) count sort a's tiles, blocks [ntiles, 2 ntiles) sort b's.
template <int D>
__global__ __launch_bounds__(kSortBlock) void rs_hist(SortIO a, SortIO b, const SortTile* __restrict__ tiles,
                                                      uint32_t ntiles, uint32_t shift) {
    constexpr uint32_t R = 1u << D;
    constexpr int kWaves = kSortBlock / 64;
    __shared__ uint32_t hist[kWaves][R];  // one histogram per wave: fewer atomic collisions
    const bool second = blockIdx.x >= ntiles;
    const uint32_t* __restrict__ keys = second ? b.keysIn : a.keysIn;
    uint32_t* __restrict__ counts = second ? b.counts : a.counts;
    const SortTile t = tiles[blockIdx.x - (second ? ntiles : 0u)];
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
// Blocks [0, nscenes R) scan sort a's counts, the next nscenes R sort b's.
template <int D>
__global__ __launch_bounds__(64) void rs_scan(SortIO a, SortIO b, const SortScene* __restrict__ scenes,
                                              uint32_t nscenes) {
    constexpr uint32_t R = 1u << D;
    constexpr uint32_t kPer = 8;  // consecutive tiles per lane: one memory round trip per 512
    const bool second = blockIdx.x >= nscenes * R;
    uint32_t* __restrict__ counts = second ? b.counts : a.counts;
    uint32_t* __restrict__ totals = second ? b.totals : a.totals;
    const uint32_t bi = blockIdx.x - (second ? nscenes * R : 0u);
    const uint32_t s = bi / R, d = bi % R;
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

// ---- single-pass scatter with decoupled look-back (small sorts) ----
// Look-back words: the top two bits say what the low 30 hold.
constexpr uint32_t kLookAggregate = 1u << 30;  // the tile's own count of the digit
constexpr uint32_t kLookPrefix = 2u << 30;     // the count in this and every earlier tile of the scene
constexpr uint32_t kLookValue = (1u << 30) - 1u;
// Bound on the unanswered polls of one thread's whole look-back (each followed by a short sleep,
// together well under a second).  A consistent launch never comes near it: every awaited tile has
// started, and publishes its count before it waits on anything.  It only guarantees that the
// wait ends.
constexpr uint32_t kLookSpinLimit = 1u << 20;
// Earlier tiles' words read per round of the look-back.
constexpr uint32_t kLookBatch = 16;

// Clears the digit totals of every pass, the first pass's look-back words and the tile tickets
// of sort a, and of sort b when `both`.
__global__ __launch_bounds__(kSortBlock) void rs_init(SortIO a, uint32_t nTotalsA, SortIO b, uint32_t nTotalsB,
                                                      uint32_t nLook, bool both) {
    const uint32_t i = blockIdx.x * kSortBlock + threadIdx.x;
    const uint32_t stride = gridDim.x * kSortBlock;
    for (int side = 0; side < (both ? 2 : 1); ++side) {
        const SortIO& io = side ? b : a;
        const uint32_t nTotals = side ? nTotalsB : nTotalsA;
        for (uint32_t k = i; k < nTotals; k += stride) io.totals[k] = 0u;
        for (uint32_t k = i; k < nLook; k += stride) io.look[k] = 0u;
        if (i < (uint32_t)kMaxPasses) io.ticket[i] = 0u;
    }
}

// Per tile, the digit histogram of every pass at once (the counts do not depend on the order),
// added to totals[pass][scene][digit].
// Blocks [0, ntiles) count sort a's tiles over passesA passes, the next ntiles sort b's over
// passesB (totals: every pass's, from the SortIO's totals).
template <int D>
__global__ __launch_bounds__(kSortBlock) void rs_hist_all(SortIO a, uint32_t passesA, SortIO b, uint32_t passesB,
                                                          const SortTile* __restrict__ tiles, uint32_t ntiles,
                                                          uint32_t numScenes) {
    constexpr uint32_t R = 1u << D;
    __shared__ uint32_t hist[kMaxPasses * R];
    const bool second = blockIdx.x >= ntiles;
    const uint32_t* __restrict__ keys = second ? b.keysIn : a.keysIn;
    uint32_t* __restrict__ totals = second ? b.totals : a.totals;
    const uint32_t passes = second ? passesB : passesA;
    const SortTile t = tiles[blockIdx.x - (second ? ntiles : 0u)];
    for (uint32_t k = threadIdx.x; k < passes * R; k += kSortBlock) hist[k] = 0u;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const uint32_t idx = t.begin + it * kSortBlock + threadIdx.x;
        if (idx < t.end) {
            const uint32_t key = keys[idx];
            for (uint32_t p = 0; p < passes; ++p) atomicAdd(&hist[p * R + ((key >> (p * D)) & (R - 1))], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < passes * R; k += kSortBlock) {
        const uint32_t c = hist[k];
        if (c) atomicAdd(&totals[((size_t)(k / R) * numScenes + t.scene) * R + k % R], c);
    }
}

__device__ __forceinline__ uint32_t look_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void look_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The count of digit d in the scene's tiles before tile ti (the scene's first tile is
// ti - localTile), from their look-back words, kLookBatch words loaded at a time, adding counts until
// a prefix word; then this tile's own prefix word.
__device__ __forceinline__ uint32_t look_back(uint32_t* look, uint32_t R, uint32_t ti, uint32_t localTile,
                                              uint32_t d, uint32_t mine) {
    uint32_t before = 0;
    if (localTile > 0) {
        const uint32_t first = ti - localTile;
        uint32_t spin = 0;
        uint32_t j = ti;  // tiles [j, ti) are accounted for
        bool done = false;
        while (!done) {
            const uint32_t m = min(kLookBatch, j - first);
            uint32_t v[kLookBatch];
#pragma unroll
            for (uint32_t u = 0; u < kLookBatch; ++u)
                v[u] = u < m ? look_load(look + (size_t)(j - 1u - u) * R + d) : kLookPrefix;
#pragma unroll
            for (uint32_t u = 0; u < kLookBatch; ++u) {
                if (done || u >= m) continue;
                while (!(v[u] & (kLookAggregate | kLookPrefix)) && spin < kLookSpinLimit) {
                    ++spin;
                    __builtin_amdgcn_s_sleep(1);
                    v[u] = look_load(look + (size_t)(j - 1u - u) * R + d);
                }
                before += v[u] & kLookValue;
                done = (v[u] & kLookPrefix) != 0u;
            }
            j -= m;
            done = done || j == first;
        }
        look_store(look + (size_t)ti * R + d, kLookPrefix | (before + mine));
    }
    return before;
}

// kLookBack: tiles are taken in ticket order (look-back words in `look`, the next pass's cleared in
// `lookNext`, the counts unused); otherwise tile blockIdx.x with its offsets from rs_scan.  Blocks
// [0, ntiles) move sort a's keys, blocks [ntiles, 2 ntiles) sort b's, each taking its own sort's
// tickets.
template <int D, bool kIdentityVals, bool kLookBack>
__global__ __launch_bounds__(kSortBlock) void rs_scatter(SortIO a, SortIO b, const SortTile* __restrict__ tiles,
                                                         uint32_t shift, uint32_t limit, uint32_t ntiles) {
    const bool second = blockIdx.x >= ntiles;
    const uint32_t* __restrict__ keysIn = second ? b.keysIn : a.keysIn;
    const uint32_t* __restrict__ valsIn = second ? b.valsIn : a.valsIn;
    uint32_t* __restrict__ keysOut = second ? b.keysOut : a.keysOut;
    uint32_t* __restrict__ valsOut = second ? b.valsOut : a.valsOut;
    const uint32_t* __restrict__ offsets = second ? b.counts : a.counts;
    const uint32_t* __restrict__ totals = second ? b.totals : a.totals;
    uint32_t* look = second ? b.look : a.look;
    uint32_t* __restrict__ lookNext = second ? b.lookNext : a.lookNext;
    uint32_t* ticket = second ? b.ticket : a.ticket;
    constexpr uint32_t R = 1u << D;
    constexpr int kWaves = kSortBlock / 64;
    constexpr int kRounds = kSortTile / kSortBlock;  // rounds of 64 keys per wave
    __shared__ uint32_t sKey[kSortTile];
    __shared__ uint32_t sVal[kSortTile];
    __shared__ uint32_t waveCnt[kWaves][R];
    __shared__ uint32_t tileStart[R];
    __shared__ uint32_t globalBase[R];
    __shared__ uint32_t scanTmp[4];
    __shared__ uint32_t sTicket;

    if (kLookBack && threadIdx.x == 0) sTicket = atomicAdd(ticket, 1u);
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t d = threadIdx.x; d < R; d += kSortBlock)
        for (int w = 0; w < kWaves; ++w) waveCnt[w][d] = 0;
    __syncthreads();
    const uint32_t ti = kLookBack ? sTicket : blockIdx.x - (second ? ntiles : 0u);
    if (ti >= ntiles) return;  // one ticket per workgroup of an ntiles grid: never taken
    const SortTile t = tiles[ti];
    if (kLookBack)
        for (uint32_t d = threadIdx.x; d < R; d += kSortBlock) lookNext[(size_t)ti * R + d] = 0u;

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
            // publish this tile's count before waiting on any other tile
            if (kLookBack)
                look_store(look + (size_t)ti * R + d, (t.localTile == 0 ? kLookPrefix : kLookAggregate) | tot);
        }
        uint32_t start = block_exclusive_scan_256(d < R ? tot : 0u, scanTmp);
        // the scene's keys with a smaller digit
        const uint32_t below = block_exclusive_scan_256(d < R ? totals[t.scene * R + d] : 0u, scanTmp);
        if (d < R) {
            tileStart[d] = start;
            const uint32_t sceneBegin = t.begin - t.localTile * (uint32_t)kSortTile;
            const uint32_t earlier = kLookBack ? look_back(look, R, ti, t.localTile, d, tot)
                                               : offsets[t.countBase * R + d * t.sceneTiles + t.localTile];
            globalBase[d] = sceneBegin + below + earlier;
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

// Sort `buf`'s buffers for pass p: the keys are read from (keysA, valsA) in even passes.
SortIO pass_io(const SortBuffers& buf, int p, uint32_t R, uint32_t nscenes, uint32_t ntiles, bool oneSweep) {
    const bool even = (p % 2) == 0;
    SortIO io{};
    io.keysIn = even ? buf.keysA : buf.keysB;
    io.valsIn = even ? buf.valsA : buf.valsB;
    io.keysOut = even ? buf.keysB : buf.keysA;
    io.valsOut = even ? buf.valsB : buf.valsA;
    io.counts = buf.counts;
    io.totals = buf.totals + (size_t)p * nscenes * R;
    if (oneSweep) {  // look-back words alternate between the two halves of buf.look
        io.look = buf.look + (size_t)(p & 1) * ntiles * R;
        io.lookNext = buf.look + (size_t)((p + 1) & 1) * ntiles * R;
        io.ticket = buf.tickets + p;
    }
    return io;
}

// Sort A (PA passes) and, when B is given, sort B (PB passes) over the same tiles, every pass of
// both in shared launches.  oneSweep: sorts of at most kOneSweepMaxTiles tiles, the totals of every
// pass at once, then one look-back scatter per pass; otherwise three launches per pass.
template <int D>
void run_sorts(hipStream_t st, const SortBuffers& A, int PA, const SortBuffers* B, int PB,
               const SortTile* tiles, uint32_t ntiles, const SortScene* scenes, uint32_t nscenes,
               uint32_t limit, bool oneSweep, bool zeroed, LaunchHook hook, void* ctx) {
    constexpr uint32_t R = 1u << D;
    const int P = std::max(PA, B ? PB : 0);
    if (oneSweep) {
        const SortIO a0 = pass_io(A, 0, R, nscenes, ntiles, true);
        const SortIO b0 = B ? pass_io(*B, 0, R, nscenes, ntiles, true) : a0;
        const uint32_t nTotA = (uint32_t)PA * nscenes * R, nTotB = B ? (uint32_t)PB * nscenes * R : 0u;
        const uint32_t nLook = ntiles * R;
        const uint32_t most = std::max(std::max(nTotA, nTotB), nLook);
        if (hook) hook(ctx, "rs_hist", 0);
        if (!zeroed)
            rs_init<<<std::min<uint32_t>((most + kSortBlock - 1) / kSortBlock, 64u), kSortBlock, 0, st>>>(
                a0, nTotA, b0, nTotB, nLook, B != nullptr);
        rs_hist_all<D><<<(B ? 2u : 1u) * ntiles, kSortBlock, 0, st>>>(a0, (uint32_t)PA, b0, B ? (uint32_t)PB : 0u,
                                                                       tiles, ntiles, nscenes);
        if (hook) hook(ctx, "rs_hist", 1);
    }
    for (int p = 0; p < P; ++p) {
        const bool inA = p < PA, inB = B && p < PB;
        const SortIO x = inA ? pass_io(A, p, R, nscenes, ntiles, oneSweep) : pass_io(*B, p, R, nscenes, ntiles, oneSweep);
        const SortIO y = inA && inB ? pass_io(*B, p, R, nscenes, ntiles, oneSweep) : x;
        const uint32_t sorts = inA && inB ? 2u : 1u;
        const uint32_t shift = (uint32_t)(p * D);
        if (oneSweep) {
            if (hook) hook(ctx, "rs_onesweep", 0);
            if (p == 0)
                rs_scatter<D, true, true><<<sorts * ntiles, kSortBlock, 0, st>>>(x, y, tiles, shift, limit, ntiles);
            else
                rs_scatter<D, false, true><<<sorts * ntiles, kSortBlock, 0, st>>>(x, y, tiles, shift, limit, ntiles);
            if (hook) hook(ctx, "rs_onesweep", 1);
        } else {
            if (hook) hook(ctx, "rs_hist", 0);
            rs_hist<D><<<sorts * ntiles, kSortBlock, 0, st>>>(x, y, tiles, ntiles, shift);
            if (hook) hook(ctx, "rs_hist", 1);
            if (hook) hook(ctx, "rs_scan", 0);
            rs_scan<D><<<sorts * nscenes * R, 64, 0, st>>>(x, y, scenes, nscenes);
            if (hook) hook(ctx, "rs_scan", 1);
            if (hook) hook(ctx, "rs_scatter", 0);
            if (p == 0)
                rs_scatter<D, true, false><<<sorts * ntiles, kSortBlock, 0, st>>>(x, y, tiles, shift, limit, ntiles);
            else
                rs_scatter<D, false, false><<<sorts * ntiles, kSortBlock, 0, st>>>(x, y, tiles, shift, limit, ntiles);
            if (hook) hook(ctx, "rs_scatter", 1);
        }
    }
}

void dispatch_sorts(int digitBits, hipStream_t st, const SortBuffers& A, int PA, const SortBuffers* B, int PB,
                    const SortTile* tiles, uint32_t ntiles, const SortScene* scenes, uint32_t nscenes,
                    uint32_t limit, bool oneSweep, bool zeroed, LaunchHook hook, void* ctx) {
    switch (digitBits) {
        case 4: run_sorts<4>(st, A, PA, B, PB, tiles, ntiles, scenes, nscenes, limit, oneSweep, zeroed, hook, ctx); break;
        case 5: run_sorts<5>(st, A, PA, B, PB, tiles, ntiles, scenes, nscenes, limit, oneSweep, zeroed, hook, ctx); break;
        case 6: run_sorts<6>(st, A, PA, B, PB, tiles, ntiles, scenes, nscenes, limit, oneSweep, zeroed, hook, ctx); break;
        case 7: run_sorts<7>(st, A, PA, B, PB, tiles, ntiles, scenes, nscenes, limit, oneSweep, zeroed, hook, ctx); break;
        default: run_sorts<8>(st, A, PA, B, PB, tiles, ntiles, scenes, nscenes, limit, oneSweep, zeroed, hook, ctx); break;
    }
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
    const bool oneSweep = numTiles <= kOneSweepMaxTiles && buf.look && buf.tickets;
    dispatch_sorts(plan.digitBits, stream, buf, P, nullptr, 0, dTiles, numTiles, dScenes, numScenes, limit,
                   oneSweep, false, hook, hookCtx);
    return (P % 2 == 0) ? 0 : 1;
}

bool radix_sort_pairs2(hipStream_t stream, const SortPlan& planA, const SortBuffers& bufA,
                       const SortPlan& planB, const SortBuffers& bufB, const SortTile* dTiles,
                       uint32_t numTiles, const SortScene* dScenes, uint32_t numScenes, uint32_t limit,
                       LaunchHook hook, void* hookCtx, int& resultA, int& resultB, bool scratchZeroed) {
    if (planA.digitBits != planB.digitBits) return false;
    const int PA = planA.passes(), PB = planB.passes();
    const bool oneSweep = numTiles <= kOneSweepMaxTiles && bufA.look && bufA.tickets && bufB.look && bufB.tickets;
    dispatch_sorts(planA.digitBits, stream, bufA, PA, &bufB, PB, dTiles, numTiles, dScenes, numScenes, limit,
                   oneSweep, scratchZeroed, hook, hookCtx);
    resultA = (PA % 2 == 0) ? 0 : 1;
    resultB = (PB % 2 == 0) ? 0 : 1;
    return true;
}

SortZeroSpans radix_sort_zero_spans(const SortPlan& planA, const SortBuffers& bufA, const SortPlan& planB,
                                    const SortBuffers& bufB, uint32_t numTiles, uint32_t numScenes) {
    SortZeroSpans z{};
    const bool oneSweep = numTiles <= kOneSweepMaxTiles && bufA.look && bufA.tickets && bufB.look && bufB.tickets;
    if (!oneSweep || planA.digitBits != planB.digitBits) return z;
    const uint32_t R = 1u << planA.digitBits;
    const SortBuffers* bufs[2] = {&bufA, &bufB};
    const int passes[2] = {planA.passes(), planB.passes()};
    for (int i = 0; i < 2; ++i) {
        z.p[z.count] = bufs[i]->totals;
        z.n[z.count++] = (uint32_t)passes[i] * numScenes * R;
        z.p[z.count] = bufs[i]->look;
        z.n[z.count++] = numTiles * R;
        z.p[z.count] = bufs[i]->tickets;
        z.n[z.count++] = (uint32_t)kMaxPasses;
    }
    return z;
}

}  // namespace akua
