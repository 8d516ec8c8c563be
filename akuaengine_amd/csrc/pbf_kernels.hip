// pbf_kernels.hip -- the PBF step on gfx950, one thread per particle slot.
//
// Stage map (reference kernel -> kernel here):
//   kernel_predict_position + kernel_compute_hashes     -> k_predict_hash
//   thrust::sort of the records                         -> radix sort by hash (logical order),
//                                                          k_cell_keys, radix sort by super-cell
//                                                          (physical order), k_gather
//   table fill + kernel_build_hash_table                -> k_cell_keys (cell-table inserts)
//   kernel_find_neighbours                              -> k_neighbours
//   kernel_calculate_densities + kernel_calculate_lambdas   -> k_density_lambda (one sweep)
//   kernel_calculate_position_delta + kernel_correct_position -> k_delta_correct
//   ... last iteration also kernel_update_position_and_velocity
//       + kernel_apply_boundary_velocity_damping        -> k_delta_correct<final>
//   kernel_compute_vorticities                          -> k_vorticity
//   kernel_apply_vorticity_confinement                  -> k_confinement
//   kernel_apply_xsph_viscosity (Jacobi)                -> k_xsph
// Every kernel reads the previous kernel's complete output (Jacobi), like the reference's
// launch-then-synchronise sequence, except XSPH, which reads a copy of the velocities instead of
// racing on them.
//
// Kernels that walk neighbours map workgroups to XCDs by contiguous ranges of physical slots:
// the hardware places workgroup b on XCD b % 8, so workgroup b takes chunk (b % 8) * (nb / 8) +
// b / 8 and each XCD's L2 serves one spatially coherent eighth of the particles.
#include "pbf_kernels.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include "boundary.h"
#include "pbf_math.h"

namespace akua {

namespace {

constexpr uint32_t kXcds = kXcdCount;

__device__ __forceinline__ const SceneDev& scene_of(const DeviceState& d, uint32_t i) {
    return d.scenes[d.chunkScene[i >> 8]];
}

// The neighbour list in chunks of four entries: chunk c (entries 4c..4c+3) of slot i is the
// uint4 at list_base(i) + c * 64, so the 64 lanes of a wave read chunk c of their lists as one
// 1 KB block (one 16-byte load per lane for four entries).
__device__ __forceinline__ uint32_t list_chunks(int maxN) { return ((uint32_t)maxN + 3u) >> 2; }
__device__ __forceinline__ size_t list_base(uint32_t i, int maxN) {
    return (size_t)(i >> 6) * (size_t)list_chunks(maxN) * 64u + (i & 63u);
}

// Visits the n entries of a neighbour list in order.  Memory requests are issued kListBatch at a
// time (the neighbours' data through load(j)), then body(value) runs for each entry in list order,
// so every per-particle sum keeps its reference order.  The list chunks of the next batch are
// loaded while the current batch's data is in flight, so a batch's gathers wait only for
// themselves, not for a list load before them.
constexpr int kListBatch = 4;
// per-kernel batch: 4 where the 8-entry batch's registers would cost occupancy
constexpr int kBatchDelta = 4, kBatchVort = 4, kBatchConf = 4, kBatchXsph = 8;

template <class T, int B = kListBatch, class Load, class Body>
__device__ __forceinline__ void walk_list(const uint4* __restrict__ lp, uint32_t n, Load load,
                                          Body body) {
    static_assert(B == 4 || B == 8, "one or two chunks per batch");
    if (n == 0u) return;
    uint4 a = lp[0];
    uint4 b = (B == 8 && n > 4u) ? lp[64] : a;
    for (uint32_t k0 = 0; k0 < n; k0 += B) {
        uint32_t j[8];
        j[0] = a.x; j[1] = a.y; j[2] = a.z; j[3] = a.w;
        if (B == 8) {
            j[4] = b.x; j[5] = b.y; j[6] = b.z; j[7] = b.w;
        }
        // entries past n are not list entries: load a valid one in their place (discarded)
#pragma unroll
        for (int u = 1; u < B; ++u) j[u] = k0 + u < n ? j[u] : j[0];
        T v[B];
#pragma unroll
        for (int u = 0; u < B; ++u) v[u] = load(j[u]);
        const uint32_t k1 = k0 + B;
        if (k1 < n) {
            a = lp[(size_t)(k1 >> 2) * 64u];
            if (B == 8) b = k1 + 4u < n ? lp[(size_t)((k1 >> 2) + 1u) * 64u] : a;
        }
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (k0 + u < n) body(v[u]);
    }
}

struct XLam {
    float4 x;
    float lam;
};
struct XV {
    float4 x;
    float4 v;
};
struct XW {
    float4 x;
    float w;
};

// Element j of a per-slot array through a 32-bit byte offset from its base: slot counts stay
// below kMaxSlots (2^28), so j * sizeof(T) <= 16 j fits in 32 bits, and the gather then takes the
// base from scalar registers and one 32-bit lane offset instead of a 64-bit address per lane.
template <class T>
__device__ __forceinline__ T ld(const T* __restrict__ base, uint32_t j) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (uint32_t)(j * (uint32_t)sizeof(T)));
}

// Physical slot of this thread under the XCD-contiguous workgroup order (the grid is a multiple
// of 8 workgroups of BS threads).
template <int BS>
__device__ __forceinline__ uint32_t xcd_slot() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    return ((b % kXcds) * (nb / kXcds) + b / kXcds) * BS + threadIdx.x;
}

// The same slot for virtual workgroup vb of nvb (a kernel that loops over virtual workgroups
// with a smaller grid).
template <int BS>
__device__ __forceinline__ uint32_t xcd_slot_of(uint32_t vb, uint32_t nvb) {
    return ((vb % kXcds) * (nvb / kXcds) + vb / kXcds) * BS + threadIdx.x;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u4(uint4* __restrict__ p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_uint4(v, v, v, v);
}

__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0xFF0000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Morton code of the particle's super-cell (2^L cells per axis, L = layoutOrigin[3]), relative to
// the scene's layout origin and clamped to the key range.  It only decides where the particle's
// data is stored.
__device__ __forceinline__ uint32_t layout_key(int cx, int cy, int cz, const SceneDev& sc) {
    const int hi = (1 << sc.layoutBits) - 1;
    const int L = sc.layoutOrigin[3];
    int u = min(max((cx >> L) - sc.layoutOrigin[0], 0), hi);
    int v = min(max((cy >> L) - sc.layoutOrigin[1], 0), hi);
    int w = min(max((cz >> L) - sc.layoutOrigin[2], 0), hi);
    return spread3((uint32_t)u) | (spread3((uint32_t)v) << 1) | (spread3((uint32_t)w) << 2);
}

// kernel_predict_position (IntegrationCUDA.cu:27-36) + kernel_compute_hashes
// (NeighbourSearchCUDA.cu:36-44) with smoothRadius as the cell size (NeighbourSearchCUDA.cu:163).
// The hash, and the layout key with the particle's current physical slot, are stored at its
// previous logical index, the first sort's input order.
// layoutPrev: the layout sort runs over the previous logical order too, in the hash sort's
// launches (single-cell super-cells, where the previous and the new logical order agree within
// every layout key): its keys go straight into its input (keysM) and the record keeps only the
// slot, to which k_cell_keys adds the new logical index.
// Blocks from predictBlocks on clear the cell table instead (every record empty), and the last
// kZeroBlocks zero the sorts' scratch spans: neither is touched by the prediction, so one launch
// serves all three.
constexpr uint32_t kZeroBlocks = 8;

__global__ __launch_bounds__(kBlock) void k_predict_hash(DeviceState d, SphParams P, StepParams S,
                                                         bool layoutPrev, uint32_t predictBlocks,
                                                         uint32_t clearBlocks, SortZeroSpans zero) {
    if (blockIdx.x >= predictBlocks + clearBlocks) {
        const uint32_t i = (blockIdx.x - predictBlocks - clearBlocks) * kBlock + threadIdx.x;
#pragma unroll
        for (uint32_t s = 0; s < 6; ++s) {
            if (s >= zero.count) break;
            for (uint32_t k = i; k < zero.n[s]; k += kZeroBlocks * kBlock) zero.p[s][k] = 0u;
        }
        return;
    }
    if (blockIdx.x >= predictBlocks) {
        const uint32_t i = (blockIdx.x - predictBlocks) * kBlock + threadIdx.x;
        if (i < d.tableTotal) d.table[i] = make_uint4(kEmptyKey, kEmptyKey, kEmptyKey, kEmptyKey);
        return;
    }
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= d.nslots) return;
    const SceneDev& sc = scene_of(d, p);
    if (p >= sc.end) return;
    float4 x = d.X[p], v = d.V[p];
    f3 xs = predict_x(make3(x), make3(v), S);
    int cx, cy, cz;
    discretize(xs, P.cellHash, cx, cy, cz);
    const uint32_t k = d.lid[p];
    d.keysL[k] = P.searchCells == 2 ? cell_key_fine(cx, cy, cz) : cell_hash(cx, cy, cz, sc);
    if (layoutPrev) {
        d.keysM[k] = layout_key(cx, cy, cz, sc);
        d.layPhys[k] = make_uint2(0u, p);
    } else {
        d.layPhys[k] = make_uint2(layout_key(cx, cy, cz, sc), p);
    }
}

__device__ __forceinline__ uint32_t table_insert(uint4* table, const SceneDev& sc, uint32_t h) {
    const uint32_t mask = (1u << sc.tabLog2) - 1u;
    uint32_t s = table_slot(h, sc.tabLog2);
    // The table holds at least twice as many slots as the scene has particles, so an empty
    // slot is always found; the probe count is bounded all the same.
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        uint32_t prev = atomicCAS(&table[sc.tabOffset + s].x, kEmptyKey, h);
        if (prev == kEmptyKey || prev == h) return sc.tabOffset + s;
        s = (s + 1u) & mask;
    }
    return kEmptyKey;
}

// Per logical index k: the layout key (second sort's input), the previous physical slot of the
// particle, and the cell-table entry of k's bucket (kernel_build_hash_table,
// NeighbourSearchCUDA.cu:52-65, plus the bucket end the reference finds by scanning).
// kKeep: the physical layout is kept this step (every particle stays in its slot), so instead of
// the layout key, the inverse map physical slot -> new logical index is written directly.
// layoutPrev (not with kKeep): the layout was sorted over the previous logical order, so the
// record at the previous logical index receives the new one instead (read back by k_gather).
template <bool kKeep>
__global__ __launch_bounds__(kBlock) void k_cell_keys(DeviceState d, SphParams P, StepParams S,
                                                      bool layoutPrev) {
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= d.nslots) return;
    const SceneDev& sc = scene_of(d, k);
    if (k >= sc.end) return;
    const uint32_t prevLogical = d.perm1[k];
    if (prevLogical >= d.nslots) return;  // a consistent sort never produces this
    if (!kKeep && layoutPrev) {
        reinterpret_cast<uint32_t*>(d.layPhys)[2u * prevLogical] = k;
    } else {
        const uint2 lp = d.layPhys[prevLogical];  // (layout key, previous physical slot)
        const uint32_t src = lp.y;
        if (kKeep) {
            if (src < d.nslots) d.perm2keep[src] = k;
        } else {
            d.keysM[k] = lp.x;
        }
        d.srcL[k] = src;
    }
    const uint32_t h = d.skeys[k];
    const bool first = (k == sc.begin) || (d.skeys[k - 1] != h);
    const bool last = (k + 1 == sc.end) || (d.skeys[k + 1] != h);
    if (first || last) {
        uint32_t slot = table_insert(d.table, sc, h);
        if (slot != kEmptyKey) {
            if (first) d.table[slot].y = k;
            if (last) d.table[slot].z = k + 1;
        }
        if (first) d.bslot[k] = slot;
    }
}

// Moves each particle to its new physical slot (the record move of the reference's thrust::sort).
// x* is recomputed from the gathered x and v with the same operations as k_predict_hash, so it is
// bit-identical to the value that was hashed.
// kCarry: also the previous density and lambda (read by a step without iterations and by the
// learned features; a step with iterations rewrites both before anything reads them).
// kCarryDelta: also the position delta (a step without iterations keeps the previous one).
// layoutPrev: the layout was sorted over the previous logical order, so perm2 gives the previous
// logical index, whose record holds the new logical index and the previous physical slot.
template <bool kCarry, bool kCarryDelta>
__global__ __launch_bounds__(kBlock) void k_gather(DeviceState d, SphParams P, StepParams S,
                                                   bool layoutPrev) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= d.nslots) return;
    const SceneDev& sc = scene_of(d, p);
    if (p >= sc.end) return;
    uint32_t k, src;
    if (layoutPrev) {
        const uint32_t prev = d.perm2[p];
        if (prev >= d.nslots) return;  // consistent sorts never produce these
        const uint2 r = d.layPhys[prev];
        k = r.x;
        src = r.y;
    } else {
        k = d.perm2[p];
        if (k >= d.nslots) return;
        src = d.srcL[k];
    }
    if (k >= d.nslots || src >= d.nslots) return;
    float4 x = d.X[src], v = d.V[src];
    f3 xs = predict_x(make3(x), make3(v), S);
    const float4 xs4 = make_float4(xs.x, xs.y, xs.z, x.w);
    d.X0[p] = x;
    d.XsA[p] = xs4;
    if (kCarry) {
        d.Tdens[p] = d.dens[src];
        d.Tlam[p] = d.lam[src];
    }
    if (kCarryDelta) d.TDP[p] = d.DP[src];
    d.Tid[p] = d.id[src];
    d.lidT[p] = k;
    d.physT[k] = p;
}

// The physical slot of each logical index from the step's persistent lid (outside a step, where
// physT no longer holds it: rendering draws in logical order).
__global__ __launch_bounds__(kBlock) void k_invert_order(DeviceState d, uint32_t* __restrict__ phys) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= d.nslots) return;
    const SceneDev& sc = scene_of(d, p);
    if (p >= sc.end) return;
    const uint32_t k = d.lid[p];
    if (k < d.nslots) phys[k] = p;
}

// Per bucket: where it starts in physical order and whether its particles occupy consecutive
// physical slots in logical order (always true for a bucket of one cell, which lies inside one
// super-cell; buckets of more than 256 particles are not marked).  One thread per logical index k,
// all reads in logical order: the bucket's first index ANDs (first physical slot | run bit) into
// the record's w, which the table fill left all ones, and every index whose logical successor in
// the bucket is not in the next physical slot clears the run bit (ANDs commute, so the order of
// the two does not matter).  Only hash-collision buckets take that second path.
__device__ __forceinline__ uint32_t table_slot_of(const uint4* __restrict__ table, const SceneDev& sc,
                                                  uint32_t h) {
    const uint32_t mask = (1u << sc.tabLog2) - 1u;
    uint32_t s = table_slot(h, sc.tabLog2);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const uint32_t key = table[sc.tabOffset + s].x;
        if (key == h) return sc.tabOffset + s;
        if (key == kEmptyKey) return kEmptyKey;
        s = (s + 1u) & mask;
    }
    return kEmptyKey;
}

__global__ __launch_bounds__(kBlock) void k_bucket_phys(DeviceState d) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= d.nslots) return;
    const SceneDev& sc = scene_of(d, k);
    if (k >= sc.end) return;
    const uint32_t h = d.skeys[k];
    const uint32_t p = d.physT[k];
    if (k == sc.begin || d.skeys[k - 1] != h) {
        const uint32_t slot = d.bslot[k];
        if (slot < d.tableTotal) {
            // at most 256 indices: the one 256 ahead is past the scene or in another bucket (the
            // sorted hashes, read nearby, instead of the record's end through the table)
            const bool small = k + 256u >= sc.end || d.skeys[k + 256u] != h;
            atomicAnd(&d.table[slot].w, p | (small ? 0x80000000u : 0u));
        }
    }
    if (k + 1u < sc.end && d.skeys[k + 1u] == h && d.physT[k + 1u] != p + 1u) {
        const uint32_t slot = table_slot_of(d.table, sc, h);
        if (slot != kEmptyKey) atomicAnd(&d.table[slot].w, 0x7FFFFFFFu);
    }
}

// A bucket's entry (first logical index, one past the last, first physical slot, contiguity)
// from its table record, with its ranges clamped to the scene (a consistent table never needs
// it, but no entry can then index outside the scene).
__device__ __forceinline__ uint4 clamp_entry(uint4 r, const SceneDev& sc) {
    uint4 e = make_uint4(r.y, r.z, r.w & 0x7FFFFFFFu, r.w >> 31);
    e.x = max(e.x, sc.begin);
    e.y = min(e.y, sc.end);
    if (e.y < e.x) e.y = e.x;
    if (e.w && (e.z < sc.begin || e.z + (e.y - e.x) > sc.end)) e.w = 0u;
    return e;
}

__device__ __forceinline__ bool table_find(const uint4* __restrict__ table, const SceneDev& sc,
                                           uint32_t h, uint4& entry) {
    const uint32_t mask = (1u << sc.tabLog2) - 1u;
    uint32_t s = table_slot(h, sc.tabLog2);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const uint4 r = table[sc.tabOffset + s];
        if (r.x == h) {
            entry = clamp_entry(r, sc);
            return true;
        }
        if (r.x == kEmptyKey) return false;
        s = (s + 1u) & mask;
    }
    return false;
}

// kernel_find_neighbours (NeighbourSearchCUDA.cu:72-130): stencil cells in (dx, dy, dz) order,
// each bucket walked in logical order, self skipped, r^2 < h^2 kept, at most maxNeighbours.  A
// bucket reached from two stencil cells with equal hashes is walked twice, as in the reference.
// The list stores physical slots.
//
// The walk is latency-bound, so memory requests are issued in batches: the nine table probes of
// one dx slab together, then each bucket's candidates eight at a time (position and physical
// slot), after which the candidates are tested and appended strictly in order.
constexpr int kWalkBatch = 8;
constexpr int kProbeGroup = 3;

// Physical slots of bucket positions m0 .. m0 + N - 1 (clamped to the bucket's last, b - 1): by
// the bucket's offset when its slots are consecutive (run), else from physT.  The physT loads sit
// on a branch of their own, so the walks' position loads stay one batch of independent requests.
template <int N>
__device__ __forceinline__ void batch_slots(const DeviceState& d, bool run, uint32_t m0, uint32_t b,
                                            int32_t shift, uint32_t (&ph)[N]) {
    if (__builtin_expect(run, 1)) {
#pragma unroll
        for (int u = 0; u < N; ++u) ph[u] = (uint32_t)((int32_t)min(m0 + u, b - 1u) + shift);
    } else {
#pragma unroll
        for (int u = 0; u < N; ++u) ph[u] = ld(d.physT, min(m0 + u, b - 1u));
    }
}

__device__ __forceinline__ void neighbours_slot(const DeviceState& d, const SphParams& P,
                                                uint32_t p) {
    uint32_t count = 0;
    if (p < d.nslots) {
        const SceneDev& sc = scene_of(d, p);
        if (p < sc.end) {
            const f3 xi = make3(d.XsA[p]);
            int cx, cy, cz;
            discretize(xi, P.cellSearch, cx, cy, cz);
            // entry k of this list: element k % 4 of chunk k / 4, at byte offset
            // (p % 64) * 16 + (k / 4) * 1024 + (k % 4) * 4 from the wave's list block, whose
            // address is the same in every lane (the wave's 64 slots share p / 64)
            char* const waveOut = reinterpret_cast<char*>(
                d.nbr4 + (size_t)__builtin_amdgcn_readfirstlane(p >> 6) * list_chunks(P.maxN) * 64u);
            const uint32_t laneOff = (p & 63u) * 16u;
            const uint32_t maxN = (uint32_t)P.maxN;
            // stencil cells c = 0..26 in (dx, dy, dz) order, probed kProbeGroup at a time
            for (int g = 0; g < 27 / kProbeGroup && count < maxN; ++g) {
                uint32_t hq[kProbeGroup], sl[kProbeGroup];
                uint4 rq[kProbeGroup], e[kProbeGroup];
#pragma unroll
                for (int q = 0; q < kProbeGroup; ++q) {
                    const int c = g * kProbeGroup + q;
                    hq[q] = cell_hash((int)((uint32_t)cx + (uint32_t)(c / 9 - 1)),
                                      (int)((uint32_t)cy + (uint32_t)((c / 3) % 3 - 1)),
                                      (int)((uint32_t)cz + (uint32_t)(c % 3 - 1)), sc);
                    sl[q] = sc.tabOffset + table_slot(hq[q], sc.tabLog2);
                }
#pragma unroll
                for (int q = 0; q < kProbeGroup; ++q) rq[q] = d.table[sl[q]];
#pragma unroll
                for (int q = 0; q < kProbeGroup; ++q) {
                    e[q] = make_uint4(0u, 0u, 0u, 0u);
                    if (rq[q].x == hq[q]) {
                        e[q] = clamp_entry(rq[q], sc);
                    } else if (rq[q].x != kEmptyKey) {
                        if (!table_find(d.table, sc, hq[q], e[q])) e[q] = make_uint4(0u, 0u, 0u, 0u);
                    }
                }
#pragma unroll
                for (int q = 0; q < kProbeGroup; ++q) {
                    const uint32_t a = e[q].x, b = e[q].y;
                    // bucket of consecutive physical slots: read the cell-ordered positions
                    // directly; otherwise (hash collisions) find each slot through physT first
                    const bool run = e[q].w != 0u;
                    const int32_t shift = (int32_t)e[q].z - (int32_t)a;  // physical - logical
                    for (uint32_t m0 = a; m0 < b && count < maxN; m0 += kWalkBatch) {
                        float4 c[kWalkBatch];
                        uint32_t ph[kWalkBatch];
                        batch_slots<kWalkBatch>(d, run, m0, b, shift, ph);
#pragma unroll
                        for (int u = 0; u < kWalkBatch; ++u) c[u] = ld(d.XsA, ph[u]);
                        // test the whole batch without branches, then append in order
                        uint32_t acc = 0u;
#pragma unroll
                        for (int u = 0; u < kWalkBatch; ++u) {
                            const f3 sep = xi - make3(c[u]);
                            const bool ok = (dot3(sep, sep) < P.h2) & (ph[u] != p) & (m0 + u < b);
                            acc |= (uint32_t)ok << u;
                        }
                        const uint32_t room = maxN - count;
                        while ((uint32_t)__popc(acc) > room) acc &= ~(1u << (31 - __clz(acc)));
                        if (acc) {
                            // append in order: the u-th candidate goes to position count + (the
                            // accepted candidates before it)
#pragma unroll
                            for (int u = 0; u < kWalkBatch; ++u) {
                                if (acc & (1u << u)) {
                                    const uint32_t k = count + (uint32_t)__popc(acc & ((1u << u) - 1u));
                                    *reinterpret_cast<uint32_t*>(waveOut + (laneOff + ((k & ~3u) << 8) +
                                                                            ((k & 3u) << 2))) = ph[u];
                                }
                            }
                            count += (uint32_t)__popc(acc);
                        }
                    }
                }
            }
        }
        d.cnt[p] = count;
    }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_neighbours(DeviceState d, SphParams P) {
    neighbours_slot(d, P, xcd_slot<BS>());
}

// The same lists with one lane per stencil cell, for scenes too small to fill the chip: 32 lanes
// per particle (lanes 27-31 idle), lane c on stencil cell c.  Each lane walks its bucket twice:
// first counting its accepted candidates, then, after an exclusive prefix of the 27 counts in
// stencil order, storing them at their list positions; positions at or past maxN are dropped, as
// the one-lane walk stops there.  A bucket reached from two stencil cells is walked by both lanes,
// as in the reference.
template <int BS>
__global__ __launch_bounds__(BS) void k_neighbours_cells(DeviceState d, SphParams P) {
    const uint32_t T = xcd_slot<BS>();
    const uint32_t p = T >> 5, c = T & 31u;
    if (p >= d.nslots) return;  // whole groups
    const SceneDev& sc = scene_of(d, p);
    if (p >= sc.end) return;
    const f3 xi = make3(d.XsA[p]);
    int cx, cy, cz;
    discretize(xi, P.cellSearch, cx, cy, cz);
    uint4 e = make_uint4(0u, 0u, 0u, 0u);
    if (c < 27u) {
        const uint32_t h = cell_hash((int)((uint32_t)cx + (c / 9u - 1u)), (int)((uint32_t)cy + ((c / 3u) % 3u - 1u)),
                                     (int)((uint32_t)cz + (c % 3u - 1u)), sc);
        if (!table_find(d.table, sc, h, e)) e = make_uint4(0u, 0u, 0u, 0u);
    }
    const uint32_t a = e.x, b = e.y;
    const bool run = e.w != 0u;
    const int32_t shift = (int32_t)e.z - (int32_t)a;  // physical - logical
    // the accepted candidates of bucket positions [m0, m0 + kWalkBatch), one bit each
    const auto batch = [&](uint32_t m0, uint32_t (&ph)[kWalkBatch]) {
        float4 cand[kWalkBatch];
        batch_slots<kWalkBatch>(d, run, m0, b, shift, ph);
#pragma unroll
        for (int u = 0; u < kWalkBatch; ++u) cand[u] = ld(d.XsA, ph[u]);
        uint32_t acc = 0u;
#pragma unroll
        for (int u = 0; u < kWalkBatch; ++u) {
            const f3 sep = xi - make3(cand[u]);
            const bool ok = (dot3(sep, sep) < P.h2) & (ph[u] != p) & (m0 + u < b);
            acc |= (uint32_t)ok << u;
        }
        return acc;
    };
    uint32_t mine = 0;
    for (uint32_t m0 = a; m0 < b; m0 += kWalkBatch) {
        uint32_t ph[kWalkBatch];
        mine += (uint32_t)__popc(batch(m0, ph));
    }
    uint32_t incl = mine;  // inclusive prefix over the group's 32 lanes (stencil order)
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 32);
        if ((int)c >= off) incl += y;
    }
    const uint32_t total = __shfl(incl, 31, 32);
    const uint32_t maxN = (uint32_t)P.maxN;
    uint32_t k = incl - mine;  // this cell's first list position
    if (mine && k < maxN) {
        uint32_t* out = reinterpret_cast<uint32_t*>(d.nbr4 + list_base(p, P.maxN));
        for (uint32_t m0 = a; m0 < b && k < maxN; m0 += kWalkBatch) {
            uint32_t ph[kWalkBatch];
            const uint32_t acc = batch(m0, ph);
#pragma unroll
            for (int u = 0; u < kWalkBatch; ++u) {
                if ((acc >> u) & 1u) {
                    if (k < maxN) out[((size_t)(k >> 2) * 64u) * 4u + (k & 3u)] = ph[u];
                    ++k;
                }
            }
        }
    }
    if (c == 0u) d.cnt[p] = min(total, maxN);
}

// The balanced work order (launch_balance): one workgroup per tile of T physical slots.  The
// slots are ranked by the key (scene rank within the tile, 254 - min(count, 254)), padding slots
// past a scene's end last within their scene; a slot's work position is the tile base plus its
// rank.  A scene owns whole 256-slot chunks, so its slots keep exactly the positions of its own
// chunks and every wave of work positions belongs to one scene.  Equal keys are ranked by LDS
// atomics, in no fixed order: the order only decides which lane handles which particle, never a
// result.  Then each lane copies the list of one work position into nbrW, one 16-byte chunk per
// four entries, so a wave writes each chunk row as one 1 KB block.
constexpr uint32_t kBalanceThreads = 256;
constexpr uint32_t kBalanceBins = (kBalanceMaxTile / 256) * 256;  // scene ranks x count classes

__global__ __launch_bounds__(kBalanceThreads) void k_balance(DeviceState d, SphParams P, uint32_t T,
                                                             uint32_t* __restrict__ work,
                                                             uint32_t* __restrict__ cntW,
                                                             uint4* __restrict__ nbrW) {
    __shared__ uint32_t bins[kBalanceBins];
    __shared__ uint32_t scan[kBalanceThreads];
    __shared__ uint16_t keyOf[kBalanceMaxTile];
    __shared__ uint16_t rankOf[kBalanceMaxTile];
    __shared__ uint32_t srcOf[kBalanceMaxTile];
    const uint32_t base = blockIdx.x * T;
    if (base >= d.nslots) return;
    const uint32_t n = min(T, d.nslots - base);
    const uint32_t tid = threadIdx.x;
    for (uint32_t b = tid; b < kBalanceBins; b += kBalanceThreads) bins[b] = 0u;
    __syncthreads();
    const uint32_t s0 = d.chunkScene[base >> 8];
    for (uint32_t e = tid; e < n; e += kBalanceThreads) {
        const uint32_t p = base + e;
        const uint32_t sc = d.chunkScene[p >> 8];
        const bool valid = p < d.scenes[sc].end;
        const uint32_t cls = valid ? 254u - min(d.cnt[p], 254u) : 255u;
        const uint32_t key = min(sc - s0, kBalanceMaxTile / 256 - 1) * 256u + cls;
        keyOf[e] = (uint16_t)key;
        rankOf[e] = (uint16_t)atomicAdd(&bins[key], 1u);
    }
    __syncthreads();
    // exclusive scan of the bins: each thread sums a run of kBalanceBins / kBalanceThreads bins
    constexpr uint32_t kRun = kBalanceBins / kBalanceThreads;
    uint32_t run = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRun; ++k) run += bins[tid * kRun + k];
    scan[tid] = run;
    __syncthreads();
    for (uint32_t off = 1; off < kBalanceThreads; off <<= 1) {
        const uint32_t v = tid >= off ? scan[tid - off] : 0u;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    uint32_t acc = scan[tid] - run;
#pragma unroll
    for (uint32_t k = 0; k < kRun; ++k) {
        const uint32_t c = bins[tid * kRun + k];
        bins[tid * kRun + k] = acc;
        acc += c;
    }
    __syncthreads();
    for (uint32_t e = tid; e < n; e += kBalanceThreads) {
        const uint32_t r = bins[keyOf[e]] + rankOf[e];
        srcOf[r] = base + e;
    }
    __syncthreads();
    const uint32_t chunks = list_chunks(P.maxN);
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    for (uint32_t w = wave; w * 64u < n; w += kBalanceThreads / 64u) {
        const uint32_t r = w * 64u + lane;
        if (r >= n) continue;
        const uint32_t t = base + r, p = srcOf[r];
        const uint32_t c = d.cnt[p];
        work[t] = p;
        cntW[t] = c;
        const uint4* src = d.nbr4 + list_base(p, P.maxN);
        uint4* dst = nbrW + list_base(t, P.maxN);
        const uint32_t nc = min((c + 3u) >> 2, chunks);
        for (uint32_t k = 0; k < nc; ++k) dst[(size_t)k * 64u] = src[(size_t)k * 64u];
    }
}

// Physical slot of work position t of a neighbour sweep (t itself without balancing).
__device__ __forceinline__ uint32_t work_slot(const DeviceState& d, uint32_t t) {
    return (d.work != nullptr && t < d.nslots) ? d.work[t] : t;
}

// The list build walking each dx slab of the stencil (its nine cells in (dy, dz) order) as one
// stream of candidates.  The nine buckets' entries (first slot, length, contiguity) sit in a
// register queue; a batch of B stream positions may span several buckets, so a lane tests about as
// many candidates as its slab holds instead of a whole batch per bucket (per-wave candidate slots
// drop by about half on the 1M dam break).  Each lane first maps its B positions to buckets by
// shifting the queue past exhausted buckets, then issues the B loads, then tests and appends in
// stream order, which is the reference's list order.  Lists equal k_neighbours' entry for entry.
// Held to 7 waves per SIMD (at most 72 VGPRs, a few spilled): the 256-scene batch's list build
// measured 3.26 -> 3.10 ms against 6 waves.
template <int BS, int B>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(7)))
void k_neighbours_slab(DeviceState d, SphParams P) {
    const uint32_t p = xcd_slot<BS>();
    uint32_t count = 0;
    if (p < d.nslots) {
        const SceneDev& sc = scene_of(d, p);
        if (p < sc.end) {
            const f3 xi = make3(d.XsA[p]);
            int cx, cy, cz;
            discretize(xi, P.cellSearch, cx, cy, cz);
            uint32_t* out = reinterpret_cast<uint32_t*>(d.nbr4 + list_base(p, P.maxN));
            const uint32_t maxN = (uint32_t)P.maxN;
            constexpr uint32_t kRun = 0x80000000u, kLen = 0x7FFFFFFFu;
            for (int g = 0; g < 3 && count < maxN; ++g) {
                // the slab's nine buckets: base (physical start of a contiguous bucket, else the
                // logical start) and length | contiguity flag
                uint32_t qb[9], ql[9];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    uint32_t hq[3];
                    uint4 rq[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const int c = g * 9 + r * 3 + q;
                        hq[q] = cell_hash((int)((uint32_t)cx + (uint32_t)(c / 9 - 1)),
                                          (int)((uint32_t)cy + (uint32_t)((c / 3) % 3 - 1)),
                                          (int)((uint32_t)cz + (uint32_t)(c % 3 - 1)), sc);
                        rq[q] = d.table[sc.tabOffset + table_slot(hq[q], sc.tabLog2)];
                    }
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        uint4 e = make_uint4(0u, 0u, 0u, 0u);
                        if (rq[q].x == hq[q]) {
                            e = clamp_entry(rq[q], sc);
                        } else if (rq[q].x != kEmptyKey) {
                            if (!table_find(d.table, sc, hq[q], e)) e = make_uint4(0u, 0u, 0u, 0u);
                        }
                        qb[r * 3 + q] = e.w ? e.z : e.x;
                        ql[r * 3 + q] = (e.y - e.x) | (e.w ? kRun : 0u);
                    }
                }
                uint32_t total = 0;
#pragma unroll
                for (int q = 0; q < 9; ++q) total += ql[q] & kLen;
                uint32_t segStart = 0, segEnd = ql[0] & kLen, left = 8;
                for (uint32_t t0 = 0; t0 < total && count < maxN; t0 += B) {
                    uint32_t mm[B];
                    bool rr[B];
#pragma unroll
                    for (int u = 0; u < B; ++u) {
                        const uint32_t t = t0 + u;
                        // shift the queue past exhausted (and empty) buckets; the lengths sum
                        // to total, so a position below total always stops in a bucket
                        while (t < total && t >= segEnd && left > 0) {
#pragma unroll
                            for (int k = 0; k < 8; ++k) {
                                qb[k] = qb[k + 1];
                                ql[k] = ql[k + 1];
                            }
                            ql[8] = 0u;
                            --left;
                            segStart = segEnd;
                            segEnd += ql[0] & kLen;
                        }
                        if (t < total) {
                            mm[u] = qb[0] + (t - segStart);
                            rr[u] = (ql[0] & kRun) != 0u;
                        } else {  // past the stream: repeat the previous position (discarded)
                            mm[u] = u ? mm[u - 1] : qb[0];
                            rr[u] = u ? rr[u - 1] : true;
                        }
                    }
                    float4 c[B];
                    uint32_t ph[B];
                    bool anyLogical = false;
#pragma unroll
                    for (int u = 0; u < B; ++u) {
                        ph[u] = mm[u];
                        anyLogical |= !rr[u];
                    }
                    if (__builtin_expect(anyLogical, 0)) {  // hash collisions: slots through physT
#pragma unroll
                        for (int u = 0; u < B; ++u)
                            if (!rr[u]) ph[u] = ld(d.physT, mm[u]);
                    }
#pragma unroll
                    for (int u = 0; u < B; ++u) c[u] = ld(d.XsA, ph[u]);
                    uint32_t acc = 0u;
#pragma unroll
                    for (int u = 0; u < B; ++u) {
                        const f3 sep = xi - make3(c[u]);
                        const bool ok = (dot3(sep, sep) < P.h2) & (ph[u] != p) & (t0 + u < total);
                        acc |= (uint32_t)ok << u;
                    }
                    const uint32_t room = maxN - count;
                    while ((uint32_t)__popc(acc) > room) acc &= ~(1u << (31 - __clz(acc)));
                    if (acc) {
#pragma unroll
                        for (int u = 0; u < B; ++u) {
                            if (acc & (1u << u)) {
                                const uint32_t k = count + (uint32_t)__popc(acc & ((1u << u) - 1u));
                                out[((size_t)(k >> 2) * 64u) * 4u + (k & 3u)] = ph[u];
                            }
                        }
                        count += (uint32_t)__popc(acc);
                    }
                }
            }
        }
        d.cnt[p] = count;
    }
}

// The non-parity search (SphParams::searchCells == 2): cells of c = h / 2, keyed by their packed
// coordinates (cell_key_fine), so every bucket is one cell and none is walked twice.  A cell at
// offset (dx, dy, dz) from the particle's cell, |d| <= 2, can hold a particle within h only if its
// box comes within h of x_i: with f the particle's position in its cell (cell units), the box lies
// d - f away along an axis for d >= 1, f - d - 1 for d <= -1 and 0 for d = 0, and it is walked when
// the squared distances sum below (h / c)^2 = 4 (with a margin of 0.25% for the rounding of f).
// Walked cells go in (dx, dy, dz) order, each bucket in logical order, self skipped, r^2 < h^2
// kept, at most maxNeighbours: the list holds every particle within h (up to the cap).
template <int BS>
__global__ __launch_bounds__(BS) void k_neighbours_fine(DeviceState d, SphParams P) {
    const uint32_t p = xcd_slot<BS>();
    uint32_t count = 0;
    if (p < d.nslots) {
        const SceneDev& sc = scene_of(d, p);
        if (p < sc.end) {
            const f3 xi = make3(d.XsA[p]);
            const float qx = xi.x / P.cellSearch, qy = xi.y / P.cellSearch, qz = xi.z / P.cellSearch;
            const int cx = to_int_sat(floorf(qx)), cy = to_int_sat(floorf(qy)), cz = to_int_sat(floorf(qz));
            const float fx = qx - floorf(qx), fy = qy - floorf(qy), fz = qz - floorf(qz);
            constexpr float kReach2 = 4.01f;
            uint32_t* out = reinterpret_cast<uint32_t*>(d.nbr4 + list_base(p, P.maxN));
            const uint32_t maxN = (uint32_t)P.maxN;
            auto gap = [](int dd, float f) {
                return dd > 0 ? (float)dd - f : dd < 0 ? f - (float)(dd + 1) : 0.0f;
            };
            for (int dx = -2; dx <= 2 && count < maxN; ++dx) {
                const float gx = gap(dx, fx), gx2 = gx * gx;
                if (gx2 >= kReach2) continue;
                for (int dy = -2; dy <= 2 && count < maxN; ++dy) {
                    const float gy = gap(dy, fy), gxy = gx2 + gy * gy;
                    if (gxy >= kReach2) continue;
                    for (int dz = -2; dz <= 2 && count < maxN; ++dz) {
                        const float gz = gap(dz, fz);
                        if (gxy + gz * gz >= kReach2) continue;
                        const uint32_t h = cell_key_fine(cx + dx, cy + dy, cz + dz);
                        uint4 e;
                        if (!table_find(d.table, sc, h, e)) continue;
                        const uint32_t a = e.x, b = e.y;
                        const bool run = e.w != 0u;
                        const int32_t shift = (int32_t)e.z - (int32_t)a;
                        for (uint32_t m0 = a; m0 < b && count < maxN; m0 += 4) {
                            float4 c[4];
                            uint32_t ph[4];
                            batch_slots<4>(d, run, m0, b, shift, ph);
#pragma unroll
                            for (int u = 0; u < 4; ++u) c[u] = ld(d.XsA, ph[u]);
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const f3 sep = xi - make3(c[u]);
                                if ((m0 + u < b) & (dot3(sep, sep) < P.h2) & (ph[u] != p) & (count < maxN)) {
                                    out[((size_t)(count >> 2) * 64u) * 4u + (count & 3u)] = ph[u];
                                    ++count;
                                }
                            }
                        }
                    }
                }
            }
        }
        d.cnt[p] = count;
    }
}

// Sum of the neighbour counts of every particle slot (statistics, on demand): one partial per
// workgroup, summed by the host.
__global__ __launch_bounds__(256) void k_count_sum(const DeviceState d, unsigned long long* part) {
    __shared__ unsigned long long sm[256];
    unsigned long long t = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < d.nslots; i += gridDim.x * 256)
        if (i < scene_of(d, i).end) t += d.cnt[i];
    sm[threadIdx.x] = t;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// kernel_calculate_densities (ConstraintSolverCUDA.cu:16-42) and kernel_calculate_lambdas
// (ConstraintSolverCUDA.cu:51-97) in one neighbour sweep.  Each sum keeps the reference's order
// (the list order), so fusing the loops changes no rounding.
// Residual contribution of one particle: max(0, C), with a non-finite C counted as infinite.
__device__ __forceinline__ float positive_part(float C) {
    return C > 0.0f ? C : (C == C ? 0.0f : __builtin_inff());
}

// Residual statistics of one check: the wave's max and sum of non-negative values, written as
// the wave's partial (no atomics; k_resid_reduce combines the partials of each scene in a fixed
// order).  Every wave with a slot below nslots writes its partial.
__device__ __forceinline__ void wave_resid_partial(const DeviceState& d, int slot, uint32_t i,
                                                   float v) {
    float m = v, t = v;
    for (int off = 32; off > 0; off >>= 1) {
        m = fmaxf(m, __shfl_xor(m, off, 64));
        t += __shfl_xor(t, off, 64);
    }
    if (__lane_id() == 0) d.residPart[(size_t)slot * d.nWaves + (i >> 6)] = make_float2(m, t);
}

// kFast (SphParams::fast): unit masses, range-tested quotients bounded in advance.
// kBnd: boundary samples (SphParams::bd): their density and gradient sums join rho and grad_i C
// after the neighbour sweep; the gradient sum is kept in Bg for the position-delta pass, which
// runs at the same x*.
template <int BS, bool kFast, bool kBnd>
__device__ __forceinline__ void density_lambda_slot(const DeviceState& d, const SphParams& P,
                                                    const float4* __restrict__ Xs, int residSlot,
                                                    int gated, uint32_t t) {
    const uint32_t t0 = t & ~63u;
    if (t0 >= d.nslots) return;  // the whole wave is past the last slot
    const uint32_t sIdx = d.chunkScene[t0 >> 8];  // a wave never spans two scenes
    if (gated && !d.gate[sIdx]) return;
    const uint32_t i = work_slot(d, t);
    const bool valid = t < d.nslots && i < d.scenes[sIdx].end;
    float C = 0.0f;
    if (valid) {
        const float4 xi4 = Xs[i];
        const f3 xi = make3(xi4);
        const uint32_t n = d.cntW[t];
        const uint4* lp = d.nbrW + list_base(t, P.maxN);
        float rho = xi4.w * P.w0;
        f3 gi{0.0f, 0.0f, 0.0f};
        float sum = 0.0f;
        const float negInv = -P.invRest;
        walk_list<float4>(
            lp, n, [&](uint32_t j) { return ld(Xs, j); },
            [&](const float4& xj4) {
                f3 sep = xi - make3(xj4);
                float d2 = dot3(sep, sep);
                // m_j == 1 under kFast: every product by m_j is the other factor itself
                rho = rho + (kFast ? poly6(d2, P) : xj4.w * poly6(d2, P));
                // a zero gradient would add zeros to gi and sum, which change neither (both start
                // at +0, and x + (-x) rounds to +0, so neither is ever -0): it is skipped
                f3 g;
                if (grad_spiky_nonzero<kFast>(sep, P, g)) {
                    if (kFast) {
                        gi = gi + g;
                        f3 gj = negInv * g;
                        sum = sum + dot3(gj, gj);
                    } else {
                        gi = gi + xj4.w * g;
                        f3 gj = (negInv * xj4.w) * g;
                        sum = sum + dot3(gj, gj);
                    }
                }
            });
        if (kBnd) {
            float rb = 0.0f;
            f3 gb{0.0f, 0.0f, 0.0f};
            if (boundary_near(P.bd, xi.x, xi.y, xi.z, P.h2)) {
                boundary_visit(P.bd, xi.x, xi.y, xi.z, P.h, P.h2,
                               [&](uint32_t idx, float s0, float s1, float s2, float d2) {
                                   const float psi = P.bd.psi[idx];
                                   rb = rb + psi * poly6(d2, P);
                                   gb = gb + psi * grad_spiky<kFast>(f3{s0, s1, s2}, P);
                               });
                d.Bg[i] = make_float4(gb.x, gb.y, gb.z, 0.0f);
            }
            rho = rho + rb;
            gi = gi + gb;
        }
        gi = gi * P.invRest;
        C = rho * P.invRest - 1.0f;
        d.dens[i] = rho;
        const float lam = -C / ((sum + dot3(gi, gi)) + P.relax);
        d.lam[i] = lam;
        // unit masses: (x*, lambda) side by side for the position-delta pass, in W (free until
        // the vorticity pass)
        if (kFast) d.W[i] = make_float4(xi.x, xi.y, xi.z, lam);
    }
    if (residSlot >= 0) wave_resid_partial(d, residSlot, t, valid ? positive_part(C) : 0.0f);
}

// One workgroup per virtual workgroup (the loop runs once); the loop form lets a launch use any
// grid size.
template <int BS, bool kFast, bool kBnd>
__global__ __launch_bounds__(BS) void k_density_lambda(DeviceState d, SphParams P,
                                                           const float4* __restrict__ Xs,
                                                           int residSlot, int gated, uint32_t nvb) {
    for (uint32_t vb = blockIdx.x; vb < nvb; vb += gridDim.x)
        density_lambda_slot<BS, kFast, kBnd>(d, P, Xs, residSlot, gated, xcd_slot_of<BS>(vb, nvb));
}

// The list build followed, in the same thread, by the first Jacobi iteration's density/lambda
// sweep at x*_pred over the list just written (each lane reads back only its own list, so no
// synchronisation is needed; x*_pred is read-only in this kernel).  Results are those of
// k_neighbours then k_density_lambda; the neighbours' positions are still in the caches.
// Physical work order only (no balanced order), no boundary samples.
// Held to the list build's 7 waves per SIMD (at most 72 VGPRs).
template <int BS, bool kFast>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(7)))
void k_neighbours_density(DeviceState d, SphParams P) {
    const uint32_t p = xcd_slot<BS>();
    neighbours_slot(d, P, p);
    density_lambda_slot<BS, kFast, false>(d, P, d.XsA, -1, 0, p);
}

// Features of the learned warm start at x*_pred, one neighbour sweep (list order):
//   0 C = rho / rho0 - 1 and 1 lambda, as the first Jacobi iteration computes them
//   2 neighbour count, 3 lambda of the previous step
//   4-6 (x*_pred - x) / h
//   7-9 2 lambda (1/rho0) sum_j m_j grad W_ij / h: the correction if every neighbour had this
//       particle's lambda
//   10-12 (1/rho0) sum_j m_j s_corr,ij grad W_ij / h: the artificial-pressure correction
//   13-15 wall proximity per axis, clamp((x - lo) / h, 0, 2) - clamp((hi - x) / h, 0, 2)
__device__ __forceinline__ void features_slot(const DeviceState& d, const SphParams& P, int record,
                                              uint32_t t) {
    const uint32_t t0 = t & ~63u;
    if (t0 >= d.nslots) return;
    const uint32_t sIdx = d.chunkScene[t0 >> 8];
    const SceneDev& sc = d.scenes[sIdx];
    const uint32_t i = work_slot(d, t);
    const bool valid = t < d.nslots && i < sc.end;
    float C = 0.0f;
    if (valid) {
        const float4 xi4 = d.XsA[i];
        const f3 xi = make3(xi4);
        const uint32_t n = d.cntW[t];
        const uint4* lp = d.nbrW + list_base(t, P.maxN);
        float rho = xi4.w * P.w0;
        f3 gi{0.0f, 0.0f, 0.0f}, ac{0.0f, 0.0f, 0.0f};
        float sum = 0.0f;
        const float negInv = -P.invRest;
        const Recip Rdq = recip(P.wdq);
        walk_list<float4>(
            lp, n, [&](uint32_t j) { return ld(d.XsA, j); },
            [&](const float4& xj4) {
                f3 sep = xi - make3(xj4);
                float d2 = dot3(sep, sep);
                const float w = poly6(d2, P);
                rho = rho + xj4.w * w;
                f3 g = grad_spiky(sep, P);
                gi = gi + xj4.w * g;
                f3 gj = (negInv * xj4.w) * g;
                sum = sum + dot3(gj, gj);
                if (P.corrOn) {
                    const float lc = P.negCorrK * corr_pow(div_rn(w, Rdq), P);
                    ac = ac + (lc * xj4.w) * g;
                }
            });
        gi = gi * P.invRest;
        C = rho * P.invRest - 1.0f;
        const float lam1 = -C / ((sum + dot3(gi, gi)) + P.relax);
        const float invH = 1.0f / P.h;
        const float4 x0 = d.X0[i];
        const f3 sg = gi * (2.0f * lam1 * invH);
        const f3 sa = ac * (P.invRest * invH);
        float wall[3];
        const float xa[3] = {xi.x, xi.y, xi.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float lo = fminf(fmaxf((xa[a] - sc.bmin[a]) * invH, 0.0f), 2.0f);
            const float hi = fminf(fmaxf((sc.bmax[a] - xa[a]) * invH, 0.0f), 2.0f);
            wall[a] = lo - hi;
        }
        float4* f = reinterpret_cast<float4*>(d.feat + (size_t)i * kMlpFeatures);
        f[0] = make_float4(C, lam1, (float)n, d.Tlam[i]);
        f[1] = make_float4((xi.x - x0.x) * invH, (xi.y - x0.y) * invH, (xi.z - x0.z) * invH, sg.x);
        f[2] = make_float4(sg.y, sg.z, sa.x, sa.y);
        f[3] = make_float4(sa.z, wall[0], wall[1], wall[2]);
        if (record) d.XsP[i] = xi4;
    }
    wave_resid_partial(d, 0, t, valid ? positive_part(C) : 0.0f);
}

template <int BS>
__global__ __launch_bounds__(BS) void k_features(DeviceState d, SphParams P, int record) {
    features_slot(d, P, record, xcd_slot<BS>());
}

// The list build followed, in the same thread, by the feature sweep over the list just written
// (as k_neighbours_density does for the density/lambda sweep).
template <int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(7)))
void k_neighbours_features(DeviceState d, SphParams P, int record) {
    const uint32_t p = xcd_slot<BS>();
    neighbours_slot(d, P, p);
    features_slot(d, P, record, p);
}

// Residual statistics per scene in two fixed-order stages (deterministic sums, no atomics).
// Stage 1: workgroup (b, s) reduces up to kResidWavesPerBlock wave partials of scene s for each
// check in slotMask into one block partial.
constexpr int kReduceBlock = 256;
constexpr uint32_t kResidWavesPerBlock = 8u * kReduceBlock;

__device__ __forceinline__ void block_max_sum(float& m, float& t, float* sm, float* st) {
    for (int off = 32; off > 0; off >>= 1) {
        m = fmaxf(m, __shfl_xor(m, off, 64));
        t += __shfl_xor(t, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        sm[w] = m;
        st[w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kReduceBlock / 64; ++k) {
            sm[0] = fmaxf(sm[0], sm[k]);
            st[0] += st[k];
        }
    }
    __syncthreads();
    m = sm[0];
    t = st[0];
    __syncthreads();
}

__global__ __launch_bounds__(kReduceBlock) void k_resid_blocks(DeviceState d, int slotMask) {
    __shared__ float sm[kReduceBlock / 64], st[kReduceBlock / 64];
    const uint32_t b = blockIdx.x, s = blockIdx.y;
    const SceneDev& sc = d.scenes[s];
    const uint32_t w0 = (sc.begin >> 6) + b * kResidWavesPerBlock;
    const uint32_t w1 = min((sc.end + 63) >> 6, w0 + kResidWavesPerBlock);
    for (int k = 0; k < 3; ++k) {
        if (!((slotMask >> k) & 1)) continue;
        float m = 0.0f, t = 0.0f;
        const float2* part = d.residPart + (size_t)k * d.nWaves;
#pragma unroll
        for (uint32_t r = 0; r < kResidWavesPerBlock / kReduceBlock; ++r) {
            const uint32_t w = w0 + r * kReduceBlock + threadIdx.x;
            if (w < w1) {
                const float2 p = part[w];
                m = fmaxf(m, p.x);
                t += p.y;
            }
        }
        block_max_sum(m, t, sm, st);
        if (threadIdx.x == 0)
            d.residBlk[((size_t)s * d.residBlocks + b) * 3 + k] = make_float2(m, t);
    }
}

// Stage 2, one wave per scene: the scene's block partials into its record; with gate, also the
// fallback decision from checks 0 (x*_pred) and 1 (x*_w).
__global__ __launch_bounds__(64) void k_resid_reduce(DeviceState d, int slotMask, int gate,
                                                     int statistic, float ratio, float tolerance) {
    const uint32_t s = blockIdx.x;
    const SceneDev& sc = d.scenes[s];
    const uint32_t nb = ((((sc.end + 63) >> 6) - (sc.begin >> 6)) + kResidWavesPerBlock - 1) /
                        kResidWavesPerBlock;
    unsigned int* rec = d.resid + s * kResidStride;
    for (int k = 0; k < 3; ++k) {
        if (!((slotMask >> k) & 1)) continue;
        float m = 0.0f, t = 0.0f;
        for (uint32_t b = __lane_id(); b < nb; b += 64) {
            const float2 p = d.residBlk[((size_t)s * d.residBlocks + b) * 3 + k];
            m = fmaxf(m, p.x);
            t += p.y;
        }
        for (int off = 32; off > 0; off >>= 1) {
            m = fmaxf(m, __shfl_xor(m, off, 64));
            t += __shfl_xor(t, off, 64);
        }
        if (__lane_id() == 0) {
            rec[k] = __float_as_uint(m);
            rec[kResidSum + k] = __float_as_uint(t);
        }
    }
    if (gate && __lane_id() == 0) {
        float rp, rw;
        if (statistic == 0) {
            rp = __uint_as_float(rec[0]);
            rw = __uint_as_float(rec[1]);
        } else {
            const float inv = 1.0f / (float)max(1u, sc.end - sc.begin);
            rp = __uint_as_float(rec[kResidSum]) * inv;
            rw = __uint_as_float(rec[kResidSum + 1]) * inv;
        }
        const uint32_t g = (rw <= ratio * rp + tolerance) ? 0u : 1u;
        d.gate[s] = g;
        d.fallbacks[s] += g;
    }
}

// Targets: x* before the last Jacobi iteration (XsLast) relative to x*_pred, in units of h, and
// the last iteration's lambda.
__global__ void k_targets(DeviceState d, SphParams P, const float4* __restrict__ XsLast) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nslots) return;
    if (i >= scene_of(d, i).end) return;
    const float invH = 1.0f / P.h;
    const float4 x = XsLast[i], p = d.XsP[i];
    reinterpret_cast<float4*>(d.targ)[i] =
        make_float4((x.x - p.x) * invH, (x.y - p.y) * invH, (x.z - p.z) * invH, d.lam[i]);
}

__global__ void k_sample(DeviceState d, uint64_t seed, uint32_t count, float* __restrict__ fo,
                         float* __restrict__ to) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    const SceneDev& sc = d.scenes[0];
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(b + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const uint32_t slot = sc.begin + (uint32_t)(z % (uint64_t)(sc.end - sc.begin));
    const float4* f = reinterpret_cast<const float4*>(d.feat + (size_t)slot * kMlpFeatures);
    float4* g = reinterpret_cast<float4*>(fo + (size_t)b * kMlpFeatures);
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = f[k];
    reinterpret_cast<float4*>(to)[b] = reinterpret_cast<const float4*>(d.targ)[slot];
}

// resolve_collision over the six walls in order (IntegrationCUDA.cu:51-102), restitution 0 and
// friction 0.95 (PBFSolver.cpp:64); vector form as written, so every rounding matches.
__device__ __forceinline__ f3 damp_walls(f3 x, f3 v, const SceneDev& sc) {
    const float minDist = 0.025f, restitution = 0.0f, friction = 0.95f;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int a = k >> 1;
        const float sgn = (k & 1) ? -1.0f : 1.0f;
        const float b = (k & 1) ? sc.bmax[a] : sc.bmin[a];
        f3 pt{a == 0 ? b : 0.0f, a == 1 ? b : 0.0f, a == 2 ? b : 0.0f};
        f3 nm{a == 0 ? sgn : 0.0f, a == 1 ? sgn : 0.0f, a == 2 ? sgn : 0.0f};
        float distance = dot3(x - pt, nm);
        float approaching = dot3(v, nm);
        if (distance < minDist) {
            f3 vn = approaching * nm;
            f3 vt = v - vn;
            if (approaching < 0.0f) {
                v = (-restitution) * vn + (1.0f - friction) * vt;
            } else if (fabsf(approaching) < 1e-5f) {
                v = (1.0f - friction) * vt;
            }
        }
    }
    return v;
}

// kernel_correct_position with handle_particle_collision (ConstraintSolverCUDA.cu:136-169) after
// the neighbour sum: dp becomes sum / rho0 and the corrected position is clamped to the walls.
__device__ __forceinline__ f3 correct_position(f3 xi, f3& dp, const SphParams& P, const SceneDev& sc) {
    dp = dp * P.invRest;
    f3 q = xi + dp;
    const float minDist = 0.025f, stiff = 0.5f;
    float qa[3] = {q.x, q.y, q.z};
    float c[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (qa[a] < sc.bmin[a] + minDist) c[a] = c[a] + stiff * (sc.bmin[a] + minDist - qa[a]);
        if (qa[a] > sc.bmax[a] - minDist) c[a] = c[a] + stiff * (sc.bmax[a] - minDist - qa[a]);
    }
    return q + f3{c[0], c[1], c[2]};
}

// kernel_update_position_and_velocity (IntegrationCUDA.cu:38-49) and the wall damping at the last
// iteration's corrected position q, and the step's writes of slot i.
template <bool kApplyDelta>
__device__ __forceinline__ void finish_step(const DeviceState& d, const StepParams& S,
                                            const SceneDev& sc, uint32_t i, float4 xi4, f3 q, f3 dp) {
    const float4 x0 = d.X0[i];
    f3 v = div3_rn(q - make3(x0), S.dt);
    d.VS[i] = make_float4(v.x, v.y, v.z, 0.0f);  // the record's new_velocity (IntegrationCUDA.cu:45)
    v = damp_walls(q, v, sc);
    d.X[i] = make_float4(q.x, q.y, q.z, xi4.w);
    d.V[i] = make_float4(v.x, v.y, v.z, 0.0f);
    d.id[i] = d.Tid[i];
    d.lid[i] = d.lidT[i];
    if (kApplyDelta) {  // the record's position_delta (ConstraintSolverCUDA.cu:129)
        d.DP[i] = make_float4(dp.x, dp.y, dp.z, 0.0f);
    } else {  // no iteration ran: density, lambda and the position delta carry over
        d.dens[i] = d.Tdens[i];
        d.lam[i] = d.Tlam[i];
        d.DP[i] = d.TDP[i];
    }
}

// kernel_calculate_position_delta (ConstraintSolverCUDA.cu:99-130), then kernel_correct_position
// with handle_particle_collision (ConstraintSolverCUDA.cu:136-169).  x* is double-buffered: the
// kernel reads XsIn and writes XsOut, so every particle sees the positions of the same iteration.
// kFinal: also kernel_update_position_and_velocity (IntegrationCUDA.cu:38-49) and the wall
// damping, writing the step's positions X, velocities V, ids and the two order maps.
// gateMode 1: only scenes whose gate is set run; 2: scenes whose gate is clear read XsAlt.
template <int BS, bool kFinal, bool kApplyDelta, bool kFast, bool kBnd>
__device__ __forceinline__ void delta_correct_slot(const DeviceState& d, const SphParams& P,
                                                   const StepParams& S,
                                                   const float4* __restrict__ XsSet,
                                                   float4* __restrict__ XsOut,
                                                   const float4* __restrict__ XsAlt, int gateMode,
                                                   uint32_t t) {
    if (t >= d.nslots) return;
    const uint32_t i = work_slot(d, t);
    const uint32_t sIdx = d.chunkScene[i >> 8];
    const SceneDev& sc = d.scenes[sIdx];
    if (i >= sc.end) return;
    if (gateMode == 1 && !d.gate[sIdx]) return;
    const float4* __restrict__ XsIn = (gateMode == 2 && !d.gate[sIdx]) ? XsAlt : XsSet;
    const float4 xi4 = XsIn[i];
    const f3 xi = make3(xi4);
    f3 q = xi;
    f3 dp{0.0f, 0.0f, 0.0f};
    if (kApplyDelta) {
        const float li = d.lam[i];
        const uint32_t n = d.cntW[t];
        const uint4* lp = d.nbrW + list_base(t, P.maxN);
        const Recip Rdq = recip(P.wdq);
        // unit masses: x*_j and lambda_j from one float4 written by the density pass (W)
        const auto load = [&](uint32_t j) {
            if (kFast) {
                const float4 a = ld(d.W, j);
                return XLam{a, a.w};
            }
            return XLam{ld(XsIn, j), ld(d.lam, j)};
        };
        // the walk once with and once without the artificial pressure s_corr, chosen per launch
        const auto walk = [&](auto corrOn) {
            constexpr bool kCorr = decltype(corrOn)::value;
            walk_list<XLam, kBatchDelta>(lp, n, load, [&](const XLam& nb) {
                f3 sep = xi - make3(nb.x);
                float d2 = dot3(sep, sep);
                float coef = li + nb.lam;
                if (kCorr) {  // w = poly6 >= +0
                    const float w = poly6(d2, P);
                    coef = coef + P.negCorrK * corr_pow(kFast ? div_rn_small_nonneg(w, Rdq) : div_rn(w, Rdq), P);
                }
                const float c = kFast ? coef : coef * nb.x.w;  // m_j == 1 under kFast
                f3 g;
                if (__builtin_expect(grad_spiky_nonzero<kFast>(sep, P, g), 1)) {
                    dp = dp + c * g;
                } else {  // the zero gradient's term c * 0: zeros, or NaN where c is not finite
                    const f3 z{0.0f, 0.0f, 0.0f};
                    dp = dp + c * z;
                }
            });
        };
        if (P.corrOn)
            walk(std::true_type{});
        else
            walk(std::false_type{});
        if (kBnd) {  // lambda_i times the boundary gradient sum of the density pass (same x*)
            f3 gb{0.0f, 0.0f, 0.0f};
            if (boundary_near(P.bd, xi.x, xi.y, xi.z, P.h2)) gb = make3(d.Bg[i]);
            dp = dp + li * gb;
        }
        q = correct_position(xi, dp, P, sc);
    }
    if (!kFinal) {
        XsOut[i] = make_float4(q.x, q.y, q.z, xi4.w);
        return;
    }
    finish_step<kApplyDelta>(d, S, sc, i, xi4, q, dp);
}

// Unit masses: held to 8 waves per SIMD (its scalar registers, over the 8-wave budget, spill to
// vector lanes); the general kernels would spill vector registers there.
template <int BS, bool kFinal, bool kApplyDelta, bool kFast, bool kBnd>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(kFast && !kBnd ? 8 : 1)))
void k_delta_correct(DeviceState d, SphParams P, StepParams S,
                                                          const float4* __restrict__ XsSet,
                                                          float4* __restrict__ XsOut,
                                                          const float4* __restrict__ XsAlt,
                                                          int gateMode, uint32_t nvb) {
    for (uint32_t vb = blockIdx.x; vb < nvb; vb += gridDim.x)
        delta_correct_slot<BS, kFinal, kApplyDelta, kFast, kBnd>(d, P, S, XsSet, XsOut, XsAlt, gateMode,
                                                                 xcd_slot_of<BS>(vb, nvb));
}

// kernel_compute_vorticities (IntegrationCUDA.cu:104-128).  W.w keeps |omega| for the next pass,
// which also finds it beside the position in XsB (free between the last Jacobi iteration and the
// next step's): (x, |omega|), one gather per neighbour there.
template <bool kFast>
__device__ __forceinline__ void vorticity_slot(const DeviceState& d, const SphParams& P, uint32_t t) {
    if (t >= d.nslots) return;
    const uint32_t i = work_slot(d, t);
    if (i >= scene_of(d, i).end) return;
    const f3 xi = make3(d.X[i]);
    const f3 vi = make3(d.V[i]);
    const uint32_t n = d.cntW[t];
    const uint4* lp = d.nbrW + list_base(t, P.maxN);
    f3 w{0.0f, 0.0f, 0.0f};
    walk_list<XV, kBatchVort>(
        lp, n, [&](uint32_t j) { return XV{ld(d.X, j), ld(d.V, j)}; },
        [&](const XV& nb) {
            f3 vij = make3(nb.v) - vi;
            f3 sep = xi - make3(nb.x);
            const float mj = kFast ? -1.0f : -nb.x.w;  // m_j == 1 under kFast
            f3 g;
            if (__builtin_expect(grad_spiky_nonzero<kFast>(sep, P, g), 1))
                w = w + mj * cross3(vij, g);
            else  // the zero gradient's term: zeros, or NaN where v_ij is not finite
                w = w + mj * cross3(vij, f3{0.0f, 0.0f, 0.0f});
        });
    const float wl = length3(w);
    d.W[i] = make_float4(w.x, w.y, w.z, wl);
    d.XsB[i] = make_float4(xi.x, xi.y, xi.z, wl);
}

template <int BS, bool kFast>
__global__ __launch_bounds__(BS) void k_vorticity(DeviceState d, SphParams P) {
    vorticity_slot<kFast>(d, P, xcd_slot<BS>());
}

// The confinement force from the summed eta (IntegrationCUDA.cu:146-163) and the write of Vc.
__device__ __forceinline__ void confine_finish(const DeviceState& d, const SphParams& P,
                                               const StepParams& S, uint32_t i, float xi4w, float4 wi,
                                               f3 eta, float invD) {
    eta = eta * invD;
    float el = length3(eta);
    float4 v = d.V[i];
    if (!(el < 1e-5f)) {
        f3 N = div3_rn(eta, el);
        f3 f = P.vortEps * cross3(N, make3(wi));
        f3 vn = make3(v) + S.dt * f;
        v = make_float4(vn.x, vn.y, vn.z, 0.0f);
    }
    v.w = xi4w / d.dens[i];
    d.Vc[i] = v;
}

// kernel_apply_vorticity_confinement (IntegrationCUDA.cu:130-165), writing Vc: the confined
// velocity, with the particle's m / rho in .w for the XSPH pass (the quotient the reference
// forms per neighbour, IntegrationCUDA.cu:184, computed once per particle instead).
template <bool kFast>
__device__ __forceinline__ void confinement_slot(const DeviceState& d, const SphParams& P,
                                                 const StepParams& S, uint32_t t) {
    if (t >= d.nslots) return;
    const uint32_t i = work_slot(d, t);
    if (i >= scene_of(d, i).end) return;
    const float4 xi4 = d.X[i];
    const f3 xi = make3(xi4);
    const float xi4w = xi4.w;
    const float4 wi = d.W[i];
    const float invD = 1.0f / d.dens[i];
    const uint32_t n = d.cntW[t];
    const uint4* lp = d.nbrW + list_base(t, P.maxN);
    f3 eta{0.0f, 0.0f, 0.0f};
    // XW.x: (x_j, |omega_j|); XW.w: m_j (read only without unit masses)
    walk_list<XW, kBatchConf>(
        lp, n, [&](uint32_t j) { return XW{ld(d.XsB, j), kFast ? 1.0f : ld(d.X, j).w}; },
        [&](const XW& nb) {
            float diff = wi.w - nb.x.w;
            f3 sep = xi - make3(nb.x);
            const float c = kFast ? diff : nb.w * diff;  // m_j == 1 under kFast
            f3 g;
            if (__builtin_expect(grad_spiky_nonzero<kFast>(sep, P, g), 1))
                eta = eta + c * g;
            else  // the zero gradient's term c * 0: zeros, or NaN where c is not finite
                eta = eta + c * f3{0.0f, 0.0f, 0.0f};
        });
    confine_finish(d, P, S, i, xi4w, wi, eta, invD);
}

template <int BS, bool kFast>
__global__ __launch_bounds__(BS) void k_confinement(DeviceState d, SphParams P, StepParams S) {
    confinement_slot<kFast>(d, P, S, xcd_slot<BS>());
}

// kernel_apply_xsph_viscosity (IntegrationCUDA.cu:167-195), Jacobi: reads Vc, writes V.
__device__ __forceinline__ void xsph_slot(const DeviceState& d, const SphParams& P, uint32_t t) {
    if (t >= d.nslots) return;
    const uint32_t i = work_slot(d, t);
    if (i >= scene_of(d, i).end) return;
    const f3 xi = make3(d.X[i]);
    const f3 vi = make3(d.Vc[i]);
    const uint32_t n = d.cntW[t];
    const uint4* lp = d.nbrW + list_base(t, P.maxN);
    f3 dv{0.0f, 0.0f, 0.0f};
    walk_list<XV, kBatchXsph>(
        lp, n, [&](uint32_t j) { return XV{ld(d.X, j), ld(d.Vc, j)}; },
        [&](const XV& nb) {
            f3 vij = make3(nb.v) - vi;
            f3 sep = xi - make3(nb.x);
            float w = poly6(dot3(sep, sep), P);
            dv = dv + (nb.v.w * vij) * w;  // nb.v.w = m_j / rho_j
        });
    f3 vn = vi + P.visc * dv;
    d.V[i] = make_float4(vn.x, vn.y, vn.z, 0.0f);
}

template <int BS>
__global__ __launch_bounds__(BS) void k_xsph(DeviceState d, SphParams P) {
    xsph_slot(d, P, xcd_slot<BS>());
}

// ---- work queue: the sweeps as persistent kernels over 256-slot tiles, heaviest first -------
// A batch of uneven scenes (and the uneven lists of one scene) leave each sweep's last waves
// running on a few CUs.  With the work queue, the tiles of each XCD's contiguous range (the range
// xcd_slot gives it) are ordered by the sum of their neighbour counts, heaviest first, once per
// step after the list build; each sweep then runs a resident grid whose workgroups take tiles of
// their XCD's queue until it is empty: in the dynamic mode the next free position (through
// counters), in the static mode every R-th position from their own rank (R resident workgroups
// per XCD).  Which workgroup runs which tile decides nothing: every slot is computed as in the
// one-launch sweep.
constexpr uint32_t kQueueTile = kQueueTileSlots;
constexpr uint32_t kQueueBuckets = 512;
constexpr uint32_t kQueueThreads = 1024;

// Per tile: the sum of its slots' neighbour counts (one wave per tile).
__global__ __launch_bounds__(256) void k_tile_cost(DeviceState d, uint32_t ntiles,
                                                   uint32_t* __restrict__ cost) {
    const uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (tile >= ntiles) return;  // whole waves
    uint32_t c = 0;
    for (uint32_t k = __lane_id(); k < kQueueTile; k += 64u) {
        const uint32_t s = tile * kQueueTile + k;
        if (s < d.nslots) c += d.cntW[s];
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if (__lane_id() == 0) cost[tile] = c;
}

// One workgroup per XCD x: its tiles [x * T8, min((x + 1) * T8, ntiles)) by cost class, heaviest
// first (a counting sort over kQueueBuckets classes of cost >> shift; the order within a class is
// whatever the LDS atomics give), their number, and the XCD's pass counters cleared.
__global__ __launch_bounds__(kQueueThreads) void k_tile_queue(const uint32_t* __restrict__ cost,
                                                             uint32_t ntiles, uint32_t T8,
                                                             uint32_t shift, uint32_t* __restrict__ queue,
                                                             uint32_t* __restrict__ qcount,
                                                             uint32_t* __restrict__ heads,
                                                             uint32_t passes) {
    __shared__ uint32_t start[kQueueBuckets];
    __shared__ uint32_t wsum[kQueueThreads / 64];
    const uint32_t x = blockIdx.x, tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    const uint32_t t0 = x * T8, t1 = min(t0 + T8, ntiles);
    for (uint32_t k = tid; k < kQueueBuckets; k += kQueueThreads) start[k] = 0u;
    for (uint32_t k = tid; k < passes * kQueueSubs; k += kQueueThreads)
        heads[(((size_t)(k / kQueueSubs) * kXcds + x) * kQueueSubs + k % kQueueSubs) * kQueueHeadStride] = 0u;
    __syncthreads();
    for (uint32_t t = t0 + tid; t < t1; t += kQueueThreads)
        atomicAdd(&start[min(cost[t] >> shift, kQueueBuckets - 1u)], 1u);
    __syncthreads();
    // exclusive prefix in descending class order: thread r < kQueueBuckets holds class B - 1 - r
    const uint32_t v = tid < kQueueBuckets ? start[kQueueBuckets - 1u - tid] : 0u;
    uint32_t inc = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, 64);
        if ((int)lane >= off) inc += y;
    }
    if (lane == 63u) wsum[wave] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
    __syncthreads();
    if (tid < kQueueBuckets) start[kQueueBuckets - 1u - tid] = base + inc - v;
    __syncthreads();
    for (uint32_t t = t0 + tid; t < t1; t += kQueueThreads) {
        const uint32_t pos = atomicAdd(&start[min(cost[t] >> shift, kQueueBuckets - 1u)], 1u);
        queue[t0 + pos] = t;
    }
    if (tid == 0) qcount[x] = t1 > t0 ? t1 - t0 : 0u;
}

// Counter of sub-queue `sub` of XCD x in pass `pass` (the dynamic mode).
__device__ __forceinline__ uint32_t* queue_head(const DeviceState& d, uint32_t pass, uint32_t x, uint32_t sub) {
    return d.heads + (((size_t)pass * kXcds + x) * kQueueSubs + sub) * kQueueHeadStride;
}

// The persistent loop: workgroup b serves XCD b % 8 (where the hardware places it); f(slot) runs
// once per thread and tile.  Static: rank r = b / 8 of R takes positions r, r + R, ...  Dynamic:
// the positions sub + 8 m form eight sub-queues (one counter each, so that no single counter
// takes every request); the workgroup takes the next m of sub-queue (b / 8) % 8 until it is empty,
// then of the seven others in turn.  Every workgroup leaves once each sub-queue is empty.
template <class F>
__device__ __forceinline__ void queue_loop(const DeviceState& d, uint32_t pass, uint32_t& sq, F f) {
    const uint32_t x = blockIdx.x % kXcds;
    const uint32_t n = d.qcount[x];
    const uint32_t* q = d.queue + (size_t)x * d.queueT8;
    if (d.queueMode == 2) {
        const uint32_t R = gridDim.x / kXcds;
        for (uint32_t k = blockIdx.x / kXcds; k < n; k += R) f(q[k] * kQueueTile + threadIdx.x);
        return;
    }
    const uint32_t sub0 = (blockIdx.x / kXcds) % kQueueSubs;
    for (uint32_t u = 0; u < kQueueSubs; ++u) {
        const uint32_t sub = (sub0 + u) % kQueueSubs;
        const uint32_t len = n > sub ? (n - sub + kQueueSubs - 1u) / kQueueSubs : 0u;
        uint32_t* head = queue_head(d, pass, x, sub);
        for (;;) {
            if (threadIdx.x == 0) sq = atomicAdd(head, 1u);
            __syncthreads();
            const uint32_t m = sq;
            __syncthreads();  // every thread has read sq before thread 0 takes the next position
            if (m >= len) break;
            f(q[sub + kQueueSubs * m] * kQueueTile + threadIdx.x);
        }
    }
}

template <bool kFast, bool kBnd>
__global__ __launch_bounds__(kQueueTile) void k_density_lambda_queue(DeviceState d, SphParams P,
                                                                     const float4* __restrict__ Xs,
                                                                     uint32_t pass) {
    __shared__ uint32_t sq;
    queue_loop(d, pass, sq, [&](uint32_t t) { density_lambda_slot<kQueueTile, kFast, kBnd>(d, P, Xs, -1, 0, t); });
}

template <bool kFinal, bool kFast, bool kBnd>
__global__ __launch_bounds__(kQueueTile) __attribute__((amdgpu_waves_per_eu(8)))
void k_delta_correct_queue(DeviceState d, SphParams P, StepParams S,
                                                                    const float4* __restrict__ XsIn,
                                                                    float4* __restrict__ XsOut, uint32_t pass) {
    __shared__ uint32_t sq;
    queue_loop(d, pass, sq, [&](uint32_t t) {
        delta_correct_slot<kQueueTile, kFinal, true, kFast, kBnd>(d, P, S, XsIn, XsOut, nullptr, 0, t);
    });
}

template <bool kFast>
__global__ __launch_bounds__(kQueueTile) void k_vorticity_queue(DeviceState d, SphParams P, uint32_t pass) {
    __shared__ uint32_t sq;
    queue_loop(d, pass, sq, [&](uint32_t t) { vorticity_slot<kFast>(d, P, t); });
}

template <bool kFast>
__global__ __launch_bounds__(kQueueTile) void k_confinement_queue(DeviceState d, SphParams P, StepParams S,
                                                                  uint32_t pass) {
    __shared__ uint32_t sq;
    queue_loop(d, pass, sq, [&](uint32_t t) { confinement_slot<kFast>(d, P, S, t); });
}

__global__ __launch_bounds__(kQueueTile) void k_xsph_queue(DeviceState d, SphParams P, uint32_t pass) {
    __shared__ uint32_t sq;
    queue_loop(d, pass, sq, [&](uint32_t t) { xsph_slot(d, P, t); });
}

// ---- lane groups: L lanes per particle, for scenes too small to fill the chip ------------------
// Thread T of a lane-group launch works on particle T / L (physical order).  The group's L lanes
// take the list entries L at a time, lane r entry k0 + r, kLaneBatch rounds of loads in flight,
// and compute each entry's terms; each lane stores its terms in its wave's rows of LDS, and the
// group's first lane adds the L terms of a round in entry order, so each sum is the one-lane
// sweep's sum, term for term and in list order (only that lane's sums are used).  Entries past the list's end contribute +0: no sum here starts at or reaches -0 (each
// starts at +0 or a positive value, and x + (-x) rounds to +0), so adding +0 changes nothing.
// Only plain steps (no learned gate, no residual check), the physical order and no wall samples.
constexpr int kLaneBatch = 4;

template <int L>
struct LaneGroup {
    uint32_t i;     // particle slot
    uint32_t r;     // lane within the group
    uint32_t base;  // first lane of the group within the wave
};

template <int L>
__device__ __forceinline__ LaneGroup<L> lane_group(uint32_t T) {
    static_assert(L >= 2 && L <= 64 && (L & (L - 1)) == 0, "a power of two that divides 64");
    return {T / (uint32_t)L, T % (uint32_t)L, __lane_id() & ~(uint32_t)(L - 1)};
}

__device__ __forceinline__ uint32_t list_entry(const uint4* __restrict__ lp, uint32_t k) {
    return reinterpret_cast<const uint32_t*>(lp + (size_t)(k >> 2) * 64u)[k & 3u];
}

// sm: this wave's NA rows of 64 floats in LDS.  LDS instructions of one wave complete in issue
// order, so a lane reads the terms the other lanes stored before the wave barrier; the second
// barrier keeps the next round's stores after this round's reads.
template <int L, int NA, class T, class Load, class Term>
__device__ __forceinline__ void walk_split(const uint4* __restrict__ lp, uint32_t n,
                                           const LaneGroup<L>& g, Load load, Term term,
                                           float (&acc)[NA], float* sm) {
    const uint32_t lane = __lane_id();
    for (uint32_t k0 = 0; k0 < n; k0 += (uint32_t)(L * kLaneBatch)) {
        T v[kLaneBatch];
        bool ok[kLaneBatch];
#pragma unroll
        for (int b = 0; b < kLaneBatch; ++b) {
            const uint32_t k = k0 + (uint32_t)(b * L) + g.r;
            ok[b] = k < n;
            v[b] = load(list_entry(lp, ok[b] ? k : 0u));  // past the end: a valid entry, discarded
        }
#pragma unroll
        for (int b = 0; b < kLaneBatch; ++b) {
            if (k0 + (uint32_t)(b * L) >= n) break;  // the same for every lane of the group
            float t[NA];
            term(v[b], t);
#pragma unroll
            for (int c = 0; c < NA; ++c) sm[c * 64 + lane] = ok[b] ? t[c] : 0.0f;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (g.r == 0) {
#pragma unroll
                for (int c = 0; c < NA; ++c)
#pragma unroll
                    for (int s = 0; s < L; ++s) acc[c] = acc[c] + sm[c * 64 + g.base + s];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// This wave's NA rows of 64 floats in the workgroup's LDS block.
template <int NA>
__device__ __forceinline__ float* wave_rows(float* block) {
    return block + (threadIdx.x >> 6) * (NA * 64);
}

template <int BS, int L, bool kFast>
__global__ __launch_bounds__(BS) void k_density_lambda_group(DeviceState d, SphParams P,
                                                             const float4* __restrict__ Xs) {
    __shared__ float lds[(BS / 64) * 5 * 64];
    const LaneGroup<L> g = lane_group<L>(xcd_slot<BS>());
    if (g.i >= d.nslots || g.i >= scene_of(d, g.i).end) return;  // whole groups
    const float4 xi4 = Xs[g.i];
    const f3 xi = make3(xi4);
    const float negInv = -P.invRest;
    float acc[5] = {xi4.w * P.w0, 0.0f, 0.0f, 0.0f, 0.0f};  // rho, grad_i C (x, y, z), sum |grad_j C|^2
    walk_split<L, 5, float4>(
        d.nbr4 + list_base(g.i, P.maxN), d.cnt[g.i], g, [&](uint32_t j) { return ld(Xs, j); },
        [&](const float4& xj4, float (&t)[5]) {
            f3 sep = xi - make3(xj4);
            float d2 = dot3(sep, sep);
            f3 gr = grad_spiky<kFast>(sep, P);
            if (kFast) {
                t[0] = poly6(d2, P);
                t[1] = gr.x; t[2] = gr.y; t[3] = gr.z;
                f3 gj = negInv * gr;
                t[4] = dot3(gj, gj);
            } else {
                t[0] = xj4.w * poly6(d2, P);
                f3 mg = xj4.w * gr;
                t[1] = mg.x; t[2] = mg.y; t[3] = mg.z;
                f3 gj = (negInv * xj4.w) * gr;
                t[4] = dot3(gj, gj);
            }
        },
        acc, wave_rows<5>(lds));
    if (g.r != 0) return;
    f3 gi{acc[1], acc[2], acc[3]};
    gi = gi * P.invRest;
    const float C = acc[0] * P.invRest - 1.0f;
    d.dens[g.i] = acc[0];
    const float lam = -C / ((acc[4] + dot3(gi, gi)) + P.relax);
    d.lam[g.i] = lam;
    if (kFast) d.W[g.i] = make_float4(xi.x, xi.y, xi.z, lam);
}

template <int BS, int L, bool kFinal, bool kFast>
__global__ __launch_bounds__(BS) void k_delta_correct_group(DeviceState d, SphParams P, StepParams S,
                                                            const float4* __restrict__ XsIn,
                                                            float4* __restrict__ XsOut) {
    __shared__ float lds[(BS / 64) * 3 * 64];
    const LaneGroup<L> g = lane_group<L>(xcd_slot<BS>());
    if (g.i >= d.nslots) return;
    const SceneDev& sc = scene_of(d, g.i);
    if (g.i >= sc.end) return;
    const float4 xi4 = XsIn[g.i];
    const f3 xi = make3(xi4);
    const float li = d.lam[g.i];
    const Recip Rdq = recip(P.wdq);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    walk_split<L, 3, XLam>(
        d.nbr4 + list_base(g.i, P.maxN), d.cnt[g.i], g,
        [&](uint32_t j) {
            if (kFast) {
                const float4 a = ld(d.W, j);
                return XLam{a, a.w};
            }
            return XLam{ld(XsIn, j), ld(d.lam, j)};
        },
        [&](const XLam& nb, float (&t)[3]) {
            f3 sep = xi - make3(nb.x);
            float d2 = dot3(sep, sep);
            float coef = li + nb.lam;
            if (P.corrOn) {
                const float w = poly6(d2, P);
                coef = coef + P.negCorrK * corr_pow(kFast ? div_rn_small(w, Rdq) : div_rn(w, Rdq), P);
            }
            f3 gr = grad_spiky<kFast>(sep, P);
            const f3 u = kFast ? coef * gr : (coef * nb.x.w) * gr;
            t[0] = u.x; t[1] = u.y; t[2] = u.z;
        },
        acc, wave_rows<3>(lds));
    if (g.r != 0) return;
    f3 dp{acc[0], acc[1], acc[2]};
    const f3 q = correct_position(xi, dp, P, sc);
    if (!kFinal) {
        XsOut[g.i] = make_float4(q.x, q.y, q.z, xi4.w);
        return;
    }
    finish_step<true>(d, S, sc, g.i, xi4, q, dp);
}

template <int BS, int L, bool kFast>
__global__ __launch_bounds__(BS) void k_vorticity_group(DeviceState d, SphParams P) {
    __shared__ float lds[(BS / 64) * 3 * 64];
    const LaneGroup<L> g = lane_group<L>(xcd_slot<BS>());
    if (g.i >= d.nslots || g.i >= scene_of(d, g.i).end) return;
    const f3 xi = make3(d.X[g.i]);
    const f3 vi = make3(d.V[g.i]);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    walk_split<L, 3, XV>(
        d.nbr4 + list_base(g.i, P.maxN), d.cnt[g.i], g,
        [&](uint32_t j) { return XV{ld(d.X, j), ld(d.V, j)}; },
        [&](const XV& nb, float (&t)[3]) {
            f3 vij = make3(nb.v) - vi;
            f3 sep = xi - make3(nb.x);
            const f3 c = cross3(vij, grad_spiky<kFast>(sep, P));
            const f3 u = (kFast ? -1.0f : -nb.x.w) * c;
            t[0] = u.x; t[1] = u.y; t[2] = u.z;
        },
        acc, wave_rows<3>(lds));
    if (g.r != 0) return;
    const f3 w{acc[0], acc[1], acc[2]};
    const float wl = length3(w);
    d.W[g.i] = make_float4(w.x, w.y, w.z, wl);
    d.XsB[g.i] = make_float4(xi.x, xi.y, xi.z, wl);
}

template <int BS, int L, bool kFast>
__global__ __launch_bounds__(BS) void k_confinement_group(DeviceState d, SphParams P, StepParams S) {
    __shared__ float lds[(BS / 64) * 3 * 64];
    const LaneGroup<L> g = lane_group<L>(xcd_slot<BS>());
    if (g.i >= d.nslots || g.i >= scene_of(d, g.i).end) return;
    const float4 xi4 = d.X[g.i];
    const f3 xi = make3(xi4);
    const float4 wi = d.W[g.i];
    const float invD = 1.0f / d.dens[g.i];
    float acc[3] = {0.0f, 0.0f, 0.0f};
    walk_split<L, 3, XW>(
        d.nbr4 + list_base(g.i, P.maxN), d.cnt[g.i], g,
        [&](uint32_t j) { return XW{ld(d.XsB, j), kFast ? 1.0f : ld(d.X, j).w}; },
        [&](const XW& nb, float (&t)[3]) {
            float diff = wi.w - nb.x.w;
            f3 sep = xi - make3(nb.x);
            const f3 gr = grad_spiky<kFast>(sep, P);
            const f3 u = (kFast ? diff : nb.w * diff) * gr;
            t[0] = u.x; t[1] = u.y; t[2] = u.z;
        },
        acc, wave_rows<3>(lds));
    if (g.r != 0) return;
    confine_finish(d, P, S, g.i, xi4.w, wi, f3{acc[0], acc[1], acc[2]}, invD);
}

template <int BS, int L>
__global__ __launch_bounds__(BS) void k_xsph_group(DeviceState d, SphParams P) {
    __shared__ float lds[(BS / 64) * 3 * 64];
    const LaneGroup<L> g = lane_group<L>(xcd_slot<BS>());
    if (g.i >= d.nslots || g.i >= scene_of(d, g.i).end) return;
    const f3 xi = make3(d.X[g.i]);
    const f3 vi = make3(d.Vc[g.i]);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    walk_split<L, 3, XV>(
        d.nbr4 + list_base(g.i, P.maxN), d.cnt[g.i], g,
        [&](uint32_t j) { return XV{ld(d.X, j), ld(d.Vc, j)}; },
        [&](const XV& nb, float (&t)[3]) {
            f3 vij = make3(nb.v) - vi;
            f3 sep = xi - make3(nb.x);
            float w = poly6(dot3(sep, sep), P);
            const f3 u = (nb.v.w * vij) * w;
            t[0] = u.x; t[1] = u.y; t[2] = u.z;
        },
        acc, wave_rows<3>(lds));
    if (g.r != 0) return;
    f3 vn = vi + P.visc * f3{acc[0], acc[1], acc[2]};
    d.V[g.i] = make_float4(vn.x, vn.y, vn.z, 0.0f);
}

// Density at the current positions over every particle within h, each counted once: the walk of
// k_neighbours with the hashing cell size, each distinct bucket of the stencil visited once (no
// duplicates from stencil hash collisions) and no neighbour cap.  Accumulates, per scene, the
// maximum and the sum of max(0, rho / rho0 - 1) (the density-constraint error).
template <int BS>
__global__ __launch_bounds__(BS) void k_density_probe(DeviceState d, SphParams P,
                                                          float* __restrict__ rhoOut,
                                                          unsigned int* __restrict__ errMaxBits,
                                                          double* __restrict__ errSum) {
    const uint32_t i = xcd_slot<BS>();
    float err = 0.0f;
    uint32_t scene = 0;
    bool active = false;
    if (i < d.nslots) {
        scene = d.chunkScene[i >> 8];
        const SceneDev& sc = d.scenes[scene];
        if (i < sc.end) {
            active = true;
            const uint32_t self = d.lidT[i];
            const float4 xi4 = d.XsA[i];
            const f3 xi = make3(xi4);
            int cx, cy, cz;
            discretize(xi, P.cellHash, cx, cy, cz);
            float rho = xi4.w * P.w0;
            uint32_t hs[27];
#pragma unroll
            for (int q = 0; q < 27; ++q)
                hs[q] = cell_hash((int)((uint32_t)cx + (uint32_t)(q / 9 - 1)),
                                  (int)((uint32_t)cy + (uint32_t)((q / 3) % 3 - 1)),
                                  (int)((uint32_t)cz + (uint32_t)(q % 3 - 1)), sc);
#pragma unroll
            for (int q = 0; q < 27; ++q) {
                bool dup = false;
#pragma unroll
                for (int r = 0; r < q; ++r) dup |= hs[r] == hs[q];
                uint4 e;
                if (dup || !table_find(d.table, sc, hs[q], e)) continue;
                const int32_t shift = (int32_t)e.z - (int32_t)e.x;
                for (uint32_t m = e.x; m < e.y; ++m) {
                    if (m == self) continue;
                    float4 xk = ld(d.XsA, e.w ? (uint32_t)((int32_t)m + shift) : ld(d.physT, m));
                    f3 sep = xi - make3(xk);
                    float d2 = dot3(sep, sep);
                    if (d2 < P.h2) rho = rho + xk.w * poly6(d2, P);
                }
            }
            if (rhoOut) rhoOut[self] = rho;
            err = fmaxf(0.0f, rho * P.invRest - 1.0f);
        }
    }
    // one atomic per wave (a wave never spans two scenes: scenes start 256-aligned)
    float m = err;
    double sum = err;
    for (int off = 32; off > 0; off >>= 1) {
        m = fmaxf(m, __shfl_down(m, off, 64));
        sum += __shfl_down(sum, off, 64);
    }
    const uint64_t any = __ballot(active);
    const uint32_t s0 = __shfl(scene, 0, 64);
    if (__lane_id() == 0 && any) {
        atomicMax(&errMaxBits[s0], __float_as_uint(m));  // non-negative floats order as integers
        atomicAdd(&errSum[s0], sum);
    }
}

inline unsigned grid(uint32_t n) { return (n + kBlock - 1) / kBlock; }
// Neighbour kernels: a whole number of workgroups of bs threads per XCD.
inline unsigned grid_xcd(uint32_t n, int bs) {
    const unsigned g = (n + (unsigned)bs - 1) / (unsigned)bs;
    return (g + kXcds - 1) / kXcds * kXcds;
}

#define AKUA_BY_BLOCK(bs, CALL)         \
    do {                                \
        if ((bs) == 128) {              \
            constexpr int BS = 128;     \
            CALL;                       \
        } else if ((bs) == 512) {       \
            constexpr int BS = 512;     \
            CALL;                       \
        } else {                        \
            constexpr int BS = 256;     \
            CALL;                       \
        }                               \
    } while (0)

// Lane-group launches: L in {2, 4, 8, 16}, nslots * L threads.
#define AKUA_BY_LANES(lanes, CALL)      \
    do {                                \
        if ((lanes) == 2) {             \
            constexpr int L = 2;        \
            CALL;                       \
        } else if ((lanes) == 4) {      \
            constexpr int L = 4;        \
            CALL;                       \
        } else if ((lanes) == 8) {      \
            constexpr int L = 8;        \
            CALL;                       \
        } else {                        \
            constexpr int L = 16;       \
            CALL;                       \
        }                               \
    } while (0)

inline unsigned grid_lanes(uint32_t nslots, int lanes, int bs) {
    return grid_xcd(nslots * (uint32_t)lanes, bs);
}

}  // namespace

void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t n, uint32_t v) {
    if (n) k_fill_u32<<<grid(n), kBlock, 0, st>>>(p, n, v);
}
void launch_clear_table(hipStream_t st, uint4* table, uint32_t n) {
    if (n) k_fill_u4<<<grid(n), kBlock, 0, st>>>(table, n, kEmptyKey);
}
void launch_predict_hash(hipStream_t st, const DeviceState& d, const SphParams& P,
                         const StepParams& S, bool layoutPrev, bool clearTable, const SortZeroSpans* zero) {
    const uint32_t predict = grid(d.nslots);
    const uint32_t clear = clearTable && d.tableTotal ? grid(d.tableTotal) : 0u;
    SortZeroSpans z{};
    if (zero) z = *zero;
    const uint32_t zeroBlocks = z.count ? kZeroBlocks : 0u;
    k_predict_hash<<<predict + clear + zeroBlocks, kBlock, 0, st>>>(d, P, S, layoutPrev, predict, clear, z);
}
void launch_cell_keys(hipStream_t st, const DeviceState& d, const SphParams& P,
                      const StepParams& S, bool keepLayout, bool layoutPrev) {
    if (keepLayout)
        k_cell_keys<true><<<grid(d.nslots), kBlock, 0, st>>>(d, P, S, false);
    else
        k_cell_keys<false><<<grid(d.nslots), kBlock, 0, st>>>(d, P, S, layoutPrev);
}
void launch_gather(hipStream_t st, const DeviceState& d, const SphParams& P, const StepParams& S,
                   bool carry, bool carryDelta, bool layoutPrev) {
    if (carryDelta)
        k_gather<true, true><<<grid(d.nslots), kBlock, 0, st>>>(d, P, S, layoutPrev);
    else if (carry)
        k_gather<true, false><<<grid(d.nslots), kBlock, 0, st>>>(d, P, S, layoutPrev);
    else
        k_gather<false, false><<<grid(d.nslots), kBlock, 0, st>>>(d, P, S, layoutPrev);
}
void launch_invert_order(hipStream_t st, const DeviceState& d, uint32_t* phys) {
    k_invert_order<<<grid(d.nslots), kBlock, 0, st>>>(d, phys);
}

void launch_bucket_phys(hipStream_t st, const DeviceState& d) {
    k_bucket_phys<<<grid(d.nslots), kBlock, 0, st>>>(d);
}
void launch_neighbours(hipStream_t st, const DeviceState& d, const SphParams& P, int bs,
                       int slabBatch, bool density, int features, bool cells) {
    if (cells) {
        AKUA_BY_BLOCK(bs, (k_neighbours_cells<BS><<<grid_lanes(d.nslots, 32, BS), BS, 0, st>>>(d, P)));
    } else if (features >= 0) {
        AKUA_BY_BLOCK(bs, (k_neighbours_features<BS><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P, features)));
    } else if (density) {
        if (P.fast)
            AKUA_BY_BLOCK(bs, (k_neighbours_density<BS, true><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
        else
            AKUA_BY_BLOCK(bs, (k_neighbours_density<BS, false><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
    } else if (P.searchCells == 2)
        AKUA_BY_BLOCK(bs, (k_neighbours_fine<BS><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
    else if (slabBatch == 4)
        AKUA_BY_BLOCK(bs, (k_neighbours_slab<BS, 4><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
    else if (slabBatch == 8)
        AKUA_BY_BLOCK(bs, (k_neighbours_slab<BS, 8><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
    else
        AKUA_BY_BLOCK(bs, (k_neighbours<BS><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
}
template <int BS>
void density_lambda_bs(hipStream_t st, const DeviceState& d, const SphParams& P, const float4* Xs,
                       int residSlot, bool gated) {
    const unsigned nvb = grid_xcd(d.nslots, BS);
    const int g = gated ? 1 : 0;
    if (P.bd.on) {
        if (P.fast)
            k_density_lambda<BS, true, true><<<nvb, BS, 0, st>>>(d, P, Xs, residSlot, g, nvb);
        else
            k_density_lambda<BS, false, true><<<nvb, BS, 0, st>>>(d, P, Xs, residSlot, g, nvb);
    } else if (P.fast) {
        k_density_lambda<BS, true, false><<<nvb, BS, 0, st>>>(d, P, Xs, residSlot, g, nvb);
    } else {
        k_density_lambda<BS, false, false><<<nvb, BS, 0, st>>>(d, P, Xs, residSlot, g, nvb);
    }
}
void launch_density_lambda(hipStream_t st, const DeviceState& d, const SphParams& P,
                           const float4* Xs, int bs, int residSlot, bool gated, int lanes) {
    if (lanes > 1) {
        const unsigned g = grid_lanes(d.nslots, lanes, bs);
        if (P.fast)
            AKUA_BY_BLOCK(bs, AKUA_BY_LANES(lanes, (k_density_lambda_group<BS, L, true><<<g, BS, 0, st>>>(d, P, Xs))));
        else
            AKUA_BY_BLOCK(bs, AKUA_BY_LANES(lanes, (k_density_lambda_group<BS, L, false><<<g, BS, 0, st>>>(d, P, Xs))));
        return;
    }
    AKUA_BY_BLOCK(bs, (density_lambda_bs<BS>(st, d, P, Xs, residSlot, gated)));
}
void launch_balance(hipStream_t st, const DeviceState& d, const SphParams& P, uint32_t tile,
                    uint32_t* work, uint32_t* cntW, uint4* nbrW) {
    if (!d.nslots) return;
    const uint32_t blocks = (d.nslots + tile - 1) / tile;
    k_balance<<<blocks, kBalanceThreads, 0, st>>>(d, P, tile, work, cntW, nbrW);
}
void launch_count_sum(hipStream_t st, const DeviceState& d, unsigned long long* part, uint32_t blocks) {
    k_count_sum<<<blocks, 256, 0, st>>>(d, part);
}
void launch_features(hipStream_t st, const DeviceState& d, const SphParams& P, bool record, int bs) {
    AKUA_BY_BLOCK(bs, (k_features<BS><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P, record ? 1 : 0)));
}
void launch_gate(hipStream_t st, const DeviceState& d, uint32_t nscenes, int statistic,
                 float ratio, float tolerance) {
    k_resid_blocks<<<dim3(d.residBlocks, nscenes), kReduceBlock, 0, st>>>(d, 3);
    k_resid_reduce<<<nscenes, 64, 0, st>>>(d, 3, 1, statistic, ratio, tolerance);
}
void launch_resid_reduce(hipStream_t st, const DeviceState& d, uint32_t nscenes, int slotMask) {
    k_resid_blocks<<<dim3(d.residBlocks, nscenes), kReduceBlock, 0, st>>>(d, slotMask);
    k_resid_reduce<<<nscenes, 64, 0, st>>>(d, slotMask, 0, 0, 0.0f, 0.0f);
}
uint32_t resid_blocks_for(uint32_t maxSceneSlots) {
    const uint32_t waves = (maxSceneSlots + 63) / 64 + 1;  // +1: a scene may start mid-wave-row
    return std::max(1u, (waves + kResidWavesPerBlock - 1) / kResidWavesPerBlock);
}
void launch_targets(hipStream_t st, const DeviceState& d, const SphParams& P, const float4* XsLast) {
    k_targets<<<grid(d.nslots), kBlock, 0, st>>>(d, P, XsLast);
}
void launch_sample(hipStream_t st, const DeviceState& d, uint64_t seed, uint32_t count,
                   float* featOut, float* targOut) {
    if (count) k_sample<<<(count + 255) / 256, 256, 0, st>>>(d, seed, count, featOut, targOut);
}
template <int BS>
void delta_correct_bs(hipStream_t st, const DeviceState& d, const SphParams& P,
                      const StepParams& S, const float4* XsIn, float4* XsOut, bool final,
                      bool applyDelta, int gate, const float4* XsAlt) {
    const unsigned nvb = grid_xcd(d.nslots, BS);
    const unsigned g = nvb;
#define AKUA_DC(F, A, K, B) \
    k_delta_correct<BS, F, A, K, B><<<g, BS, 0, st>>>(d, P, S, XsIn, XsOut, XsAlt, gate, nvb)
    if (final && !applyDelta) {
        AKUA_DC(true, false, false, false);  // a step without iterations
    } else if (P.bd.on) {
        if (final) {
            if (P.fast) AKUA_DC(true, true, true, true); else AKUA_DC(true, true, false, true);
        } else {
            if (P.fast) AKUA_DC(false, true, true, true); else AKUA_DC(false, true, false, true);
        }
    } else if (final) {
        if (P.fast) AKUA_DC(true, true, true, false); else AKUA_DC(true, true, false, false);
    } else {
        if (P.fast) AKUA_DC(false, true, true, false); else AKUA_DC(false, true, false, false);
    }
#undef AKUA_DC
}
void launch_delta_correct(hipStream_t st, const DeviceState& d, const SphParams& P,
                          const StepParams& S, const float4* XsIn, float4* XsOut, bool final,
                          bool applyDelta, int bs, int gate, const float4* XsAlt, int lanes) {
    if (lanes > 1 && applyDelta) {
        const unsigned g = grid_lanes(d.nslots, lanes, bs);
#define AKUA_DCG(F, K) \
    AKUA_BY_BLOCK(bs, AKUA_BY_LANES(lanes, (k_delta_correct_group<BS, L, F, K><<<g, BS, 0, st>>>(d, P, S, XsIn, XsOut))))
        if (final) {
            if (P.fast) AKUA_DCG(true, true); else AKUA_DCG(true, false);
        } else {
            if (P.fast) AKUA_DCG(false, true); else AKUA_DCG(false, false);
        }
#undef AKUA_DCG
        return;
    }
    AKUA_BY_BLOCK(bs, (delta_correct_bs<BS>(st, d, P, S, XsIn, XsOut, final, applyDelta, gate,
                                            XsAlt)));
}
void launch_vorticity(hipStream_t st, const DeviceState& d, const SphParams& P, int bs, int lanes) {
    if (lanes > 1) {
        const unsigned g = grid_lanes(d.nslots, lanes, bs);
        if (P.fast)
            AKUA_BY_BLOCK(bs, AKUA_BY_LANES(lanes, (k_vorticity_group<BS, L, true><<<g, BS, 0, st>>>(d, P))));
        else
            AKUA_BY_BLOCK(bs, AKUA_BY_LANES(lanes, (k_vorticity_group<BS, L, false><<<g, BS, 0, st>>>(d, P))));
        return;
    }
    if (P.fast)
        AKUA_BY_BLOCK(bs, (k_vorticity<BS, true><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
    else
        AKUA_BY_BLOCK(bs, (k_vorticity<BS, false><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
}
void launch_confinement(hipStream_t st, const DeviceState& d, const SphParams& P,
                        const StepParams& S, int bs, int lanes) {
    if (lanes > 1) {
        const unsigned g = grid_lanes(d.nslots, lanes, bs);
        if (P.fast)
            AKUA_BY_BLOCK(bs, AKUA_BY_LANES(lanes, (k_confinement_group<BS, L, true><<<g, BS, 0, st>>>(d, P, S))));
        else
            AKUA_BY_BLOCK(bs, AKUA_BY_LANES(lanes, (k_confinement_group<BS, L, false><<<g, BS, 0, st>>>(d, P, S))));
        return;
    }
    if (P.fast)
        AKUA_BY_BLOCK(bs, (k_confinement<BS, true><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P, S)));
    else
        AKUA_BY_BLOCK(bs, (k_confinement<BS, false><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P, S)));
}
void launch_xsph(hipStream_t st, const DeviceState& d, const SphParams& P, int bs, int lanes) {
    if (lanes > 1) {
        const unsigned g = grid_lanes(d.nslots, lanes, bs);
        AKUA_BY_BLOCK(bs, AKUA_BY_LANES(lanes, (k_xsph_group<BS, L><<<g, BS, 0, st>>>(d, P))));
        return;
    }
    AKUA_BY_BLOCK(bs, (k_xsph<BS><<<grid_xcd(d.nslots, BS), BS, 0, st>>>(d, P)));
}
namespace {
// Workgroups of 256 threads that the device holds at once for `kernel`, per XCD (cached per
// kernel and device).
unsigned resident_per_xcd(const void* kernel) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, unsigned> cache;
    auto check = [](hipError_t e, const char* what) {
        if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    };
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(kernel, dev);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int perCu = 0, cus = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, (int)kQueueTile, 0),
          "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    const unsigned v = std::max(1u, (unsigned)std::max(perCu, 1) * (unsigned)std::max(cus, 1) / kXcds);
    cache.emplace(key, v);
    return v;
}
// The grid of a queued sweep: a multiple of 8 workgroups, no more per XCD than its queue holds.
template <class K>
unsigned queue_grid(K kernel, const DeviceState& d) {
    return kXcds * std::max(1u, std::min(resident_per_xcd(reinterpret_cast<const void*>(kernel)), d.queueT8));
}
}  // namespace

void launch_tile_queue(hipStream_t st, const DeviceState& d, const SphParams& P, uint32_t* cost,
                       uint32_t* queue, uint32_t* qcount, uint32_t* heads, uint32_t passes) {
    const uint32_t ntiles = (d.nslots + kQueueTile - 1) / kQueueTile;
    if (ntiles) k_tile_cost<<<(ntiles + 3) / 4, 256, 0, st>>>(d, ntiles, cost);
    // cost classes: the largest possible tile cost (256 full lists) spans kQueueBuckets classes
    const uint64_t maxCost = (uint64_t)kQueueTile * (uint64_t)std::max(P.maxN, 1);
    uint32_t shift = 0;
    while ((maxCost >> shift) >= kQueueBuckets) ++shift;
    k_tile_queue<<<kXcds, kQueueThreads, 0, st>>>(cost, ntiles, d.queueT8, shift, queue, qcount, heads, passes);
}
void launch_density_lambda_queue(hipStream_t st, const DeviceState& d, const SphParams& P,
                                 const float4* Xs, uint32_t pass) {
#define AKUA_DLQ(K, B) \
    k_density_lambda_queue<K, B><<<queue_grid(k_density_lambda_queue<K, B>, d), kQueueTile, 0, st>>>(d, P, Xs, pass)
    if (P.bd.on) {
        if (P.fast) AKUA_DLQ(true, true); else AKUA_DLQ(false, true);
    } else {
        if (P.fast) AKUA_DLQ(true, false); else AKUA_DLQ(false, false);
    }
#undef AKUA_DLQ
}
void launch_delta_correct_queue(hipStream_t st, const DeviceState& d, const SphParams& P,
                                const StepParams& S, const float4* XsIn, float4* XsOut, bool final,
                                uint32_t pass) {
#define AKUA_DCQ(F, K, B)                                                                             \
    k_delta_correct_queue<F, K, B><<<queue_grid(k_delta_correct_queue<F, K, B>, d), kQueueTile, 0, st>>>( \
        d, P, S, XsIn, XsOut, pass)
    if (P.bd.on) {
        if (final) {
            if (P.fast) AKUA_DCQ(true, true, true); else AKUA_DCQ(true, false, true);
        } else {
            if (P.fast) AKUA_DCQ(false, true, true); else AKUA_DCQ(false, false, true);
        }
    } else if (final) {
        if (P.fast) AKUA_DCQ(true, true, false); else AKUA_DCQ(true, false, false);
    } else {
        if (P.fast) AKUA_DCQ(false, true, false); else AKUA_DCQ(false, false, false);
    }
#undef AKUA_DCQ
}
void launch_vorticity_queue(hipStream_t st, const DeviceState& d, const SphParams& P, uint32_t pass) {
    if (P.fast)
        k_vorticity_queue<true><<<queue_grid(k_vorticity_queue<true>, d), kQueueTile, 0, st>>>(d, P, pass);
    else
        k_vorticity_queue<false><<<queue_grid(k_vorticity_queue<false>, d), kQueueTile, 0, st>>>(d, P, pass);
}
void launch_confinement_queue(hipStream_t st, const DeviceState& d, const SphParams& P,
                              const StepParams& S, uint32_t pass) {
    if (P.fast)
        k_confinement_queue<true><<<queue_grid(k_confinement_queue<true>, d), kQueueTile, 0, st>>>(d, P, S, pass);
    else
        k_confinement_queue<false><<<queue_grid(k_confinement_queue<false>, d), kQueueTile, 0, st>>>(d, P, S, pass);
}
void launch_xsph_queue(hipStream_t st, const DeviceState& d, const SphParams& P, uint32_t pass) {
    k_xsph_queue<<<queue_grid(k_xsph_queue, d), kQueueTile, 0, st>>>(d, P, pass);
}
void launch_density_probe(hipStream_t st, const DeviceState& d, const SphParams& P, float* rho,
                          unsigned int* errMaxBits, double* errSum) {
    k_density_probe<256><<<grid_xcd(d.nslots, 256), 256, 0, st>>>(d, P, rho, errMaxBits, errSum);
}

}  // namespace akua
