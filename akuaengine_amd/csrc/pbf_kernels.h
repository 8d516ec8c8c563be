// pbf_kernels.h -- host launchers of the PBF step kernels (pbf_kernels.hip).
//
// Two orders of the particles:
//   logical  -- the reference's order: after every step, the stable sort by cell hash of the
//               previous logical order (NeighbourSearchCUDA.cu:166-170).  Bucket walks, the
//               neighbour-list order, the permutation tap and akua_solver_get_state use it.
//   physical -- where the particle's data lives in HBM: sorted by the Morton code of its
//               2x2x2 super-cell (ties in logical order), so the 64 particles of a wave are
//               spatial neighbours and their neighbour gathers hit the same cache lines.
// lid[p] is the logical index of physical slot p (kept by the step); its inverse, the physical slot
// of each logical index, is physT during a step and is rebuilt from lid where needed outside it.
// Every per-particle sum runs over the neighbour list in the reference's order, so the physical
// layout changes no result, only memory locality.
//
// Data layout in HBM (physical order):
//   X    float4 (x, y, z, mass)   positions after the last step        16 B
//   V    float4 (vx, vy, vz, 0)   velocities after the last step       16 B
//   dens, lam  float              density and lambda of the last step   4 B each
//   id   uint32                   original particle index               4 B
//   lid  uint32                   logical index                         4 B
//   VS   float4 (xyz, 0)          the step's velocity before the wall damping, vorticity and
//                                 XSPH: the record's new_velocity       16 B
//   DP   float4 (xyz, 0)          the last iteration's position delta: the record's
//                                 position_delta                        16 B
// Per-step scratch: X0 (old positions), XsA/XsB (x* ping-pong; after the last iteration XsB holds
// (x, |omega|) from the vorticity pass for the confinement pass), W (vorticity and its length),
// Vc (velocity after vorticity confinement), two sorts' keys/values, the cell table, and the
// neighbour list.  The list holds physical indices in chunks of four entries, wave-interleaved:
// chunk c (entries 4c..4c+3) of slot p is the uint4 at ((p / 64) * ceil(maxN / 4) + c) * 64 +
// p % 64, so the 64 lanes of a wave read chunk c of their 64 lists as one 1 KB block.
#pragma once

#include <hip/hip_runtime.h>

#include "pbf_types.h"
#include "radix_sort.h"

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
    uint32_t* lid;             // logical index of each physical slot (persistent)
    float4* VS;                // new_velocity of the last step (persistent)
    float4* DP;                // position_delta of the last iteration (persistent)
    float4* TDP;               // DP moved to the new physical slots (a step without iterations)
    float4* X0;
    float4* XsA;
    float4* XsB;
    float4* W;
    float4* Vc;
    float4* Bg;                // boundary gradient sum of each slot's last density pass (boundary on)
    float* Tdens;
    float* Tlam;
    uint32_t* Tid;
    uint32_t* lidT;            // this step's logical index of each physical slot
    uint32_t* physT;           // this step's physical slot of each logical index
    uint32_t* srcL;            // previous physical slot of each logical index
    uint2* layPhys;            // at each previous logical index: (layout key, physical slot at the
                               // step start), or with layoutPrev (new logical index, physical
                               // slot at the step start)
    uint32_t* keysL;           // hashes at their previous logical index (first sort's input)
    const uint32_t* skeys;     // hashes in logical order (first sort's output)
    const uint32_t* perm1;     // logical index -> previous logical index
    uint32_t* keysM;           // super-cell Morton keys in logical order (second sort's input)
    const uint32_t* perm2;     // physical slot -> logical index
    uint32_t* perm2keep;       // the same map for a step that keeps the physical layout
    uint4* table;              // cell table records (all scenes): (hash key, first logical
                               // index, one past the last, first physical slot | 2^31 if the
                               // bucket's physical slots are consecutive in logical order)
    uint32_t* bslot;           // table slot of the bucket a logical index starts (first only)
    uint32_t tableTotal;
    uint4* nbr4;               // neighbour lists in four-entry chunks (see list_base)
    uint32_t* cnt;
    // The neighbour sweeps' work order.  Work position t (the t-th thread of a sweep) handles
    // physical slot work[t]; its count is cntW[t] and its list the chunks of position t in nbrW
    // (same chunk layout as nbr4).  Without balancing, work is null, cntW = cnt, nbrW = nbr4.
    const uint32_t* work;
    const uint32_t* cntW;
    const uint4* nbrW;
    // the sweeps' work queue (tuning workQueue): per XCD x, queue[x * queueT8 + k] is its k-th
    // 256-slot tile (heaviest first) and qcount[x] how many it holds; heads holds the counters of
    // the dynamic mode (queue_head); queueMode 1: dynamic, 2: static
    const uint32_t* queue;
    const uint32_t* qcount;
    uint32_t* heads;
    uint32_t queueT8;
    uint32_t queueMode;
    unsigned long long* nbrSum;  // per-workgroup partial sums of neighbour counts (statistics)
    // learned warm start (null when unused)
    float* feat;               // 16 features per slot at x*_pred
    float* targ;               // training targets per slot: (x*_final - x*_pred) / h, lambda
    float4* XsP;               // copy of x*_pred (training)
    float4* XsW;               // warm-started x*
    float* lamHat;             // the network's lambda output
    unsigned int* resid;       // per scene, kResidStride words: max over particles of
                               // max(0, rho / rho0 - 1) as float bits at [k], and its sum as a
                               // float at [kResidSum + k], for k = 0 at x*_pred, 1 at x*_w, 2 at
                               // the last plain density pass
    float2* residPart;         // per check (3) and wave: (max, sum) of max(0, C) over its lanes
    uint32_t nWaves;           // waves of 64 slots: ceil(nslots / 64)
    float2* residBlk;          // per scene, block and check: the first reduction stage
    uint32_t residBlocks;      // blocks per scene of the first stage
    uint32_t* gate;            // per scene: 1 when this step falls back to the plain solve
    uint32_t* fallbacks;       // per scene: running count of fallback steps
};

constexpr int kMlpFeatures = 16;
constexpr int kResidStride = 8;
constexpr int kResidSum = 4;

constexpr int kBlock = 256;

void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t n, uint32_t v);
// Every record of the cell table empty.
void launch_clear_table(hipStream_t st, uint4* table, uint32_t n);
// layoutPrev (the three launches of a step alike): the layout keys are sorted in the previous
// logical order, in the same launches as the hashes, instead of in the new logical order after
// them.  Only for single-cell super-cells (layout level 0), where both orders agree within each key.
// clearTable: the same launch also empties every record of the cell table (launch_clear_table);
// zero (may be NULL): and zeroes those spans of the sorts' scratch.
void launch_predict_hash(hipStream_t st, const DeviceState& d, const SphParams& P,
                         const StepParams& S, bool layoutPrev, bool clearTable,
                         const SortZeroSpans* zero = nullptr);
// keepLayout: the step keeps every particle in its physical slot (no layout sort); the map
// physical slot -> logical index is then written to perm2keep.
void launch_cell_keys(hipStream_t st, const DeviceState& d, const SphParams& P,
                      const StepParams& S, bool keepLayout, bool layoutPrev);
// carry: also move the previous density and lambda (needed by a step without iterations and by
// the learned features); carryDelta: also the position delta (a step without iterations keeps
// it, as the reference's record does).
void launch_gather(hipStream_t st, const DeviceState& d, const SphParams& P, const StepParams& S,
                   bool carry, bool carryDelta, bool layoutPrev);
// Physical start and contiguity of every bucket (after the gather).
void launch_bucket_phys(hipStream_t st, const DeviceState& d);
// phys[lid[p]] = p over every scene's slots.
void launch_invert_order(hipStream_t st, const DeviceState& d, uint32_t* phys);
// bs: workgroup size of the neighbour kernels (128, 256 or 512; results do not depend on it).
// slabBatch: 0 walks each stencil bucket in batches of eight candidates; 4 or 8 walks each dx slab
// of nine buckets as one stream in batches of that many (identical lists).
// density: the same launch also runs the first density/lambda sweep at x*_pred (d.XsA), as
// launch_density_lambda(d.XsA) would; features >= 0: it runs the feature sweep instead, as
// launch_features(record = features) would.  Either only with the physical work order
// (d.work == nullptr), slabBatch 0, searchCells 1 and no boundary samples.
// cells: the same lists with one lane per stencil cell (32 per particle), for small scenes; with
// searchCells 1, slabBatch 0 and neither sweep.
void launch_neighbours(hipStream_t st, const DeviceState& d, const SphParams& P, int bs,
                       int slabBatch, bool density = false, int features = -1, bool cells = false);
// The balanced work order of the neighbour sweeps: within each tile of `tile` physical slots
// (a multiple of 256, at most kBalanceMaxTile), slots sorted by scene, then by neighbour count
// (descending), so the 64 lanes of a wave walk lists of similar length; writes work, cntW and
// the lists in work order (nbrW).  Each wave of work positions stays within one scene.
constexpr uint32_t kBalanceMaxTile = 2048;
void launch_balance(hipStream_t st, const DeviceState& d, const SphParams& P, uint32_t tile,
                    uint32_t* work, uint32_t* cntW, uint4* nbrW);
// Partial sums of the neighbour counts (blocks partials into part).
void launch_count_sum(hipStream_t st, const DeviceState& d, unsigned long long* part, uint32_t blocks);
// residSlot >= 0: also the per-scene max and sum of max(0, C) into check residSlot of d.resid.
// gated: only scenes whose d.gate is set run (the fallback's iterations).
// lanes (4, 8 or 16; 1: one thread per particle): a lane group per particle in the neighbour
// sweeps, for scenes too small to fill the chip, with results identical to one lane per particle;
// only for plain steps (residSlot -1, no gate), the physical order and no wall samples.
void launch_density_lambda(hipStream_t st, const DeviceState& d, const SphParams& P,
                           const float4* Xs, int bs, int residSlot = -1, bool gated = false,
                           int lanes = 1);
// Position correction with the wall clamp.  final == true also runs the velocity update and the
// wall damping and writes X, V, id, lid and phys (the last iteration); applyDelta == false skips
// the correction (a step with zero iterations).
// gate == 1: only scenes whose d.gate is set run.  gate == 2 (final only): scenes whose d.gate
// is clear read XsAlt instead of XsIn (the learned step's final correction).
void launch_delta_correct(hipStream_t st, const DeviceState& d, const SphParams& P,
                          const StepParams& S, const float4* XsIn, float4* XsOut, bool final,
                          bool applyDelta, int bs, int gate = 0, const float4* XsAlt = nullptr,
                          int lanes = 1);
// Features of the learned warm start at x*_pred (XsA) into d.feat, the per-scene residual at
// x*_pred into resid slot 0, and with record also x*_pred into d.XsP.
void launch_features(hipStream_t st, const DeviceState& d, const SphParams& P, bool record, int bs);
// Fallback decision per scene: gate = r(x*_w) > ratio * r(x*_pred) + tolerance, r the max
// (statistic 0) or the mean (statistic 1) over particles of max(0, rho / rho0 - 1).
void launch_gate(hipStream_t st, const DeviceState& d, uint32_t nscenes, int statistic,
                 float ratio, float tolerance);
// Per-scene max and sum of the residual checks in slotMask (bit k: check k) into d.resid.
void launch_resid_reduce(hipStream_t st, const DeviceState& d, uint32_t nscenes, int slotMask);
// First-stage blocks per scene for scenes of at most maxSceneSlots slots.
uint32_t resid_blocks_for(uint32_t maxSceneSlots);
// Training targets of every slot after a recorded plain step; XsLast holds x* before the last
// iteration.
void launch_targets(hipStream_t st, const DeviceState& d, const SphParams& P, const float4* XsLast);
// Gathers count random samples (slots of scene 0) of features and targets.
void launch_sample(hipStream_t st, const DeviceState& d, uint64_t seed, uint32_t count,
                   float* featOut, float* targOut);
void launch_vorticity(hipStream_t st, const DeviceState& d, const SphParams& P, int bs, int lanes = 1);
void launch_confinement(hipStream_t st, const DeviceState& d, const SphParams& P,
                        const StepParams& S, int bs, int lanes = 1);
void launch_xsph(hipStream_t st, const DeviceState& d, const SphParams& P, int bs, int lanes = 1);

// The work queue of the sweeps (DeviceState::queue): per-tile neighbour counts into cost, each
// XCD's tiles heaviest first, and the pass counters of `passes` persistent launches cleared.
// The queued sweeps run as many 256-thread workgroups as the device holds at once (a multiple
// of 8, at most 8 per queued tile of an XCD); `pass` numbers the launch within the step (its
// counters in DeviceState::heads).
constexpr uint32_t kQueueTileSlots = 256;
constexpr uint32_t kXcdCount = 8;  // XCDs of the MI355X; workgroup b runs on XCD b % 8
// Dynamic mode: each XCD's queue is served through kQueueSubs counters, each on its own 128-byte
// line (kQueueHeadStride counters apart); heads holds passes * kXcdCount * kQueueSubs of them.
constexpr uint32_t kQueueSubs = 8;
constexpr uint32_t kQueueHeadStride = 32;
inline size_t queue_heads_size(uint32_t passes) {
    return (size_t)passes * kXcdCount * kQueueSubs * kQueueHeadStride;
}
void launch_tile_queue(hipStream_t st, const DeviceState& d, const SphParams& P, uint32_t* cost,
                       uint32_t* queue, uint32_t* qcount, uint32_t* heads, uint32_t passes);
void launch_density_lambda_queue(hipStream_t st, const DeviceState& d, const SphParams& P,
                                 const float4* Xs, uint32_t pass);
void launch_delta_correct_queue(hipStream_t st, const DeviceState& d, const SphParams& P,
                                const StepParams& S, const float4* XsIn, float4* XsOut, bool final,
                                uint32_t pass);
void launch_vorticity_queue(hipStream_t st, const DeviceState& d, const SphParams& P, uint32_t pass);
void launch_confinement_queue(hipStream_t st, const DeviceState& d, const SphParams& P,
                              const StepParams& S, uint32_t pass);
void launch_xsph_queue(hipStream_t st, const DeviceState& d, const SphParams& P, uint32_t pass);
// Exact SPH density at the positions held in XsA; rho by logical index, per-scene error max and
// sum.
void launch_density_probe(hipStream_t st, const DeviceState& d, const SphParams& P, float* rho,
                          unsigned int* errMaxBits, double* errSum);

}  // namespace akua
