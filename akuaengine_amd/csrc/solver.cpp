// solver.cpp -- the C ABI (include/akua.h) over the gfx950 PBF kernels.
//
// akua_solver is the headless counterpart of AkuaEngine::PBFSolver (PBFSolver.h:12-31): it owns
// every device buffer, sequences the stages of PBFSolver::step (PBFSolver.cpp:22-78) on one HIP
// stream without host synchronisation, and replays the whole step as one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/akua.h"
#include "aos.h"
#include "../../include/akua_io.h"
#include "boundary.h"
#include "last_error.h"
#include "render.h"
#include "mlp.h"
#include "pbf_kernels.h"
#include "pbf_types.h"
#include "radix_sort.h"

using namespace akua;

namespace {

thread_local std::string g_lastError;

struct AkuaError : std::runtime_error {
    akua_status code;
    AkuaError(akua_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                    \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            throw AkuaError(_e == hipErrorOutOfMemory ? AKUA_ERR_OUT_OF_MEMORY : AKUA_ERR_HIP, \
                            std::string(#expr) + ": " + hipGetErrorString(_e));            \
    } while (0)

void fail(akua_status c, const std::string& m) { throw AkuaError(c, m); }

template <class F>
akua_status guarded(F&& f) {
    try {
        f();
        g_lastError.clear();
        return AKUA_OK;
    } catch (const AkuaError& e) {
        g_lastError = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_lastError = "host allocation failed";
        return AKUA_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return AKUA_ERR_HIP;
    }
}

// The constants of the SPH kernels, computed with the float expressions of
// SmoothingKernelsCUDA.h:16-28 (poly6 and spiky factors) and ConstraintSolverCUDA.cu:201, :123-124.
float poly6_host(float d2, float h2, float cp) {
    if (d2 > h2) return 0.0f;
    float t = h2 - d2;
    return cp * ((t * t) * t);
}

SphParams make_sph(const akua_pbf_config& c, const akua_lambda_corr& corr) {
    SphParams P{};
    P.h = c.smoothRadius;
    P.h2 = P.h * P.h;
    P.cp = 315.0f / (64.0f * 3.14f * powf(P.h, 9.0f));
    P.cs = -45.0f / (3.14f * powf(P.h, 6.0f));
    P.w0 = poly6_host(0.0f, P.h2, P.cp);
    float dq2 = corr.delta_q * corr.delta_q;
    P.wdq = poly6_host(dq2, P.h2, P.cp);
    P.invRest = 1.0f / c.restDensity;
    P.relax = c.relaxation;
    P.negCorrK = -corr.k;
    P.corrOn = corr.enabled != 0 ? 1 : 0;
    P.corrN = corr.n;
    P.corrNInt = (corr.n >= 0.0f && corr.n <= 16.0f && corr.n == floorf(corr.n)) ? (int)corr.n : -1;
    P.cellHash = c.smoothRadius;
    P.cellSearch = c.spatialHashCellSize;
    P.searchCells = 1;
    P.vortEps = c.vorticityEpsilon;
    P.visc = c.viscosity;
    P.maxN = c.maxNeighbours;
    return P;
}

// Whether the neighbour sweeps' quotients need no upper-bound or divisor range test: the spiky
// gradient divides numerators of magnitude at most |cs| h^3 by |r| in [1e-5, h]; the artificial
// pressure divides poly6 values of at most cp h^6 by poly6(delta_q^2).  Both must lie inside
// [2^-40, 2^40] (the exact-division range), with a margin for rounding.
bool quotients_bounded(const SphParams& P) {
    const double lo = std::ldexp(1.0, -40), hi = std::ldexp(1.0, 40), margin = 1.01;
    const double h = P.h, cs = std::fabs((double)P.cs), cp = std::fabs((double)P.cp);
    if (!(h > 0.0 && h <= hi && 1e-5 >= lo)) return false;
    if (!(cs * h * h * h * margin <= hi)) return false;
    const double h6 = h * h * h * h * h * h;
    if (!(cp * h6 * margin <= hi)) return false;
    if (!P.corrOn) return true;  // no artificial-pressure quotient
    const double wdq = P.wdq;
    return wdq >= lo && wdq <= hi;
}

uint32_t ceil_log2(uint64_t v) {
    uint32_t r = 0;
    while ((1ull << r) < v) ++r;
    return r;
}

// Owning device allocation (not copyable; alloc replaces any earlier allocation).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void alloc(size_t count) {
        release();
        if (count) HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

constexpr uint32_t kCountBlocks = 512;

// Four-entry chunks per neighbour list (the kernels' list_chunks).
inline size_t list_chunk_count(int maxN) { return ((size_t)maxN + 3) / 4; }

struct TimedLaunch {
    const char* name;
    hipEvent_t a, b;
};

}  // namespace

struct akua_solver {
    akua_pbf_config cfg{};
    akua_lambda_corr corr{};
    SphParams sph{};
    int device = 0;
    hipStream_t stream = nullptr;

    // scenes
    std::vector<SceneDev> scenes;
    std::vector<uint32_t> chunkSceneHost;
    uint32_t nslots = 0;
    uint32_t nparticles = 0;
    uint32_t tableTotal = 0;
    SortPlan plan1;          // hash sort (logical order)
    int hashBits = 1;        // bits of the reference's hashes (log2 of the largest table size)
    SortPlan plan2;          // super-cell sort (physical layout)
    // block, layout level, digit bits (0: balanced), layout interval, balance tile
    akua_tuning tuning{256, 1, 0, 0, 1, 0, 0, 0, 0, 0};
    int layoutPhase = 0;           // steps since the physical layout was last re-sorted (0: now)
    std::vector<SortTile> tilesHost;
    std::vector<SortScene> sortScenesHost;
    uint32_t numTiles = 0;
    bool scenesDirty = true;

    // device buffers
    DevBuf<SceneDev> dScenes;
    DevBuf<uint32_t> dChunkScene;
    DevBuf<float4> X, V, VS, DP, TDP, X0, XsA, XsB, XsC, W, Vc, color;
    DevBuf<float> dens, lam, Tdens, Tlam, size;
    DevBuf<uint32_t> id, Tid, lid, phys, lidT, physT, srcL, bslot, nbr, cnt, perm2keep;
    DevBuf<uint2> layPhys;
    DevBuf<uint32_t> s1KA, s1VA, s1KB, s1VB, s2KA, s2VA, s2KB, s2VB, counts, totals, look, tickets;
    DevBuf<uint32_t> counts2, totals2, look2, tickets2;  // the layout sort's, beside the hash sort
    DevBuf<uint4> table;
    DevBuf<uint32_t> work, cntW, nbrW;  // the sweeps' balanced work order (allocated when used)
    DevBuf<uint32_t> qcost, queue, qcount, heads;  // the sweeps' work queue (allocated when used)
    DevBuf<SortTile> dTiles;
    DevBuf<SortScene> dSortScenes;
    DevBuf<unsigned long long> nbrSum;
    DevBuf<float> probeRho;
    DevBuf<unsigned int> probeMax;
    DevBuf<double> probeSum;
    bool hasState = false;
    bool exact = true;             // the autotuner keeps the reference's search (akua_solver_set_exact)
    bool unitMass = false;         // every particle's mass is exactly 1 (known from the upload)
    bool boundedQuotients = false; // quotients_bounded(sph)

    // learned warm start
    int mode = 0;                  // 0: plain solver, 1: learned warm start with fallback
    int gateStat = 0;              // 0: max over particles, 1: mean
    int learnedIters = 1;          // Jacobi iterations after the warm start
    int trainIters = 1;            // recorded steps: targets are x* before the last trainIters
                                   // iterations (akua_solver_train's targetIterations)
    float gateRatio = 1.0f, gateTol = 0.01f;
    bool haveNet = false;
    // boundary samples on the box walls (akua_solver_set_boundary)
    int bndLayers = 0;             // 0: off
    float bndSpacing = 0.0f;       // <= 0: the configuration's particle spacing
    bool bndDirty = true;          // rebuild before the next step (settings or box changed)
    bool boxSeen = false;          // a step has passed a box
    DevBuf<float> bndPsi;
    DevBuf<float4> Bg;
    // headless rendering (akua_render): per-pixel keys and the RGB8 image
    DevBuf<unsigned long long> renderKeys;
    DevBuf<unsigned char> renderRgb;
    DevBuf<float> net, feat, targ, lamHat;
    DevBuf<float4> XsP, XsW;
    DevBuf<unsigned int> resid;
    DevBuf<float2> residPart, residBlk;
    uint32_t residBlocks = 1;
    DevBuf<uint32_t> gate, fallbacks;
    int64_t learnedSteps = 0;
    // training
    DevBuf<float> adamM, adamV, part, batchF, batchT, lossBuf;
    uint32_t batchCap = 0, lossCap = 0;
    int64_t adamT = 0;

    // step graphs: [0] re-sorts the physical layout, [1] keeps it
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        StepParams params{};
        int plan2 = -1;
        akua_tuning tuning{};
        bool dirty = true;
    };
    bool useGraph = true;
    StepGraph graphs[2];
    bool graphDirty = false;
    uint64_t stepsSinceStats = 0;

    // step_device: orders the solver's stream with the caller's stream (created on first use)
    hipEvent_t orderEvent = nullptr;
    // record export staging (get_particles), allocated on first use
    DevBuf<AosParticle> aosStage;

    // timing
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> eventPool;
    size_t eventNext = 0;
    std::map<std::string, std::pair<double, int64_t>> timeAcc;
    std::vector<std::string> timeOrder;

    DeviceState devState(bool keepLayout = false) {
        DeviceState d{};
        d.nslots = nslots;
        d.scenes = dScenes.p;
        d.chunkScene = dChunkScene.p;
        d.X = X.p; d.V = V.p; d.dens = dens.p; d.lam = lam.p; d.id = id.p;
        d.VS = VS.p; d.DP = DP.p; d.TDP = TDP.p;
        d.lid = lid.p;
        d.X0 = X0.p; d.XsA = XsA.p; d.XsB = XsB.p; d.W = W.p; d.Vc = Vc.p; d.Bg = Bg.p;
        d.Tdens = Tdens.p; d.Tlam = Tlam.p; d.Tid = Tid.p;
        d.lidT = lidT.p; d.physT = physT.p;
        d.srcL = srcL.p; d.layPhys = layPhys.p;
        const bool r1 = plan1.passes() % 2 == 1, r2 = plan2.passes() % 2 == 1;
        d.keysL = s1KA.p;
        d.skeys = r1 ? s1KB.p : s1KA.p;
        d.perm1 = r1 ? s1VB.p : s1VA.p;
        d.keysM = s2KA.p;
        d.perm2 = keepLayout ? perm2keep.p : (r2 ? s2VB.p : s2VA.p);
        d.perm2keep = perm2keep.p;
        d.table = table.p; d.bslot = bslot.p; d.tableTotal = tableTotal;
        d.nbr4 = reinterpret_cast<uint4*>(nbr.p); d.cnt = cnt.p; d.nbrSum = nbrSum.p;
        if (tuning.balanceTile > 0 && nbrW.p) {
            d.work = work.p; d.cntW = cntW.p; d.nbrW = reinterpret_cast<const uint4*>(nbrW.p);
        } else {
            d.work = nullptr; d.cntW = cnt.p; d.nbrW = d.nbr4;
        }
        d.queue = queue.p; d.qcount = qcount.p; d.heads = heads.p;
        d.queueMode = (uint32_t)tuning.workQueue;
        d.queueT8 = ((nslots + kQueueTileSlots - 1) / kQueueTileSlots + kXcdCount - 1) / kXcdCount;
        d.feat = feat.p; d.targ = targ.p; d.XsP = XsP.p; d.XsW = XsW.p; d.lamHat = lamHat.p;
        d.resid = resid.p; d.gate = gate.p; d.fallbacks = fallbacks.p;
        d.residPart = residPart.p; d.nWaves = (nslots + 63) / 64;
        d.residBlk = residBlk.p; d.residBlocks = residBlocks;
        return d;
    }

    // Teardown: wait for the stream's work, free the device buffers, then destroy the graphs,
    // the events and last the stream.
    ~akua_solver() {
        if (stream) (void)hipStreamSynchronize(stream);
        release_buffers();
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        for (auto e : eventPool) (void)hipEventDestroy(e);
        if (orderEvent) (void)hipEventDestroy(orderEvent);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void release_buffers() {
        dScenes.release(); dChunkScene.release();
        for (auto* b : {&X, &V, &VS, &DP, &TDP, &X0, &XsA, &XsB, &XsC, &W, &Vc, &color, &Bg, &XsP, &XsW})
            b->release();
        for (auto* b : {&dens, &lam, &Tdens, &Tlam, &size, &probeRho, &bndPsi, &net, &feat, &targ,
                        &lamHat, &adamM, &adamV, &part, &batchF, &batchT, &lossBuf})
            b->release();
        for (auto* b : {&id, &Tid, &lid, &phys, &lidT, &physT, &srcL, &bslot, &nbr, &cnt,
                        &perm2keep, &s1KA, &s1VA, &s1KB, &s1VB, &s2KA, &s2VA, &s2KB, &s2VB, &counts,
                        &totals, &look, &tickets, &counts2, &totals2, &look2, &tickets2, &gate, &fallbacks, &work, &cntW, &nbrW, &qcost,
                        &queue, &qcount, &heads})
            b->release();
        layPhys.release(); table.release(); dTiles.release(); dSortScenes.release(); nbrSum.release();
        probeMax.release(); probeSum.release(); renderKeys.release(); renderRgb.release();
        resid.release(); residPart.release(); residBlk.release(); aosStage.release();
    }
};

namespace {

void update_layout(akua_solver* s);

// The multiply-high modulo must agree with % for this divisor; checked on edge values and a
// pseudo-random sample before any kernel uses it.
void check_fast_mod(const FastMod& f) {
    uint32_t xs[] = {0u, 1u, f.d - 1u, f.d, f.d + 1u, 2u * f.d - 1u, 2u * f.d, 0x7FFFFFFFu,
                     0x80000000u, 0xFFFFFFFEu, 0xFFFFFFFFu};
    uint64_t z = 0x9E3779B97F4A7C15ull ^ f.d;
    for (uint32_t x : xs)
        if (fast_mod(x, f.d, f.magic, f.shift) != x % f.d)
            fail(AKUA_ERR_UNSUPPORTED, "fast modulo self-check failed for table size " + std::to_string(f.d));
    for (int i = 0; i < 4096; ++i) {
        z = z * 6364136223846793005ull + 1442695040888963407ull;
        uint32_t x = (uint32_t)(z >> 32);
        if (fast_mod(x, f.d, f.magic, f.shift) != x % f.d)
            fail(AKUA_ERR_UNSUPPORTED, "fast modulo self-check failed for table size " + std::to_string(f.d));
    }
}

void validate_config(const akua_pbf_config* cfg, const akua_lambda_corr* corr) {
    if (!cfg || !corr) fail(AKUA_ERR_INVALID_ARG, "config pointers must not be NULL");
    if (!(cfg->smoothRadius > 0.0f) || !std::isfinite(cfg->smoothRadius))
        fail(AKUA_ERR_INVALID_ARG, "smoothRadius must be a positive finite number");
    if (!(cfg->spatialHashCellSize > 0.0f) || !std::isfinite(cfg->spatialHashCellSize))
        fail(AKUA_ERR_INVALID_ARG, "spatialHashCellSize must be a positive finite number");
    if (!(cfg->restDensity > 0.0f) || !std::isfinite(cfg->restDensity))
        fail(AKUA_ERR_INVALID_ARG, "restDensity must be a positive finite number");
    if (cfg->maxNeighbours < 1) fail(AKUA_ERR_INVALID_ARG, "maxNeighbours must be >= 1");
    if (cfg->solverIterations < 0) fail(AKUA_ERR_INVALID_ARG, "solverIterations must be >= 0");
    if (!std::isfinite(cfg->relaxation) || !std::isfinite(cfg->viscosity) ||
        !std::isfinite(cfg->vorticityEpsilon) || !std::isfinite(corr->k) ||
        !std::isfinite(corr->n) || !std::isfinite(corr->delta_q))
        fail(AKUA_ERR_INVALID_ARG, "non-finite solver parameter");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(cfg->gravity[k])) fail(AKUA_ERR_INVALID_ARG, "non-finite gravity");
}

void check_box(const float* bmin, const float* bmax) {
    if (!bmin || !bmax) fail(AKUA_ERR_INVALID_ARG, "box pointers must not be NULL");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(bmin[k]) || !std::isfinite(bmax[k]))
            fail(AKUA_ERR_INVALID_ARG, "non-finite box bound");
}

void setup(akua_solver* s, const std::vector<uint32_t>& counts) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        fail(AKUA_ERR_NO_DEVICE, "no HIP device is visible");
    HIP_CHECK(hipGetDevice(&s->device));
    HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));

    // scene slot layout
    uint32_t slot = 0, tab = 0;
    uint32_t maxTable = 1;
    std::vector<uint32_t> sb, se;
    for (size_t i = 0; i < counts.size(); ++i) {
        uint64_t n = counts[i];
        uint64_t T = (uint64_t)s->cfg.maxNeighbours * n;
        if (T > 0xFFFFFFFFull)
            fail(AKUA_ERR_UNSUPPORTED, "maxNeighbours * numParticles must fit in 32 bits");
        SceneDev sc{};
        sc.begin = slot;
        sc.end = slot + (uint32_t)n;
        sc.tableSize = (uint32_t)T;
        const FastMod fm = make_fast_mod(sc.tableSize);
        sc.modMagic = fm.magic;
        sc.modShift = fm.shift;
        check_fast_mod(fm);
        uint32_t lg = std::max<uint32_t>(6u, ceil_log2(2ull * n));
        sc.tabOffset = tab;
        sc.tabLog2 = lg;
        tab += 1u << lg;
        for (int k = 0; k < 4; ++k) { sc.bmin[k] = 0.0f; sc.bmax[k] = 1.0f; }
        s->scenes.push_back(sc);
        sb.push_back(sc.begin);
        se.push_back(sc.end);
        maxTable = std::max(maxTable, sc.tableSize);
        uint64_t next = (uint64_t)slot + n;
        next = (next + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
        if (counts.size() == 1) next = slot + n;  // a single scene needs no padding
        if (next > std::min<uint64_t>(kSortMaxSlots, kMaxSlots))
            fail(AKUA_ERR_UNSUPPORTED, "too many particles (at most 2^28 slots)");
        slot = (uint32_t)next;
    }
    s->nslots = slot;
    s->tableTotal = tab;
    s->nparticles = 0;
    for (auto c : counts) s->nparticles += c;
    s->hashBits = (int)ceil_log2((uint64_t)maxTable);
    s->plan1 = SortPlan::balanced(s->hashBits);
    s->plan2 = SortPlan::balanced(3);
    s->numTiles = build_sort_tiles(sb, se, s->tilesHost, s->sortScenesHost);

    const uint32_t nChunks = (s->nslots + 255) / 256;
    s->chunkSceneHost.assign(nChunks, 0);
    for (size_t i = 0; i < s->scenes.size(); ++i) {
        const SceneDev& sc = s->scenes[i];
        for (uint32_t c = sc.begin / 256; c < nChunks && c * 256 < std::max(sc.end, sc.begin + 1); ++c)
            s->chunkSceneHost[c] = (uint32_t)i;
    }

    const size_t ns = s->nslots;
    const size_t nsList = (ns + 63) / 64 * 64;
    s->dScenes.alloc(s->scenes.size());
    s->dChunkScene.alloc(nChunks);
    s->X.alloc(ns); s->V.alloc(ns); s->X0.alloc(ns); s->XsA.alloc(ns); s->XsB.alloc(ns);
    s->VS.alloc(ns); s->DP.alloc(ns); s->TDP.alloc(ns);
    s->W.alloc(ns); s->Vc.alloc(ns); s->color.alloc(ns);
    s->dens.alloc(ns); s->lam.alloc(ns); s->Tdens.alloc(ns); s->Tlam.alloc(ns); s->size.alloc(ns);
    s->id.alloc(ns); s->Tid.alloc(ns);
    s->lid.alloc(ns); s->phys.alloc(ns); s->lidT.alloc(ns); s->physT.alloc(ns);
    s->XsC.alloc(ns); s->srcL.alloc(ns); s->layPhys.alloc(ns); s->perm2keep.alloc(ns);
    s->s1KA.alloc(ns); s->s1VA.alloc(ns); s->s1KB.alloc(ns); s->s1VB.alloc(ns);
    s->s2KA.alloc(ns); s->s2VA.alloc(ns); s->s2KB.alloc(ns); s->s2VB.alloc(ns);
    for (auto* c : {&s->counts, &s->counts2}) c->alloc((size_t)s->numTiles * (1u << kMaxDigitBits));
    for (auto* t : {&s->totals, &s->totals2})
        t->alloc((size_t)kMaxPasses * (1u << kMaxDigitBits) * s->scenes.size());
    if (s->numTiles <= kOneSweepMaxTiles) {
        for (auto* l : {&s->look, &s->look2}) l->alloc((size_t)2 * s->numTiles * (1u << kMaxDigitBits));
        for (auto* t : {&s->tickets, &s->tickets2}) t->alloc((size_t)kMaxPasses);
    }
    s->table.alloc(s->tableTotal);
    s->bslot.alloc(ns);
    s->nbr.alloc(nsList * (size_t)list_chunk_count(s->cfg.maxNeighbours) * 4u);
    s->cnt.alloc(ns);
    s->dTiles.alloc(s->tilesHost.size());
    s->dSortScenes.alloc(s->sortScenesHost.size());
    s->nbrSum.alloc(kCountBlocks);
    s->probeRho.alloc(ns);
    s->probeMax.alloc(s->scenes.size());
    s->probeSum.alloc(s->scenes.size());
    HIP_CHECK(hipMemcpy(s->dChunkScene.p, s->chunkSceneHost.data(), nChunks * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->dTiles.p, s->tilesHost.data(), s->tilesHost.size() * sizeof(SortTile),
                        hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->dSortScenes.p, s->sortScenesHost.data(),
                        s->sortScenesHost.size() * sizeof(SortScene), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemset(s->nbrSum.p, 0, kCountBlocks * sizeof(unsigned long long)));
    HIP_CHECK(hipMemset(s->W.p, 0, ns * sizeof(float4)));
    HIP_CHECK(hipMemset(s->s1KA.p, 0, ns * 4));
    HIP_CHECK(hipMemset(s->lidT.p, 0, ns * 4));
    HIP_CHECK(hipMemset(s->srcL.p, 0, ns * 4));
    HIP_CHECK(hipMemset(s->perm2keep.p, 0, ns * 4));
    HIP_CHECK(hipMemset(s->bslot.p, 0xFF, ns * 4));
    HIP_CHECK(hipMemset(s->table.p, 0xFF, (size_t)s->tableTotal * sizeof(uint4)));
    HIP_CHECK(hipMemset(s->layPhys.p, 0, ns * sizeof(uint2)));
    HIP_CHECK(hipMemset(s->XsC.p, 0, ns * sizeof(float4)));
    HIP_CHECK(hipMemset(s->physT.p, 0, ns * 4));
    HIP_CHECK(hipMemset(s->lid.p, 0, ns * 4));
    HIP_CHECK(hipMemset(s->phys.p, 0, ns * 4));
    HIP_CHECK(hipMemset(s->s2KA.p, 0, ns * 4));
    HIP_CHECK(hipMemset(s->X.p, 0, ns * sizeof(float4)));
    HIP_CHECK(hipMemset(s->V.p, 0, ns * sizeof(float4)));
    HIP_CHECK(hipMemset(s->VS.p, 0, ns * sizeof(float4)));
    HIP_CHECK(hipMemset(s->DP.p, 0, ns * sizeof(float4)));
    HIP_CHECK(hipMemset(s->TDP.p, 0, ns * sizeof(float4)));
    HIP_CHECK(hipMemset(s->cnt.p, 0, ns * 4));
    s->sph = make_sph(s->cfg, s->corr);
    s->boundedQuotients = quotients_bounded(s->sph);
    update_layout(s);
}

// SphParams::fast from the masses and the parameters; the step graphs capture the kernels it
// selects, so a change re-captures them.
void update_fast(akua_solver* s) {
    const int f = (s->unitMass && s->boundedQuotients) ? 1 : 0;
    if (s->sph.fast != f) {
        s->sph.fast = f;
        s->graphDirty = true;
    }
}

void timing_hook(void* ctx, const char* name, int phase);

struct Launcher {
    akua_solver* s;
    hipStream_t st;
    bool timed;
    void before(const char* name) {
        if (timed) timing_hook(s, name, 0);
    }
    void after(const char* name) {
        if (timed) timing_hook(s, name, 1);
    }
};

hipEvent_t next_event(akua_solver* s) {
    if (s->eventNext == s->eventPool.size()) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        s->eventPool.push_back(e);
    }
    return s->eventPool[s->eventNext++];
}

void timing_hook(void* ctx, const char* name, int phase) {
    auto* s = static_cast<akua_solver*>(ctx);
    if (phase == 0) {
        TimedLaunch t{name, next_event(s), nullptr};
        HIP_CHECK(hipEventRecord(t.a, s->stream));
        s->pending.push_back(t);
    } else {
        TimedLaunch& t = s->pending.back();
        t.b = next_event(s);
        HIP_CHECK(hipEventRecord(t.b, s->stream));
    }
}

void collect_timing(akua_solver* s) {
    if (s->pending.empty()) return;
    HIP_CHECK(hipStreamSynchronize(s->stream));
    for (auto& t : s->pending) {
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
        auto it = s->timeAcc.find(t.name);
        if (it == s->timeAcc.end()) {
            s->timeOrder.push_back(t.name);
            s->timeAcc[t.name] = {ms, 1};
        } else {
            it->second.first += ms;
            it->second.second += 1;
        }
    }
    s->pending.clear();
    s->eventNext = 0;
}

// The logical order: hashes stably sorted over the previous logical order.
void sort_hash(akua_solver* s, hipStream_t st, bool timed) {
    SortBuffers b{s->s1KA.p, s->s1VA.p, s->s1KB.p, s->s1VB.p, s->counts.p, s->totals.p, s->look.p, s->tickets.p};
    radix_sort_pairs(st, s->plan1, b, s->dTiles.p, s->numTiles, s->dSortScenes.p,
                     (uint32_t)s->scenes.size(), s->nslots, timed ? timing_hook : nullptr, s);
}

// The two sorts' buffers and scratch.
SortBuffers hash_sort_buffers(akua_solver* s) {
    return SortBuffers{s->s1KA.p, s->s1VA.p, s->s1KB.p, s->s1VB.p, s->counts.p, s->totals.p, s->look.p, s->tickets.p};
}
SortBuffers layout_sort_buffers(akua_solver* s) {
    return SortBuffers{s->s2KA.p, s->s2VA.p, s->s2KB.p, s->s2VB.p, s->counts2.p, s->totals2.p, s->look2.p, s->tickets2.p};
}
SortZeroSpans sort_both_zero_spans(akua_solver* s) {
    return radix_sort_zero_spans(s->plan1, hash_sort_buffers(s), s->plan2, layout_sort_buffers(s), s->numTiles,
                                 (uint32_t)s->scenes.size());
}
// Both sorts in shared launches, the layout keys in the previous logical order (single-cell
// super-cells with equal digit widths); false if the widths differ (nothing launched).
// zeroed: the spans of sort_both_zero_spans were zeroed since the last sort.
bool sort_both(akua_solver* s, hipStream_t st, bool timed, bool zeroed) {
    int r1 = 0, r2 = 0;
    return radix_sort_pairs2(st, s->plan1, hash_sort_buffers(s), s->plan2, layout_sort_buffers(s), s->dTiles.p,
                             s->numTiles, s->dSortScenes.p, (uint32_t)s->scenes.size(), s->nslots,
                             timed ? timing_hook : nullptr, s, r1, r2, zeroed);
}

// The physical layout: super-cell Morton keys stably sorted over the new logical order.
void sort_layout(akua_solver* s, hipStream_t st, bool timed) {
    SortBuffers b{s->s2KA.p, s->s2VA.p, s->s2KB.p, s->s2VB.p, s->counts.p, s->totals.p, s->look.p, s->tickets.p};
    radix_sort_pairs(st, s->plan2, b, s->dTiles.p, s->numTiles, s->dSortScenes.p,
                     (uint32_t)s->scenes.size(), s->nslots, timed ? timing_hook : nullptr, s);
}

// PBFSolver::step (PBFSolver.cpp:22-78) as one stream-ordered sequence of launches.
// record: also the features and training targets of the step (a plain step).  In learned mode
// the step is the warm start, one Jacobi iteration and, per scene, the plain iterations again
// from x*_pred when the residual check fails.
// Persistent launches per step with the work queue: every density/lambda and position-delta
// sweep, then vorticity, confinement and XSPH.
uint32_t queue_passes(const akua_solver* s) { return 2u * (uint32_t)s->cfg.solverIterations + 3u; }

void enqueue_step(akua_solver* s, hipStream_t st, const StepParams& S, bool timed,
                  bool record = false, bool keepLayout = false) {
    Launcher L{s, st, timed};
    DeviceState d = s->devState(keepLayout);
    const SphParams& P = s->sph;
    // single-cell super-cells: the layout keys are sorted in the previous logical order, which
    // agrees with the new one within each key, so both sorts share their launches
    const bool together = !keepLayout && s->tuning.layoutLevel == 0 && s->plan1.digitBits == s->plan2.digitBits;
    // the prediction's launch also clears the cell table and, for paired look-back sorts, their
    // scratch
    const SortZeroSpans zero = together ? sort_both_zero_spans(s) : SortZeroSpans{};
    L.before("predict_hash");
    launch_predict_hash(st, d, P, S, together, true, &zero);
    L.after("predict_hash");
    if (together) {
        if (!sort_both(s, st, timed, zero.count > 0)) fail(AKUA_ERR_STATE, "paired sorts need equal digit widths");
        L.before("cell_keys");
        launch_cell_keys(st, d, P, S, false, true);
        L.after("cell_keys");
    } else {
        sort_hash(s, st, timed);
        L.before("cell_keys");
        launch_cell_keys(st, d, P, S, keepLayout, false);
        L.after("cell_keys");
        if (!keepLayout) sort_layout(s, st, timed);
    }
    L.before("gather");
    launch_gather(st, d, P, S,
                  s->cfg.solverIterations == 0 || record || (s->mode == 1 && s->haveNet),
                  s->cfg.solverIterations == 0, together);
    L.after("gather");
    L.before("bucket_phys");
    launch_bucket_phys(st, d);
    L.after("bucket_phys");
    const int bs = s->tuning.block;
    const int I = s->cfg.solverIterations;
    const bool learned = s->mode == 1 && !record;
    const bool plainOrder = !d.work && s->tuning.slabBatch == 0 && P.searchCells == 1 && !P.bd.on;
    // small scenes: lane groups in the sweeps of a plain step, one lane per stencil cell in the
    // list build
    const int lanesSet = s->tuning.laneGroup > 1 && plainOrder && !learned && !record ? s->tuning.laneGroup : 1;
    // with the work queue too, the queue runs the list build, the Jacobi sweeps and confinement, and
    // the lane groups run vorticity and XSPH (two gathered records per neighbour, whose latency
    // more lanes per particle hide better)
    const bool queueSet = s->tuning.workQueue > 0 && !learned && !record && s->queue.p &&
                          s->heads.n >= queue_heads_size(queue_passes(s));
    const int lanes = queueSet ? 1 : lanesSet;
    const int postLanes = lanesSet;
    // the first plain density/lambda sweep, or the learned step's feature sweep, runs inside
    // the list build's launch
    const bool fusable = s->tuning.fuseSweep == 1 && plainOrder && lanes == 1;
    const bool fusedDensity = fusable && I > 0 && !learned && !record;
    const bool fusedFeatures = fusable && (learned || record);
    // the sweeps of a plain step as persistent launches over each XCD's queue of tiles
    const bool queued = queueSet;
    uint32_t pass = 0;
    const char* nbrName = fusedDensity ? "neighbours_density" : fusedFeatures ? "neighbours_features"
                                                                              : "neighbours";
    L.before(nbrName);
    launch_neighbours(st, d, P, bs, s->tuning.slabBatch, fusedDensity,
                      fusedFeatures ? (record ? 1 : 0) : -1, lanes > 1);
    L.after(nbrName);
    if (d.work) {
        L.before("balance");
        launch_balance(st, d, P, (uint32_t)s->tuning.balanceTile, s->work.p, s->cntW.p,
                       reinterpret_cast<uint4*>(s->nbrW.p));
        L.after("balance");
    }
    if (queued) {
        L.before("tile_queue");
        launch_tile_queue(st, d, P, s->qcost.p, s->queue.p, s->qcount.p, s->heads.p, queue_passes(s));
        L.after("tile_queue");
    }
    const int J = learned ? std::min(s->learnedIters, I) : 0;
    // record: targets are x* before the last Jt iterations and the lambda computed there
    const int Jt = std::min(std::max(s->trainIters, 1), I);
    float4* xs[2] = {s->XsA.p, s->XsB.p};
    // the warm chain: x*_w stays in XsW (for the taps); later iterates alternate between XsC
    // (scratch) and XsB (which the fallback may overwrite only
    // in scenes whose gate is set, where the warm chain's result is discarded)
    auto ws = [&](int j) { return j == 0 ? s->XsW.p : (j & 1) ? s->XsC.p : s->XsB.p; };
    if ((learned || record) && !fusedFeatures) {
        L.before("features");
        launch_features(st, d, P, record, bs);
        L.after("features");
    }
    if (learned) {
        L.before("mlp_apply");
        launch_mlp_apply(st, d, s->net.p, P.h, s->XsA.p, s->XsW.p, s->lamHat.p);
        L.after("mlp_apply");
        // J iterations from the warm start; the first density pass is the residual check
        for (int j = 0; j < J; ++j) {
            const float4* in = ws(j);
            L.before("density_lambda_warm");
            launch_density_lambda(st, d, P, in, bs, j == 0 ? 1 : -1, false);
            L.after("density_lambda_warm");
            if (j == 0) launch_gate(st, d, (uint32_t)s->scenes.size(), s->gateStat, s->gateRatio, s->gateTol);
            if (j + 1 < J) {
                L.before("delta_correct_warm");
                launch_delta_correct(st, d, P, S, in, ws(j + 1), false, true, bs);
                L.after("delta_correct_warm");
            }
        }
    }
    if (I == 0) {
        L.before("update_damping");
        launch_delta_correct(st, d, P, S, xs[0], nullptr, true, false, bs);
        L.after("update_damping");
    }
    for (int it = 0; it < I; ++it) {
        const float4* in = xs[it & 1];
        float4* out = xs[(it + 1) & 1];
        const bool last = it + 1 == I;
        if (!(fusedDensity && it == 0)) {
            L.before(learned ? "density_lambda_fallback" : "density_lambda");
            if (queued)
                launch_density_lambda_queue(st, d, P, in, pass++);
            else
                launch_density_lambda(st, d, P, in, bs, record && it == I - Jt ? 2 : -1, learned, lanes);
            L.after(learned ? "density_lambda_fallback" : "density_lambda");
        }
        if (record && it == I - Jt) {
            L.before("targets");
            launch_targets(st, d, P, in);
            launch_resid_reduce(st, d, (uint32_t)s->scenes.size(), 0x5);
            L.after("targets");
        }
        const char* name = last ? "delta_correct_update" : learned ? "delta_correct_fallback" : "delta_correct";
        L.before(name);
        // learned: the fallback's iterations run only where the gate is set; the final
        // correction reads the warm chain's last iterate where it is clear
        if (queued)
            launch_delta_correct_queue(st, d, P, S, in, out, last, pass++);
        else
            launch_delta_correct(st, d, P, S, in, out, last, true, bs, learned ? (last ? 2 : 1) : 0,
                                 learned ? ws(J - 1) : nullptr, lanes);
        L.after(name);
    }
    L.before("vorticity");
    if (queued && postLanes == 1)
        launch_vorticity_queue(st, d, P, pass++);
    else
        launch_vorticity(st, d, P, bs, postLanes);
    L.after("vorticity");
    L.before("confinement");
    if (queued)
        launch_confinement_queue(st, d, P, S, pass++);
    else
        launch_confinement(st, d, P, S, bs, lanes);
    L.after("confinement");
    L.before("xsph");
    if (queued && postLanes == 1)
        launch_xsph_queue(st, d, P, pass++);
    else
        launch_xsph(st, d, P, bs, postLanes);
    L.after("xsph");
}

void upload_scenes(akua_solver* s) {
    if (!s->scenesDirty) return;
    HIP_CHECK(hipStreamSynchronize(s->stream));
    HIP_CHECK(hipMemcpy(s->dScenes.p, s->scenes.data(), s->scenes.size() * sizeof(SceneDev),
                        hipMemcpyHostToDevice));
    s->scenesDirty = false;
}

SortPlan make_plan(int bits, int digitBits) {
    SortPlan p = SortPlan::balanced(bits);
    if (digitBits >= 4 && digitBits <= kMaxDigitBits) p.digitBits = digitBits;
    return p;
}

// Super-cell layout of the scenes from their boxes: bits per axis (at most 10, the same for every
// scene) cover the longest axis with kLayoutMarginCells hash cells (at least one super-cell) to
// spare beyond each wall, and each axis's key range is centred on the box, so predicted positions
// that overshoot a wall keep keys of their own.  Positions beyond the range are clamped to its edge,
// which puts several cells under one key: that costs locality only, and with the layout sorted in
// the previous logical order it also splits their buckets (a key's cells then interleave).
constexpr int kLayoutMarginCells = 4;

void update_layout(akua_solver* s) {
    int bits = 1;
    const int L = s->tuning.layoutLevel;
    const float span = (float)(1 << L);
    const int margin = std::max(1, (kLayoutMarginCells + (1 << L) - 1) >> L);  // in super-cells
    for (SceneDev& sc : s->scenes) {
        int ext = 1;
        for (int a = 0; a < 3; ++a) {
            const int lo = (int)std::floor(std::floor(sc.bmin[a] / s->sph.cellHash) / span);
            const int hi = (int)std::floor(std::floor(sc.bmax[a] / s->sph.cellHash) / span);
            ext = std::max(ext, hi - lo + 1 + 2 * margin);
        }
        bits = std::max(bits, (int)std::min<uint32_t>(10u, std::max<uint32_t>(1u, ceil_log2((uint64_t)ext))));
    }
    for (SceneDev& sc : s->scenes) {
        for (int a = 0; a < 3; ++a) {
            const int lo = (int)std::floor(std::floor(sc.bmin[a] / s->sph.cellHash) / span);
            const int hi = (int)std::floor(std::floor(sc.bmax[a] / s->sph.cellHash) / span);
            sc.layoutOrigin[a] = lo - ((1 << bits) - (hi - lo + 1)) / 2;
        }
        sc.layoutOrigin[3] = L;
    }
    for (SceneDev& sc : s->scenes) sc.layoutBits = (uint32_t)bits;
    s->plan1 = make_plan(s->plan1.bits, s->tuning.hashDigitBits);
    // single-cell super-cells with the layout digit width left to the solver: the hash sort's
    // width, so that the two sorts share their launches
    const int layoutDigits = L == 0 && s->tuning.layoutDigitBits == 0 ? s->plan1.digitBits : s->tuning.layoutDigitBits;
    s->plan2 = make_plan(3 * bits, layoutDigits);
    s->scenesDirty = true;
}

// Neighbour search cells per smoothing radius (SphParams::searchCells): 1 the reference's, 2 the
// non-parity search over h / 2 cells.  The hash cell size follows; the step graphs re-capture.
void apply_search(akua_solver* s, int cells) {
    if (s->sph.searchCells == cells) return;
    const float h = s->cfg.smoothRadius;
    s->sph.searchCells = cells;
    s->sph.cellHash = cells == 2 ? h * 0.5f : h;
    s->sph.cellSearch = cells == 2 ? h * 0.5f : s->cfg.spatialHashCellSize;
    // the packed cell keys of the h / 2 search take 30 bits
    s->plan1 = make_plan(cells == 2 ? 30 : s->hashBits, s->tuning.hashDigitBits);
    s->graphDirty = true;
    update_layout(s);
}

void set_box(akua_solver* s, size_t scene, const float* bmin, const float* bmax) {
    SceneDev& sc = s->scenes[scene];
    s->boxSeen = true;
    bool changed = false;
    for (int k = 0; k < 3; ++k) {
        if (sc.bmin[k] != bmin[k] || sc.bmax[k] != bmax[k]) changed = true;
        sc.bmin[k] = bmin[k];
        sc.bmax[k] = bmax[k];
    }
    if (changed) {
        s->scenesDirty = true;
        s->bndDirty = true;
        update_layout(s);
    }
}

// Boundary samples for the current box (single-scene solvers): rebuilt when the settings or the
// box changed; the step graphs capture SphParams::bd, so a rebuild re-captures them.
void ensure_boundary(akua_solver* s) {
    if (s->bndLayers <= 0) {
        if (s->sph.bd.on) {
            s->sph.bd = BoundaryDev{};
            s->graphDirty = true;
        }
        return;
    }
    if (!s->bndDirty && s->sph.bd.on) return;
    const SceneDev& sc = s->scenes[0];
    BoundaryDev B{};
    std::vector<float> psi;
    std::string err;
    const float sp = s->bndSpacing > 0.0f ? s->bndSpacing : s->cfg.particle_spacing;
    if (!build_boundary(sc.bmin, sc.bmax, sp, s->bndLayers, s->sph, s->cfg.restDensity, B, psi, err))
        fail(AKUA_ERR_INVALID_ARG, err);
    HIP_CHECK(hipStreamSynchronize(s->stream));  // earlier steps may still read the old weights
    if (s->bndPsi.n < psi.size()) s->bndPsi.alloc(psi.size());
    HIP_CHECK(hipMemcpy(s->bndPsi.p, psi.data(), psi.size() * sizeof(float), hipMemcpyHostToDevice));
    if (s->Bg.n < s->nslots) s->Bg.alloc(s->nslots);
    B.psi = s->bndPsi.p;
    s->sph.bd = B;
    s->bndDirty = false;
    s->graphDirty = true;
}

// Buffers of the balanced work order, allocated the first time a setting uses them (before any
// capture: a graph records no allocation).
void ensure_balance(akua_solver* s) {
    if (s->tuning.balanceTile <= 0 || s->nbrW.p) return;
    HIP_CHECK(hipStreamSynchronize(s->stream));
    s->work.alloc(s->nslots);
    s->cntW.alloc(s->nslots);
    s->nbrW.alloc(s->nbr.n);
    s->graphDirty = true;
}

// Buffers of the sweeps' work queue, allocated the first time a setting uses them (before any
// capture) and again when more passes are needed.
void ensure_queue(akua_solver* s) {
    if (s->tuning.workQueue <= 0) return;
    const uint32_t passes = queue_passes(s);
    if (s->queue.p && s->heads.n >= queue_heads_size(passes)) return;
    HIP_CHECK(hipStreamSynchronize(s->stream));
    const uint32_t ntiles = (s->nslots + kQueueTileSlots - 1) / kQueueTileSlots;
    const uint32_t t8 = (ntiles + kXcdCount - 1) / kXcdCount;
    s->qcost.alloc(std::max<uint32_t>(ntiles, 1));
    s->queue.alloc((size_t)std::max<uint32_t>(t8, 1) * kXcdCount);
    s->qcount.alloc(kXcdCount);
    s->heads.alloc(queue_heads_size(passes));
    s->graphDirty = true;
}

void run_steps(akua_solver* s, float dt, int nsteps) {
    if (!s->hasState) fail(AKUA_ERR_STATE, "set the particles before stepping");
    if (!(dt > 0.0f) || !std::isfinite(dt)) fail(AKUA_ERR_INVALID_ARG, "deltaTime must be > 0");
    upload_scenes(s);
    ensure_boundary(s);
    ensure_balance(s);
    ensure_queue(s);
    StepParams S{dt, {s->cfg.gravity[0], s->cfg.gravity[1], s->cfg.gravity[2]}};
    if (s->graphDirty) {
        for (auto& g : s->graphs) g.dirty = true;
        s->graphDirty = false;
    }
    for (int k = 0; k < nsteps; ++k) {
        // the physical layout is re-sorted every layoutInterval steps; it never changes a result
        const bool keep = s->layoutPhase != 0;
        if (s->timing) {
            enqueue_step(s, s->stream, S, true, false, keep);
            collect_timing(s);
        } else if (s->useGraph) {
            auto& G = s->graphs[keep ? 1 : 0];
            const bool same = G.exec && !G.dirty && std::memcmp(&S, &G.params, sizeof(S)) == 0 &&
                              G.plan2 == s->plan2.bits &&
                              std::memcmp(&G.tuning, &s->tuning, sizeof(akua_tuning)) == 0;
            if (!same) {
                if (G.exec) {
                    HIP_CHECK(hipGraphExecDestroy(G.exec));
                    G.exec = nullptr;
                }
                hipGraph_t g;
                HIP_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                enqueue_step(s, s->stream, S, false, false, keep);
                HIP_CHECK(hipStreamEndCapture(s->stream, &g));
                hipError_t e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIP_CHECK(e);
                G.params = S;
                G.plan2 = s->plan2.bits;
                G.tuning = s->tuning;
                G.dirty = false;
            }
            HIP_CHECK(hipGraphLaunch(G.exec, s->stream));
        } else {
            enqueue_step(s, s->stream, S, false, false, keep);
        }
        HIP_CHECK(hipGetLastError());
        s->layoutPhase = (s->layoutPhase + 1) % std::max(1, s->tuning.layoutInterval);
        s->stepsSinceStats++;
        if (s->mode == 1) s->learnedSteps++;
    }
}

// Exact density at the current positions (k_density_probe): hashes and sorts the positions as a
// step with dt = 0 would, then walks each distinct stencil bucket once.  Touches only step scratch
// (the sort buffers, X0, XsA, the cell table), never the particle state.
void run_density_probe(akua_solver* s) {
    if (!s->hasState) fail(AKUA_ERR_STATE, "set the particles first");
    upload_scenes(s);
    const StepParams S0{0.0f, {0.0f, 0.0f, 0.0f}};
    hipStream_t st = s->stream;
    DeviceState d = s->devState();
    // the probe's own walk uses cells of h whatever search the steps use
    SphParams P0 = s->sph;
    P0.cellHash = s->cfg.smoothRadius;
    P0.cellSearch = s->cfg.smoothRadius;
    P0.searchCells = 1;
    HIP_CHECK(hipMemsetAsync(s->probeMax.p, 0, s->scenes.size() * sizeof(unsigned int), st));
    HIP_CHECK(hipMemsetAsync(s->probeSum.p, 0, s->scenes.size() * sizeof(double), st));
    launch_clear_table(st, s->table.p, s->tableTotal);
    launch_predict_hash(st, d, P0, S0, false, false);
    sort_hash(s, st, false);
    launch_cell_keys(st, d, P0, S0, false, false);
    sort_layout(s, st, false);
    launch_gather(st, d, P0, S0, false, false, false);
    launch_bucket_phys(st, d);
    launch_density_probe(st, d, P0, s->probeRho.p, s->probeMax.p, s->probeSum.p);
    HIP_CHECK(hipGetLastError());
}

void check_n(akua_solver* s, int32_t n) {
    if (n != (int32_t)s->nparticles)
        fail(AKUA_ERR_INVALID_ARG, "n does not match the solver's particle count");
}

void upload_soa(akua_solver* s, const float* pos, const float* vel, const float* mass,
                const float* dens, const float* lam, const float* color, const float* size,
                bool fromAos, const akua_particle* aos) {
    const size_t n = s->nparticles;
    std::vector<float4> X(s->nslots, make_float4(0, 0, 0, 0)), V(s->nslots, make_float4(0, 0, 0, 0));
    std::vector<float4> VS(s->nslots, make_float4(0, 0, 0, 0)), DP(s->nslots, make_float4(0, 0, 0, 0));
    std::vector<float4> C(s->nslots, make_float4(0, 0, 1, 1));
    std::vector<float> D(s->nslots, 0.0f), L(s->nslots, 0.0f), Z(s->nslots, 0.0f);
    std::vector<uint32_t> I(s->nslots, 0);
    // particles fill the scenes' slot ranges in scene order
    size_t k = 0;
    bool unit = true;
    for (const SceneDev& sc : s->scenes) {
        for (uint32_t slot = sc.begin; slot < sc.end; ++slot, ++k) {
            float p[3], v[3], m, dd, ll;
            if (fromAos) {
                const akua_particle& a = aos[k];
                for (int c = 0; c < 3; ++c) { p[c] = a.position[c]; v[c] = a.velocity[c]; }
                m = a.mass; dd = a.density; ll = a.lambda;
                VS[slot] = make_float4(a.new_velocity[0], a.new_velocity[1], a.new_velocity[2], 0.0f);
                DP[slot] = make_float4(a.position_delta[0], a.position_delta[1], a.position_delta[2], 0.0f);
                C[k] = make_float4(a.color[0], a.color[1], a.color[2], a.color[3]);
                Z[k] = a.size;
            } else {
                for (int c = 0; c < 3; ++c) { p[c] = pos[3 * k + c]; v[c] = vel[3 * k + c]; }
                m = mass ? mass[k] : 1.0f;
                dd = dens ? dens[k] : 0.0f;
                ll = lam ? lam[k] : 0.0f;
                if (color) C[k] = make_float4(color[4 * k], color[4 * k + 1], color[4 * k + 2], color[4 * k + 3]);
                if (size) Z[k] = size[k];
            }
            for (int c = 0; c < 3; ++c)
                if (!std::isfinite(p[c]) || !std::isfinite(v[c]))
                    fail(AKUA_ERR_NON_FINITE, "particle " + std::to_string(k) + " has a non-finite position or velocity");
            if (!std::isfinite(m)) fail(AKUA_ERR_NON_FINITE, "non-finite mass");
            unit = unit && m == 1.0f;
            X[slot] = make_float4(p[0], p[1], p[2], m);
            V[slot] = make_float4(v[0], v[1], v[2], 0.0f);
            D[slot] = dd;
            L[slot] = ll;
            I[slot] = (uint32_t)k;
        }
    }
    (void)n;
    HIP_CHECK(hipStreamSynchronize(s->stream));
    HIP_CHECK(hipMemcpy(s->X.p, X.data(), X.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->V.p, V.data(), V.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->VS.p, VS.data(), VS.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->DP.p, DP.data(), DP.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->dens.p, D.data(), D.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->lam.p, L.data(), L.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->id.p, I.data(), I.size() * 4, hipMemcpyHostToDevice));
    std::vector<uint32_t> ident(s->nslots);
    for (uint32_t i = 0; i < s->nslots; ++i) ident[i] = i;
    HIP_CHECK(hipMemcpy(s->lid.p, ident.data(), ident.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->color.p, C.data(), C.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(s->size.p, Z.data(), Z.size() * 4, hipMemcpyHostToDevice));
    s->hasState = true;
    s->layoutPhase = 0;  // the uploaded order is not cell ordered: re-sort on the next step
    s->unitMass = unit;
    update_fast(s);
}

template <class T>
std::vector<T> download(akua_solver* s, const T* p, size_t n) {
    std::vector<T> h(n);
    HIP_CHECK(hipStreamSynchronize(s->stream));
    if (n) HIP_CHECK(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

// ---- tuning ---------------------------------------------------------------------------------

void validate_tuning(const akua_tuning& t) {
    if (t.block != 128 && t.block != 256 && t.block != 512)
        fail(AKUA_ERR_INVALID_ARG, "tuning.block must be 128, 256 or 512");
    if (t.layoutLevel < 0 || t.layoutLevel > 3)
        fail(AKUA_ERR_INVALID_ARG, "tuning.layoutLevel must be in [0, 3]");
    for (int b : {t.hashDigitBits, t.layoutDigitBits})
        if (b != 0 && (b < 5 || b > kMaxDigitBits))
            fail(AKUA_ERR_INVALID_ARG, "tuning digit widths must be 0 or 5..8");
    if (t.layoutInterval < 1 || t.layoutInterval > 8)
        fail(AKUA_ERR_INVALID_ARG, "tuning.layoutInterval must be in [1, 8]");
    if (t.balanceTile != 0 && (t.balanceTile < 256 || t.balanceTile > (int)kBalanceMaxTile ||
                               t.balanceTile % 256 != 0))
        fail(AKUA_ERR_INVALID_ARG, "tuning.balanceTile must be 0 or a multiple of 256 in [256, 2048]");
    if (t.slabBatch != 0 && t.slabBatch != 4 && t.slabBatch != 8)
        fail(AKUA_ERR_INVALID_ARG, "tuning.slabBatch must be 0, 4 or 8");
    if (t.fuseSweep != 0 && t.fuseSweep != 1) fail(AKUA_ERR_INVALID_ARG, "tuning.fuseSweep must be 0 or 1");
    if (t.laneGroup != 0 && t.laneGroup != 2 && t.laneGroup != 4 && t.laneGroup != 8 && t.laneGroup != 16)
        fail(AKUA_ERR_INVALID_ARG, "tuning.laneGroup must be 0, 2, 4, 8 or 16");
    if (t.workQueue < 0 || t.workQueue > 2) fail(AKUA_ERR_INVALID_ARG, "tuning.workQueue must be 0, 1 or 2");
}

void apply_tuning(akua_solver* s, const akua_tuning& t) {
    validate_tuning(t);
    if (t.layoutInterval != s->tuning.layoutInterval) s->layoutPhase = 0;
    s->tuning = t;
    update_layout(s);
}

// The per-slot state a step reads: a trial restarts from it and the tuner restores it.
struct Snapshot {
    DevBuf<float4> X, V, W, VS, DP;
    DevBuf<float> dens, lam;
    DevBuf<uint32_t> id, lid;
};

template <class T>
void copy_dev(T* dst, const T* src, size_t n, hipStream_t st) {
    HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, st));
}

void take_snapshot(akua_solver* s, Snapshot& q) {
    const size_t n = s->nslots;
    q.X.alloc(n); q.V.alloc(n); q.W.alloc(n); q.VS.alloc(n); q.DP.alloc(n);
    q.dens.alloc(n); q.lam.alloc(n);
    q.id.alloc(n); q.lid.alloc(n);
    hipStream_t st = s->stream;
    copy_dev(q.X.p, s->X.p, n, st); copy_dev(q.V.p, s->V.p, n, st); copy_dev(q.W.p, s->W.p, n, st);
    copy_dev(q.VS.p, s->VS.p, n, st); copy_dev(q.DP.p, s->DP.p, n, st);
    copy_dev(q.dens.p, s->dens.p, n, st); copy_dev(q.lam.p, s->lam.p, n, st);
    copy_dev(q.id.p, s->id.p, n, st); copy_dev(q.lid.p, s->lid.p, n, st);
}

void restore_snapshot(akua_solver* s, const Snapshot& q) {
    const size_t n = s->nslots;
    s->layoutPhase = 0;
    hipStream_t st = s->stream;
    copy_dev(s->X.p, q.X.p, n, st); copy_dev(s->V.p, q.V.p, n, st); copy_dev(s->W.p, q.W.p, n, st);
    copy_dev(s->VS.p, q.VS.p, n, st); copy_dev(s->DP.p, q.DP.p, n, st);
    copy_dev(s->dens.p, q.dens.p, n, st); copy_dev(s->lam.p, q.lam.p, n, st);
    copy_dev(s->id.p, q.id.p, n, st); copy_dev(s->lid.p, q.lid.p, n, st);
}

// Device milliseconds per step of `steps` graph-launched steps from the snapshot, after one
// untimed step that instantiates the graph for the setting.
// Milliseconds per step of the current setting from the snapshot: one untimed step (which
// captures the setting's graphs), then the faster of two timed runs of `steps` steps.
double time_trial(akua_solver* s, const Snapshot& q, float dt, int steps) {
    restore_snapshot(s, q);
    run_steps(s, dt, 1);
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    double best = 0.0;
    hipError_t e = hipSuccess;
    for (int r = 0; r < 2 && e == hipSuccess; ++r) {
        HIP_CHECK(hipEventRecord(a, s->stream));
        run_steps(s, dt, steps);
        HIP_CHECK(hipEventRecord(b, s->stream));
        HIP_CHECK(hipEventSynchronize(b));
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, a, b);
        if (r == 0 || ms < best) best = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    HIP_CHECK(e);
    return best / steps;
}

std::string tuning_key(akua_solver* s) {
    hipDeviceProp_t prop{};
    HIP_CHECK(hipGetDeviceProperties(&prop, s->device));
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    std::ostringstream k;
    k << arch << " n2^" << ceil_log2(std::max<uint64_t>(1, s->nparticles)) << " scenes2^"
      << ceil_log2(std::max<uint64_t>(1, s->scenes.size())) << " cap" << s->cfg.maxNeighbours
      << " it" << s->cfg.solverIterations << (s->exact ? "" : " inexact");
    return k.str();
}

// Cache lines: "<key> | block level hashBits layoutBits interval balanceTile slabBatch
// searchCells fuseSweep laneGroup workQueue".  The last line for a key wins; a line without the last fields
// (written before they existed) reads them as 0.
bool cache_lookup(const char* path, const std::string& key, akua_tuning& t, int& cells) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        const size_t bar = line.find(" | ");
        if (bar == std::string::npos || line.compare(0, bar, key) != 0 || bar != key.size()) continue;
        std::istringstream v(line.substr(bar + 3));
        akua_tuning c{};
        int sc = 1;
        if (!(v >> c.block >> c.layoutLevel >> c.hashDigitBits >> c.layoutDigitBits >> c.layoutInterval >>
              c.balanceTile >> c.slabBatch >> sc))
            continue;
        if (!(v >> c.fuseSweep)) c.fuseSweep = 0;
        if (!(v >> c.laneGroup)) c.laneGroup = 0;
        if (!(v >> c.workQueue)) c.workQueue = 0;
        try {
            validate_tuning(c);
        } catch (const AkuaError&) {
            continue;
        }
        if (sc != 1 && sc != 2) continue;
        t = c;
        cells = sc;
        found = true;
    }
    return found;
}

void cache_store(const char* path, const std::string& key, const akua_tuning& t, int cells) {
    std::ofstream f(path, std::ios::app);
    if (!f) fail(AKUA_ERR_INVALID_ARG, std::string("cannot append to the tuning cache ") + path);
    f << key << " | " << t.block << ' ' << t.layoutLevel << ' ' << t.hashDigitBits << ' '
      << t.layoutDigitBits << ' ' << t.layoutInterval << ' ' << t.balanceTile << ' ' << t.slabBatch
      << ' ' << cells << ' ' << t.fuseSweep << ' ' << t.laneGroup << ' ' << t.workQueue << '\n';
}

// ---- learned warm start ----------------------------------------------------------------------

void ensure_learned(akua_solver* s) {
    if (s->net.p) return;
    const size_t ns = s->nslots, nsc = s->scenes.size();
    s->net.alloc(kMlpParams);
    s->feat.alloc(ns * kMlpFeatures);
    s->lamHat.alloc(ns);
    s->XsW.alloc(ns);
    s->resid.alloc(kResidStride * nsc);
    s->residPart.alloc(3 * ((ns + 63) / 64));
    HIP_CHECK(hipMemset(s->residPart.p, 0, 3 * ((ns + 63) / 64) * sizeof(float2)));
    uint32_t maxScene = 1;
    for (const SceneDev& sc : s->scenes) maxScene = std::max(maxScene, sc.end - sc.begin);
    s->residBlocks = resid_blocks_for(maxScene);
    s->residBlk.alloc((size_t)nsc * s->residBlocks * 3);
    HIP_CHECK(hipMemset(s->residBlk.p, 0, (size_t)nsc * s->residBlocks * 3 * sizeof(float2)));
    s->gate.alloc(nsc);
    s->fallbacks.alloc(nsc);
    HIP_CHECK(hipMemset(s->net.p, 0, kMlpParams * sizeof(float)));
    HIP_CHECK(hipMemset(s->feat.p, 0, ns * kMlpFeatures * sizeof(float)));
    HIP_CHECK(hipMemset(s->lamHat.p, 0, ns * sizeof(float)));
    HIP_CHECK(hipMemset(s->XsW.p, 0, ns * sizeof(float4)));
    HIP_CHECK(hipMemset(s->resid.p, 0, kResidStride * nsc * sizeof(unsigned int)));
    HIP_CHECK(hipMemset(s->gate.p, 0, nsc * sizeof(uint32_t)));
    HIP_CHECK(hipMemset(s->fallbacks.p, 0, nsc * sizeof(uint32_t)));
    s->graphDirty = true;
}

void reset_adam(akua_solver* s) {
    s->adamT = 0;
    if (s->adamM.p) {
        HIP_CHECK(hipMemset(s->adamM.p, 0, kMlpTrainable * sizeof(float)));
        HIP_CHECK(hipMemset(s->adamV.p, 0, kMlpTrainable * sizeof(float)));
    }
}

void ensure_train(akua_solver* s, uint32_t batch, uint32_t losses) {
    ensure_learned(s);
    const size_t ns = s->nslots;
    if (!s->XsP.p) {
        s->XsP.alloc(ns);
        s->targ.alloc(ns * 4);
        s->adamM.alloc(kMlpTrainable);
        s->adamV.alloc(kMlpTrainable);
        s->part.alloc((size_t)kMlpPartRows * kMlpPartRow);
        HIP_CHECK(hipMemset(s->XsP.p, 0, ns * sizeof(float4)));
        HIP_CHECK(hipMemset(s->targ.p, 0, ns * 4 * sizeof(float)));
        HIP_CHECK(hipMemset(s->part.p, 0, (size_t)kMlpPartRows * kMlpPartRow * sizeof(float)));
        reset_adam(s);
    }
    if (batch > s->batchCap) {
        s->batchF.release();
        s->batchT.release();
        s->batchF.alloc((size_t)batch * kMlpFeatures);
        s->batchT.alloc((size_t)batch * 4);
        s->batchCap = batch;
    }
    if (losses > s->lossCap) {
        s->lossBuf.release();
        s->lossBuf.alloc(losses);
        s->lossCap = losses;
    }
}

void check_network(const float* p) {
    for (int k = 0; k < kMlpParams; ++k)
        if (!std::isfinite(p[k])) fail(AKUA_ERR_NON_FINITE, "network parameter " + std::to_string(k) + " is not finite");
    for (int k = 0; k < kMlpOut; ++k)
        if (p[kOffYmul + k] == 0.0f) fail(AKUA_ERR_INVALID_ARG, "network output scale is zero");
}

void upload_network(akua_solver* s, const float* p) {
    check_network(p);
    ensure_learned(s);
    HIP_CHECK(hipStreamSynchronize(s->stream));
    HIP_CHECK(hipMemcpy(s->net.p, p, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    s->haveNet = true;
    reset_adam(s);
}

// splitmix64 stream and a normal deviate by Box-Muller.
struct HostRng {
    uint64_t z;
    uint64_t next() {
        uint64_t x = (z += 0x9E3779B97F4A7C15ull);
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return x ^ (x >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

// He-normal weights, zero biases, the last layer scaled down so that the untrained network's
// warm start is close to none.  Normalisation: feature means and standard deviations, and the
// targets' root mean squares, from the given FP64 sums over count samples.
struct SampleStats {
    double count = 0.0;
    double fs[kMlpIn] = {}, fq[kMlpIn] = {}, tq[kMlpOut] = {};
    void add(const std::vector<float>& F, const std::vector<float>& T, uint32_t n) {
        for (uint32_t b = 0; b < n; ++b) {
            for (int f = 0; f < kMlpIn; ++f) {
                const double x = F[(size_t)b * kMlpIn + f];
                fs[f] += x;
                fq[f] += x * x;
            }
            for (int o = 0; o < kMlpOut; ++o) {
                const double t = T[(size_t)b * kMlpOut + o];
                tq[o] += t * t;
            }
        }
        count += n;
    }
};

std::vector<float> initial_network(uint64_t seed, const SampleStats& st) {
    std::vector<float> p(kMlpParams, 0.0f);
    HostRng rng{seed ^ 0x6A09E667F3BCC909ull};
    for (int k = 0; k < kMlpIn * kMlpHid; ++k) p[kOffW1 + k] = (float)(rng.normal() * std::sqrt(2.0 / kMlpIn));
    for (int k = 0; k < kMlpHid * kMlpHid; ++k) p[kOffW2 + k] = (float)(rng.normal() * std::sqrt(2.0 / kMlpHid));
    for (int k = 0; k < kMlpHid * kMlpOut; ++k) p[kOffW3 + k] = (float)(rng.normal() * 0.1 * std::sqrt(1.0 / kMlpHid));
    for (int f = 0; f < kMlpIn; ++f) {
        const double m = st.fs[f] / st.count;
        const double var = std::max(0.0, st.fq[f] / st.count - m * m);
        const double sd = std::sqrt(var);
        p[kOffMean + f] = (float)m;
        p[kOffInv + f] = sd > 1e-30 && std::isfinite(sd) ? (float)(1.0 / sd) : 1.0f;
    }
    for (int o = 0; o < kMlpOut; ++o) {
        const double rms = std::sqrt(st.tq[o] / st.count);
        p[kOffYmul + o] = rms > 1e-30 && std::isfinite(rms) ? (float)rms : 1.0f;
    }
    return p;
}

void validate_train_config(const akua_train_config& c) {
    if (c.samplesPerStep < 16 || c.samplesPerStep > (1 << 22))
        fail(AKUA_ERR_INVALID_ARG, "samplesPerStep must be in [16, 2^22]");
    if (c.minibatch < 16 || c.minibatch > c.samplesPerStep)
        fail(AKUA_ERR_INVALID_ARG, "minibatch must be in [16, samplesPerStep]");
    if (c.epochs < 1 || c.epochs > 64) fail(AKUA_ERR_INVALID_ARG, "epochs must be in [1, 64]");
    if (c.normSteps < 1 || c.normSteps > 100000) fail(AKUA_ERR_INVALID_ARG, "normSteps must be in [1, 10^5]");
    if (!(c.learningRate > 0.0f) || !std::isfinite(c.learningRate))
        fail(AKUA_ERR_INVALID_ARG, "learningRate must be > 0");
    for (float w : c.lossWeights)
        if (!(w >= 0.0f) || !std::isfinite(w)) fail(AKUA_ERR_INVALID_ARG, "loss weights must be >= 0");
}

// Online training on the solver's own plain trajectory: each step is a plain step that also
// records features and targets; samplesPerStep random particles of it feed epochs passes of
// Adam updates over minibatches.  A solver without a network first gathers normalisation
// statistics over normSteps steps (no updates), then starts from He-normal weights.
// resid[4k..4k+3] of step k: max and mean over particles of max(0, C) at x*_pred, then max and
// mean at the last plain density pass.
void train(akua_solver* s, float dt, const float* bmin, const float* bmax, int steps,
           const akua_train_config& c, float* lossOut, float* residOut) {
    if (s->scenes.size() != 1) fail(AKUA_ERR_UNSUPPORTED, "training runs on a single-scene solver");
    if (!s->hasState) fail(AKUA_ERR_STATE, "set the particles before training");
    if (!(dt > 0.0f) || !std::isfinite(dt)) fail(AKUA_ERR_INVALID_ARG, "deltaTime must be > 0");
    if (steps < 1 || steps > 1000000) fail(AKUA_ERR_INVALID_ARG, "steps must be in [1, 10^6]");
    if (s->cfg.solverIterations < 1) fail(AKUA_ERR_UNSUPPORTED, "training needs solverIterations >= 1");
    validate_train_config(c);
    if (c.targetIterations < 1 || c.targetIterations > s->cfg.solverIterations)
        fail(AKUA_ERR_INVALID_ARG, "targetIterations must be in [1, solverIterations]");
    s->trainIters = c.targetIterations;
    check_box(bmin, bmax);
    set_box(s, 0, bmin, bmax);
    upload_scenes(s);
    const uint32_t B = (uint32_t)c.samplesPerStep, M = (uint32_t)c.minibatch;
    const uint32_t nb = B / M;
    const uint32_t perStep = nb * (uint32_t)c.epochs;
    ensure_train(s, B, perStep * (uint32_t)steps);
    ensure_balance(s);
    HIP_CHECK(hipMemset(s->lossBuf.p, 0, (size_t)perStep * steps * sizeof(float)));
    hipStream_t st = s->stream;
    const StepParams S{dt, {s->cfg.gravity[0], s->cfg.gravity[1], s->cfg.gravity[2]}};
    MlpLossWeights lw{{c.lossWeights[0], c.lossWeights[1], c.lossWeights[2], c.lossWeights[3]}};
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const double count = (double)s->nparticles;
    std::vector<unsigned int> res(kResidStride);
    SampleStats stats;
    int normLeft = s->haveNet ? 0 : c.normSteps;
    std::vector<char> trained(steps, 0);
    for (int k = 0; k < steps; ++k) {
        enqueue_step(s, st, S, false, true);
        s->stepsSinceStats++;
        DeviceState d = s->devState();
        launch_sample(st, d, c.seed + 0x2545F4914F6CDD1Dull * (uint64_t)(s->adamT + k + 1), B,
                      s->batchF.p, s->batchT.p);
        HIP_CHECK(hipGetLastError());
        if (normLeft > 0) {
            std::vector<float> F((size_t)B * kMlpFeatures), T((size_t)B * 4);
            HIP_CHECK(hipStreamSynchronize(st));
            HIP_CHECK(hipMemcpy(F.data(), s->batchF.p, F.size() * 4, hipMemcpyDeviceToHost));
            HIP_CHECK(hipMemcpy(T.data(), s->batchT.p, T.size() * 4, hipMemcpyDeviceToHost));
            stats.add(F, T, B);
            if (--normLeft == 0) {
                const std::vector<float> p = initial_network(c.seed, stats);
                upload_network(s, p.data());
            }
        } else {
            for (int e = 0; e < c.epochs; ++e) {
                for (uint32_t mb = 0; mb < nb; ++mb) {
                    launch_mlp_grad(st, s->net.p, s->batchF.p + (size_t)mb * M * kMlpFeatures,
                                    s->batchT.p + (size_t)mb * M * 4, M, lw, s->part.p);
                    s->adamT++;
                    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)s->adamT)));
                    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)s->adamT)));
                    launch_mlp_adam(st, s->part.p, s->net.p, s->adamM.p, s->adamV.p,
                                    c.learningRate, b1, b2, eps, c1, c2,
                                    s->lossBuf.p + (size_t)k * perStep + (size_t)e * nb + mb);
                }
            }
            trained[k] = 1;
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(st));
        HIP_CHECK(hipMemcpy(res.data(), s->resid.p, kResidStride * sizeof(unsigned int), hipMemcpyDeviceToHost));
        if (residOut) {
            float v[kResidStride];
            std::memcpy(v, res.data(), sizeof(v));
            residOut[4 * k + 0] = v[0];
            residOut[4 * k + 1] = (float)(v[kResidSum] / count);
            residOut[4 * k + 2] = v[2];
            residOut[4 * k + 3] = (float)(v[kResidSum + 2] / count);
        }
    }
    std::vector<float> L((size_t)perStep * steps);
    HIP_CHECK(hipMemcpy(L.data(), s->lossBuf.p, L.size() * 4, hipMemcpyDeviceToHost));
    if (lossOut)
        for (int k = 0; k < steps; ++k) {
            double a = 0.0;
            for (uint32_t u = 0; u < perStep; ++u) a += L[(size_t)k * perStep + u];
            lossOut[k] = trained[k] ? (float)(a / perStep) : std::nanf("");
        }
}

void need_device() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        fail(AKUA_ERR_NO_DEVICE, "no HIP device is visible");
}

template <class T>
void to_device(DevBuf<T>& b, const T* h, size_t n) {
    b.alloc(n);
    if (n) HIP_CHECK(hipMemcpy(b.p, h, n * sizeof(T), hipMemcpyHostToDevice));
}

void autotune(akua_solver* s, float dt, const float* bmin, const float* bmax, int trialSteps,
              const char* cachePath) {
    if (!s->hasState) fail(AKUA_ERR_STATE, "set the particles before tuning");
    if (!(dt > 0.0f) || !std::isfinite(dt)) fail(AKUA_ERR_INVALID_ARG, "deltaTime must be > 0");
    if (trialSteps < 1 || trialSteps > 1000) fail(AKUA_ERR_INVALID_ARG, "trialSteps must be in [1, 1000]");
    if ((bmin == nullptr) != (bmax == nullptr)) fail(AKUA_ERR_INVALID_ARG, "give both box corners or neither");
    if (bmin) {
        check_box(bmin, bmax);
        for (size_t i = 0; i < s->scenes.size(); ++i) set_box(s, i, bmin, bmax);
    }
    std::string key;
    if (cachePath) {
        key = tuning_key(s);
        akua_tuning c{};
        int cells = 1;
        if (cache_lookup(cachePath, key, c, cells)) {
            apply_tuning(s, c);
            if (!s->exact) apply_search(s, cells);
            return;
        }
    }
    const bool timing = s->timing;
    const bool graph = s->useGraph;
    const akua_tuning before = s->tuning;
    const int cellsBefore = s->sph.searchCells;
    Snapshot q;
    take_snapshot(s, q);
    s->timing = false;
    s->useGraph = true;
    // on any failure the solver is left as it was found: state, setting and launch mode
    struct Restore {
        akua_solver* s;
        const Snapshot& q;
        bool timing, graph;
        akua_tuning tuning;
        int cells;
        bool armed = true;
        ~Restore() {
            if (!armed) return;
            try {
                s->timing = timing;
                s->useGraph = graph;
                apply_search(s, cells);
                apply_tuning(s, tuning);
                restore_snapshot(s, q);
                (void)hipStreamSynchronize(s->stream);
            } catch (...) {
            }
        }
    } guard{s, q, timing, graph, before, cellsBefore};
    akua_tuning best = s->tuning;
    apply_tuning(s, best);
    double bestMs = time_trial(s, q, dt, trialSteps);
    const auto sweep = [&](int akua_tuning::*field, std::initializer_list<int> values) {
        for (int v : values) {
            if (best.*field == v) continue;
            akua_tuning t = best;
            t.*field = v;
            apply_tuning(s, t);
            const double ms = time_trial(s, q, dt, trialSteps);
            if (ms < bestMs) {
                bestMs = ms;
                best = t;
            }
        }
    };
    sweep(&akua_tuning::block, {128, 256, 512});
    sweep(&akua_tuning::layoutLevel, {0, 1, 2});
    sweep(&akua_tuning::hashDigitBits, {0, 5, 6, 8});
    sweep(&akua_tuning::layoutDigitBits, {0, 5, 8});
    sweep(&akua_tuning::layoutInterval, {1, 2, 4});
    sweep(&akua_tuning::balanceTile, {0, 512, 1024, 2048});
    sweep(&akua_tuning::slabBatch, {0, 4, 8});
    // the sweeps' launch forms interact (lane groups beside the work queue run only vorticity and
    // XSPH, without it every sweep and then no fused first sweep), so they are timed together
    {
        const akua_tuning base = best;
        for (int f : {0, 1})
            for (int wq : {0, 1, 2})
                for (int l : {0, 2, 4, 8, 16}) {
                    if (l > 1 && wq == 0 && f == 1) continue;  // the same launches as f == 0
                    if (f == base.fuseSweep && wq == base.workQueue && l == base.laneGroup) continue;
                    akua_tuning t = base;
                    t.fuseSweep = f;
                    t.workQueue = wq;
                    t.laneGroup = l;
                    apply_tuning(s, t);
                    const double ms = time_trial(s, q, dt, trialSteps);
                    if (ms < bestMs) {
                        bestMs = ms;
                        best = t;
                    }
                }
    }
    // the non-parity search over h / 2 cells, only when the caller gave up bit-exactness
    int bestCells = s->sph.searchCells;
    if (!s->exact) {
        apply_tuning(s, best);
        for (int cells : {1, 2}) {
            if (cells == bestCells) continue;
            apply_search(s, cells);
            const double ms = time_trial(s, q, dt, trialSteps);
            if (ms < bestMs) {
                bestMs = ms;
                bestCells = cells;
            }
        }
        apply_search(s, bestCells);
    }
    apply_tuning(s, best);
    restore_snapshot(s, q);
    HIP_CHECK(hipStreamSynchronize(s->stream));
    guard.armed = false;
    s->timing = timing;
    s->useGraph = graph;
    if (cachePath) cache_store(cachePath, key, best, bestCells);
}

}  // namespace

void akua::set_last_error(const std::string& message) { g_lastError = message; }

extern "C" {

int32_t akua_abi_version(void) { return AKUA_ABI_VERSION; }

const char* akua_last_error(void) { return g_lastError.c_str(); }

void akua_default_config(akua_pbf_config* c, akua_lambda_corr* k) {
    if (c) {
        c->restDensity = 7600.0f;
        c->particle_spacing = 0.05f;
        c->smoothRadius = 0.1f;
        c->spatialHashCellSize = 0.1f;
        c->relaxation = 600.0f;
        c->vorticityEpsilon = 0.00001f;
        c->viscosity = 0.01f;
        c->maxNeighbours = 128;
        c->solverIterations = 4;
        c->gravity[0] = 0.0f;
        c->gravity[1] = -9.8f;
        c->gravity[2] = 0.0f;
    }
    if (k) {
        k->enabled = 1;
        k->k = 0.0001f;
        k->n = 4.0f;
        k->delta_q = 0.03f;
    }
}

akua_status akua_solver_create(int32_t numParticles, const akua_pbf_config* cfg,
                               const akua_lambda_corr* corr, akua_solver** out) {
    return guarded([&] {
        if (!out) fail(AKUA_ERR_INVALID_ARG, "out must not be NULL");
        *out = nullptr;
        validate_config(cfg, corr);
        if (numParticles < 1) fail(AKUA_ERR_INVALID_ARG, "numParticles must be >= 1");
        auto* s = new akua_solver();
        s->cfg = *cfg;
        s->corr = *corr;
        try {
            setup(s, std::vector<uint32_t>{(uint32_t)numParticles});
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

akua_status akua_batch_create(int32_t numScenes, const int32_t* counts, const akua_pbf_config* cfg,
                              const akua_lambda_corr* corr, akua_solver** out) {
    return guarded([&] {
        if (!out || !counts) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        *out = nullptr;
        validate_config(cfg, corr);
        if (numScenes < 1) fail(AKUA_ERR_INVALID_ARG, "numScenes must be >= 1");
        std::vector<uint32_t> c((size_t)numScenes);
        uint64_t total = 0;
        for (int32_t i = 0; i < numScenes; ++i) {
            if (counts[i] < 1) fail(AKUA_ERR_INVALID_ARG, "every scene needs at least one particle");
            c[(size_t)i] = (uint32_t)counts[i];
            total += (uint64_t)counts[i];
        }
        if (total > 0x7FFFFFFFull) fail(AKUA_ERR_UNSUPPORTED, "too many particles");
        auto* s = new akua_solver();
        s->cfg = *cfg;
        s->corr = *corr;
        try {
            setup(s, c);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

akua_status akua_batch_num_scenes(const akua_solver* s, int32_t* n) {
    return guarded([&] {
        if (!s || !n) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        *n = (int32_t)s->scenes.size();
    });
}

akua_status akua_batch_set_box(akua_solver* s, int32_t scene, const float boxMin[3],
                               const float boxMax[3]) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        if (scene < 0 || (size_t)scene >= s->scenes.size()) fail(AKUA_ERR_INVALID_ARG, "scene out of range");
        check_box(boxMin, boxMax);
        set_box(s, (size_t)scene, boxMin, boxMax);
    });
}

akua_status akua_batch_step(akua_solver* s, float dt, int32_t nsteps) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        if (nsteps < 0) fail(AKUA_ERR_INVALID_ARG, "nsteps must be >= 0");
        run_steps(s, dt, nsteps);
    });
}

akua_status akua_batch_density_error(akua_solver* s, float* maxError, double* meanError) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        run_density_probe(s);
        auto mx = download(s, s->probeMax.p, s->scenes.size());
        auto sm = download(s, s->probeSum.p, s->scenes.size());
        for (size_t i = 0; i < mx.size(); ++i) {
            float v;
            std::memcpy(&v, &mx[i], sizeof(v));
            if (maxError) maxError[i] = v;
            if (meanError) meanError[i] = sm[i] / (double)(s->scenes[i].end - s->scenes[i].begin);
        }
    });
}

akua_status akua_solver_destroy(akua_solver* s) {
    return guarded([&] {
        if (!s) return;
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        delete s;
    });
}

akua_status akua_solver_set_particles(akua_solver* s, const akua_particle* particles, int32_t n) {
    return guarded([&] {
        if (!s || !particles) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_n(s, n);
        upload_soa(s, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, true, particles);
    });
}

akua_status akua_solver_set_state(akua_solver* s, const float* pos, const float* vel,
                                  const float* mass, const float* density, const float* lambda,
                                  int32_t n) {
    return guarded([&] {
        if (!s || !pos || !vel) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_n(s, n);
        upload_soa(s, pos, vel, mass, density, lambda, nullptr, nullptr, false, nullptr);
    });
}

akua_status akua_solver_get_state(akua_solver* s, float* pos, float* vel, float* mass,
                                  float* density, float* lambda, uint32_t* ids, int32_t n) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        check_n(s, n);
        auto X = download(s, s->X.p, s->nslots);
        auto V = download(s, s->V.p, s->nslots);
        auto D = download(s, s->dens.p, s->nslots);
        auto L = download(s, s->lam.p, s->nslots);
        auto I = download(s, s->id.p, s->nslots);
        auto LI = download(s, s->lid.p, s->nslots);
        size_t base = 0;
        for (const SceneDev& sc : s->scenes) {
            for (uint32_t slot = sc.begin; slot < sc.end; ++slot) {
                const size_t k = base + (LI[slot] - sc.begin);  // logical position
                if (pos) { pos[3 * k] = X[slot].x; pos[3 * k + 1] = X[slot].y; pos[3 * k + 2] = X[slot].z; }
                if (vel) { vel[3 * k] = V[slot].x; vel[3 * k + 1] = V[slot].y; vel[3 * k + 2] = V[slot].z; }
                if (mass) mass[k] = X[slot].w;
                if (density) density[k] = D[slot];
                if (lambda) lambda[k] = L[slot];
                if (ids) ids[k] = I[slot];
            }
            base += sc.end - sc.begin;
        }
    });
}

akua_status akua_solver_get_particles(akua_solver* s, akua_particle* out, int32_t n) {
    return guarded([&] {
        if (!s || !out) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_n(s, n);
        if (s->scenes.size() != 1) fail(AKUA_ERR_UNSUPPORTED, "record export is per single-scene solver");
        if (s->aosStage.n < (size_t)n) s->aosStage.alloc((size_t)n);
        DeviceState ds = s->devState();
        launch_export_aos(s->stream, s->aosStage.p, (uint32_t)n, s->X.p, s->V.p, s->VS.p, s->DP.p,
                          s->dens.p, s->lam.p, s->id.p, s->lid.p, s->W.p, ds.skeys, s->color.p,
                          s->size.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(s->stream));
        HIP_CHECK(hipMemcpy(out, s->aosStage.p, (size_t)n * sizeof(AosParticle), hipMemcpyDeviceToHost));
    });
}

akua_status akua_solver_step(akua_solver* s, float dt, const float boxMin[3], const float boxMax[3]) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        check_box(boxMin, boxMax);
        set_box(s, 0, boxMin, boxMax);
        run_steps(s, dt, 1);
    });
}

akua_status akua_solver_step_n(akua_solver* s, float dt, const float boxMin[3],
                               const float boxMax[3], int32_t nsteps) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        if (nsteps < 0) fail(AKUA_ERR_INVALID_ARG, "nsteps must be >= 0");
        check_box(boxMin, boxMax);
        for (size_t i = 0; i < s->scenes.size(); ++i) set_box(s, i, boxMin, boxMax);
        run_steps(s, dt, nsteps);
    });
}

akua_status akua_solver_step_device(akua_solver* s, void* dParticles, int32_t n, float dt,
                                    const float boxMin[3], const float boxMax[3], void* stream) {
    return guarded([&] {
        if (!s || !dParticles) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_n(s, n);
        if (s->scenes.size() != 1) fail(AKUA_ERR_UNSUPPORTED, "record stepping is per single-scene solver");
        check_box(boxMin, boxMax);
        set_box(s, 0, boxMin, boxMax);
        // NULL is the null stream: the solver's own stream does not synchronise with it, so the
        // order is made explicit for every caller stream
        hipStream_t user = static_cast<hipStream_t>(stream);
        if (!s->orderEvent) HIP_CHECK(hipEventCreateWithFlags(&s->orderEvent, hipEventDisableTiming));
        hipEvent_t ev = s->orderEvent;
        // the solver's stream waits for the caller's earlier work on the buffer
        HIP_CHECK(hipEventRecord(ev, user));
        HIP_CHECK(hipStreamWaitEvent(s->stream, ev, 0));
        auto* buf = static_cast<AosParticle*>(dParticles);
        launch_import_aos(s->stream, buf, (uint32_t)n, s->X.p, s->V.p, s->VS.p, s->DP.p, s->dens.p,
                          s->lam.p, s->id.p, s->lid.p, s->phys.p, s->color.p, s->size.p);
        s->hasState = true;
        s->layoutPhase = 0;
        s->unitMass = false;  // the masses came from device memory: not known here
        update_fast(s);
        run_steps(s, dt, 1);
        DeviceState ds = s->devState();
        launch_export_aos(s->stream, buf, (uint32_t)n, s->X.p, s->V.p, s->VS.p, s->DP.p, s->dens.p,
                          s->lam.p, s->id.p, s->lid.p, s->W.p, ds.skeys, s->color.p, s->size.p);
        // the caller's later work on its stream waits for the written-back records
        HIP_CHECK(hipEventRecord(ev, s->stream));
        HIP_CHECK(hipStreamWaitEvent(user, ev, 0));
        HIP_CHECK(hipGetLastError());
    });
}

akua_status akua_solver_set_gravity(akua_solver* s, const float g[3]) {
    return guarded([&] {
        if (!s || !g) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(g[k])) fail(AKUA_ERR_INVALID_ARG, "non-finite gravity");
        for (int k = 0; k < 3; ++k) s->cfg.gravity[k] = g[k];
    });
}

akua_status akua_solver_synchronize(akua_solver* s) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

akua_status akua_solver_num_particles(const akua_solver* s, int32_t* n) {
    return guarded([&] {
        if (!s || !n) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        *n = (int32_t)s->nparticles;
    });
}

akua_status akua_solver_set_graph(akua_solver* s, int32_t enable) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        s->useGraph = enable != 0;
    });
}

akua_status akua_solver_set_tuning(akua_solver* s, const akua_tuning* t) {
    return guarded([&] {
        if (!s || !t) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        apply_tuning(s, *t);
    });
}

akua_status akua_solver_set_search(akua_solver* s, int32_t cellsPerH) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        if (cellsPerH != 1 && cellsPerH != 2) fail(AKUA_ERR_INVALID_ARG, "cellsPerH must be 1 or 2");
        apply_search(s, cellsPerH);
    });
}

akua_status akua_solver_get_search(const akua_solver* s, int32_t* cellsPerH) {
    return guarded([&] {
        if (!s || !cellsPerH) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        *cellsPerH = s->sph.searchCells;
    });
}

akua_status akua_solver_set_exact(akua_solver* s, int32_t exact) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        s->exact = exact != 0;
    });
}

akua_status akua_solver_get_tuning(const akua_solver* s, akua_tuning* t) {
    return guarded([&] {
        if (!s || !t) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        *t = s->tuning;
    });
}

akua_status akua_solver_autotune(akua_solver* s, float dt, const float boxMin[3],
                                 const float boxMax[3], int32_t trialSteps, const char* cachePath,
                                 akua_tuning* best) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        autotune(s, dt, boxMin, boxMax, trialSteps, cachePath);
        if (best) *best = s->tuning;
    });
}

akua_status akua_solver_set_network(akua_solver* s, const float* params, int32_t count) {
    return guarded([&] {
        if (!s || !params) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        if (count != kMlpParams) fail(AKUA_ERR_INVALID_ARG, "count must be AKUA_MLP_PARAMS");
        upload_network(s, params);
    });
}

akua_status akua_solver_get_network(akua_solver* s, float* params, int32_t count) {
    return guarded([&] {
        if (!s || !params) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        if (count != kMlpParams) fail(AKUA_ERR_INVALID_ARG, "count must be AKUA_MLP_PARAMS");
        if (!s->haveNet) fail(AKUA_ERR_STATE, "the solver has no network");
        auto h = download(s, s->net.p, kMlpParams);
        std::memcpy(params, h.data(), kMlpParams * sizeof(float));
    });
}

akua_status akua_solver_set_learned(akua_solver* s, int32_t enable, int32_t iterations,
                                    int32_t statistic, float ratio, float tolerance) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        if (enable) {
            if (!s->haveNet) fail(AKUA_ERR_STATE, "set or train a network first");
            if (s->bndLayers > 0)
                fail(AKUA_ERR_UNSUPPORTED, "the learned step does not support boundary samples");
            if (s->cfg.solverIterations < 1)
                fail(AKUA_ERR_UNSUPPORTED, "the learned step needs solverIterations >= 1");
            if (!std::isfinite(ratio) || !std::isfinite(tolerance) || ratio < 0.0f)
                fail(AKUA_ERR_INVALID_ARG, "ratio must be >= 0 and tolerance finite");
            if (statistic != 0 && statistic != 1)
                fail(AKUA_ERR_INVALID_ARG, "statistic must be 0 (max) or 1 (mean)");
            if (iterations < 1 || iterations > s->cfg.solverIterations)
                fail(AKUA_ERR_INVALID_ARG, "iterations must be in [1, solverIterations]");
            s->learnedIters = iterations;
        }
        s->mode = enable ? 1 : 0;
        s->gateStat = statistic;
        s->gateRatio = ratio;
        s->gateTol = tolerance;
        s->graphDirty = true;
    });
}

akua_status akua_solver_set_boundary(akua_solver* s, int32_t layers, float spacing) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        if (layers < 0 || layers > kMaxBoundaryLayers) fail(AKUA_ERR_INVALID_ARG, "layers must be 0..4");
        if (!std::isfinite(spacing) || spacing < 0.0f) fail(AKUA_ERR_INVALID_ARG, "spacing must be >= 0");
        if (layers > 0 && s->scenes.size() != 1)
            fail(AKUA_ERR_UNSUPPORTED, "boundary samples are per single-scene solver");
        if (layers > 0 && s->mode == 1)
            fail(AKUA_ERR_UNSUPPORTED, "the learned step does not support boundary samples");
        s->bndLayers = layers;
        s->bndSpacing = spacing;
        s->bndDirty = true;
        if (layers == 0) s->sph.bd = BoundaryDev{};
        s->graphDirty = true;
    });
}

akua_status akua_debug_boundary(akua_solver* s, float* pos, float* psi, int32_t capacity, int32_t* count) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        if (capacity < 0) fail(AKUA_ERR_INVALID_ARG, "capacity must be >= 0");
        if (s->bndLayers <= 0) fail(AKUA_ERR_STATE, "boundary samples are off");
        if (!s->boxSeen) fail(AKUA_ERR_STATE, "no box yet: the samples follow the box of a step");
        ensure_boundary(s);
        const BoundaryDev& B = s->sph.bd;
        const uint32_t m = B.base[6];
        if (count) *count = (int32_t)m;
        const uint32_t k = std::min<uint32_t>(m, (uint32_t)capacity);
        if (pos)
            for (uint32_t i = 0; i < k; ++i) boundary_position(B, i, pos + 3 * (size_t)i);
        if (psi && k) {
            HIP_CHECK(hipStreamSynchronize(s->stream));
            HIP_CHECK(hipMemcpy(psi, s->bndPsi.p, (size_t)k * sizeof(float), hipMemcpyDeviceToHost));
        }
    });
}

akua_status akua_boundary_samples(const akua_pbf_config* cfg, const float boxMin[3],
                                  const float boxMax[3], int32_t layers, float spacing, float* pos,
                                  float* psi, int32_t capacity, int32_t* count) {
    return guarded([&] {
        if (!cfg || !boxMin || !boxMax) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        if (layers < 1 || layers > kMaxBoundaryLayers) fail(AKUA_ERR_INVALID_ARG, "layers must be 1..4");
        if (capacity < 0 || !std::isfinite(spacing) || spacing < 0.0f)
            fail(AKUA_ERR_INVALID_ARG, "capacity and spacing must be >= 0");
        akua_lambda_corr corr{1, 0.0001f, 4.0f, 0.03f};
        const SphParams P = make_sph(*cfg, corr);
        BoundaryDev B{};
        std::vector<float> w;
        std::string err;
        const float sp = spacing > 0.0f ? spacing : cfg->particle_spacing;
        if (!build_boundary(boxMin, boxMax, sp, layers, P, cfg->restDensity, B, w, err))
            fail(AKUA_ERR_INVALID_ARG, err);
        const uint32_t m = B.base[6];
        if (count) *count = (int32_t)m;
        const uint32_t k = std::min<uint32_t>(m, (uint32_t)capacity);
        for (uint32_t i = 0; i < k; ++i) {
            if (pos) boundary_position(B, i, pos + 3 * (size_t)i);
            if (psi) psi[i] = w[i];
        }
    });
}

void akua_default_camera(akua_camera* c) {
    if (!c) return;
    c->position[0] = 3.0f;
    c->position[1] = 2.0f;
    c->position[2] = 6.0f;
    c->yaw = -90.0f;
    c->pitch = 0.0f;
    c->fovY = 45.0f;
    c->nearZ = 0.1f;
    c->farZ = 160.0f;
}

namespace {
void check_camera(const akua_camera* cam, int32_t width, int32_t height) {
    if (!cam) fail(AKUA_ERR_INVALID_ARG, "NULL camera");
    if (width < 1 || height < 1 || (int64_t)width * height > (1 << 26))
        fail(AKUA_ERR_INVALID_ARG, "width and height must be positive, at most 2^26 pixels");
    const float v[8] = {cam->position[0], cam->position[1], cam->position[2], cam->yaw, cam->pitch,
                        cam->fovY, cam->nearZ, cam->farZ};
    for (float x : v)
        if (!std::isfinite(x)) fail(AKUA_ERR_INVALID_ARG, "non-finite camera parameter");
    if (!(cam->fovY > 0.0f && cam->fovY < 180.0f) || !(cam->nearZ > 0.0f) || !(cam->farZ > cam->nearZ) ||
        !(std::fabs(cam->pitch) < 90.0f))
        fail(AKUA_ERR_INVALID_ARG, "camera needs 0 < fovY < 180, 0 < nearZ < farZ and |pitch| < 90");
}
}  // namespace

akua_status akua_camera_matrix(const akua_camera* cam, int32_t width, int32_t height, float vp[16]) {
    return guarded([&] {
        check_camera(cam, width, height);
        if (!vp) fail(AKUA_ERR_INVALID_ARG, "NULL matrix");
        RenderParams R{};
        camera_setup(cam->position, cam->yaw, cam->pitch, cam->fovY, cam->nearZ, cam->farZ, width, height, R);
        for (int k = 0; k < 16; ++k) vp[k] = R.vp[k];
    });
}

akua_status akua_render(akua_solver* s, const akua_camera* cam, int32_t width, int32_t height,
                        float pointSize, int32_t drawFloor, uint8_t* rgb) {
    return guarded([&] {
        if (!s || !rgb) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_camera(cam, width, height);
        if (!std::isfinite(pointSize)) fail(AKUA_ERR_INVALID_ARG, "non-finite point size");
        if (s->scenes.size() != 1) fail(AKUA_ERR_UNSUPPORTED, "rendering is per single-scene solver");
        if (!s->hasState) fail(AKUA_ERR_STATE, "set the particles first");
        RenderParams R{};
        camera_setup(cam->position, cam->yaw, cam->pitch, cam->fovY, cam->nearZ, cam->farZ, width, height, R);
        R.pointSize = pointSize;
        R.drawFloor = drawFloor ? 1 : 0;
        const size_t pixels = (size_t)width * (size_t)height;
        HIP_CHECK(hipStreamSynchronize(s->stream));
        if (s->renderKeys.n < pixels) s->renderKeys.alloc(pixels);
        if (s->renderRgb.n < 3 * pixels) s->renderRgb.alloc(3 * pixels);
        upload_scenes(s);  // the order inversion reads the scene ranges
        launch_invert_order(s->stream, s->devState(), s->phys.p);
        launch_render(s->stream, s->X.p, s->color.p, s->size.p, s->id.p, s->lid.p, s->phys.p,
                      s->scenes[0].end, R, s->renderKeys.p, s->renderRgb.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(s->stream));
        HIP_CHECK(hipMemcpy(rgb, s->renderRgb.p, 3 * pixels, hipMemcpyDeviceToHost));
    });
}

akua_status akua_solver_learned_stats(akua_solver* s, int64_t* learnedSteps, int64_t* fallbacks,
                                      float* residual) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        int64_t fb = 0;
        float r[kResidStride] = {};
        if (s->fallbacks.p) {
            auto f = download(s, s->fallbacks.p, s->scenes.size());
            for (auto v : f) fb += v;
            auto q = download(s, s->resid.p, kResidStride);
            std::memcpy(r, q.data(), sizeof(r));
        }
        if (learnedSteps) *learnedSteps = s->learnedSteps;
        if (fallbacks) *fallbacks = fb;
        if (residual) {
            const SceneDev& sc = s->scenes[0];
            const float cnt = (float)std::max<uint32_t>(1u, sc.end - sc.begin);
            for (int k = 0; k < 3; ++k) {
                residual[k] = r[k];
                residual[3 + k] = r[kResidSum + k] / cnt;
            }
        }
    });
}

akua_status akua_solver_train(akua_solver* s, float dt, const float boxMin[3], const float boxMax[3],
                              int32_t steps, const akua_train_config* config, float* loss,
                              float* residuals) {
    return guarded([&] {
        if (!s || !config || !boxMin || !boxMax) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        if (s->bndLayers > 0) fail(AKUA_ERR_UNSUPPORTED, "training does not support boundary samples");
        train(s, dt, boxMin, boxMax, steps, *config, loss, residuals);
    });
}

akua_status akua_debug_features(akua_solver* s, float* features, float* targets, int32_t n) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        check_n(s, n);
        if (!s->feat.p) fail(AKUA_ERR_STATE, "no learned or recorded step has run");
        if (targets && !s->targ.p) fail(AKUA_ERR_STATE, "no recorded step has run");
        auto lid = download(s, s->lid.p, s->nslots);
        auto F = download(s, s->feat.p, (size_t)s->nslots * kMlpFeatures);
        std::vector<float> T;
        if (targets) T = download(s, s->targ.p, (size_t)s->nslots * 4);
        for (const SceneDev& sc : s->scenes)
            for (uint32_t p = sc.begin; p < sc.end; ++p) {
                const uint32_t k = lid[p];
                if (features) std::memcpy(features + (size_t)k * kMlpFeatures, &F[(size_t)p * kMlpFeatures], kMlpFeatures * 4);
                if (targets) std::memcpy(targets + (size_t)k * 4, &T[(size_t)p * 4], 16);
            }
    });
}

akua_status akua_debug_warm_start(akua_solver* s, float* xPred, float* xWarm, float* lambdaHat,
                                  int32_t n) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        check_n(s, n);
        if (!s->XsW.p) fail(AKUA_ERR_STATE, "no learned step has run");
        auto lid = download(s, s->lid.p, s->nslots);
        auto A = download(s, s->XsA.p, s->nslots);
        auto Wm = download(s, s->XsW.p, s->nslots);
        auto L = download(s, s->lamHat.p, s->nslots);
        for (const SceneDev& sc : s->scenes)
            for (uint32_t p = sc.begin; p < sc.end; ++p) {
                const uint32_t k = lid[p];
                if (xPred) { xPred[3 * k] = A[p].x; xPred[3 * k + 1] = A[p].y; xPred[3 * k + 2] = A[p].z; }
                if (xWarm) { xWarm[3 * k] = Wm[p].x; xWarm[3 * k + 1] = Wm[p].y; xWarm[3 * k + 2] = Wm[p].z; }
                if (lambdaHat) lambdaHat[k] = L[p];
            }
    });
}

akua_status akua_mlp_forward(const float* params, const float* features, int32_t n, float* outputs) {
    return guarded([&] {
        if (!params || !features || !outputs || n < 1) fail(AKUA_ERR_INVALID_ARG, "NULL argument or n < 1");
        need_device();
        DevBuf<float> P, F, Y;
        to_device(P, params, kMlpParams);
        to_device(F, features, (size_t)n * kMlpFeatures);
        Y.alloc((size_t)n * 4);
        launch_mlp_eval(nullptr, P.p, F.p, (uint32_t)n, Y.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(outputs, Y.p, (size_t)n * 16, hipMemcpyDeviceToHost));
    });
}

akua_status akua_mlp_gradient(const float* params, const float* features, const float* targets,
                              int32_t n, const float lossWeights[4], float* grad, float* loss) {
    return guarded([&] {
        if (!params || !features || !targets || !lossWeights || !grad || !loss || n < 1)
            fail(AKUA_ERR_INVALID_ARG, "NULL argument or n < 1");
        need_device();
        DevBuf<float> P, F, T, part, G, L;
        to_device(P, params, kMlpParams);
        to_device(F, features, (size_t)n * kMlpFeatures);
        to_device(T, targets, (size_t)n * 4);
        part.alloc((size_t)kMlpPartRows * kMlpPartRow);
        G.alloc(kMlpTrainable);
        L.alloc(1);
        MlpLossWeights lw{{lossWeights[0], lossWeights[1], lossWeights[2], lossWeights[3]}};
        launch_mlp_grad(nullptr, P.p, F.p, T.p, (uint32_t)n, lw, part.p);
        launch_mlp_reduce(nullptr, part.p, G.p, L.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(grad, G.p, kMlpTrainable * sizeof(float), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(loss, L.p, sizeof(float), hipMemcpyDeviceToHost));
    });
}

akua_status akua_mlp_adam_step(float* params, float* m, float* v, int64_t t, float lr,
                               const float* features, const float* targets, int32_t n,
                               const float lossWeights[4], float* loss) {
    return guarded([&] {
        if (!params || !m || !v || !features || !targets || !lossWeights || !loss || n < 1 || t < 1)
            fail(AKUA_ERR_INVALID_ARG, "NULL argument, n < 1 or t < 1");
        need_device();
        DevBuf<float> P, M, V, F, T, part, L;
        to_device(P, params, kMlpParams);
        to_device(M, m, kMlpTrainable);
        to_device(V, v, kMlpTrainable);
        to_device(F, features, (size_t)n * kMlpFeatures);
        to_device(T, targets, (size_t)n * 4);
        part.alloc((size_t)kMlpPartRows * kMlpPartRow);
        L.alloc(1);
        MlpLossWeights lw{{lossWeights[0], lossWeights[1], lossWeights[2], lossWeights[3]}};
        const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)t)));
        launch_mlp_grad(nullptr, P.p, F.p, T.p, (uint32_t)n, lw, part.p);
        launch_mlp_adam(nullptr, part.p, P.p, M.p, V.p, lr, 0.9f, 0.999f, 1e-8f, c1, c2, L.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(params, P.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(m, M.p, kMlpTrainable * sizeof(float), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(v, V.p, kMlpTrainable * sizeof(float), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(loss, L.p, sizeof(float), hipMemcpyDeviceToHost));
    });
}

akua_status akua_debug_hashes(akua_solver* s, uint32_t* out, int32_t n) {
    return guarded([&] {
        if (!s || !out) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_n(s, n);
        auto h = download(s, s->devState().skeys, s->nslots);
        size_t k = 0;
        for (const SceneDev& sc : s->scenes)
            for (uint32_t slot = sc.begin; slot < sc.end; ++slot) out[k++] = h[slot];
    });
}

akua_status akua_debug_permutation(akua_solver* s, uint32_t* out, int32_t n) {
    return guarded([&] {
        if (!s || !out) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_n(s, n);
        auto h = download(s, s->devState().perm1, s->nslots);
        size_t k = 0;
        for (const SceneDev& sc : s->scenes)
            for (uint32_t slot = sc.begin; slot < sc.end; ++slot) out[k++] = h[slot] - sc.begin;
    });
}

akua_status akua_debug_physical_order(akua_solver* s, uint32_t* out, int32_t n) {
    return guarded([&] {
        if (!s || !out) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_n(s, n);
        auto L = download(s, s->lid.p, s->nslots);
        size_t k = 0;
        for (const SceneDev& sc : s->scenes)
            for (uint32_t slot = sc.begin; slot < sc.end; ++slot) out[k++] = L[slot] - sc.begin;
    });
}

akua_status akua_debug_cell_ranges(akua_solver* s, uint32_t* hash, uint32_t* start, uint32_t* end,
                                   int32_t cap, int32_t* ncells) {
    return guarded([&] {
        if (!s || !ncells) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        if (s->scenes.size() != 1) fail(AKUA_ERR_UNSUPPORTED, "cell ranges are per single-scene solver");
        auto keys = download(s, s->devState().skeys, s->nslots);
        auto tr = download(s, s->table.p, s->tableTotal);
        const SceneDev& sc = s->scenes[0];
        std::vector<uint32_t> H, S, E;
        for (uint32_t k = sc.begin; k < sc.end; ++k) {
            if (k != sc.begin && keys[k] == keys[k - 1]) continue;
            uint32_t h = keys[k];
            // probe exactly as the device lookup does
            uint32_t mask = (1u << sc.tabLog2) - 1u;
            uint32_t slot = (h * 2654435761u) >> (32u - sc.tabLog2);
            uint32_t st = 0xFFFFFFFFu, en = 0xFFFFFFFFu;
            for (uint32_t probe = 0; probe <= mask; ++probe) {
                uint32_t kk = tr[sc.tabOffset + slot].x;
                if (kk == h) { st = tr[sc.tabOffset + slot].y; en = tr[sc.tabOffset + slot].z; break; }
                if (kk == kEmptyKey) break;
                slot = (slot + 1u) & mask;
            }
            H.push_back(h); S.push_back(st); E.push_back(en);
        }
        *ncells = (int32_t)H.size();
        if ((int32_t)H.size() > cap) fail(AKUA_ERR_INVALID_ARG, "cap is smaller than the number of cells");
        for (size_t c = 0; c < H.size(); ++c) {
            if (hash) hash[c] = H[c];
            if (start) start[c] = S[c] - sc.begin;
            if (end) end[c] = E[c] - sc.begin;
        }
    });
}

akua_status akua_debug_neighbours(akua_solver* s, uint32_t* nbrOut, uint32_t* cntOut, int32_t n) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        check_n(s, n);
        const int maxN = s->cfg.maxNeighbours;
        auto L = download(s, s->nbr.p, s->nbr.n);
        auto C = download(s, s->cnt.p, s->nslots);
        auto LI = download(s, s->lid.p, s->nslots);
        size_t base0 = 0;
        for (const SceneDev& sc : s->scenes) {
            for (uint32_t i = sc.begin; i < sc.end; ++i) {
                const size_t k = base0 + (LI[i] - sc.begin);  // the row of particle i's logical index
                if (cntOut) cntOut[k] = C[i];
                if (!nbrOut) continue;
                // chunk c of slot i: the uint4 at ((i / 64) * chunks + c) * 64 + i % 64
                const size_t chunks = list_chunk_count(maxN);
                const size_t base = (size_t)(i >> 6) * chunks * 64 + (i & 63);
                for (int j = 0; j < maxN; ++j)
                    nbrOut[k * maxN + j] =
                        (uint32_t)j < C[i] ? LI[L[(base + (size_t)(j >> 2) * 64) * 4 + (j & 3)]] - sc.begin
                                           : 0u;
            }
            base0 += sc.end - sc.begin;
        }
    });
}

akua_status akua_solver_set_timing(akua_solver* s, int32_t enable) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        s->timing = enable != 0;
    });
}

akua_status akua_solver_reset_timing(akua_solver* s) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        collect_timing(s);
        s->timeAcc.clear();
        s->timeOrder.clear();
    });
}

akua_status akua_solver_get_timing(akua_solver* s, char* nameBuf, int32_t nameLen, double* ms,
                                   int64_t* launches, int32_t cap, int32_t* count) {
    return guarded([&] {
        if (!s || !count) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        collect_timing(s);
        *count = (int32_t)s->timeOrder.size();
        for (int32_t i = 0; i < cap && i < *count; ++i) {
            const std::string& nm = s->timeOrder[i];
            const auto& acc = s->timeAcc[nm];
            if (nameBuf && nameLen > 0) {
                std::strncpy(nameBuf + (size_t)i * nameLen, nm.c_str(), nameLen - 1);
                nameBuf[(size_t)i * nameLen + nameLen - 1] = 0;
            }
            if (ms) ms[i] = acc.first;
            if (launches) launches[i] = acc.second;
        }
    });
}

akua_status akua_solver_density_error(akua_solver* s, float* maxError, double* meanError) {
    return guarded([&] {
        if (!s) fail(AKUA_ERR_INVALID_ARG, "NULL solver");
        run_density_probe(s);
        auto mx = download(s, s->probeMax.p, s->scenes.size());
        auto sm = download(s, s->probeSum.p, s->scenes.size());
        float m = 0.0f;
        double tot = 0.0;
        for (size_t i = 0; i < mx.size(); ++i) {
            float v;
            std::memcpy(&v, &mx[i], sizeof(v));
            m = std::max(m, v);
            tot += sm[i];
        }
        if (maxError) *maxError = m;
        if (meanError) *meanError = tot / (double)s->nparticles;
    });
}

akua_status akua_debug_probe_density(akua_solver* s, float* rho, int32_t n) {
    return guarded([&] {
        if (!s || !rho) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        check_n(s, n);
        auto r = download(s, s->probeRho.p, s->nslots);  // indexed by logical index
        size_t k = 0;
        for (const SceneDev& sc : s->scenes)
            for (uint32_t slot = sc.begin; slot < sc.end; ++slot) rho[k++] = r[slot];
    });
}

akua_status akua_solver_mean_neighbours(akua_solver* s, double* mean) {
    return guarded([&] {
        if (!s || !mean) fail(AKUA_ERR_INVALID_ARG, "NULL argument");
        if (s->stepsSinceStats == 0) {
            *mean = 0.0;
            return;
        }
        launch_count_sum(s->stream, s->devState(), s->nbrSum.p, kCountBlocks);
        HIP_CHECK(hipGetLastError());
        auto v = download(s, s->nbrSum.p, kCountBlocks);
        unsigned long long t = 0;
        for (auto x : v) t += x;
        *mean = (double)t / (double)s->nparticles;
    });
}

}  // extern "C"
