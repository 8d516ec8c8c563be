/* akua.h -- C ABI of the MI355X-native Position-Based Fluids solver (libakua.so).
 *
 * The ABI replaces the reference's Simulation-layer API and its CUDA op wrappers:
 *   AkuaEngine::PBFSolver (PBFSolver.h:12-31, PBFSolver.cpp:13-78), whose step() maps an
 *   OpenGL-owned particle buffer through the CUDA-GL interop (InteropResource.cpp:9-35) and calls
 *   IntegrationCUDA / NeighbourSearchCUDA / ConstraintSolverCUDA (IntegrationCUDA.h:12-32,
 *   NeighbourSearchCUDA.h:12-21, ConstraintSolverCUDA.h:14-27).
 * Here the solver owns its device memory, runs headless, and every call returns an akua_status.
 * No C++ exception crosses the ABI.  Signatures use only plain C types and pointers.
 *
 * Threading: one solver runs on one HIP stream.  Calls on one solver must be serialised by the
 * caller; different solvers may be used from different threads.  akua_solver_step() is
 * asynchronous with respect to the host; every get/debug call synchronises the solver's stream.
 */
#ifndef AKUA_H
#define AKUA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AKUA_ABI_VERSION 1

typedef enum akua_status {
    AKUA_OK = 0,
    AKUA_ERR_INVALID_ARG = 1,   /* null pointer, size mismatch, out-of-range parameter */
    AKUA_ERR_OUT_OF_MEMORY = 2, /* device allocation failed */
    AKUA_ERR_HIP = 3,           /* a HIP runtime call failed; akua_last_error() names it */
    AKUA_ERR_NON_FINITE = 4,    /* NaN or Inf in caller-provided particle data */
    AKUA_ERR_STATE = 5,         /* call out of order (e.g. step before set_particles) */
    AKUA_ERR_UNSUPPORTED = 6,   /* configuration this build does not handle */
    AKUA_ERR_NO_DEVICE = 7      /* no HIP device is visible */
} akua_status;

/* Mirrors AkuaEngine::PBFConfig field for field (PBFConfig.h:18-29), glm::vec3 -> float[3].
 * Defaults (akua_default_config): restDensity 7600, particle_spacing 0.05, smoothRadius 0.1,
 * spatialHashCellSize 0.1, relaxation 600, vorticityEpsilon 1e-5, viscosity 0.01,
 * maxNeighbours 128, solverIterations 4, gravity (0, -9.8, 0). */
typedef struct akua_pbf_config {
    float restDensity;
    float particle_spacing;
    float smoothRadius;          /* h; also the hashing cell size (NeighbourSearchCUDA.cu:163) */
    float spatialHashCellSize;   /* cell size of the neighbour search (NeighbourSearchCUDA.cu:178) */
    float relaxation;            /* epsilon in the lambda denominator */
    float vorticityEpsilon;
    float viscosity;             /* XSPH coefficient c */
    int32_t maxNeighbours;       /* neighbour-list cap; the hash table size is maxNeighbours * N */
    int32_t solverIterations;    /* Jacobi iterations per step */
    float gravity[3];
} akua_pbf_config;

/* Mirrors AkuaEngine::LambdaCorrParams (PBFConfig.h:10-15).  enabled == 0 drops the
 * artificial-pressure term s_corr from the position delta.  The reference never reads the flag
 * (ConstraintSolverCUDA.cu:99-130 always applies the term): results are the reference's whenever
 * it is set.  Defaults: enabled 1, k 1e-4, n 4, delta_q 0.03. */
typedef struct akua_lambda_corr {
    int32_t enabled;
    float k;
    float n;
    float delta_q;
} akua_lambda_corr;

/* The reference's 108-byte interleaved particle record (Particle.h:8-31), field order and offsets
 * kept so a renderer's vertex buffer (position at offset 0, stride 108) can be filled directly. */
typedef struct akua_particle {
    float position[3];
    float velocity[3];
    float new_position[3];
    float new_velocity[3];
    float position_delta[3];
    float vorticity[3];
    float mass;
    float density;
    float lambda;
    uint32_t hash;
    float color[4];
    float size;
} akua_particle;

typedef struct akua_solver akua_solver;

/* ---- library ------------------------------------------------------------------------------ */
int32_t akua_abi_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char* akua_last_error(void);
void akua_default_config(akua_pbf_config* cfg, akua_lambda_corr* corr);

/* ---- single-scene solver (PBFSolver) ------------------------------------------------------ */
/* PBFSolver::PBFSolver(int, const PBFConfig&, const LambdaCorrParams&) (PBFSolver.h:14,
 * PBFSolver.cpp:13-20).  numParticles >= 1; maxNeighbours * numParticles must fit in uint32. */
akua_status akua_solver_create(int32_t numParticles, const akua_pbf_config* cfg,
                               const akua_lambda_corr* corr, akua_solver** out);
akua_status akua_solver_destroy(akua_solver* s);

/* Upload particles as the reference's AoS records (Renderer::uploadParticles,
 * Renderer.cpp:185-219).  Reads position, velocity, new_velocity, position_delta, mass, density,
 * lambda, color and size (new_velocity and position_delta are returned unchanged until a step
 * rewrites them; a step without iterations keeps position_delta, as the reference's does). */
akua_status akua_solver_set_particles(akua_solver* s, const akua_particle* particles, int32_t n);
/* Download in the solver's current (sorted) order, as the reference's buffer is physically
 * reordered by each step (NeighbourSearchCUDA.cu:166-170).  Fills every field from the last
 * step: position, velocity, new_position (= position), new_velocity ((x* - x) / dt, the velocity
 * before the wall damping, vorticity and XSPH, IntegrationCUDA.cu:45), position_delta (the last
 * iteration's delta, ConstraintSolverCUDA.cu:129), vorticity, mass, density, lambda, hash. */
akua_status akua_solver_get_particles(akua_solver* s, akua_particle* particles, int32_t n);

/* Structure-of-arrays upload: pos/vel are 3*n floats (xyz interleaved per particle); mass,
 * density and lambda n floats each (NULL: mass 1, density 0, lambda 0).  Density and lambda only
 * matter to a step with solverIterations == 0, which reuses them as the reference reuses the
 * record's fields.  Particle ids are assigned 0..n-1 in the given order. */
akua_status akua_solver_set_state(akua_solver* s, const float* pos, const float* vel,
                                  const float* mass, const float* density, const float* lambda,
                                  int32_t n);
/* Structure-of-arrays download in the current order; any pointer may be NULL.  ids[k] is the
 * original index of the particle now in slot k. */
akua_status akua_solver_get_state(akua_solver* s, float* pos, float* vel, float* mass,
                                  float* density, float* lambda, uint32_t* ids, int32_t n);

/* PBFSolver::step (PBFSolver.h:17, PBFSolver.cpp:22-78) on the solver-owned buffers. */
akua_status akua_solver_step(akua_solver* s, float deltaTime, const float boxMin[3],
                             const float boxMax[3]);
/* Run nsteps steps with the same arguments (one graph launch per step). */
akua_status akua_solver_step_n(akua_solver* s, float deltaTime, const float boxMin[3],
                               const float boxMax[3], int32_t nsteps);
/* The GL-interop replacement: the caller owns a device buffer of n akua_particle records (the
 * role of the mapped VBO, InteropResource.cpp:9-35).  The call imports it, runs one step and
 * writes the records back in the new sorted order, ordered on the caller's `stream` (NULL: the
 * null stream): the import waits for the work enqueued on `stream` before the call, and work
 * enqueued on `stream` after the call waits for the written-back records.  The host does not
 * wait.  The buffer must stay valid until that work has completed. */
akua_status akua_solver_step_device(akua_solver* s, void* d_particles, int32_t n, float deltaTime,
                                    const float boxMin[3], const float boxMax[3], void* stream);
/* PBFSolver::setGravity (PBFSolver.h:18, :29-31). */
akua_status akua_solver_set_gravity(akua_solver* s, const float gravity[3]);
akua_status akua_solver_synchronize(akua_solver* s);
akua_status akua_solver_num_particles(const akua_solver* s, int32_t* n);

/* Use a hipGraph per step (default 1).  0 launches the kernels directly each step. */
akua_status akua_solver_set_graph(akua_solver* s, int32_t enable);

/* ---- tuning ---------------------------------------------------------------------------------- */
/* Launch and layout choices that never change a result: every setting gives bit-identical
 * particle states.  No reference counterpart (the reference fixes 256-thread blocks and a
 * thrust sort, NeighbourSearchCUDA.cu:150-172). */
typedef struct akua_tuning {
    int32_t block;            /* workgroup size of the neighbour-walking kernels: 128, 256 or 512 */
    int32_t layoutLevel;      /* physical layout: super-cells of 2^level hash cells per axis, 0..3 */
    int32_t hashDigitBits;    /* radix digit width of the hash sort: 0 (balanced) or 5..8 */
    int32_t layoutDigitBits;  /* radix digit width of the layout sort: 0 (balanced) or 5..8 */
    int32_t layoutInterval;   /* the physical layout is re-sorted every layoutInterval steps (1..8);
                                 in between, particles keep their memory slots */
    int32_t balanceTile;      /* 0 (the default): the neighbour sweeps run one thread per physical
                                 slot; 256..2048 (a multiple of 256): in a per-step work order, the
                                 slots of each tile of this many sorted by neighbour count, so that
                                 the lanes of a wave walk lists of similar length (their gathers
                                 then spread over the tile) */
    int32_t slabBatch;        /* 0: the list build walks each of the 27 stencil buckets in batches
                                 of 8 candidates; 4 or 8: it walks each dx slab of nine buckets as
                                 one stream of candidates in batches of this many */
    int32_t fuseSweep;        /* 0 (the default): the list build is a launch of its own; 1: the
                                 same launch then runs, in each thread over the list it just
                                 wrote, the first density/lambda sweep (or, in a learned or
                                 training step, the feature sweep); applies with balanceTile 0,
                                 slabBatch 0, the reference's search and no boundary samples */
    int32_t laneGroup;        /* 0 (the default): one thread per particle; 2, 4, 8 or 16: for scenes
                                 too small to fill the chip, the neighbour sweeps of a plain step
                                 run that many lanes per particle (each list's terms in parallel,
                                 summed in list order) and the list build one lane per stencil cell;
                                 applies with balanceTile 0, slabBatch 0, the reference's search and
                                 no boundary samples, and takes precedence over fuseSweep.  With
                                 workQueue set too, only the vorticity and XSPH sweeps run in lane
                                 groups; the queue runs the others */
    int32_t workQueue;        /* 0 (the default): each neighbour sweep after the list build is one
                                 launch with a workgroup per block of slots; 1 or 2: after the list
                                 build, the 256-slot tiles each XCD serves are queued heaviest first
                                 (by their neighbour counts) and the sweeps of a plain step run as
                                 resident grids whose workgroups take tiles from their XCD's queue
                                 until it is empty: 1 the next free position (through counters),
                                 2 every R-th position from the workgroup's rank among the R of
                                 its XCD; with laneGroup set too, vorticity and XSPH run in lane
                                 groups instead */
} akua_tuning;

akua_status akua_solver_set_tuning(akua_solver* s, const akua_tuning* t);
/* Neighbour search cells per smoothing radius h.  1 (the default) is the reference's search
 * (NeighbourSearchCUDA.cu:72-130, :163, :178): hash cells of smoothRadius, the 27-cell stencil of
 * spatialHashCellSize cells; every result is the reference's, bit for bit.  2 is a search that
 * does not keep bit-exactness: cells of h / 2 keyed by their packed coordinates (no hash
 * collisions), walking the cells of the 5x5x5 block around the particle's cell whose boxes come
 * within h of it (about 15.6 h^3 of candidate volume instead of 27 h^3).  Its lists hold exactly
 * the particles within h, in another order; the reference's lists also repeat the particles of a
 * bucket that two stencil cells hash to (NeighbourSearchCUDA.cu:95-127), which this search does
 * not.  Every neighbour sum is therefore rounded in another order (and a capped list may keep
 * other neighbours), so results equal the reference's only to rounding where no list of the
 * reference repeats a neighbour.  The exact density probe (akua_solver_density_error) keeps cells
 * of h either way.  No reference counterpart. */
akua_status akua_solver_set_search(akua_solver* s, int32_t cellsPerH);
akua_status akua_solver_get_search(const akua_solver* s, int32_t* cellsPerH);
/* exact 1 (the default): akua_solver_autotune keeps the search as it is.  0: the caller gives up
 * bit-exactness, and the autotuner also times the search over h / 2 cells and keeps the faster
 * (cache entries of such tunings are keyed apart). */
akua_status akua_solver_set_exact(akua_solver* s, int32_t exact);
akua_status akua_solver_get_tuning(const akua_solver* s, akua_tuning* t);
/* Times short trials of candidate settings by coordinate descent (block, then layout level, then
 * the two digit widths, then the layout interval, the balance tile, the slab batch and the lane
 * group, then, without lane groups, the fused first sweep and the work queue) and keeps the
 * fastest.  Each trial restarts from a copy of the current
 * particle state, which is restored afterwards, so the call leaves the state as it found it.
 * boxMin/boxMax: the box the trials step in (NULL: each scene's current box).  trialSteps: timed
 * steps per candidate (1..1000), run twice; the faster run counts.  cachePath (NULL: no cache): a text file of earlier results keyed
 * by device architecture, log2 of the particle count, log2 of the scene count, neighbour cap and
 * iteration count; a hit applies the cached setting without trials, a miss appends the result.
 * best (may be NULL) receives the setting now in use. */
akua_status akua_solver_autotune(akua_solver* s, float deltaTime, const float boxMin[3],
                                 const float boxMax[3], int32_t trialSteps, const char* cachePath,
                                 akua_tuning* best);

/* ---- boundary samples on the walls ------------------------------------------------------------- */
/* Static samples on the six walls of the box passed to each step (Akinci et al. 2012, in the PBF
 * form of Bender's PositionBasedDynamics): each sample b carries psi_b = rho0 / sum_k W(x_b - x_k)
 * over the samples k; it adds psi_b W to a particle's density, psi_b grad W / rho0 to grad_i C_i,
 * and lambda_i psi_b grad W / rho0 to the position delta.  The soft clamp and the velocity damping
 * stay.  layers: 0 (off, the default and the reference's behaviour) or 1..4 planes per wall, spaced
 * `spacing` apart outward (0: particle_spacing).  No reference counterpart: the reference plans
 * "virtual particles" for this (ConstraintSolverCUDA.cu:132-135).  The samples are rebuilt when the
 * box changes (a host computation and a graph re-capture).  Single-scene solvers; not with the
 * learned step or training.  akua_solver_density_error counts fluid particles only. */
akua_status akua_solver_set_boundary(akua_solver* s, int32_t layers, float spacing);
/* The samples for the current box (valid once a step has passed a box): positions (3 floats per
 * sample) and psi, up to `capacity` samples in index order; *count receives the total.  Any
 * pointer may be NULL. */
akua_status akua_debug_boundary(akua_solver* s, float* pos, float* psi, int32_t capacity,
                                int32_t* count);
/* The same samples without a solver (host only, no GPU): those a solver with this configuration
 * would build for this box.  spacing 0: cfg->particle_spacing. */
akua_status akua_boundary_samples(const akua_pbf_config* cfg, const float boxMin[3],
                                  const float boxMax[3], int32_t layers, float spacing, float* pos,
                                  float* psi, int32_t capacity, int32_t* count);

/* ---- learned warm start ------------------------------------------------------------------------ */
/* A per-particle network (16 features -> 64 -> 64 -> 4, ReLU, BF16 weights and activations on the
 * matrix cores, FP32 accumulation) predicts the position correction of the plain solve from
 * features at the predicted positions x*_pred.  A learned step moves x* to the prediction x*_w,
 * runs J Jacobi iterations from there (the last J of the plain solve), and per scene falls back
 * to the plain solve from x*_pred (all solverIterations iterations, bit-identical to a plain
 * step) when
 *     r(x*_w)  >  ratio * r(x*_pred) + tolerance,
 * r the max (statistic 0) or the mean (statistic 1) over particles of max(0, rho_i / rho0 - 1).
 * No reference counterpart (the reference always runs solverIterations iterations,
 * ConstraintSolverCUDA.cu:173-222).
 *
 * Parameter blob (AKUA_MLP_PARAMS floats): W1[16][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4]
 * (weights input-major), then fmean[16] finv[16] (features enter as (f - fmean) * finv) and
 * ymul[4] (outputs leave multiplied by ymul).  Features, per particle at x*_pred:
 *   0 rho / rho0 - 1, 1 the first iteration's lambda, 2 neighbour count, 3 previous lambda,
 *   4-6 (x*_pred - x) / h, 7-9 2 lambda (1/rho0) sum_j m_j grad W_ij / h,
 *   10-12 (1/rho0) sum_j m_j s_corr,ij grad W_ij / h,
 *   13-15 per axis clamp((x - lo) / h, 0, 2) - clamp((hi - x) / h, 0, 2).
 * Outputs: (x*_{I-J} - x*_pred) / h of the plain solve with I iterations, x*_{I-J} the positions
 * its last J iterations start from (3), and the lambda computed there (1).  Normalised targets
 * are clipped to [-8, 8] in the training loss. */
#define AKUA_MLP_PARAMS 5544
#define AKUA_MLP_TRAINABLE 5508

typedef struct akua_train_config {
    int32_t samplesPerStep;   /* random particles recorded per plain step */
    int32_t minibatch;        /* samples per Adam update */
    int32_t epochs;           /* passes over each step's samples */
    int32_t normSteps;        /* a solver without a network: steps whose samples set the
                                 normalisation before training starts */
    int32_t targetIterations; /* J: the network learns x* before the last J iterations */
    float learningRate;       /* Adam, beta1 0.9, beta2 0.999, eps 1e-8 */
    uint64_t seed;            /* sampling and (for a solver without a network) initial weights */
    float lossWeights[4];     /* per output, on outputs normalised by ymul */
} akua_train_config;

akua_status akua_solver_set_network(akua_solver* s, const float* params, int32_t count);
akua_status akua_solver_get_network(akua_solver* s, float* params, int32_t count);
/* enable 1: learned steps with the fallback rule above (iterations = J, 1..solverIterations,
 * the J the network was trained for); 0: plain steps. */
akua_status akua_solver_set_learned(akua_solver* s, int32_t enable, int32_t iterations,
                                    int32_t statistic, float ratio, float tolerance);
/* Steps run in learned mode, scene-steps that fell back, and the last step's residuals of scene 0:
 * residual[0..2] the max at x*_pred, at x*_w and at the last plain density pass of a recorded
 * step; residual[3..5] the means. */
akua_status akua_solver_learned_stats(akua_solver* s, int64_t* learnedSteps, int64_t* fallbacks,
                                      float residual[6]);
/* Trains the solver's network on its own plain trajectory (single-scene solvers without
 * boundary samples, AKUA_ERR_UNSUPPORTED otherwise; the learned-step settings of
 * akua_solver_set_learned are left as they are): `steps` plain
 * steps (the state advances exactly as akua_solver_step would), each recording features and
 * targets of config->samplesPerStep random particles and running config->epochs passes of Adam
 * over them in minibatches.  A solver without a network first takes the feature and target
 * scales from the samples of config->normSteps steps, then starts from He-normal weights.
 * loss[k] (may be NULL): mean minibatch loss of step k (NaN for a normalisation step);
 * residuals[4k..4k+3] (may be NULL): max and mean of max(0, C) at x*_pred, then at x*_{I-J},
 * of step k. */
akua_status akua_solver_train(akua_solver* s, float deltaTime, const float boxMin[3],
                              const float boxMax[3], int32_t steps, const akua_train_config* config,
                              float* loss, float* residuals);
/* Features (16 per particle) and, after a training step, targets (4 per particle) of the last
 * learned or recorded step, by logical index (the order of akua_solver_get_state). */
akua_status akua_debug_features(akua_solver* s, float* features, float* targets, int32_t n);
/* x*_pred, x*_w (3 floats per particle) and the network's lambda output of the last learned step,
 * by logical index.  x*_pred is only valid when the step did not fall back. */
akua_status akua_debug_warm_start(akua_solver* s, float* xPred, float* xWarm, float* lambdaHat,
                                  int32_t n);
/* Device evaluation of the network on n feature rows (outputs: 4 per row, denormalised). */
akua_status akua_mlp_forward(const float* params, const float* features, int32_t n, float* outputs);
/* Gradient (AKUA_MLP_TRAINABLE floats) and value of the training loss
 * (1/n) sum_rows sum_o lossWeights[o] (y_o - targets_o / ymul_o)^2, y the normalised outputs. */
akua_status akua_mlp_gradient(const float* params, const float* features, const float* targets,
                              int32_t n, const float lossWeights[4], float* grad, float* loss);
/* One Adam step (t >= 1 is the step number for the bias corrections) on the loss over all n rows;
 * params (AKUA_MLP_PARAMS), m and v (AKUA_MLP_TRAINABLE each) are updated in place. */
akua_status akua_mlp_adam_step(float* params, float* m, float* v, int64_t t, float learningRate,
                               const float* features, const float* targets, int32_t n,
                               const float lossWeights[4], float* loss);
/* akua_mlp_adam_step on device memory, allocating nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight training or
 * serving loops).  params, m, v, features, targets (n rows) and loss are device pointers, updated
 * as akua_mlp_adam_step updates them; lossWeights is host memory read during the call.  The work
 * is enqueued on `stream` (NULL: the null stream) and the call returns without waiting.  The
 * gradient partials live in one fixed array of the library per device, 256 x 5509 floats
 * (5.64 MB, held until the process ends), whatever n is, so no n is refused for its size; every
 * call rewrites all of it, and calls that overlap on one device race on it. */
akua_status akua_mlp_adam_step_device(float* params, float* m, float* v, int64_t t,
                                      float learningRate, const float* features,
                                      const float* targets, int32_t n, const float lossWeights[4],
                                      float* loss, void* stream);

/* ---- batched scenes ---------------------------------------------------------------------------- */
/* Many independent scenes in one solver: every kernel of a step covers all scenes in one launch
 * (each 256-slot chunk of particle slots belongs to one scene), and the whole batched step is one
 * hipGraph.  Scene s has counts[s] particles, its own hash table of maxNeighbours * counts[s]
 * entries and its own box; particles never interact across scenes.  No reference counterpart
 * (the reference steps one scene).  set_state/get_state take and return the scenes' particles
 * concatenated in scene order; akua_solver_step_n applies one box to all scenes. */
akua_status akua_batch_create(int32_t numScenes, const int32_t* counts, const akua_pbf_config* cfg,
                              const akua_lambda_corr* corr, akua_solver** out);
akua_status akua_batch_num_scenes(const akua_solver* s, int32_t* numScenes);
akua_status akua_batch_set_box(akua_solver* s, int32_t scene, const float boxMin[3],
                               const float boxMax[3]);
/* nsteps steps of every scene, each with its own box. */
akua_status akua_batch_step(akua_solver* s, float deltaTime, int32_t nsteps);
/* Per-scene max and mean density-constraint error (as akua_solver_density_error). */
akua_status akua_batch_density_error(akua_solver* s, float* maxError, double* meanError);

/* ---- parity taps: the last step's intermediate results, copied to host ---------------------- */
/* Sorted hash keys (sorted_hash[k] is the hash of the particle in slot k after the sort). */
akua_status akua_debug_hashes(akua_solver* s, uint32_t* sorted_hash, int32_t n);
/* perm[k] = slot, in the previous order, of the particle now in slot k. */
akua_status akua_debug_permutation(akua_solver* s, uint32_t* perm, int32_t n);
/* order[p] = logical index (the reference's array position) of the particle stored in physical
 * slot p, per scene (p and the index counted from the scene's first slot).  The physical order is
 * the library's own memory layout; no result depends on it.  No reference counterpart: a check
 * of the layout, for tests. */
akua_status akua_debug_physical_order(akua_solver* s, uint32_t* order, int32_t n);
/* Occupied buckets in sorted order, looked up through the device cell table: for each bucket c
 * (c < *ncells), its hash, first slot and one-past-last slot.  With cap too small, *ncells is
 * still set and AKUA_ERR_INVALID_ARG returned. */
akua_status akua_debug_cell_ranges(akua_solver* s, uint32_t* hash, uint32_t* start, uint32_t* end,
                                   int32_t cap, int32_t* ncells);
/* Neighbour lists in the reference's row-major layout nbr[i * maxNeighbours + k]
 * (NeighbourSearchCUDA.cu:94, :119) and counts cnt[i]; entries past cnt[i] are 0. */
akua_status akua_debug_neighbours(akua_solver* s, uint32_t* nbr, uint32_t* cnt, int32_t n);

/* ---- timing ---------------------------------------------------------------------------------- */
/* With timing on, the solver records HIP events on its stream around each kernel of a step
 * (graphs are then not used).  names/ms: per-kernel accumulated device time and launch count. */
akua_status akua_solver_set_timing(akua_solver* s, int32_t enable);
akua_status akua_solver_reset_timing(akua_solver* s);
/* Fills up to cap entries; *count gets the number of timed kernels.  name_buf receives cap names
 * of name_len bytes each. */
akua_status akua_solver_get_timing(akua_solver* s, char* name_buf, int32_t name_len, double* ms,
                                   int64_t* launches, int32_t cap, int32_t* count);
/* Density-constraint error at the current positions: the exact SPH density of every particle
 * over all particles within h, each counted once (no neighbour cap, no duplicates from hash
 * collisions), then max and mean over particles of max(0, density / restDensity - 1).  Runs on
 * the device; overwrites the step's scratch (so the parity taps then describe this evaluation). */
akua_status akua_solver_density_error(akua_solver* s, float* maxError, double* meanError);
/* Densities of the last akua_solver_density_error evaluation, in the order it sorted the
 * particles (hash order of the current positions). */
akua_status akua_debug_probe_density(akua_solver* s, float* rho, int32_t n);

/* Achievable device bandwidth: a grid-stride float4 copy of `bytes` bytes, `reps` times, timed
 * with HIP events; *gbps counts bytes read plus bytes written.  Also the known-size kernel that
 * calibrates hardware byte counters. */
akua_status akua_stream_copy_bandwidth(int64_t bytes, int32_t reps, double* gbps);

/* Device self-test of the written-out correctly rounded division and square root the kernels use
 * (exact equality with '/' and sqrtf over n random inputs per check).  mismatches[0]: division
 * and its shared reciprocal (equal to 1.0f / b over divisors in [2^-40, 2^40)),
 * [1]: square root over [2^-96, 2^128), [2]: square roots below 2^-96 that reach 2^-47,
 * [3]: square root over squared lengths in [0, 2^-2).  All four are 0 on a conforming device. */
akua_status akua_selftest_arith(uint64_t n, uint64_t seed, uint64_t mismatches[4]);

/* Mean neighbour count over the particles of the last step (device reduction, on demand), for
 * the byte model; 0 before the first step. */
akua_status akua_solver_mean_neighbours(akua_solver* s, double* mean);

#ifdef __cplusplus
}
#endif

#endif /* AKUA_H */
