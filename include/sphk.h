/*
 * sphk.h -- C ABI of the MI355X SPH step library (libsphk.so).
 *
 * The reference has no FFI: its boundary is the C++ SPHEngine class
 * (src/sph_engine.h:91-140) and the pybind11 module built from
 * python/bindings.cpp.  Each entry point below replaces one piece of that
 * class; the C++ SPHEngine and the `sph` Python module of this project are
 * thin layers over these calls.
 *
 * Conventions
 *  - Every function returns int status: 0 = ok, negative = error (SPHK_E_*).
 *    The message of the last error on the calling thread is available from
 *    sphk_last_error().  No C++ exception crosses this boundary.
 *  - Host pointers are plain arrays; xyz quantities are interleaved (n*3 floats).
 *  - One engine owns one HIP stream on one device.  Calls on one engine must
 *    not overlap from several threads.  sphk_step() enqueues work and returns;
 *    every read-back call synchronises the engine's stream first.
 *  - Particle order on input and output is the caller's order (ids 0..n-1),
 *    as in ParticleSystem (src/particle.cpp:26-33: id = index).
 */
#ifndef SPHK_H
#define SPHK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHK_OK 0
#define SPHK_E_INVALID (-1)   /* bad argument */
#define SPHK_E_HIP (-2)       /* HIP runtime error */
#define SPHK_E_CAPACITY (-3)  /* particle count above capacity */
#define SPHK_E_DOMAIN (-4)    /* positions need a cell grid beyond the supported size */
#define SPHK_E_NODEVICE (-5)  /* no GPU visible */

typedef struct sphk_engine sphk_engine;

/* Same fields and defaults as SPHParameters (src/sph_engine.h:14-34), flattened.
 * `formulation` is an addition: 0 (default) = the reference's formulas,
 * 1 = consistent SPH (kernel normalised by 3/(2 pi h^3), acceleration = F / rho_i). */
typedef struct {
    float rest_density;
    float gas_constant;
    float viscosity;
    float smoothing_length;
    float particle_mass;
    float timestep;
    float gravity;
    float damping;
    float CFL_factor;
    float xmin, xmax, ymin, ymax, zmin, zmax;
    float neighbor_search_radius;
    int32_t formulation;
} sphk_params;

/* Phase timings and neighbour statistics of the steps since the last reset,
 * mirroring SPHEngine::PerformanceStats (src/sph_engine.h:51-59).  Times in seconds,
 * measured with HIP events recorded on the engine's stream around each phase when
 * timing is enabled (no synchronisation per step; events are read by sphk_get_stats).
 * neighbor_search_time covers cell keys, sort, cell table and gather;
 * force_computation_time covers the fused force + leapfrog + clamp kernel, so
 * integration_time stays 0. */
typedef struct {
    double total_time;
    double neighbor_search_time;
    double density_computation_time;
    double force_computation_time;
    double integration_time;
    uint64_t max_neighbors;
    uint64_t total_neighbor_queries;
    uint64_t total_neighbors;  /* sum of neighbour-list lengths (incl. self) */
    uint64_t steps;
    uint64_t timed_steps;      /* steps whose phases were timed (timing enabled) */
} sphk_stats;

/* read-back selectors for sphk_download */
#define SPHK_POSITIONS 0     /* n*3 */
#define SPHK_VELOCITIES 1    /* n*3 */
#define SPHK_DENSITIES 2     /* capacity entries (SPHEngine::get_densities, src/sph_engine.h:135) */
#define SPHK_PRESSURES 3     /* capacity entries (src/sph_engine.h:136) */
#define SPHK_ACCELERATIONS 4 /* n*3, the accelerations_ buffer (src/sph_engine.h:64) */
#define SPHK_MASSES 5        /* n */

/* Default parameter values of SPHParameters{} (src/sph_engine.h:14-34). */
int sphk_default_params(sphk_params* out);

/* Lifecycle.  Replaces SPHEngine(size_t max_particles) (src/sph_engine.cpp:13-18).
 * capacity: 1 .. 2^27 - 2 particles (the walks address neighbour rows of 32 bytes with
 * 32-bit byte offsets); larger requests fail with SPHK_E_INVALID. */
int sphk_create(size_t capacity, sphk_engine** out);
int sphk_destroy(sphk_engine* e);

/* Replaces SPHEngine::initialize / set_parameters / set_boundaries
 * (src/sph_engine.cpp:20-33,152-172): stores the parameters used by later steps. */
int sphk_set_params(sphk_engine* e, const sphk_params* p);
int sphk_get_params(const sphk_engine* e, sphk_params* out);

/* Writes particles [offset, offset+n) and sets the particle count to offset+n.
 * offset must not exceed the current count.  vel may be NULL (zero velocity);
 * mass may be NULL (params.particle_mass).  Replaces ParticleSystem::add_particles
 * (src/particle.cpp:35-39) when offset == current count, and clear+add when 0.
 * Particles beyond capacity are dropped and SPHK_E_CAPACITY is returned after the
 * ones that fit are stored, as ParticleSystem::add_particle warns and drops
 * (src/particle.cpp:26-33). */
int sphk_upload(sphk_engine* e, size_t offset, size_t n, const float* pos_xyz,
                const float* vel_xyz, const float* mass);
/* Batched scenes.  Splits the particles (caller order) into nscenes independent scenes:
 * scene k holds particles [offsets[k], offsets[k+1]) and is clamped to its own walls
 * bounds[6k .. 6k+5] = (xmin, xmax, ymin, ymax, zmin, zmax) (NULL: params' bounds for all).
 * offsets[0] == 0 and offsets[nscenes] must equal the particle count when stepping.  No
 * neighbour search crosses two scenes; each scene gets its own dense cell grid, and one step
 * runs every scene in the same launches.  nscenes == 0 returns to one scene holding every
 * particle.  No reference counterpart (the reference runs one scene per engine). */
int sphk_set_scenes(sphk_engine* e, size_t nscenes, const uint64_t* offsets, const float* bounds);
int sphk_scene_count(const sphk_engine* e, size_t* out);

/* Steps are captured into a HIP graph and replayed while the step's shape (particle count,
 * grids, launch shapes, dt mode) is unchanged.  Steps with timing enabled run eagerly.
 * No reference counterpart. */
int sphk_set_graph(sphk_engine* e, int enabled);

/* Particle count (ParticleSystem::size). */
int sphk_size(const sphk_engine* e, size_t* out);
/* Resets the particle count, simulated time and step counter
 * (SPHEngine::clear_particles, src/sph_engine.cpp:87-91). */
int sphk_clear(sphk_engine* e);

/* One SPH step (SPHEngine::step, src/sph_engine.cpp:93-144): cell grid build,
 * density, pressure, forces, leapfrog, wall clamp.  dt <= 0 selects the CFL
 * timestep (src/sph_engine.cpp:312-333), computed on the device.  Enqueued on
 * the engine's stream.  A no-op when the engine holds no particles. */
int sphk_step(sphk_engine* e, float dt);
/* Runs num_steps steps (SPHEngine::run_steps, src/sph_engine.cpp:146-150). */
int sphk_run_steps(sphk_engine* e, size_t num_steps, int adaptive_timestep);

/* Renderer instance data.  Replaces the per-frame host repack of
 * Renderer::update_particle_data (src/renderer.cpp:279-311): writes, on the device, 9 floats
 * per particle in caller order -- position (3), velocity (3), colour (3) -- into
 * out_instances, a device buffer of at least 9 n floats (for example a mapped OpenGL vertex
 * buffer).  colors_rgb: optional device array of n * 3 floats; NULL gives every particle the
 * default colour of Particle (0, 0.5, 1) (src/particle.h:42).  Synchronises the stream. */
int sphk_render_pack(sphk_engine* e, const float* colors_rgb, float* out_instances,
                     size_t count);

/* Read-backs (synchronise the stream). */
int sphk_download(sphk_engine* e, int which, float* out, size_t count);
int sphk_get_time(sphk_engine* e, float* out);
int sphk_get_step_count(const sphk_engine* e, uint64_t* out);
/* get_total_mass (sum of densities * h^3), get_total_energy (kinetic),
 * compute_conservation_errors (src/sph_engine.cpp:178-200). */
int sphk_total_mass(sphk_engine* e, float* out);
int sphk_total_energy(sphk_engine* e, float* out);
int sphk_conservation_errors(sphk_engine* e, float* mass_error, float* energy_error);
/* The kinetic energy of sphk_total_energy without allocating: no hipMalloc, hipFree or host
 * buffer on any call, for callers that cannot allocate (code under stream capture, serving
 * loops, harnesses that forbid it).  No counterpart in the reference.  Each term is the same
 * FP32 (0.5 m) |v|^2; the terms are summed on the device in FP64 in a fixed order (the
 * result differs from sphk_total_energy's sequential FP32 sum only by that sum's rounding).
 * Partial sums use a device scratch of 1,025 doubles per GPU, shared by every engine on that
 * GPU and kept until the process ends; any particle count is handled (the first pass strides
 * over the particles).  Calls that overlap on one GPU race on that scratch. */
int sphk_total_energy_noalloc(sphk_engine* e, float* out);

/* Performance stats (SPHEngine::get_performance_stats / reset_performance_stats). */
int sphk_set_timing(sphk_engine* e, int enabled);
int sphk_get_stats(sphk_engine* e, sphk_stats* out);
int sphk_reset_stats(sphk_engine* e);

/* Device-resident snapshot of the particle state (positions, velocities,
 * accelerations, time) and its restore, both device-to-device on the stream.
 * No reference counterpart: used for reset-and-step measurement and for
 * surrogate training trajectories. */
int sphk_snapshot_save(sphk_engine* e);
int sphk_snapshot_restore(sphk_engine* e);

/* Synchronises the engine's stream. */
int sphk_synchronize(sphk_engine* e);

/* ---- parity/debug read-backs (run the cell-grid build for the current state) ----
 * sphk_debug_cell_coords: per particle (caller order) the integer cell coordinates
 *   floor(p * (1/cell)) of SpatialHash::get_grid_coords (src/spatial_hash.h:30-36),
 *   n*3 int32; non-finite coordinates give INT32_MIN like x86's conversion.
 * sphk_debug_cell_keys: the 63-bit SpatialHash::hash_position keys
 *   (src/spatial_hash.h:20-27), n uint64.
 * sphk_debug_sorted: the cell-sorted order (caller ids, ascending within a cell,
 *   as SpatialHash::build appends them, src/spatial_hash.cpp:19-24) and the
 *   per-particle dense cell index; n entries each.
 * sphk_debug_neighbors: neighbour lists in SpatialHash::query_squared order
 *   (src/spatial_hash.cpp:31-57), CSR over caller ids; call with idx == NULL to
 *   get the total in *total, then again with buffers of that size. */
int sphk_debug_cell_coords(sphk_engine* e, int32_t* coords_xyz, size_t n);
int sphk_debug_cell_keys(sphk_engine* e, uint64_t* keys, size_t n);
int sphk_debug_sorted(sphk_engine* e, uint32_t* order, uint32_t* cell_of_sorted, size_t n);
int sphk_debug_neighbors(sphk_engine* e, int64_t* offsets, int64_t* idx, size_t cap,
                         size_t* total);
/* The candidate words the density walk of the production step stores for the force walk
 * (the matrix-core filter's output), produced for the current state with the engine's
 * density launch shape; nothing but work arrays is written.  Per chunk (64 consecutive
 * sorted slots of one grid column): items[c] its first sorted slot, nwords[c] its words,
 * wj0[c * 256 + w] the first sorted slot of word w, mask[(c * 256 + w) * 64 + l] lane l's
 * bits (candidate k in bit 31 - k).  The engine stores W words per chunk (*words_stored,
 * if given: 256, or 128 where the word buffers would pass 16 GiB); chunks of more than W
 * words store only their first W (the force walk tests such chunks again), and words W to
 * 255 of every row are left as they were.  Call with items == NULL to get the chunk count
 * in *nchunks, then with arrays of that many chunks.  Replaces no reference function: it
 * exposes the filter that stands in for SpatialHash::query_squared
 * (src/spatial_hash.cpp:31-57) so its pairs can be checked one by one. */
int sphk_debug_walk_words(sphk_engine* e, uint32_t* items, uint32_t* nwords, uint32_t* wj0,
                          uint32_t* mask, size_t chunk_cap, size_t* nchunks,
                          size_t* words_stored);
/* Grid shape of the last build: origin cell (3), dims (3), number of cells. */
int sphk_debug_grid(sphk_engine* e, int32_t* origin3, int32_t* dims3, uint64_t* ncells);

/* Launch shapes of the two neighbour walks (LDS words per segment x wavefronts per
 * workgroup; "_split" shapes walk each chunk with three wavefronts, one per part of its
 * candidate columns, and sum the parts in a second kernel; "pp" shapes hold two segments,
 * so a lane through one goes on into the next), indexed 0..sphk_walk_shape_count()-1.
 * Results do not depend on the shape (each particle's sums run in the same order in every
 * shape); only speed does.  Until set, an engine uses seg32_waves4_split below 65,536
 * particles and seg24_waves4 (density) with pp16_waves4 (force) above (chosen again whenever
 * the particle count changes).  No reference counterpart: chosen by the autotuner of the
 * `sph` package. */
int sphk_set_walk_shapes(sphk_engine* e, int density_shape, int force_shape);
int sphk_get_walk_shapes(const sphk_engine* e, int* density_shape, int* force_shape);
int sphk_walk_shape_count(void);
const char* sphk_walk_shape_name(int shape);

/* ---- learned macro-step surrogate ------------------------------------------------------
 * No counterpart in the reference (its repository holds no learned component).  A
 * message-passing MLP over the step's own neighbour lists predicts the acceleration of one
 * macro-step of `substeps` plain steps:
 *   edge e_ij = [r_ij/R, (v_i - v_j)/v_s, |r_ij|/R, W(|r_ij|)/W(0),
 *                (rho_j - rho0) k_rho]                                      (9)
 *   S_i = c sum_{j != i} ReLU(W1 e_ij + b1)      m_i = W2 S_i + c n_i b2     (128, 128)
 *   z_i = ReLU(Wn1 [m_i, (rho_i - rho0) k_rho, v_i / v_s] + bn1)           (132 -> 128)
 *   a_i = a_s (Wn2 z_i + bn2)                                                (3)
 * The macro-step over T = substeps * dt is v' = d^substeps v + a T, x' = x + T (v + v') / 2
 * followed by the wall clamp (d = damping).  Training targets are
 * a* = (v(t+T) - d^substeps v(t)) / T from the engine's own plain steps.  Dense layers run
 * on bf16 MFMA with FP32 accumulation; parameters are FP32 (Adam) and rounded to bf16 for
 * every use. */
typedef struct {
    int32_t substeps;     /* plain steps replaced by one macro-step (4) */
    int32_t judge_every;  /* macro-steps per judge run (16) */
    float tolerance;      /* fallback above this mean |rho - rho_SPH| in kg/m^3 (10 = 1% of 1000) */
    float inv_v;          /* 1/v_s, velocity feature scale (1 s/m) */
    float rho_scale;      /* k_rho, density feature scale (0.01 m^3/kg) */
    float msg_scale;      /* c, message sum scale (1/64) */
    float acc_scale;      /* a_s, output scale (10 m/s^2) */
} sphk_surrogate_config;

/* Training on trajectories the engine itself generates.  Each snapshot is one macro-step
 * of plain SPH from the current state: its targets come from the plain run, then
 * `iters_per_snapshot` Adam steps train on random batches of chunks (64 cell-sorted
 * particles each) of the snapshot's starting state, and the run continues from the plain
 * result.  The engine ends `snapshots` macro-steps later in simulated time. */
typedef struct {
    uint64_t snapshots;          /* 2000 */
    uint32_t iters_per_snapshot; /* 1 */
    uint32_t batch_chunks;       /* 1024 chunks = 65,536 particles */
    float dt;                    /* plain step of the generating run */
    float lr;                    /* Adam learning rate (1e-3; beta 0.9, 0.999, eps 1e-8) */
    uint64_t seed;               /* batch selection */
} sphk_train_config;

/* Outcome of sphk_surrogate_run. */
typedef struct {
    uint64_t macro_steps;        /* macro-steps advanced by the network */
    uint64_t plain_macro_steps;  /* macro-steps advanced by plain SPH (judge runs, fallback) */
    uint64_t judges;             /* judge runs */
    float last_error;            /* mean |rho - rho_SPH| of the last judge run, kg/m^3 */
    float max_error;             /* largest mean error of any judge run */
    float max_abs_error;         /* largest per-particle |rho - rho_SPH| of any judge run */
    int32_t fell_back;           /* 1 once the run switched to plain SPH */
    uint64_t fallback_at;        /* macro-step index of the switch */
    float error_bound;           /* accumulated density error bound of the run, kg/m^3 */
} sphk_surrogate_report;

int sphk_surrogate_default_config(sphk_surrogate_config* out);
/* Number of FP32 parameters (35,203). */
int sphk_surrogate_param_count(size_t* out);
/* Creates (or resets) the engine's network: Xavier-uniform weights from `seed`, zero
 * biases, zero Adam state.  NULL cfg = defaults. */
int sphk_surrogate_init(sphk_engine* e, const sphk_surrogate_config* cfg, uint64_t seed);
int sphk_surrogate_get_config(sphk_engine* e, sphk_surrogate_config* out);
/* Flat parameters W1[128][9] b1[128] W2[128][128] b2[128] Wn1[128][132] bn1[128]
 * Wn2[3][128] bn2[3].  set_params resets the Adam state. */
int sphk_surrogate_get_params(sphk_engine* e, float* out, size_t count);
int sphk_surrogate_set_params(sphk_engine* e, const float* params, size_t count);
/* Network acceleration (m/s^2) of every particle of the current state, caller order,
 * count >= 3 n floats.  The state is not advanced. */
int sphk_surrogate_predict(sphk_engine* e, float* out, size_t count);
/* Targets of the current state: advances the engine by one macro-step of plain SPH at dt
 * and writes a* (m/s^2, caller order, 3 n floats). */
int sphk_surrogate_targets(sphk_engine* e, float dt, float* out, size_t count);
/* Mean-square loss (normalised units) and its gradient for the given targets (m/s^2, caller
 * order) on a batch of `batch_chunks` random chunks of the current state (all chunks when
 * larger than their number).  in_batch (optional, n bytes) marks the batch particles. */
int sphk_surrogate_loss_grad(sphk_engine* e, const float* target, size_t count,
                             uint64_t batch_seed, uint32_t batch_chunks, float* loss,
                             float* grad, size_t grad_count, uint8_t* in_batch);
/* One Adam step with the given gradient (count = parameter count). */
int sphk_surrogate_adam(sphk_engine* e, const float* grad, size_t count, float lr);
/* Generates and trains (see sphk_train_config).  loss_curve (optional) receives the mean
 * batch loss of each snapshot, up to curve_len values. */
int sphk_surrogate_train(sphk_engine* e, const sphk_train_config* tc, float* loss_curve,
                         size_t curve_len);
/* One macro-step of the network (no judge). */
int sphk_surrogate_macro_step(sphk_engine* e, float dt);
/* Runs macro_steps macro-steps with the judge.  Macro-steps 1, 1 + judge_every,
 * 1 + 2 judge_every, ... are judge runs: the network and plain SPH advance from the same
 * state, the mean |rho - rho_SPH| of the two results is the network's divergence per
 * macro-step, and the plain result is kept.  The other macro-steps are run by the network.
 * The judge bounds the run's density error against a plain run of the same length N: with e
 * the largest divergence measured so far, velocity errors that add up like a random walk give
 * at most (4/3) e (N - 1)^(3/2) at the end of the run once any network step has run
 * (error_bound); the network runs only while that bound is within the tolerance, after which
 * the run continues with plain SPH only. */
int sphk_surrogate_run(sphk_engine* e, uint64_t macro_steps, float dt,
                       sphk_surrogate_report* rep);
/* Densities of the current state (one density pass, caller order, n floats). */
int sphk_surrogate_densities(sphk_engine* e, float* out, size_t count);

/* Last error message of the calling thread; returns its length. */
int sphk_last_error(char* buf, size_t len);
/* Library version string ("1.0.0", the reference module's __version__). */
const char* sphk_version(void);

#ifdef __cplusplus
}
#endif
#endif
