/*
 * sph_oracle.h -- CPU restatement of the reference SPH step (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP path, not part of the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * It restates, operation for operation in IEEE FP32 (no contraction, no
 * reassociation), the reference hot path:
 *   SpatialHash::get_grid_coords/hash_position  src/spatial_hash.h:20-36
 *   SpatialHash::build                          src/spatial_hash.cpp:15-25
 *   SpatialHash::query_squared                  src/spatial_hash.cpp:31-57, src/spatial_hash.h:70-73
 *   CubicSplineKernel W/gradW/laplacianW (3-D)  src/kernels.cpp:23-36,58-66,96-108,130-153
 *   equations::compute_density .. acceleration  src/sph_engine.cpp:370-443
 *   SPHEngine::step                             src/sph_engine.cpp:93-144
 *   SPHEngine::integrate_leapfrog               src/sph_engine.cpp:290-310
 *   ParticleSystem::apply_boundary_conditions   src/particle.cpp:122-153
 *   SPHEngine::compute_cfl_timestep             src/sph_engine.cpp:312-333
 *   get_total_mass / get_total_energy           src/sph_engine.cpp:187-200
 * plus an FP64 evaluation of density and acceleration over the same neighbour
 * lists (the error yardstick for the FP32 paths).
 *
 * The reference ships no tests or golden vectors for this path and cannot be compiled here
 * (its sources need GLM, absent), so every function is pinned by tests/test_oracle.py:
 *   grid coords / keys    closed form on random points, specials (NaN, negatives, rounding)
 *   neighbour query       brute force O(n^2), symmetry, lexicographic order, the FP32 pair
 *                         two cells apart that a 27-cell scan would miss
 *   W, gradW, laplacianW  closed forms, FP64 finite differences, integral 2/3
 *   density               perfect-lattice bulk value (FP64 666.7195) and neighbour count 257
 *   forces                FP64 lattice interior a = g, pressure-term sign, viscosity sign
 *   leapfrog, clamp, CFL  closed forms
 * Unpinned: the FP32 bit patterns the reference binary would give under -ffast-math; the
 * restatement follows the source in IEEE arithmetic.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field order and defaults follow SPHParameters (src/sph_engine.h:14-34). */
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
    int32_t formulation; /* 0 = reference formulas, 1 = consistent (sigma=3/(2 pi h^3), a=F/rho_i) */
} or_params;

typedef struct or_state or_state;

/* which-selectors for or_get */
enum { OR_POS = 0, OR_VEL = 1, OR_DENSITY = 2, OR_PRESSURE = 3, OR_ACCEL = 4, OR_MASS = 5 };

void      or_default_params(or_params* p);
or_state* or_create(size_t capacity);
void      or_destroy(or_state* s);
void      or_set_params(or_state* s, const or_params* p);
/* Replaces the particle set (ids = index).  pos/vel are xyz-interleaved, n*3 floats. */
int       or_set_particles(or_state* s, size_t n, const float* pos, const float* vel, const float* mass);
size_t    or_size(const or_state* s);
/* Full reference step; dt <= 0 selects the CFL timestep.  Returns the dt used. */
float     or_step(or_state* s, float dt);
/* Sub-phases, in step order (for phase-by-phase parity). */
void      or_update_neighbor_lists(or_state* s);
void      or_compute_densities(or_state* s);
void      or_compute_pressures(or_state* s);
void      or_compute_forces(or_state* s);
void      or_integrate_leapfrog(or_state* s, float dt);
void      or_apply_boundary_conditions(or_state* s);
float     or_compute_cfl_timestep(const or_state* s);
/* Copies N rows (3 floats per row for POS/VEL/ACCEL, 1 otherwise). */
void      or_get(const or_state* s, int which, float* out);
/* Neighbour lists in reference order (CSR). offsets has n+1 entries. */
size_t    or_neighbor_total(const or_state* s);
void      or_get_neighbors(const or_state* s, int64_t* offsets, int64_t* idx);
size_t    or_max_neighbors(const or_state* s);
float     or_total_mass(const or_state* s);
float     or_total_energy(const or_state* s);
/* Number of hash buckets (occupied cells) after the last build. */
size_t    or_num_cells(const or_state* s);

/* Standalone pieces (pinned by closed forms). */
uint64_t  or_hash_position(int32_t x, int32_t y, int32_t z);
void      or_grid_coords(float inv_cell, const float* p, int32_t* c);
void      or_cell_keys(size_t n, const float* pos, float cell_size, int32_t* coords, uint64_t* keys);
float     or_W(const float* r, float h, int32_t formulation);
void      or_gradW(const float* r, float h, int32_t formulation, float* out);
float     or_laplacianW(const float* r, float h, int32_t formulation);
/* Leapfrog + wall clamp on caller arrays (n particles), in place. */
void      or_integrate_clamp(size_t n, float* pos, float* vel, const float* acc, float dt,
                             const or_params* p);

/* FP64 density and acceleration over given neighbour lists (reference order).
 * pos/vel/mass are FP32 inputs widened to double. */
void      or64_density_accel(size_t n, const float* pos, const float* vel, const float* mass,
                             const int64_t* offsets, const int64_t* idx, const or_params* p,
                             double* rho, double* acc);
/* Per-particle sum of |pressure+viscosity force terms| / m (the conditioning yardstick for
 * acceleration comparisons), FP64 over the given lists. */
void      or64_accel_term_scale(size_t n, const float* pos, const float* vel, const float* mass,
                                const float* rho, const int64_t* offsets, const int64_t* idx,
                                const or_params* p, double* scale);
/* Sampled evaluation for scenes too large for a neighbour vector per particle.
 * or_count_neighbors: builds the hash over all particles and runs the query of every
 *   particle counting only; returns the sum of the list lengths, the longest in *max_out.
 * or_sample_step0: builds the hash over all particles, queries the sampled particles and
 *   their neighbours, and evaluates densities (FP32, for the sample and its neighbours) and
 *   accelerations (FP32, for the sample) exactly as the full phases would; per sample k:
 *   list length, density, acceleration (3), and over the same lists the FP64 density,
 *   FP64 acceleration (3) and the conditioning scale of or64_accel_term_scale.  Any output
 *   may be NULL.  Returns -1 for an index outside [0, n).
 * or_sample_neighbors: the lists of the sample after or_sample_step0 (CSR, ns + 1 offsets;
 *   idx NULL for the offsets only). */
uint64_t  or_count_neighbors(or_state* s, uint64_t* max_out);
int       or_sample_step0(or_state* s, size_t ns, const int64_t* sample, int64_t* counts,
                          float* rho, float* acc, double* rho64, double* acc64, double* scale);
void      or_sample_neighbors(const or_state* s, size_t ns, const int64_t* sample,
                              int64_t* offsets, int64_t* idx);
/* Brute-force O(n^2) neighbour sets with the same d^2 operation order (ascending j). */
size_t    or_brute_neighbors(size_t n, const float* pos, float r2, int64_t* offsets, int64_t* idx,
                             size_t cap);

int       or_num_threads(void);

/* ---- learned macro-step surrogate (surrogate_oracle.c) ---- */
#define OR_SUR_NF 9 /* edge features */
typedef struct {
    int32_t h1, h2, h3; /* widths: message hidden, message output, node hidden */
    int32_t bf16;       /* 1: round to bfloat16 where the device does */
    float inv_r;        /* 1/R, scale of r and |r| */
    float inv_v;        /* 1/v_s, scale of velocities */
    float inv_h;        /* 1/h, kernel argument q = |r|/h */
    float w_norm;       /* 1/w_shape(0) */
    float rho0;         /* rest density */
    float rho_scale;    /* node feature (rho - rho0) * rho_scale */
    float msg_scale;    /* S_i = msg_scale * sum_j ReLU(u_ij) */
    float acc_scale;    /* a = acc_scale * out */
} or_sur_cfg;

float     or_bf16_round(float f);
size_t    or_sur_nparams(const or_sur_cfg* c);
/* Network output (normalised units, n*3) over CSR neighbour lists that may include self.
 * S_out (n*h1) and count_out (n) are optional. */
void      or_sur_forward(const or_sur_cfg* c, const float* params, size_t n, const float* pos,
                         const float* vel, const float* rho, const int64_t* off,
                         const int64_t* idx, double* out, double* S_out, int64_t* count_out);
/* Mean-square loss over the particles with in_batch[i] != 0 and its gradient (np values). */
double    or_sur_loss_grad(const or_sur_cfg* c, const float* params, size_t n, const float* pos,
                           const float* vel, const float* rho, const int64_t* off,
                           const int64_t* idx, const uint8_t* in_batch, const float* target,
                           double* grad);
/* One Adam step (t >= 1) in place, FP64 arithmetic. */
void      or_sur_adam(size_t np, float* w, float* m, float* v, const float* g, float lr,
                      float beta1, float beta2, float eps, int32_t t);

#ifdef __cplusplus
}
#endif
#endif
