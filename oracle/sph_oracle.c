/*
 * sph_oracle.c -- CPU restatement of the reference SPH step (TEST INFRASTRUCTURE ONLY).
 *
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Never linked into the product library.
 *
 * Floating-point contract: every expression below is written in the evaluation
 * order of the reference source (GLM vec3 operators are component-wise; glm::dot
 * is (x*x + y*y) + z*z; glm::length is sqrt(dot)).  The parity build is compiled
 * with -ffp-contract=off and without -ffast-math so that each operation rounds
 * once, as written.  The baseline build of the same file uses the reference's
 * release flags (-O3 -mavx2 -ffast-math -fopenmp, CMakeLists.txt:148).
 */
#include "sph_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI_F 3.14159265358979323846f
#define OR_PI_D 3.14159265358979323846

/* ---------------------------------------------------------------------------
 * Small growable arrays standing in for std::vector<size_t>.
 * ------------------------------------------------------------------------- */
typedef struct {
    size_t* v;
    size_t n, cap;
} vec_sz;

static void vsz_push(vec_sz* a, size_t x) {
    if (a->n == a->cap) {
        size_t nc = a->cap ? a->cap * 2 : 8;
        a->v = (size_t*)realloc(a->v, nc * sizeof(size_t));
        a->cap = nc;
    }
    a->v[a->n++] = x;
}

/* ---------------------------------------------------------------------------
 * Spatial hash: std::unordered_map<uint64_t, std::vector<size_t>> restated as an
 * open-addressing table of buckets.  Bucket contents are appended in ascending
 * particle index, as SpatialHash::build does (src/spatial_hash.cpp:19-24).
 * ------------------------------------------------------------------------- */
typedef struct {
    uint64_t key;
    int used;
    vec_sz items;
} bucket;

typedef struct {
    bucket* slots;
    size_t nslots; /* power of two */
    size_t count;
} hashmap;

static uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

static void hm_clear(hashmap* m) {
    for (size_t i = 0; i < m->nslots; ++i) {
        free(m->slots[i].items.v);
    }
    free(m->slots);
    m->slots = NULL;
    m->nslots = 0;
    m->count = 0;
}

static void hm_init(hashmap* m, size_t expect) {
    size_t ns = 16;
    while (ns < expect * 2 + 16) ns <<= 1;
    m->slots = (bucket*)calloc(ns, sizeof(bucket));
    m->nslots = ns;
    m->count = 0;
}

static bucket* hm_find(const hashmap* m, uint64_t key) {
    if (!m->nslots) return NULL;
    size_t mask = m->nslots - 1;
    size_t i = (size_t)mix64(key) & mask;
    for (;;) {
        bucket* b = &m->slots[i];
        if (!b->used) return NULL;
        if (b->key == key) return b;
        i = (i + 1) & mask;
    }
}

/* Doubles the table, moving every bucket (its items keep their order). */
static void hm_grow(hashmap* m) {
    const size_t ns = m->nslots * 2;
    bucket* slots = (bucket*)calloc(ns, sizeof(bucket));
    const size_t mask = ns - 1;
    for (size_t k = 0; k < m->nslots; ++k) {
        const bucket* b = &m->slots[k];
        if (!b->used) continue;
        size_t i = (size_t)mix64(b->key) & mask;
        while (slots[i].used) i = (i + 1) & mask;
        slots[i] = *b;
    }
    free(m->slots);
    m->slots = slots;
    m->nslots = ns;
}

/* The table is kept at most half full (a sparse scene has about as many cells as particles),
 * so a probe always reaches a free slot. */
static bucket* hm_get_or_insert(hashmap* m, uint64_t key) {
    if (2 * (m->count + 1) > m->nslots) hm_grow(m);
    size_t mask = m->nslots - 1;
    size_t i = (size_t)mix64(key) & mask;
    for (;;) {
        bucket* b = &m->slots[i];
        if (!b->used) {
            b->used = 1;
            b->key = key;
            m->count++;
            return b;
        }
        if (b->key == key) return b;
        i = (i + 1) & mask;
    }
}

/* ---------------------------------------------------------------------------
 * Hash key and grid coordinates (src/spatial_hash.h:20-36).
 * static_cast<int>(std::floor(v)) is undefined in C++ for NaN or out-of-range v;
 * on x86 (cvttss2si) it yields INT_MIN, which is what is restated here.
 * ------------------------------------------------------------------------- */
uint64_t or_hash_position(int32_t x, int32_t y, int32_t z) {
    uint64_t hash = 0;
    hash |= ((uint64_t)((uint32_t)x & 0x1FFFFFu) << 42);
    hash |= ((uint64_t)((uint32_t)y & 0x1FFFFFu) << 21);
    hash |= ((uint64_t)((uint32_t)z & 0x1FFFFFu));
    return hash;
}

static int32_t x86_floor_to_int(float v) {
    float f = floorf(v);
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT_MIN;
    return (int32_t)f;
}

void or_grid_coords(float inv_cell, const float* p, int32_t* c) {
    c[0] = x86_floor_to_int(p[0] * inv_cell);
    c[1] = x86_floor_to_int(p[1] * inv_cell);
    c[2] = x86_floor_to_int(p[2] * inv_cell);
}

void or_cell_keys(size_t n, const float* pos, float cell_size, int32_t* coords, uint64_t* keys) {
    /* SpatialHash::set_cell_size: inv_cell_size_ = 1.0f / cell_size_ (src/spatial_hash.h:63-66) */
    float inv = 1.0f / cell_size;
    for (size_t i = 0; i < n; ++i) {
        int32_t c[3];
        or_grid_coords(inv, pos + 3 * i, c);
        if (coords) {
            coords[3 * i + 0] = c[0];
            coords[3 * i + 1] = c[1];
            coords[3 * i + 2] = c[2];
        }
        if (keys) keys[i] = or_hash_position(c[0], c[1], c[2]);
    }
}

/* within_radius_squared (src/spatial_hash.h:70-73): diff = pos1 - pos2; dot(diff,diff) <= r2 */
static inline int within_r2(const float* a, const float* b, float r2) {
    float dx = a[0] - b[0];
    float dy = a[1] - b[1];
    float dz = a[2] - b[2];
    float tx = dx * dx, ty = dy * dy, tz = dz * dz;
    float d2 = (tx + ty) + tz;
    return d2 <= r2;
}

/* ---------------------------------------------------------------------------
 * Cubic spline kernel, 3-D (src/kernels.cpp:23-36,58-66,96-108,130-153).
 * formulation 1 multiplies sigma by 3/2 (normalised kernel); the formulas are
 * otherwise unchanged.
 * ------------------------------------------------------------------------- */
typedef struct {
    float h, h_inv, h_sq, sigma;
} kern;

static kern make_kernel(float h, int formulation) {
    kern k;
    k.h = h;
    k.h_inv = 1.0f / h;
    k.h_sq = h * h;
    if (formulation == 1)
        k.sigma = 3.0f / (2.0f * OR_PI_F * h * h * h);
    else
        k.sigma = 1.0f / (OR_PI_F * h * h * h);
    return k;
}

static inline float vlen(const float* r) {
    float tx = r[0] * r[0], ty = r[1] * r[1], tz = r[2] * r[2];
    return sqrtf((tx + ty) + tz);
}

static inline float W3(const kern* k, float q) {
    if (q >= 0.0f && q <= 1.0f) {
        float a = 2.0f / 3.0f - q * q;
        float b = 0.5f * q;
        b = b * q;
        b = b * q;
        return k->sigma * (a + b);
    } else if (q > 1.0f && q <= 2.0f) {
        float tmp = 2.0f - q;
        float c = (1.0f / 6.0f) * tmp;
        c = c * tmp;
        c = c * tmp;
        return k->sigma * c;
    }
    return 0.0f;
}

static inline float kW(const kern* k, const float* r) {
    float q = vlen(r) / k->h; /* kernel_utils::compute_q (src/kernels.h:117-119) */
    return W3(k, q);
}

static inline void kgradW(const kern* k, const float* r, float* g) {
    float q = vlen(r) / k->h;
    float r_len = vlen(r);
    g[0] = g[1] = g[2] = 0.0f;
    if (r_len < 1e-6f) return;
    float rn[3] = {r[0] / r_len, r[1] / r_len, r[2] / r_len};
    float s;
    if (q >= 0.0f && q <= 1.0f) {
        float poly = -2.0f * q;
        float t = 1.5f * q;
        t = t * q;
        poly = poly + t;
        s = k->sigma * poly;
    } else if (q > 1.0f && q <= 2.0f) {
        float tmp = 2.0f - q;
        float t = 0.5f * tmp;
        t = t * tmp;
        s = (-k->sigma) * t;
    } else {
        /* grad stays vec3(0); returned as 0/h */
        g[0] = 0.0f / k->h;
        g[1] = 0.0f / k->h;
        g[2] = 0.0f / k->h;
        return;
    }
    g[0] = (s * rn[0]) / k->h;
    g[1] = (s * rn[1]) / k->h;
    g[2] = (s * rn[2]) / k->h;
}

static inline float klapW(const kern* k, const float* r) {
    float q = vlen(r) / k->h;
    if (q >= 0.0f && q <= 1.0f) {
        float t = 3.0f * q;
        t = -2.0f + t;
        return (k->sigma * t) / k->h_sq;
    } else if (q > 1.0f && q <= 2.0f) {
        float tmp = 2.0f - q;
        return (k->sigma * tmp) / k->h_sq;
    }
    return 0.0f;
}

float or_W(const float* r, float h, int32_t formulation) {
    kern k = make_kernel(h, formulation);
    return kW(&k, r);
}

void or_gradW(const float* r, float h, int32_t formulation, float* out) {
    kern k = make_kernel(h, formulation);
    kgradW(&k, r, out);
}

float or_laplacianW(const float* r, float h, int32_t formulation) {
    kern k = make_kernel(h, formulation);
    return klapW(&k, r);
}

/* ---------------------------------------------------------------------------
 * Engine state.  AoS-free: the restatement keeps SoA arrays in original index
 * order; values and operation order are what matters for parity.
 * Buffers sized to capacity as in SPHEngine::initialize (src/sph_engine.cpp:25-29).
 * ------------------------------------------------------------------------- */
struct or_state {
    size_t cap, n;
    or_params p;
    float* pos;   /* n*3 */
    float* vel;   /* n*3 */
    float* mass;  /* n */
    float* dens;  /* cap */
    float* pres;  /* cap */
    float* acc;   /* cap*3 */
    vec_sz* nbr;  /* cap */
    hashmap grid;
    float inv_cell;
    size_t max_nbr;
};

void or_default_params(or_params* p) {
    p->rest_density = 1000.0f;
    p->gas_constant = 2000.0f;
    p->viscosity = 0.001f;
    p->smoothing_length = 0.02f;
    p->particle_mass = 0.001f;
    p->timestep = 0.001f;
    p->gravity = -9.81f;
    p->damping = 0.99f;
    p->CFL_factor = 0.4f;
    p->xmin = -1.0f; p->xmax = 1.0f;
    p->ymin = -1.0f; p->ymax = 1.0f;
    p->zmin = -1.0f; p->zmax = 1.0f;
    p->neighbor_search_radius = 0.04f;
    p->formulation = 0;
}

or_state* or_create(size_t capacity) {
    or_state* s = (or_state*)calloc(1, sizeof(or_state));
    s->cap = capacity;
    or_default_params(&s->p);
    s->inv_cell = 1.0f / s->p.neighbor_search_radius;
    s->pos = (float*)calloc(capacity * 3 + 3, sizeof(float));
    s->vel = (float*)calloc(capacity * 3 + 3, sizeof(float));
    s->mass = (float*)calloc(capacity + 1, sizeof(float));
    s->dens = (float*)calloc(capacity + 1, sizeof(float));
    s->pres = (float*)calloc(capacity + 1, sizeof(float));
    s->acc = (float*)calloc(capacity * 3 + 3, sizeof(float));
    s->nbr = (vec_sz*)calloc(capacity + 1, sizeof(vec_sz));
    return s;
}

void or_destroy(or_state* s) {
    if (!s) return;
    hm_clear(&s->grid);
    for (size_t i = 0; i < s->cap; ++i) free(s->nbr[i].v);
    free(s->nbr);
    free(s->pos); free(s->vel); free(s->mass);
    free(s->dens); free(s->pres); free(s->acc);
    free(s);
}

void or_set_params(or_state* s, const or_params* p) {
    /* SPHEngine::initialize: hash cell = params.neighbor_search_radius (src/sph_engine.cpp:20-23) */
    s->p = *p;
    s->inv_cell = 1.0f / p->neighbor_search_radius;
}

int or_set_particles(or_state* s, size_t n, const float* pos, const float* vel, const float* mass) {
    if (n > s->cap) n = s->cap; /* ParticleSystem::add_particle drops beyond capacity (src/particle.cpp:26-33) */
    s->n = n;
    memcpy(s->pos, pos, n * 3 * sizeof(float));
    if (vel)
        memcpy(s->vel, vel, n * 3 * sizeof(float));
    else
        memset(s->vel, 0, n * 3 * sizeof(float));
    if (mass)
        memcpy(s->mass, mass, n * sizeof(float));
    else
        for (size_t i = 0; i < n; ++i) s->mass[i] = s->p.particle_mass;
    return (int)n;
}

size_t or_size(const or_state* s) { return s->n; }

/* SpatialHash::build (src/spatial_hash.cpp:15-25): every particle appended, in ascending
 * index, to the bucket of its cell key. */
static void build_hash(or_state* s) {
    size_t n = s->n;
    hm_clear(&s->grid);
    hm_init(&s->grid, n / 8 + 16);
    for (size_t i = 0; i < n; ++i) {
        int32_t c[3];
        or_grid_coords(s->inv_cell, s->pos + 3 * i, c);
        bucket* b = hm_get_or_insert(&s->grid, or_hash_position(c[0], c[1], c[2]));
        vsz_push(&b->items, i);
    }
}

/* SpatialHash::query_squared (src/spatial_hash.cpp:31-57) for particle i: the
 * (2 cr + 1)^3 cells around its cell in (dx, dy, dz) lexicographic order, each bucket in
 * ascending index.  Appends to `out` when given; returns the count either way. */
static size_t query_one(const or_state* s, size_t i, vec_sz* out) {
    float r2 = s->p.neighbor_search_radius * s->p.neighbor_search_radius;
    int cr = (int)ceilf(sqrtf(r2) * s->inv_cell) + 1;
    const float* pi = s->pos + 3 * i;
    int32_t c[3];
    or_grid_coords(s->inv_cell, pi, c);
    size_t cnt = 0;
    for (int dx = -cr; dx <= cr; ++dx) {
        for (int dy = -cr; dy <= cr; ++dy) {
            for (int dz = -cr; dz <= cr; ++dz) {
                /* ivec3 addition wraps in two's complement */
                int32_t cx = (int32_t)((uint32_t)c[0] + (uint32_t)dx);
                int32_t cy = (int32_t)((uint32_t)c[1] + (uint32_t)dy);
                int32_t cz = (int32_t)((uint32_t)c[2] + (uint32_t)dz);
                bucket* b = hm_find(&s->grid, or_hash_position(cx, cy, cz));
                if (!b) continue;
                for (size_t k = 0; k < b->items.n; ++k) {
                    size_t j = b->items.v[k];
                    if (within_r2(pi, s->pos + 3 * j, r2)) {
                        if (out) vsz_push(out, j);
                        ++cnt;
                    }
                }
            }
        }
    }
    return cnt;
}

/* SPHEngine::update_neighbor_lists (src/sph_engine.cpp:335-353) ->
 * SpatialHash::build (src/spatial_hash.cpp:15-25) + query_squared (:31-57). */
void or_update_neighbor_lists(or_state* s) {
    size_t n = s->n;
    build_hash(s);
    size_t maxn = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(max : maxn)
#endif
    for (long long ii = 0; ii < (long long)n; ++ii) {
        size_t i = (size_t)ii;
        vec_sz* out = &s->nbr[i];
        out->n = 0;
        query_one(s, i, out);
        if (out->n > maxn) maxn = out->n;
    }
    s->max_nbr = maxn;
}

/* equations::compute_density (src/sph_engine.cpp:370-383) over particle i's list */
static float density_one(const or_state* s, const kern* k, float w0, size_t i) {
    const float* pi = s->pos + 3 * i;
    float density = s->mass[i] * w0;
    const vec_sz* nb = &s->nbr[i];
    for (size_t t = 0; t < nb->n; ++t) {
        size_t j = nb->v[t];
        if (i == j) continue;
        const float* pj = s->pos + 3 * j;
        float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
        density += s->mass[j] * kW(k, r);
    }
    return density;
}

/* compute_densities + equations::compute_density (src/sph_engine.cpp:203-212,370-383) */
void or_compute_densities(or_state* s) {
    kern k = make_kernel(s->p.smoothing_length, s->p.formulation);
    float zero[3] = {0.0f, 0.0f, 0.0f};
    float w0 = kW(&k, zero);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long long ii = 0; ii < (long long)s->n; ++ii) {
        s->dens[ii] = density_one(s, &k, w0, (size_t)ii);
    }
}

/* compute_pressures + equations::compute_pressure (src/sph_engine.cpp:214-224,385-388) */
void or_compute_pressures(or_state* s) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long long ii = 0; ii < (long long)s->n; ++ii) {
        s->pres[ii] = s->p.gas_constant * (s->dens[ii] - s->p.rest_density);
    }
}

/* compute_forces (src/sph_engine.cpp:226-244) with compute_pressure_force (:390-412),
 * compute_viscosity_force (:414-432), compute_external_forces (:434-436) and
 * compute_acceleration (:438-443). */
static void force_one(or_state* s, const kern* kp, size_t i) {
    const kern k = *kp;
    const float visc = s->p.viscosity;
    const float grav = s->p.gravity;
    const int consistent = s->p.formulation == 1;
    const float* pi = s->pos + 3 * i;
    const float* vi = s->vel + 3 * i;
    const vec_sz* nb = &s->nbr[i];
    float fp[3] = {0.0f, 0.0f, 0.0f};
    for (size_t t = 0; t < nb->n; ++t) {
        size_t j = nb->v[t];
        if (i == j) continue;
        const float* pj = s->pos + 3 * j;
        float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
        float r_len = vlen(r);
        if (r_len < 1e-6f) continue;
        float pressure_term = (s->pres[i] + s->pres[j]) / (2.0f * s->dens[j]);
        float g[3];
        kgradW(&k, r, g);
        float c = s->mass[j] * pressure_term;
        fp[0] = fp[0] - c * g[0];
        fp[1] = fp[1] - c * g[1];
        fp[2] = fp[2] - c * g[2];
    }
    float fv[3] = {0.0f, 0.0f, 0.0f};
    for (size_t t = 0; t < nb->n; ++t) {
        size_t j = nb->v[t];
        if (i == j) continue;
        const float* pj = s->pos + 3 * j;
        const float* vj = s->vel + 3 * j;
        float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
        float vij[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
        float lap = klapW(&k, r);
        float c = (s->mass[j] / s->dens[j]) * visc;
        fv[0] = fv[0] + (c * vij[0]) * lap;
        fv[1] = fv[1] + (c * vij[1]) * lap;
        fv[2] = fv[2] + (c * vij[2]) * lap;
    }
    float ext[3] = {0.0f, grav, 0.0f};
    float m = consistent ? s->dens[i] : s->mass[i];
    for (int d = 0; d < 3; ++d) {
        s->acc[3 * i + d] = ((fp[d] + fv[d]) + ext[d] * m) / m;
    }
}

void or_compute_forces(or_state* s) {
    kern k = make_kernel(s->p.smoothing_length, s->p.formulation);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long long ii = 0; ii < (long long)s->n; ++ii) force_one(s, &k, (size_t)ii);
}

static inline void integrate_one(float* x, float* v, const float* a, float dt, float damping) {
    /* integrate_leapfrog body (src/sph_engine.cpp:296-308) */
    for (int d = 0; d < 3; ++d) {
        float half = 0.5f * a[d];
        v[d] = v[d] + half * dt;
    }
    for (int d = 0; d < 3; ++d) x[d] = x[d] + v[d] * dt;
    for (int d = 0; d < 3; ++d) {
        float half = 0.5f * a[d];
        v[d] = v[d] + half * dt;
    }
    for (int d = 0; d < 3; ++d) v[d] = v[d] * damping;
}

static inline void clamp_one(float* x, float* v, const or_params* p) {
    /* ParticleSystem::apply_boundary_conditions (src/particle.cpp:122-153) */
    const float damping = 0.8f;
    const float lo[3] = {p->xmin, p->ymin, p->zmin};
    const float hi[3] = {p->xmax, p->ymax, p->zmax};
    for (int d = 0; d < 3; ++d) {
        if (x[d] < lo[d]) {
            x[d] = lo[d];
            v[d] = v[d] * -damping;
        } else if (x[d] > hi[d]) {
            x[d] = hi[d];
            v[d] = v[d] * -damping;
        }
    }
}

void or_integrate_leapfrog(or_state* s, float dt) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long long ii = 0; ii < (long long)s->n; ++ii) {
        integrate_one(s->pos + 3 * ii, s->vel + 3 * ii, s->acc + 3 * ii, dt, s->p.damping);
    }
}

void or_apply_boundary_conditions(or_state* s) {
    for (size_t i = 0; i < s->n; ++i) clamp_one(s->pos + 3 * i, s->vel + 3 * i, &s->p);
}

void or_integrate_clamp(size_t n, float* pos, float* vel, const float* acc, float dt,
                        const or_params* p) {
    for (size_t i = 0; i < n; ++i) integrate_one(pos + 3 * i, vel + 3 * i, acc + 3 * i, dt, p->damping);
    for (size_t i = 0; i < n; ++i) clamp_one(pos + 3 * i, vel + 3 * i, p);
}

/* SPHEngine::compute_cfl_timestep (src/sph_engine.cpp:312-333).  The force-based bound
 * reads accelerations_[0] only, as the reference does. */
float or_compute_cfl_timestep(const or_state* s) {
    float max_velocity = 0.0f;
    for (size_t i = 0; i < s->n; ++i) {
        float v = vlen(s->vel + 3 * i);
        max_velocity = (max_velocity < v) ? v : max_velocity; /* std::max(a, b): a < b ? b : a */
    }
    float dt_cfl = (s->p.CFL_factor * s->p.smoothing_length) / (max_velocity + 1e-6f);
    float dt_force = s->p.CFL_factor * sqrtf(s->p.smoothing_length / (vlen(s->acc) + 1e-6f));
    /* std::min({a, b, c}) keeps the first of equal minima */
    float m = dt_cfl;
    if (dt_force < m) m = dt_force;
    if (s->p.timestep < m) m = s->p.timestep;
    return m;
}

float or_step(or_state* s, float dt) {
    if (s->n == 0) return dt;
    if (dt <= 0.0f) dt = or_compute_cfl_timestep(s);
    or_update_neighbor_lists(s);
    or_compute_densities(s);
    or_compute_pressures(s);
    or_compute_forces(s);
    or_integrate_leapfrog(s, dt);
    or_apply_boundary_conditions(s);
    return dt;
}

void or_get(const or_state* s, int which, float* out) {
    size_t n = s->n;
    switch (which) {
        case OR_POS: memcpy(out, s->pos, n * 3 * sizeof(float)); break;
        case OR_VEL: memcpy(out, s->vel, n * 3 * sizeof(float)); break;
        case OR_DENSITY: memcpy(out, s->dens, n * sizeof(float)); break;
        case OR_PRESSURE: memcpy(out, s->pres, n * sizeof(float)); break;
        case OR_ACCEL: memcpy(out, s->acc, n * 3 * sizeof(float)); break;
        case OR_MASS: memcpy(out, s->mass, n * sizeof(float)); break;
        default: break;
    }
}

size_t or_neighbor_total(const or_state* s) {
    size_t t = 0;
    for (size_t i = 0; i < s->n; ++i) t += s->nbr[i].n;
    return t;
}

void or_get_neighbors(const or_state* s, int64_t* offsets, int64_t* idx) {
    int64_t o = 0;
    for (size_t i = 0; i < s->n; ++i) {
        offsets[i] = o;
        for (size_t t = 0; t < s->nbr[i].n; ++t) idx[o++] = (int64_t)s->nbr[i].v[t];
    }
    offsets[s->n] = o;
}

size_t or_max_neighbors(const or_state* s) { return s->max_nbr; }
size_t or_num_cells(const or_state* s) { return s->grid.count; }

/* get_total_mass (src/sph_engine.cpp:187-190): std::accumulate in float, then * h^3 */
float or_total_mass(const or_state* s) {
    float acc = 0.0f;
    for (size_t i = 0; i < s->n; ++i) acc = acc + s->dens[i];
    float h = s->p.smoothing_length;
    return acc * ((h * h) * h);
}

/* get_total_energy (src/sph_engine.cpp:192-200) */
float or_total_energy(const or_state* s) {
    float ke = 0.0f;
    for (size_t i = 0; i < s->n; ++i) {
        const float* v = s->vel + 3 * i;
        float tx = v[0] * v[0], ty = v[1] * v[1], tz = v[2] * v[2];
        float speed_sq = (tx + ty) + tz;
        ke += (0.5f * s->mass[i]) * speed_sq;
    }
    return ke;
}

/* ---------------------------------------------------------------------------
 * FP64 yardstick over given neighbour lists.
 * ------------------------------------------------------------------------- */
typedef struct {
    double h, h_sq, sigma;
} kern64;

static kern64 make_kernel64(double h, int formulation) {
    kern64 k;
    k.h = h;
    k.h_sq = h * h;
    k.sigma = (formulation == 1 ? 1.5 : 1.0) / (OR_PI_D * h * h * h);
    return k;
}

static double W64(const kern64* k, double q) {
    if (q >= 0.0 && q <= 1.0) return k->sigma * (2.0 / 3.0 - q * q + 0.5 * q * q * q);
    if (q > 1.0 && q <= 2.0) {
        double t = 2.0 - q;
        return k->sigma * (t * t * t / 6.0);
    }
    return 0.0;
}

static double dW64(const kern64* k, double q) { /* dW/dq */
    if (q >= 0.0 && q <= 1.0) return k->sigma * (-2.0 * q + 1.5 * q * q);
    if (q > 1.0 && q <= 2.0) {
        double t = 2.0 - q;
        return -k->sigma * 0.5 * t * t;
    }
    return 0.0;
}

static double L64(const kern64* k, double q) {
    if (q >= 0.0 && q <= 1.0) return k->sigma * (-2.0 + 3.0 * q) / k->h_sq;
    if (q > 1.0 && q <= 2.0) return k->sigma * (2.0 - q) / k->h_sq;
    return 0.0;
}

/* FP64 density of particle i over its list (self skipped, added as m_i W(0)). */
static double dens64_one(const kern64* k, size_t i, const float* pos, const float* mass,
                         const int64_t* list, int64_t len) {
    double d = (double)mass[i] * W64(k, 0.0);
    for (int64_t t = 0; t < len; ++t) {
        size_t j = (size_t)list[t];
        if (j == i) continue;
        double r[3] = {(double)pos[3 * i] - pos[3 * j], (double)pos[3 * i + 1] - pos[3 * j + 1],
                       (double)pos[3 * i + 2] - pos[3 * j + 2]};
        double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        d += (double)mass[j] * W64(k, rl / k->h);
    }
    return d;
}

/* FP64 acceleration of particle i over its list, densities rho (indexed like pos). */
static void acc64_one(const kern64* k, size_t i, const float* pos, const float* vel,
                      const float* mass, const int64_t* list, int64_t len, const double* rho,
                      const or_params* p, double* acc) {
    double Pi = (double)p->gas_constant * (rho[i] - (double)p->rest_density);
    double fp[3] = {0, 0, 0}, fv[3] = {0, 0, 0};
    for (int64_t t = 0; t < len; ++t) {
        size_t j = (size_t)list[t];
        if (j == i) continue;
        double r[3] = {(double)pos[3 * i] - pos[3 * j], (double)pos[3 * i + 1] - pos[3 * j + 1],
                       (double)pos[3 * i + 2] - pos[3 * j + 2]};
        double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        double q = rl / k->h;
        double Pj = (double)p->gas_constant * (rho[j] - (double)p->rest_density);
        if (rl >= 1e-6) {
            double pt = (Pi + Pj) / (2.0 * rho[j]);
            double gs = dW64(k, q) / (rl * k->h);
            for (int c = 0; c < 3; ++c) fp[c] -= (double)mass[j] * pt * gs * r[c];
        }
        double lap = L64(k, q);
        double cc = (double)mass[j] / rho[j] * (double)p->viscosity * lap;
        for (int c = 0; c < 3; ++c)
            fv[c] += cc * ((double)vel[3 * j + c] - (double)vel[3 * i + c]);
    }
    double m = p->formulation == 1 ? rho[i] : (double)mass[i];
    double ext[3] = {0.0, (double)p->gravity, 0.0};
    for (int c = 0; c < 3; ++c) acc[c] = (fp[c] + fv[c] + ext[c] * m) / m;
}

/* Sum of the magnitudes of particle i's pressure and viscosity terms over its list, divided
 * by m_i (rho_i in the consistent formulation), plus |g|. */
static double scale_one(const kern64* k, size_t i, const float* pos, const float* vel,
                        const float* mass, const float* rho, const int64_t* list, int64_t len,
                        const or_params* p) {
    double Pi = (double)p->gas_constant * ((double)rho[i] - (double)p->rest_density);
    double sum = 0.0;
    for (int64_t t = 0; t < len; ++t) {
        size_t j = (size_t)list[t];
        if (j == i) continue;
        double r[3] = {(double)pos[3 * i] - pos[3 * j], (double)pos[3 * i + 1] - pos[3 * j + 1],
                       (double)pos[3 * i + 2] - pos[3 * j + 2]};
        double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        double q = rl / k->h;
        double Pj = (double)p->gas_constant * ((double)rho[j] - (double)p->rest_density);
        if (rl >= 1e-6) {
            double pt = (Pi + Pj) / (2.0 * (double)rho[j]);
            sum += fabs((double)mass[j] * pt * dW64(k, q) / k->h);
        }
        double lap = L64(k, q);
        double dv[3] = {(double)vel[3 * j] - vel[3 * i], (double)vel[3 * j + 1] - vel[3 * i + 1],
                        (double)vel[3 * j + 2] - vel[3 * i + 2]};
        double dvl = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        sum += fabs((double)mass[j] / (double)rho[j] * (double)p->viscosity * lap) * dvl;
    }
    double m = p->formulation == 1 ? (double)rho[i] : (double)mass[i];
    return sum / fabs(m) + fabs((double)p->gravity);
}

void or64_density_accel(size_t n, const float* pos, const float* vel, const float* mass,
                        const int64_t* offsets, const int64_t* idx, const or_params* p,
                        double* rho, double* acc) {
    kern64 k = make_kernel64((double)p->smoothing_length, p->formulation);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long long ii = 0; ii < (long long)n; ++ii) {
        size_t i = (size_t)ii;
        rho[i] = dens64_one(&k, i, pos, mass, idx + offsets[i], offsets[i + 1] - offsets[i]);
    }
    if (!acc) return;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long long ii = 0; ii < (long long)n; ++ii) {
        size_t i = (size_t)ii;
        acc64_one(&k, i, pos, vel, mass, idx + offsets[i], offsets[i + 1] - offsets[i], rho, p,
                  acc + 3 * i);
    }
}

void or64_accel_term_scale(size_t n, const float* pos, const float* vel, const float* mass,
                           const float* rho, const int64_t* offsets, const int64_t* idx,
                           const or_params* p, double* scale) {
    kern64 k = make_kernel64((double)p->smoothing_length, p->formulation);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long long ii = 0; ii < (long long)n; ++ii) {
        size_t i = (size_t)ii;
        scale[i] = scale_one(&k, i, pos, vel, mass, rho, idx + offsets[i],
                             offsets[i + 1] - offsets[i], p);
    }
}

/* ---------------------------------------------------------------------------
 * Sampled evaluation for scenes too large for a neighbour vector per particle.
 * ------------------------------------------------------------------------- */
uint64_t or_count_neighbors(or_state* s, uint64_t* max_out) {
    build_hash(s);
    uint64_t total = 0, mx = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : total) reduction(max : mx)
#endif
    for (long long ii = 0; ii < (long long)s->n; ++ii) {
        uint64_t c = (uint64_t)query_one(s, (size_t)ii, NULL);
        total += c;
        if (c > mx) mx = c;
    }
    if (max_out) *max_out = mx;
    return total;
}

int or_sample_step0(or_state* s, size_t ns, const int64_t* sample, int64_t* counts, float* rho,
                    float* acc, double* rho64, double* acc64, double* scale) {
    const size_t n = s->n;
    for (size_t k = 0; k < ns; ++k)
        if (sample[k] < 0 || (size_t)sample[k] >= n) return -1;
    build_hash(s);
    /* need: bit 0 density (the sample and its neighbours), bit 1 sample, bit 2 queried */
    uint8_t* need = (uint8_t*)calloc(n ? n : 1, 1);
    size_t* uniq = (size_t*)malloc((ns ? ns : 1) * sizeof(size_t));
    size_t nu = 0;
    for (size_t k = 0; k < ns; ++k) {
        size_t i = (size_t)sample[k];
        if (!(need[i] & 2)) uniq[nu++] = i;
        need[i] |= 3;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (long long u = 0; u < (long long)nu; ++u) {
        size_t i = uniq[u];
        s->nbr[i].n = 0;
        query_one(s, i, &s->nbr[i]);
    }
    for (size_t u = 0; u < nu; ++u) {
        const vec_sz* nb = &s->nbr[uniq[u]];
        need[uniq[u]] |= 4;
        for (size_t t = 0; t < nb->n; ++t) need[nb->v[t]] |= 1;
    }
    size_t na = 0;
    for (size_t i = 0; i < n; ++i) na += need[i] & 1;
    size_t* A = (size_t*)malloc((na ? na : 1) * sizeof(size_t));
    na = 0;
    for (size_t i = 0; i < n; ++i)
        if (need[i] & 1) A[na++] = i;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (long long a = 0; a < (long long)na; ++a) {
        size_t i = A[a];
        if (need[i] & 4) continue;
        s->nbr[i].n = 0;
        query_one(s, i, &s->nbr[i]);
    }
    kern k = make_kernel(s->p.smoothing_length, s->p.formulation);
    float zero[3] = {0.0f, 0.0f, 0.0f};
    float w0 = kW(&k, zero);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (long long a = 0; a < (long long)na; ++a) {
        size_t i = A[a];
        s->dens[i] = density_one(s, &k, w0, i);
        s->pres[i] = s->p.gas_constant * (s->dens[i] - s->p.rest_density);
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (long long u = 0; u < (long long)nu; ++u) force_one(s, &k, uniq[u]);
    /* FP64 over the same lists: densities of A, then the samples' accelerations */
    kern64 k64 = make_kernel64((double)s->p.smoothing_length, s->p.formulation);
    double* r64 = NULL;
    if (rho64 || acc64) {
        r64 = (double*)calloc(n ? n : 1, sizeof(double));
        int64_t** lists = (int64_t**)calloc(na ? na : 1, sizeof(int64_t*));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
        for (long long a = 0; a < (long long)na; ++a) {
            size_t i = A[a];
            const vec_sz* nb = &s->nbr[i];
            int64_t* l = (int64_t*)malloc((nb->n ? nb->n : 1) * sizeof(int64_t));
            for (size_t t = 0; t < nb->n; ++t) l[t] = (int64_t)nb->v[t];
            r64[i] = dens64_one(&k64, i, s->pos, s->mass, l, (int64_t)nb->n);
            lists[a] = l;
        }
        for (size_t a = 0; a < na; ++a) free(lists[a]);
        free(lists);
    }
    for (size_t kk = 0; kk < ns; ++kk) {
        size_t i = (size_t)sample[kk];
        const vec_sz* nb = &s->nbr[i];
        if (counts) counts[kk] = (int64_t)nb->n;
        if (rho) rho[kk] = s->dens[i];
        if (acc)
            for (int d = 0; d < 3; ++d) acc[3 * kk + d] = s->acc[3 * i + d];
        if (!rho64 && !acc64 && !scale) continue;
        int64_t* l = (int64_t*)malloc((nb->n ? nb->n : 1) * sizeof(int64_t));
        for (size_t t = 0; t < nb->n; ++t) l[t] = (int64_t)nb->v[t];
        if (rho64) rho64[kk] = r64[i];
        if (acc64) acc64_one(&k64, i, s->pos, s->vel, s->mass, l, (int64_t)nb->n, r64, &s->p,
                             acc64 + 3 * kk);
        if (scale)
            scale[kk] = scale_one(&k64, i, s->pos, s->vel, s->mass, s->dens, l, (int64_t)nb->n,
                                  &s->p);
        free(l);
    }
    /* the lists of the non-sample particles are dropped; the samples' stay readable through
     * or_sample_neighbors */
    for (size_t a = 0; a < na; ++a) {
        size_t i = A[a];
        if (need[i] & 2) continue;
        free(s->nbr[i].v);
        s->nbr[i].v = NULL;
        s->nbr[i].n = s->nbr[i].cap = 0;
    }
    free(r64);
    free(A);
    free(uniq);
    free(need);
    return 0;
}

void or_sample_neighbors(const or_state* s, size_t ns, const int64_t* sample, int64_t* offsets,
                         int64_t* idx) {
    int64_t o = 0;
    for (size_t k = 0; k < ns; ++k) {
        const vec_sz* nb = &s->nbr[(size_t)sample[k]];
        offsets[k] = o;
        if (idx)
            for (size_t t = 0; t < nb->n; ++t) idx[o + (int64_t)t] = (int64_t)nb->v[t];
        o += (int64_t)nb->n;
    }
    offsets[ns] = o;
}

size_t or_brute_neighbors(size_t n, const float* pos, float r2, int64_t* offsets, int64_t* idx,
                          size_t cap) {
    size_t o = 0;
    for (size_t i = 0; i < n; ++i) {
        offsets[i] = (int64_t)o;
        for (size_t j = 0; j < n; ++j) {
            if (within_r2(pos + 3 * i, pos + 3 * j, r2)) {
                if (o < cap) idx[o] = (int64_t)j;
                ++o;
            }
        }
    }
    offsets[n] = (int64_t)o;
    return o;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
