/* Sanitizer run of the CPU oracle (test infrastructure): built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by tests/native/Makefile and run by test_native_sanitizers.py.
 * Exercises the step (both formulations), neighbour lists against brute force, non-finite
 * and far-away particles, the CFL step, the FP64 evaluation and the surrogate's forward,
 * gradient and Adam step on a small network.  Prints "sanitize ok" and exits 0 when every
 * check holds; a sanitizer report aborts the run. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sph_oracle.h"

static uint64_t rng = 0x9e3779b97f4a7c15ull;
static float uniform01(void) {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (float)((rng >> 40) * (1.0 / 16777216.0));
}

#define CHECK(cond, msg)                              \
    do {                                              \
        if (!(cond)) {                                \
            fprintf(stderr, "check failed: %s\n", msg); \
            return 1;                                 \
        }                                             \
    } while (0)

static int run_scene(int formulation, size_t n) {
    float* pos = malloc(3 * n * sizeof(float));
    float* vel = malloc(3 * n * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        size_t a = i % 13, b = (i / 13) % 13, c = i / 169;
        pos[3 * i + 0] = -0.2f + 0.01f * (float)a + 0.002f * uniform01();
        pos[3 * i + 1] = -0.3f + 0.01f * (float)b + 0.002f * uniform01();
        pos[3 * i + 2] = 0.1f + 0.01f * (float)c + 0.002f * uniform01();
        vel[3 * i + 0] = 0.1f * (uniform01() - 0.5f);
        vel[3 * i + 1] = 0.1f * (uniform01() - 0.5f);
        vel[3 * i + 2] = 0.1f * (uniform01() - 0.5f);
    }
    pos[0] = NAN;              /* a non-finite particle */
    pos[3 * (n - 1)] = 0.9f;   /* a lone particle near the wall */
    or_params p;
    or_default_params(&p);
    p.formulation = formulation;
    or_state* s = or_create(n + 8);
    CHECK(s != NULL, "create");
    or_set_params(s, &p);
    CHECK(or_set_particles(s, n, pos, vel, NULL) == (int)n, "set_particles");
    or_update_neighbor_lists(s);
    const size_t tot = or_neighbor_total(s);
    int64_t* off = malloc((n + 1) * sizeof(int64_t));
    int64_t* idx = malloc((tot ? tot : 1) * sizeof(int64_t));
    or_get_neighbors(s, off, idx);
    const float r2 = p.neighbor_search_radius * p.neighbor_search_radius;
    int64_t* boff = malloc((n + 1) * sizeof(int64_t));
    const size_t btot = or_brute_neighbors(n, pos, r2, boff, NULL, 0);
    int64_t* bidx = malloc((btot ? btot : 1) * sizeof(int64_t));
    or_brute_neighbors(n, pos, r2, boff, bidx, btot);
    CHECK(btot == tot, "neighbour total vs brute force");
    for (int step = 0; step < 3; ++step) {
        const float dt = formulation == 1 ? 1e-4f : or_compute_cfl_timestep(s);
        or_step(s, dt);
    }
    float* rho = malloc(n * sizeof(float));
    float* acc = malloc(3 * n * sizeof(float));
    or_get(s, OR_DENSITY, rho);
    or_get(s, OR_ACCEL, acc);
    CHECK(or_max_neighbors(s) > 0, "max neighbours");
    (void)or_total_mass(s);
    (void)or_total_energy(s);
    double* rho64 = malloc(n * sizeof(double));
    double* acc64 = malloc(3 * n * sizeof(double));
    float* mass = malloc(n * sizeof(float));
    float* cur = malloc(3 * n * sizeof(float));
    float* curv = malloc(3 * n * sizeof(float));
    or_get(s, OR_MASS, mass);
    or_get(s, OR_POS, cur);
    or_get(s, OR_VEL, curv);
    or_update_neighbor_lists(s);
    const size_t tot2 = or_neighbor_total(s);
    int64_t* off2 = malloc((n + 1) * sizeof(int64_t));
    int64_t* idx2 = malloc((tot2 ? tot2 : 1) * sizeof(int64_t));
    or_get_neighbors(s, off2, idx2);
    or64_density_accel(n, cur, curv, mass, off2, idx2, &p, rho64, acc64);
    free(pos); free(vel); free(off); free(idx); free(boff); free(bidx); free(rho); free(acc);
    free(rho64); free(acc64); free(mass); free(cur); free(curv); free(off2); free(idx2);
    or_destroy(s);
    return 0;
}

static int run_surrogate(void) {
    const size_t n = 300;
    float* pos = malloc(3 * n * sizeof(float));
    float* vel = malloc(3 * n * sizeof(float));
    float* rho = malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            pos[3 * i + d] = 0.1f * uniform01();
            vel[3 * i + d] = uniform01() - 0.5f;
        }
        rho[i] = 1000.0f + 10.0f * (uniform01() - 0.5f);
    }
    const float r2 = 0.04f * 0.04f;
    int64_t* off = malloc((n + 1) * sizeof(int64_t));
    const size_t tot = or_brute_neighbors(n, pos, r2, off, NULL, 0);
    int64_t* idx = malloc(tot * sizeof(int64_t));
    or_brute_neighbors(n, pos, r2, off, idx, tot);
    or_sur_cfg c = {16, 16, 16, 1, 25.0f, 1.0f, 50.0f, 1.5f, 1000.0f, 0.01f, 1.0f / 64.0f, 10.0f};
    const size_t np = or_sur_nparams(&c);
    float* w = malloc(np * sizeof(float));
    float* m = calloc(np, sizeof(float));
    float* v = calloc(np, sizeof(float));
    float* g = malloc(np * sizeof(float));
    double* gd = malloc(np * sizeof(double));
    for (size_t k = 0; k < np; ++k) w[k] = 0.2f * (uniform01() - 0.5f);
    double* out = malloc(3 * n * sizeof(double));
    double* S = malloc(n * (size_t)c.h1 * sizeof(double));
    int64_t* cnt = malloc(n * sizeof(int64_t));
    or_sur_forward(&c, w, n, pos, vel, rho, off, idx, out, S, cnt);
    uint8_t* batch = malloc(n);
    float* target = malloc(3 * n * sizeof(float));
    for (size_t i = 0; i < n; ++i) batch[i] = (uint8_t)(i % 3 == 0);
    for (size_t i = 0; i < 3 * n; ++i) target[i] = uniform01() - 0.5f;
    const double loss = or_sur_loss_grad(&c, w, n, pos, vel, rho, off, idx, batch, target, gd);
    CHECK(isfinite(loss), "surrogate loss");
    for (size_t k = 0; k < np; ++k) g[k] = (float)gd[k];
    or_sur_adam(np, w, m, v, g, 1e-3f, 0.9f, 0.999f, 1e-8f, 1);
    free(pos); free(vel); free(rho); free(off); free(idx); free(w); free(m); free(v); free(g);
    free(gd); free(out); free(S); free(cnt); free(batch); free(target);
    return 0;
}

int main(void) {
    int32_t coords[3];
    uint64_t keys[4];
    const float special[12] = {0.0f, -1e-10f, 0.04f, NAN, 1.0f, -3.0f, INFINITY, 0.0f, 2.0f,
                               1e30f, -1e30f, 0.5f};
    int32_t cc[12];
    or_cell_keys(4, special, 0.04f, cc, keys);
    or_grid_coords(25.0f, special, coords);
    if (run_scene(0, 2000) || run_scene(1, 2000) || run_surrogate()) return 1;
    printf("sanitize ok\n");
    return 0;
}
