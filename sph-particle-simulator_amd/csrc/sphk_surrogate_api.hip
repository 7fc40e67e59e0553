// sphk_surrogate_api.hip -- C ABI of the learned macro-step surrogate: network state,
// forward / backward / Adam orchestration, the training-data generator and the main run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sphk_engine.h"
#include "sphk_surrogate.h"

namespace sphk {

struct SurrogateState {
    sphk_surrogate_config cfg{};
    float* params = nullptr;  // FP32 master copy
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;
    uint32_t adam_t = 0;
    float* tsave = nullptr;  // clock saved across a judge run
    __bf16* pk = nullptr;  // packed bf16 operands
    float* pf = nullptr;   // bf16-rounded FP32 copies
    bool packed = false;
    // activation rows
    uint64_t rows_cap = 0;
    __bf16 *S = nullptr, *X = nullptr, *Z = nullptr;
    float* OUT = nullptr;
    float* cnt = nullptr;
    uint32_t* rowslot = nullptr;
    // backward rows (batch)
    uint64_t brows_cap = 0;
    __bf16 *DOUT = nullptr, *DZ = nullptr, *DM = nullptr, *DS = nullptr;
    // weight-gradient partials and reductions
    uint64_t part_cap = 0;
    float* part = nullptr;
    uint64_t red_cap = 0;
    float* red = nullptr;
    std::vector<float> h_red;
    // per-particle buffers (caller order)
    uint64_t ncap = 0;
    float4 *target = nullptr, *x0 = nullptr, *v0 = nullptr, *x1 = nullptr, *v1 = nullptr;
    float *rho_a = nullptr, *rho_b = nullptr;
    // batch
    uint64_t batch_cap = 0;
    uint32_t* d_batch = nullptr;
    std::vector<uint32_t> h_batch;
};

void surrogate_free(SurrogateState* s) {
    if (!s) return;
    void* ptrs[] = {s->params, s->adam_m, s->adam_v, s->grad, s->tsave, s->pk, s->pf, s->S, s->X, s->Z,
                    s->OUT, s->cnt, s->rowslot, s->DOUT, s->DZ, s->DM, s->DS, s->part, s->red,
                    s->target, s->x0, s->v0, s->x1, s->v1, s->rho_a, s->rho_b, s->d_batch};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete s;
}

namespace {

constexpr uint32_t kRowsPerSlab = 1024;

uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// U[0,1) from (seed, counter), the scene generators' counter RNG.
double uniform01(uint64_t seed, uint64_t counter) {
    const uint64_t u = splitmix64(counter ^ splitmix64(seed));
    return (double)(u >> 11) * (1.0 / 9007199254740992.0);
}

void default_config(sphk_surrogate_config* c) {
    c->substeps = 4;
    c->judge_every = 16;
    c->tolerance = 10.0f;
    c->inv_v = 1.0f;
    c->rho_scale = 0.01f;
    c->msg_scale = 1.0f / 64.0f;
    c->acc_scale = 10.0f;
}

int check_config(const sphk_surrogate_config* c) {
    if (c->substeps < 1 || c->judge_every < 1 || !(c->tolerance >= 0.0f) ||
        !std::isfinite(c->inv_v) || !std::isfinite(c->rho_scale) ||
        !(c->msg_scale > 0.0f) || !(c->acc_scale > 0.0f))
        return set_error(SPHK_E_INVALID, "invalid surrogate configuration");
    return SPHK_OK;
}

SurScales scales_of(const sphk_engine* e, const SurrogateState* s) {
    SurScales sc;
    sc.inv_r = 1.0f / e->params.neighbor_search_radius;
    sc.inv_v = s->cfg.inv_v;
    sc.inv_h = 1.0f / e->params.smoothing_length;
    sc.w_norm = 1.5f;  // 1 / W(0) without sigma = 1 / (2/3)
    sc.rho0 = e->params.rest_density;
    sc.rho_scale = s->cfg.rho_scale;
    sc.msg_scale = s->cfg.msg_scale;
    sc.acc_scale = s->cfg.acc_scale;
    return sc;
}

template <class T>
int grow_buf(T** p, uint64_t& cap, uint64_t want) {
    if (want <= cap && *p) return SPHK_OK;
    if (*p) HIP_TRY(hipFree(*p));
    *p = nullptr;
    HIP_TRY(dalloc(p, want));
    cap = want;
    return SPHK_OK;
}

int get_state(sphk_engine* e, SurrogateState** out) {
    if (int rc = check_engine(e)) return rc;
    if (!e->sur) return set_error(SPHK_E_INVALID, "surrogate not initialised (sphk_surrogate_init)");
    if (int rc = set_device(e)) return rc;
    *out = e->sur;
    return SPHK_OK;
}

// Activation rows for at least `rows` rows (multiple of 64), zero-filled when (re)allocated.
int ensure_rows(SurrogateState* s, uint64_t rows, hipStream_t st) {
    rows = (rows + 63) / 64 * 64;
    if (rows <= s->rows_cap && s->S) return SPHK_OK;
    for (__bf16** p : {&s->S, &s->X, &s->Z}) {
        if (*p) HIP_TRY(hipFree(*p));
        *p = nullptr;
        HIP_TRY(dalloc(p, rows * kSurRow));
        HIP_TRY(hipMemsetAsync(*p, 0, rows * kSurRow * sizeof(__bf16), st));
    }
    if (s->OUT) HIP_TRY(hipFree(s->OUT));
    if (s->cnt) HIP_TRY(hipFree(s->cnt));
    if (s->rowslot) HIP_TRY(hipFree(s->rowslot));
    s->OUT = nullptr;
    s->cnt = nullptr;
    s->rowslot = nullptr;
    HIP_TRY(dalloc(&s->OUT, rows * 4));
    HIP_TRY(dalloc(&s->cnt, rows));
    HIP_TRY(dalloc(&s->rowslot, rows));
    HIP_TRY(hipMemsetAsync(s->OUT, 0, rows * 4 * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(s->cnt, 0, rows * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(s->rowslot, 0xff, rows * sizeof(uint32_t), st));
    s->rows_cap = rows;
    return SPHK_OK;
}

int ensure_batch_rows(SurrogateState* s, uint64_t rows) {
    if (rows <= s->brows_cap && s->DOUT) return SPHK_OK;
    for (__bf16** p : {&s->DOUT, &s->DZ, &s->DM, &s->DS}) {
        if (*p) HIP_TRY(hipFree(*p));
        *p = nullptr;
    }
    HIP_TRY(dalloc(&s->DOUT, rows * kSurOutRow));
    HIP_TRY(dalloc(&s->DZ, rows * 128));
    HIP_TRY(dalloc(&s->DM, rows * 128));
    HIP_TRY(dalloc(&s->DS, rows * 128));
    s->brows_cap = rows;
    return SPHK_OK;
}

int ensure_particles(SurrogateState* s, uint64_t n) {
    n = std::max<uint64_t>(n, 1);
    if (n <= s->ncap && s->target) return SPHK_OK;
    for (float4** p : {&s->target, &s->x0, &s->v0, &s->x1, &s->v1}) {
        if (*p) HIP_TRY(hipFree(*p));
        *p = nullptr;
        HIP_TRY(dalloc(p, n));
    }
    for (float** p : {&s->rho_a, &s->rho_b}) {
        if (*p) HIP_TRY(hipFree(*p));
        *p = nullptr;
        HIP_TRY(dalloc(p, n));
    }
    s->ncap = n;
    return SPHK_OK;
}

int ensure_red(SurrogateState* s, uint64_t count) {
    if (int rc = grow_buf(&s->red, s->red_cap, count)) return rc;
    if (s->h_red.size() < count) s->h_red.resize(count);
    return SPHK_OK;
}

int pack(sphk_engine* e, SurrogateState* s) {
    if (s->packed) return SPHK_OK;
    HIP_TRY(launch_sur_pack(s->params, s->pk, s->pf, e->stream));
    s->packed = true;
    return SPHK_OK;
}

// Sorted state + density pass (rho_i, neighbour words) of the current positions.
int prepare_density(sphk_engine* e) {
    if (int rc = prepare_grid(e)) return rc;
    if (int rc = build_sorted(e)) return rc;
    HIP_TRY(launch_density_walk(e->walk, e->spm, e->sq, e->svr, e->frows, e->skey, e->cell_start,
                                (uint32_t)e->n, e->consts, e->d_scenes, e->shape_density,
                                e->stream));
    return SPHK_OK;
}

SurRows rows_of(SurrogateState* s) {
    SurRows r;
    r.S = s->S;
    r.X = s->X;
    r.Z = s->Z;
    r.OUT = reinterpret_cast<float4*>(s->OUT);
    r.cnt = s->cnt;
    r.rowslot = s->rowslot;
    return r;
}

// Network forward for every particle of the prepared state: OUT rows by sorted slot.
int forward_all(sphk_engine* e, SurrogateState* s) {
    const uint32_t n = (uint32_t)e->n;
    hipStream_t st = e->stream;
    if (int rc = ensure_rows(s, n, st)) return rc;
    if (int rc = pack(e, s)) return rc;
    HIP_TRY(launch_sur_forward(false, e->walk, nullptr, 0, max_chunks(n, e->consts), e->spm,
                               e->svr, e->skey, e->sid, e->cell_start, n, e->consts,
                               e->d_scenes, s->pk, s->pf, scales_of(e, s), rows_of(s),
                               SurStep{}, st));
    return SPHK_OK;
}

// Picks min(batch_chunks, chunks) distinct chunk items (partial Fisher-Yates).
int choose_batch(sphk_engine* e, SurrogateState* s, uint64_t seed, uint32_t batch_chunks,
                 uint32_t* nbatch) {
    uint32_t nitems = 0;
    HIP_TRY(hipMemcpyAsync(&nitems, e->walk.nitems, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const uint32_t nb = std::min(batch_chunks, nitems);
    std::vector<uint32_t> all(nitems);
    for (uint32_t i = 0; i < nitems; ++i) all[i] = i;
    for (uint32_t i = 0; i < nb; ++i) {
        const uint64_t r = (uint64_t)(uniform01(seed, i) * (double)(nitems - i));
        std::swap(all[i], all[i + std::min<uint64_t>(r, nitems - i - 1)]);
    }
    s->h_batch.assign(all.begin(), all.begin() + nb);
    std::sort(s->h_batch.begin(), s->h_batch.end());
    if (int rc = grow_buf(&s->d_batch, s->batch_cap, std::max<uint32_t>(nb, 1))) return rc;
    if (nb) HIP_TRY(hipMemcpy(s->d_batch, s->h_batch.data(), nb * 4, hipMemcpyHostToDevice));
    *nbatch = nb;
    return SPHK_OK;
}

// Forward + loss + backward on the chosen batch of the prepared state; s->grad holds the
// gradient of the mean loss afterwards.  Returns the loss through *loss.
int loss_grad_batch(sphk_engine* e, SurrogateState* s, uint32_t nb, float* loss) {
    hipStream_t st = e->stream;
    const uint32_t R = 64u * nb;
    if (int rc = ensure_rows(s, R, st)) return rc;
    if (int rc = ensure_batch_rows(s, std::max<uint32_t>(R, 64))) return rc;
    if (int rc = pack(e, s)) return rc;
    const SurScales sc = scales_of(e, s);
    if (nb == 0) {
        HIP_TRY(hipMemsetAsync(s->grad, 0, kSurParams * sizeof(float), st));
        *loss = 0.0f;
        return SPHK_OK;
    }
    HIP_TRY(launch_sur_forward(false, e->walk, s->d_batch, nb, 0, e->spm, e->svr, e->skey,
                               e->sid, e->cell_start, (uint32_t)e->n, e->consts, e->d_scenes,
                               s->pk, s->pf, sc, rows_of(s), SurStep{}, st));
    // loss and its row gradients
    const uint32_t blocks = (R + 255u) / 256u;
    if (int rc = ensure_red(s, 2ull * blocks)) return rc;
    uint32_t nparts = 0;
    HIP_TRY(launch_sur_loss(s->OUT, s->rowslot, e->sid, s->target, R, s->DOUT, s->red, &nparts,
                            st));
    HIP_TRY(hipMemcpyAsync(s->h_red.data(), s->red, 2ull * nparts * sizeof(float),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double sq = 0.0, rows = 0.0;
    for (uint32_t b = 0; b < nparts; ++b) {
        sq += s->h_red[2 * b];
        rows += s->h_red[2 * b + 1];
    }
    if (rows <= 0.0) {
        HIP_TRY(hipMemsetAsync(s->grad, 0, kSurParams * sizeof(float), st));
        *loss = 0.0f;
        return SPHK_OK;
    }
    *loss = (float)(sq / (3.0 * rows));
    const float gscale = (float)(1.0 / (3.0 * rows));
    // backward through the node MLP and the hoisted message layer
    SurEpiArgs ep{};
    ep.out = s->DZ;
    ep.z = s->Z;
    HIP_TRY(launch_sur_gemm(kEpiDZ, s->DOUT, kSurOutRow, s->pk + kPkWn2T, 32, 32, 128, R, ep, st));
    ep = SurEpiArgs{};
    ep.out = s->DM;
    HIP_TRY(launch_sur_gemm(kEpiDM, s->DZ, 128, s->pk + kPkWn1T, 128, 128, 128, R, ep, st));
    ep = SurEpiArgs{};
    ep.out = s->DS;
    ep.scale = s->cfg.msg_scale;
    HIP_TRY(launch_sur_gemm(kEpiDS, s->DM, 128, s->pk + kPkW2T, 128, 128, 128, R, ep, st));
    // weight gradients (partials per slab of rows, then a fixed-order reduction)
    const uint32_t slabs = (R + kRowsPerSlab - 1) / kRowsPerSlab;
    const uint64_t need =
        std::max<uint64_t>((uint64_t)slabs * 128 * kSurRow, 2ull * nb * 128 * 16);
    if (int rc = grow_buf(&s->part, s->part_cap, need)) return rc;
    HIP_TRY(launch_sur_wgrad(s->DOUT, kSurOutRow, 3, s->Z, kSurRow, 129, R, kRowsPerSlab, slabs,
                             s->part, st));
    HIP_TRY(launch_sur_reduce(s->part, slabs, 3, 129, 3, 128, 128, kOffWn2, kOffBn2, gscale,
                              s->grad, st));
    HIP_TRY(launch_sur_wgrad(s->DZ, 128, 128, s->X, kSurRow, kSurNodeIn + 1, R, kRowsPerSlab,
                             slabs, s->part, st));
    HIP_TRY(launch_sur_reduce(s->part, slabs, 128, kSurNodeIn + 1, 128, kSurNodeIn, kSurNodeIn,
                              kOffWn1, kOffBn1, gscale, s->grad, st));
    HIP_TRY(launch_sur_wgrad(s->DM, 128, 128, s->S, kSurRow, 129, R, kRowsPerSlab, slabs, s->part,
                             st));
    HIP_TRY(launch_sur_reduce(s->part, slabs, 128, 129, 128, 128, 128, kOffW2, kOffB2, gscale,
                              s->grad, st));
    HIP_TRY(launch_sur_edges_bwd(e->walk, s->d_batch, nb, e->spm, e->svr, e->skey, e->cell_start,
                                 e->consts, e->d_scenes, s->pk, sc, s->DS, s->part, st));
    HIP_TRY(launch_sur_reduce(s->part, 2u * nb, 128, 16, 128, kSurNF, kSurNF, kOffW1, kOffB1,
                              gscale, s->grad, st));
    return SPHK_OK;
}

int adam_step(sphk_engine* e, SurrogateState* s, float lr) {
    s->adam_t++;
    const double b1 = 0.9, b2 = 0.999;
    const float bc1 = (float)(1.0 - std::pow(b1, (double)s->adam_t));
    const float bc2 = (float)(1.0 - std::pow(b2, (double)s->adam_t));
    HIP_TRY(launch_sur_adam(s->params, s->adam_m, s->adam_v, s->grad, kSurParams, lr, 0.9f,
                            0.999f, 1e-8f, bc1, bc2, e->stream));
    s->packed = false;
    return pack(e, s);
}

// Copies the caller-order state (pm, vel) between the engine and a save slot.
int save_state(sphk_engine* e, float4* x, float4* v) {
    HIP_TRY(hipMemcpyAsync(x, e->pm, e->n * sizeof(float4), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(v, e->vel, e->n * sizeof(float4), hipMemcpyDeviceToDevice, e->stream));
    return SPHK_OK;
}

int load_state(sphk_engine* e, const float4* x, const float4* v) {
    HIP_TRY(hipMemcpyAsync(e->pm, x, e->n * sizeof(float4), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->vel, v, e->n * sizeof(float4), hipMemcpyDeviceToDevice, e->stream));
    // a saved state need not lie inside the walls (an uploaded one, before any clamp): the
    // next grid is sized from the measured extent
    e->sorted_valid = false;
    e->ext_unknown = true;
    return SPHK_OK;
}

// One plain step at fixed dt with the engine's bookkeeping.
int plain_step(sphk_engine* e, float dt) {
    if (int rc = prepare_grid(e)) return rc;
    if (int rc = enqueue_step(e, dt, false, nullptr)) return rc;
    e->step_count++;
    e->stats.total_neighbor_queries += e->n;
    e->stats.steps++;
    after_positions_moved(e);
    return SPHK_OK;
}

int plain_macro(sphk_engine* e, SurrogateState* s, float dt) {
    for (int k = 0; k < s->cfg.substeps; ++k)
        if (int rc = plain_step(e, dt)) return rc;
    return SPHK_OK;
}

float damping_power(const sphk_engine* e, const SurrogateState* s) {
    float dk = 1.0f;
    for (int k = 0; k < s->cfg.substeps; ++k) dk *= e->params.damping;
    return dk;
}

// Targets of the current state into s->target; the engine advances one plain macro-step.
int make_targets(sphk_engine* e, SurrogateState* s, float dt) {
    if (int rc = ensure_particles(s, e->n)) return rc;
    if (int rc = save_state(e, s->x0, s->v0)) return rc;
    if (int rc = plain_macro(e, s, dt)) return rc;
    const float T = dt * (float)s->cfg.substeps;
    HIP_TRY(launch_sur_targets(s->v0, e->vel, (uint32_t)e->n, damping_power(e, s),
                               1.0f / (T * s->cfg.acc_scale), s->target, e->stream));
    return SPHK_OK;
}

// One network macro-step of the current state (forward and integration in one pass).
int macro_step(sphk_engine* e, SurrogateState* s, float dt) {
    if (int rc = prepare_density(e)) return rc;
    if (int rc = pack(e, s)) return rc;
    const uint32_t n = (uint32_t)e->n;
    SurStep step;
    step.T = dt * (float)s->cfg.substeps;
    step.dk = damping_power(e, s);
    step.pm = e->pm;
    step.vel = e->vel;
    step.acc = e->acc;
    step.rho = e->rho;
    step.prs = e->prs;
    SurRows none{};
    HIP_TRY(launch_sur_forward(true, e->walk, nullptr, 0, max_chunks(n, e->consts), e->spm,
                               e->svr, e->skey, e->sid, e->cell_start, n, e->consts,
                               e->d_scenes, s->pk, s->pf, scales_of(e, s), none, step,
                               e->stream));
    HIP_TRY(launch_advance_time(e->fscal + 1, nullptr, dt, e->stream, s->cfg.substeps));
    e->step_count += (uint64_t)s->cfg.substeps;
    after_positions_moved(e);
    return SPHK_OK;
}

// Densities of the current state in caller order into dst.
int densities_into(sphk_engine* e, float* dst) {
    if (int rc = prepare_density(e)) return rc;
    HIP_TRY(launch_sur_rho_home(e->spm, e->svr, e->sid, e->cell_start, (uint32_t)e->n, e->consts,
                                dst, e->stream));
    return SPHK_OK;
}

// mean and max of |a - b| over the n particles
int rho_error(sphk_engine* e, SurrogateState* s, const float* a, const float* b, double* mean,
              double* mx) {
    const uint32_t n = (uint32_t)e->n;
    const uint32_t blocks = (n + 255u) / 256u;
    if (int rc = ensure_red(s, 2ull * std::max<uint32_t>(blocks, 1))) return rc;
    uint32_t nparts = 0;
    HIP_TRY(launch_sur_rho_diff(a, b, n, s->red, &nparts, e->stream));
    HIP_TRY(hipMemcpyAsync(s->h_red.data(), s->red, 2ull * nparts * sizeof(float),
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    double sum = 0.0, m = 0.0;
    for (uint32_t i = 0; i < nparts; ++i) {
        sum += s->h_red[2 * i];
        m = std::max(m, (double)s->h_red[2 * i + 1]);
    }
    *mean = n ? sum / n : 0.0;
    *mx = m;
    return SPHK_OK;
}

int check_dt(float dt) {
    if (!(dt > 0.0f) || !std::isfinite(dt)) return set_error(SPHK_E_INVALID, "dt must be > 0");
    return SPHK_OK;
}

}  // namespace

}  // namespace sphk

using namespace sphk;

extern "C" {

int sphk_surrogate_default_config(sphk_surrogate_config* out) {
    if (!out) return set_error(SPHK_E_INVALID, "null config");
    default_config(out);
    return SPHK_OK;
}

int sphk_surrogate_param_count(size_t* out) {
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    *out = kSurParams;
    return SPHK_OK;
}

int sphk_surrogate_init(sphk_engine* e, const sphk_surrogate_config* cfg, uint64_t seed) {
    if (int rc = check_engine(e)) return rc;
    sphk_surrogate_config c;
    default_config(&c);
    if (cfg) c = *cfg;
    if (int rc = check_config(&c)) return rc;
    if (int rc = set_device(e)) return rc;
    if (!e->sur) {
        SurrogateState* s = new SurrogateState();
        e->sur = s;
        HIP_TRY(dalloc(&s->params, kSurParams));
        HIP_TRY(dalloc(&s->adam_m, kSurParams));
        HIP_TRY(dalloc(&s->adam_v, kSurParams));
        HIP_TRY(dalloc(&s->grad, kSurParams));
        HIP_TRY(dalloc(&s->tsave, 1));
        HIP_TRY(dalloc(&s->pk, kPkTotal));
        HIP_TRY(dalloc(&s->pf, kFTotal));
    }
    SurrogateState* s = e->sur;
    s->cfg = c;
    // Xavier-uniform weights, U(-a, a) with a = sqrt(6 / (fan_in + fan_out)); zero biases
    std::vector<float> p(kSurParams, 0.0f);
    struct Layer {
        uint32_t off, fan_out, fan_in;
    } layers[] = {{kOffW1, 128, kSurNF}, {kOffW2, 128, 128}, {kOffWn1, 128, kSurNodeIn},
                  {kOffWn2, 3, 128}};
    for (const Layer& L : layers) {
        const double a = std::sqrt(6.0 / (double)(L.fan_in + L.fan_out));
        for (uint32_t k = 0; k < L.fan_out * L.fan_in; ++k)
            p[L.off + k] = (float)((2.0 * uniform01(seed, L.off + k) - 1.0) * a);
    }
    HIP_TRY(hipMemcpy(s->params, p.data(), kSurParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(s->adam_m, 0, kSurParams * sizeof(float)));
    HIP_TRY(hipMemset(s->adam_v, 0, kSurParams * sizeof(float)));
    HIP_TRY(hipMemset(s->grad, 0, kSurParams * sizeof(float)));
    s->adam_t = 0;
    s->packed = false;
    return pack(e, s);
}

int sphk_surrogate_get_config(sphk_engine* e, sphk_surrogate_config* out) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    *out = s->cfg;
    return SPHK_OK;
}

int sphk_surrogate_get_params(sphk_engine* e, float* out, size_t count) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (!out || count < kSurParams) return set_error(SPHK_E_INVALID, "need %u floats", kSurParams);
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(out, s->params, kSurParams * sizeof(float), hipMemcpyDeviceToHost));
    return SPHK_OK;
}

int sphk_surrogate_set_params(sphk_engine* e, const float* params, size_t count) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (!params || count != kSurParams)
        return set_error(SPHK_E_INVALID, "expected %u parameters", kSurParams);
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(s->params, params, kSurParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(s->adam_m, 0, kSurParams * sizeof(float)));
    HIP_TRY(hipMemset(s->adam_v, 0, kSurParams * sizeof(float)));
    s->adam_t = 0;
    s->packed = false;
    return pack(e, s);
}

int sphk_surrogate_predict(sphk_engine* e, float* out, size_t count) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (!out || count < 3 * e->n) return set_error(SPHK_E_INVALID, "need 3 n floats");
    if (e->n == 0) return SPHK_OK;
    if (int rc = ensure_particles(s, e->n)) return rc;
    if (int rc = prepare_density(e)) return rc;
    if (int rc = forward_all(e, s)) return rc;
    HIP_TRY(launch_sur_scatter_out(s->OUT, e->sid, e->cell_start, (uint32_t)e->n, e->consts,
                                   s->cfg.acc_scale, s->x1, e->stream));
    std::vector<float4> h(e->n);
    HIP_TRY(hipMemcpyAsync(h.data(), s->x1, e->n * sizeof(float4), hipMemcpyDeviceToHost,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < e->n; ++i) {
        out[3 * i] = h[i].x;
        out[3 * i + 1] = h[i].y;
        out[3 * i + 2] = h[i].z;
    }
    return SPHK_OK;
}

int sphk_surrogate_targets(sphk_engine* e, float dt, float* out, size_t count) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (int rc = check_dt(dt)) return rc;
    if (!out || count < 3 * e->n) return set_error(SPHK_E_INVALID, "need 3 n floats");
    if (e->n == 0) return SPHK_OK;
    if (int rc = make_targets(e, s, dt)) return rc;
    std::vector<float4> h(e->n);
    HIP_TRY(hipMemcpyAsync(h.data(), s->target, e->n * sizeof(float4), hipMemcpyDeviceToHost,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const float as = s->cfg.acc_scale;
    for (size_t i = 0; i < e->n; ++i) {
        out[3 * i] = h[i].x * as;
        out[3 * i + 1] = h[i].y * as;
        out[3 * i + 2] = h[i].z * as;
    }
    return SPHK_OK;
}

int sphk_surrogate_loss_grad(sphk_engine* e, const float* target, size_t count,
                             uint64_t batch_seed, uint32_t batch_chunks, float* loss,
                             float* grad, size_t grad_count, uint8_t* in_batch) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (!target || count < 3 * e->n || !loss)
        return set_error(SPHK_E_INVALID, "need 3 n target floats and a loss output");
    if (grad && grad_count < kSurParams)
        return set_error(SPHK_E_INVALID, "need %u gradient floats", kSurParams);
    if (e->n == 0) return set_error(SPHK_E_INVALID, "no particles");
    if (int rc = ensure_particles(s, e->n)) return rc;
    std::vector<float4> h(e->n);
    const float inv = 1.0f / s->cfg.acc_scale;
    for (size_t i = 0; i < e->n; ++i)
        h[i] = make_float4(target[3 * i] * inv, target[3 * i + 1] * inv, target[3 * i + 2] * inv,
                           0.0f);
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(s->target, h.data(), e->n * sizeof(float4), hipMemcpyHostToDevice));
    if (int rc = prepare_density(e)) return rc;
    uint32_t nb = 0;
    if (int rc = choose_batch(e, s, batch_seed, batch_chunks, &nb)) return rc;
    if (int rc = loss_grad_batch(e, s, nb, loss)) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (grad) HIP_TRY(hipMemcpy(grad, s->grad, kSurParams * sizeof(float), hipMemcpyDeviceToHost));
    if (in_batch) {
        memset(in_batch, 0, e->n);
        const uint64_t R = 64ull * nb;
        std::vector<uint32_t> slots(R), sid(e->n);
        if (R) HIP_TRY(hipMemcpy(slots.data(), s->rowslot, R * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(sid.data(), e->sid, e->n * 4, hipMemcpyDeviceToHost));
        for (uint64_t r = 0; r < R; ++r)
            if (slots[r] != kSurNoSlot) in_batch[sid[slots[r]]] = 1;
    }
    return SPHK_OK;
}

int sphk_surrogate_adam(sphk_engine* e, const float* grad, size_t count, float lr) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (!grad || count != kSurParams)
        return set_error(SPHK_E_INVALID, "expected %u gradient values", kSurParams);
    if (!(lr >= 0.0f)) return set_error(SPHK_E_INVALID, "lr must be >= 0");
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(s->grad, grad, kSurParams * sizeof(float), hipMemcpyHostToDevice));
    if (int rc = adam_step(e, s, lr)) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_surrogate_train(sphk_engine* e, const sphk_train_config* tc, float* loss_curve,
                         size_t curve_len) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (!tc) return set_error(SPHK_E_INVALID, "null training config");
    if (int rc = check_dt(tc->dt)) return rc;
    if (tc->batch_chunks == 0 || !(tc->lr >= 0.0f))
        return set_error(SPHK_E_INVALID, "invalid training config");
    if (e->n == 0) return set_error(SPHK_E_INVALID, "no particles");
    if (int rc = ensure_particles(s, e->n)) return rc;
    for (uint64_t snap = 0; snap < tc->snapshots; ++snap) {
        // plain macro-step from the snapshot state: targets, then keep the result aside
        if (int rc = make_targets(e, s, tc->dt)) return rc;
        if (int rc = save_state(e, s->x1, s->v1)) return rc;
        if (int rc = load_state(e, s->x0, s->v0)) return rc;
        if (int rc = prepare_density(e)) return rc;
        double lsum = 0.0;
        for (uint32_t it = 0; it < tc->iters_per_snapshot; ++it) {
            uint32_t nb = 0;
            const uint64_t bseed = splitmix64(tc->seed ^ splitmix64(snap * 1024u + it));
            if (int rc = choose_batch(e, s, bseed, tc->batch_chunks, &nb)) return rc;
            float l = 0.0f;
            if (int rc = loss_grad_batch(e, s, nb, &l)) return rc;
            if (int rc = adam_step(e, s, tc->lr)) return rc;
            lsum += l;
        }
        if (loss_curve && snap < curve_len)
            loss_curve[snap] = tc->iters_per_snapshot ? (float)(lsum / tc->iters_per_snapshot) : 0.0f;
        if (int rc = load_state(e, s->x1, s->v1)) return rc;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_surrogate_macro_step(sphk_engine* e, float dt) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (int rc = check_dt(dt)) return rc;
    if (e->n == 0) return SPHK_OK;
    return macro_step(e, s, dt);
}

int sphk_surrogate_run(sphk_engine* e, uint64_t macro_steps, float dt,
                       sphk_surrogate_report* rep) {
    SurrogateState* s;
    if (int rc = get_state(e, &s)) return rc;
    if (int rc = check_dt(dt)) return rc;
    sphk_surrogate_report r{};
    if (e->n == 0) {
        if (rep) *rep = r;
        return SPHK_OK;
    }
    if (int rc = ensure_particles(s, e->n)) return rc;
    const uint64_t J = (uint64_t)s->cfg.judge_every;
    // Error model of the run against a plain run.  A network macro-step leaves an acceleration
    // error; velocity errors add up step after step like a random walk and positions integrate
    // them, so the position error -- and with it the density error -- after m macro-steps is
    // about (4/3) m^(3/2) times that of one macro-step.  Errors persist after the network stops
    // (plain steps carry them along), so once any network step has run the error at the end of
    // the run is bounded by (4/3) e (N - 1)^(3/2), e the largest divergence per macro-step the
    // judge runs measured, N the run's macro-steps (the first one is a judge run).  Network
    // steps are allowed only while that bound stays within the tolerance.
    double emax = 0.0;
    const double span = macro_steps > 1 ? (double)(macro_steps - 1) : 0.0;
    const double growth = (4.0 / 3.0) * span * std::sqrt(span);
    for (uint64_t k = 1; k <= macro_steps; ++k) {
        if (r.fell_back) {
            if (int rc = plain_macro(e, s, dt)) return rc;
            r.plain_macro_steps++;
            continue;
        }
        if ((k - 1) % J != 0) {
            if (int rc = macro_step(e, s, dt)) return rc;
            r.macro_steps++;
            continue;
        }
        // judge: the network and plain SPH from the same state; the plain result is kept
        if (int rc = save_state(e, s->x0, s->v0)) return rc;
        const uint64_t steps0 = e->step_count;
        HIP_TRY(hipMemcpyAsync(s->tsave, e->fscal + 1, sizeof(float), hipMemcpyDeviceToDevice,
                               e->stream));
        if (int rc = macro_step(e, s, dt)) return rc;
        if (int rc = densities_into(e, s->rho_a)) return rc;
        // back to the saved state and clock for the plain run
        if (int rc = load_state(e, s->x0, s->v0)) return rc;
        HIP_TRY(hipMemcpyAsync(e->fscal + 1, s->tsave, sizeof(float), hipMemcpyDeviceToDevice,
                               e->stream));
        e->step_count = steps0;
        if (int rc = plain_macro(e, s, dt)) return rc;
        if (int rc = densities_into(e, s->rho_b)) return rc;
        double mean = 0.0, mx = 0.0;
        if (int rc = rho_error(e, s, s->rho_a, s->rho_b, &mean, &mx)) return rc;
        emax = std::max(emax, mean);
        r.judges++;
        r.plain_macro_steps++;
        r.last_error = (float)mean;
        r.max_error = std::max(r.max_error, (float)mean);
        r.max_abs_error = std::max(r.max_abs_error, (float)mx);
        if (!(emax * growth <= (double)s->cfg.tolerance)) {
            r.fell_back = 1;
            r.fallback_at = k;
        }
    }
    if (r.macro_steps > 0) r.error_bound = (float)(emax * growth);
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (rep) *rep = r;
    return SPHK_OK;
}

int sphk_surrogate_densities(sphk_engine* e, float* out, size_t count) {
    if (int rc = check_engine(e)) return rc;
    if (!out || count < e->n) return set_error(SPHK_E_INVALID, "need n floats");
    if (e->n == 0) return SPHK_OK;
    if (int rc = set_device(e)) return rc;
    float* tmp = nullptr;
    HIP_TRY(dalloc(&tmp, e->n));
    int rc = densities_into(e, tmp);
    if (rc == SPHK_OK) {
        hipError_t h1 = hipMemcpyAsync(out, tmp, e->n * sizeof(float), hipMemcpyDeviceToHost,
                                       e->stream);
        hipError_t h2 = hipStreamSynchronize(e->stream);
        if (h1 != hipSuccess || h2 != hipSuccess)
            rc = set_error(SPHK_E_HIP, "density read-back failed");
    }
    (void)hipFree(tmp);
    return rc;
}

}  // extern "C"
