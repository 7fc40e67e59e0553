// sphk_api.hip -- C ABI of libsphk: engine lifecycle, cell-grid sizing, step orchestration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sphk_engine.h"

using namespace sphk;

namespace sphk {

thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

const std::string& last_error_text() { return g_last_error; }
void set_last_error_text(const std::string& s) { g_last_error = s; }

}  // namespace sphk

namespace {

// Largest dense grid (cells incl. padding) an engine will build.  4 bytes per cell.
constexpr uint64_t kMaxCells = 1ull << 28;
// Cell coordinates beyond +-2^30 cannot be represented in the dense grid.
constexpr double kMaxCellCoord = 1073741824.0;

}  // namespace

namespace sphk {

void fill_default(sphk_params* p) {
    p->rest_density = 1000.0f;
    p->gas_constant = 2000.0f;
    p->viscosity = 0.001f;
    p->smoothing_length = 0.02f;
    p->particle_mass = 0.001f;
    p->timestep = 0.001f;
    p->gravity = -9.81f;
    p->damping = 0.99f;
    p->CFL_factor = 0.4f;
    p->xmin = -1.0f;
    p->xmax = 1.0f;
    p->ymin = -1.0f;
    p->ymax = 1.0f;
    p->zmin = -1.0f;
    p->zmax = 1.0f;
    p->neighbor_search_radius = 0.04f;
    p->formulation = 0;
}

// Host-side constants, in the reference's FP32 expression order.
StepConsts make_consts(const sphk_params& p) {
#pragma clang fp contract(off)
    StepConsts c{};
    const float pi_f = 3.14159265358979323846f;
    const float h = p.smoothing_length;
    c.inv_cell = 1.0f / p.neighbor_search_radius;  // SpatialHash::set_cell_size (src/spatial_hash.h:63-66)
    c.r2 = p.neighbor_search_radius * p.neighbor_search_radius;  // src/sph_engine.cpp:347
    c.r2_filter = c.r2;
    for (int k = 0; k < 8; ++k) c.r2_filter = std::nextafter(c.r2_filter, INFINITY);
    c.h = h;
    c.inv_h = 1.0f / h;
    if (p.formulation == 1)
        c.sigma = 3.0f / (2.0f * pi_f * h * h * h);
    else
        c.sigma = 1.0f / (pi_f * h * h * h);  // src/kernels.cpp:27
    float q0 = 0.0f;
    float w0 = (2.0f / 3.0f - q0 * q0) + ((0.5f * q0) * q0) * q0;
    c.w0 = c.sigma * w0;
    // the force walk sums dW/dq / 4 and half the Laplacian term; the factors 4 and 2 scale
    // exactly
    c.grad_scale = 4.0f * (c.sigma / h);
    float h_sq = h * h;
    c.lap_scale = 2.0f * ((c.sigma * p.viscosity) / h_sq);
    c.k_gas = p.gas_constant;
    c.rho0 = p.rest_density;
    c.grav = p.gravity;
    c.damping = p.damping;
    c.formulation = p.formulation;
    // the reference counts a pair only when d^2 <= r^2 (src/sph_engine.cpp:347,
    // src/spatial_hash.h:70-73); with a radius below the support 2h a pair beyond the radius
    // would still carry kernel weight, so the walks then test it
    c.cut = p.neighbor_search_radius < 2.0f * h ? 1 : 0;
    // Quantized candidate filter.  X = floor(x * qscale) per axis (FP64) differs from
    // (x_a - x_b) qscale by under 1 + 2^-12, so a pair that passes the FP32 test (true
    // distance below sqrt(r2) (1 + 3 * 2^-24)) has |dX| < bound and dX^2 + dY^2 + dZ^2 <= T.
    c.qscale = 256.0 * (double)c.inv_cell;
    const double bound =
        std::sqrt((double)c.r2) * (1.0 + 1e-6) * c.qscale + std::sqrt(3.0) * 1.001;
    c.qneg = -(int32_t)(std::floor(bound * bound) + 1.0);
    return c;
}

int floor_cell(float x, float inv, int64_t* out) {
    float s = x * inv;
    double f = std::floor((double)s);
    if (!(f >= -kMaxCellCoord && f < kMaxCellCoord)) return -1;
    *out = (int64_t)f;
    return 0;
}

void ext_add_box(SceneHost& sc, const float lo[3], const float hi[3]) {
    for (int d = 0; d < 3; ++d) {
        if (!sc.ext_any) {
            sc.ext_lo[d] = lo[d];
            sc.ext_hi[d] = hi[d];
        } else {
            sc.ext_lo[d] = std::min(sc.ext_lo[d], lo[d]);
            sc.ext_hi[d] = std::max(sc.ext_hi[d], hi[d]);
        }
    }
    sc.ext_any = true;
}

// Box of the wall bounds (finite positions lie inside it after every step's clamp).
int bounds_box(const float b[6], float lo[3], float hi[3]) {
    for (int d = 0; d < 3; ++d) {
        const float a = b[2 * d], c = b[2 * d + 1];
        if (!std::isfinite(a) || !std::isfinite(c)) return -1;
        lo[d] = std::min(a, c);
        hi[d] = std::max(a, c);
    }
    return 0;
}

void params_bounds(const sphk_params& p, float b[6]) {
    b[0] = p.xmin;
    b[1] = p.xmax;
    b[2] = p.ymin;
    b[3] = p.ymax;
    b[4] = p.zmin;
    b[5] = p.zmax;
}

// Dense grid covering every cell a finite position in the scene's extent can fall in,
// padded by two cells per side so the +-2 walk never leaves it.
int grid_for_ext(const SceneHost& sc, float radius, Grid* g) {
    const float inv = 1.0f / radius;
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (sc.ext_any) {
        for (int d = 0; d < 3; ++d) {
            if (floor_cell(sc.ext_lo[d], inv, &lo[d]) || floor_cell(sc.ext_hi[d], inv, &hi[d]))
                return set_error(SPHK_E_DOMAIN,
                                 "particle coordinates need cell indices beyond +-2^30");
        }
    }
    int64_t dims[3];
    uint64_t G = 1;
    for (int d = 0; d < 3; ++d) {
        dims[d] = hi[d] - lo[d] + 1 + 4;
        G *= (uint64_t)dims[d];
        if (G > kMaxCells)
            return set_error(SPHK_E_DOMAIN,
                             "positions span more than %llu cells of size %g; the dense cell "
                             "grid supports at most that many",
                             (unsigned long long)kMaxCells, (double)radius);
    }
    g->ox = (int32_t)(lo[0] - 2);
    g->oy = (int32_t)(lo[1] - 2);
    g->oz = (int32_t)(lo[2] - 2);
    g->nx = (int32_t)dims[0];
    g->ny = (int32_t)dims[1];
    g->nz = (int32_t)dims[2];
    g->G = G;
    return SPHK_OK;
}

__global__ __launch_bounds__(256) void k_bbox(const float4* __restrict__ pm, uint32_t begin,
                                              uint32_t end, uint32_t* __restrict__ out7) {
    // out7 = {min x,y,z, max x,y,z as order-preserving uints, any}
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool found = false;
    for (uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x; i < end;
         i += gridDim.x * blockDim.x) {
        float4 p = pm[i];
        if (__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z)) {
            lo[0] = fminf(lo[0], p.x);
            lo[1] = fminf(lo[1], p.y);
            lo[2] = fminf(lo[2], p.z);
            hi[0] = fmaxf(hi[0], p.x);
            hi[1] = fmaxf(hi[1], p.y);
            hi[2] = fmaxf(hi[2], p.z);
            found = true;
        }
    }
    auto ord = [](float f) {
        unsigned b = __float_as_uint(f);
        return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    };
    // wavefront reduction first: one set of atomics per wavefront, not per thread
    unsigned olo[3], ohi[3];
    for (int d = 0; d < 3; ++d) {
        olo[d] = ord(lo[d]);
        ohi[d] = ord(hi[d]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int d = 0; d < 3; ++d) {
            const unsigned a = __shfl_xor(olo[d], off), b = __shfl_xor(ohi[d], off);
            olo[d] = a < olo[d] ? a : olo[d];
            ohi[d] = b > ohi[d] ? b : ohi[d];
        }
    }
    // then across the workgroup's wavefronts in LDS: one set of atomics per workgroup (all
    // workgroups update the same seven words, so their number bounds the kernel's time)
    __shared__ unsigned red[4][7];
    const unsigned wave = threadIdx.x >> 6;
    const bool any = __ballot(found) != 0ull;
    if ((threadIdx.x & 63u) == 0u) {
        for (int d = 0; d < 3; ++d) {
            red[wave][d] = olo[d];
            red[wave][3 + d] = ohi[d];
        }
        red[wave][6] = any ? 1u : 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nw = blockDim.x >> 6;
        for (unsigned w = 1; w < nw; ++w) {
            for (int d = 0; d < 3; ++d) {
                red[0][d] = red[w][d] < red[0][d] ? red[w][d] : red[0][d];
                red[0][3 + d] = red[w][3 + d] > red[0][3 + d] ? red[w][3 + d] : red[0][3 + d];
            }
            red[0][6] |= red[w][6];
        }
        if (red[0][6]) {
            for (int d = 0; d < 3; ++d) {
                atomicMin(&out7[d], red[0][d]);
                atomicMax(&out7[3 + d], red[0][3 + d]);
            }
            atomicOr(&out7[6], 1u);
        }
    }
}

float unord(unsigned b) {
    unsigned r = (b & 0x80000000u) ? (b & 0x7fffffffu) : ~b;
    float f;
    memcpy(&f, &r, 4);
    return f;
}

// Device bounding box of each scene's finite positions (synchronous; used after a restore
// or a scene change, or when the wall bounds are too wide to bound the grid).
int device_bbox(sphk_engine* e) {
    const size_t S = e->scenes.size();
    uint32_t* buf = nullptr;
    HIP_TRY(dalloc(&buf, 8 * S));
    std::vector<uint32_t> init(8 * S, 0u);
    for (size_t k = 0; k < S; ++k)
        for (int d = 0; d < 3; ++d) init[8 * k + d] = 0xffffffffu;
    HIP_TRY(hipMemcpyAsync(buf, init.data(), init.size() * 4, hipMemcpyHostToDevice, e->stream));
    for (size_t k = 0; k < S; ++k) {
        const uint64_t b = e->scenes[k].begin, en = std::min<uint64_t>(e->scenes[k].end, e->n);
        if (en > b) {
            uint32_t blocks = (uint32_t)std::min<uint64_t>((en - b + 255) / 256, 256);
            hipLaunchKernelGGL(k_bbox, dim3(blocks), dim3(256), 0, e->stream, e->pm, (uint32_t)b,
                               (uint32_t)en, buf + 8 * k);
            HIP_TRY(hipGetLastError());
        }
    }
    std::vector<uint32_t> out(8 * S);
    HIP_TRY(hipMemcpyAsync(out.data(), buf, out.size() * 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipFree(buf));
    for (size_t k = 0; k < S; ++k) {
        SceneHost& sc = e->scenes[k];
        sc.ext_any = out[8 * k + 6] != 0;
        for (int d = 0; d < 3; ++d) {
            sc.ext_lo[d] = sc.ext_any ? unord(out[8 * k + d]) : 0.0f;
            sc.ext_hi[d] = sc.ext_any ? unord(out[8 * k + 3 + d]) : 0.0f;
        }
    }
    e->ext_unknown = false;
    return SPHK_OK;
}

template <class T>
int grow(T** p, uint64_t& cap, uint64_t want, uint64_t& gen) {
    if (want <= cap) return SPHK_OK;
    if (*p) HIP_TRY(hipFree(*p));
    *p = nullptr;
    uint64_t c = std::max(want, cap * 5 / 4);
    HIP_TRY(dalloc(p, (size_t)c));
    cap = c;
    ++gen;
    return SPHK_OK;
}

// Scene ranges: one implicit scene [0, n) with the parameter bounds, or the explicit ones.
int check_scenes(sphk_engine* e) {
    if (!e->scenes_explicit) {
        if (e->scenes.size() != 1) e->scenes.resize(1);
        e->scenes[0].begin = 0;
        e->scenes[0].end = e->n;
        params_bounds(e->params, e->scenes[0].bounds);
        return SPHK_OK;
    }
    if (e->scenes.empty() || e->scenes.back().end != e->n)
        return set_error(SPHK_E_INVALID, "scene ranges cover %llu particles, the engine holds %zu",
                         (unsigned long long)(e->scenes.empty() ? 0 : e->scenes.back().end), e->n);
    return SPHK_OK;
}

int ensure_grid(sphk_engine* e) {
    if (int rc = check_scenes(e)) return rc;
    if (e->ext_unknown) {
        if (int rc = device_bbox(e)) return rc;
    }
    const size_t S = e->scenes.size();
    uint64_t G = 0, ncols = 0;
    e->table.resize(S);
    for (size_t k = 0; k < S; ++k) {
        SceneHost& sc = e->scenes[k];
        if (int rc = grid_for_ext(sc, e->params.neighbor_search_radius, &sc.grid)) return rc;
        SceneGrid& t = e->table[k];
        t.ox = sc.grid.ox;
        t.oy = sc.grid.oy;
        t.oz = sc.grid.oz;
        t.nx = sc.grid.nx;
        t.ny = sc.grid.ny;
        t.nz = sc.grid.nz;
        t.cell_base = (uint32_t)G;
        t.col_base = (uint32_t)ncols;
        t.pbegin = (uint32_t)sc.begin;
        t.pend = (uint32_t)sc.end;
        for (int d = 0; d < 3; ++d) {
            t.lo[d] = sc.bounds[2 * d];
            t.hi[d] = sc.bounds[2 * d + 1];
        }
        G += sc.grid.G;
        ncols += (uint64_t)sc.grid.nx * (uint64_t)sc.grid.ny;
        if (G > kMaxCells)
            return set_error(SPHK_E_DOMAIN, "scene grids need %llu cells; at most %llu supported",
                             (unsigned long long)G, (unsigned long long)kMaxCells);
    }
    if (int rc = grow(&e->cell_start, e->cell_start_cap, G + 2, e->buffer_gen)) return rc;
    uint64_t scap = e->d_scenes_cap;
    if (int rc = grow(&e->d_scenes, scap, S, e->buffer_gen)) return rc;
    e->d_scenes_cap = scap;
    if (e->table_uploaded_gen != e->buffer_gen || e->table_uploaded.size() != S ||
        memcmp(e->table_uploaded.data(), e->table.data(), S * sizeof(SceneGrid)) != 0) {
        HIP_TRY(hipMemcpyAsync(e->d_scenes, e->table.data(), S * sizeof(SceneGrid),
                               hipMemcpyHostToDevice, e->stream));
        e->table_uploaded = e->table;
        e->table_uploaded_gen = e->buffer_gen;
    }
    e->consts = make_consts(e->params);
    e->consts.G = (uint32_t)G;
    e->consts.ncols = (uint32_t)ncols;
    e->consts.nscenes = (uint32_t)S;
    e->consts.pad_slot = (uint32_t)e->cap;
    // chunk work lists and neighbour words: at most ceil(C/64) + one partial chunk per column
    const uint64_t chunks = (e->cap + 63) / 64 + ncols;
    WalkBuffers& wb = e->walk;
    if (chunks > wb.chunk_cap) {
        // every member is cleared as it is freed, so a failed allocation below leaves no
        // freed pointer behind (the next call or sphk_destroy frees only live ones)
        for (uint32_t** p : {&wb.items, &wb.mask, &wb.wj0, &wb.nwords, &wb.pstarts}) {
            if (*p) HIP_TRY(hipFree(*p));
            *p = nullptr;
        }
        const size_t want = std::max<size_t>(chunks, wb.chunk_cap * 5 / 4);
        // kMaxWords words (64 KiB of bits) per chunk while the buffers stay within the budget
        const uint32_t mw =
            (uint64_t)want * kMaxWords * 64 * 4 <= kWordBufferBudget ? kMaxWords : kMinWords;
        wb.chunk_cap = 0;
        wb.max_words = 0;
        HIP_TRY(dalloc(&wb.items, want));
        HIP_TRY(dalloc(&wb.mask, want * mw * 64));
        HIP_TRY(dalloc(&wb.wj0, want * mw));
        HIP_TRY(dalloc(&wb.nwords, want));
        HIP_TRY(dalloc(&wb.pstarts, want));
        wb.chunk_cap = want;
        wb.max_words = mw;
        ++e->buffer_gen;
    }
    e->consts.max_words = wb.max_words;
    if (!e->shapes_explicit) {
        const bool small = e->n < kSplitBelow;
        e->shape_density = small ? kShapeSmall : kShapeLargeDensity;
        e->shape_force = small ? kShapeSmall : kShapeLargeForce;
    }
    // per-part sums of the split walk shapes: three sums per slot, only while selected
    if ((walk_shape_split(e->shape_density) || walk_shape_split(e->shape_force)) &&
        wb.part_cap < e->cap) {
        if (wb.part_acc) HIP_TRY(hipFree(wb.part_acc));
        wb.part_acc = nullptr;
        if (wb.part_cnt) HIP_TRY(hipFree(wb.part_cnt));
        wb.part_cnt = nullptr;
        if (wb.part_force) HIP_TRY(hipFree(wb.part_force));
        wb.part_force = nullptr;
        wb.part_cap = 0;
        HIP_TRY(dalloc(&wb.part_acc, kParts * e->cap));
        HIP_TRY(dalloc(&wb.part_cnt, kParts * e->cap));
        HIP_TRY(dalloc(&wb.part_force, 2 * kParts * e->cap));
        wb.part_cap = e->cap;
        ++e->buffer_gen;
    }
    uint64_t ccap = wb.col_cap;
    if (int rc = grow(&wb.colcnt, ccap, ncols + 1, e->buffer_gen)) return rc;
    wb.col_cap = ccap;
    if (!wb.nitems) {
        HIP_TRY(dalloc(&wb.nitems, 1));
        ++e->buffer_gen;
    }
    const size_t hist_words = ((size_t)1 << kRadixMaxBits) * ((e->cap + kRadixTile - 1) / kRadixTile);
    uint64_t sscap = e->scan_scratch_cap;
    if (int rc = grow(&e->scan_scratch, sscap,
                      scan_scratch_words(std::max<size_t>(hist_words, ncols + 1)), e->buffer_gen))
        return rc;
    e->scan_scratch_cap = sscap;
    uint64_t lcap = e->lookback_cap;
    if (int rc = grow(&e->lookback, lcap, lookback_state_words(hist_words), e->buffer_gen))
        return rc;
    if (lcap != e->lookback_cap) HIP_TRY(hipMemsetAsync(e->lookback, 0, lcap * 4, e->stream));
    e->lookback_cap = lcap;
    e->grid_valid = true;
    return SPHK_OK;
}

int total_key_bits(uint64_t G) {
    int b = 1;
    while ((1ull << b) < G + 1) ++b;
    return b;
}

// keys -> sort -> cell_start -> gather -> chunk list.  Leaves e->skey/e->sid/e->spm/e->svr
// valid.  With `tail` (a step's build), the chunk-list launch also steps the particles of the
// non-finite bin and advances the clock (the fields filled here are taken from the engine).
int build_sorted(sphk_engine* e, const StepTail* tail) {
    const uint32_t n = (uint32_t)e->n;
    hipStream_t st = e->stream;
    uint32_t* ko = nullptr;
    uint32_t* vo = nullptr;
    HIP_TRY(radix_sort_keys(e->pm, e->consts, e->d_scenes, e->uscal, e->keys, e->keys_alt, e->vals,
                            e->vals_alt, e->hist, e->lookback, e->scan_scratch, n,
                            total_key_bits(e->consts.G), &ko, &vo, st));
    e->skey = ko;
    e->sid = vo;
    HIP_TRY(launch_gather_cells(e->pm, e->vel, e->sid, e->skey, n, e->consts.G, e->consts.qscale,
                                e->sq, e->spm, e->svr, e->cell_start, st));
    StepTail t;
    if (tail) {
        t = *tail;
        t.spm = e->spm;
        t.svr = e->svr;
        t.sid = e->sid;
        t.n = n;
    }
    HIP_TRY(launch_chunks(e->cell_start, e->consts, e->d_scenes, e->walk.colcnt, e->scan_scratch,
                          e->walk.items, e->walk.nitems, tail ? &t : nullptr, st));
    e->sorted_valid = true;
    return SPHK_OK;
}

// A finite particle outside the dense grid is routed to the non-finite bin and
// flagged (bit 0); the grid sizing above makes this unreachable, so report it loudly.  Bit 1:
// a look-back of the sort's scan gave up waiting (its result is not to be trusted).
int check_grid_flag(sphk_engine* e) {
    uint32_t flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, e->uscal, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (flag & 2u) return set_error(SPHK_E_DOMAIN, "the cell sort's scan timed out waiting");
    if (flag) return set_error(SPHK_E_DOMAIN, "a finite particle fell outside the cell grid");
    return SPHK_OK;
}

// Folds the events of completed timed steps into the phase totals.
int drain_timing(sphk_engine* e) {
    for (auto& set : e->pending) {
        HIP_TRY(hipEventSynchronize(set.ev[3]));
        float t01, t12, t23, t03;
        HIP_TRY(hipEventElapsedTime(&t01, set.ev[0], set.ev[1]));
        HIP_TRY(hipEventElapsedTime(&t12, set.ev[1], set.ev[2]));
        HIP_TRY(hipEventElapsedTime(&t23, set.ev[2], set.ev[3]));
        HIP_TRY(hipEventElapsedTime(&t03, set.ev[0], set.ev[3]));
        e->stats.neighbor_search_time += t01 * 1e-3;
        e->stats.density_computation_time += t12 * 1e-3;
        // the force kernel also integrates and clamps (one fused pass)
        e->stats.force_computation_time += t23 * 1e-3;
        e->stats.total_time += t03 * 1e-3;
        e->stats.timed_steps++;
        e->free_sets.push_back(set);
    }
    e->pending.clear();
    return SPHK_OK;
}

int check_engine(const sphk_engine* e) {
    if (!e) return set_error(SPHK_E_INVALID, "null engine");
    return SPHK_OK;
}

int set_device(const sphk_engine* e) {
    HIP_TRY(hipSetDevice(e->device));
    return SPHK_OK;
}

}  // namespace sphk

namespace sphk {

int prepare_grid(sphk_engine* e) {
    if (!e->grid_valid || e->ext_unknown || !e->scenes_explicit) {
        if (int rc = check_scenes(e)) return rc;
    }
    if (!e->grid_valid || e->ext_unknown) {
        if (int rc = ensure_grid(e)) return rc;
    }
    return SPHK_OK;
}

void after_positions_moved(sphk_engine* e) {
    e->sorted_valid = false;
    // after the wall clamp every finite position of a scene lies in its bounds box
    bool changed = false;
    for (auto& sc : e->scenes) {
        float lo[3], hi[3];
        if (bounds_box(sc.bounds, lo, hi) != 0) {
            e->ext_unknown = true;
            continue;
        }
        SceneHost next = sc;
        next.ext_any = true;
        for (int d = 0; d < 3; ++d) {
            next.ext_lo[d] = lo[d];
            next.ext_hi[d] = hi[d];
        }
        std::string keep = g_last_error;
        Grid g;
        if (grid_for_ext(next, e->params.neighbor_search_radius, &g) != SPHK_OK) {
            g_last_error = keep;  // walls too wide to bound the grid: measure the live bbox
            e->ext_unknown = true;
            continue;
        }
        if (g.G != sc.grid.G || g.ox != sc.grid.ox || g.oy != sc.grid.oy || g.oz != sc.grid.oz ||
            g.nx != sc.grid.nx || g.ny != sc.grid.ny || g.nz != sc.grid.nz)
            changed = true;
        for (int d = 0; d < 3; ++d) {
            sc.ext_lo[d] = lo[d];
            sc.ext_hi[d] = hi[d];
        }
        sc.ext_any = true;
    }
    if (changed) e->grid_valid = false;
}

int enqueue_step(sphk_engine* e, float dt, bool adaptive, hipEvent_t* ev) {
    hipStream_t st = e->stream;
    const uint32_t n = (uint32_t)e->n;
    if (ev) HIP_TRY(hipEventRecord(ev[0], st));
    if (adaptive) HIP_TRY(launch_cfl(e->vel, e->acc, n, e->params, e->uscal + 1, e->fscal, st));
    StepTail tail{};
    tail.dt_dev = adaptive ? e->fscal : nullptr;
    tail.dt = dt;
    tail.pm = e->pm;
    tail.vel = e->vel;
    tail.acc = e->acc;
    tail.rho = e->rho;
    tail.prs = e->prs;
    tail.time_dev = e->fscal + 1;
    if (int rc = build_sorted(e, &tail)) return rc;
    if (ev) HIP_TRY(hipEventRecord(ev[1], st));
    HIP_TRY(launch_density_walk(e->walk, e->spm, e->sq, e->svr, e->frows, e->skey, e->cell_start,
                                n, e->consts, e->d_scenes, e->shape_density, st));
    if (ev) HIP_TRY(hipEventRecord(ev[2], st));
    HIP_TRY(launch_force_walk(e->walk, e->spm, e->sq, e->svr, e->frows, e->skey, e->sid,
                              e->cell_start, n,
                              e->consts, e->d_scenes, adaptive ? e->fscal : nullptr, dt, e->pm,
                              e->vel, e->acc, e->rho, e->prs, e->shape_force, e->nstats, st));
    if (ev) HIP_TRY(hipEventRecord(ev[3], st));
    return SPHK_OK;
}

// Captures the step into a graph (or reuses the one captured for the same key) and
// launches it.
int graph_step(sphk_engine* e, float dt, bool adaptive) {
    GraphKey key;
    memset(&key, 0, sizeof(key));
    key.consts = e->consts;
    key.n = (uint32_t)e->n;
    key.shape_density = e->shape_density;
    key.shape_force = e->shape_force;
    key.adaptive = adaptive;
    key.dt = adaptive ? 0.0f : dt;
    key.buffer_gen = e->buffer_gen;
    if (!e->graph_valid || memcmp(&key, &e->graph_key, sizeof(key)) != 0) {
        if (e->graph_exec) {
            HIP_TRY(hipGraphExecDestroy(e->graph_exec));
            e->graph_exec = nullptr;
        }
        e->graph_valid = false;
        HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_step(e, dt, adaptive, nullptr);
        hipGraph_t graph = nullptr;
        hipError_t ec = hipStreamEndCapture(e->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIP_TRY(ec);
        hipError_t ei = hipGraphInstantiate(&e->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ei);
        memcpy(&e->graph_key, &key, sizeof(key));
        e->graph_valid = true;
    }
    HIP_TRY(hipGraphLaunch(e->graph_exec, e->stream));
    return SPHK_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free kinetic energy (sphk_total_energy_noalloc).  The partial sums live in
// `scratch`, a file-scope device array: one FP64 sum per workgroup of the first pass and,
// after them, the FP32 result of the second.  Every engine on a GPU shares that GPU's copy;
// its address on a device is taken once (hipGetSymbolAddress, at the first engine created on
// the device) and host code uses only that address.  Each term is the reference's FP32
// (0.5 m) |v|^2 (src/sph_engine.cpp:193-200); the terms are summed in FP64 in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kEnergyBlocks = 1024;  // workgroups of the first pass, at most
constexpr int kEnergyThreads = 256;
static __device__ __align__(256) double scratch[kEnergyBlocks + 1];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// Address of this device's scratch (device `dev` current); taken once per device.
int scratch_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(SPHK_E_INVALID, "device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return SPHK_OK;
    void* ptr = nullptr;
    hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !ptr)
        return set_error(SPHK_E_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                         hipGetErrorString(err));
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
    return SPHK_OK;
}

double* scratch_of(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

// Fixed-order sum of a workgroup's FP64 values: each wavefront by xor-shuffles (lane order
// fixed), then the four wavefront sums in order.
__device__ __forceinline__ double block_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __shared__ double wsum[kEnergyThreads / 64];
    const int wave = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63u) == 0) wsum[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kEnergyThreads / 64; ++w) t += wsum[w];
    return t;
}

__global__ __launch_bounds__(kEnergyThreads) void k_energy_partial(const float4* __restrict__ pm,
                                                                  const float4* __restrict__ vel,
                                                                  uint32_t n,
                                                                  double* __restrict__ part) {
#pragma clang fp contract(off)
    double acc = 0.0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float4 v = vel[i];
        const float tx = v.x * v.x, ty = v.y * v.y, tz = v.z * v.z;
        const float sp = (tx + ty) + tz;
        acc += (double)((0.5f * pm[i].w) * sp);
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kEnergyThreads) void k_energy_final(double* __restrict__ part,
                                                                uint32_t nparts) {
    double acc = 0.0;
    for (uint32_t b = threadIdx.x; b < nparts; b += blockDim.x) acc += part[b];
    const double t = block_sum(acc);
    if (threadIdx.x == 0) reinterpret_cast<float*>(part + kEnergyBlocks)[0] = (float)t;
}

}  // namespace sphk

extern "C" {

int sphk_default_params(sphk_params* out) {
    if (!out) return set_error(SPHK_E_INVALID, "null params");
    fill_default(out);
    return SPHK_OK;
}

static int create_buffers(sphk_engine* e, size_t capacity);

int sphk_create(size_t capacity, sphk_engine** out) {
    if (!out) return set_error(SPHK_E_INVALID, "null output handle");
    *out = nullptr;
    if (capacity == 0) return set_error(SPHK_E_INVALID, "capacity must be positive");
    // the walks address neighbour rows (32 bytes, capacity + 1 of them) by 32-bit byte offsets
    if (capacity > kMaxCapacity)
        return set_error(SPHK_E_INVALID, "capacity above %llu", (unsigned long long)kMaxCapacity);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(SPHK_E_NODEVICE, "no HIP device visible");
    sphk_engine* e = new sphk_engine();
    if (int rc = create_buffers(e, capacity)) {
        sphk_destroy(e);  // frees whatever was allocated before the failure
        return rc;
    }
    *out = e;
    return SPHK_OK;
}

// Device set-up of a new engine; on failure the caller destroys the partial engine.
static int create_buffers(sphk_engine* e, size_t capacity) {
    e->cap = capacity;
    fill_default(&e->params);
    e->scenes.resize(1);
    HIP_TRY(hipGetDevice(&e->device));
    if (int rc = scratch_setup(e->device)) return rc;
    HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    const size_t C = capacity;
    HIP_TRY(dalloc(&e->pm, C));
    HIP_TRY(dalloc(&e->vel, C));
    HIP_TRY(dalloc(&e->acc, C));
    HIP_TRY(dalloc(&e->rho, C));
    HIP_TRY(dalloc(&e->prs, C));
    HIP_TRY(dalloc(&e->snap_pm, C));
    HIP_TRY(dalloc(&e->snap_vel, C));
    HIP_TRY(dalloc(&e->snap_acc, C));
    HIP_TRY(dalloc(&e->keys, C));
    HIP_TRY(dalloc(&e->keys_alt, C));
    HIP_TRY(dalloc(&e->vals, C));
    HIP_TRY(dalloc(&e->vals_alt, C));
    HIP_TRY(dalloc(&e->hist, ((size_t)1 << kRadixMaxBits) * ((C + kRadixTile - 1) / kRadixTile)));
    HIP_TRY(dalloc(&e->spm, C + 1));
    HIP_TRY(dalloc(&e->svr, C + 1));
    HIP_TRY(dalloc(&e->frows, C + 1));
    HIP_TRY(dalloc(&e->sq, C + 1));
    {
        // padding slot: zero mass and weights, 1e15 away from any particle
        const float4 far_row = make_float4(1e15f, 1e15f, 1e15f, 0.0f);
        const float4 zero_row = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        HIP_TRY(hipMemcpy(e->spm + C, &far_row, sizeof(float4), hipMemcpyHostToDevice));
        const ForceRow pad_row{far_row, zero_row};
        HIP_TRY(hipMemcpy(e->frows + C, &pad_row, sizeof(ForceRow), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->svr + C, &zero_row, sizeof(float4), hipMemcpyHostToDevice));
    }
    HIP_TRY(dalloc(&e->fscal, 4));
    HIP_TRY(dalloc(&e->uscal, 4));
    HIP_TRY(dalloc(&e->nstats, 2));
    HIP_TRY(hipMemsetAsync(e->acc, 0, C * sizeof(float4), e->stream));
    HIP_TRY(hipMemsetAsync(e->vel, 0, C * sizeof(float4), e->stream));
    HIP_TRY(hipMemsetAsync(e->rho, 0, C * sizeof(float), e->stream));
    HIP_TRY(hipMemsetAsync(e->prs, 0, C * sizeof(float), e->stream));
    HIP_TRY(hipMemsetAsync(e->fscal, 0, 4 * sizeof(float), e->stream));
    HIP_TRY(hipMemsetAsync(e->uscal, 0, 4 * sizeof(uint32_t), e->stream));
    HIP_TRY(hipMemsetAsync(e->nstats, 0, 2 * sizeof(unsigned long long), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_destroy(sphk_engine* e) {
    if (!e) return SPHK_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    void* ptrs[] = {e->pm,   e->vel,      e->acc,  e->rho,      e->prs,  e->snap_pm,
                    e->snap_vel, e->snap_acc, e->keys, e->keys_alt, e->vals, e->vals_alt,
                    e->hist, e->spm,      e->svr,  e->cell_start, e->fscal, e->uscal,
                    e->nstats, e->frows, e->sq};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (void* p : {(void*)e->walk.items, (void*)e->walk.nitems, (void*)e->walk.colcnt,
                    (void*)e->walk.mask, (void*)e->walk.wj0, (void*)e->walk.nwords,
                    (void*)e->walk.pstarts, (void*)e->walk.part_acc, (void*)e->walk.part_cnt,
                    (void*)e->walk.part_force, (void*)e->scan_scratch, (void*)e->d_scenes,
                    (void*)e->lookback})
        if (p) (void)hipFree(p);
    if (e->sur) surrogate_free(e->sur);
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    for (auto* v : {&e->pending, &e->free_sets})
        for (auto& set : *v)
            for (auto& ev : set.ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return SPHK_OK;
}

int sphk_set_params(sphk_engine* e, const sphk_params* p) {
    if (int rc = check_engine(e)) return rc;
    if (!p) return set_error(SPHK_E_INVALID, "null params");
    if (!(p->neighbor_search_radius > 0.0f) || !std::isfinite(p->neighbor_search_radius))
        return set_error(SPHK_E_INVALID, "neighbor_search_radius must be positive and finite");
    if (!(p->smoothing_length > 0.0f) || !std::isfinite(p->smoothing_length))
        return set_error(SPHK_E_INVALID, "smoothing_length must be positive and finite");
    if (p->formulation != 0 && p->formulation != 1)
        return set_error(SPHK_E_INVALID, "formulation must be 0 or 1");
    if (p->neighbor_search_radius != e->params.neighbor_search_radius) e->ext_unknown = true;
    e->params = *p;
    if (!e->scenes_explicit) {
        e->scenes.resize(1);
        params_bounds(e->params, e->scenes[0].bounds);
    }
    e->grid_valid = false;
    e->sorted_valid = false;
    return SPHK_OK;
}

int sphk_get_params(const sphk_engine* e, sphk_params* out) {
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null params");
    *out = e->params;
    return SPHK_OK;
}

int sphk_upload(sphk_engine* e, size_t offset, size_t n, const float* pos, const float* vel,
                const float* mass) {
    if (int rc = check_engine(e)) return rc;
    if (offset > e->n) return set_error(SPHK_E_INVALID, "offset %zu beyond count %zu", offset, e->n);
    if (n && !pos) return set_error(SPHK_E_INVALID, "null positions");
    if (int rc = set_device(e)) return rc;
    size_t fit = std::min(n, e->cap - offset);
    std::vector<float4> hp(fit), hv(fit);
    const float inv = 1.0f / e->params.neighbor_search_radius;
    for (size_t i = 0; i < fit; ++i) {
        float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        float m = mass ? mass[i] : e->params.particle_mass;
        hp[i] = make_float4(x, y, z, m);
        hv[i] = vel ? make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f)
                    : make_float4(0.f, 0.f, 0.f, 0.f);
        if (std::isfinite(x) && std::isfinite(y) && std::isfinite(z)) {
            int64_t cc;
            if (floor_cell(x, inv, &cc) || floor_cell(y, inv, &cc) || floor_cell(z, inv, &cc))
                return set_error(SPHK_E_DOMAIN,
                                 "particle %zu at (%g, %g, %g) needs a cell index beyond +-2^30",
                                 offset + i, (double)x, (double)y, (double)z);
        }
    }
    if (fit) {
        HIP_TRY(hipMemcpyAsync(e->pm + offset, hp.data(), fit * sizeof(float4),
                               hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->vel + offset, hv.data(), fit * sizeof(float4),
                               hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    e->n = offset + fit;
    if (!e->scenes_explicit) {
        // one implicit scene [0, n): grow its extent by the uploaded positions
        e->scenes.resize(1);
        SceneHost& sc = e->scenes[0];
        if (offset == 0) sc.ext_any = false;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        bool any = false;
        for (size_t i = 0; i < fit; ++i) {
            const float4& q = hp[i];
            if (std::isfinite(q.x) && std::isfinite(q.y) && std::isfinite(q.z)) {
                lo[0] = std::min(lo[0], q.x);
                lo[1] = std::min(lo[1], q.y);
                lo[2] = std::min(lo[2], q.z);
                hi[0] = std::max(hi[0], q.x);
                hi[1] = std::max(hi[1], q.y);
                hi[2] = std::max(hi[2], q.z);
                any = true;
            }
        }
        if (any) ext_add_box(sc, lo, hi);
        if (offset == 0) e->ext_unknown = false;
    } else {
        // explicit scenes: measure every scene's extent on the device before the next build
        e->ext_unknown = true;
    }
    e->grid_valid = false;
    e->sorted_valid = false;
    if (fit < n)
        return set_error(SPHK_E_CAPACITY, "particle capacity %zu exceeded: %zu particles dropped",
                         e->cap, n - fit);
    return SPHK_OK;
}

int sphk_set_scenes(sphk_engine* e, size_t nscenes, const uint64_t* offsets, const float* bounds) {
    if (int rc = check_engine(e)) return rc;
    if (nscenes == 0) {
        e->scenes_explicit = false;
        e->scenes.assign(1, SceneHost{});
        e->ext_unknown = true;
        e->grid_valid = false;
        e->sorted_valid = false;
        return SPHK_OK;
    }
    if (!offsets) return set_error(SPHK_E_INVALID, "null scene offsets");
    if (offsets[0] != 0) return set_error(SPHK_E_INVALID, "scene offsets must start at 0");
    for (size_t k = 0; k < nscenes; ++k)
        if (offsets[k + 1] < offsets[k])
            return set_error(SPHK_E_INVALID, "scene offsets must not decrease");
    if (offsets[nscenes] > e->cap)
        return set_error(SPHK_E_CAPACITY, "scenes hold %llu particles, capacity %zu",
                         (unsigned long long)offsets[nscenes], e->cap);
    std::vector<SceneHost> sc(nscenes);
    for (size_t k = 0; k < nscenes; ++k) {
        sc[k].begin = offsets[k];
        sc[k].end = offsets[k + 1];
        if (bounds)
            for (int d = 0; d < 6; ++d) sc[k].bounds[d] = bounds[6 * k + d];
        else
            params_bounds(e->params, sc[k].bounds);
    }
    e->scenes = std::move(sc);
    e->scenes_explicit = true;
    e->ext_unknown = true;
    e->grid_valid = false;
    e->sorted_valid = false;
    return SPHK_OK;
}

int sphk_scene_count(const sphk_engine* e, size_t* out) {
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    *out = e->scenes.size();
    return SPHK_OK;
}

int sphk_set_graph(sphk_engine* e, int enabled) {
    if (int rc = check_engine(e)) return rc;
    e->use_graph = enabled != 0;
    return SPHK_OK;
}

int sphk_size(const sphk_engine* e, size_t* out) {
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    *out = e->n;
    return SPHK_OK;
}

int sphk_clear(sphk_engine* e) {
    if (int rc = check_engine(e)) return rc;
    if (int rc = set_device(e)) return rc;
    e->n = 0;
    e->step_count = 0;
    for (auto& sc : e->scenes) sc.ext_any = false;
    e->ext_unknown = false;
    e->grid_valid = false;
    e->sorted_valid = false;
    HIP_TRY(hipMemsetAsync(e->fscal + 1, 0, sizeof(float), e->stream));
    return SPHK_OK;
}

// Enqueues the kernels of one step on the engine's stream (eagerly or into a capture).
int sphk_step(sphk_engine* e, float dt) {
    if (int rc = check_engine(e)) return rc;
    if (e->n == 0) return SPHK_OK;  // SPHEngine::step returns when empty (src/sph_engine.cpp:94)
    if (int rc = set_device(e)) return rc;
    if (int rc = prepare_grid(e)) return rc;
    const bool adaptive = !(dt > 0.0f);
    hipEvent_t* ev = nullptr;
    // bound the events waiting to be read: fold the oldest in every 256 timed steps
    if (e->timing && e->pending.size() >= 256) {
        if (int rc = drain_timing(e)) return rc;
    }
    if (e->timing) {
        if (e->free_sets.empty()) {
            sphk_engine::EventSet set;
            // timing only: without the system-scope fence (a cache write-back per record,
            // which idled the GPU ~5 us at each of the four records of a step)
            for (auto& x : set.ev) HIP_TRY(hipEventCreateWithFlags(&x, hipEventDisableSystemFence));
            e->free_sets.push_back(set);
        }
        e->pending.push_back(e->free_sets.back());
        e->free_sets.pop_back();
        ev = e->pending.back().ev;
    }
    if (e->use_graph && !ev) {
        if (int rc = graph_step(e, dt, adaptive)) return rc;
    } else {
        if (int rc = enqueue_step(e, dt, adaptive, ev)) return rc;
    }
    e->step_count++;
    e->stats.total_neighbor_queries += e->n;
    e->stats.steps++;
    after_positions_moved(e);
    return SPHK_OK;
}

int sphk_run_steps(sphk_engine* e, size_t num_steps, int adaptive_timestep) {
    if (int rc = check_engine(e)) return rc;
    for (size_t i = 0; i < num_steps; ++i) {
        int rc = sphk_step(e, adaptive_timestep ? 0.0f : e->params.timestep);
        if (rc) return rc;
    }
    return SPHK_OK;
}

int sphk_synchronize(sphk_engine* e) {
    if (int rc = check_engine(e)) return rc;
    if (int rc = set_device(e)) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_download(sphk_engine* e, int which, float* out, size_t count) {
    if (int rc = check_engine(e)) return rc;
    if (!out && count) return set_error(SPHK_E_INVALID, "null output");
    if (int rc = set_device(e)) return rc;
    if (int rc = check_grid_flag(e)) return rc;
    const size_t n = e->n;
    switch (which) {
        case SPHK_POSITIONS:
        case SPHK_VELOCITIES:
        case SPHK_ACCELERATIONS: {
            if (count < n * 3) return set_error(SPHK_E_INVALID, "need %zu floats", n * 3);
            const float4* src = which == SPHK_POSITIONS ? e->pm
                                : which == SPHK_VELOCITIES ? e->vel
                                                           : e->acc;
            std::vector<float4> tmp(n);
            if (n) HIP_TRY(hipMemcpyAsync(tmp.data(), src, n * sizeof(float4), hipMemcpyDeviceToHost,
                                          e->stream));
            HIP_TRY(hipStreamSynchronize(e->stream));
            for (size_t i = 0; i < n; ++i) {
                out[3 * i] = tmp[i].x;
                out[3 * i + 1] = tmp[i].y;
                out[3 * i + 2] = tmp[i].z;
            }
            return SPHK_OK;
        }
        case SPHK_MASSES: {
            if (count < n) return set_error(SPHK_E_INVALID, "need %zu floats", n);
            std::vector<float4> tmp(n);
            if (n) HIP_TRY(hipMemcpyAsync(tmp.data(), e->pm, n * sizeof(float4), hipMemcpyDeviceToHost,
                                          e->stream));
            HIP_TRY(hipStreamSynchronize(e->stream));
            for (size_t i = 0; i < n; ++i) out[i] = tmp[i].w;
            return SPHK_OK;
        }
        case SPHK_DENSITIES:
        case SPHK_PRESSURES: {
            size_t m = std::min(count, e->cap);
            HIP_TRY(hipMemcpyAsync(out, which == SPHK_DENSITIES ? e->rho : e->prs, m * sizeof(float),
                                   hipMemcpyDeviceToHost, e->stream));
            HIP_TRY(hipStreamSynchronize(e->stream));
            return SPHK_OK;
        }
        default:
            return set_error(SPHK_E_INVALID, "unknown download selector %d", which);
    }
}

int sphk_render_pack(sphk_engine* e, const float* colors_rgb, float* out_instances,
                     size_t count) {
    // Renderer::update_particle_data (src/renderer.cpp:279-311) packs position, velocity and
    // colour of every particle on the host each frame; here the same 9-float layout is written
    // on the device into a caller buffer (a mapped GL/VK buffer or any device allocation).
    if (int rc = check_engine(e)) return rc;
    if (!out_instances || count < 9 * e->n)
        return set_error(SPHK_E_INVALID, "need a device buffer of %zu floats", 9 * e->n);
    if (e->n == 0) return SPHK_OK;
    if (int rc = set_device(e)) return rc;
    HIP_TRY(launch_render_pack(e->pm, e->vel, colors_rgb, (uint32_t)e->n, out_instances,
                               e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_get_time(sphk_engine* e, float* out) {
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    if (int rc = set_device(e)) return rc;
    HIP_TRY(hipMemcpyAsync(out, e->fscal + 1, sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_get_step_count(const sphk_engine* e, uint64_t* out) {
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    *out = e->step_count;
    return SPHK_OK;
}

int sphk_total_mass(sphk_engine* e, float* out) {
    // get_total_mass: std::accumulate(densities_[0..N)) * h^3 (src/sph_engine.cpp:187-190)
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    std::vector<float> d(e->n);
    if (e->n) {
        int rc = sphk_download(e, SPHK_DENSITIES, d.data(), e->n);
        if (rc) return rc;
    }
    float acc = 0.0f;
    for (float v : d) acc = acc + v;
    float h = e->params.smoothing_length;
    *out = acc * ((h * h) * h);
    return SPHK_OK;
}

int sphk_total_energy(sphk_engine* e, float* out) {
    // get_total_energy: sum 0.5 m |v|^2 (src/sph_engine.cpp:192-200)
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    if (int rc = set_device(e)) return rc;
    const size_t n = e->n;
    std::vector<float4> v(n), pm(n);
    if (n) {
        HIP_TRY(hipMemcpyAsync(v.data(), e->vel, n * sizeof(float4), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(pm.data(), e->pm, n * sizeof(float4), hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    float ke = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float tx = v[i].x * v[i].x, ty = v[i].y * v[i].y, tz = v[i].z * v[i].z;
        float sp = (tx + ty) + tz;
        ke += (0.5f * pm[i].w) * sp;
    }
    *out = ke;
    return SPHK_OK;
}

int sphk_total_energy_noalloc(sphk_engine* e, float* out) {
    // sphk_total_energy without an allocation: two kernels over the engine's velocities and
    // masses, partial sums in the shared device scratch, one float copied back.
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    if (int rc = set_device(e)) return rc;
    double* sc = scratch_of(e->device);
    if (!sc) return set_error(SPHK_E_INVALID, "no scratch on device %d", e->device);
    const uint32_t n = (uint32_t)e->n;
    uint32_t blocks = (n + kEnergyThreads - 1) / kEnergyThreads;
    blocks = std::max<uint32_t>(1u, std::min<uint32_t>(blocks, (uint32_t)kEnergyBlocks));
    if (blocks > (uint32_t)kEnergyBlocks)
        return set_error(SPHK_E_INVALID, "%u partial sums exceed the scratch of %d", blocks,
                         kEnergyBlocks);
    hipLaunchKernelGGL(k_energy_partial, dim3(blocks), dim3(kEnergyThreads), 0, e->stream,
                       e->pm, e->vel, n, sc);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kEnergyThreads), 0, e->stream, sc, blocks);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, sc + kEnergyBlocks, sizeof(float), hipMemcpyDeviceToHost,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_conservation_errors(sphk_engine* e, float* mass_error, float* energy_error) {
    // compute_conservation_errors (src/sph_engine.cpp:178-185); energy_error is 0 there too
    if (int rc = check_engine(e)) return rc;
    float tm = 0.0f;
    if (int rc = sphk_total_mass(e, &tm)) return rc;
    float initial = (float)e->n * e->params.particle_mass;
    if (mass_error) *mass_error = std::fabs(tm - initial) / initial;
    if (energy_error) *energy_error = 0.0f;
    return SPHK_OK;
}

int sphk_set_timing(sphk_engine* e, int enabled) {
    if (int rc = check_engine(e)) return rc;
    e->timing = enabled != 0;
    return SPHK_OK;
}

int sphk_get_stats(sphk_engine* e, sphk_stats* out) {
    if (int rc = check_engine(e)) return rc;
    if (!out) return set_error(SPHK_E_INVALID, "null output");
    if (int rc = set_device(e)) return rc;
    unsigned long long ns[2];
    HIP_TRY(hipMemcpyAsync(ns, e->nstats, sizeof(ns), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (int rc = drain_timing(e)) return rc;
    *out = e->stats;
    out->max_neighbors = ns[0];
    out->total_neighbors = ns[1];
    return SPHK_OK;
}

int sphk_reset_stats(sphk_engine* e) {
    if (int rc = check_engine(e)) return rc;
    if (int rc = set_device(e)) return rc;
    if (int rc = drain_timing(e)) return rc;
    e->stats = sphk_stats{};
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(e->nstats), 0u, 4, e->stream));
    return SPHK_OK;
}

int sphk_snapshot_save(sphk_engine* e) {
    if (int rc = check_engine(e)) return rc;
    if (int rc = set_device(e)) return rc;
    const size_t n = e->n;
    HIP_TRY(launch_copy_state(e->pm, e->vel, e->acc, e->snap_pm, e->snap_vel, e->snap_acc, n,
                              e->fscal + 1, e->fscal + 2, e->stream));
    e->snap_n = n;
    e->snap_steps = e->step_count;
    e->snap_scenes = e->scenes;
    e->snap_ext_unknown = e->ext_unknown;
    e->has_snapshot = true;
    return SPHK_OK;
}

int sphk_snapshot_restore(sphk_engine* e) {
    if (int rc = check_engine(e)) return rc;
    if (!e->has_snapshot) return set_error(SPHK_E_INVALID, "no snapshot saved");
    if (int rc = set_device(e)) return rc;
    const size_t n = e->snap_n;
    HIP_TRY(launch_copy_state(e->snap_pm, e->snap_vel, e->snap_acc, e->pm, e->vel, e->acc, n,
                              e->fscal + 2, e->fscal + 1, e->stream));
    e->n = n;
    e->step_count = e->snap_steps;
    e->sorted_valid = false;
    if (e->snap_scenes.size() == e->scenes.size()) {
        for (size_t k = 0; k < e->scenes.size(); ++k) {
            SceneHost& sc = e->scenes[k];
            const SceneHost& sn = e->snap_scenes[k];
            for (int d = 0; d < 3; ++d) {
                sc.ext_lo[d] = sn.ext_lo[d];
                sc.ext_hi[d] = sn.ext_hi[d];
            }
            sc.ext_any = sn.ext_any;
        }
        e->ext_unknown = e->snap_ext_unknown;
    } else {
        e->ext_unknown = true;
    }
    e->grid_valid = false;
    return SPHK_OK;
}

// ---- parity/debug ----

static int prepare_sorted(sphk_engine* e);

int sphk_debug_cell_coords(sphk_engine* e, int32_t* coords, size_t count) {
    // Cell coordinates as the GPU binned them: dense index of each particle decoded back to
    // absolute cells.  Particles in the non-finite bin get the x86 conversion of
    // static_cast<int>(std::floor(p * inv)) per axis (INT_MIN where non-finite or out of range),
    // as SpatialHash::get_grid_coords yields on x86 (src/spatial_hash.h:30-36).
    if (int rc = check_engine(e)) return rc;
    if (!coords || count < e->n) return set_error(SPHK_E_INVALID, "need %zu rows", e->n);
    if (int rc = prepare_sorted(e)) return rc;
    const size_t n = e->n;
    std::vector<uint32_t> sid(n), skey(n);
    if (n) {
        HIP_TRY(hipMemcpyAsync(sid.data(), e->sid, n * 4, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(skey.data(), e->skey, n * 4, hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<float> pos(n * 3);
    if (int rc = sphk_download(e, SPHK_POSITIONS, pos.data(), pos.size())) return rc;
    const float inv = 1.0f / e->params.neighbor_search_radius;
    for (size_t s = 0; s < n; ++s) {
        uint32_t id = sid[s];
        uint32_t L = skey[s];
        int32_t* c = coords + 3 * (size_t)id;
        if (L < e->consts.G) {
            size_t k = 0;
            while (k + 1 < e->table.size() && e->table[k + 1].cell_base <= L) ++k;
            const SceneGrid& g = e->table[k];
            uint32_t local = L - g.cell_base;
            uint32_t rz = local % (uint32_t)g.nz;
            uint32_t t = local / (uint32_t)g.nz;
            uint32_t ry = t % (uint32_t)g.ny;
            uint32_t rx = t / (uint32_t)g.ny;
            c[0] = g.ox + (int32_t)rx;
            c[1] = g.oy + (int32_t)ry;
            c[2] = g.oz + (int32_t)rz;
        } else {
            for (int d = 0; d < 3; ++d) {
                float f = std::floor(pos[3 * id + d] * inv);
                c[d] = (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : INT32_MIN;
            }
        }
    }
    return SPHK_OK;
}

int sphk_debug_cell_keys(sphk_engine* e, uint64_t* keys, size_t count) {
    if (int rc = check_engine(e)) return rc;
    if (!keys || count < e->n) return set_error(SPHK_E_INVALID, "need %zu keys", e->n);
    std::vector<int32_t> c(e->n * 3);
    if (int rc = sphk_debug_cell_coords(e, c.data(), e->n)) return rc;
    for (size_t i = 0; i < e->n; ++i) {
        uint64_t hsh = 0;
        hsh |= ((uint64_t)((uint32_t)c[3 * i] & 0x1FFFFFu) << 42);
        hsh |= ((uint64_t)((uint32_t)c[3 * i + 1] & 0x1FFFFFu) << 21);
        hsh |= ((uint64_t)((uint32_t)c[3 * i + 2] & 0x1FFFFFu));
        keys[i] = hsh;
    }
    return SPHK_OK;
}

static int prepare_sorted(sphk_engine* e) {
    if (int rc = set_device(e)) return rc;
    if (int rc = check_scenes(e)) return rc;
    if (!e->grid_valid || e->ext_unknown) {
        if (int rc = ensure_grid(e)) return rc;
    }
    if (!e->sorted_valid && e->n) {
        if (int rc = build_sorted(e)) return rc;
    }
    return SPHK_OK;
}

int sphk_debug_sorted(sphk_engine* e, uint32_t* order, uint32_t* cell_of_sorted, size_t count) {
    if (int rc = check_engine(e)) return rc;
    if (count < e->n) return set_error(SPHK_E_INVALID, "need %zu entries", e->n);
    if (int rc = prepare_sorted(e)) return rc;
    if (e->n) {
        if (order)
            HIP_TRY(hipMemcpyAsync(order, e->sid, e->n * 4, hipMemcpyDeviceToHost, e->stream));
        if (cell_of_sorted)
            HIP_TRY(hipMemcpyAsync(cell_of_sorted, e->skey, e->n * 4, hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_debug_neighbors(sphk_engine* e, int64_t* offsets, int64_t* idx, size_t cap,
                         size_t* total) {
    if (int rc = check_engine(e)) return rc;
    if (int rc = prepare_sorted(e)) return rc;
    const size_t n = e->n;
    uint32_t* dcnt = nullptr;
    HIP_TRY(dalloc(&dcnt, n));
    HIP_TRY(launch_fill_u32(dcnt, 0u, n, e->stream));
    HIP_TRY(launch_nbr_lists(e->walk, e->spm, e->sq, e->skey, e->sid, e->cell_start, (uint32_t)n,
                             e->consts, e->d_scenes, 0, dcnt, nullptr, nullptr, e->shape_density,
                             e->stream));
    std::vector<uint32_t> cnt(n);
    if (n) HIP_TRY(hipMemcpyAsync(cnt.data(), dcnt, n * 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipFree(dcnt));
    std::vector<int64_t> offs(n + 1);
    int64_t o = 0;
    for (size_t i = 0; i < n; ++i) {
        offs[i] = o;
        o += cnt[i];
    }
    offs[n] = o;
    if (total) *total = (size_t)o;
    if (offsets) memcpy(offsets, offs.data(), (n + 1) * sizeof(int64_t));
    if (!idx) return SPHK_OK;
    if (cap < (size_t)o) return set_error(SPHK_E_INVALID, "need %lld index slots", (long long)o);
    int64_t* doffs = nullptr;
    int64_t* didx = nullptr;
    HIP_TRY(dalloc(&doffs, n + 1));
    HIP_TRY(dalloc(&didx, (size_t)o));
    HIP_TRY(hipMemcpyAsync(doffs, offs.data(), (n + 1) * 8, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(launch_nbr_lists(e->walk, e->spm, e->sq, e->skey, e->sid, e->cell_start, (uint32_t)n,
                             e->consts, e->d_scenes, 1, nullptr, doffs, didx, e->shape_density,
                             e->stream));
    if (o) HIP_TRY(hipMemcpyAsync(idx, didx, (size_t)o * 8, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipFree(doffs));
    HIP_TRY(hipFree(didx));
    return SPHK_OK;
}

int sphk_debug_walk_words(sphk_engine* e, uint32_t* items, uint32_t* nwords, uint32_t* wj0,
                          uint32_t* mask, size_t chunk_cap, size_t* nchunks,
                          size_t* words_stored) {
    if (int rc = check_engine(e)) return rc;
    if (!nchunks) return set_error(SPHK_E_INVALID, "null chunk count");
    *nchunks = 0;
    if (int rc = prepare_sorted(e)) return rc;
    const uint32_t n = (uint32_t)e->n;
    if (n == 0) return SPHK_OK;
    HIP_TRY(launch_density_walk(e->walk, e->spm, e->sq, e->svr, e->frows, e->skey, e->cell_start,
                                n, e->consts, e->d_scenes, e->shape_density, e->stream));
    uint32_t nc = 0;
    HIP_TRY(hipMemcpyAsync(&nc, e->walk.nitems, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *nchunks = nc;
    const size_t mw = e->walk.max_words;
    if (words_stored) *words_stored = mw;
    if (!items) return SPHK_OK;
    if (chunk_cap < nc) return set_error(SPHK_E_INVALID, "need %u chunks", nc);
    if (!nwords || !wj0 || !mask) return set_error(SPHK_E_INVALID, "null output array");
    HIP_TRY(hipMemcpyAsync(items, e->walk.items, (size_t)nc * 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(nwords, e->walk.nwords, (size_t)nc * 4, hipMemcpyDeviceToHost,
                           e->stream));
    // the caller's rows hold kMaxWords words per chunk; the engine's hold mw
    if (nc) {
        HIP_TRY(hipMemcpy2DAsync(wj0, kMaxWords * 4, e->walk.wj0, mw * 4, mw * 4, nc,
                                 hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpy2DAsync(mask, kMaxWords * 64 * 4, e->walk.mask, mw * 64 * 4, mw * 64 * 4,
                                 nc, hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return SPHK_OK;
}

int sphk_debug_grid(sphk_engine* e, int32_t* origin3, int32_t* dims3, uint64_t* ncells) {
    if (int rc = check_engine(e)) return rc;
    if (int rc = prepare_sorted(e)) return rc;
    if (int rc = check_grid_flag(e)) return rc;
    const Grid& g = e->scenes[0].grid;
    if (origin3) {
        origin3[0] = g.ox;
        origin3[1] = g.oy;
        origin3[2] = g.oz;
    }
    if (dims3) {
        dims3[0] = g.nx;
        dims3[1] = g.ny;
        dims3[2] = g.nz;
    }
    if (ncells) *ncells = e->consts.G;
    return SPHK_OK;
}

int sphk_set_walk_shapes(sphk_engine* e, int density_shape, int force_shape) {
    if (int rc = check_engine(e)) return rc;
    if (density_shape < 0 || density_shape >= kNumWalkShapes || force_shape < 0 ||
        force_shape >= kNumWalkShapes)
        return set_error(SPHK_E_INVALID, "walk shapes must be in [0, %d)", kNumWalkShapes);
    e->shape_density = density_shape;
    e->shape_force = force_shape;
    e->shapes_explicit = true;
    // a split shape needs the per-part sums, sized with the grid buffers
    if ((walk_shape_split(density_shape) || walk_shape_split(force_shape)) &&
        e->walk.part_cap < e->cap)
        e->grid_valid = false;
    return SPHK_OK;
}

int sphk_get_walk_shapes(const sphk_engine* e, int* density_shape, int* force_shape) {
    if (int rc = check_engine(e)) return rc;
    if (density_shape) *density_shape = e->shape_density;
    if (force_shape) *force_shape = e->shape_force;
    return SPHK_OK;
}

int sphk_walk_shape_count(void) { return kNumWalkShapes; }

const char* sphk_walk_shape_name(int shape) {
    return (shape >= 0 && shape < kNumWalkShapes) ? walk_shape_name(shape) : "";
}

int sphk_last_error(char* buf, size_t len) {
    if (buf && len) {
        size_t m = std::min(len - 1, g_last_error.size());
        memcpy(buf, g_last_error.data(), m);
        buf[m] = 0;
    }
    return (int)g_last_error.size();
}

const char* sphk_version(void) { return "1.0.0"; }

}  // extern "C"
