// sphk_surrogate.hip -- kernels of the learned macro-step surrogate: operand packing, the
// edge pass (message layer 1 on MFMA over the stored neighbour words), the row GEMMs of the
// node MLP and of the backward pass, weight-gradient reductions, loss, Adam, and the
// macro-step integration.  Network and layouts: sphk_surrogate.h.
#include <hip/hip_runtime.h>

#include "sphk_surrogate.h"
#include "sphk_walk_common.h"

namespace sphk {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load8(const __bf16* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
    return z;
}

__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
#pragma unroll
    for (int i = 0; i < 16; ++i) z[i] = 0.0f;
    return z;
}

// Row of accumulator register v for lane half h.
__device__ __forceinline__ uint32_t acc_row(int v, uint32_t h) {
    return (uint32_t)((v >> 2) * 8 + (v & 3)) + 4u * h;
}

// Cubic-spline shape in the operation order of the oracle (no contraction).
__device__ __forceinline__ float w_shape_ieee(float q) {
#pragma clang fp contract(off)
    const float q2 = q * q;
    const float win = (2.0f / 3.0f - q2) + 0.5f * q2 * q;
    const float t = 2.0f - q;
    const float wout = (1.0f / 6.0f) * t * t * t;
    return q <= 1.0f ? win : (q <= 2.0f ? wout : 0.0f);
}

// Edge features of one pair in FP32 (any values for a non-edge: the caller zeroes them).
__device__ __forceinline__ void edge_values(float4 pi, float4 vi, float4 pj, float4 vj,
                                            float d2, const SurScales& sc, float (&f)[9]) {
#pragma clang fp contract(off)
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float rl = __builtin_amdgcn_sqrtf(d2);
    const float w = w_shape_ieee(rl * sc.inv_h);
    f[0] = rx * sc.inv_r;
    f[1] = ry * sc.inv_r;
    f[2] = rz * sc.inv_r;
    f[3] = (vi.x - vj.x) * sc.inv_v;
    f[4] = (vi.y - vj.y) * sc.inv_v;
    f[5] = (vi.z - vj.z) * sc.inv_v;
    f[6] = rl * sc.inv_r;
    f[7] = w * sc.w_norm;
    f[8] = (vj.w - sc.rho0) * sc.rho_scale;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Edge features of one pair rounded to bf16 (round to nearest even, two per conversion),
// zero when `edge` is false: features 0..7 as one operand vector, feature 8 (neighbour
// density) apart.
__device__ __forceinline__ void edge_features(bool edge, float4 pi, float4 vi, float4 pj,
                                              float4 vj, float d2, const SurScales& sc,
                                              bf16x8& e, __bf16& e8) {
    float f[9];
    edge_values(pi, vi, pj, vj, d2, sc, f);
    uint4 w;
    uint32_t* wp = &w.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const f32x2 v = {f[2 * k], f[2 * k + 1]};
        const uint32_t bits = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
        wp[k] = edge ? bits : 0u;
    }
    e = __builtin_bit_cast(bf16x8, w);
    e8 = edge ? (__bf16)f[8] : (__bf16)0.0f;
}

__device__ __forceinline__ uint32_t bits16(__bf16 x) {
    return (uint32_t)__builtin_bit_cast(uint16_t, x);
}

// B operands of message layer 1 for hidden tile hb (n = 32 hb + l % 32).  The K range of an
// MFMA is 16: one half of the lanes supplies W1 columns 0..7, the other half [b1, W1 column 8]
// (k = 8, 9 for block 0; k = 0, 1 for block 1), so W1 e + b1 is one product.
struct Layer1B {
    bf16x8 blk0[4], blk1[4];
};

__device__ __forceinline__ void load_layer1(const __bf16* __restrict__ pk, uint32_t lane,
                                            Layer1B& B) {
    const uint32_t col = lane & 31u;
    const bool hi = lane >= 32u;
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
        const uint32_t n = 32u * hb + col;
        const bf16x8 w = load8(pk + kPkW1 + n * 8u);
        bf16x8 b = zero8();
        b[0] = pk[kPkB1 + n];
        b[1] = pk[kPkW1r + n];
        B.blk0[hb] = hi ? b : w;
        B.blk1[hb] = hi ? w : b;
    }
}

// A operands: row r of block b is the edge of lane 32 b + r.  Its owner supplies features
// 0..7; the partner lane (l ^ 32) supplies [1, feature 8] (zeros for a padding row, which is
// then exactly zero).
__device__ __forceinline__ void layer1_a(bf16x8 fe, __bf16 e8, bool edge, uint32_t lane,
                                         bf16x8& a0, bf16x8& a1) {
    const unsigned long long bal = __ballot(edge);
    const bool pv = ((bal >> (lane ^ 32u)) & 1ull) != 0ull;
    const uint32_t e8p = (uint32_t)__shfl_xor((int)bits16(e8), 32);
    bf16x8 other = zero8();
    other[0] = (__bf16)(pv ? 1.0f : 0.0f);
    other[1] = __builtin_bit_cast(__bf16, (uint16_t)e8p);
    const bool hi = lane >= 32u;
    a0 = hi ? other : fe;
    a1 = hi ? fe : other;
}

constexpr int kEdgeWaves = 2;   // one chunk per workgroup: 32 targets per wavefront
constexpr int kStage = 136;     // LDS row stride (bf16) of staged 128-wide rows
constexpr int kStageX = 152;    // LDS row stride (bf16) of staged node-input rows

// One macro-step of a particle: v' = d^k v + a T, x' = x + T (v + v') / 2, then the wall
// clamp of the plain step (src/particle.cpp:122-153).
__device__ __forceinline__ void macro_integrate(float4& p, float3& v1, float3 v0, float3 a,
                                                float T, float dk, const SceneGrid& g) {
#pragma clang fp contract(off)
    v1 = make_float3(dk * v0.x + a.x * T, dk * v0.y + a.y * T, dk * v0.z + a.z * T);
    p.x = p.x + (0.5f * T) * (v0.x + v1.x);
    p.y = p.y + (0.5f * T) * (v0.y + v1.y);
    p.z = p.z + (0.5f * T) * (v0.z + v1.z);
    const float wd = -0.8f;
    if (p.x < g.lo[0]) { p.x = g.lo[0]; v1.x = v1.x * wd; } else if (p.x > g.hi[0]) { p.x = g.hi[0]; v1.x = v1.x * wd; }
    if (p.y < g.lo[1]) { p.y = g.lo[1]; v1.y = v1.y * wd; } else if (p.y > g.hi[1]) { p.y = g.hi[1]; v1.y = v1.y * wd; }
    if (p.z < g.lo[2]) { p.z = g.lo[2]; v1.z = v1.z * wd; } else if (p.z > g.hi[2]) { p.z = g.hi[2]; v1.z = v1.z * wd; }
}

// Row-tile GEMM of one wavefront: D[t][n] (+)= sum_k A[t][k] Bt[n][k] for the 32 target rows
// of A (LDS, row stride lda) and the 4 output tiles of Bt (global, row stride ldb), K = 16 ks.
__device__ __forceinline__ void tile_gemm(f32x16 (&d)[4], const __bf16* A, int lda,
                                          const __bf16* __restrict__ Bt, int ldb, int ks,
                                          int ntiles, uint32_t lane) {
    const uint32_t col = lane & 31u, h = lane >> 5;
    for (int k = 0; k < ks; ++k) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + col * lda + 16 * k + 8 * h);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (t < ntiles)
                d[t] = mfma16(a, load8(Bt + (size_t)(32 * t + col) * ldb + 16 * k + 8 * h), d[t]);
    }
}

// Activation staging of one wavefront (reuses the list once the walk is done).
struct EpiLds {
    __bf16 s[32 * kStage];      // S, later z (row stride kStage)
    __bf16 x[32 * kStageX];     // node input (row stride kStageX)
};

static_assert(sizeof(EpiLds) <= sizeof(WalkLds), "activation staging reuses the list");

}  // namespace

// ---------------------------------------------------------------------------
// operand packing
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sur_pack(const float* __restrict__ p,
                                                  __bf16* __restrict__ pk,
                                                  float* __restrict__ pf) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < kPkTotal) {
        float v;
        if (i < kPkB1) {
            v = p[kOffW1 + (i / 8u) * kSurNF + i % 8u];
        } else if (i < kPkW1r) {
            v = p[kOffB1 + (i - kPkB1)];
        } else if (i < kPkW2) {
            v = p[kOffW1 + (i - kPkW1r) * kSurNF + 8u];
        } else if (i < kPkW2T) {
            v = p[kOffW2 + (i - kPkW2)];
        } else if (i < kPkWn1) {
            const uint32_t t = i - kPkW2T, in = t / 128u, out = t % 128u;
            v = p[kOffW2 + out * 128u + in];
        } else if (i < kPkWn1T) {
            const uint32_t t = i - kPkWn1, out = t / kSurRow, in = t % kSurRow;
            v = in < (uint32_t)kSurNodeIn ? p[kOffWn1 + out * kSurNodeIn + in] : 0.0f;
        } else if (i < kPkWn2) {
            const uint32_t t = i - kPkWn1T, in = t / 128u, out = t % 128u;
            v = in < (uint32_t)kSurNodeIn ? p[kOffWn1 + out * kSurNodeIn + in] : 0.0f;
        } else if (i < kPkWn2T) {
            const uint32_t t = i - kPkWn2, out = t / 128u, in = t % 128u;
            v = out < 3u ? p[kOffWn2 + out * 128u + in] : 0.0f;
        } else {
            const uint32_t t = i - kPkWn2T, in = t / 32u, out = t % 32u;
            v = out < 3u ? p[kOffWn2 + out * 128u + in] : 0.0f;
        }
        pk[i] = (__bf16)v;
    }
    if (i < kFTotal) {
        float v;
        if (i < kFB1) v = p[kOffW1 + i];
        else if (i < kFB2) v = p[kOffB1 + (i - kFB1)];
        else if (i < kFBn1) v = p[kOffB2 + (i - kFB2)];
        else if (i < kFBn2) v = p[kOffBn1 + (i - kFBn1)];
        else v = (i - kFBn2) < 3u ? p[kOffBn2 + (i - kFBn2)] : 0.0f;
        pf[i] = (float)(__bf16)v;
    }
}

// ---------------------------------------------------------------------------
// forward pass.  Per chunk (two wavefronts of 32 targets):
//   edge walk: S_i = c * sum_j ReLU(W1 e_ij + b1) through
//       sum ReLU(u) = (sum u + sum |u|) / 2,  sum_j u_ij = W1 (sum_j e_ij) + n_i b1,
//     so the per-edge work after the MFMA is one add with an |.| modifier per value.  The
//     two MFMA blocks of a step hold the two lanes' edges of the same 32 targets.
//   node MLP on MFMA from LDS: m = W2 S + c n b2, z = ReLU(Wn1 x + bn1), out = Wn2 z + bn2.
// kIntegrate: the macro-step is applied and written to the caller-order state.  Otherwise
// the S, X, Z and OUT rows are written for the backward pass (batch: rows 64 b + lane;
// no batch: rows = sorted slots).
// ---------------------------------------------------------------------------
template <bool kIntegrate>
__global__ __launch_bounds__(64 * kEdgeWaves) __attribute__((amdgpu_waves_per_eu(2))) void
k_sur_forward(WalkBuffers wb, const uint32_t* __restrict__ batch, uint32_t nbatch,
              const float4* __restrict__ spm, const float4* __restrict__ svr,
              const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sid,
              const uint32_t* __restrict__ cs, StepConsts c, const SceneGrid* __restrict__ scenes,
              const __bf16* __restrict__ pk, const float* __restrict__ pf, SurScales sc,
              SurRows rows, SurStep step) {
    __shared__ WalkLds lds_all[kEdgeWaves];
    __shared__ float lout_all[kEdgeWaves][32][4];
    __shared__ float lcnt_all[kEdgeWaves][32];
    const uint32_t half = uni(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t col = lane & 31u, h = lane >> 5;
    WalkLds& L = lds_all[half];
    const uint32_t b_idx = xcd_unit(batch ? nbatch : uni(wb.nitems[0]), 1, 0);
    if (b_idx == ~0u) return;
    const uint32_t item = batch ? uni(batch[b_idx]) : b_idx;
    const uint32_t clane = 32u * half + col;  // this lane's target within the chunk
    Chunk ch;
    uint32_t s;
    bool active;
    float4 pi;
    if (!load_chunk(item, wb.items, wb.nitems, skey, cs, spm, c, scenes, clane, ch, s, active,
                    pi))
        return;
    const float4 vi = active ? svr[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    Layer1B B;
    load_layer1(pk, lane, B);
    f32x16 acc[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) acc[hb] = zero16();
    float se[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) se[k] = 0.0f;
    float ne = 0.0f;
    walk_pairs(item, ch, s, active, pi, clane, wb, spm, svr, cs, c, L, lane,
               [&](bool edge, float4 pj, float4 vj, float d2) {
                   bf16x8 fe;
                   __bf16 e8;
                   edge_features(edge, pi, vi, pj, vj, d2, sc, fe, e8);
                   // the sums must see the values the MFMA sees (sum u is subtracted from
                   // sum |u| for hidden units that are mostly inactive)
#pragma unroll
                   for (int k = 0; k < 8; ++k) se[k] += (float)fe[k];
                   se[8] += (float)e8;
                   ne += edge ? 1.0f : 0.0f;
                   bf16x8 a0, a1;
                   layer1_a(fe, e8, edge, lane, a0, a1);
#pragma unroll
                   for (int hb = 0; hb < 4; ++hb) {
                       const f32x16 d0 = mfma16(a0, B.blk0[hb], zero16());
                       const f32x16 d1 = mfma16(a1, B.blk1[hb], zero16());
#pragma unroll
                       for (int v = 0; v < 16; ++v) acc[hb][v] += fabsf(d0[v]) + fabsf(d1[v]);
                   }
               });
    // both lanes of a target hold half of its sums
#pragma unroll
    for (int k = 0; k < 9; ++k) se[k] += __shfl_xor(se[k], 32);
    ne += __shfl_xor(ne, 32);
    float* lcnt = lcnt_all[half];
    if (h == 0u) lcnt[col] = ne;
    // sum u + sum |u| = [sum e, n] W1aug^T + acc, on MFMA with the accumulators as C.  The
    // sums enter as three bf16 parts (x = x0 + x1 + x2 to about 2^-27 relative): sum u is
    // cancelled against sum |u| for hidden units that are mostly inactive.
    {
        // lanes 0..31 supply k = 0..7 (feature sums 0..7), lanes 32..63 k = 8, 9 (the edge
        // count, which meets b1, and feature sum 8, which meets W1 column 8)
        float x[8];
        if (h == 0u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) x[k] = se[k];
        } else {
            x[0] = ne;
            x[1] = se[8];
#pragma unroll
            for (int k = 2; k < 8; ++k) x[k] = 0.0f;
        }
#pragma unroll
        for (int part = 0; part < 3; ++part) {
            bf16x8 a;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                a[k] = (__bf16)x[k];
                x[k] = x[k] - (float)a[k];
            }
#pragma unroll
            for (int hb = 0; hb < 4; ++hb) acc[hb] = mfma16(a, B.blk0[hb], acc[hb]);
        }
    }
    wave_sync();  // the list is free from here on
    EpiLds& E = *reinterpret_cast<EpiLds*>(&L);
    const float half_c = 0.5f * sc.msg_scale;
#pragma unroll
    for (int hb = 0; hb < 4; ++hb)
#pragma unroll
        for (int v = 0; v < 16; ++v)
            E.s[acc_row(v, h) * kStage + 32u * hb + col] = (__bf16)(half_c * acc[hb][v]);
    // node input tail: [rho, v, 1, 0 ...] (column 132 = 1 is the bias column of dbn1)
    if (h == 0u) {
#pragma clang fp contract(off)
        const bool a = active;
        __bf16* xr = E.x + col * kStageX + 128;
        xr[0] = (__bf16)(a ? (vi.w - sc.rho0) * sc.rho_scale : 0.0f);
        xr[1] = (__bf16)(a ? vi.x * sc.inv_v : 0.0f);
        xr[2] = (__bf16)(a ? vi.y * sc.inv_v : 0.0f);
        xr[3] = (__bf16)(a ? vi.z * sc.inv_v : 0.0f);
        xr[4] = (__bf16)(a ? 1.0f : 0.0f);
#pragma unroll
        for (int k = 5; k < 16; ++k) xr[k] = (__bf16)0.0f;
    }
    wave_sync();
    // m = W2 S + c n b2 -> X[:, :128]
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = zero16();
    tile_gemm(acc, E.s, kStage, pk + kPkW2, 128, 8, 4, lane);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint32_t n = 32u * t + col;
        const float b2 = pf[kFB2 + n];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const uint32_t r = acc_row(v, h);
            E.x[r * kStageX + n] = (__bf16)(acc[t][v] + sc.msg_scale * lcnt[r] * b2);
        }
    }
    wave_sync();
    // rows for the backward pass: S (and its c n column) before z overwrites it
    const uint32_t len = ch.s_end - ch.s0;
    const bool all_rows = batch != nullptr;
    const size_t row0 = batch ? 64u * (size_t)b_idx : (size_t)ch.s0;
    if constexpr (!kIntegrate) {
        for (uint32_t idx = lane; idx < 32u * 18u; idx += 64u) {
            const uint32_t tl = idx / 18u, part = idx % 18u;
            const uint32_t cl = 32u * half + tl;
            if (!all_rows && cl >= len) continue;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (part < 16u) {
                val = *reinterpret_cast<const uint4*>(E.s + tl * kStage + part * 8u);
            } else if (part == 16u && cl < len) {
                val.x = bits16((__bf16)(sc.msg_scale * lcnt[tl]));
            }
            *reinterpret_cast<uint4*>(rows.S + (row0 + cl) * kSurRow + part * 8u) = val;
            uint4 xv = make_uint4(0u, 0u, 0u, 0u);
            if (cl < len) xv = *reinterpret_cast<const uint4*>(E.x + tl * kStageX + part * 8u);
            *reinterpret_cast<uint4*>(rows.X + (row0 + cl) * kSurRow + part * 8u) = xv;
        }
        wave_sync();
    }
    // z = ReLU(Wn1 x + bn1) -> E.s
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = zero16();
    tile_gemm(acc, E.x, kStageX, pk + kPkWn1, kSurRow, 9, 4, lane);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const uint32_t n = 32u * t + col;
        const float bn1 = pf[kFBn1 + n];
#pragma unroll
        for (int v = 0; v < 16; ++v)
            E.s[acc_row(v, h) * kStage + n] = (__bf16)fmaxf(acc[t][v] + bn1, 0.0f);
    }
    wave_sync();
    // out = Wn2 z + bn2
    acc[0] = zero16();
    tile_gemm(acc, E.s, kStage, pk + kPkWn2, 128, 8, 1, lane);
    float(*lout)[4] = lout_all[half];
    if (col < 3u) {
        const float bn2 = pf[kFBn2 + col];
#pragma unroll
        for (int v = 0; v < 16; ++v) lout[acc_row(v, h)][col] = acc[0][v] + bn2;
    }
    wave_sync();
    if constexpr (kIntegrate) {
        if (h == 0u && active) {
            const float3 a = make_float3(sc.acc_scale * lout[col][0], sc.acc_scale * lout[col][1],
                                         sc.acc_scale * lout[col][2]);
            float4 p = pi;
            float3 v1;
            const SceneGrid& g = scenes[ch.scene];
            macro_integrate(p, v1, make_float3(vi.x, vi.y, vi.z), a, step.T, step.dk, g);
            const uint32_t id = sid[s];
            step.pm[id] = p;
            step.vel[id] = make_float4(v1.x, v1.y, v1.z, 0.0f);
            step.acc[id] = make_float4(a.x, a.y, a.z, 0.0f);
            step.rho[id] = vi.w;
            step.prs[id] = c.k_gas * (vi.w - c.rho0);
        }
    } else {
        // Z rows (+ bias column) and OUT rows
        for (uint32_t idx = lane; idx < 32u * 18u; idx += 64u) {
            const uint32_t tl = idx / 18u, part = idx % 18u;
            const uint32_t cl = 32u * half + tl;
            if (!all_rows && cl >= len) continue;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (cl < len) {
                if (part < 16u) val = *reinterpret_cast<const uint4*>(E.s + tl * kStage + part * 8u);
                else if (part == 16u) val.x = bits16((__bf16)1.0f);
            }
            *reinterpret_cast<uint4*>(rows.Z + (row0 + cl) * kSurRow + part * 8u) = val;
        }
        if (h == 0u && (all_rows || active)) {
            const size_t row = row0 + clane;
            rows.OUT[row] = active ? make_float4(lout[col][0], lout[col][1], lout[col][2], 0.0f)
                                   : make_float4(0.f, 0.f, 0.f, 0.f);
            rows.cnt[row] = active ? sc.msg_scale * ne : 0.0f;
            rows.rowslot[row] = active ? s : kSurNoSlot;
        }
    }
}

// ---------------------------------------------------------------------------
// row GEMM: one wavefront per 32 rows, all N columns (tiles of 32)
// ---------------------------------------------------------------------------
template <int kEpi, int kNT>
__global__ __launch_bounds__(256) void k_sur_gemm(const __bf16* __restrict__ A, int lda,
                                                  const __bf16* __restrict__ Bt, int ldb, int K,
                                                  uint32_t R, SurEpiArgs ep) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t row0 = (blockIdx.x * 4u + wave) * 32u;
    if (row0 >= R) return;
    const uint32_t h = lane >> 5, col = lane & 31u;
    f32x16 acc[kNT];
#pragma unroll
    for (int t = 0; t < kNT; ++t) acc[t] = zero16();
    const __bf16* ap = A + (size_t)(row0 + col) * lda + 8u * h;
    const __bf16* bp = Bt + (size_t)col * ldb + 8u * h;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 a = load8(ap + k0);
#pragma unroll
        for (int t = 0; t < kNT; ++t) acc[t] = mfma16(a, load8(bp + (size_t)t * 32 * ldb + k0), acc[t]);
    }
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
        const uint32_t n = 32u * t + col;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const size_t row = row0 + acc_row(v, h);
            const float x = acc[t][v];
            if constexpr (kEpi == kEpiM) {
                static_cast<__bf16*>(ep.out)[row * kSurRow + n] =
                    (__bf16)(x + ep.cnt[row] * ep.bias[n]);
            } else if constexpr (kEpi == kEpiZ) {
                static_cast<__bf16*>(ep.out)[row * kSurRow + n] = (__bf16)fmaxf(x + ep.bias[n], 0.0f);
            } else if constexpr (kEpi == kEpiOut) {
                if (n < 3u) static_cast<float*>(ep.out)[row * 4 + n] = x + ep.bias[n];
            } else if constexpr (kEpi == kEpiDZ) {
                const bool on = (float)ep.z[row * kSurRow + n] > 0.0f;
                static_cast<__bf16*>(ep.out)[row * 128 + n] = (__bf16)(on ? x : 0.0f);
            } else if constexpr (kEpi == kEpiDM) {
                static_cast<__bf16*>(ep.out)[row * 128 + n] = (__bf16)x;
            } else {
                static_cast<__bf16*>(ep.out)[row * 128 + n] = (__bf16)(ep.scale * x);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// weight-gradient partials: one wavefront per (32 x 32 tile, row slab)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_sur_wgrad(const __bf16* __restrict__ P, int ldp, int M,
                                                  const __bf16* __restrict__ Q, int ldq, int N,
                                                  uint32_t R, uint32_t rps,
                                                  float* __restrict__ part) {
    const int ntn = (N + 31) / 32;
    const int tm = blockIdx.x / ntn, tn = blockIdx.x % ntn;
    const uint32_t slab = blockIdx.y;
    const uint32_t lane = threadIdx.x, h = lane >> 5, col = lane & 31u;
    const int m = 32 * tm + (int)col, n = 32 * tn + (int)col;
    const uint32_t r_begin = slab * rps;
    const uint32_t r_end = min(R, r_begin + rps);
    f32x16 acc = zero16();
    for (uint32_t r0 = r_begin; r0 < r_end; r0 += 16u) {
        bf16x8 a, b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t r = r0 + 8u * h + i;
            const bool in = r < r_end;
            a[i] = (in && m < M) ? P[(size_t)r * ldp + m] : (__bf16)0.0f;
            b[i] = (in && n < N) ? Q[(size_t)r * ldq + n] : (__bf16)0.0f;
        }
        acc = mfma16(a, b, acc);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int mm = 32 * tm + (int)acc_row(v, h);
        if (mm < M && n < N) part[((size_t)slab * M + mm) * N + n] = acc[v];
    }
}

// ---------------------------------------------------------------------------
// edge pass (backward): d(W1|b1) = sum_i sum_j (dS_i * [u_ij > 0]) (e_ij | 1)^T.  The masked
// gradient is formed in the accumulator layout of the recomputed layer-1 product and fed
// back as an A operand (K = edges); the edge features reach the B operand through LDS.  One
// workgroup per batch chunk, 32 targets per wavefront.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64 * kEdgeWaves) void k_sur_edges_bwd(
    WalkBuffers wb, const uint32_t* __restrict__ batch, uint32_t nbatch,
    const float4* __restrict__ spm, const float4* __restrict__ svr,
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ cs, StepConsts c,
    const SceneGrid* __restrict__ scenes, const __bf16* __restrict__ pk, SurScales sc,
    const __bf16* __restrict__ DS, float* __restrict__ part) {
    __shared__ WalkLds lds_all[kEdgeWaves];
    __shared__ __bf16 le_all[kEdgeWaves][64 * 16];
    const uint32_t half = uni(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t col = lane & 31u, h = lane >> 5;
    WalkLds& L = lds_all[half];
    __bf16* le = le_all[half];
    const uint32_t b = blockIdx.x;
    if (b >= nbatch) return;
    const uint32_t item = uni(batch[b]);
    const uint32_t clane = 32u * half + col;
    Chunk ch;
    uint32_t s;
    bool active;
    float4 pi;
    if (!load_chunk(item, wb.items, wb.nitems, skey, cs, spm, c, scenes, clane, ch, s, active,
                    pi))
        return;
    const float4 vi = active ? svr[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    Layer1B B;
    load_layer1(pk, lane, B);
    float ds[4][16];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const size_t row = 64u * (size_t)b + 32u * half + acc_row(v, h);
            ds[hb][v] = (float)DS[row * 128 + 32u * hb + col];
        }
    f32x16 accw[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) accw[hb] = zero16();
    walk_pairs(item, ch, s, active, pi, clane, wb, spm, svr, cs, c, L, lane,
               [&](bool edge, float4 pj, float4 vj, float d2) {
                   bf16x8 fe;
                   __bf16 e8;
                   edge_features(edge, pi, vi, pj, vj, d2, sc, fe, e8);
                   // E row of this lane's edge: [features 0..8, 1, 0 ...]
                   bf16x8 ext = zero8();
                   ext[0] = e8;
                   ext[1] = (__bf16)(edge ? 1.0f : 0.0f);
                   *reinterpret_cast<bf16x8*>(le + lane * 16) = fe;
                   *reinterpret_cast<bf16x8*>(le + lane * 16 + 8) = ext;
                   wave_sync();
                   bf16x8 a[2];
                   layer1_a(fe, e8, edge, lane, a[0], a[1]);
#pragma unroll
                   for (int blk = 0; blk < 2; ++blk) {
                       bf16x8 e[2];
#pragma unroll
                       for (int g = 0; g < 2; ++g)
#pragma unroll
                           for (int i = 0; i < 8; ++i) {
                               const uint32_t er = 32u * blk + 16u * g + acc_row(i, h);
                               e[g][i] = col < 10u ? le[er * 16 + col] : (__bf16)0.0f;
                           }
#pragma unroll
                       for (int hb = 0; hb < 4; ++hb) {
                           const f32x16 d =
                               mfma16(a[blk], blk == 0 ? B.blk0[hb] : B.blk1[hb], zero16());
                           bf16x8 g0, g1;
#pragma unroll
                           for (int i = 0; i < 8; ++i) {
                               g0[i] = (__bf16)(d[i] > 0.0f ? ds[hb][i] : 0.0f);
                               g1[i] = (__bf16)(d[8 + i] > 0.0f ? ds[hb][8 + i] : 0.0f);
                           }
                           accw[hb] = mfma16(g0, e[0], accw[hb]);
                           accw[hb] = mfma16(g1, e[1], accw[hb]);
                       }
                   }
                   wave_sync();
               });
    if (col < 16u) {
        const size_t pbase = (2u * (size_t)b + half) * 128u;
#pragma unroll
        for (int hb = 0; hb < 4; ++hb)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint32_t m = 32u * hb + acc_row(v, h);
                part[(pbase + m) * 16u + col] = accw[hb][v];
            }
    }
}

// ---------------------------------------------------------------------------
// small element-wise kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sur_reduce(const float* __restrict__ part,
                                                    uint32_t slabs, int M_all, int ld, int M,
                                                    int ncols, int bias_col, uint32_t dst,
                                                    uint32_t bias_dst, float scale,
                                                    float* __restrict__ grad) {
    const int per_row = ncols + (bias_col >= 0 ? 1 : 0);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= M * per_row) return;
    const int m = i / per_row, k = i % per_row;
    const int n = k < ncols ? k : bias_col;
    float sum = 0.0f;
    for (uint32_t sl = 0; sl < slabs; ++sl) sum += part[((size_t)sl * M_all + m) * ld + n];
    if (k < ncols) grad[dst + (uint32_t)(m * ncols + n)] = scale * sum;
    else grad[bias_dst + (uint32_t)m] = scale * sum;
}

__device__ __forceinline__ float block_sum_256(float x, float* red) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_sur_loss(const float* __restrict__ out,
                                                  const uint32_t* __restrict__ rowslot,
                                                  const uint32_t* __restrict__ sid,
                                                  const float4* __restrict__ target, uint32_t R,
                                                  float scale, __bf16* __restrict__ dout,
                                                  float* __restrict__ loss_part) {
    __shared__ float red[4];
    __shared__ float redc[4];
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    float l = 0.0f, cnt = 0.0f;
    if (r < R) {
        const uint32_t slot = rowslot[r];
        float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
        if (slot != kSurNoSlot) {
            cnt = 1.0f;
            const float4 t = target[sid[slot]];
            d0 = out[4 * (size_t)r] - t.x;
            d1 = out[4 * (size_t)r + 1] - t.y;
            d2 = out[4 * (size_t)r + 2] - t.z;
            l = (d0 * d0 + d1 * d1) + d2 * d2;
        }
        const __bf16 g0 = (__bf16)(scale * d0), g1 = (__bf16)(scale * d1),
                     g2 = (__bf16)(scale * d2);
        uint4 v = make_uint4((uint32_t)__builtin_bit_cast(uint16_t, g0) |
                                 ((uint32_t)__builtin_bit_cast(uint16_t, g1) << 16),
                             (uint32_t)__builtin_bit_cast(uint16_t, g2), 0u, 0u);
        uint4* o = reinterpret_cast<uint4*>(dout + (size_t)r * kSurOutRow);
        o[0] = v;
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
        o[1] = z;
        o[2] = z;
        o[3] = z;
    }
    const float tot = block_sum_256(l, red);
    const float tc = block_sum_256(cnt, redc);
    if (threadIdx.x == 0) {
        loss_part[2 * blockIdx.x] = tot;
        loss_part[2 * blockIdx.x + 1] = tc;
    }
}

__global__ __launch_bounds__(256) void k_sur_targets(const float4* __restrict__ v0,
                                                     const float4* __restrict__ v4, uint32_t n,
                                                     float dk, float inv_Tas,
                                                     float4* __restrict__ target) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 a = v0[i], b = v4[i];
    target[i] = make_float4((b.x - dk * a.x) * inv_Tas, (b.y - dk * a.y) * inv_Tas,
                            (b.z - dk * a.z) * inv_Tas, 0.0f);
}

// Particles with a non-finite coordinate (dense index G) have no neighbours: free fall over
// the macro-step, as in the plain step.
__global__ __launch_bounds__(256) void k_sur_nonfinite(const float4* __restrict__ spm,
                                                       const float4* __restrict__ svr,
                                                       const uint32_t* __restrict__ sid,
                                                       const uint32_t* __restrict__ cs,
                                                       uint32_t n, StepConsts c,
                                                       const SceneGrid* __restrict__ scenes,
                                                       SurStep step) {
    const uint32_t s = cs[c.G] + blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    const float4 vr = svr[s];
    const uint32_t id = sid[s];
    float4 p = spm[s];
    const float3 a = make_float3(0.0f, c.grav, 0.0f);
    float3 v1;
    macro_integrate(p, v1, make_float3(vr.x, vr.y, vr.z), a, step.T, step.dk,
                    scenes[scene_of_particle(scenes, c.nscenes, id)]);
    const float rho_i = p.w * c.w0;
    step.pm[id] = p;
    step.vel[id] = make_float4(v1.x, v1.y, v1.z, 0.0f);
    step.acc[id] = make_float4(a.x, a.y, a.z, 0.0f);
    step.rho[id] = rho_i;
    step.prs[id] = c.k_gas * (rho_i - c.rho0);
}

__global__ __launch_bounds__(256) void k_sur_adam(float* __restrict__ w, float* __restrict__ m,
                                                  float* __restrict__ v,
                                                  const float* __restrict__ g, uint32_t np,
                                                  float lr, float beta1, float beta2, float eps,
                                                  float bc1, float bc2) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.0f - beta2) * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    const float mh = __fdiv_rn(mi, bc1), vh = __fdiv_rn(vi, bc2);
    w[i] = w[i] - __fdiv_rn(lr * mh, __fsqrt_rn(vh) + eps);
}

__global__ __launch_bounds__(256) void k_sur_rho_diff(const float* __restrict__ a,
                                                      const float* __restrict__ b, uint32_t n,
                                                      float* __restrict__ part) {
    __shared__ float red[4];
    __shared__ float redm[4];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const float d = i < n ? fabsf(a[i] - b[i]) : 0.0f;
    float mx = d;
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63u) == 0) redm[threadIdx.x >> 6] = mx;
    const float tot = block_sum_256(d, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = tot;
        part[2 * blockIdx.x + 1] = fmaxf(fmaxf(redm[0], redm[1]), fmaxf(redm[2], redm[3]));
    }
}

// Densities of the sorted state in caller order (non-finite particles: m W(0)).
__global__ __launch_bounds__(256) void k_sur_rho_home(const float4* __restrict__ spm,
                                                      const float4* __restrict__ svr,
                                                      const uint32_t* __restrict__ sid,
                                                      const uint32_t* __restrict__ cs,
                                                      uint32_t n, StepConsts c,
                                                      float* __restrict__ out) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    out[sid[s]] = s < cs[c.G] ? svr[s].w : spm[s].w * c.w0;
}

// Network accelerations a = a_s out in caller order (non-finite particles: (0, g, 0)).
__global__ __launch_bounds__(256) void k_sur_scatter_out(const float* __restrict__ out,
                                                         const uint32_t* __restrict__ sid,
                                                         const uint32_t* __restrict__ cs,
                                                         uint32_t n, StepConsts c,
                                                         float acc_scale,
                                                         float4* __restrict__ dst) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    float4 a = make_float4(0.0f, c.grav, 0.0f, 0.0f);
    if (s < cs[c.G])
        a = make_float4(acc_scale * out[4 * (size_t)s], acc_scale * out[4 * (size_t)s + 1],
                        acc_scale * out[4 * (size_t)s + 2], 0.0f);
    dst[sid[s]] = a;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_sur_rho_home(const float4* spm, const float4* svr, const uint32_t* sid,
                               const uint32_t* cs, uint32_t n, const StepConsts& c, float* out,
                               hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_rho_home, dim3((n + 255) / 256), dim3(256), 0, st, spm, svr, sid, cs,
                       n, c, out);
    return hipGetLastError();
}

hipError_t launch_sur_scatter_out(const float* out, const uint32_t* sid, const uint32_t* cs,
                                  uint32_t n, const StepConsts& c, float acc_scale, float4* dst,
                                  hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_scatter_out, dim3((n + 255) / 256), dim3(256), 0, st, out, sid, cs,
                       n, c, acc_scale, dst);
    return hipGetLastError();
}

hipError_t launch_sur_pack(const float* params, __bf16* pk, float* pf, hipStream_t st) {
    hipLaunchKernelGGL(k_sur_pack, dim3((kPkTotal + 255) / 256), dim3(256), 0, st, params, pk, pf);
    return hipGetLastError();
}

hipError_t launch_sur_forward(bool integrate, const WalkBuffers& wb, const uint32_t* batch,
                              uint32_t nbatch, uint32_t max_items, const float4* spm,
                              const float4* svr, const uint32_t* skey, const uint32_t* sid,
                              const uint32_t* cs, uint32_t n, const StepConsts& c,
                              const SceneGrid* scenes, const __bf16* pk, const float* pf,
                              const SurScales& sc, const SurRows& rows, const SurStep& step,
                              hipStream_t st) {
    const uint32_t chunks = batch ? nbatch : max_items;
    if (chunks == 0) return hipSuccess;
    const uint32_t blocks = (chunks + kXcds - 1) / kXcds * kXcds;  // xcd_unit() needs it
    if (integrate) {
        hipLaunchKernelGGL(k_sur_forward<true>, dim3(blocks), dim3(64 * kEdgeWaves), 0, st, wb,
                           batch, nbatch, spm, svr, skey, sid, cs, c, scenes, pk, pf, sc, rows,
                           step);
        // the non-finite bin holds at most n particles; the kernel reads its start on the device
        hipLaunchKernelGGL(k_sur_nonfinite, dim3((n + 255) / 256), dim3(256), 0, st, spm, svr,
                           sid, cs, n, c, scenes, step);
    } else {
        hipLaunchKernelGGL(k_sur_forward<false>, dim3(blocks), dim3(64 * kEdgeWaves), 0, st, wb,
                           batch, nbatch, spm, svr, skey, sid, cs, c, scenes, pk, pf, sc, rows,
                           step);
    }
    return hipGetLastError();
}

hipError_t launch_sur_gemm(int epi, const __bf16* A, int lda, const __bf16* Bt, int ldb, int K,
                           int N, uint32_t R, const SurEpiArgs& ep, hipStream_t st) {
    if (R == 0) return hipSuccess;
    const dim3 grid((R / 32u + 3u) / 4u), block(256);
#define SUR_GEMM(E, NT) \
    hipLaunchKernelGGL((k_sur_gemm<E, NT>), grid, block, 0, st, A, lda, Bt, ldb, K, R, ep)
    switch (epi) {
        case kEpiM: SUR_GEMM(kEpiM, 4); break;
        case kEpiZ: SUR_GEMM(kEpiZ, 4); break;
        case kEpiOut: SUR_GEMM(kEpiOut, 1); break;
        case kEpiDZ: SUR_GEMM(kEpiDZ, 4); break;
        case kEpiDM: SUR_GEMM(kEpiDM, 4); break;
        default: SUR_GEMM(kEpiDS, 4); break;
    }
#undef SUR_GEMM
    (void)N;
    return hipGetLastError();
}

hipError_t launch_sur_wgrad(const __bf16* P, int ldp, int M, const __bf16* Q, int ldq, int N,
                            uint32_t R, uint32_t rows_per_slab, uint32_t slabs, float* part,
                            hipStream_t st) {
    const int tiles = ((M + 31) / 32) * ((N + 31) / 32);
    hipLaunchKernelGGL(k_sur_wgrad, dim3(tiles, slabs), dim3(64), 0, st, P, ldp, M, Q, ldq, N, R,
                       rows_per_slab, part);
    return hipGetLastError();
}

hipError_t launch_sur_edges_bwd(const WalkBuffers& wb, const uint32_t* batch, uint32_t nbatch,
                                const float4* spm, const float4* svr, const uint32_t* skey,
                                const uint32_t* cs, const StepConsts& c,
                                const SceneGrid* scenes, const __bf16* pk, const SurScales& sc,
                                const __bf16* DS, float* part, hipStream_t st) {
    if (nbatch == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_edges_bwd, dim3(nbatch), dim3(64 * kEdgeWaves), 0, st, wb, batch,
                       nbatch, spm, svr, skey, cs, c, scenes, pk, sc, DS, part);
    return hipGetLastError();
}

hipError_t launch_sur_reduce(const float* part, uint32_t slabs, int M_all, int ld, int M,
                             int ncols, int bias_col, uint32_t dst, uint32_t bias_dst,
                             float scale, float* grad, hipStream_t st) {
    const int total = M * (ncols + (bias_col >= 0 ? 1 : 0));
    hipLaunchKernelGGL(k_sur_reduce, dim3((total + 255) / 256), dim3(256), 0, st, part, slabs,
                       M_all, ld, M, ncols, bias_col, dst, bias_dst, scale, grad);
    return hipGetLastError();
}

hipError_t launch_sur_loss(const float* out, const uint32_t* rowslot, const uint32_t* sid,
                           const float4* target, uint32_t R, __bf16* dout, float* loss_part,
                           uint32_t* nparts, hipStream_t st) {
    const uint32_t blocks = (R + 255u) / 256u;
    *nparts = blocks;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_loss, dim3(blocks), dim3(256), 0, st, out, rowslot, sid, target, R,
                       2.0f, dout, loss_part);
    return hipGetLastError();
}

hipError_t launch_sur_targets(const float4* v0, const float4* v4, uint32_t n, float dk,
                              float inv_Tas, float4* target, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_targets, dim3((n + 255) / 256), dim3(256), 0, st, v0, v4, n, dk,
                       inv_Tas, target);
    return hipGetLastError();
}

hipError_t launch_sur_adam(float* w, float* m, float* v, const float* g, uint32_t np, float lr,
                           float beta1, float beta2, float eps, float bc1, float bc2,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_sur_adam, dim3((np + 255) / 256), dim3(256), 0, st, w, m, v, g, np, lr,
                       beta1, beta2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t launch_sur_rho_diff(const float* a, const float* b, uint32_t n, float* part,
                               uint32_t* nparts, hipStream_t st) {
    const uint32_t blocks = (n + 255u) / 256u;
    *nparts = blocks;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_rho_diff, dim3(blocks), dim3(256), 0, st, a, b, n, part);
    return hipGetLastError();
}

}  // namespace sphk
