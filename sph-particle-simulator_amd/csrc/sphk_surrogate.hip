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

// Edge features of one pair (zero when `edge` is false), rounded to bf16.
__device__ __forceinline__ bf16x8 edge_features(bool edge, float4 pi, float4 vi, float4 pj,
                                                float4 vj, float d2, const SurScales& sc) {
#pragma clang fp contract(off)
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float rl = __builtin_amdgcn_sqrtf(d2);
    const float w = w_shape_ieee(rl * sc.inv_h);
    float f[8];
    f[0] = rx * sc.inv_r;
    f[1] = ry * sc.inv_r;
    f[2] = rz * sc.inv_r;
    f[3] = (vi.x - vj.x) * sc.inv_v;
    f[4] = (vi.y - vj.y) * sc.inv_v;
    f[5] = (vi.z - vj.z) * sc.inv_v;
    f[6] = rl * sc.inv_r;
    f[7] = w * sc.w_norm;
    bf16x8 e;
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = (__bf16)(edge ? f[k] : 0.0f);
    return e;
}

// B operands of message layer 1 for hidden tile hb (n = 32 hb + l % 32).  The K range of an
// MFMA is 16: one half of the lanes supplies the 8 edge features, the other half a bias
// slot (k = 8 for block 0, k = 0 for block 1) so that W1 e + b1 is one product.
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
        B.blk0[hb] = hi ? b : w;
        B.blk1[hb] = hi ? w : b;
    }
}

// A operands: row r of block b is the edge of lane 32 b + r.  The lane that owns the edge
// supplies its features; the partner lane (l ^ 32) supplies the bias slot (1 for a real
// edge, 0 for padding, so padding rows are exactly zero).
__device__ __forceinline__ void layer1_a(bf16x8 fe, bool edge, uint32_t lane, bf16x8& a0,
                                         bf16x8& a1) {
    const unsigned long long bal = __ballot(edge);
    const bool pv = ((bal >> (lane ^ 32u)) & 1ull) != 0ull;
    bf16x8 other = zero8();
    other[0] = (__bf16)(pv ? 1.0f : 0.0f);
    const bool hi = lane >= 32u;
    a0 = hi ? other : fe;
    a1 = hi ? fe : other;
}

// Wave max of a per-lane count.
__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(x, off);
        x = o > x ? o : x;
    }
    return uni(x);
}

// Walks every neighbour j != i of each lane's particle in uniform steps over the wave and
// calls body(edge, pj, vj, d2) once per step; lanes with nothing left get edge = false.
// The neighbour words come from the density pass (re-tested for chunks beyond kMaxWords).
template <int kSeg, class Body>
__device__ __forceinline__ void for_each_edge(uint32_t item, const Chunk& ch, uint32_t s,
                                              bool active, float4 pi,
                                              const WalkBuffers& wb, const float4* spm,
                                              const float4* svr, const uint32_t* cs,
                                              const StepConsts& c, uint32_t* lmask,
                                              uint32_t* lwj0, uint32_t lane, Body&& body) {
    const int total_words = (int)uni(wb.nwords[item]);
    const bool stored = total_words <= kMaxWords;
    const uint32_t s_safe = active ? s : ch.s0;
    Cursor cu;
    if (!stored) cursor_begin(cu, ch, cs, c);
    int seg_base = 0;
    for (;;) {
        int nw;
        SegLane sl;
        if (stored) {
            nw = min(kSeg, total_words - seg_base);
            if (nw <= 0) break;
            const size_t base = (size_t)item * kMaxWords + seg_base;
            sl.nonempty = 0;
            sl.count = 0;
            for (int w = 0; w < nw; ++w) {
                const uint32_t word = wb.mask[(base + w) * 64 + lane];
                lmask[w * 64 + lane] = word;
                sl.nonempty |= (word != 0u ? 1u : 0u) << w;
                sl.count += (uint32_t)__builtin_popcount(word);
            }
            for (int w = (int)lane; w < nw; w += 64) lwj0[w] = wb.wj0[base + w];
        } else {
            nw = phase1_segment<kSeg>(cu, ch, spm, cs, c, pi.x, pi.y, pi.z, lmask, lwj0, lane,
                                      sl);
            if (nw == 0) break;
        }
        wave_sync();
        const uint32_t tmax = wave_max(sl.count);
        BitWalk<kSeg> bw;
        bw.ne = sl.nonempty;
        for (uint32_t t = 0; t < tmax; ++t) {
            const bool more = t < sl.count;
            uint32_t j = s_safe;
            if (more) j = bw.next(lmask, lwj0, lane);
            const float4 pj = spm[j];
            const float4 vj = svr[j];
            const float d2 = dist2_exact(pi.x, pi.y, pi.z, pj.x, pj.y, pj.z);
            const bool edge = more && j != s && d2 <= c.r2;  // the reference's test, self excluded
            body(edge, pj, vj, d2);
        }
        seg_base += nw;
        wave_sync();
        if (!stored && cu.done) break;
    }
}

constexpr int kEdgeSeg = 8;
constexpr int kEdgeWaves = 2;
constexpr int kStage = 136;  // LDS row stride (bf16) of the staged S rows

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
            v = p[kOffW1 + i];
        } else if (i < kPkW2) {
            v = p[kOffB1 + (i - kPkB1)];
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
// edge pass (forward): S_i = c * sum_j ReLU(W1 e_ij + b1) through
//   sum ReLU(u) = (sum u + sum |u|) / 2,  sum_j u_ij = W1 (sum_j e_ij) + n_i b1,
// so the per-edge work after the MFMA is one add with an |.| modifier per value.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64 * kEdgeWaves) void k_sur_edges(
    WalkBuffers wb, const uint32_t* __restrict__ batch, uint32_t nbatch,
    const float4* __restrict__ spm, const float4* __restrict__ svr,
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ cs, StepConsts c,
    const SceneGrid* __restrict__ scenes, const __bf16* __restrict__ pk,
    const float* __restrict__ pf, SurScales sc, __bf16* __restrict__ S, __bf16* __restrict__ X,
    __bf16* __restrict__ Z, float* __restrict__ cnt, uint32_t* __restrict__ rowslot) {
    __shared__ uint32_t lmask_all[kEdgeWaves][kEdgeSeg * 64];
    __shared__ uint32_t lwj0_all[kEdgeWaves][kEdgeSeg];
    __shared__ float lse_all[kEdgeWaves][64][9];
    __shared__ __bf16 stage_all[kEdgeWaves][64 * kStage];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* lmask = lmask_all[wave];
    uint32_t* lwj0 = lwj0_all[wave];
    const uint32_t w_idx = swizzled_block() * kEdgeWaves + wave;
    uint32_t item;
    if (batch) {
        if (w_idx >= nbatch) return;
        item = uni(batch[w_idx]);
    } else {
        item = w_idx;
    }
    Chunk ch;
    uint32_t s;
    bool active;
    float4 pi;
    if (!load_chunk(item, wb.items, wb.nitems, skey, cs, spm, c, scenes, lane, ch, s, active,
                    pi))
        return;
    const float4 vi = active ? svr[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    Layer1B B;
    load_layer1(pk, lane, B);
    f32x16 acc0[4], acc1[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
        acc0[hb] = zero16();
        acc1[hb] = zero16();
    }
    float se[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) se[k] = 0.0f;
    float ne = 0.0f;
    for_each_edge<kEdgeSeg>(
        item, ch, s, active, pi, wb, spm, svr, cs, c, lmask, lwj0, lane,
        [&](bool edge, float4 pj, float4 vj, float d2) {
            const bf16x8 fe = edge_features(edge, pi, vi, pj, vj, d2, sc);
#pragma unroll
            for (int k = 0; k < 8; ++k) se[k] += (float)fe[k];
            ne += edge ? 1.0f : 0.0f;
            bf16x8 a0, a1;
            layer1_a(fe, edge, lane, a0, a1);
#pragma unroll
            for (int hb = 0; hb < 4; ++hb) {
                const f32x16 d0 = mfma16(a0, B.blk0[hb], zero16());
                const f32x16 d1 = mfma16(a1, B.blk1[hb], zero16());
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    acc0[hb][v] += fabsf(d0[v]);
                    acc1[hb][v] += fabsf(d1[v]);
                }
            }
        });
    // per-target feature sums -> LDS, then S = c (sum u + sum |u|) / 2 staged as bf16 rows
    float(*lse)[9] = lse_all[wave];
    __bf16* stage = stage_all[wave];
#pragma unroll
    for (int k = 0; k < 8; ++k) lse[lane][k] = se[k];
    lse[lane][8] = ne;
    wave_sync();
    const uint32_t col = lane & 31u, h = lane >> 5;
    float w1f[4][8], b1f[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
        const uint32_t n = 32u * hb + col;
#pragma unroll
        for (int k = 0; k < 8; ++k) w1f[hb][k] = pf[kFW1 + n * 8u + k];
        b1f[hb] = pf[kFB1 + n];
    }
    const float half_c = 0.5f * sc.msg_scale;
#pragma unroll
    for (int blk = 0; blk < 2; ++blk) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const uint32_t tl = 32u * blk + acc_row(v, h);
            float et[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) et[k] = lse[tl][k];
#pragma unroll
            for (int hb = 0; hb < 4; ++hb) {
                float su = b1f[hb] * et[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) su += w1f[hb][k] * et[k];
                const float a = blk == 0 ? acc0[hb][v] : acc1[hb][v];
                stage[tl * kStage + 32u * hb + col] = (__bf16)(half_c * (su + a));
            }
        }
    }
    wave_sync();
    // rows out: S (128 + bias column block), X and Z tails
    const uint32_t len = ch.s_end - ch.s0;
    const uint32_t row0 = batch ? 64u * w_idx : ch.s0;
    const bool all_rows = batch != nullptr;
    for (uint32_t idx = lane; idx < 64u * 18u; idx += 64u) {
        const uint32_t tl = idx / 18u, part = idx % 18u;
        if (!all_rows && tl >= len) continue;
        const size_t row = (size_t)row0 + tl;
        uint4 val = make_uint4(0u, 0u, 0u, 0u);
        if (part < 16u) {
            val = *reinterpret_cast<const uint4*>(stage + tl * kStage + part * 8u);
        } else if (part == 16u && tl < len) {
            const __bf16 cn = (__bf16)(sc.msg_scale * lse[tl][8]);
            val.x = (uint32_t)__builtin_bit_cast(uint16_t, cn);
        }
        *reinterpret_cast<uint4*>(S + row * kSurRow + part * 8u) = val;
    }
    if (all_rows || active) {
        const size_t row = (size_t)row0 + lane;
        uint4 xt0 = make_uint4(0u, 0u, 0u, 0u), xt1 = xt0, zt0 = xt0;
        if (active) {
#pragma clang fp contract(off)
            const __bf16 f0 = (__bf16)((vi.w - sc.rho0) * sc.rho_scale);
            const __bf16 f1 = (__bf16)(vi.x * sc.inv_v);
            const __bf16 f2 = (__bf16)(vi.y * sc.inv_v);
            const __bf16 f3 = (__bf16)(vi.z * sc.inv_v);
            const uint32_t one = (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)1.0f);
            xt0.x = (uint32_t)__builtin_bit_cast(uint16_t, f0) |
                    ((uint32_t)__builtin_bit_cast(uint16_t, f1) << 16);
            xt0.y = (uint32_t)__builtin_bit_cast(uint16_t, f2) |
                    ((uint32_t)__builtin_bit_cast(uint16_t, f3) << 16);
            xt0.z = one;  // column 132: bias column of Wn1
            zt0.x = one;  // column 128 of Z: bias column of Wn2
        }
        *reinterpret_cast<uint4*>(X + row * kSurRow + 128) = xt0;
        *reinterpret_cast<uint4*>(X + row * kSurRow + 136) = xt1;
        *reinterpret_cast<uint4*>(Z + row * kSurRow + 128) = zt0;
        *reinterpret_cast<uint4*>(Z + row * kSurRow + 136) = xt1;
        cnt[row] = active ? sc.msg_scale * ne : 0.0f;
        rowslot[row] = active ? s : kSurNoSlot;
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
// edge pass (backward): d(W1|b1) = sum_i sum_j (dS_i * [u_ij > 0]) (e_ij | 1)^T.  The
// masked gradient is formed in the accumulator layout of the recomputed layer-1 product
// and fed back as an A operand (K = edges); the edge features reach the B operand through
// LDS.  One wavefront per batch chunk.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_sur_edges_bwd(
    WalkBuffers wb, const uint32_t* __restrict__ batch, uint32_t nbatch,
    const float4* __restrict__ spm, const float4* __restrict__ svr,
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ cs, StepConsts c,
    const SceneGrid* __restrict__ scenes, const __bf16* __restrict__ pk, SurScales sc,
    const __bf16* __restrict__ DS, float* __restrict__ part) {
    __shared__ uint32_t lmask[kEdgeSeg * 64];
    __shared__ uint32_t lwj0[kEdgeSeg];
    __shared__ __bf16 le[64 * 16];
    const uint32_t lane = threadIdx.x;
    const uint32_t b = blockIdx.x;
    if (b >= nbatch) return;
    const uint32_t item = uni(batch[b]);
    Chunk ch;
    uint32_t s;
    bool active;
    float4 pi;
    if (!load_chunk(item, wb.items, wb.nitems, skey, cs, spm, c, scenes, lane, ch, s, active,
                    pi))
        return;
    const float4 vi = active ? svr[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    Layer1B B;
    load_layer1(pk, lane, B);
    const uint32_t col = lane & 31u, h = lane >> 5;
    float ds[2][4][16];
#pragma unroll
    for (int blk = 0; blk < 2; ++blk)
#pragma unroll
        for (int hb = 0; hb < 4; ++hb)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const size_t row = 64u * b + 32u * blk + acc_row(v, h);
                ds[blk][hb][v] = (float)DS[row * 128 + 32u * hb + col];
            }
    f32x16 accw[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) accw[hb] = zero16();
    for_each_edge<kEdgeSeg>(
        item, ch, s, active, pi, wb, spm, svr, cs, c, lmask, lwj0, lane,
        [&](bool edge, float4 pj, float4 vj, float d2) {
            const bf16x8 fe = edge_features(edge, pi, vi, pj, vj, d2, sc);
            bf16x8 ext = zero8();
            ext[0] = (__bf16)(edge ? 1.0f : 0.0f);
            *reinterpret_cast<bf16x8*>(le + lane * 16) = fe;
            *reinterpret_cast<bf16x8*>(le + lane * 16 + 8) = ext;
            wave_sync();
            bf16x8 a[2];
            layer1_a(fe, edge, lane, a[0], a[1]);
#pragma unroll
            for (int blk = 0; blk < 2; ++blk) {
                bf16x8 e[2];
#pragma unroll
                for (int g = 0; g < 2; ++g)
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const uint32_t er = 32u * blk + 16u * g + acc_row(i, h);
                        e[g][i] = col < 9u ? le[er * 16 + col] : (__bf16)0.0f;
                    }
#pragma unroll
                for (int hb = 0; hb < 4; ++hb) {
                    const f32x16 d = mfma16(a[blk], blk == 0 ? B.blk0[hb] : B.blk1[hb], zero16());
                    bf16x8 g0, g1;
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        g0[i] = (__bf16)(d[i] > 0.0f ? ds[blk][hb][i] : 0.0f);
                        g1[i] = (__bf16)(d[8 + i] > 0.0f ? ds[blk][hb][8 + i] : 0.0f);
                    }
                    accw[hb] = mfma16(g0, e[0], accw[hb]);
                    accw[hb] = mfma16(g1, e[1], accw[hb]);
                }
            }
            wave_sync();
        });
    if (col < 16u) {
#pragma unroll
        for (int hb = 0; hb < 4; ++hb)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint32_t m = 32u * hb + acc_row(v, h);
                part[((size_t)b * 128u + m) * 16u + col] = accw[hb][v];
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

__global__ __launch_bounds__(256) void k_sur_integrate(
    const float* __restrict__ out, const float4* __restrict__ spm,
    const float4* __restrict__ svr, const uint32_t* __restrict__ sid,
    const uint32_t* __restrict__ cs, uint32_t n, StepConsts c,
    const SceneGrid* __restrict__ scenes, float T, float dk, float acc_scale,
    float4* __restrict__ pm, float4* __restrict__ vel, float4* __restrict__ acc,
    float* __restrict__ rho, float* __restrict__ prs) {
#pragma clang fp contract(off)
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= n) return;
    const uint32_t first_nonfinite = cs[c.G];
    const float4 p = spm[s];
    const float4 vr = svr[s];
    const uint32_t id = sid[s];
    float3 a;
    float rho_i;
    if (s < first_nonfinite) {
        a = make_float3(acc_scale * out[4 * (size_t)s], acc_scale * out[4 * (size_t)s + 1],
                        acc_scale * out[4 * (size_t)s + 2]);
        rho_i = vr.w;
    } else {  // no neighbours, not even itself: free fall, as in the plain step
        a = make_float3(0.0f, c.grav, 0.0f);
        rho_i = p.w * c.w0;
    }
    float3 v1 = make_float3(dk * vr.x + a.x * T, dk * vr.y + a.y * T, dk * vr.z + a.z * T);
    float4 q = make_float4(p.x + (0.5f * T) * (vr.x + v1.x), p.y + (0.5f * T) * (vr.y + v1.y),
                           p.z + (0.5f * T) * (vr.z + v1.z), p.w);
    const SceneGrid& g = scenes[scene_of_particle(scenes, c.nscenes, id)];
    const float wd = -0.8f;
    if (q.x < g.lo[0]) { q.x = g.lo[0]; v1.x = v1.x * wd; } else if (q.x > g.hi[0]) { q.x = g.hi[0]; v1.x = v1.x * wd; }
    if (q.y < g.lo[1]) { q.y = g.lo[1]; v1.y = v1.y * wd; } else if (q.y > g.hi[1]) { q.y = g.hi[1]; v1.y = v1.y * wd; }
    if (q.z < g.lo[2]) { q.z = g.lo[2]; v1.z = v1.z * wd; } else if (q.z > g.hi[2]) { q.z = g.hi[2]; v1.z = v1.z * wd; }
    pm[id] = q;
    vel[id] = make_float4(v1.x, v1.y, v1.z, 0.0f);
    acc[id] = make_float4(a.x, a.y, a.z, 0.0f);
    rho[id] = rho_i;
    prs[id] = c.k_gas * (rho_i - c.rho0);
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

hipError_t launch_sur_edges(const WalkBuffers& wb, const uint32_t* batch, uint32_t nbatch,
                            uint32_t max_items, const float4* spm, const float4* svr,
                            const uint32_t* skey, const uint32_t* cs, const StepConsts& c,
                            const SceneGrid* scenes, const __bf16* pk, const float* pf,
                            const SurScales& sc, __bf16* S, __bf16* X, __bf16* Z, float* cnt,
                            uint32_t* rowslot, hipStream_t st) {
    const uint32_t waves = batch ? nbatch : max_items;
    if (waves == 0) return hipSuccess;
    uint32_t blocks = (waves + kEdgeWaves - 1) / kEdgeWaves;
    blocks = (blocks + kXcds - 1) / kXcds * kXcds;
    hipLaunchKernelGGL(k_sur_edges, dim3(blocks), dim3(64 * kEdgeWaves), 0, st, wb, batch,
                       nbatch, spm, svr, skey, cs, c, scenes, pk, pf, sc, S, X, Z, cnt, rowslot);
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
    hipLaunchKernelGGL(k_sur_edges_bwd, dim3(nbatch), dim3(64), 0, st, wb, batch, nbatch, spm,
                       svr, skey, cs, c, scenes, pk, sc, DS, part);
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

hipError_t launch_sur_integrate(const float* out, const float4* spm, const float4* svr,
                                const uint32_t* sid, const uint32_t* cs, uint32_t n,
                                const StepConsts& c, const SceneGrid* scenes, float T, float dk,
                                float acc_scale, float4* pm, float4* vel, float4* acc,
                                float* rho, float* prs, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_integrate, dim3((n + 255) / 256), dim3(256), 0, st, out, spm, svr,
                       sid, cs, n, c, scenes, T, dk, acc_scale, pm, vel, acc, rho, prs);
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
