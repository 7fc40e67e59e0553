// sphk_walk.hip -- the two neighbour walks of the SPH step (density; forces + integration).
//
// Work item ("chunk"): up to 64 consecutive cell-sorted particles of one (x,y) column of
// the dense grid, one per lane of a 64-wide wavefront.  Within a column the sort order is
// ascending z cell, so a chunk covers z cells [zmin, zmax] and all of its candidate
// neighbours lie in the 3x3 neighbouring columns between z cells zmin-1 and zmax+1 (widened
// to +-2 on an axis when `axis_edge_flags` proves a pair two cells apart possible for some
// lane of the chunk).  Each column contributes one contiguous range of sorted slots.
//
// Phase 1 (test): the chunk's candidate words (32 consecutive slots of one column) are
//   visited in the cursor's order (parts, passes, columns: see kPasses), and each lane gets
//   one bit per candidate from the conservative integer filter: on the matrix cores
//   (mfma_filter_word, 32 candidates x 64 targets per word) where the chunk's coordinates fit
//   its exact range, else one candidate at a time for the whole wave (push_quant, candidate
//   data through the scalar cache).  `wj0[w]` is the sorted slot of the first candidate of
//   word w.  Segments hold up to kSeg - 1 words and never span two parts.
// Phase 2 (evaluate): every lane walks only its own set bits and evaluates the kernel for
//   those pairs (pairs the filter let through beyond the radius weigh exactly zero).
// The density pass stores the words of each chunk in HBM (masks, wj0, word count, part
// starts) and the force rows of each particle; the force pass reads them back instead of
// testing again.  Chunks with more words than the engine stores (StepConsts::max_words: 256,
// or 128 for engines whose word buffers would pass 16 GiB; pile-ups of thousands of particles
// per cell) are re-tested by the force pass.  The neighbour-list export walks the reference's
// (dx, dy, dz, id) order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "sphk_walk_common.h"

namespace sphk {

// ---------------------------------------------------------------------------
// chunk list: per column, ceil(len/64) chunks of consecutive sorted slots
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_chunk_count(const uint32_t* __restrict__ cs,
                                                     StepConsts c,
                                                     const SceneGrid* __restrict__ scenes,
                                                     uint32_t* __restrict__ cnt) {
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col > c.ncols) return;
    if (col == c.ncols) {
        cnt[col] = 0;
        return;
    }
    const SceneGrid g = scenes[scene_of_col(scenes, c.nscenes, col)];
    const uint32_t first = g.cell_base + (col - g.col_base) * (uint32_t)g.nz;
    const uint32_t a = cs[first], b = cs[first + (uint32_t)g.nz];
    cnt[col] = (b - a + 63u) / 64u;
}

__global__ __launch_bounds__(256) void k_chunk_fill(const uint32_t* __restrict__ cs, StepConsts c,
                                                    const SceneGrid* __restrict__ scenes,
                                                    const uint32_t* __restrict__ offs,
                                                    uint32_t* __restrict__ items,
                                                    uint32_t* __restrict__ nitems) {
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col == 0) nitems[0] = offs[c.ncols];
    if (col >= c.ncols) return;
    const SceneGrid g = scenes[scene_of_col(scenes, c.nscenes, col)];
    const uint32_t first = g.cell_base + (col - g.col_base) * (uint32_t)g.nz;
    const uint32_t a = cs[first], b = cs[first + (uint32_t)g.nz];
    uint32_t o = offs[col];
    for (uint32_t s = a; s < b; s += 64u) items[o++] = s;
}

uint32_t max_chunks(uint32_t n, const StepConsts& c) { return (n + 63u) / 64u + c.ncols; }

// ---------------------------------------------------------------------------
// Launch shapes of the walks: words per LDS segment x wavefronts per workgroup.  Every
// wavefront works on its own chunk; the waves of a workgroup never synchronise.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// density: rho_i = m_i W(0) + sum_{j != i} m_j W(r_ij)   (compute_density,
// src/sph_engine.cpp:370-383).  The self pair is one of the words' bits (d = 0), so the sum
// runs over the set bits, self included.  Stores the neighbour words of every chunk and the
// force walk's neighbour rows of every particle.
// ---------------------------------------------------------------------------
// The cubic spline (CubicSplineKernel::W_3d, src/kernels.cpp:58-66) as one expression:
// 6 W(q) / sigma = (2 - q)^3 - 4 (1 - q)^3 for q <= 1
// and (2 - q)^3 for 1 < q <= 2, i.e. t^3 - 4 s^3 with t = max(2 - q, 0), s = max(1 - q, 0).
// The walk sums an eighth of it, u^3 + s^2 (-s / 2) with u = t / 2 = max(1 - q / 2, 0): every
// rounding is that of the t-form scaled by 2^-3 (terms below 2^-126 aside), so the sums are
// the same bits / 8; u and s lie in [0, 1], where the max is the instructions' clamp bit
// (seven operations, no select).  The sum is scaled by 8 sigma / 6 at the end.
__device__ __forceinline__ float w6_shape8(float r, float half_inv_h, float inv_h) {
    const float u = __builtin_amdgcn_fmed3f(__builtin_fmaf(-half_inv_h, r, 1.0f), 0.0f, 1.0f);
    const float s = __builtin_amdgcn_fmed3f(__builtin_fmaf(-inv_h, r, 1.0f), 0.0f, 1.0f);
    return __builtin_fmaf(s * s, -0.5f * s, u * u * u);
}

// One neighbour's m_j (6 W / 8 sigma).  Pairs the filter let through beyond the radius: with
// a radius of at least 2h they have q >= 2 (up to the rounding of sqrt, which leaves a shape
// below 2^-75) and weigh nothing, so d^2 may be formed with fused multiply-adds; with a smaller
// radius (kCut) the reference's test d^2 <= r^2, on the reference's unfused d^2, drops them.
// (The neighbour count, which needs the reference's test on every pair, is taken by the force
// walk: this walk is bound by VALU issue, the force walk by its row gathers.)
template <bool kCut>
__device__ __forceinline__ float density_pair(const float4& pi, const float4& q,
                                              const StepConsts& c, float half_inv_h, float acc) {
    float d2;
    if (kCut) {
        d2 = dist2_exact(pi.x, pi.y, pi.z, q.x, q.y, q.z);
    } else {
        const float dx = pi.x - q.x, dy = pi.y - q.y, dz = pi.z - q.z;
        d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    }
    float w = w6_shape8(__builtin_amdgcn_sqrtf(d2), half_inv_h, c.inv_h);
    if (kCut) w = d2 <= c.r2 ? w : 0.0f;
    return __builtin_fmaf(q.w, w, acc);
}

// A value computed here, before whatever follows.  Pinning each pair's sum keeps it ahead of
// the next step's refill branch, so the row it used is dead when the row after it is requested
// and the two share registers; left to the compiler, the sums sink past the branches and the
// loop copies a row set at its back edge.  Measured on the 1M dam break: density walk 0.429
// -> 0.423 ms; at 10,000 particles the split shapes' walks (both pinned) 1-2 % faster.  The
// ping-pong force walk keeps its sums unpinned: its steps wait on row gathers, and pinned sums
// (which hold the next request back until they are formed) took it 0.581 -> 0.594 ms.
__device__ __forceinline__ float pinned(float x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Neighbour positions are requested two steps ahead, in three rotating register sets.
template <bool kCut, int kSeg>
__device__ __forceinline__ float density_bits(BitWalk<kSeg>& bw, uint32_t steps,
                                              const uint32_t* __restrict__ lmask,
                                              const uint32_t* __restrict__ lwj0, uint32_t lane,
                                              const float4* __restrict__ spm, const float4& pi,
                                              const StepConsts& c, float acc) {
    if (steps == 0) return acc;
    const float hh = 0.5f * c.inv_h;
    float4 q0 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
    float4 q1 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
    uint32_t t = 0;
    for (; t + 3 <= steps; t += 3) {
        const float4 q2 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
        acc = pinned(density_pair<kCut>(pi, q0, c, hh, acc));
        q0 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
        acc = pinned(density_pair<kCut>(pi, q1, c, hh, acc));
        q1 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
        acc = pinned(density_pair<kCut>(pi, q2, c, hh, acc));
    }
    if (t < steps) acc = density_pair<kCut>(pi, q0, c, hh, acc);
    if (t + 1 < steps) acc = density_pair<kCut>(pi, q1, c, hh, acc);
    return acc;
}

// The same walk with kAhead rows in flight, for the split shapes, whose wavefronts are few
// (below one per SIMD at 10,000 particles) and hide little of each other's gather latency.
// Rows rotate through kAhead + 1 register sets, each refilled right after its row is used;
// the pairs are summed in the same order as above.  Measured at 10,000 particles: each walk
// ~1 us (2 %) faster than with two rows in flight, so the small-scene walks are not bound by
// gather latency either.
template <bool kCut, int kSeg, int kAhead>
__device__ __forceinline__ float density_bits_deep(BitWalk<kSeg>& bw, uint32_t steps,
                                                   const uint32_t* __restrict__ lmask,
                                                   const uint32_t* __restrict__ lwj0,
                                                   uint32_t lane, const float4* __restrict__ spm,
                                                   const float4& pi, const StepConsts& c,
                                                   float acc) {
    if (steps == 0) return acc;
    const float hh = 0.5f * c.inv_h;
    float4 q[kAhead + 1];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) q[k] = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
    uint32_t t = 0;
    for (; t + kAhead + 1 <= steps; t += kAhead + 1) {
#pragma unroll
        for (int k = 0; k <= kAhead; ++k) {
            q[(k + kAhead) % (kAhead + 1)] = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
            acc = pinned(density_pair<kCut>(pi, q[k], c, hh, acc));
        }
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k)
        if (t + k < steps) acc = density_pair<kCut>(pi, q[k], c, hh, acc);
    return acc;
}

// Rows in flight per wavefront in the split shapes' walks, and per lane in the ping-pong force
// walk.  Measured on the 1M dam break (pp16_waves4): 2 rows 0.578 ms, 3 rows 0.574, 4 rows
// 0.566, 5 rows 0.561, 6 rows 0.560 (158 VGPRs: three wavefronts per SIMD where the LDS allows
// four), 8 rows 0.624 (two wavefronts per SIMD): the force walk is bound by its gathers'
// latency more than by wavefronts in flight.
constexpr int kDeepAhead = 5;
constexpr int kPPAhead = 6;

// The ping-pong shapes' walk, which goes on across calls: on entry and on return the rows of
// the lane's next two pairs are in flight in q0 and q1, and a call takes exactly `steps` bits
// and sums exactly `steps` pairs (the two in flight at a part's end are summed by the caller).
template <bool kCut, int kW>
__device__ __forceinline__ float density_run(BitWalk<kW>& bw, float4& q0, float4& q1,
                                             uint32_t steps, const uint32_t* __restrict__ lmask,
                                             const uint32_t* __restrict__ lwj0, uint32_t lane,
                                             const float4* __restrict__ spm, const float4& pi,
                                             const StepConsts& c, float acc) {
    const float hh = 0.5f * c.inv_h;
    uint32_t t = 0;
    for (; t + 3 <= steps; t += 3) {
        const float4 q2 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
        acc = pinned(density_pair<kCut>(pi, q0, c, hh, acc));
        q0 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
        acc = pinned(density_pair<kCut>(pi, q1, c, hh, acc));
        q1 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
        acc = pinned(density_pair<kCut>(pi, q2, c, hh, acc));
    }
    for (; t < steps; ++t) {
        const float4 q2 = row_at(spm, bw.template next<4>(lmask, lwj0, lane));
        acc = density_pair<kCut>(pi, q0, c, hh, acc);
        q0 = q1;
        q1 = q2;
    }
    return acc;
}

// Density, the force rows of the particle (B = m/rho and C = B P / 4, so that
// m_j (P_i + P_j) / (2 rho_j) = 2 ((P_i / 4) B_j + C_j); the 2 rides on dW/dq) and rho in
// svr.w.
__device__ __forceinline__ void density_epilogue(float sum, const float4& pi, uint32_t s,
                                                 const StepConsts& c, float4* __restrict__ svr,
                                                 ForceRow* __restrict__ rows) {
    const float rho = c.sigma * (sum * (8.0f / 6.0f));  // 8.0f / 6.0f == 8 * (1.0f / 6.0f)
    float4 v = svr[s];
    v.w = rho;
    svr[s] = v;
    const float B = pi.w / rho;
    const float P = c.k_gas * (rho - c.rho0);
    rows[s].a = make_float4(pi.x, pi.y, pi.z, B);
    rows[s].b = make_float4(v.x, v.y, v.z, 0.25f * B * P);
}

// Density pass over one chunk (kSplit false: all parts; true: the part of this wavefront,
// three wavefronts per chunk, the sums finished by k_density_finish).  Seven wavefronts per
// SIMD: the 15-word shapes' LDS allows it, and the walk is VALU-bound (a few spills outside
// the walk loops; measured faster at 1M than the 6 the default register allocation gives).
// The split shapes serve small scenes, whose wavefronts are few: no occupancy target, so
// their deeper read-ahead stays in registers.  The ping-pong shapes aim at what their LDS
// allows (6 with two 12-word buffers, 4 with two 16-word ones).
template <class Shape, bool kSplit, bool kCut>
__global__ __launch_bounds__(Shape::threads) __attribute__((amdgpu_waves_per_eu(kSplit ? 1 : Shape::density_waves, 8))) void k_density_walk(
    const uint32_t* __restrict__ items, const uint32_t* __restrict__ nitems,
    const float4* __restrict__ spm, const uint2* __restrict__ sq, float4* __restrict__ svr,
    ForceRow* __restrict__ rows, const uint32_t* __restrict__ skey,
    const uint32_t* __restrict__ cs, StepConsts c, const SceneGrid* __restrict__ scenes,
    uint32_t* __restrict__ gmask, uint32_t* __restrict__ gwj0, uint32_t* __restrict__ gnw,
    uint32_t* __restrict__ gpart, PartSums ps) {
    constexpr int kSeg = Shape::seg;
    constexpr int kLds = Shape::lds_words;
    static_assert(!(kSplit && Shape::pp), "ping-pong shapes walk a chunk with one wavefront");
    __shared__ uint32_t lmask_all[Shape::waves][kLds * 64];
    __shared__ uint32_t lwj0_all[Shape::waves][kLds];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* lmask = lmask_all[wave];
    uint32_t* lwj0 = lwj0_all[wave];
    const uint32_t units = kSplit ? uni(nitems[0]) * kParts : uni(nitems[0]);
    const uint32_t wid = xcd_unit(units, Shape::waves, wave);
    if (wid == ~0u) return;
    const uint32_t item = kSplit ? wid / kParts : wid;
    const int part = kSplit ? (int)(wid % kParts) : -1;
    Chunk ch;
    uint32_t s;
    bool active;
    float4 pi;
    if (!load_chunk(item, items, nitems, skey, cs, spm, c, scenes, lane, ch, s, active, pi)) return;
    const uint2 qi = active ? sq[s] : make_uint2(0u, 0u);
    init_sentinel<kLds>(lmask, lwj0, lane, c);
    MfmaFilter mf;
    mfma_filter_begin(mf, ch, sq, qi, lane, c.qneg);
    WordTable wt;
    const bool tab = word_table_build(wt, ch, cs, c, part, lane);
    if (lane == 0 && part <= 0) {
        gpart[item] = pack_parts(wt.ps1, wt.ps2);
        gnw[item] = (uint32_t)wt.total;
    }
    Cursor cu;
    if (!tab) cursor_begin(cu, ch, cs, c, 0, part);
    // per-part running sum `cur`, parts folded into `total` as (p0 + p1) + p2
    float total = 0.0f, cur = 0.0f;
    int cur_part = 0;
    int seg_base = wt.first;
    int ti = 0;  // next table entry
    if constexpr (Shape::pp) {
        // Segments are made in the newer buffer.  A segment of the part being walked joins
        // the walk at once; one of the next part waits there until every lane is through the
        // current part (the parts' sums are folded between them, in the same order as in
        // every other shape).
        uint32_t* const lm_new = lmask + kSeg * 64;
        uint32_t* const lj_new = lwj0 + kSeg;
        const float hh = 0.5f * c.inv_h;
        PingPong<kSeg> pp;
        pp.init(c.pad_slot);
        float4 q0, q1;  // rows of the next two pairs, requested (within a phase)
        bool more = true, pend = false;  // pend: the newer buffer holds the next part's segment
        int pend_part = 0, pend_nw = 0, walk_part = -1;
        for (;;) {
            if (pp.nw_new == 0 && !pend && more) {
                SegLane sl;
                int seg_part;
                const int nw = tab ? phase1_table<kSeg>(wt, ti, sq, c, qi, active, lm_new, lj_new,
                                                        lane, sl, seg_part, mf)
                                   : phase1_segment<kSeg, true>(cu, ch, sq, cs, c, qi, active,
                                                                lm_new, lj_new, lane, sl,
                                                                seg_part, &mf);
                if (nw == 0) {
                    more = false;
                } else {
                    wave_sync();
                    store_words(item, c.max_words, seg_base, nw, lm_new, lj_new, gmask, gwj0,
                                lane);
                    seg_base += nw;
                    if (tab ? ti >= wt.n : cu.done) more = false;
                    if (walk_part < 0 || seg_part == walk_part) {
                        if (walk_part < 0) {  // the walk's first segment
                            for (; cur_part < seg_part; ++cur_part) {
                                total = cur_part == 0 ? cur : total + cur;
                                cur = 0.0f;
                            }
                            walk_part = seg_part;
                        }
                        pp.add_newer(sl, nw, lmask, lwj0, lane, c.pad_slot);
                    } else {
                        pend = true;
                        pend_part = seg_part;
                        pend_nw = nw;
                    }
                }
            }
            // A phase: walk until every lane is through the older buffer (or, with nothing to
            // overlap, through all it holds).  Its read-ahead starts at the phase's start and its
            // last two rows are summed at its end, in walk order; nothing stays in flight while
            // the next segment is made.
            const bool overlap = pp.nw_new != 0;
            if (!overlap && walk_part < 0) break;  // no segment at all
            if ((overlap ? pp.older_steps() : pp.all_steps()) != 0) {
                q0 = row_at(spm, pp.bw.template next<4>(lmask, lwj0, lane));
                q1 = row_at(spm, pp.bw.template next<4>(lmask, lwj0, lane));
                pp.walked(2);
                const uint32_t steps = overlap ? pp.older_steps() : pp.all_steps();
                cur = density_run<kCut>(pp.bw, q0, q1, steps, lmask, lwj0, lane, spm, pi, c, cur);
                pp.walked(steps);
                cur = density_pair<kCut>(pi, q0, c, hh, cur);
                cur = density_pair<kCut>(pi, q1, c, hh, cur);
            }
            if (overlap) {
                pp.shift_down(lmask, lwj0, lane);
                continue;
            }
            if (!pend) break;
            for (; cur_part < pend_part; ++cur_part) {
                total = cur_part == 0 ? cur : total + cur;
                cur = 0.0f;
            }
            walk_part = pend_part;
            pend = false;
            pp.add_held(pend_nw, lmask, lwj0, lane, c.pad_slot);
        }
    } else {
        for (;;) {
            SegLane sl;
            int seg_part;
            const int nw =
                tab ? phase1_table<kSeg>(wt, ti, sq, c, qi, active, lmask, lwj0, lane, sl, seg_part,
                                         mf)
                    : phase1_segment<kSeg, true>(cu, ch, sq, cs, c, qi, active, lmask, lwj0, lane,
                                                 sl, seg_part, &mf);
            if (nw == 0) break;
            if (!kSplit) {
                for (; cur_part < seg_part; ++cur_part) {
                    total = cur_part == 0 ? cur : total + cur;
                    cur = 0.0f;
                }
            }
            wave_sync();
            store_words(item, c.max_words, seg_base, nw, lmask, lwj0, gmask, gwj0, lane);
            BitWalk<kSeg> bw;
            bw.begin(sl.nonempty, lmask, lwj0, lane);
            const uint32_t steps = wave_max(sl.count);
            if constexpr (kSplit)
                cur = density_bits_deep<kCut, kSeg, kDeepAhead>(bw, steps, lmask, lwj0, lane, spm,
                                                                pi, c, cur);
            else
                cur = density_bits<kCut>(bw, steps, lmask, lwj0, lane, spm, pi, c, cur);
            seg_base += nw;
            wave_sync();
            if (tab ? ti >= wt.n : cu.done) break;
        }
    }
    if (kSplit) {
        if (active) {
            ps.acc[(size_t)part * ps.cap + s] = cur;
        }
        return;
    }
    for (; cur_part < kParts; ++cur_part) {
        total = cur_part == 0 ? cur : total + cur;
        cur = 0.0f;
    }
    if (active) density_epilogue(total, pi, s, c, svr, rows);
}

// Finishes split density passes: per finite particle (sorted slots below the non-finite
// bin), the three part sums in order, then the same epilogue as the single-wavefront pass.
__global__ __launch_bounds__(256) void k_density_finish(
    const float4* __restrict__ spm, float4* __restrict__ svr, ForceRow* __restrict__ rows,
    const uint32_t* __restrict__ cs, StepConsts c, PartSums ps) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= cs[c.G]) return;
    const float a0 = ps.acc[s], a1 = ps.acc[ps.cap + s], a2 = ps.acc[2 * ps.cap + s];
    density_epilogue((a0 + a1) + a2, spm[s], s, c, svr, rows);
}

// ---------------------------------------------------------------------------
// forces + leapfrog + wall clamp
//   F_p = -sum_{j != i, |r| >= 1e-6} m_j (P_i + P_j) / (2 rho_j) gradW(r_ij)
//   F_v =  sum_{j != i} (m_j / rho_j) mu (v_j - v_i) laplacianW(r_ij)
//   a_i = (F_p + F_v + g m_i) / m_i        (rho_i instead of m_i: consistent formulation)
// (src/sph_engine.cpp:390-443), integrate_leapfrog (:290-310) and
// apply_boundary_conditions (src/particle.cpp:122-153).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void integrate_clamp(float4& p, float3& v, float3 a, float dt,
                                                const StepConsts& c, const float* lo,
                                                const float* hi) {
#pragma clang fp contract(off)
    float hx = 0.5f * a.x, hy = 0.5f * a.y, hz = 0.5f * a.z;
    v.x = v.x + hx * dt;
    v.y = v.y + hy * dt;
    v.z = v.z + hz * dt;
    p.x = p.x + v.x * dt;
    p.y = p.y + v.y * dt;
    p.z = p.z + v.z * dt;
    v.x = v.x + hx * dt;
    v.y = v.y + hy * dt;
    v.z = v.z + hz * dt;
    v.x = v.x * c.damping;
    v.y = v.y * c.damping;
    v.z = v.z * c.damping;
    const float wd = -0.8f;
    if (p.x < lo[0]) {
        p.x = lo[0];
        v.x = v.x * wd;
    } else if (p.x > hi[0]) {
        p.x = hi[0];
        v.x = v.x * wd;
    }
    if (p.y < lo[1]) {
        p.y = lo[1];
        v.y = v.y * wd;
    } else if (p.y > hi[1]) {
        p.y = hi[1];
        v.y = v.y * wd;
    }
    if (p.z < lo[2]) {
        p.z = lo[2];
        v.z = v.z * wd;
    } else if (p.z > hi[2]) {
        p.z = hi[2];
        v.z = v.z * wd;
    }
}

__device__ __forceinline__ float3 final_accel(float3 fp, float3 fv, float m_i, float rho_i,
                                              const StepConsts& c) {
    // explicit fused forms: left to contraction, the choice of which product to fuse can
    // differ between the kernels that inline this, and so would their bits
    const float fx = __builtin_fmaf(fp.x, c.grad_scale, fv.x * c.lap_scale);
    const float fy = __builtin_fmaf(fp.y, c.grad_scale, fv.y * c.lap_scale);
    const float fz = __builtin_fmaf(fp.z, c.grad_scale, fv.z * c.lap_scale);
    const float m = c.formulation == 1 ? rho_i : m_i;
    const float inv_m = 1.0f / m;
    return make_float3(fx * inv_m, __builtin_fmaf(c.grav, m, fy) * inv_m, fz * inv_m);
}

// Pressure and viscosity sums of one target over neighbour rows (ForceRow a, b).
// The self pair (r = 0, v_j = v_i) and the padding slot (B = C = 0) add exactly zero.  Pairs
// the filter let through beyond the radius: with a radius of at least 2h they have q >= 2 (up
// to rounding, which leaves a Laplacian factor below 3e-8 of a neighbour's) and add nothing;
// with a smaller radius (kCut) the reference's test on its own d^2 drops them: d^2 is pushed
// to 1e30, where both kernel factors clamp to zero.
struct ForceSums {
    float3 fp = make_float3(0.f, 0.f, 0.f);
    float3 fv = make_float3(0.f, 0.f, 0.f);
    template <bool kCut>
    // `cnt` counts the pairs that pass the reference's test d^2 <= r^2 on the reference's
    // unfused d^2 (the particle itself included): the engine's neighbour statistics.
    __device__ __forceinline__ void add(const float4& pi, const float3& vi, float p_quarter,
                                        const float4& a, const float4& b, const StepConsts& c,
                                        uint32_t& cnt) {
        const float rx = pi.x - a.x, ry = pi.y - a.y, rz = pi.z - a.z;
        float d2 = dist2_exact(pi.x, pi.y, pi.z, a.x, a.y, a.z);
        const bool in = d2 <= c.r2;
        cnt += in ? 1u : 0u;
        if (kCut) d2 = in ? d2 : 1e30f;
        const float rinv = __builtin_amdgcn_rsqf(fmaxf(d2, 1e-30f));
        const float rl = d2 * rinv;
        // u = (2 - q) / 2 and s = 1 - q, clamped to [0, 1] by the instructions' clamp bit
        // (seven operations, no select): dwq = (s - u)(s + u) = s^2 - u^2 is q (3q - 4) / 4
        // for q <= 1 and -(2 - q)^2 / 4 beyond, proportional to dW/dq; lap = u - 2 s is
        // (3q - 2) / 2 and (2 - q) / 2.  grad_scale and lap_scale carry the constant factors.
        const float u =
            __builtin_amdgcn_fmed3f(__builtin_fmaf(-0.5f * c.inv_h, rl, 1.0f), 0.0f, 1.0f);
        const float s = __builtin_amdgcn_fmed3f(__builtin_fmaf(-c.inv_h, rl, 1.0f), 0.0f, 1.0f);
        const float dwq = (s - u) * (s + u);
        const float lap = __builtin_fmaf(-2.0f, s, u);
        // gradW is zero below |r| = 1e-6 (src/kernels.cpp:96-98)
        const float dw = rl >= 1e-6f ? dwq : 0.0f;
        const float coef = __builtin_fmaf(p_quarter, a.w, b.w) * dw * rinv;
        fp.x -= coef * rx;
        fp.y -= coef * ry;
        fp.z -= coef * rz;
        const float cv = a.w * lap;
        fv.x += cv * (b.x - vi.x);
        fv.y += cv * (b.y - vi.y);
        fv.z += cv * (b.z - vi.z);
    }
    // the sums (and the count) formed here, before whatever follows (see pinned)
    __device__ __forceinline__ void pin(uint32_t& cnt) {
        asm volatile("" : "+v"(fp.x), "+v"(fp.y), "+v"(fp.z), "+v"(fv.x), "+v"(fv.y),
                     "+v"(fv.z), "+v"(cnt));
    }
};

// The wavefront runs its largest lane count; rows are requested two neighbours ahead, in
// three rotating register sets (a gather from L2 takes several hundred cycles, one step of
// pair arithmetic about a hundred).
template <int kSeg, bool kCut>
__device__ __forceinline__ void force_bits(BitWalk<kSeg>& bw, uint32_t steps,
                                           const uint32_t* __restrict__ lmask,
                                           const uint32_t* __restrict__ lwj0, uint32_t lane,
                                           const ForceRow* __restrict__ rows, const float4& pi,
                                           const float3& vi, float p_quarter, const StepConsts& c,
                                           ForceSums& f, uint32_t& cnt) {
    if (steps == 0) return;
    static_assert(sizeof(ForceRow) == 32, "rows are 32-byte blocks");
    uint32_t o = bw.template next<5>(lmask, lwj0, lane);
    float4 a0 = row_at(rows, o).a, b0 = row_at(rows, o).b;
    o = bw.template next<5>(lmask, lwj0, lane);
    float4 a1 = row_at(rows, o).a, b1 = row_at(rows, o).b;
    uint32_t t = 0;
    for (; t + 3 <= steps; t += 3) {
        o = bw.template next<5>(lmask, lwj0, lane);
        const float4 a2 = row_at(rows, o).a, b2 = row_at(rows, o).b;
        f.template add<kCut>(pi, vi, p_quarter, a0, b0, c, cnt);
        o = bw.template next<5>(lmask, lwj0, lane);
        a0 = row_at(rows, o).a;
        b0 = row_at(rows, o).b;
        f.template add<kCut>(pi, vi, p_quarter, a1, b1, c, cnt);
        o = bw.template next<5>(lmask, lwj0, lane);
        a1 = row_at(rows, o).a;
        b1 = row_at(rows, o).b;
        f.template add<kCut>(pi, vi, p_quarter, a2, b2, c, cnt);
    }
    if (t < steps) f.template add<kCut>(pi, vi, p_quarter, a0, b0, c, cnt);
    if (t + 1 < steps) f.template add<kCut>(pi, vi, p_quarter, a1, b1, c, cnt);
}

// force_bits with kAhead rows in flight (the split shapes; see density_bits_deep).
template <int kSeg, bool kCut, int kAhead>
__device__ __forceinline__ void force_bits_deep(BitWalk<kSeg>& bw, uint32_t steps,
                                                const uint32_t* __restrict__ lmask,
                                                const uint32_t* __restrict__ lwj0, uint32_t lane,
                                                const ForceRow* __restrict__ rows,
                                                const float4& pi, const float3& vi,
                                                float p_quarter, const StepConsts& c,
                                                ForceSums& f, uint32_t& cnt) {
    if (steps == 0) return;
    float4 a[kAhead + 1], b[kAhead + 1];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
        const uint32_t o = bw.template next<5>(lmask, lwj0, lane);
        a[k] = row_at(rows, o).a;
        b[k] = row_at(rows, o).b;
    }
    uint32_t t = 0;
    for (; t + kAhead + 1 <= steps; t += kAhead + 1) {
#pragma unroll
        for (int k = 0; k <= kAhead; ++k) {
            const uint32_t o = bw.template next<5>(lmask, lwj0, lane);
            a[(k + kAhead) % (kAhead + 1)] = row_at(rows, o).a;
            b[(k + kAhead) % (kAhead + 1)] = row_at(rows, o).b;
            f.template add<kCut>(pi, vi, p_quarter, a[k], b[k], c, cnt);
            f.pin(cnt);
        }
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k)
        if (t + k < steps) f.template add<kCut>(pi, vi, p_quarter, a[k], b[k], c, cnt);
}

// force_bits for the ping-pong shapes: the walk goes on across calls with the rows of the
// lane's next kAhead pairs in flight (slots 0 .. kAhead - 1 on entry and on return, slot
// kAhead free); a call takes exactly `steps` bits and adds exactly `steps` pairs.  Each step
// requests into the free slot and sums the oldest, so over a block of kAhead + 1 steps the
// slots come back to their order and no row set is copied between blocks.
template <int kW, bool kCut, int kAhead>
__device__ __forceinline__ void force_run_n(BitWalk<kW>& bw, float4 (&a)[kAhead + 1],
                                            float4 (&b)[kAhead + 1], uint32_t steps,
                                            const uint32_t* __restrict__ lmask,
                                            const uint32_t* __restrict__ lwj0, uint32_t lane,
                                            const ForceRow* __restrict__ rows, const float4& pi,
                                            const float3& vi, float p_quarter,
                                            const StepConsts& c, ForceSums& f, uint32_t& cnt) {
    uint32_t t = 0;
    for (; t + kAhead + 1 <= steps; t += kAhead + 1) {
#pragma unroll
        for (int k = 0; k <= kAhead; ++k) {
            const uint32_t o = bw.template next<5>(lmask, lwj0, lane);
            a[(k + kAhead) % (kAhead + 1)] = row_at(rows, o).a;
            b[(k + kAhead) % (kAhead + 1)] = row_at(rows, o).b;
            f.template add<kCut>(pi, vi, p_quarter, a[k], b[k], c, cnt);
        }
    }
    for (; t < steps; ++t) {
        const uint32_t o = bw.template next<5>(lmask, lwj0, lane);
        a[kAhead] = row_at(rows, o).a;
        b[kAhead] = row_at(rows, o).b;
        f.template add<kCut>(pi, vi, p_quarter, a[0], b[0], c, cnt);
#pragma unroll
        for (int k = 0; k < kAhead; ++k) {
            a[k] = a[k + 1];
            b[k] = b[k + 1];
        }
    }
}

// Folds the sums of finished parts into the running total, which starts at +0:
// ((0 + part 0) + part 1) + part 2.  Closing `n` parts at once adds `f` for the first and
// +0 for each further (empty) one, as the three-wavefront finish does.  The single-wavefront
// walk keeps the total in LDS (`t`, six rows of 64 lanes): six more live registers would cost
// it a wavefront per SIMD.
__device__ __forceinline__ void fold_parts(float* __restrict__ t, uint32_t lane, ForceSums& f,
                                           int n) {
    if (n <= 0) return;
    const float v[6] = {f.fp.x, f.fp.y, f.fp.z, f.fv.x, f.fv.y, f.fv.z};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float x = t[k * 64 + lane] + v[k];
        if (n >= 2) x = x + 0.0f;
        t[k * 64 + lane] = x;
    }
    f = ForceSums();
}

// The same fold with the running total in registers (shapes whose occupancy LDS limits).
__device__ __forceinline__ void fold_parts_reg(float (&t)[6], ForceSums& f, int n) {
    if (n <= 0) return;
    const float v[6] = {f.fp.x, f.fp.y, f.fp.z, f.fv.x, f.fv.y, f.fv.z};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float x = t[k] + v[k];
        if (n >= 2) x = x + 0.0f;
        t[k] = x;
    }
    f = ForceSums();
}

// Acceleration, leapfrog, clamp and the caller-order writes of one target.
__device__ __forceinline__ void force_epilogue(const ForceSums& f, float4 pi, float3 vi,
                                               float rho_i, float P_i, uint32_t s,
                                               const SceneGrid& g, const StepConsts& c,
                                               const uint32_t* __restrict__ sid,
                                               const float* __restrict__ dt_dev, float dt_arg,
                                               float4* __restrict__ pm, float4* __restrict__ vel,
                                               float4* __restrict__ acc_out,
                                               float* __restrict__ rho_out,
                                               float* __restrict__ prs_out) {
    const float3 a = final_accel(f.fp, f.fv, pi.w, rho_i, c);
    const float dt = dt_dev ? *dt_dev : dt_arg;
    float3 v = vi;
    float4 p = pi;
    integrate_clamp(p, v, a, dt, c, g.lo, g.hi);
    const uint32_t id = sid[s];
    pm[id] = p;
    vel[id] = make_float4(v.x, v.y, v.z, 0.0f);
    acc_out[id] = make_float4(a.x, a.y, a.z, 0.0f);
    rho_out[id] = rho_i;
    prs_out[id] = P_i;
}

// Force pass over one chunk (kSplit false: all parts; true: the part of this wavefront,
// three wavefronts per chunk, finished by k_force_finish).  Shapes of 24 words and more keep
// the parts' running total in registers: their LDS, not their registers, limits occupancy,
// and the 1.5 KB per wavefront it frees lets the 28-word shape run 5 wavefronts per SIMD.
template <class Shape, bool kSplit, bool kCut>
__global__ __launch_bounds__(Shape::threads) __attribute__((amdgpu_waves_per_eu(Shape::force_waves)))
void k_force_walk(
    const uint32_t* __restrict__ items, const uint32_t* __restrict__ nitems,
    const float4* __restrict__ spm, const uint2* __restrict__ sq,
    const float4* __restrict__ svr, const ForceRow* __restrict__ rows,
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sid,
    const uint32_t* __restrict__ cs, StepConsts c, const SceneGrid* __restrict__ scenes,
    const uint32_t* __restrict__ gmask, const uint32_t* __restrict__ gwj0,
    const uint32_t* __restrict__ gnw, const uint32_t* __restrict__ gpart, PartSums ps,
    const float* __restrict__ dt_dev, float dt_arg,
    float4* __restrict__ pm,
    float4* __restrict__ vel, float4* __restrict__ acc_out, float* __restrict__ rho_out,
    float* __restrict__ prs_out, unsigned long long* __restrict__ stats) {
    constexpr int kSeg = Shape::seg;
    constexpr int kLds = Shape::lds_words;
    static_assert(!(kSplit && Shape::pp), "ping-pong shapes walk a chunk with one wavefront");
    __shared__ uint32_t lmask_all[Shape::waves][kLds * 64];
    __shared__ uint32_t lwj0_all[Shape::waves][kLds];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    constexpr bool kRegTot = !kSplit && !Shape::pp && kSeg >= 24;
    constexpr bool kLdsTot = !kSplit && !kRegTot;
    __shared__ float ltot_all[kLdsTot ? Shape::waves : 1][kLdsTot ? 6 * 64 : 1];
    float rtot[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t* lmask = lmask_all[wave];
    uint32_t* lwj0 = lwj0_all[wave];
    float* ltot = ltot_all[kLdsTot ? wave : 0];
    const uint32_t units = kSplit ? uni(nitems[0]) * kParts : uni(nitems[0]);
    const uint32_t wid = xcd_unit(units, Shape::waves, wave);
    if (wid == ~0u) return;
    const uint32_t item = kSplit ? wid / kParts : wid;
    const int part = kSplit ? (int)(wid % kParts) : -1;
    Chunk ch;
    uint32_t s;
    bool active;
    float4 pi;
    if (!load_chunk(item, items, nitems, skey, cs, spm, c, scenes, lane, ch, s, active, pi)) return;
    const uint2 qi = active ? sq[s] : make_uint2(0u, 0u);
    init_sentinel<kLds>(lmask, lwj0, lane, c);
    if (kLdsTot) {
#pragma unroll
        for (int k = 0; k < 6; ++k) ltot[k * 64 + lane] = 0.0f;
    }
    const float4 vr = active ? svr[s] : make_float4(0.f, 0.f, 0.f, 1.f);
    const float3 vi = make_float3(vr.x, vr.y, vr.z);
    const float rho_i = vr.w;
    const float P_i = c.k_gas * (rho_i - c.rho0);
    const float p_quarter = 0.25f * P_i;
    ForceSums f;
    uint32_t cnt = 0;  // neighbours of this wavefront's part(s)
    int cur_part = 0;
    const int total_words = (int)uni(gnw[item]);
    const bool stored = total_words <= (int)c.max_words;
    const uint32_t pstarts = stored ? uni(gpart[item]) : 0u;
    Cursor cu;
    if (!stored) cursor_begin(cu, ch, cs, c, 0, part);
    // stored words of this wavefront: all, or those of its part
    int seg_base = 0, words_end = total_words;
    if (kSplit) {
        seg_base = part_word_start(pstarts, part);
        words_end = part_word_end(pstarts, part, total_words);
    }
    if constexpr (Shape::pp) {
        // As in the density walk: segments are loaded into the newer buffer; one of the part
        // being walked joins the walk at once, one of the next part waits until every lane
        // is through the current part.
        uint32_t* const lm_new = lmask + kSeg * 64;
        uint32_t* const lj_new = lwj0 + kSeg;
        PingPong<kSeg> pp;
        pp.init(c.pad_slot);
        float4 a[kPPAhead + 1], b[kPPAhead + 1];  // rows requested ahead (within a phase)
        bool more = true, pend = false;  // pend: the newer buffer holds the next part's segment
        int pend_part = 0, pend_nw = 0, walk_part = -1;
        auto request_ahead = [&]() {
#pragma unroll
            for (int k = 0; k < kPPAhead; ++k) {
                const uint32_t o = pp.bw.template next<5>(lmask, lwj0, lane);
                a[k] = row_at(rows, o).a;
                b[k] = row_at(rows, o).b;
            }
            pp.walked(kPPAhead);
        };
        for (;;) {
            if (pp.nw_new == 0 && !pend && more) {
                int nw = 0, seg_part = 0;
                SegLane sl = {0u, 0u};
                if (stored) {
                    if (seg_base < words_end) {
                        seg_part = part_of_word(seg_base, pstarts);
                        const int pend_end = part_word_end(pstarts, seg_part, total_words);
                        nw = min(kSeg - 1, pend_end - seg_base);
                        const size_t base = (size_t)item * c.max_words + seg_base;
                        uint32_t wv[kSeg - 1];
#pragma unroll
                        for (int w = 0; w < kSeg - 1; ++w)
                            wv[w] = __builtin_nontemporal_load(
                                &gmask[(base + (size_t)min(w, nw - 1)) * 64 + lane]);
#pragma unroll
                        for (int w = 0; w < kSeg - 1; ++w) {
                            if (w < nw) {
                                lm_new[w * 64 + lane] = wv[w];
                                sl.nonempty |= (wv[w] != 0u ? 1u : 0u) << w;
                                sl.count += (uint32_t)__builtin_popcount(wv[w]);
                            }
                        }
                        for (int w = (int)lane; w < nw; w += 64) lj_new[w] = gwj0[base + w];
                        seg_base += nw;
                    }
                    if (seg_base >= words_end) more = false;
                } else {
                    nw = phase1_segment<kSeg>(cu, ch, sq, cs, c, qi, active, lm_new, lj_new, lane,
                                              sl, seg_part);
                    if (cu.done) more = false;
                }
                if (nw == 0) {
                    more = false;
                } else {
                    wave_sync();
                    if (walk_part < 0 || seg_part == walk_part) {
                        if (walk_part < 0) {  // the walk's first segment
                            fold_parts(ltot, lane, f, seg_part - cur_part);
                            cur_part = seg_part;
                            walk_part = seg_part;
                        }
                        pp.add_newer(sl, nw, lmask, lwj0, lane, c.pad_slot);
                    } else {
                        pend = true;
                        pend_part = seg_part;
                        pend_nw = nw;
                    }
                }
            }
            // A phase, as in the density walk: nothing stays in flight between phases.
            const bool overlap = pp.nw_new != 0;
            if (!overlap && walk_part < 0) break;  // no segment at all
            if ((overlap ? pp.older_steps() : pp.all_steps()) != 0) {
                request_ahead();
                const uint32_t steps = overlap ? pp.older_steps() : pp.all_steps();
                force_run_n<kLds, kCut, kPPAhead>(pp.bw, a, b, steps, lmask, lwj0, lane, rows,
                                                  pi, vi, p_quarter, c, f, cnt);
                pp.walked(steps);
#pragma unroll
                for (int k = 0; k < kPPAhead; ++k)
                    f.template add<kCut>(pi, vi, p_quarter, a[k], b[k], c, cnt);
            }
            if (overlap) {
                pp.shift_down(lmask, lwj0, lane);
                continue;
            }
            if (!pend) break;
            fold_parts(ltot, lane, f, pend_part - cur_part);
            cur_part = pend_part;
            walk_part = pend_part;
            pend = false;
            pp.add_held(pend_nw, lmask, lwj0, lane, c.pad_slot);
        }
    } else {
        for (;;) {
            int nw, seg_part;
            SegLane sl;
            if (stored) {
                if (seg_base >= words_end) break;
                seg_part = part_of_word(seg_base, pstarts);
                const int pend = part_word_end(pstarts, seg_part, total_words);
                nw = min(kSeg - 1, pend - seg_base);
                const size_t base = (size_t)item * c.max_words + seg_base;
                sl.nonempty = 0;
                sl.count = 0;
                // all of the segment's word loads are issued before the first is used (a loop
                // over nw would wait for each load of its remainder in turn); slots past nw
                // re-read word nw - 1 and are dropped.  The words are read once: nontemporal
                // loads keep them from evicting the neighbour rows from L2
                uint32_t wv[kSeg - 1];
    #pragma unroll
                for (int w = 0; w < kSeg - 1; ++w)
                    wv[w] = __builtin_nontemporal_load(
                        &gmask[(base + (size_t)min(w, nw - 1)) * 64 + lane]);
    #pragma unroll
                for (int w = 0; w < kSeg - 1; ++w) {
                    if (w < nw) {
                        lmask[w * 64 + lane] = wv[w];
                        sl.nonempty |= (wv[w] != 0u ? 1u : 0u) << w;
                        sl.count += (uint32_t)__builtin_popcount(wv[w]);
                    }
                }
                for (int w = (int)lane; w < nw; w += 64) lwj0[w] = gwj0[base + w];
            } else {
                nw = phase1_segment<kSeg>(cu, ch, sq, cs, c, qi, active, lmask, lwj0, lane, sl,
                                          seg_part);
                if (nw == 0) break;
            }
            if (!kSplit) {
                if (kRegTot)
                    fold_parts_reg(rtot, f, seg_part - cur_part);
                else
                    fold_parts(ltot, lane, f, seg_part - cur_part);
                cur_part = seg_part;
            }
            wave_sync();
            BitWalk<kSeg> bw;
            bw.begin(sl.nonempty, lmask, lwj0, lane);
            if constexpr (kSplit)
                force_bits_deep<kSeg, kCut, kDeepAhead>(bw, wave_max(sl.count), lmask, lwj0, lane,
                                                        rows, pi, vi, p_quarter, c, f, cnt);
            else
                force_bits<kSeg, kCut>(bw, wave_max(sl.count), lmask, lwj0, lane, rows, pi, vi,
                                       p_quarter, c, f, cnt);
            seg_base += nw;
            wave_sync();
            if (!stored && cu.done) break;
        }
    }
    if (kSplit) {
        if (active) {
            float4* out = ps.force + 2 * ((size_t)part * ps.cap + s);
            out[0] = make_float4(f.fp.x, f.fp.y, f.fp.z, 0.0f);
            out[1] = make_float4(f.fv.x, f.fv.y, f.fv.z, 0.0f);
            ps.cnt[(size_t)part * ps.cap + s] = cnt;
        }
        return;
    }
    if (stats) add_nbr_stats(cnt, stats, lane);
    if (kRegTot)
        fold_parts_reg(rtot, f, kParts - cur_part);
    else
        fold_parts(ltot, lane, f, kParts - cur_part);
    if (!active) return;
    ForceSums tot;
    if (kRegTot) {
        tot.fp = make_float3(rtot[0], rtot[1], rtot[2]);
        tot.fv = make_float3(rtot[3], rtot[4], rtot[5]);
    } else {
        tot.fp = make_float3(ltot[0 * 64 + lane], ltot[1 * 64 + lane], ltot[2 * 64 + lane]);
        tot.fv = make_float3(ltot[3 * 64 + lane], ltot[4 * 64 + lane], ltot[5 * 64 + lane]);
    }
    force_epilogue(tot, pi, vi, rho_i, P_i, s, scenes[ch.scene], c, sid, dt_dev, dt_arg, pm, vel,
                   acc_out, rho_out, prs_out);
}

// Finishes split force passes: per finite particle, the three part sums in order, then the
// same epilogue as the single-wavefront pass.
__global__ __launch_bounds__(256) void k_force_finish(
    const float4* __restrict__ spm, const float4* __restrict__ svr,
    const uint32_t* __restrict__ sid, const uint32_t* __restrict__ cs, StepConsts c,
    const SceneGrid* __restrict__ scenes, PartSums ps, const float* __restrict__ dt_dev,
    float dt_arg, float4* __restrict__ pm, float4* __restrict__ vel,
    float4* __restrict__ acc_out, float* __restrict__ rho_out, float* __restrict__ prs_out,
    unsigned long long* __restrict__ stats) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = s < cs[c.G];
    if (stats) {
        const uint32_t cnt =
            active ? ps.cnt[s] + ps.cnt[ps.cap + s] + ps.cnt[2 * ps.cap + s] : 0u;
        add_nbr_stats(cnt, stats, threadIdx.x & 63u);
    }
    if (!active) return;
    // ((0 + part 0) + part 1) + part 2, as the single-wavefront walk folds them
    float t[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int p = 0; p < kParts; ++p) {
        const float4* in = ps.force + 2 * ((size_t)p * ps.cap + s);
        const float4 a = in[0], b = in[1];
        t[0] = t[0] + a.x;
        t[1] = t[1] + a.y;
        t[2] = t[2] + a.z;
        t[3] = t[3] + b.x;
        t[4] = t[4] + b.y;
        t[5] = t[5] + b.z;
    }
    ForceSums tot;
    tot.fp = make_float3(t[0], t[1], t[2]);
    tot.fv = make_float3(t[3], t[4], t[5]);
    const float4 pi = spm[s];
    const float4 vr = svr[s];
    const float P_i = c.k_gas * (vr.w - c.rho0);
    const uint32_t id = sid[s];
    force_epilogue(tot, pi, make_float3(vr.x, vr.y, vr.z), vr.w, P_i, s,
                   scenes[scene_of_particle(scenes, c.nscenes, id)], c, sid, dt_dev, dt_arg, pm,
                   vel, acc_out, rho_out, prs_out);
}

// Particles with a non-finite coordinate (dense index G) have no neighbour, not even
// themselves: rho = m W(0), a = (0, g, 0); they are integrated and clamped like the rest.  The
// bin holds the sorted slots [cs[G], n); workgroup `b` of `nb` takes every nb-th run of 256.
// Workgroup 0 also advances the simulation time (time_dev, if given) by dt.
__device__ __forceinline__ void nonfinite_particles(const StepTail& t,
                                                    const uint32_t* __restrict__ cs,
                                                    const StepConsts& c,
                                                    const SceneGrid* __restrict__ scenes,
                                                    uint32_t b, uint32_t nb) {
    const float dt = t.dt_dev ? *t.dt_dev : t.dt;
    if (t.time_dev && b == 0 && threadIdx.x == 0) *t.time_dev = *t.time_dev + dt;
    for (uint32_t s = cs[c.G] + b * blockDim.x + threadIdx.x; s < t.n; s += nb * blockDim.x) {
        const float4 pi = t.spm[s];
        const float4 vi = t.svr[s];
        const float rho_i = pi.w * c.w0;
        const float P_i = c.k_gas * (rho_i - c.rho0);
        const float3 a = final_accel(make_float3(0.f, 0.f, 0.f), make_float3(0.f, 0.f, 0.f),
                                     pi.w, rho_i, c);
        float3 v = make_float3(vi.x, vi.y, vi.z);
        float4 p = pi;
        const uint32_t id = t.sid[s];
        const SceneGrid& g = scenes[scene_of_particle(scenes, c.nscenes, id)];
        integrate_clamp(p, v, a, dt, c, g.lo, g.hi);
        t.pm[id] = p;
        t.vel[id] = make_float4(v.x, v.y, v.z, 0.0f);
        t.acc[id] = make_float4(a.x, a.y, a.z, 0.0f);
        t.rho[id] = rho_i;
        t.prs[id] = P_i;
    }
}

__global__ __launch_bounds__(256) void k_nonfinite_step(StepTail t, const uint32_t* __restrict__ cs,
                                                        StepConsts c,
                                                        const SceneGrid* __restrict__ scenes) {
    nonfinite_particles(t, cs, c, scenes, blockIdx.x, gridDim.x);
}

// Grids of at most this many (x, y) columns get their chunk list from one launch: every
// workgroup counts the chunks of all columns and scans the counts in LDS, then writes the
// items of its own 256 columns.  Workgroups from `fill_blocks` on do the non-finite bin's step
// (kTail).
constexpr uint32_t kChunkListLocal = 8192;

// Sorted slots [x, y) of column `col` (its cells are consecutive, z fastest).  kOne: a
// one-scene grid, whose geometry g0 is loaded once instead of looked up per column.
template <bool kOne>
__device__ __forceinline__ uint2 column_slots(const uint32_t* __restrict__ cs,
                                              const SceneGrid* __restrict__ scenes,
                                              const StepConsts& c, const SceneGrid& g0,
                                              uint32_t col) {
    const SceneGrid& g = kOne ? g0 : scenes[scene_of_col(scenes, c.nscenes, col)];
    const uint32_t nz = (uint32_t)g.nz;
    const uint32_t first = g.cell_base + (col - g.col_base) * nz;
    return make_uint2(cs[first], cs[first + nz]);
}

template <bool kTail>
__global__ __launch_bounds__(256) void k_chunk_list(const uint32_t* __restrict__ cs, StepConsts c,
                                                    const SceneGrid* __restrict__ scenes,
                                                    uint32_t* __restrict__ items,
                                                    uint32_t* __restrict__ nitems,
                                                    uint32_t fill_blocks, StepTail t) {
    if (blockIdx.x >= fill_blocks) {
        if constexpr (kTail)
            nonfinite_particles(t, cs, c, scenes, blockIdx.x - fill_blocks,
                                gridDim.x - fill_blocks);
        return;
    }
    constexpr int kPer = (int)(kChunkListLocal / 256u);
    __shared__ uint32_t off[kChunkListLocal];
    __shared__ uint32_t lw[4];
    const uint32_t ncols = c.ncols;
    // thread k counts the columns [k*per, (k+1)*per), all its loads in flight at once
    const uint32_t per = (ncols + 255u) / 256u;
    const uint32_t beg = threadIdx.x * per;
    const SceneGrid g0 = scenes[0];
    const bool one = c.nscenes == 1;
    uint32_t cnt[kPer];
    uint32_t sum = 0;
    if (one) {
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t col = beg + (uint32_t)j;
            const uint2 r = (uint32_t)j < per && col < ncols
                                ? column_slots<true>(cs, scenes, c, g0, col)
                                : make_uint2(0u, 0u);
            cnt[j] = (r.y - r.x + 63u) / 64u;
            sum += cnt[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const uint32_t col = beg + (uint32_t)j;
            const uint2 r = (uint32_t)j < per && col < ncols
                                ? column_slots<false>(cs, scenes, c, g0, col)
                                : make_uint2(0u, 0u);
            cnt[j] = (r.y - r.x + 63u) / 64u;
            sum += cnt[j];
        }
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan_256(sum, lw, &total);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t col = beg + (uint32_t)j;
        if ((uint32_t)j < per && col < ncols) off[col] = run;
        run += cnt[j];
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) nitems[0] = total;
    const uint32_t col = blockIdx.x * 256u + threadIdx.x;
    if (col >= ncols) return;
    const uint2 r = one ? column_slots<true>(cs, scenes, c, g0, col)
                        : column_slots<false>(cs, scenes, c, g0, col);
    uint32_t o = off[col];
    for (uint32_t s = r.x; s < r.y; s += 64u) items[o++] = s;
}

hipError_t launch_chunks(const uint32_t* cell_start, const StepConsts& c, const SceneGrid* scenes,
                         uint32_t* cnt_scratch, uint32_t* scan_scratch, uint32_t* items,
                         uint32_t* nitems, const StepTail* tail, hipStream_t st) {
    // the non-finite bin holds at most n particles; its start is read on the device
    const uint32_t tail_blocks =
        tail ? std::max<uint32_t>(1u, std::min<uint32_t>((tail->n + 255u) / 256u, 64u)) : 0u;
    if (c.ncols <= kChunkListLocal) {
        const uint32_t fill_blocks = std::max<uint32_t>(1u, (c.ncols + 255u) / 256u);
        if (tail)
            hipLaunchKernelGGL(k_chunk_list<true>, dim3(fill_blocks + tail_blocks), dim3(256), 0,
                               st, cell_start, c, scenes, items, nitems, fill_blocks, *tail);
        else
            hipLaunchKernelGGL(k_chunk_list<false>, dim3(fill_blocks), dim3(256), 0, st,
                               cell_start, c, scenes, items, nitems, fill_blocks, StepTail{});
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_chunk_count, dim3((c.ncols + 1 + 255) / 256), dim3(256), 0, st,
                       cell_start, c, scenes, cnt_scratch);
    hipError_t e = scan_exclusive(cnt_scratch, c.ncols + 1, scan_scratch, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_chunk_fill, dim3((c.ncols + 255) / 256 + 1), dim3(256), 0, st, cell_start,
                       c, scenes, cnt_scratch, items, nitems);
    if (tail)
        hipLaunchKernelGGL(k_nonfinite_step, dim3(tail_blocks), dim3(256), 0, st, *tail,
                           cell_start, c, scenes);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// neighbour lists for parity checks: the production phase-1 test, expanded to CSR over
// caller ids in reference order.  pass 0 counts, pass 1 writes.
// ---------------------------------------------------------------------------
template <class Shape>
__global__ __launch_bounds__(Shape::threads) void k_nbr_lists(
    const uint32_t* __restrict__ items, const uint32_t* __restrict__ nitems,
    const float4* __restrict__ spm, const uint2* __restrict__ sq,
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sid,
    const uint32_t* __restrict__ cs, StepConsts c, const SceneGrid* __restrict__ scenes,
    int pass,
    uint32_t* __restrict__ cnt_by_id, const int64_t* __restrict__ offs_by_id,
    int64_t* __restrict__ idx) {
    constexpr int kSeg = Shape::seg;
    __shared__ uint32_t lmask_all[Shape::waves][kSeg * 64];
    __shared__ uint32_t lwj0_all[Shape::waves][kSeg];
    const uint32_t wave = uni(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* lmask = lmask_all[wave];
    uint32_t* lwj0 = lwj0_all[wave];
    Chunk ch;
    uint32_t s;
    bool active;
    float4 pi;
    const uint32_t wid = xcd_unit(uni(nitems[0]), Shape::waves, wave);
    if (wid == ~0u) return;
    if (!load_chunk(wid, items, nitems, skey, cs, spm, c, scenes,
                    lane, ch, s, active, pi))
        return;
    const uint2 qi = active ? sq[s] : make_uint2(0u, 0u);
    Cursor cu;
    cursor_begin(cu, ch, cs, c, 1);
    uint32_t cnt = 0;
    int64_t o = (pass == 1 && active) ? offs_by_id[sid[s]] : 0;
    for (;;) {
        SegLane sl;
        int part;  // always 0 in the reference order
        const int nw =
            phase1_segment<kSeg>(cu, ch, sq, cs, c, qi, active, lmask, lwj0, lane, sl, part);
        if (nw == 0) break;
        wave_sync();
        BitWalk<kSeg> bw;
        bw.begin(sl.nonempty, lmask, lwj0, lane);
        for (uint32_t t = 0; t < sl.count; ++t) {
            const uint32_t j = bw.next(lmask, lwj0, lane);
            const float4 q = spm[j];
            if (dist2_exact(pi.x, pi.y, pi.z, q.x, q.y, q.z) <= c.r2) {
                if (pass == 1) idx[o++] = sid[j];
                ++cnt;
            }
        }
        wave_sync();
        if (cu.done) break;
    }
    if (pass == 0 && active) cnt_by_id[sid[s]] = cnt;
}

// ---------------------------------------------------------------------------
// launch shape table (index = WalkTuning value)
// ---------------------------------------------------------------------------
using Shape0 = WalkShape<16, 4>;
using Shape1 = WalkShape<32, 4>;
using Shape2 = WalkShape<28, 4>;
using Shape3 = WalkShape<32, 1>;
using Shape4 = WalkShape<24, 4>;
using Shape5 = WalkShape<12, 4, true>;
using Shape6 = WalkShape<16, 4, true>;

const char* walk_shape_name(int v) {
    switch (v) {
        case 0: return "seg16_waves4";
        case 1: return "seg32_waves4";
        case 2: return "seg28_waves4";
        case 3: return "seg32_waves1";
        case 4: return "seg24_waves4";
        case 5: return "seg16_waves4_split";
        case 6: return "seg32_waves4_split";
        case 7: return "pp12_waves4";
        case 8: return "pp16_waves4";
        default: return "?";
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  The grid is
// padded to a multiple of 8 and sized for the most chunks a step can have (the host does not
// know the count, which the chunk list leaves on the device); each XCD walks one contiguous
// eighth of the chunks that exist (xcd_unit), whose candidate columns overlap.  A split shape
// runs three wavefronts per chunk (consecutive, so mostly in one workgroup).
template <class Shape>
static uint32_t grid_for(uint32_t chunks, bool split = false) {
    const uint32_t waves = split ? chunks * (uint32_t)kParts : chunks;
    uint32_t blocks = (waves + Shape::waves - 1) / Shape::waves;
    return (blocks + kXcds - 1) / kXcds * kXcds;
}

// CALL sees the shape as type S and the split flag as constexpr kS.
#define SPHK_DISPATCH_SHAPE(v, CALL)     \
    switch (v) {                         \
        case 0: { using S = Shape0; constexpr bool kS = false; CALL; break; } \
        case 1: { using S = Shape1; constexpr bool kS = false; CALL; break; } \
        case 2: { using S = Shape2; constexpr bool kS = false; CALL; break; } \
        case 3: { using S = Shape3; constexpr bool kS = false; CALL; break; } \
        case 5: { using S = Shape0; constexpr bool kS = true; CALL; break; } \
        case 6: { using S = Shape1; constexpr bool kS = true; CALL; break; } \
        case 7: { using S = Shape5; constexpr bool kS = false; CALL; break; } \
        case 8: { using S = Shape6; constexpr bool kS = false; CALL; break; } \
        default: { using S = Shape4; constexpr bool kS = false; CALL; break; } \
    }

static hipError_t part_sums(const WalkBuffers& wb, uint32_t n, int shape, PartSums* ps) {
    *ps = PartSums{wb.part_acc, wb.part_cnt, wb.part_force, wb.part_cap};
    if (walk_shape_split(shape) && (!wb.part_acc || wb.part_cap < n))
        return hipErrorInvalidValue;  // split shapes need per-part sums for every particle
    return hipSuccess;
}

hipError_t launch_density_walk(const WalkBuffers& wb, const float4* spm, const uint2* sq,
                               float4* svr,
                               ForceRow* rows, const uint32_t* skey,
                               const uint32_t* cell_start, uint32_t n,
                               const StepConsts& c, const SceneGrid* scenes, int shape,
                               hipStream_t st) {
    if (n == 0) return hipSuccess;
    PartSums ps;
    if (hipError_t err = part_sums(wb, n, shape, &ps)) return err;
    const uint32_t chunks = max_chunks(n, c);
#define SPHK_DENSITY_LAUNCH(CUT)                                                            \
    SPHK_DISPATCH_SHAPE(shape, hipLaunchKernelGGL((k_density_walk<S, kS, CUT>),               \
                                                  dim3(grid_for<S>(chunks, kS)),             \
                                                  dim3(S::threads), 0, st, wb.items,         \
                                                  wb.nitems, spm, sq, svr, rows, skey,       \
                                                  cell_start, c, scenes, wb.mask, wb.wj0,    \
                                                  wb.nwords, wb.pstarts, ps))
    if (c.cut) {
        SPHK_DENSITY_LAUNCH(true);
    } else {
        SPHK_DENSITY_LAUNCH(false);
    }
#undef SPHK_DENSITY_LAUNCH
    if (walk_shape_split(shape))
        hipLaunchKernelGGL(k_density_finish, dim3((n + 255) / 256), dim3(256), 0, st, spm, svr,
                           rows, cell_start, c, ps);
    return hipGetLastError();
}

hipError_t launch_force_walk(const WalkBuffers& wb, const float4* spm, const uint2* sq,
                             const float4* svr,
                             const ForceRow* rows, const uint32_t* skey,
                             const uint32_t* sid,
                             const uint32_t* cell_start, uint32_t n, const StepConsts& c,
                             const SceneGrid* scenes, const float* dt_dev, float dt, float4* pm,
                             float4* vel, float4* acc, float* rho, float* prs, int shape,
                             unsigned long long* nbr_stats, hipStream_t st) {
    if (n == 0) return hipSuccess;
    PartSums ps;
    if (hipError_t err = part_sums(wb, n, shape, &ps)) return err;
    const uint32_t chunks = max_chunks(n, c);
#define SPHK_FORCE_LAUNCH(CUT)                                                              \
    SPHK_DISPATCH_SHAPE(shape, hipLaunchKernelGGL((k_force_walk<S, kS, CUT>),                 \
                                                  dim3(grid_for<S>(chunks, kS)),             \
                                                  dim3(S::threads), 0, st, wb.items,         \
                                                  wb.nitems, spm, sq, svr, rows, skey, sid,  \
                                                  cell_start, c, scenes, wb.mask, wb.wj0,    \
                                                  wb.nwords, wb.pstarts, ps, dt_dev, dt, pm, \
                                                  vel, acc, rho, prs, nbr_stats))
    if (c.cut) {
        SPHK_FORCE_LAUNCH(true);
    } else {
        SPHK_FORCE_LAUNCH(false);
    }
#undef SPHK_FORCE_LAUNCH
    if (walk_shape_split(shape))
        hipLaunchKernelGGL(k_force_finish, dim3((n + 255) / 256), dim3(256), 0, st, spm, svr,
                           sid, cell_start, c, scenes, ps, dt_dev, dt, pm, vel, acc, rho, prs,
                           nbr_stats);
    return hipGetLastError();
}

hipError_t launch_nbr_lists(const WalkBuffers& wb, const float4* spm, const uint2* sq,
                            const uint32_t* skey,
                            const uint32_t* sid, const uint32_t* cell_start, uint32_t n,
                            const StepConsts& c, const SceneGrid* scenes, int pass,
                            uint32_t* count_by_id, const int64_t* offsets_by_id, int64_t* idx,
                            int shape, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t chunks = max_chunks(n, c);
    // one wavefront per chunk in every shape (a split shape lists with its unsplit twin, a
    // ping-pong shape with the single-buffer shape of the same LDS size)
    const int base = shape == 5 ? 0 : shape == 6 ? 1 : shape == 7 ? 4 : shape == 8 ? 1 : shape;
    SPHK_DISPATCH_SHAPE(base, hipLaunchKernelGGL(k_nbr_lists<S>, dim3(grid_for<S>(chunks)),
                                                  dim3(S::threads), 0, st, wb.items, wb.nitems,
                                                  spm, sq, skey, sid, cell_start, c, scenes,
                                                  pass, count_by_id, offsets_by_id, idx));
    return hipGetLastError();
}

}  // namespace sphk
