// sphk_walk_common.h -- device helpers of the chunked neighbour walks (chunk description,
// candidate cursor, phase-1 test words, per-lane bit walk), shared by the SPH walks and the
// surrogate's edge walks.  See sphk_walk.hip for the walk design.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "sphk_internal.h"

namespace sphk {

namespace {

// LDS visibility between lanes of one wavefront: LDS operations of a wave complete in
// order, so only the compiler must be kept from reordering them.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float next_up(float x) {
    if (x == 0.0f) return __int_as_float(1);
    int i = __float_as_int(x);
    return __int_as_float(x > 0.0f ? i + 1 : i - 1);
}

__device__ __forceinline__ float next_down(float x) {
    if (x == 0.0f) return -__int_as_float(1);
    int i = __float_as_int(x);
    return __int_as_float(x > 0.0f ? i - 1 : i + 1);
}

// (dx*dx + dy*dy) + dz*dz with dx = a - b, every operation rounded once
// (within_radius_squared / glm::dot, src/spatial_hash.h:70-73).
__device__ __forceinline__ float dist2_exact(float ax, float ay, float az, float bx, float by,
                                             float bz) {
#pragma clang fp contract(off)
    float dx = ax - bx;
    float dy = ay - by;
    float dz = az - bz;
    float tx = dx * dx;
    float ty = dy * dy;
    float tz = dz * dz;
    return (tx + ty) + tz;
}

// Contracted d^2 for the FP32 candidate filter of the surrogate's in-kernel list building
// (the SPH walks use the integer filter, push_quant): fl(dx^2 + fl(dy^2 + fl(dz^2))) with
// two fused multiply-adds.  Both roundings stay within 3 units of 2^-24 of the true sum, so the two
// values differ by under 7 units of 2^-24 relative, and
// testing it against r2 raised by 8 ulp (StepConsts::r2_filter) keeps every pair the exact
// test keeps; the few extra pairs are dropped by the exact test where the pair is used.
__device__ __forceinline__ float dist2_filter(float ax, float ay, float az, float bx, float by,
                                              float bz) {
    const float dx = ax - bx;
    const float dy = ay - by;
    const float dz = az - bz;
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}

// Can a particle at coordinate x (cell c = floor(fl(x*inv))) have a neighbour in cell c+2
// (bit 1) or c-2 (bit 0)?  A neighbour needs fl(fl(b-x)^2) <= r2 on this axis alone, and
// fl(b-x) grows with b, so it suffices to test the float b nearest to x that lands in cell
// c+2 (resp. c-2).  Exact; conservative (flag set) if the bracket search does not settle.
__device__ __forceinline__ int axis_edge_flags(float x, float inv, float r2) {
#pragma clang fp contract(off)
    float c = floorf(x * inv);
    int flags = 0;
    {
        float t = c + 2.0f;
        float b = t / inv;
        int ok = 0;
        for (int k = 0; k < 6; ++k) {
            if (b * inv >= t) {
                ok = 1;
                break;
            }
            b = next_up(b);
        }
        if (ok) {
            for (int k = 0; k < 6; ++k) {
                float bd = next_down(b);
                if (bd * inv >= t) {
                    b = bd;
                } else {
                    ok = 2;
                    break;
                }
            }
        }
        float d = b - x;
        if (ok != 2 || d * d <= r2) flags |= 2;
    }
    {
        float t = c - 1.0f;
        float b = t / inv;
        int ok = 0;
        for (int k = 0; k < 6; ++k) {
            if (b * inv < t) {
                ok = 1;
                break;
            }
            b = next_down(b);
        }
        if (ok) {
            for (int k = 0; k < 6; ++k) {
                float bu = next_up(b);
                if (bu * inv < t) {
                    b = bu;
                } else {
                    ok = 2;
                    break;
                }
            }
        }
        float d = x - b;
        if (ok != 2 || d * d <= r2) flags |= 1;
    }
    return flags;
}

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

constexpr uint32_t kXcds = 8;

// Work unit of this wavefront (a chunk, or a chunk's part in the split walks) after the XCD
// split: workgroups are dealt to the 8 XCDs in turn (block b runs on XCD b % 8), and XCD x
// takes the x-th eighth of the `units` units that exist (one contiguous run, so neighbouring
// chunks share its L2).  The grid is sized for the most units a step can have (see grid_for),
// which gives every XCD at least an eighth of them; ~0u for a wavefront past its XCD's run.
__device__ __forceinline__ uint32_t xcd_unit(uint32_t units, uint32_t waves_per_block,
                                             uint32_t wave) {
    const uint32_t run = (units + kXcds - 1) / kXcds;
    const uint32_t j = (blockIdx.x / kXcds) * waves_per_block + wave;
    if (j >= run) return ~0u;
    const uint32_t u = (blockIdx.x % kXcds) * run + j;
    return u < units ? u : ~0u;
}
__device__ __forceinline__ int unii(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave-uniform description of one chunk and its candidate columns.
struct Chunk {
    uint32_t s0, s_end;   // sorted slots [s0, s_end)
    uint32_t scene;       // scene whose grid holds the chunk
    uint32_t cell_base;   // that grid's first global cell
    uint32_t ny, nz;      // that grid's dims (y, z)
    int rx, ry;           // column (grid-relative)
    int dx0, dx1, dy0, dy1;
    int zlo, zhi;         // candidate z cells (inclusive), grid-relative
    int zfirst;           // z cell of the first particle, grid-relative
};

// Candidate columns in walk order, as (dx + 2) * 5 + (dy + 2).  Row 0, balanced order: the
// four corner columns, the centre, the four edge columns, then the ring two columns out
// (walked only when axis_edge_flags asks for it).  Row 1: the reference's (dx, dy)
// lexicographic order (SpatialHash::query_squared), for the neighbour-list export.
__constant__ uint8_t kColumnOrder[2][25] = {
    {6, 18, 8, 16, 12, 7, 17, 11, 13, 0, 1, 2, 3, 4, 5, 9, 10, 14, 15, 19, 20, 21, 22, 23, 24},
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24}};

// Parts of a walk (balanced order): the corner columns with the centre (positions 0-4), the
// edge columns (5-8), the outer ring (9-24).  Sums over a chunk's neighbours are formed per
// part, each in candidate order, and combined as (0 + part 0) + part 1) + part 2, so a walk
// split over three wavefronts (one part each) gives the same bits as one wavefront walking
// all three.  Part starts are segment breaks.
__device__ __forceinline__ int part_begin(int p) { return p == 0 ? 0 : p == 1 ? 5 : 9; }
__device__ __forceinline__ int part_end(int p) { return p == 0 ? 5 : p == 1 ? 9 : 25; }
__device__ __forceinline__ int part_of_position(int k) { return k < 5 ? 0 : k < 9 ? 1 : 2; }

// Candidate order.  A word is 32 consecutive sorted slots of one candidate column (the last
// word of a column may be shorter), candidate k of a word in bit 31 - k.  In the balanced
// order each part is walked in kPasses passes: pass s visits the part's columns in order and
// takes words s, s + kPasses, s + 2 kPasses, ... of each.  A column's words run up in z, so
// every segment mixes all columns and z ranges of the part, and the neighbour counts of the
// 64 lanes in a segment come out even (the wavefront runs the largest): a replay of the 1M
// dam break gives ~300 steps per chunk with 31-word segments against ~326 walking each
// column's words in turn (254 for one continuous walk).  The reference order walks every
// word of each column in turn.
constexpr int kPasses = 4;

// Cursor over the candidate words of a chunk.
struct Cursor {
    int k;            // position of the current column in the order
    int kend;         // end of the positions of the current part
    int part;         // current part (0 in the reference order)
    int last_part;    // last part walked
    int pass;         // current pass
    int order;        // 0 balanced, 1 reference
    bool fresh;       // the next word opens a part
    uint32_t j, end;  // next word's first slot; end of the current column range
    uint32_t step;    // slots from one word of a column to the next walked one
    bool done;
    int pw0, pw1, pw2;  // words of each part
};

__device__ __forceinline__ void column_range(const Chunk& ch, int dx, int dy,
                                             const uint32_t* __restrict__ cs, const StepConsts& c,
                                             uint32_t& a, uint32_t& b) {
    uint32_t base = ch.cell_base +
                    ((uint32_t)(ch.rx + dx) * ch.ny + (uint32_t)(ch.ry + dy)) * ch.nz;
    a = uni(cs[base + (uint32_t)ch.zlo]);
    b = uni(cs[base + (uint32_t)ch.zhi + 1u]);
}

// Column at order position k, if the chunk walks it.
__device__ __forceinline__ bool column_at(const Chunk& ch, int order, int k, int& dx, int& dy) {
    const int code = (int)kColumnOrder[order][k];
    dx = code / 5 - 2;
    dy = code % 5 - 2;
    return dx >= ch.dx0 && dx <= ch.dx1 && dy >= ch.dy0 && dy <= ch.dy1;
}

// Moves to the first column at or after position k of the current pass that has a word in
// this pass, continuing with the next pass and the next part (done when none is left).
__device__ __forceinline__ void cursor_seek(Cursor& cu, const Chunk& ch,
                                            const uint32_t* __restrict__ cs, const StepConsts& c,
                                            int k) {
    for (;;) {
        for (; k < cu.kend; ++k) {
            int dx, dy;
            if (!column_at(ch, cu.order, k, dx, dy)) continue;
            uint32_t a, b;
            column_range(ch, dx, dy, cs, c, a, b);
            const uint32_t j = a + 32u * (uint32_t)cu.pass;
            if (j >= b) continue;
            cu.k = k;
            cu.j = j;
            cu.end = b;
            return;
        }
        if (cu.order == 0 && cu.pass + 1 < kPasses) {
            ++cu.pass;
            k = part_begin(cu.part);
            continue;
        }
        if (cu.order == 0 && cu.part < cu.last_part) {
            ++cu.part;
            cu.pass = 0;
            cu.kend = part_end(cu.part);
            cu.fresh = true;
            k = part_begin(cu.part);
            continue;
        }
        cu.done = true;
        return;
    }
}

// order 0: balanced (the SPH walks); 1: the reference's (neighbour-list export).  A part
// p >= 0 (balanced order) limits the walk to that part; the word counts of all parts are
// counted either way.
__device__ __forceinline__ void cursor_begin(Cursor& cu, const Chunk& ch,
                                             const uint32_t* __restrict__ cs, const StepConsts& c,
                                             int order = 0, int part = -1) {
    cu.order = order;
    cu.part = part >= 0 ? part : 0;
    cu.last_part = order != 0 ? 0 : part >= 0 ? part : 2;
    cu.kend = order != 0 ? 25 : part_end(cu.part);
    cu.pass = 0;
    cu.step = order != 0 ? 32u : 32u * kPasses;
    cu.done = false;
    cu.fresh = true;
    cu.pw0 = cu.pw1 = cu.pw2 = 0;
    if (order == 0) {
        for (int k = 0; k < 25; ++k) {
            int dx, dy;
            if (!column_at(ch, 0, k, dx, dy)) continue;
            uint32_t a, b;
            column_range(ch, dx, dy, cs, c, a, b);
            const int w = (int)((b - a + 31u) / 32u);
            const int p = part_of_position(k);
            cu.pw0 += p == 0 ? w : 0;
            cu.pw1 += p == 1 ? w : 0;
            cu.pw2 += p == 2 ? w : 0;
        }
    }
    cursor_seek(cu, ch, cs, c, order != 0 ? 0 : part_begin(cu.part));
}

// Words of the chunk before part p (where its words start in the stored order).
__device__ __forceinline__ int part_word_base(const Cursor& cu, int p) {
    return (p > 0 ? cu.pw0 : 0) + (p > 1 ? cu.pw1 : 0);
}

// Moves on when the current column has no word left in this pass; false when all are done.
__device__ __forceinline__ bool cursor_settle(Cursor& cu, const Chunk& ch,
                                              const uint32_t* __restrict__ cs,
                                              const StepConsts& c) {
    while (!cu.done && cu.j >= cu.end) cursor_seek(cu, ch, cs, c, cu.k + 1);
    return !cu.done;
}

// Part starts of a chunk's stored words, packed 16 bits each (words before part 1, part 2).
__device__ __forceinline__ uint32_t pack_parts(int ps1, int ps2) {
    return (uint32_t)ps1 | (uint32_t)ps2 << 16;
}
// First stored word of part p, and the end of its words.
__device__ __forceinline__ int part_word_start(uint32_t parts, int p) {
    return p == 0 ? 0 : (int)((parts >> (16 * (p - 1))) & 0xffffu);
}
__device__ __forceinline__ int part_word_end(uint32_t parts, int p, int total) {
    return p == 2 ? total : part_word_start(parts, p + 1);
}
// Part of the stored word at `pos`.
__device__ __forceinline__ int part_of_word(int pos, uint32_t parts) {
    return pos < part_word_start(parts, 1) ? 0 : pos < part_word_start(parts, 2) ? 1 : 2;
}

// Appends one test result to a 32-bit word: word = 2*word + (d2 <= r2).  One compare into
// VCC and one add-with-carry; the first candidate of a word ends up in bit 31.
__device__ __forceinline__ uint32_t push_bit(uint32_t word, float d2, float r2) {
    uint32_t out;
    asm("v_cmp_ge_f32 vcc, %2, %3\n\tv_addc_co_u32 %0, vcc, %1, %1, vcc"
        : "=v"(out)
        : "v"(word), "s"(r2), "v"(d2)
        : "vcc");
    return out;
}

// ---------------------------------------------------------------------------
// Word table: the words of a chunk's walked range in walk order (balanced order, parts,
// passes), built once per chunk so that phase 1 reads each word's (first slot, count) with
// one lane read instead of stepping the cursor (whose column visits cost a scalar-load round
// trip each, four per column with the passes).  Lane l holds entries l and 64 + l; an entry
// is j0 | (cnt - 1) << 27 (slots stay below 2^27).  Ranges of more than kTableWords words use
// the cursor; both give the same sequence.
// ---------------------------------------------------------------------------
constexpr int kTableWords = 128;

struct WordTable {
    uint32_t e0, e1;  // this lane's entries (l, 64 + l)
    int n;            // uniform: words in the table
    int first;        // uniform: stored-word index of entry 0
    int ps1, ps2;     // uniform: stored-word index where parts 1 and 2 start
    int total;        // uniform: words of the whole chunk
};

// Words of a column of nw words in pass s: s, s + kPasses, ...
__device__ __forceinline__ int pass_words(int nw, int s) {
    return nw > s ? (nw - s + kPasses - 1) / kPasses : 0;
}

// Counts the chunk's words; fills the table for the range of `part` (all parts if < 0) when
// it holds at most kTableWords words.  Returns whether the table is filled.
__device__ __forceinline__ bool word_table_build(WordTable& t, const Chunk& ch,
                                                 const uint32_t* __restrict__ cs,
                                                 const StepConsts& c, int part, uint32_t lane) {
    int nw[25];
    int pw[3] = {0, 0, 0};
    int W[3][kPasses] = {};
#pragma unroll
    for (int k = 0; k < 25; ++k) {
        int dx, dy;
        nw[k] = 0;
        if (column_at(ch, 0, k, dx, dy)) {
            uint32_t a, b;
            column_range(ch, dx, dy, cs, c, a, b);
            nw[k] = (int)((b - a + 31u) / 32u);
        }
        const int p = part_of_position(k);
        pw[p] += nw[k];
#pragma unroll
        for (int s = 0; s < kPasses; ++s) W[p][s] += pass_words(nw[k], s);
    }
    t.ps1 = pw[0];
    t.ps2 = pw[0] + pw[1];
    t.total = t.ps2 + pw[2];
    t.first = part > 0 ? (part == 1 ? t.ps1 : t.ps2) : 0;
    t.n = part >= 0 ? pw[part] : t.total;
    t.e0 = t.e1 = 0u;
    if (t.n > kTableWords) return false;
    const int i0 = t.first + (int)lane, i1 = i0 + 64;  // stored-word indices of this lane's entries
    int base = 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        int pre[kPasses];
        int pb[kPasses];  // stored-word index where pass s of part p starts
#pragma unroll
        for (int s = 0; s < kPasses; ++s) {
            pre[s] = 0;
            pb[s] = s == 0 ? base : pb[s - 1] + W[p][s - 1];
        }
        if (part < 0 || part == p) {
#pragma unroll
            for (int k = part_begin(p); k < part_end(p); ++k) {
                if (nw[k] == 0) continue;
                int dx, dy;
                column_at(ch, 0, k, dx, dy);
                uint32_t a, b;
                column_range(ch, dx, dy, cs, c, a, b);
#pragma unroll
                for (int s = 0; s < kPasses; ++s) {
                    const int m = pass_words(nw[k], s);
                    const int lo = pb[s] + pre[s];
                    // entry for stored-word index i: word s + kPasses (i - lo) of the column
                    if (i0 >= lo && i0 < lo + m) {
                        const uint32_t j0 = a + 32u * (uint32_t)(s + kPasses * (i0 - lo));
                        t.e0 = j0 | (min(32u, b - j0) - 1u) << 27;
                    }
                    if (i1 >= lo && i1 < lo + m) {
                        const uint32_t j0 = a + 32u * (uint32_t)(s + kPasses * (i1 - lo));
                        t.e1 = j0 | (min(32u, b - j0) - 1u) << 27;
                    }
                    pre[s] += m;
                }
            }
        }
        base += pw[p];
    }
    return true;
}

// Entry i (uniform, < n) of the table: first slot and count.
__device__ __forceinline__ void word_table_entry(const WordTable& t, int i, uint32_t& j0,
                                                 uint32_t& cnt) {
    const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)(i < 64 ? t.e0 : t.e1), i & 63);
    j0 = e & ((1u << 27) - 1u);
    cnt = (e >> 27) + 1u;
}

// Per-lane summary of one LDS segment: which words hold a neighbour (bit w) and how many
// neighbours in total.
struct SegLane {
    uint32_t nonempty;
    uint32_t count;
};

// Appends the quantized filter's verdict on one candidate to a 32-bit word: with
// dX, dY (16-bit wrap-around) and dZ (24-bit) the differences of the quantized coordinates,
// t = dX^2 + dY^2 + dZ^2 + qneg is negative exactly when the squared distance is at most T,
// and word = 2 * word + sign(t).  Five VALU: packed 16-bit subtract, subtract, 24-bit
// multiply-add, dot2, align-bit.  Wrap-around can only turn a far candidate into a passing
// one, never a neighbour into a failing one.
typedef short sphk_short2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t push_quant(uint32_t word, uint2 a, uint2 b, int32_t qneg) {
    const sphk_short2 dxy =
        __builtin_bit_cast(sphk_short2, a.x) - __builtin_bit_cast(sphk_short2, b.x);
    const uint32_t dz = (uint32_t)(((int32_t)((a.y - b.y) << 8)) >> 8);
    const int32_t t = __builtin_amdgcn_sdot2(dxy, dxy, (int32_t)(dz * dz + (uint32_t)qneg), false);
    return __builtin_amdgcn_alignbit(word, (uint32_t)t, 31);
}

// Phase 1: fills up to kSeg - 1 words of filter bits into LDS (masks[w*64 + lane], wj0[w]):
// a superset of the neighbours (see push_quant; lanes past the chunk get empty words).
// Candidate k of a word sits in bit 31-k (words of fewer than 32 candidates are
// left-aligned).  Word kSeg - 1 is the sentinel of BitWalk and is never written here.
// Returns the number of words written.
// ---------------------------------------------------------------------------
// The same quantized test on the matrix cores.  With a and b the target's and the
// candidate's quantized coordinates relative to an origin near the chunk,
//   |a - b|^2 - (T + 1) = b.b - 2 a.b + (a.a - T - 1),
// one v_mfma_f32_32x32x16_f16 evaluates it for 32 candidates (rows) x 32 targets (columns):
// K slots 0-2 hold b and -2a, slots 3-4 b.b as lo + 2048 hi against (1, 2048), slots 5-6
// (1, 2048) against a.a - T - 1 as lo + 2048 hi, the rest zero.  Every operand is an integer of
// magnitude <= 2048 (exact in f16) and every partial sum an integer below 2^24 (exact in the
// FP32 accumulator) when all coordinates lie within +-1024 of the origin: the first target's
// x, y (candidates are at most 2 cells + 1 unit away: |.| <= 769) and, in z, the middle of the
// candidate cells (|.| <= 128 (zhi - zlo + 1) + 257: at most 5 candidate z cells).  Chunks with
// more use the VALU test.  The verdicts equal the VALU test's wherever that one does not wrap
// around, which only adds far candidates; either way only such candidates differ, and they
// add exactly zero.
// ---------------------------------------------------------------------------
typedef _Float16 sphk_half8 __attribute__((ext_vector_type(8)));
typedef float sphk_f32x16 __attribute__((ext_vector_type(16)));

struct MfmaFilter {
    bool on;             // uniform: the chunk fits the exact range
    uint32_t ox, oz;     // origin: packed x | y << 16 (16-bit wrap-around), z
    sphk_half8 b1, b2;   // B operands: targets 0-31 and 32-63 (lanes 0-31 hold them)
    uint32_t crow;       // candidate of this lane's A row: 16 ((m >> 2) & 1) + 4 (m >> 3) + (m & 3)
    uint32_t pref_j;     // uniform: first slot of the word whose row is in `pref` (~0: none)
    uint2 pref;          // this lane's row of the next word of the current column
};

__device__ __forceinline__ void mfma_filter_begin(MfmaFilter& mf, const Chunk& ch,
                                                  const uint2* __restrict__ sq, uint2 qi,
                                                  uint32_t lane, int32_t qneg) {
    mf.on = ch.zhi - ch.zlo <= 4;
    const uint32_t qfx = uni(sq[ch.s0].x), qfz = uni(sq[ch.s0].y);
    mf.ox = qfx;
    mf.oz = qfz + (uint32_t)(128 * (ch.zlo + ch.zhi + 1 - 2 * ch.zfirst));
    const sphk_short2 a = __builtin_bit_cast(sphk_short2, qi.x) - __builtin_bit_cast(sphk_short2, mf.ox);
    // (lanes without a target compute garbage in their own columns only; unsigned
    // arithmetic keeps that free of overflow)
    const uint32_t az = qi.y - mf.oz;
    const int32_t aa = __builtin_amdgcn_sdot2(a, a, (int32_t)(az * az + (uint32_t)qneg), false);
    sphk_half8 p;
    p[0] = (_Float16)(-2 * (int32_t)a.x);
    p[1] = (_Float16)(-2 * (int32_t)a.y);
    p[2] = (_Float16)(int32_t)(0u - 2u * az);
    p[3] = (_Float16)1.0f;
    p[4] = (_Float16)2048.0f;
    p[5] = (_Float16)(aa & 2047);
    p[6] = (_Float16)(aa >> 11);
    p[7] = (_Float16)0.0f;
    sphk_half8 q;
    const uint4 pu = __builtin_bit_cast(uint4, p);
    uint4 qu;
    qu.x = (uint32_t)__shfl_xor((int)pu.x, 32);
    qu.y = (uint32_t)__shfl_xor((int)pu.y, 32);
    qu.z = (uint32_t)__shfl_xor((int)pu.z, 32);
    qu.w = (uint32_t)__shfl_xor((int)pu.w, 32);
    q = __builtin_bit_cast(sphk_half8, qu);
    const bool lo = lane < 32u;
    const sphk_half8 zero = {};
    mf.b1 = lo ? p : zero;
    mf.b2 = lo ? q : zero;
    const uint32_t m = lane & 31u;
    mf.crow = 16u * ((m >> 2) & 1u) + 4u * (m >> 3) + (m & 3u);
    mf.pref_j = ~0u;
    mf.pref = make_uint2(0u, 0u);
}

// Verdict word of this lane's target over candidates j0 .. j0 + cnt - 1 (cnt in 1..32),
// candidate k at bit 31 - k, like the VALU test.  The rows of the next word to be tested
// (jn .. jn + cntn - 1; jn = ~0: none) are requested before this word's arithmetic.
__device__ __forceinline__ uint32_t mfma_filter_word(MfmaFilter& mf,
                                                     const uint2* __restrict__ sq, uint32_t j0,
                                                     uint32_t cnt, uint32_t jn, uint32_t cntn,
                                                     uint32_t lane) {
    const uint2 qb = mf.pref_j == j0 ? mf.pref : sq[j0 + min(mf.crow, cnt - 1u)];
    const sphk_short2 b = __builtin_bit_cast(sphk_short2, qb.x) - __builtin_bit_cast(sphk_short2, mf.ox);
    // candidate offsets from the origin are below 2^11 per axis (the candidates lie within
    // the chunk's two-cell reach), so bz^2 is a 24-bit product and every A entry converts
    // from 16 bits in one instruction
    const int16_t bz = (int16_t)(qb.y - mf.oz);
    const int32_t bb = __builtin_amdgcn_sdot2(b, b, (int32_t)bz * (int32_t)bz, false);
    sphk_half8 a;
    a[0] = (_Float16)b.x;
    a[1] = (_Float16)b.y;
    a[2] = (_Float16)bz;
    a[3] = (_Float16)(int16_t)(bb & 2047);
    a[4] = (_Float16)(int16_t)(bb >> 11);
    a[5] = (_Float16)1.0f;
    a[6] = (_Float16)2048.0f;
    a[7] = (_Float16)0.0f;
    // lanes 32-63 supply K slots 8-15, whose B rows are zero: their A entries (the same
    // candidate's, finite) add exactly 0 and need no masking
    // next word's rows, requested once this word's are used up (the empty asm pins the A
    // operand before the request, so this word does not wait for the next one's load)
    asm volatile("" ::"v"(a));
    if (jn != ~0u) {
        mf.pref = sq[jn + min(mf.crow, cntn - 1u)];
        mf.pref_j = jn;
    }
    const sphk_f32x16 z16 = {};
    const sphk_f32x16 d1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, mf.b1, z16, 0, 0, 0);
    const sphk_f32x16 d2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, mf.b2, z16, 0, 0, 0);
    uint32_t x = 0, y = 0;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        x = __builtin_amdgcn_alignbit(x, __float_as_uint(d1[v]), 31);
        y = __builtin_amdgcn_alignbit(y, __float_as_uint(d2[v]), 31);
    }
    // lanes 0-31 own targets 0-31 (candidates 0-15 in x, 16-31 in lane + 32's x); lanes 32-63
    // own targets 32-63 (candidates 0-15 in lane - 32's y, 16-31 in y).  v_permlane32_swap
    // trades the upper lanes' x for the lower lanes' y, after which every lane holds its
    // target's candidates 0-15 in x and 16-31 in y.
    const auto xy = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    const uint32_t word = xy[0] << 16 | xy[1];
    return word & (~0u << (32u - cnt));
}

template <int kSeg, bool kMfma = false>
__device__ __forceinline__ int phase1_segment(Cursor& cu, const Chunk& ch,
                                              const uint2* __restrict__ sq,
                                              const uint32_t* __restrict__ cs,
                                              const StepConsts& c, uint2 qi, bool active,
                                              uint32_t* __restrict__ lmask,
                                              uint32_t* __restrict__ lwj0, uint32_t lane,
                                              SegLane& sl, int& part,
                                              MfmaFilter* mf = nullptr) {
    static_assert(kSeg <= 32, "one nonempty bit per word");
    int w = 0;
    sl.nonempty = 0;
    sl.count = 0;
    part = 0;
    while (w < kSeg - 1 && cursor_settle(cu, ch, cs, c)) {
        if (w == 0) part = cu.part;
        if (cu.fresh) {
            if (w > 0) break;  // a new part starts the next segment
            cu.fresh = false;
        }
        const uint32_t j0 = uni(cu.j);
        const uint32_t cnt = uni(min(32u, cu.end - j0));
        uint32_t word = 0;
        const uint2* __restrict__ pb = sq + j0;  // uniform base: scalar loads
        if (kMfma && mf->on) {
            const uint32_t jn = j0 + cu.step;
            word = mfma_filter_word(*mf, sq, j0, cnt, jn < cu.end ? jn : ~0u,
                                    jn < cu.end ? min(32u, cu.end - jn) : 0u, lane);
        } else if (cnt == 32u) {
#pragma unroll
            for (int k = 0; k < 32; ++k) word = push_quant(word, qi, pb[k], c.qneg);
        } else {
            for (uint32_t k = 0; k < cnt; ++k) word = push_quant(word, qi, pb[k], c.qneg);
            word <<= (32u - cnt);
        }
        word = active ? word : 0u;
        lmask[w * 64 + lane] = word;
        if (lane == 0) lwj0[w] = j0;
        sl.nonempty |= (word != 0u ? 1u : 0u) << w;
        sl.count += (uint32_t)__builtin_popcount(word);
        cu.j = j0 + cu.step;
        ++w;
    }
    return w;
}

// Phase 1 over the word table: the next segment of up to kSeg - 1 words from local entry `i`
// (advanced), never past a part start, tested on the matrix cores (or the VALU test where the
// chunk is too wide for them).  Same words and order as phase1_segment.
template <int kSeg>
__device__ __forceinline__ int phase1_table(const WordTable& t, int& i,
                                            const uint2* __restrict__ sq, const StepConsts& c,
                                            uint2 qi, bool active, uint32_t* __restrict__ lmask,
                                            uint32_t* __restrict__ lwj0, uint32_t lane,
                                            SegLane& sl, int& part, MfmaFilter& mf) {
    static_assert(kSeg <= 32, "one nonempty bit per word");
    int w = 0;
    sl.nonempty = 0;
    sl.count = 0;
    const int g0 = t.first + i;
    part = g0 < t.ps1 ? 0 : g0 < t.ps2 ? 1 : 2;
    while (w < kSeg - 1 && i < t.n) {
        const int g = t.first + i;
        if (w > 0 && (g == t.ps1 || g == t.ps2)) break;  // a new part starts the next segment
        uint32_t j0, cnt;
        word_table_entry(t, i, j0, cnt);
        j0 = uni(j0);
        cnt = uni(cnt);
        uint32_t word = 0;
        if (mf.on) {
            uint32_t jn = ~0u, cntn = 0u;
            if (i + 1 < t.n) word_table_entry(t, i + 1, jn, cntn);
            word = mfma_filter_word(mf, sq, j0, cnt, uni(jn), uni(cntn), lane);
        } else {
            const uint2* __restrict__ pb = sq + j0;  // uniform base: scalar loads
            if (cnt == 32u) {
#pragma unroll
                for (int k = 0; k < 32; ++k) word = push_quant(word, qi, pb[k], c.qneg);
            } else {
                for (uint32_t k = 0; k < cnt; ++k) word = push_quant(word, qi, pb[k], c.qneg);
                word <<= (32u - cnt);
            }
        }
        word = active ? word : 0u;
        lmask[w * 64 + lane] = word;
        if (lane == 0) lwj0[w] = j0;
        sl.nonempty |= (word != 0u ? 1u : 0u) << w;
        sl.count += (uint32_t)__builtin_popcount(word);
        ++i;
        ++w;
    }
    return w;
}

// Builds the wave-uniform chunk description; returns false for an out-of-range item.
// Lanes past the chunk get NaN coordinates, so every test fails for them.
__device__ __forceinline__ bool load_chunk(uint32_t item, const uint32_t* __restrict__ items,
                                           const uint32_t* __restrict__ nitems,
                                           const uint32_t* __restrict__ skey,
                                           const uint32_t* __restrict__ cs,
                                           const float4* __restrict__ spm, const StepConsts& c,
                                           const SceneGrid* __restrict__ scenes, uint32_t lane,
                                           Chunk& ch, uint32_t& s, bool& active, float4& pi) {
    if (item >= uni(nitems[0])) return false;
    ch.s0 = uni(items[item]);
    const uint32_t L0 = uni(skey[ch.s0]);
    ch.scene = uni(scene_of_cell(scenes, c.nscenes, L0));
    ch.cell_base = uni(scenes[ch.scene].cell_base);
    ch.ny = uni((uint32_t)scenes[ch.scene].ny);
    ch.nz = uni((uint32_t)scenes[ch.scene].nz);
    const uint32_t local = L0 - ch.cell_base;
    const uint32_t col = local / ch.nz;
    const uint32_t col_end = uni(cs[ch.cell_base + (col + 1u) * ch.nz]);
    ch.s_end = min(ch.s0 + 64u, col_end);
    ch.rx = (int)(col / ch.ny);
    ch.ry = (int)(col % ch.ny);
    const int zmin = (int)(local % ch.nz);
    const int zmax = (int)((uni(skey[ch.s_end - 1u]) - ch.cell_base) % ch.nz);
    s = ch.s0 + lane;
    active = s < ch.s_end;
    const float qnan = __int_as_float(0x7fc00000);
    pi = active ? spm[s] : make_float4(qnan, qnan, qnan, 0.0f);
    int fx = 0, fy = 0, fz = 0;
    if (active) {
        fx = axis_edge_flags(pi.x, c.inv_cell, c.r2);
        fy = axis_edge_flags(pi.y, c.inv_cell, c.r2);
        fz = axis_edge_flags(pi.z, c.inv_cell, c.r2);
    }
    const int fxl = __ballot(fx & 1) != 0, fxh = __ballot(fx & 2) != 0;
    const int fyl = __ballot(fy & 1) != 0, fyh = __ballot(fy & 2) != 0;
    const int fzl = __ballot(fz & 1) != 0, fzh = __ballot(fz & 2) != 0;
    ch.dx0 = unii(-1 - fxl);
    ch.dx1 = unii(1 + fxh);
    ch.dy0 = unii(-1 - fyl);
    ch.dy1 = unii(1 + fyh);
    ch.zlo = unii(zmin - 1 - fzl);
    ch.zhi = unii(zmax + 1 + fzh);
    ch.zfirst = zmin;
    return true;
}

// Per-lane walk over the set bits of one segment, in candidate order; refills from the next
// non-empty word by selects (no inner loop).  Word kSeg - 1 of the segment is a sentinel
// (bit 31 in every lane, first slot = StepConsts::pad_slot, see init_sentinel): once a
// lane's bits are used up it returns the padding slot on every call, so a wavefront can run
// its largest lane count with no per-lane condition.
template <int kSeg>
struct BitWalk {
    uint32_t m = 0, jb = 0, ne = 0;
    uint32_t mq = 0, jq = 0;  // the next non-empty word (or the sentinel), read ahead
    // starts the walk of a segment whose non-empty words are `nonempty`; the first word is
    // read here, and each refill reads the one after, so a refill never waits for LDS
    // `ne` always holds the sentinel word's bit, the highest, which is never cleared
    __device__ __forceinline__ void begin(uint32_t nonempty, const uint32_t* __restrict__ lmask,
                                          const uint32_t* __restrict__ lwj0, uint32_t lane) {
        ne = nonempty | (1u << (kSeg - 1));
        m = 0u;
        fetch(lmask, lwj0, lane);
    }
    __device__ __forceinline__ void fetch(const uint32_t* __restrict__ lmask,
                                          const uint32_t* __restrict__ lwj0, uint32_t lane) {
        const uint32_t w = (uint32_t)__builtin_ctz(ne);
        mq = lmask[w * 64 + lane];
        jq = lwj0[w];
        ne = (ne & (ne - 1u)) | (1u << (kSeg - 1));
    }
    // next candidate: its slot (kShift 0) or its byte offset in rows of 2^kShift bytes, so
    // a gather needs one 32-bit add-shift and a load with the row base in SGPRs
    template <int kShift = 0>
    __device__ __forceinline__ uint32_t next(const uint32_t* __restrict__ lmask,
                                             const uint32_t* __restrict__ lwj0, uint32_t lane) {
        if (m == 0u) {
            m = mq;
            jb = jq << kShift;
            fetch(lmask, lwj0, lane);
        }
        const uint32_t k = (uint32_t)__builtin_clz(m);
        m ^= 0x80000000u >> k;
        return jb + (k << kShift);
    }
};

// Row at byte offset `off` (below 2^32) of an array in global memory.
template <class T>
__device__ __forceinline__ const T& row_at(const T* __restrict__ base, uint32_t off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

// Writes the sentinel word kSeg - 1 of a wavefront's segment buffers (once per wavefront;
// phase 1 and the stored-word loads never overwrite it).
template <int kSeg>
__device__ __forceinline__ void init_sentinel(uint32_t* __restrict__ lmask,
                                              uint32_t* __restrict__ lwj0, uint32_t lane,
                                              const StepConsts& c) {
    lmask[(kSeg - 1) * 64 + lane] = 0x80000000u;
    if (lane == 0) lwj0[kSeg - 1] = c.pad_slot;
}

__device__ __forceinline__ void add_nbr_stats(uint32_t cnt, unsigned long long* stats,
                                              uint32_t lane) {
    uint32_t mx = cnt;
    unsigned long long sm = cnt;
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
        sm += __shfl_xor(sm, off);
    }
    if (lane == 0) {
        atomicMax(&stats[0], (unsigned long long)mx);
        atomicAdd(&stats[1], sm);
    }
}

// Stores the LDS words [0, nw) of segment `seg_base` of this chunk into HBM (nontemporal: the
// force walk reads them once, long after, and the rows the walks gather should keep L2).
__device__ __forceinline__ void store_words(uint32_t item, uint32_t max_words, int seg_base,
                                            int nw,
                                            const uint32_t* __restrict__ lmask,
                                            const uint32_t* __restrict__ lwj0,
                                            uint32_t* __restrict__ gmask,
                                            uint32_t* __restrict__ gwj0, uint32_t lane) {
    const size_t base = (size_t)item * max_words;
    for (int w = 0; w < nw; ++w) {
        int gw = seg_base + w;
        if (gw >= (int)max_words) break;
        __builtin_nontemporal_store(lmask[w * 64 + lane], &gmask[(base + gw) * 64 + lane]);
    }
    for (int w = (int)lane; w < nw; w += 64) {
        int gw = seg_base + w;
        if (gw < (int)max_words) gwj0[base + gw] = lwj0[w];
    }
}


// ---------------------------------------------------------------------------
// Pair walk: two lanes per target (l and l ^ 32), 32 targets per wavefront.  The non-empty
// words of each target are compacted into an LDS list; the lower lane takes neighbour bits
// from the front, the upper one from the back, each its share of the count, so every lane
// walks continuously (no per-segment maxima) and the two never meet.
// ---------------------------------------------------------------------------
// Wave max of a per-lane count.
__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(x, off);
        x = o > x ? o : x;
    }
    return uni(x);
}

// Ping-pong segments of a wavefront: the older segment's words at LDS words [0, kSeg - 1),
// the newer one's at [kSeg, 2 kSeg - 1), the sentinel at word 2 kSeg - 1; the bit walk runs
// over both (its set holds older words in bits 0.., newer ones in bits kSeg..), so a lane
// through the older segment goes straight on into the newer.  `rem` counts a lane's bits
// not yet walked: every step takes one (the read-ahead word is always the next real one, see
// add_newer), so the wavefront knows how many steps empty the older buffer.
template <int kSeg>
struct PingPong {
    static constexpr uint32_t kSentinel = 1u << (2 * kSeg - 1);
    static constexpr uint32_t kLow = (1u << (kSeg - 1)) - 1u;
    BitWalk<2 * kSeg> bw;
    uint32_t rem = 0;      // this lane's bits not yet walked, both buffers
    uint32_t cnt_new = 0;  // this lane's bits in the newer buffer
    int nw_new = 0;        // uniform: words in the newer buffer (0: free)
    __device__ __forceinline__ void init(uint32_t pad_slot) {
        bw.m = 0u;
        bw.jb = 0u;
        bw.ne = kSentinel;
        bw.mq = 0x80000000u;  // read-ahead: the sentinel word
        bw.jq = pad_slot;
    }
    // A segment of `nw` words written to the newer buffer joins the walk; a lane whose
    // read-ahead is the sentinel reads again, so no step is lost on it.
    __device__ __forceinline__ void add_newer(const SegLane& sl, int nw,
                                              const uint32_t* __restrict__ lmask,
                                              const uint32_t* __restrict__ lwj0, uint32_t lane,
                                              uint32_t pad_slot) {
        bw.ne |= sl.nonempty << kSeg;
        rem += sl.count;
        cnt_new = sl.count;
        nw_new = nw;
        if (bw.jq == pad_slot) bw.fetch(lmask, lwj0, lane);
    }
    // add_newer for a segment already in the newer buffer whose per-lane bits were not kept
    // (a segment of the next part, held until the current part is walked).
    __device__ __forceinline__ void add_held(int nw, const uint32_t* __restrict__ lmask,
                                             const uint32_t* __restrict__ lwj0, uint32_t lane,
                                             uint32_t pad_slot) {
        SegLane sl = {0u, 0u};
#pragma unroll
        for (int w = 0; w < kSeg - 1; ++w) {
            const uint32_t word = w < nw ? lmask[(kSeg + w) * 64 + lane] : 0u;
            sl.nonempty |= (word != 0u ? 1u : 0u) << w;
            sl.count += (uint32_t)__builtin_popcount(word);
        }
        add_newer(sl, nw, lmask, lwj0, lane, pad_slot);
    }
    // Steps after which every lane is through the older buffer; steps that walk all.
    __device__ __forceinline__ uint32_t older_steps() const {
        return wave_max(rem > cnt_new ? rem - cnt_new : 0u);
    }
    __device__ __forceinline__ uint32_t all_steps() const { return wave_max(rem); }
    __device__ __forceinline__ void walked(uint32_t steps) { rem = rem > steps ? rem - steps : 0u; }
    // Every lane is through the older buffer: the newer segment's words move down to it.
    // All kSeg - 1 words are copied (those past the segment are never read: no bit of `ne`
    // names them), four at a time, each group's reads issued before its writes.
    __device__ __forceinline__ void shift_down(uint32_t* __restrict__ lmask,
                                               uint32_t* __restrict__ lwj0, uint32_t lane) {
        wave_sync();
        const uint32_t j = lwj0[kSeg + (lane < kSeg - 1 ? lane : 0u)];
#pragma unroll
        for (int w0 = 0; w0 < kSeg - 1; w0 += 4) {
            uint32_t v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (w0 + k < kSeg - 1) v[k] = lmask[(kSeg + w0 + k) * 64 + lane];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (w0 + k < kSeg - 1) lmask[(w0 + k) * 64 + lane] = v[k];
        }
        if (lane < kSeg - 1) lwj0[lane] = j;
        bw.ne = ((bw.ne >> kSeg) & kLow) | kSentinel;
        nw_new = 0;
        wave_sync();
    }
};

constexpr int kCols = 32;       // target columns of a wavefront (two lanes per target)
constexpr int kListCap = 72;    // non-empty words per target column held at once
constexpr int kBatch = 16;      // stored words requested per batch while building a list

// LDS of one wavefront's walk: per target column, its non-empty words (mask, first slot)
// in candidate order, entry i of column c at list[i * 32 + c].
struct alignas(16) WalkLds {
    uint2 list[kListCap * kCols];
};

// Appends the words of a window to the column lists, up to kListCap non-empty words per
// column.  Stored words are read from the density pass; otherwise candidates are tested here
// (phase 1 with the contracted filter).  Returns false when the chunk has no words left.
struct ListBuilder {
    int wpos = 0;        // next stored word
    uint32_t n_ne = 0;   // non-empty words of this lane's column in the window
    uint32_t cnt = 0;    // neighbour bits of this lane's column in the window
};

__device__ __forceinline__ bool build_list(ListBuilder& lb, bool stored, int total_words,
                                           const WalkBuffers& wb, uint32_t item,
                                           uint32_t clane, Cursor& cu, const Chunk& ch,
                                           const float4* spm, const uint32_t* cs_,
                                           const StepConsts& c, float4 pi, uint32_t lane,
                                           WalkLds& L) {
    const uint32_t col = lane & 31u;
    const bool writer = lane < 32u;
    lb.n_ne = 0u;
    lb.cnt = 0u;
    bool any = false;
    if (stored) {
        const size_t base = (size_t)item * c.max_words;
        // a window of up to kListCap words cannot overflow a column list
        const int end = min(total_words, lb.wpos + kListCap);
        while (lb.wpos < end) {
            const int nb = min(kBatch, end - lb.wpos);
            uint32_t wv[kBatch];
            uint32_t jv[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) {
                const int w = lb.wpos + min(k, nb - 1);
                wv[k] = wb.mask[(base + w) * 64 + clane];
                jv[k] = uni(wb.wj0[base + w]);
            }
#pragma unroll
            for (int k = 0; k < kBatch; ++k) {
                if (k < nb && wv[k] != 0u) {
                    if (writer) L.list[lb.n_ne * kCols + col] = make_uint2(wv[k], jv[k]);
                    ++lb.n_ne;
                    lb.cnt += (uint32_t)__builtin_popcount(wv[k]);
                }
            }
            lb.wpos += nb;
            any = true;
        }
    } else {
        while (__ballot(lb.n_ne >= (uint32_t)kListCap) == 0ull && cursor_settle(cu, ch, cs_, c)) {
            const uint32_t j0 = uni(cu.j);
            const uint32_t cnt = uni(min(32u, cu.end - j0));
            uint32_t word = 0;
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4 q = spm[j0 + k];
                word = push_bit(word, dist2_filter(pi.x, pi.y, pi.z, q.x, q.y, q.z), c.r2_filter);
            }
            word <<= (32u - cnt) & 31u;  // left-aligned: candidate k in bit 31 - k
            if (word != 0u) {
                if (writer) L.list[lb.n_ne * kCols + col] = make_uint2(word, j0);
                ++lb.n_ne;
                lb.cnt += (uint32_t)__builtin_popcount(word);
            }
            cu.j = j0 + cu.step;
            any = true;
        }
    }
    return any;
}

// Walk of one lane over its target's list: the lower lane of a target (p = 0) takes bits
// from the front, the upper one (p = 1) from the back, each its share of the count, so the
// two never meet.  `e` is the list entry at `idx`, read one step ahead.
struct PairWalk {
    uint32_t m = 0u, jb = 0u;
    int idx = 0, dir = 1;
    uint2 e;
    __device__ __forceinline__ void begin(uint32_t p, uint32_t n_ne) {
        m = 0u;
        jb = 0u;
        dir = p ? -1 : 1;
        idx = p ? (int)n_ne - 1 : 0;
    }
    __device__ __forceinline__ void request(const WalkLds& L, uint32_t col) {
        e = L.list[(uint32_t)max(idx, 0) * kCols + col];
    }
    __device__ __forceinline__ uint32_t next(uint32_t p) {
        const bool refill = m == 0u;
        m = refill ? e.x : m;
        jb = refill ? e.y : jb;
        idx += refill ? dir : 0;
        const uint32_t bit = p ? (uint32_t)__builtin_ctz(m) : 31u - (uint32_t)__builtin_clz(m);
        m ^= 1u << bit;
        return jb + (31u - bit);
    }
};

// Walks every neighbour j != i of the wavefront's 32 targets, two lanes per target, in
// uniform steps; body(edge, pj, vj, d2) is called once per step (edge = false on lanes with
// nothing left).  Neighbour data is requested one step ahead (vj only with kVel).
template <bool kVel = true, class Body>
__device__ __forceinline__ void walk_pairs(uint32_t item, const Chunk& ch, uint32_t s,
                                           bool active, float4 pi, uint32_t clane,
                                           const WalkBuffers& wb, const float4* spm,
                                           const float4* svr, const uint32_t* cs_,
                                           const StepConsts& c, WalkLds& L, uint32_t lane,
                                           Body&& body) {
    const uint32_t col = lane & 31u, p = lane >> 5;
    const int total_words = (int)uni(wb.nwords[item]);
    const bool stored = total_words <= (int)c.max_words;
    const uint32_t s_safe = active ? s : ch.s0;
    Cursor cu;
    if (!stored) cursor_begin(cu, ch, cs_, c);
    ListBuilder lb;
    for (;;) {
        if (!build_list(lb, stored, total_words, wb, item, clane, cu, ch, spm, cs_, c, pi, lane,
                        L))
            break;
        wave_sync();
        const uint32_t quota = p ? lb.cnt >> 1 : (lb.cnt + 1u) >> 1;
        const uint32_t tmax = wave_max(quota);
        PairWalk pw;
        pw.begin(p, lb.n_ne);
        pw.request(L, col);
        uint32_t jn = s_safe;
        if (quota > 0u) jn = pw.next(p);
        pw.request(L, col);
        float4 pn = spm[jn];
        float4 vn = kVel ? svr[jn] : make_float4(0.f, 0.f, 0.f, 0.f);
        for (uint32_t t = 0; t < tmax; ++t) {
            const uint32_t j = jn;
            const bool has = t < quota;
            const float4 pj = pn;
            const float4 vj = vn;
            jn = s_safe;
            if (t + 1u < quota) jn = pw.next(p);
            pw.request(L, col);
            pn = spm[jn];
            if (kVel) vn = svr[jn];
            const float d2 = dist2_exact(pi.x, pi.y, pi.z, pj.x, pj.y, pj.z);
            const bool edge = has && j != s && d2 <= c.r2;  // the reference's test, self excluded
            body(edge, pj, vj, d2);
        }
        wave_sync();
    }
}


// kPP: two buffers of kSeg words each (ping-pong): while the wavefront walks the older
// segment, the next one is already in the newer buffer, so a lane that is through the older
// segment goes on with the newer instead of waiting for the slowest lane.
template <int kSeg, int kWaves, bool kPP = false>
struct WalkShape {
    static constexpr int seg = kSeg;
    static constexpr int waves = kWaves;
    static constexpr int threads = 64 * kWaves;
    static constexpr bool pp = kPP;
    static constexpr int lds_words = kPP ? 2 * kSeg : kSeg;  // LDS words per wavefront
    static_assert(!kPP || 2 * kSeg <= 32, "both buffers' words fit one 32-bit set");
    // occupancy target of the force walk: 5 wavefronts per SIMD where the LDS allows it
    // (24- and 28-word segments without the LDS running total), else what the registers give
    static constexpr int force_waves = (!kPP && (kSeg == 24 || kSeg == 28)) ? 5 : 1;
    // occupancy target of the single-wavefront density walk
    static constexpr int density_waves = !kPP ? 7 : kSeg <= 12 ? 6 : 4;
};

}  // namespace

}  // namespace sphk
