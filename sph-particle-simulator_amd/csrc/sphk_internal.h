// sphk_internal.h -- device-side layout and launcher declarations of libsphk.
//
// HBM layout (one engine, capacity C particles):
//   home arrays, caller (id) order, float4 rows:
//     pm [C]  (x, y, z, mass)         vel [C] (vx, vy, vz, 0)
//     acc[C]  (ax, ay, az, 0)         rho[C], prs[C]  (capacity-length, like densities_/pressures_)
//   per-step work arrays, cell-sorted order:
//     skey[C]  dense cell index        sid[C]  caller id
//     spm [C]  (x, y, z, mass)         svr[C]  (vx, vy, vz, rho)
//     cell_start[G+2]  prefix of cell occupancy, z-fastest dense grid, G = nx*ny*nz;
//                      dense index G is the bin of particles with non-finite positions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sphk.h"

namespace sphk {

// Kernel-side constants of one step shared by every scene.  Derived on the host from
// sphk_params with the reference's own FP32 expressions (src/kernels.cpp:23-36,
// src/spatial_hash.h:63-66).
struct StepConsts {
    float inv_cell;  // 1.0f / neighbor_search_radius
    float r2;        // neighbor_search_radius^2
    float r2_filter; // r2 raised by 8 ulp: threshold of the contracted FP32 candidate filter
    float h, inv_h;
    float sigma;     // 1/(pi h^3) (x 3/2 for the consistent formulation)
    float w0;        // W(0) = sigma * (2/3)
    float grad_scale;  // 4 sigma / h          (applied once per particle)
    float lap_scale;   // 2 sigma * mu / h^2   (applied once per particle)
    float k_gas, rho0, grav, damping;
    int32_t formulation;
    int32_t cut;      // 1 when the search radius is below the kernel support 2h: the walks
                      // then drop pairs beyond the radius explicitly (else their weight is
                      // zero or below 2^-75 of a neighbour's)
    uint32_t G;       // cells of all scene grids; particles with dense index G have a
                      // non-finite coordinate
    uint32_t ncols;   // (x, y) columns of all scene grids
    uint32_t nscenes;
    uint32_t pad_slot;  // sorted slot past every particle (index = capacity): zero mass and
                        // weights, far away; the walks read it in place of a neighbour
    int32_t qneg;       // -(T + 1), T the threshold of the quantized candidate filter
    uint32_t max_words; // candidate words stored per chunk (the word buffers' stride); chunks
                        // with more are tested again by the force walk
    double qscale;      // quantized coordinate = floor(x * qscale): 256 units per cell
};

// Neighbour row of the force walk, one per sorted slot: a = (x, y, z, m/rho),
// b = (vx, vy, vz, m P / (4 rho^2)); both halves share one 32-byte block.
struct alignas(32) ForceRow {
    float4 a, b;
};

// Dense cell grid of one scene (z fastest) and its place in the global cell/column
// numbering.  Scenes are independent: no neighbour search crosses two scene grids.
struct SceneGrid {
    int32_t ox, oy, oz;  // origin cell (absolute cell coordinates)
    int32_t nx, ny, nz;  // dims, including two padding cells per side
    uint32_t cell_base;  // first global cell index
    uint32_t col_base;   // first global column index
    uint32_t pbegin, pend;  // particle (caller id) range
    float lo[3], hi[3];  // wall clamp bounds
};

struct Grid {
    int32_t ox = 0, oy = 0, oz = 0;
    int32_t nx = 0, ny = 0, nz = 0;
    uint64_t G = 0;
};

// radix sort tile geometry
constexpr int kRadixThreads = 256;
constexpr int kRadixWaves = kRadixThreads / 64;
constexpr int kRadixItems = 8;  // keys per thread
constexpr int kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixMaxBits = 11;

// Largest engine capacity: (capacity + 1) rows of 32 bytes stay below 2^32 bytes, the range
// of the walks' 32-bit row offsets.
constexpr uint64_t kMaxCapacity = (1ull << 27) - 2;

// neighbour walks: candidate words (32 candidates each) stored per chunk in HBM for the force
// walk: kMaxWords, or kMinWords where the word buffers would pass kWordBufferBudget bytes.
// Chunks next to a free surface can reach ~160 words (a sparse column beside full ones);
// one the force walk must test again outlasts the rest of its launch.
constexpr int kMaxWords = 256;
constexpr int kMinWords = 128;
constexpr uint64_t kWordBufferBudget = 16ull << 30;
constexpr int kNumWalkShapes = 9;  // launch shapes selectable per walk (see walk_shape_name)
const char* walk_shape_name(int v);
// Parts of a chunk's candidate columns (corners, edges, centre with the outer ring).
constexpr int kParts = 3;
// Shapes that walk each chunk with three wavefronts, one per part of its candidate columns;
// they need the per-part sums of WalkBuffers.
constexpr bool walk_shape_split(int v) { return v == 5 || v == 6; }
// Shapes an engine uses until the caller sets its own: below kSplitBelow particles the
// single-wavefront walk leaves most of the GPU idle and the split seg32_waves4 is faster;
// above, seg24_waves4 for the density walk and pp16_waves4 for the force walk (the fastest
// pair on the 1M dam break on MI355X).
constexpr int kShapeSmall = 6;
constexpr int kShapeLargeDensity = 4;
constexpr int kShapeLargeForce = 8;
constexpr uint64_t kSplitBelow = 65536;

// HBM buffers of the chunked walks (sized for max_chunks())
struct WalkBuffers {
    uint32_t* items = nullptr;    // first sorted slot of each chunk
    uint32_t* nitems = nullptr;   // [1] number of chunks
    uint32_t* colcnt = nullptr;   // [ncols + 1] per-column chunk counts -> offsets
    uint32_t* mask = nullptr;     // [chunks][max_words][64] neighbour bits
    uint32_t* wj0 = nullptr;      // [chunks][max_words] first slot of each word
    uint32_t* nwords = nullptr;   // [chunks] words of each chunk
    uint32_t* pstarts = nullptr;  // [chunks] first stored word of parts 1 and 2, 16 bits each
    size_t chunk_cap = 0;
    uint32_t max_words = 0;       // stride of mask and wj0 (kMaxWords or kMinWords)
    size_t col_cap = 0;
    // per-part sums of the split shapes, [3][part_cap] by sorted slot (allocated only
    // while a split shape is selected)
    float* part_acc = nullptr;     // density sums
    uint32_t* part_cnt = nullptr;  // neighbour counts
    float4* part_force = nullptr;  // [3][part_cap][2]: pressure sum, viscosity sum
    size_t part_cap = 0;
};

// Device view of the per-part sums.
struct PartSums {
    float* acc;
    uint32_t* cnt;
    float4* force;
    size_t cap;
};

// Scene lookups by binary search over the scene table (few scenes; uniform per wave where
// used by the walks).
__device__ __forceinline__ uint32_t scene_of_cell(const SceneGrid* __restrict__ sc, uint32_t ns,
                                                  uint32_t L) {
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (sc[mid].cell_base <= L) lo = mid; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t scene_of_col(const SceneGrid* __restrict__ sc, uint32_t ns,
                                                 uint32_t col) {
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (sc[mid].col_base <= col) lo = mid; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t scene_of_particle(const SceneGrid* __restrict__ sc,
                                                      uint32_t ns, uint32_t id) {
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (sc[mid].pbegin <= id) lo = mid; else hi = mid;
    }
    return lo;
}

// Exclusive prefix of v over the 256 threads of a workgroup (all threads call it); *total, if
// given, receives the workgroup's sum.  lds_waves: 4 words of LDS.
__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* lds_waves,
                                                             uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (uint32_t w = 0; w < 4u; ++w) {
        uint32_t t = lds_waves[w];
        if (w < wave) wbase += t;
        tot += t;
    }
    if (total) *total = tot;
    return wbase + x - v;
}

// Device-side arguments of the step's last part for the particles of the non-finite bin
// (they have no neighbour: rho = m W(0), a = (0, g, 0), integrated and clamped like the rest),
// done by the chunk-list launch of a step's build; it also advances the clock by dt.
struct StepTail {
    const float4* spm;
    const float4* svr;
    const uint32_t* sid;
    uint32_t n;
    const float* dt_dev;  // adaptive dt on the device, or null for `dt`
    float dt;
    float4* pm;
    float4* vel;
    float4* acc;
    float* rho;
    float* prs;
    float* time_dev;
};

// launchers; all enqueue on `st`
size_t scan_scratch_words(size_t m);
hipError_t scan_exclusive(uint32_t* data, uint32_t m, uint32_t* scratch, hipStream_t st);
// one-launch scan; `state` (lookback_state_words(m) words) is zero before and after each call;
// a look-back that waits past its bound sets bit 1 of *err
size_t lookback_state_words(size_t m);
hipError_t scan_exclusive_lookback(uint32_t* data, uint32_t m, uint32_t* state, uint32_t* err,
                                   hipStream_t st);
// stable LSD sort of the particles' dense cell indices (computed from pm in the first pass and
// left in `keys`) with their caller ids as values
hipError_t radix_sort_keys(const float4* pm, const StepConsts& c, const SceneGrid* scenes,
                           uint32_t* err_flag, uint32_t* keys, uint32_t* keys_alt, uint32_t* vals,
                           uint32_t* vals_alt, uint32_t* hist, uint32_t* lookback,
                           uint32_t* scan_scratch, uint32_t n, int total_bits,
                           uint32_t** keys_out, uint32_t** vals_out, hipStream_t st);
size_t radix_hist_words(uint32_t n, int total_bits);
// sorted copies (spm, svr, sq) and the cell table cell_start[0, G + 2), one launch
hipError_t launch_gather_cells(const float4* pm, const float4* vel, const uint32_t* sid,
                               const uint32_t* skey, uint32_t n, uint32_t G, double qscale,
                               uint2* sq, float4* spm, float4* svr, uint32_t* cell_start,
                               hipStream_t st);
uint32_t max_chunks(uint32_t n, const StepConsts& c);
// chunk list of the sorted particles; with `tail`, also the non-finite particles' step
hipError_t launch_chunks(const uint32_t* cell_start, const StepConsts& c, const SceneGrid* scenes,
                         uint32_t* cnt_scratch, uint32_t* scan_scratch, uint32_t* items,
                         uint32_t* nitems, const StepTail* tail, hipStream_t st);
hipError_t launch_density_walk(const WalkBuffers& wb, const float4* spm, const uint2* sq,
                               float4* svr,
                               ForceRow* rows, const uint32_t* skey,
                               const uint32_t* cell_start, uint32_t n,
                               const StepConsts& c, const SceneGrid* scenes, int shape,
                               hipStream_t st);
hipError_t launch_force_walk(const WalkBuffers& wb, const float4* spm, const uint2* sq,
                             const float4* svr,
                             const ForceRow* rows, const uint32_t* skey,
                             const uint32_t* sid,
                             const uint32_t* cell_start, uint32_t n, const StepConsts& c,
                             const SceneGrid* scenes, const float* dt_dev, float dt, float4* pm,
                             float4* vel, float4* acc, float* rho, float* prs, int shape,
                             unsigned long long* nbr_stats, hipStream_t st);
hipError_t launch_nbr_lists(const WalkBuffers& wb, const float4* spm, const uint2* sq,
                            const uint32_t* skey,
                            const uint32_t* sid, const uint32_t* cell_start, uint32_t n,
                            const StepConsts& c, const SceneGrid* scenes, int pass,
                            uint32_t* count_by_id, const int64_t* offsets_by_id, int64_t* idx,
                            int shape, hipStream_t st);
hipError_t launch_cfl(const float4* vel, const float4* acc, uint32_t n, const sphk_params& p,
                      uint32_t* scratch_u32, float* dt_dev, hipStream_t st);
// time += dt (or *dt_dev), `reps` times in sequence (the clock of `reps` steps)
hipError_t launch_advance_time(float* time_dev, const float* dt_dev, float dt, hipStream_t st,
                               int reps = 1);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
// renderer instance data (9 floats per particle: position, velocity, colour), caller order
hipError_t launch_render_pack(const float4* pm, const float4* vel, const float* rgb, uint32_t n,
                              float* out, hipStream_t st);
hipError_t launch_copy_state(const float4* a0, const float4* a1, const float4* a2, float4* b0,
                             float4* b1, float4* b2, size_t n, const float* t_src,
                             float* t_dst, hipStream_t st);

}  // namespace sphk
