// sphk_kernels.hip -- gfx950 kernels of the SPH step.
//
// Step pipeline (one HIP stream):
//   radix sort     stable LSD sort of (cell, caller id) -> cell-sorted order; the first pass
//                  computes the dense cell index of every particle (caller order)
//   gather+cells   cell-sorted float4 copies of (x,y,z,m) and (vx,vy,vz,-), quantized
//                  coordinates of the candidate filter; beside them (same launch) the
//                  occupancy prefix of the dense grid (binary search per cell)
//   chunks         64-particle work items per (x,y) column of the grid
//   density, force walks   (sphk_walk.hip)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "sphk_internal.h"

namespace sphk {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// ---------------------------------------------------------------------------
// dense cell index per particle (caller order)
// ---------------------------------------------------------------------------
// Dense index of particle i at p: its cell in its scene's grid, or G (the non-finite bin) for a
// non-finite coordinate.  A finite particle outside the grid's interior sets *err.
// kOne: a one-scene grid, whose geometry g0 is loaded once instead of looked up per particle.
template <bool kOne>
__device__ __forceinline__ uint32_t cell_key(const float4& p, uint32_t i, const StepConsts& c,
                                             const SceneGrid* __restrict__ scenes,
                                             const SceneGrid& g0, uint32_t* __restrict__ err) {
    uint32_t L = c.G;
    if (__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z)) {
        const SceneGrid& g = kOne ? g0 : scenes[scene_of_particle(scenes, c.nscenes, i)];
        float fx = floorf(p.x * c.inv_cell);
        float fy = floorf(p.y * c.inv_cell);
        float fz = floorf(p.z * c.inv_cell);
        bool inr = fx >= -1073741824.0f && fx < 1073741824.0f && fy >= -1073741824.0f &&
                   fy < 1073741824.0f && fz >= -1073741824.0f && fz < 1073741824.0f;
        int rx = inr ? (int)fx - g.ox : -1;
        int ry = inr ? (int)fy - g.oy : -1;
        int rz = inr ? (int)fz - g.oz : -1;
        if (rx >= 2 && rx < g.nx - 2 && ry >= 2 && ry < g.ny - 2 && rz >= 2 && rz < g.nz - 2) {
            L = g.cell_base +
                ((uint32_t)rx * (uint32_t)g.ny + (uint32_t)ry) * (uint32_t)g.nz + (uint32_t)rz;
        } else {
            atomicOr(err, 1u);
        }
    }
    return L;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of 32-bit keys with 32-bit values.  Each pass:
//   upsweep   per-tile digit histogram (LDS atomics)   -> hist[digit][tile]
//   scan      exclusive scan of hist (digit-major)     -> global base per (digit, tile)
//             (done inside the downsweep when hist is small)
//   downsweep per-wave stable ranks from ballot-matched peer groups, per-digit
//             prefix over the waves of the tile, scatter.
// Items of a tile are owned wave-contiguously (wave w owns 512 consecutive keys,
// 64 per iteration), so (wave, iteration, lane) order is input order and every
// pass is stable.
// ---------------------------------------------------------------------------
// kKeys: the first pass also computes the keys from the positions (cell_key) and stores them
// for its downsweep, so the keys need no launch of their own.
template <bool kKeys>
__global__ __launch_bounds__(kRadixThreads) void k_radix_upsweep(
    const uint32_t* __restrict__ keys, uint32_t n, int shift, uint32_t mask, uint32_t ntiles,
    uint32_t* __restrict__ hist, const float4* __restrict__ pm, StepConsts c,
    const SceneGrid* __restrict__ scenes, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ err) {
    __shared__ uint32_t h[1 << kRadixMaxBits];
    const uint32_t nb = mask + 1;
    for (uint32_t i = threadIdx.x; i < nb; i += kRadixThreads) h[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRadixTile;
    // all of a thread's loads are issued before its first histogram update
    uint32_t key[kRadixItems];
    if constexpr (kKeys) {
        float4 p[kRadixItems];
#pragma unroll
        for (int k = 0; k < kRadixItems; ++k) {
            const uint32_t idx = base + k * kRadixThreads + threadIdx.x;
            p[k] = idx < n ? pm[idx] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const SceneGrid g0 = scenes[0];
        if (c.nscenes == 1) {
#pragma unroll
            for (int k = 0; k < kRadixItems; ++k) {
                const uint32_t idx = base + k * kRadixThreads + threadIdx.x;
                if (idx < n) {
                    key[k] = cell_key<true>(p[k], idx, c, scenes, g0, err);
                    keys_out[idx] = key[k];
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kRadixItems; ++k) {
                const uint32_t idx = base + k * kRadixThreads + threadIdx.x;
                if (idx < n) {
                    key[k] = cell_key<false>(p[k], idx, c, scenes, g0, err);
                    keys_out[idx] = key[k];
                }
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kRadixItems; ++k) {
            const uint32_t idx = base + k * kRadixThreads + threadIdx.x;
            key[k] = idx < n ? keys[idx] : 0u;
        }
    }
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
        const uint32_t idx = base + k * kRadixThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[(key[k] >> shift) & mask], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += kRadixThreads) hist[i * ntiles + blockIdx.x] = h[i];
}

// Multi-block exclusive scan of uint32 (three launches: tile sums, scan of the tile sums,
// tile scans with the tile offset).  Tiles of 4096 values, 16 per thread.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanPer;

__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(const uint32_t* __restrict__ data,
                                                                  uint32_t m,
                                                                  uint32_t* __restrict__ sums) {
    __shared__ uint32_t lw[kScanThreads / 64];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        uint32_t i = base + k;
        if (i < m) acc += data[i];
    }
    uint32_t tot;
    block_exclusive_scan_256(acc, lw, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(uint32_t* __restrict__ data, uint32_t m,
                                                             const uint32_t* __restrict__ offs) {
    __shared__ uint32_t lw[kScanThreads / 64];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        uint32_t i = base + k;
        v[k] = i < m ? data[i] : 0u;
        acc += v[k];
    }
    uint32_t run = (offs ? offs[blockIdx.x] : 0u) + block_exclusive_scan_256(acc, lw, nullptr);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        uint32_t i = base + k;
        if (i < m) data[i] = run;
        run += v[k];
    }
}

// Single-launch exclusive scan of uint32 with decoupled look-back.  `state` holds a ticket
// counter, a completion counter and one status word per tile, all zero between calls.  Each
// workgroup takes the next ticket as its tile (tickets follow start order), publishes its tile
// sum (kAggregate), then wavefront 0 reads the status words of the 64 tiles before it and adds
// them back to the nearest inclusive prefix (kInclusive), moving 64 tiles further back while none
// is in sight, and publishes its own inclusive prefix.  A tile waits only for tiles that took
// their tickets before it and publish their sums without waiting on anything, so every wait
// ends; the wait is also bounded (kLookbackSpins, then *err is set and the scan is wrong but
// finishes).  The last workgroup to finish clears the counters and status words.  Values are
// below 2^30 (sums of at most 2^27 particles).
constexpr uint32_t kAggregate = 1u << 30;
constexpr uint32_t kInclusive = 2u << 30;
constexpr uint32_t kStatusValue = (1u << 30) - 1u;
constexpr uint32_t kLookbackSpins = 1u << 22;
constexpr uint32_t kLookbackMaxTiles = 64;

__device__ __forceinline__ uint32_t status_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kScanThreads) void k_scan_lookback(uint32_t* __restrict__ data,
                                                                uint32_t m, uint32_t ntiles,
                                                                uint32_t* __restrict__ state,
                                                                uint32_t* __restrict__ err) {
    __shared__ uint32_t lw[kScanThreads / 64];
    __shared__ uint32_t s_tile, s_prefix, s_last;
    uint32_t* status = state + 2;
    if (threadIdx.x == 0) s_tile = atomicAdd(&state[0], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile < ntiles) {
        const uint32_t base = tile * kScanTile + threadIdx.x * kScanPer;
        uint32_t v[kScanPer];
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const uint32_t i = base + k;
            v[k] = i < m ? data[i] : 0u;
            acc += v[k];
        }
        uint32_t tot;
        uint32_t run = block_exclusive_scan_256(acc, lw, &tot);
        if (threadIdx.x < 64) {
            const uint32_t lane = threadIdx.x;
            if (lane == 0)
                status_store(&status[tile], (tile == 0 ? kInclusive : kAggregate) | tot);
            uint32_t prefix = 0;
            int32_t hi = (int32_t)tile - 1;  // the window is tiles hi, hi - 1, ..., hi - 63
            uint32_t spins = 0;
            while (hi >= 0) {
                const int32_t j = hi - (int32_t)lane;
                // tiles before the first count as an inclusive prefix of zero
                const uint32_t w = j >= 0 ? status_load(&status[j]) : kInclusive;
                const uint64_t incl = __ballot((w & kInclusive) != 0u);
                const uint64_t wait = __ballot(w == 0u);
                // lanes up to the nearest inclusive prefix, or all 64 if none is in the window
                const uint64_t need = incl ? (~0ull >> (63 - __builtin_ctzll(incl))) : ~0ull;
                if (wait & need) {
                    if (++spins >= kLookbackSpins) {
                        if (lane == 0) atomicOr(err, 2u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t part = ((need >> lane) & 1ull) ? (w & kStatusValue) : 0u;
                for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
                prefix += part;
                if (incl) break;
                hi -= 64;
            }
            if (lane == 0) {
                if (tile > 0)
                    status_store(&status[tile], kInclusive | ((prefix + tot) & kStatusValue));
                s_prefix = prefix;
            }
        }
        __syncthreads();
        run += s_prefix;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const uint32_t i = base + k;
            if (i < m) data[i] = run;
            run += v[k];
        }
    } else if (threadIdx.x == 0) {
        // a ticket counter left non-zero (it is cleared at allocation and by every call's last
        // workgroup): flag it and touch no data; tiles waiting on a missing ticket give up at
        // their bound, and the call still clears the state below
        atomicOr(err, 2u);
    }
    // completion: the last workgroup clears the state for the next call (every other workgroup
    // is through its look-back when it counts itself done)
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&state[1], 1u) == ntiles - 1u;
    }
    __syncthreads();
    if (s_last) {
        for (uint32_t t = threadIdx.x; t < ntiles; t += kScanThreads) status_store(&status[t], 0u);
        if (threadIdx.x == 0) {
            status_store(&state[0], 0u);
            status_store(&state[1], 0u);
        }
    }
}

size_t lookback_state_words(size_t m) { return 2 + (m + kScanTile - 1) / kScanTile; }

hipError_t scan_exclusive_lookback(uint32_t* data, uint32_t m, uint32_t* state, uint32_t* err,
                                   hipStream_t st) {
    if (m == 0) return hipSuccess;
    const uint32_t tiles = (m + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_lookback, dim3(tiles), dim3(kScanThreads), 0, st, data, m, tiles,
                       state, err);
    return hipGetLastError();
}

size_t scan_scratch_words(size_t m) {
    size_t words = 0;
    while (m > kScanTile) {
        m = (m + kScanTile - 1) / kScanTile;
        words += m;
    }
    return words + 1;
}

// `scratch` holds scan_scratch_words(m) values: the tile sums of each recursion level.
hipError_t scan_exclusive(uint32_t* data, uint32_t m, uint32_t* scratch, hipStream_t st) {
    if (m == 0) return hipSuccess;
    const uint32_t tiles = (m + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, st, data, m,
                           (const uint32_t*)nullptr);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(tiles), dim3(kScanThreads), 0, st, data, m, scratch);
    hipError_t e = scan_exclusive(scratch, tiles, scratch + tiles, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, st, data, m, scratch);
    return hipGetLastError();
}

// Histograms of at most this many words (digits x tiles) are scanned by every downsweep
// workgroup for itself (kLocalScan), which saves the scan's launches on small inputs.
constexpr uint32_t kRadixLocalScanMax = 8192;

// kLocalScan: `offs` is the unscanned histogram hist[digit][tile] (at most kRadixLocalScanMax
// words), and each workgroup takes the exclusive prefix at its own tile's entries itself.
template <bool kLocalScan>
__global__ __launch_bounds__(kRadixThreads) void k_radix_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, int shift,
    uint32_t mask, int bits, uint32_t ntiles, const uint32_t* __restrict__ offs) {
    __shared__ uint32_t wc[kRadixWaves << kRadixMaxBits];
    const uint32_t nb = mask + 1;
    for (uint32_t i = threadIdx.x; i < kRadixWaves * nb; i += kRadixThreads) wc[i] = 0;
    // the tile's digit offsets (kLocalScan; declared in its branch so that the other kernel
    // keeps exactly 32 KB of LDS, five workgroups per CU)
    const uint32_t* loffs = nullptr;
    if constexpr (kLocalScan) {
        // thread t sums the words [t*per, (t+1)*per) of the digit-major histogram; the
        // exclusive prefix of the block gives each word's global prefix, kept for the words of
        // this tile (positions d * ntiles + blockIdx.x)
        __shared__ uint32_t lsw[kRadixWaves];
        __shared__ uint32_t lo[1 << kRadixMaxBits];
        loffs = lo;
        const uint32_t m = nb * ntiles;
        const uint32_t per = (m + kRadixThreads - 1) / kRadixThreads;
        const uint32_t beg = threadIdx.x * per;
        const uint32_t end = beg + per < m ? beg + per : m;
        uint32_t sum = 0;
        for (uint32_t i = beg; i < end; ++i) sum += offs[i];
        uint32_t run = block_exclusive_scan_256(sum, lsw, nullptr);
        // first position of this tile at or after beg
        uint32_t next = beg <= blockIdx.x ? blockIdx.x
                                          : blockIdx.x + ((beg - blockIdx.x + ntiles - 1) / ntiles) * ntiles;
        for (uint32_t i = beg; i < end; ++i) {
            if (i == next) {
                lo[i / ntiles] = run;
                next += ntiles;
            }
            run += offs[i];
        }
    }
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * kRadixTile + wave * (kRadixItems * 64);
    uint32_t* mywc = wc + wave * nb;
    uint32_t key[kRadixItems];
    uint32_t val[kRadixItems];
    uint32_t base_off[kRadixItems];
    uint32_t rank[kRadixItems];
    // keys, values and (unless scanned here) the digits' global offsets are all requested
    // before the ranking, which hides their latency
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
        const uint32_t idx = base + k * 64 + lane;
        key[k] = idx < n ? keys_in[idx] : 0u;
        val[k] = vals_in ? (idx < n ? vals_in[idx] : 0u) : idx;
    }
    if constexpr (!kLocalScan) {
#pragma unroll
        for (int k = 0; k < kRadixItems; ++k) {
            const uint32_t idx = base + k * 64 + lane;
            base_off[k] = idx < n ? offs[((key[k] >> shift) & mask) * ntiles + blockIdx.x] : 0u;
        }
    }
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
        uint32_t idx = base + k * 64 + lane;
        bool valid = idx < n;
        uint32_t d = (key[k] >> shift) & mask;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            uint32_t bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        if (!valid) peers = 0;
        uint32_t below = (uint32_t)__popcll(peers & lt_mask);
        uint32_t prev = valid ? mywc[d] : 0u;
        rank[k] = prev + below;
        if (valid && below == 0) mywc[d] = prev + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < nb; d += kRadixThreads) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < (uint32_t)kRadixWaves; ++w) {
            uint32_t t = wc[w * nb + d];
            wc[w * nb + d] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
        uint32_t idx = base + k * 64 + lane;
        if (idx < n) {
            uint32_t d = (key[k] >> shift) & mask;
            uint32_t dest = (kLocalScan ? loffs[d] : base_off[k]) + mywc[d] + rank[k];
            keys_out[dest] = key[k];
            vals_out[dest] = val[k];
        }
    }
}

static void radix_plan(int total_bits, int& passes, int& bits_per) {
    if (total_bits < 1) total_bits = 1;
    passes = (total_bits + kRadixMaxBits - 1) / kRadixMaxBits;
    bits_per = (total_bits + passes - 1) / passes;
}

size_t radix_hist_words(uint32_t n, int total_bits) {
    int passes, bits_per;
    radix_plan(total_bits, passes, bits_per);
    size_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    if (ntiles == 0) ntiles = 1;
    return ((size_t)1 << bits_per) * ntiles;
}

// The first pass computes the keys (cell_key of each position) in its upsweep and writes them
// to `keys`.
hipError_t radix_sort_keys(const float4* pm, const StepConsts& c, const SceneGrid* scenes,
                           uint32_t* err_flag, uint32_t* keys, uint32_t* keys_alt, uint32_t* vals,
                           uint32_t* vals_alt, uint32_t* hist, uint32_t* lookback,
                           uint32_t* scan_scratch, uint32_t n, int total_bits,
                           uint32_t** keys_out, uint32_t** vals_out, hipStream_t st) {
    int passes, bits_per;
    radix_plan(total_bits, passes, bits_per);
    if (total_bits < 1) total_bits = 1;
    uint32_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    uint32_t* src_k = keys;
    uint32_t* dst_k = keys_alt;
    const uint32_t* src_v = nullptr;
    uint32_t* dst_v = vals;
    for (int p = 0; p < passes; ++p) {
        int shift = p * bits_per;
        int b = total_bits - shift < bits_per ? total_bits - shift : bits_per;
        uint32_t mask = (1u << b) - 1u;
        uint32_t m = (mask + 1) * ntiles;
        if (p == 0)
            hipLaunchKernelGGL(k_radix_upsweep<true>, dim3(ntiles), dim3(kRadixThreads), 0, st,
                               (const uint32_t*)nullptr, n, shift, mask, ntiles, hist, pm, c,
                               scenes, src_k, err_flag);
        else
            hipLaunchKernelGGL(k_radix_upsweep<false>, dim3(ntiles), dim3(kRadixThreads), 0, st,
                               src_k, n, shift, mask, ntiles, hist, pm, c, scenes,
                               (uint32_t*)nullptr, err_flag);
        if (m <= kRadixLocalScanMax) {
            hipLaunchKernelGGL(k_radix_downsweep<true>, dim3(ntiles), dim3(kRadixThreads), 0, st,
                               src_k, src_v, dst_k, dst_v, n, shift, mask, b, ntiles, hist);
        } else {
            // one launch while few tiles poll the status words; beyond that their polling
            // costs more than the three launches (9.6M particles, 587 tiles: +50 us per pass)
            hipError_t se = m <= kLookbackMaxTiles * kScanTile
                                ? scan_exclusive_lookback(hist, m, lookback, err_flag, st)
                                : scan_exclusive(hist, m, scan_scratch, st);
            if (se != hipSuccess) return se;
            hipLaunchKernelGGL(k_radix_downsweep<false>, dim3(ntiles), dim3(kRadixThreads), 0, st,
                               src_k, src_v, dst_k, dst_v, n, shift, mask, b, ntiles, hist);
        }
        uint32_t* nk = (dst_k == keys_alt) ? keys : keys_alt;
        uint32_t* nv = (dst_v == vals) ? vals_alt : vals;
        src_k = dst_k;
        src_v = dst_v;
        dst_k = nk;
        dst_v = nv;
    }
    *keys_out = src_k;
    *vals_out = const_cast<uint32_t*>(src_v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// gather into cell-sorted order
// ---------------------------------------------------------------------------
// Quantized coordinate of the candidate filter: floor(x * qscale) modulo 2^32, computed in
// FP64 (exact for |x * qscale| < 2^52; 0 for non-finite or larger values, whose particles
// are never candidates).
__device__ __forceinline__ uint32_t quantize(float x, double qscale) {
    const double v = floor((double)x * qscale);
    if (!(fabs(v) < 4503599627370496.0)) return 0u;
    return (uint32_t)(v - 4294967296.0 * floor(v * (1.0 / 4294967296.0)));
}

// One launch, two independent parts that need only the sorted keys and ids: workgroups below
// `gather_blocks` gather particle s into sorted order, the rest fill the cell table
// cs[c] = first sorted slot with key >= c, for c in [0, G + 1] (binary search per cell).
__global__ __launch_bounds__(256) void k_gather_cells(const float4* __restrict__ pm,
                                                      const float4* __restrict__ vel,
                                                      const uint32_t* __restrict__ sid,
                                                      const uint32_t* __restrict__ skey,
                                                      uint32_t n, double qscale,
                                                      uint2* __restrict__ sq,
                                                      float4* __restrict__ spm,
                                                      float4* __restrict__ svr,
                                                      uint32_t gather_blocks, uint32_t ncs,
                                                      uint32_t* __restrict__ cs) {
    if (blockIdx.x >= gather_blocks) {
        const uint32_t cidx = (blockIdx.x - gather_blocks) * blockDim.x + threadIdx.x;
        if (cidx >= ncs) return;
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (skey[mid] < cidx)
                lo = mid + 1;
            else
                hi = mid;
        }
        cs[cidx] = lo;
        return;
    }
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t id = sid[s];
    const float4 p = pm[id];
    spm[s] = p;
    const float4 v = vel[id];
    svr[s] = make_float4(v.x, v.y, v.z, 0.0f);
    // x and y keep their low 16 bits (packed), z its low 32: differences of neighbours are
    // exact in 16- and 24-bit wrap-around arithmetic
    const uint32_t qx = quantize(p.x, qscale), qy = quantize(p.y, qscale);
    sq[s] = make_uint2((qx & 0xffffu) | (qy << 16), quantize(p.z, qscale));
}

hipError_t launch_gather_cells(const float4* pm, const float4* vel, const uint32_t* sid,
                               const uint32_t* skey, uint32_t n, uint32_t G, double qscale,
                               uint2* sq, float4* spm, float4* svr, uint32_t* cell_start,
                               hipStream_t st) {
    const uint32_t gather_blocks = (n + 255) / 256;
    const uint32_t ncs = G + 2;
    const uint32_t cell_blocks = (ncs + 255) / 256;
    hipLaunchKernelGGL(k_gather_cells, dim3(gather_blocks + cell_blocks), dim3(256), 0, st, pm,
                       vel, sid, skey, n, qscale, sq, spm, svr, gather_blocks, ncs, cell_start);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// CFL timestep (compute_cfl_timestep, src/sph_engine.cpp:312-333):
//   dt = min(CFL h / (max|v| + 1e-6), CFL sqrt(h / (|a_0| + 1e-6)), timestep)
// max|v| ignores NaN like std::max(running, v).  scratch[0] holds max|v| bits and
// is left at zero for the next call.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float len3_exact(float x, float y, float z) {
#pragma clang fp contract(off)
    float tx = x * x, ty = y * y, tz = z * z;
    return __fsqrt_rn((tx + ty) + tz);
}

__global__ __launch_bounds__(256) void k_cfl_reduce(const float4* __restrict__ vel, uint32_t n,
                                                    uint32_t* __restrict__ scratch) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    for (; i < n; i += gridDim.x * blockDim.x) {
        float4 v = vel[i];
        float l = len3_exact(v.x, v.y, v.z);
        if (m < l) m = l;
    }
    for (int off = 32; off > 0; off >>= 1) {
        float o = __shfl_xor(m, off);
        if (m < o) m = o;
    }
    // one atomic per workgroup: every workgroup updates the same word
    __shared__ float red[4];
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < (blockDim.x >> 6); ++w)
            if (m < red[w]) m = red[w];
        if (m > 0.0f) atomicMax(scratch, __float_as_uint(m));
    }
}

__global__ void k_cfl_finalize(const float4* __restrict__ acc, uint32_t n, float cfl, float h,
                               float timestep, uint32_t* __restrict__ scratch,
                               float* __restrict__ dt_dev) {
    float maxv = __uint_as_float(scratch[0]);
    float4 a0 = n > 0 ? acc[0] : make_float4(0.f, 0.f, 0.f, 0.f);
    float dt_cfl = __fdiv_rn(cfl * h, maxv + 1e-6f);
    float dt_force = cfl * __fsqrt_rn(__fdiv_rn(h, len3_exact(a0.x, a0.y, a0.z) + 1e-6f));
    float m = dt_cfl;
    if (dt_force < m) m = dt_force;
    if (timestep < m) m = timestep;
    *dt_dev = m;
    scratch[0] = 0u;
}

hipError_t launch_cfl(const float4* vel, const float4* acc, uint32_t n, const sphk_params& p,
                      uint32_t* scratch_u32, float* dt_dev, hipStream_t st) {
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 256) blocks = 256;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_cfl_reduce, dim3(blocks), dim3(256), 0, st, vel, n, scratch_u32);
    hipLaunchKernelGGL(k_cfl_finalize, dim3(1), dim3(1), 0, st, acc, n, p.CFL_factor,
                       p.smoothing_length, p.timestep, scratch_u32, dt_dev);
    return hipGetLastError();
}

// Copies three float4 arrays of n rows and one scalar in one launch (state snapshot and
// restore: positions, velocities, accelerations and the clock).
__global__ __launch_bounds__(256) void k_copy_state(const float4* __restrict__ a0,
                                                    const float4* __restrict__ a1,
                                                    const float4* __restrict__ a2,
                                                    float4* __restrict__ b0,
                                                    float4* __restrict__ b1,
                                                    float4* __restrict__ b2, size_t n,
                                                    const float* __restrict__ t_src,
                                                    float* __restrict__ t_dst) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float4 x0 = a0[i], x1 = a1[i], x2 = a2[i];
        b0[i] = x0;
        b1[i] = x1;
        b2[i] = x2;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *t_dst = *t_src;
}

hipError_t launch_copy_state(const float4* a0, const float4* a1, const float4* a2, float4* b0,
                             float4* b1, float4* b2, size_t n, const float* t_src,
                             float* t_dst, hipStream_t st) {
    const size_t blocks = std::min<size_t>(std::max<size_t>((n + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(k_copy_state, dim3((uint32_t)blocks), dim3(256), 0, st, a0, a1, a2, b0,
                       b1, b2, n, t_src, t_dst);
    return hipGetLastError();
}

__global__ void k_advance_time(float* __restrict__ t, const float* __restrict__ dt_dev,
                               float dt, int reps) {
    float d = dt_dev ? *dt_dev : dt;
    float x = *t;
    for (int k = 0; k < reps; ++k) x = x + d;
    *t = x;
}

hipError_t launch_advance_time(float* time_dev, const float* dt_dev, float dt, hipStream_t st,
                               int reps) {
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, st, time_dev, dt_dev, dt, reps);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint32_t v,
                                                  size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Renderer instance data: 9 floats per particle, caller order -- position (3), velocity (3),
// colour (3) -- the layout Renderer::update_particle_data builds on the host every frame
// (src/renderer.cpp:279-311).  One thread per output float4: reads are gathers from the home
// rows (cached), writes are coalesced 16-byte stores.
// ---------------------------------------------------------------------------
template <bool kVec4>
__global__ __launch_bounds__(256) void k_render_pack(const float4* __restrict__ pm,
                                                     const float4* __restrict__ vel,
                                                     const float* __restrict__ rgb, uint32_t n,
                                                     float* __restrict__ out) {
    const size_t total = (size_t)9 * n;
    const size_t k0 = 4 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (k0 >= total) return;
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const size_t k = k0 + (size_t)c;
        v[c] = 0.0f;
        if (k >= total) continue;
        const size_t i = k / 9;
        const uint32_t f = (uint32_t)(k - 9 * i);
        if (f < 3) {
            const float4 p = pm[i];
            v[c] = f == 0 ? p.x : f == 1 ? p.y : p.z;
        } else if (f < 6) {
            const float4 q = vel[i];
            v[c] = f == 3 ? q.x : f == 4 ? q.y : q.z;
        } else {
            // Particle::color's default (src/particle.h:42) when no colours are given
            const float def = f == 6 ? 0.0f : f == 7 ? 0.5f : 1.0f;
            v[c] = rgb ? rgb[3 * i + (f - 6)] : def;
        }
    }
    if (kVec4 && k0 + 4 <= total) {
        *reinterpret_cast<float4*>(out + k0) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (k0 + (size_t)c < total) out[k0 + (size_t)c] = v[c];
    }
}

hipError_t launch_render_pack(const float4* pm, const float4* vel, const float* rgb, uint32_t n,
                              float* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const size_t quads = ((size_t)9 * n + 3) / 4;
    const dim3 grid((unsigned)((quads + 255) / 256));
    if (reinterpret_cast<uintptr_t>(out) % 16 == 0)
        hipLaunchKernelGGL(k_render_pack<true>, grid, dim3(256), 0, st, pm, vel, rgb, n, out);
    else
        hipLaunchKernelGGL(k_render_pack<false>, grid, dim3(256), 0, st, pm, vel, rgb, n, out);
    return hipGetLastError();
}

}  // namespace sphk
