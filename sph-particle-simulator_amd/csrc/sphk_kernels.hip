// sphk_kernels.hip -- gfx950 kernels of the SPH step.
//
// Step pipeline (one HIP stream):
//   cell_keys      dense cell index of every particle, caller order
//   radix sort     stable LSD sort of (cell, caller id) -> cell-sorted order
//   cell_start     occupancy prefix of the dense grid (binary search per cell)
//   gather         cell-sorted float4 copies of (x,y,z,m) and (vx,vy,vz,-)
//   density        27-cell walk, rho_i = m_i W(0) + sum_j m_j W(r_ij)
//   force          27-cell walk, pressure + viscosity + gravity, then leapfrog and
//                  wall clamp, results scattered back to caller order
//
// Neighbour semantics follow SpatialHash::query_squared (src/spatial_hash.cpp:31-57):
// j is a neighbour of i when (dx*dx + dy*dy) + dz*dz <= r^2 in FP32 with each
// operation rounded once.  The reference scans 5x5x5 cells (cell_radius = 2 for
// cell == r); a pair whose cells differ by 2 on an axis can only pass through
// rounding of p*(1/cell), so the walk scans the 27 cells around i and widens to
// the +-2 layer on an axis only for particles where `edge_flags` proves such a
// pair possible.  Cells are visited in the reference's lexicographic (dx,dy,dz)
// order and particles within a cell in ascending caller id, so neighbour lists
// come out in the reference's order.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

#include "sphk_internal.h"

namespace sphk {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
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

// Can a particle at coordinate x (cell c = floor(fl(x*inv))) have a neighbour in
// cell c+2 (bit 1) or c-2 (bit 0)?  A neighbour needs fl(fl(b-x)^2) <= r2 on this
// axis alone, and fl(b-x) grows with b, so it suffices to test the nearest float b
// that lands in cell c+2 (resp. c-2).  Exact; conservative if the bracket search
// does not settle.
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
                if (bd * inv >= t)
                    b = bd;
                else {
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
                if (bu * inv < t)
                    b = bu;
                else {
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

struct WalkBox {
    int rx, ry, rz;  // cell relative to the grid origin
    int x0, x1, y0, y1, z0, z1;
};

__device__ __forceinline__ WalkBox walk_box(float4 p, uint32_t L, const StepConsts& c) {
    WalkBox w;
    w.rz = (int)(L % (uint32_t)c.nz);
    uint32_t t = L / (uint32_t)c.nz;
    w.ry = (int)(t % (uint32_t)c.ny);
    w.rx = (int)(t / (uint32_t)c.ny);
    int fx = axis_edge_flags(p.x, c.inv_cell, c.r2);
    int fy = axis_edge_flags(p.y, c.inv_cell, c.r2);
    int fz = axis_edge_flags(p.z, c.inv_cell, c.r2);
    w.x0 = -1 - (fx & 1);
    w.x1 = 1 + (fx >> 1);
    w.y0 = -1 - (fy & 1);
    w.y1 = 1 + (fy >> 1);
    w.z0 = -1 - (fz & 1);
    w.z1 = 1 + (fz >> 1);
    return w;
}

// Visits every neighbour j (including i itself) of the particle in slot s, in
// reference order.  fn(j, pj, d2).
template <class Fn>
__device__ __forceinline__ void walk_neighbors(const float4* __restrict__ spm,
                                               const uint32_t* __restrict__ cs, float4 pi,
                                               const WalkBox& w, const StepConsts& c, Fn&& fn) {
    for (int dx = w.x0; dx <= w.x1; ++dx) {
        for (int dy = w.y0; dy <= w.y1; ++dy) {
            uint32_t base = ((uint32_t)(w.rx + dx) * (uint32_t)c.ny + (uint32_t)(w.ry + dy)) *
                                (uint32_t)c.nz + (uint32_t)w.rz;
            uint32_t a = cs[base + w.z0];
            uint32_t b = cs[base + w.z1 + 1];
            for (uint32_t j = a; j < b; ++j) {
                float4 pj = spm[j];
                float d2 = dist2_exact(pi.x, pi.y, pi.z, pj.x, pj.y, pj.z);
                if (d2 <= c.r2) fn(j, pj, d2);
            }
        }
    }
}

// Cubic-spline shape without sigma: 2/3 - q^2 + q^3/2 (q <= 1), (2-q)^3/6 (1 < q <= 2)
// (CubicSplineKernel::W_3d, src/kernels.cpp:58-66).
__device__ __forceinline__ float w_shape(float q) {
    float q2 = q * q;
    float win = (2.0f / 3.0f - q2) + 0.5f * q2 * q;
    float t = 2.0f - q;
    float wout = (1.0f / 6.0f) * t * t * t;
    return q <= 1.0f ? win : (q <= 2.0f ? wout : 0.0f);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// ---------------------------------------------------------------------------
// K1: dense cell index per particle (caller order)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cell_keys(const float4* __restrict__ pm, uint32_t n,
                                                   StepConsts c, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ err) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pm[i];
    uint32_t L = c.G;
    if (__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z)) {
        float fx = floorf(p.x * c.inv_cell);
        float fy = floorf(p.y * c.inv_cell);
        float fz = floorf(p.z * c.inv_cell);
        bool inr = fx >= -1073741824.0f && fx < 1073741824.0f && fy >= -1073741824.0f &&
                   fy < 1073741824.0f && fz >= -1073741824.0f && fz < 1073741824.0f;
        int rx = inr ? (int)fx - c.ox : -1;
        int ry = inr ? (int)fy - c.oy : -1;
        int rz = inr ? (int)fz - c.oz : -1;
        if (rx >= 2 && rx < c.nx - 2 && ry >= 2 && ry < c.ny - 2 && rz >= 2 && rz < c.nz - 2) {
            L = ((uint32_t)rx * (uint32_t)c.ny + (uint32_t)ry) * (uint32_t)c.nz + (uint32_t)rz;
        } else {
            atomicOr(err, 1u);
        }
    }
    keys[i] = L;
}

hipError_t launch_cell_keys(const float4* pm, uint32_t n, const StepConsts& c, uint32_t* keys,
                            uint32_t* err_flag, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cell_keys, dim3((n + 255) / 256), dim3(256), 0, st, pm, n, c, keys,
                       err_flag);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of 32-bit keys with 32-bit values.  Each pass:
//   upsweep   per-tile digit histogram (LDS atomics)   -> hist[digit][tile]
//   scan      exclusive scan of hist (digit-major)     -> global base per (digit, tile)
//   downsweep per-wave stable ranks from ballot-matched peer groups, per-digit
//             prefix over the waves of the tile, scatter.
// Items of a tile are owned wave-contiguously (wave w owns 512 consecutive keys,
// 64 per iteration), so (wave, iteration, lane) order is input order and every
// pass is stable.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kRadixThreads) void k_radix_upsweep(const uint32_t* __restrict__ keys,
                                                                 uint32_t n, int shift,
                                                                 uint32_t mask, uint32_t ntiles,
                                                                 uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[1 << kRadixMaxBits];
    const uint32_t nb = mask + 1;
    for (uint32_t i = threadIdx.x; i < nb; i += kRadixThreads) h[i] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRadixTile;
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
        uint32_t idx = base + k * kRadixThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & mask], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += kRadixThreads) hist[i * ntiles + blockIdx.x] = h[i];
}

__global__ __launch_bounds__(1024) void k_scan_exclusive_1blk(uint32_t* __restrict__ data,
                                                              uint32_t m) {
    __shared__ uint32_t part[1024];
    const uint32_t per = (m + 1023) / 1024;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = min(b + per, m);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += data[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t i = b; i < e; ++i) {
        uint32_t t = data[i];
        data[i] = run;
        run += t;
    }
}

__global__ __launch_bounds__(kRadixThreads) void k_radix_downsweep(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, int shift,
    uint32_t mask, int bits, uint32_t ntiles, const uint32_t* __restrict__ offs) {
    __shared__ uint32_t wc[kRadixWaves << kRadixMaxBits];
    const uint32_t nb = mask + 1;
    for (uint32_t i = threadIdx.x; i < kRadixWaves * nb; i += kRadixThreads) wc[i] = 0;
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * kRadixTile + wave * (kRadixItems * 64);
    uint32_t* mywc = wc + wave * nb;
    uint32_t key[kRadixItems];
    uint32_t rank[kRadixItems];
#pragma unroll
    for (int k = 0; k < kRadixItems; ++k) {
        uint32_t idx = base + k * 64 + lane;
        bool valid = idx < n;
        key[k] = valid ? keys_in[idx] : 0u;
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
            uint32_t dest = offs[d * ntiles + blockIdx.x] + mywc[d] + rank[k];
            keys_out[dest] = key[k];
            vals_out[dest] = vals_in ? vals_in[idx] : idx;
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

hipError_t radix_sort_keys(uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                           uint32_t* hist, uint32_t n, int total_bits, uint32_t** keys_out,
                           uint32_t** vals_out, hipStream_t st) {
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
        hipLaunchKernelGGL(k_radix_upsweep, dim3(ntiles), dim3(kRadixThreads), 0, st, src_k, n,
                           shift, mask, ntiles, hist);
        hipLaunchKernelGGL(k_scan_exclusive_1blk, dim3(1), dim3(1024), 0, st, hist, m);
        hipLaunchKernelGGL(k_radix_downsweep, dim3(ntiles), dim3(kRadixThreads), 0, st, src_k,
                           src_v, dst_k, dst_v, n, shift, mask, b, ntiles, hist);
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
// cell_start[c] = first sorted slot with key >= c, for c in [0, G+1]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cell_start(const uint32_t* __restrict__ skey, uint32_t n,
                                                    uint32_t ncs, uint32_t* __restrict__ cs) {
    uint32_t cidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (cidx >= ncs) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (skey[mid] < cidx)
            lo = mid + 1;
        else
            hi = mid;
    }
    cs[cidx] = lo;
}

hipError_t launch_cell_start(const uint32_t* skey, uint32_t n, uint32_t G, uint32_t* cell_start,
                             hipStream_t st) {
    uint32_t ncs = G + 2;
    hipLaunchKernelGGL(k_cell_start, dim3((ncs + 255) / 256), dim3(256), 0, st, skey, n, ncs,
                       cell_start);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// gather into cell-sorted order
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather(const float4* __restrict__ pm,
                                                const float4* __restrict__ vel,
                                                const uint32_t* __restrict__ sid, uint32_t n,
                                                float4* __restrict__ spm, float4* __restrict__ svr) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t id = sid[s];
    spm[s] = pm[id];
    float4 v = vel[id];
    svr[s] = make_float4(v.x, v.y, v.z, 0.0f);
}

hipError_t launch_gather(const float4* pm, const float4* vel, const uint32_t* sid, uint32_t n,
                         float4* spm, float4* svr, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, st, pm, vel, sid, n, spm, svr);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// neighbour statistics: stats[0] = max list length (incl. self), stats[1] = sum
// ---------------------------------------------------------------------------
__device__ __forceinline__ void add_nbr_stats(uint32_t cnt, unsigned long long* stats) {
    uint32_t mx = cnt;
    unsigned long long sm = cnt;
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
        sm += __shfl_xor(sm, off);
    }
    if (lane_id() == 0) {
        atomicMax(&stats[0], (unsigned long long)mx);
        atomicAdd(&stats[1], sm);
    }
}

// ---------------------------------------------------------------------------
// density: rho_i = m_i W(0) + sum_{j != i} m_j W(r_ij)
// (compute_density, src/sph_engine.cpp:370-383)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_density(const float4* __restrict__ spm,
                                                 float4* __restrict__ svr,
                                                 const uint32_t* __restrict__ skey,
                                                 const uint32_t* __restrict__ cs, uint32_t n,
                                                 StepConsts c,
                                                 unsigned long long* __restrict__ stats) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cnt = 0;
    if (s < n) {
        float4 pi = spm[s];
        uint32_t L = skey[s];
        float rho = pi.w * c.w0;
        if (L < c.G) {
            WalkBox w = walk_box(pi, L, c);
            float acc = 0.0f;
            walk_neighbors(spm, cs, pi, w, c, [&](uint32_t j, float4 pj, float d2) {
                ++cnt;
                if (j != s) {
                    float q = __builtin_amdgcn_sqrtf(d2) * c.inv_h;
                    acc += pj.w * w_shape(q);
                }
            });
            rho = rho + c.sigma * acc;
        }
        reinterpret_cast<float*>(svr)[4 * s + 3] = rho;
    }
    if (stats) add_nbr_stats(cnt, stats);
}

hipError_t launch_density(const float4* spm, float4* svr, const uint32_t* skey,
                          const uint32_t* cell_start, uint32_t n, const StepConsts& c,
                          uint32_t* nbr_stats, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_density, dim3((n + 255) / 256), dim3(256), 0, st, spm, svr, skey,
                       cell_start, n, c, reinterpret_cast<unsigned long long*>(nbr_stats));
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// forces + leapfrog + wall clamp
//   F_p = -sum_{j != i, |r| >= 1e-6} m_j (P_i + P_j) / (2 rho_j) gradW(r_ij)
//   F_v =  sum_{j != i} (m_j / rho_j) mu (v_j - v_i) laplacianW(r_ij)
//   a_i = (F_p + F_v + g m_i) / m_i        (rho_i instead of m_i: consistent formulation)
// (src/sph_engine.cpp:390-443), then integrate_leapfrog (:290-310) and
// apply_boundary_conditions (src/particle.cpp:122-153).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void integrate_clamp(float4& p, float3& v, float3 a, float dt,
                                                const StepConsts& c) {
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
    if (p.x < c.lo[0]) {
        p.x = c.lo[0];
        v.x = v.x * wd;
    } else if (p.x > c.hi[0]) {
        p.x = c.hi[0];
        v.x = v.x * wd;
    }
    if (p.y < c.lo[1]) {
        p.y = c.lo[1];
        v.y = v.y * wd;
    } else if (p.y > c.hi[1]) {
        p.y = c.hi[1];
        v.y = v.y * wd;
    }
    if (p.z < c.lo[2]) {
        p.z = c.lo[2];
        v.z = v.z * wd;
    } else if (p.z > c.hi[2]) {
        p.z = c.hi[2];
        v.z = v.z * wd;
    }
}

__device__ __forceinline__ float3 final_accel(float3 fp, float3 fv, float m_i, float rho_i,
                                              const StepConsts& c) {
    float fx = fp.x * c.grad_scale + fv.x * c.lap_scale;
    float fy = fp.y * c.grad_scale + fv.y * c.lap_scale;
    float fz = fp.z * c.grad_scale + fv.z * c.lap_scale;
    float m = c.formulation == 1 ? rho_i : m_i;
    float inv_m = 1.0f / m;
    return make_float3(fx * inv_m, (fy + c.grav * m) * inv_m, fz * inv_m);
}

__global__ __launch_bounds__(256) void k_force_integrate(
    const float4* __restrict__ spm, const float4* __restrict__ svr,
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sid,
    const uint32_t* __restrict__ cs, uint32_t n, StepConsts c, const float* __restrict__ dt_dev,
    float dt_arg, float4* __restrict__ pm, float4* __restrict__ vel, float4* __restrict__ acc,
    float* __restrict__ rho_out, float* __restrict__ prs_out) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float4 pi = spm[s];
    float4 vi = svr[s];
    uint32_t L = skey[s];
    const float rho_i = vi.w;
    const float P_i = c.k_gas * (rho_i - c.rho0);
    float3 fp = make_float3(0.f, 0.f, 0.f);
    float3 fv = make_float3(0.f, 0.f, 0.f);
    if (L < c.G) {
        WalkBox w = walk_box(pi, L, c);
        walk_neighbors(spm, cs, pi, w, c, [&](uint32_t j, float4 pj, float d2) {
            if (j == s) return;
            float4 vj = svr[j];
            float rl = __builtin_amdgcn_sqrtf(d2);
            float q = rl * c.inv_h;
            float t = 2.0f - q;
            float inv_rho_j = __builtin_amdgcn_rcpf(vj.w);
            float dwdq = q <= 1.0f ? q * (-2.0f + 1.5f * q) : (q <= 2.0f ? -0.5f * t * t : 0.0f);
            float lapf = q <= 1.0f ? (-2.0f + 3.0f * q) : (q <= 2.0f ? t : 0.0f);
            float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            if (rl >= 1e-6f) {
                float P_j = c.k_gas * (vj.w - c.rho0);
                float coef = pj.w * (P_i + P_j) * 0.5f * inv_rho_j * dwdq * __builtin_amdgcn_rcpf(rl);
                fp.x -= coef * rx;
                fp.y -= coef * ry;
                fp.z -= coef * rz;
            }
            float cv = pj.w * inv_rho_j * lapf;
            fv.x += cv * (vj.x - vi.x);
            fv.y += cv * (vj.y - vi.y);
            fv.z += cv * (vj.z - vi.z);
        });
    }
    float3 a = final_accel(fp, fv, pi.w, rho_i, c);
    float dt = dt_dev ? *dt_dev : dt_arg;
    float3 v = make_float3(vi.x, vi.y, vi.z);
    float4 p = pi;
    integrate_clamp(p, v, a, dt, c);
    uint32_t id = sid[s];
    pm[id] = p;
    vel[id] = make_float4(v.x, v.y, v.z, 0.0f);
    acc[id] = make_float4(a.x, a.y, a.z, 0.0f);
    rho_out[id] = rho_i;
    prs_out[id] = P_i;
}

hipError_t launch_force_integrate(const float4* spm, const float4* svr, const uint32_t* skey,
                                  const uint32_t* sid, const uint32_t* cell_start, uint32_t n,
                                  const StepConsts& c, const float* dt_dev, float dt, float4* pm,
                                  float4* vel, float4* acc, float* rho, float* prs,
                                  hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_force_integrate, dim3((n + 255) / 256), dim3(256), 0, st, spm, svr, skey,
                       sid, cell_start, n, c, dt_dev, dt, pm, vel, acc, rho, prs);
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
    if (lane_id() == 0 && m > 0.0f) atomicMax(scratch, __float_as_uint(m));
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
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_cfl_reduce, dim3(blocks), dim3(256), 0, st, vel, n, scratch_u32);
    hipLaunchKernelGGL(k_cfl_finalize, dim3(1), dim3(1), 0, st, acc, n, p.CFL_factor,
                       p.smoothing_length, p.timestep, scratch_u32, dt_dev);
    return hipGetLastError();
}

__global__ void k_advance_time(float* __restrict__ t, const float* __restrict__ dt_dev,
                               float dt) {
    float d = dt_dev ? *dt_dev : dt;
    *t = *t + d;
}

hipError_t launch_advance_time(float* time_dev, const float* dt_dev, float dt, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, st, time_dev, dt_dev, dt);
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
// neighbour lists for parity checks (caller ids, reference order)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nbr_count(const float4* __restrict__ spm,
                                                   const uint32_t* __restrict__ skey,
                                                   const uint32_t* __restrict__ sid,
                                                   const uint32_t* __restrict__ cs, uint32_t n,
                                                   StepConsts c, uint32_t* __restrict__ cnt_by_id) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float4 pi = spm[s];
    uint32_t L = skey[s];
    uint32_t cnt = 0;
    if (L < c.G) {
        WalkBox w = walk_box(pi, L, c);
        walk_neighbors(spm, cs, pi, w, c, [&](uint32_t, float4, float) { ++cnt; });
    }
    cnt_by_id[sid[s]] = cnt;
}

__global__ __launch_bounds__(256) void k_nbr_write(const float4* __restrict__ spm,
                                                   const uint32_t* __restrict__ skey,
                                                   const uint32_t* __restrict__ sid,
                                                   const uint32_t* __restrict__ cs, uint32_t n,
                                                   StepConsts c, const int64_t* __restrict__ offs,
                                                   int64_t* __restrict__ idx) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float4 pi = spm[s];
    uint32_t L = skey[s];
    if (L >= c.G) return;
    int64_t o = offs[sid[s]];
    WalkBox w = walk_box(pi, L, c);
    walk_neighbors(spm, cs, pi, w, c, [&](uint32_t j, float4, float) { idx[o++] = sid[j]; });
}

hipError_t launch_nbr_count(const float4* spm, const uint32_t* skey, const uint32_t* sid,
                            const uint32_t* cell_start, uint32_t n, const StepConsts& c,
                            uint32_t* count_by_id, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nbr_count, dim3((n + 255) / 256), dim3(256), 0, st, spm, skey, sid,
                       cell_start, n, c, count_by_id);
    return hipGetLastError();
}

hipError_t launch_nbr_write(const float4* spm, const uint32_t* skey, const uint32_t* sid,
                            const uint32_t* cell_start, uint32_t n, const StepConsts& c,
                            const int64_t* offsets_by_id, int64_t* idx, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nbr_write, dim3((n + 255) / 256), dim3(256), 0, st, spm, skey, sid,
                       cell_start, n, c, offsets_by_id, idx);
    return hipGetLastError();
}

}  // namespace sphk
