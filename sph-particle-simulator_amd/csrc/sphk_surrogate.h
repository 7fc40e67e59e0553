// sphk_surrogate.h -- device layout and launchers of the learned macro-step surrogate.
//
// Network (widths fixed at 128):
//   edge features  e_ij = [r_ij/R (3), (v_i - v_j)/v_s (3), |r_ij|/R, W(|r_ij|)/W(0),
//                          (rho_j - rho0) k_rho]
//   message        S_i = c * sum_{j != i, |r_ij| <= R} ReLU(W1 e_ij + b1)
//                  m_i = W2 S_i + c n_i b2          (second layer linear, applied once)
//   node           x_i = [m_i, (rho_i - rho0) k_rho, v_i / v_s]
//                  z_i = ReLU(Wn1 x_i + bn1),  out_i = Wn2 z_i + bn2,  a_i = a_s out_i
// Flat FP32 parameters (Xavier-initialised, Adam-trained):
//   W1[128][9] b1[128] W2[128][128] b2[128] Wn1[128][132] bn1[128] Wn2[3][128] bn2[3]
//
// Dense layers run on v_mfma_f32_32x32x16_bf16 (bf16 operands, FP32 accumulation).  Its
// lane layout, for lane l (h = l / 32): A[l % 32][8h + i], B[8h + i][l % 32] (i < 8), and
// accumulator register v holds D[(v / 4) * 8 + 4h + v % 4][l % 32].
//
// Activation rows (one per particle; inference: sorted slot, training: 64 * batch chunk +
// lane), bf16, row stride kRow = 144:
//   S : [S_i (128) | c n_i | 0 ...]        the extra column makes db2 a column of dW2
//   X : [m_i (128) | rho, v (3) | 1 | 0...] the 1 makes dbn1 a column of dWn1
//   Z : [z_i (128) | 1 | 0 ...]            the 1 makes dbn2 a column of dWn2
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sphk_internal.h"

namespace sphk {

constexpr int kSurH = 128;           // message hidden, message output and node hidden width
constexpr int kSurNF = 9;            // edge features
constexpr int kSurNodeIn = 132;      // 128 + rho + v(3)
constexpr int kSurRow = 144;         // row stride of S, X, Z (multiple of 16)
constexpr int kSurOutRow = 32;       // row stride of the output-gradient rows (3 used)
constexpr uint32_t kSurNoSlot = 0xffffffffu;

// flat parameter offsets
constexpr uint32_t kOffW1 = 0;
constexpr uint32_t kOffB1 = kOffW1 + 128 * kSurNF;
constexpr uint32_t kOffW2 = kOffB1 + 128;
constexpr uint32_t kOffB2 = kOffW2 + 128 * 128;
constexpr uint32_t kOffWn1 = kOffB2 + 128;
constexpr uint32_t kOffBn1 = kOffWn1 + 128 * kSurNodeIn;
constexpr uint32_t kOffWn2 = kOffBn1 + 128;
constexpr uint32_t kOffBn2 = kOffWn2 + 3 * 128;
constexpr uint32_t kSurParams = kOffBn2 + 3;  // 35203

// packed bf16 weights (element offsets; every region 16-byte aligned): operand images
constexpr uint32_t kPkW1 = 0;                       // [128][8]  W1 columns 0..7
constexpr uint32_t kPkB1 = kPkW1 + 128 * 8;         // [128]     b1
constexpr uint32_t kPkW1r = kPkB1 + 128;            // [128]     W1 column 8 (neighbour rho)
constexpr uint32_t kPkW2 = kPkW1r + 128;            // [128 out][128 in]
constexpr uint32_t kPkW2T = kPkW2 + 128 * 128;      // [128 in][128 out]
constexpr uint32_t kPkWn1 = kPkW2T + 128 * 128;     // [128 out][144 in] (cols >= 132 zero)
constexpr uint32_t kPkWn1T = kPkWn1 + 128 * kSurRow;  // [144 in][128 out] (rows >= 132 zero)
constexpr uint32_t kPkWn2 = kPkWn1T + kSurRow * 128;  // [32 out][128 in] (rows >= 3 zero)
constexpr uint32_t kPkWn2T = kPkWn2 + 32 * 128;     // [128 in][32 out] (cols >= 3 zero)
constexpr uint32_t kPkTotal = kPkWn2T + 128 * 32;
static_assert(kPkW2 % 8 == 0 && kPkWn1 % 8 == 0 && kPkWn2T % 8 == 0, "16-byte regions");
// bf16-rounded copies as FP32 (element offsets)
constexpr uint32_t kFW1 = 0;                 // [128][9]
constexpr uint32_t kFB1 = kFW1 + 128 * kSurNF;
constexpr uint32_t kFB2 = kFB1 + 128;
constexpr uint32_t kFBn1 = kFB2 + 128;
constexpr uint32_t kFBn2 = kFBn1 + 128;
constexpr uint32_t kFTotal = kFBn2 + 4;

// Feature and output scales (see the network above).
struct SurScales {
    float inv_r, inv_v, inv_h, w_norm;
    float rho0, rho_scale, msg_scale, acc_scale;
};

// Epilogues of the row GEMM C = A * Bt^T.
enum SurEpi : int {
    kEpiM = 0,    // X[:, :128] = bf16(C + cnt[row] * b2)
    kEpiZ = 1,    // Z[:, :128] = bf16(ReLU(C + bn1))
    kEpiOut = 2,  // OUT[row][0..2] = C + bn2           (FP32, 4 floats per row)
    kEpiDZ = 3,   // DZ = Z > 0 ? bf16(C) : 0             (row stride 128)
    kEpiDM = 4,   // DM = bf16(C)                         (row stride 128)
    kEpiDS = 5,   // DS = bf16(c * C)                     (row stride 128)
};

struct SurEpiArgs {
    void* out;          // destination
    const float* bias;  // FP32 bias (bf16-rounded values)
    const float* cnt;   // c * n_i per row (kEpiM)
    const __bf16* z;    // Z rows (kEpiDZ)
    float scale;        // kEpiDS
};

// Packs FP32 parameters into the bf16 operand images and the FP32 rounded copies.
hipError_t launch_sur_pack(const float* params, __bf16* pk, float* pf, hipStream_t st);

// Activation rows written by the forward pass for the backward pass.
struct SurRows {
    __bf16* S;
    __bf16* X;
    __bf16* Z;
    float4* OUT;         // network output (normalised), 3 used
    float* cnt;          // c n_i
    uint32_t* rowslot;   // sorted slot of the row, kSurNoSlot for padding
};

// Macro-step applied by the forward pass (integrate mode).
struct SurStep {
    float T;    // macro-step length
    float dk;   // damping^substeps
    float4* pm;
    float4* vel;
    float4* acc;
    float* rho;
    float* prs;
};

// Forward pass over the stored neighbour words: edge walk (message layer 1 on MFMA) and node
// MLP.  integrate: applies the macro-step to the caller-order state (every chunk; the
// non-finite bin falls freely).  Otherwise writes the activation rows: batch == nullptr:
// every chunk, row = sorted slot (active lanes only); else chunk items batch[b], rows
// 64 b + lane (all 64 written).
hipError_t launch_sur_forward(bool integrate, const WalkBuffers& wb, const uint32_t* batch,
                              uint32_t nbatch, uint32_t max_items, const float4* spm,
                              const float4* svr, const uint32_t* skey, const uint32_t* sid,
                              const uint32_t* cs, uint32_t n, const StepConsts& c,
                              const SceneGrid* scenes, const __bf16* pk, const float* pf,
                              const SurScales& sc, const SurRows& rows, const SurStep& step,
                              hipStream_t st);

// Row GEMM: C[R x N] = A[R x K] * Bt[N x K]^T (bf16, FP32 accumulate), R multiple of 32,
// K multiple of 16, N <= 128 multiple of 32.
hipError_t launch_sur_gemm(int epi, const __bf16* A, int lda, const __bf16* Bt, int ldb, int K,
                           int N, uint32_t R, const SurEpiArgs& ep, hipStream_t st);

// Weight gradient partials: part[slab][m][n] = sum over the slab's rows r of
// P[r][m] * Q[r][n], m < M, n < N (tiles of 32).  rows_per_slab multiple of 16.
hipError_t launch_sur_wgrad(const __bf16* P, int ldp, int M, const __bf16* Q, int ldq, int N,
                            uint32_t R, uint32_t rows_per_slab, uint32_t slabs, float* part,
                            hipStream_t st);

// Edge backward over a batch: partials part[2 b + half][128][16] of d(W1|b1): columns 0..8
// are W1's, column 9 is b1.
hipError_t launch_sur_edges_bwd(const WalkBuffers& wb, const uint32_t* batch, uint32_t nbatch,
                                const float4* spm, const float4* svr, const uint32_t* skey,
                                const uint32_t* cs, const StepConsts& c,
                                const SceneGrid* scenes, const __bf16* pk, const SurScales& sc,
                                const __bf16* DS, float* part, hipStream_t st);

// Sums `slabs` partial matrices [M_all][ld] into grad: rows m < M, columns n < ncols go to
// grad[dst + m * ncols + n]; column bias_col (if >= 0) goes to grad[bias_dst + m].
// Every value is multiplied by `scale`.
hipError_t launch_sur_reduce(const float* part, uint32_t slabs, int M_all, int ld, int M,
                             int ncols, int bias_col, uint32_t dst, uint32_t bias_dst,
                             float scale, float* grad, hipStream_t st);

// Loss rows: DOUT = 2 (OUT - target) for rows with a slot (zero otherwise); per block the
// partial sum of squared errors and the number of rows with a slot, interleaved.  The
// 1/(3B) of the mean is applied when the weight gradients are reduced.
hipError_t launch_sur_loss(const float* out, const uint32_t* rowslot, const uint32_t* sid,
                           const float4* target, uint32_t R, __bf16* dout, float* loss_part,
                           uint32_t* nparts, hipStream_t st);

// Targets (normalised): t[id] = (v4 - d^k v0) / (T a_s).
hipError_t launch_sur_targets(const float4* v0, const float4* v4, uint32_t n, float dk,
                              float inv_Tas, float4* target, hipStream_t st);

hipError_t launch_sur_adam(float* w, float* m, float* v, const float* g, uint32_t np, float lr,
                           float beta1, float beta2, float eps, float bc1, float bc2,
                           hipStream_t st);

// Density comparison: per-block partial sums of |a - b| and max |a - b| over n values.
hipError_t launch_sur_rho_diff(const float* a, const float* b, uint32_t n, float* part,
                               uint32_t* nparts, hipStream_t st);

// Densities of the sorted state in caller order.
hipError_t launch_sur_rho_home(const float4* spm, const float4* svr, const uint32_t* sid,
                               const uint32_t* cs, uint32_t n, const StepConsts& c, float* out,
                               hipStream_t st);

// Network accelerations in caller order.
hipError_t launch_sur_scatter_out(const float* out, const uint32_t* sid, const uint32_t* cs,
                                  uint32_t n, const StepConsts& c, float acc_scale, float4* dst,
                                  hipStream_t st);

}  // namespace sphk
