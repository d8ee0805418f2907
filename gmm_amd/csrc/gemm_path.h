// The dense-contraction (GEMM) form of the E-step shared by the
// full-covariance path (fullcov.hip), the Gaussian-selection and
// posterior-export paths and the robust auxiliary statistics (auxstats.hip):
// frames to extended rows [1, x', 0..], component log-likelihoods
// L = Theta^T phi with the feature map phi formed from those rows on the fly
// (gemm_mfma.hip), a fused per-frame log-sum-exp / posterior pass, the
// statistics Gamma^T phi the same way, and FP64 fixed-order sums of FP32
// partials.
#pragma once

#include "ctx.h"

namespace gmmhip {

// What theta currently holds: a full model, a diagonal model over the full
// feature map (full second-order statistics of a diagonal model), or a
// diagonal model over [1, x', x'^2].
enum GemmOperand { OP_FULL = 0, OP_DIAG_FULLFEAT = 1, OP_DIAG = 2 };

struct FullState {
    int K = 0, KP = 0;           // features, padded to a multiple of 32
    float *d_theta = nullptr;    // [M][KP]
    size_t theta_cap = 0;
    uint32_t *d_pairs = nullptr; // [KP] feature pairs (see feature_pairs)
    size_t pairs_cap = 0;
    float *d_cen = nullptr;      // [D]
    size_t cen_cap = 0;
    float *d_xt = nullptr;       // [ncp][ext_row_floats(D)] extended rows of a chunk
    size_t xt_cap = 0;
    float *d_L = nullptr;        // [nc][M]: log-likelihoods, then posteriors
    size_t L_cap = 0;
    uint8_t *d_bad = nullptr;    // [nc]
    size_t bad_cap = 0;
    float *d_part = nullptr;     // [S][M][KP] FP32 statistics of one chunk's sub-chunks
    size_t part_cap = 0;
    double *d_acc = nullptr;     // [M][KP] FP64 statistics over chunks
    size_t acc_cap = 0;
    std::vector<float> h_cen;
    uint64_t operand_version = ~0ull;  // model_version theta was built for
    int operand_kind = OP_FULL;
};

}  // namespace gmmhip

namespace gmmhip_detail {

// Feature codes of phi: -1 -> 1; -2 - d -> x'_d; (i << 16) | j -> x'_i x'_j.
// full: [1, x', x'_i x'_j (i <= j, packed upper)]; diagonal: [1, x', x'_d^2].
std::vector<int> feature_codes(int D, bool full);

// The handle's GEMM-path state (created on first use).
int gemm_state(gmmhip_ctx *h, gmmhip::FullState **out);

// The GEMM operand theta of the handle's diagonal model, over the full
// feature map (zero off-diagonal precisions: full second-order statistics of
// a diagonal model) or over [1, x', x'^2] (fullcov.hip).
int diag_operand(gmmhip_ctx *h, bool full_features);

// Features padded to a multiple of 32 (the staging step of fc_ll_kernel).
inline int feature_pad(int K) { return (K + 31) / 32 * 32; }
// Floats per extended row [1, x'_1..x'_D, 0, ..]: at least one zero slot,
// a multiple of 4.
int ext_row_floats(int D);
// The (a, b) pairs of the feature codes, padded to KP with the zero slot:
// phi_k = xt[a_k] * xt[b_k].
std::vector<uint32_t> feature_pairs(const std::vector<int> &codes, int D, int KP);
hipError_t launch_fc_rows(const float *X, uint64_t n0, int nc, int ncp, int D, const float *cen,
                          float *xt, uint8_t *bad, hipStream_t s);
// L[t][m] (t < nc) = phi(xt_t) . theta_m, KP a multiple of 32.
hipError_t launch_fc_ll(const float *xt, int nc, int D, const float *theta, int M, int KP,
                        const uint32_t *pairs, float *L, hipStream_t s);
// part[s][m][k] = sum over the sub frames of sub-chunk s of g[t][m] * phi_k(t),
// rows of stride DXP (phi_k = xt[a_k] * xt[b_k]).
hipError_t launch_fc_stats(const float *xt, int DXP, const float *g, int M, int KP,
                           const uint32_t *pairs, int sub, int S, float *part, hipStream_t s);
// Sub-chunks S of the statistics of an M x KP output for chunks of at most
// nc_max frames: about 16 blocks (128 x 128 tile, sub-chunk) per CU, at
// least 256 frames per sub-chunk, at most 512 MB of FP32 partials.
int stats_subchunks(int M, int KP, int num_cus, uint64_t nc_max);
// Frames per sub-chunk of a chunk of nc frames split S ways: a multiple of
// the statistics kernel's 32-frame stage (sub * S >= nc; the padding frames
// carry zero posteriors and zero rows).
inline int stats_sub(uint64_t nc, int S) { return (int)(((nc + S - 1) / S + 31) / 32 * 32); }
// cut = 1: posteriors g > min_gamma kept, others 0 (statistics); cut = 0:
// posteriors before the cut.  Clamped / non-finite frames get 0 either way.
hipError_t launch_fc_lse(float *L, int nc, int ncp, int M, const uint8_t *bad, float min_ll,
                         float min_gamma, int cut, float *LLnat, uint64_t n0, double *blockLL,
                         hipStream_t s);
hipError_t launch_fc_acc(const float *part, size_t E, int S, int first, double *acc,
                         hipStream_t s);

}  // namespace gmmhip_detail
