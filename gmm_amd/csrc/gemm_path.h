// The dense-contraction (GEMM) form of the E-step shared by the
// full-covariance path (fullcov.hip) and the robust auxiliary statistics
// (auxstats.hip): frames expanded to a feature map phi, component
// log-likelihoods L = Theta^T phi by rocBLAS, a fused per-frame
// log-sum-exp / posterior pass, and FP64 fixed-order sums of FP32 partials.
#pragma once

#include "ctx.h"

#include <rocblas/rocblas.h>

namespace gmmhip {

// What theta currently holds: a full model, a diagonal model over the full
// feature map (full second-order statistics of a diagonal model), or a
// diagonal model over [1, x', x'^2].
enum GemmOperand { OP_FULL = 0, OP_DIAG_FULLFEAT = 1, OP_DIAG = 2 };

struct FullState {
    rocblas_handle blas = nullptr;
    int K = 0, KP = 0;           // features, padded to a multiple of 8
    float *d_theta = nullptr;    // [M][KP]
    size_t theta_cap = 0;
    int *d_pairs = nullptr;      // [K] feature code (see feature_codes)
    size_t pairs_cap = 0;
    float *d_cen = nullptr;      // [D]
    size_t cen_cap = 0;
    float *d_phi = nullptr;      // [nc][KP]
    size_t phi_cap = 0;
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

// The handle's GEMM-path state (created on first use: rocBLAS handle on the
// handle's stream, atomics off).
int gemm_state(gmmhip_ctx *h, gmmhip::FullState **out);

// The GEMM operand theta of the handle's diagonal model, over the full
// feature map (zero off-diagonal precisions: full second-order statistics of
// a diagonal model) or over [1, x', x'^2] (fullcov.hip).
int diag_operand(gmmhip_ctx *h, bool full_features);

hipError_t launch_fc_features(const float *X, uint64_t n0, int nc, int ncp, int D, int K, int KP,
                              const float *cen, const int *codes, float *phi, uint8_t *bad,
                              hipStream_t s);
// cut = 1: posteriors g > min_gamma kept, others 0 (statistics); cut = 0:
// posteriors before the cut.  Clamped / non-finite frames get 0 either way.
hipError_t launch_fc_lse(float *L, int nc, int ncp, int M, const uint8_t *bad, float min_ll,
                         float min_gamma, int cut, float *LLnat, uint64_t n0, double *blockLL,
                         hipStream_t s);
hipError_t launch_fc_acc(const float *part, size_t E, int S, int first, double *acc,
                         hipStream_t s);

#define BLASCHK(h, expr)                                                                \
    do {                                                                                \
        rocblas_status s_ = (expr);                                                     \
        if (s_ != rocblas_status_success)                                               \
            return fail((h), s_ == rocblas_status_memory_error ? GMMHIP_ERR_OOM         \
                                                               : GMMHIP_ERR_HIP,        \
                        "%s failed: %s", #expr, rocblas_status_to_string(s_));          \
    } while (0)

}  // namespace gmmhip_detail
