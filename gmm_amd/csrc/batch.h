// Internal interface of the batched per-utterance kernels (batch.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmmhip {

struct SegReduceArgs {
    const float *partials;   // [chunks][Mp][KP]
    const int *chunk_first;  // [U+1] first chunk of each utterance
    const float *off;        // [D] centring of the partials
    const float *scale;      // [D] scaling of the partials (mixed form) or null
    int U, M, Mp, D, KP;
    double *n;               // [U][M]
    double *F;               // [U][M][D]
    // MAP means of the first Umeans utterances (null: none)
    const float *mu;         // [M][D] UBM means
    int tau;                 // relevance factor
    int Umeans;
    float *means;            // [Umeans][M][D]
};

struct MllrSolveArgs {
    const double *Gp;     // [(u,i)][ldg] packed upper triangle of G
    const double *Kp;     // [(u,i)][ldk] k
    const double *Ginit;  // [D][E(E+1)/2] UBM prior, packed as Gp, or null
    const double *kinit;  // [D][E] or null
    int U, D;
    int ldg, ldk;
    int *status;          // [U] in: 0 / 7 (under-determined); out: |= 8 if G not SPD
    double *W;            // [(u,i)][E]
};

// MLLR statistics of many utterances as one GEMM per kind:
//   G[(u,i)][p] = sum_m n[u][m] ivar[m][i] P[m][p]   (B = P,  Cn = (D+1)(D+2)/2)
//   k[(u,i)][e] = sum_m F[u][m][i] ivar[m][i] Xi[m][e] (B = Xi, Cn = D+1)
// with components of zero occupancy left out; the A operand is formed while
// it is staged, so it never exists in HBM.
// Components the batched MLLR statistics support (the occupied-component
// list of an utterance lives in LDS).
constexpr int MLLR_MAX_COMPONENTS = 4096;

struct MllrGkArgs {
    const double *n;     // [U][M]
    const double *F;     // [U][M][D]
    const double *ivar;  // [M][D]
    const double *B;     // [M][ldb]
    int ldb, Cn;
    int U, M, D;
    double *C;           // [U*D][ldc]
    int ldc;
    int kmode;           // 0: G, 1: k
};

hipError_t launch_fill_u32(unsigned int *p, unsigned int v, int n, hipStream_t s);
// Utterance rows X[src[u] .. src[u+1]) copied to Xp[dst[u] ..), zero up to dst[u+1].
hipError_t launch_pad_rows(const float *X, const uint64_t *src, const uint64_t *dst, int U, int D,
                           float *Xp, hipStream_t s);
hipError_t launch_mllr_gk(const MllrGkArgs &a, hipStream_t s);
hipError_t launch_seg_reduce(const SegReduceArgs &a, hipStream_t s);
hipError_t launch_mllr_check(const double *n, int U, int M, int D, int *status, hipStream_t s);
hipError_t launch_mllr_solve(const MllrSolveArgs &a, hipStream_t s);
// out[u] = sum of v[begin[u] .. end[u]) in FP64; the summation order is
// fixed relative to begin[u] (one block per segment)
hipError_t launch_seg_sum(const float *v, const uint64_t *begin, const uint64_t *end, int U,
                          double *out, hipStream_t s);

}  // namespace gmmhip
