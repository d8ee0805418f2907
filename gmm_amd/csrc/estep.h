// Internal interface between the C ABI (gmmhip_api.hip) and the E-step
// kernels (estep_fp32.hip).  Not installed; the public surface is
// include/gmmhip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmmhip {

// Frames per E-step log-likelihood block: 4 waves x SUB sub-tiles x 16,
// SUB = 1, 2 (default) or 4.
constexpr int LL_FRAMES_PER_BLOCK = 128;
// Frames staged per iteration of the statistics kernel.
constexpr int ST_FRAMES_PER_ITER = 64;
// Components owned by one statistics block (4 waves x 16).
constexpr int ST_COMPS_PER_BLOCK = 64;

// Row stride (floats) of the LDS frame tile [1, x, x^2, 0...] for a
// contraction depth KP: a multiple of 64 plus 4, which keeps both MFMA
// operand reads of the tile free of bank conflicts.
constexpr int lds_stride(int kp) { return (kp + 63) / 64 * 64 + 4; }

// Contraction depth for D dims: [1, x(D), x^2(D)] padded to 16.
inline int kp_for_dim(int D) { return (2 * D + 1 + 15) / 16 * 16; }
constexpr int KP_MAX = 256;  // D <= 127

struct EstepArgs {
    // frames, row-major [N][ldx], resident in HBM
    const float *X;
    uint64_t ldx;
    uint64_t N;
    int D;
    int KP;
    // packed model: [Mt][KP/4][64] floats (log2 units), centring offset [D]
    const float *P;
    int M;
    int Mt;
    const float *off;
    float min_ll;
    float min_gamma;
    // per-frame outputs of the log-likelihood kernel
    float *LL2;      // [N] log2-domain LL, -inf marks "no statistics"
    float *LLnat;    // [N] natural LL (may be null)
    double *blockLL; // [ll_blocks(N, ll_sub)] per-block sums
    int ll_sub;      // log-likelihood frame tile / 64: 1, 2, 4 (0 = 2)
    // statistics
    uint64_t chunk_frames;
    int nchunks;
    int ngroups;
    float *partials; // [nchunks][Mt*16][KP]
    double *red;     // [Mt*16][KP]
    double *tot;     // [1]
    // batched mode: explicit chunk frame ranges [chunk_lo[c], chunk_hi[c])
    // (null: chunk c covers chunk_frames frames from c*chunk_frames) and a
    // device work queue of nchunks*ngroups items (null: one block per item)
    const uint64_t *chunk_lo;
    const uint64_t *chunk_hi;
    unsigned int *queue;
    int queue_blocks;
};

inline uint64_t ll_blocks(uint64_t N, int sub = LL_FRAMES_PER_BLOCK / 64) {
    const uint64_t f = 64u * (uint64_t)sub;
    return (N + f - 1) / f;
}
inline int ll_sub_of(const EstepArgs &a) { return a.ll_sub > 0 ? a.ll_sub : 2; }

hipError_t launch_loglik_fp32(const EstepArgs &a, hipStream_t s);
hipError_t launch_stats_fp32(const EstepArgs &a, hipStream_t s);
// persistent variant: a.queue_blocks blocks pull (chunk, group) items from
// *a.queue (which must hold 0 at launch)
hipError_t launch_stats_fp32_queue(const EstepArgs &a, hipStream_t s);
hipError_t launch_reduce_stats(const EstepArgs &a, hipStream_t s);
hipError_t launch_sum_blockll(const EstepArgs &a, hipStream_t s);
// out[e] = sum_c part[c*E + e] in FP64, c = 0..nchunks-1 in order
hipError_t launch_reduce_partials(const float *part, int nchunks, size_t E, double *out,
                                  hipStream_t s);
hipError_t launch_sum_doubles(const double *v, uint64_t n, double *out, hipStream_t s);

}  // namespace gmmhip
