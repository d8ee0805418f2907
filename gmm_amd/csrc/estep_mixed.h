// Internal interface of the mixed-precision E-step (estep_mixed.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmmhip {

// Contraction depth of the mixed log-likelihoods: [1, z(D), z^2(D)] padded
// to 32.
inline int kp_mixed_contract(int D) { return (2 * D + 1 + 31) / 32 * 32; }
// Layout depth of the split feature rows and the split model (KP below):
// the contraction depth rounded up to a power of two, zero columns past
// 2D+1, so that every depth takes the statistics kernel with transposing
// LDS reads.  The log-likelihood kernel reads only the first
// kp_mixed_contract(D) columns.
inline int kp_mixed_for_dim(int D) {
    const int k = kp_mixed_contract(D);
    return k <= 32 ? 32 : k <= 64 ? 64 : k <= 128 ? 128 : 256;
}

// largest feature dimension of the mixed forms (2D+1 <= 256)
constexpr int PREP_MAX_D = 127;

struct PrepArgs {
    const float *X;        // [N][D] contiguous
    uint64_t N, Npad;      // Npad = ceil32(N)
    int D, KP;
    const float *off;      // [D] centring
    const float *scale;    // [D] 1/std
    uint16_t *za_hi_raw, *za_lo_raw;  // [Npad][KP]
    uint8_t *bad;          // [N] written for every frame: 1 non-finite, 2 out of the 16-bit range, 0 ok
    unsigned int *nrange;  // count of (frame, dim) values flagged 2
    float zlimit;          // |z| above this is out of range (fp16: 240)
};

struct MixedArgs {
    uint64_t N, Npad;
    int D, KP;
    const void *za_hi, *za_lo;
    const uint8_t *bad;
    const void *P16;       // split model [Mt32][KP/16][2][64] x 16 B
    int M, Mt32;
    float min_ll, min_gamma;
    float *LL2, *LLnat;
    double *blockLL;
    int ll_nw;             // log-likelihood waves per block: 4 (128 frames) or 8 (256, the
                           // contraction depth a multiple of 64)
    int st_nw;             // statistics waves per block: 4 (128 components) or 8 (256; KP 64
                           // or 128); ngroups = ceil(Mt32 / st_nw)
    uint64_t chunk_frames;
    int nchunks, ngroups;
    float *partials;       // [nchunks][Mt32*32][KP]
    // optional explicit chunk frame ranges [chunk_lo[c], chunk_hi[c]); each
    // chunk_lo must be a multiple of 32 (null: chunk c covers chunk_frames
    // frames from c * chunk_frames)
    const uint64_t *chunk_lo, *chunk_hi;
    // persistent form (launch_stats_mixed_queue): eight item counters, zero
    // at launch, and the number of blocks
    unsigned int *queue;
    int queue_blocks;
};

// Split model of the mixed form: for the feature map z = (x - off) * scale,
// the log2-unit coefficients [const, nu*omega, -omega/2] of component m
// (nu = scale (mu - off), omega = 1 / (var scale^2)), formed in FP64 and
// split hi = T(v), lo = T(v - hi), in the MFMA B-operand layout
// [Mt32][KP/16][hi, lo][64 lanes][8]; component slots past M are zero.
struct SplitModelArgs {
    const float *mu, *var;     // [M][D]
    const double *gconst;      // [M]
    const float *off, *scale;  // [D]
    int M, D, KP, Mt32;
    uint16_t *out;             // Mt32 * 32 slots * KP * 2 (hi, lo) values
    unsigned int *overflow;    // set non-zero when a hi piece is not finite
};
// prec: GMMHIP_PREC_BF16X3 (1) or GMMHIP_PREC_FP16X3 (2)
hipError_t launch_pack_split_model(const SplitModelArgs &a, int prec, hipStream_t s);

hipError_t launch_colstats(const float *X, uint64_t N, int D, int blocks, uint64_t per_block,
                           double *part, hipStream_t s);
// prec: GMMHIP_PREC_BF16X3 (1) or GMMHIP_PREC_FP16X3 (2)
hipError_t launch_prep_frames(const PrepArgs &p, int prec, hipStream_t s);
hipError_t launch_loglik_mixed(const MixedArgs &a, int prec, hipStream_t s);
// Frames per block of the mixed log-likelihood kernel at feature dimension D
// with nw waves (4, or 8 where the contraction depth is a multiple of 64):
// the number of per-block log-likelihood sums it writes is ceil(Npad /
// that).
int mixed_ll_frames_per_block(int D, int nw);
// Statistics waves per block for a component tile of comp_tile components
// (128 or 256) at depth KP: 8 for 256 at KP 64 or 128, else 4.
int mixed_stats_waves(int KP, int comp_tile);
hipError_t launch_stats_mixed(const MixedArgs &a, int prec, hipStream_t s);
// Persistent form of the statistics kernel: queue_blocks blocks pull
// (chunk, group) items, chunk-major, from eight per-XCD lists
// (counters a.queue[0..7], zero at launch) until all are empty.  Same
// partials as launch_stats_mixed, bit for bit.
hipError_t launch_stats_mixed_queue(const MixedArgs &a, int prec, hipStream_t s);
// Blocks for the persistent form: the resident blocks of the chip (at most
// the item count rounded up to a multiple of 8).
int mixed_stats_queue_blocks(int KP, int st_nw, int num_cus, int nitems);

}  // namespace gmmhip
