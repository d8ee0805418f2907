// Handle state shared by the C ABI translation units (internal).
#pragma once

#include "../../include/gmmhip.h"
#include "estep.h"
#include "estep_mixed.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using namespace gmmhip;

struct gmmhip_batch_state;  // batched per-utterance buffers (batch_api.hip)
namespace gmmhip {
struct FullState;           // full-covariance model and buffers (fullcov.hip)
struct AuxState;            // robust auxiliary statistics buffers (auxstats.hip)
}

struct gmmhip_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    gmmhip_opts opt{};
    std::string err;

    // model
    int M = 0, D = 0, KP = 0, Mt = 0;
    float *d_P = nullptr;
    size_t P_cap = 0;
    float *d_off = nullptr;
    size_t off_cap = 0;
    std::vector<float> h_off;

    // frames
    const float *d_x = nullptr;
    uint64_t N = 0, ld = 0;
    int xD = 0;
    float *d_xown = nullptr;
    size_t xown_cap = 0;

    // scratch
    float *d_LL2 = nullptr;
    size_t LL2_cap = 0;
    float *d_LLnat = nullptr;
    size_t LLnat_cap = 0;
    double *d_blockLL = nullptr;
    size_t blockLL_cap = 0;
    float *d_part = nullptr;
    size_t part_cap = 0;
    double *d_red = nullptr;
    size_t red_cap = 0;
    double *d_tot = nullptr;
    double *h_res = nullptr;  // pinned: [Mp*KP] + 1
    size_t h_res_cap = 0;

    // host copy of the active model (for the mixed-precision packing)
    std::vector<float> m_w, m_mu, m_var;
    std::vector<double> m_gconst;
    uint64_t model_version = 0, frames_version = 0;

    // mixed-precision frame features (ZA hi/lo) and split model
    int z_prec = 0, KPm = 0;
    uint64_t z_version = 0, z_N = 0, z_Npad = 0;
    bool z_range_fallback = false;
    uint16_t *d_za_hi = nullptr, *d_za_lo = nullptr;
    size_t z_cap = 0;
    uint8_t *d_bad = nullptr;
    size_t bad_cap = 0;
    unsigned int *d_nrange = nullptr;
    double *d_colpart = nullptr;
    size_t colpart_cap = 0;
    float *d_zoff = nullptr, *d_zscale = nullptr;
    size_t zos_cap = 0, zscale_cap = 0;
    std::vector<float> h_zoff, h_zscale;
    void *d_P16 = nullptr;
    size_t P16_cap = 0;
    uint64_t p16_model = 0, p16_z = 0;
    int p16_prec = 0;
    bool p16_overflow = false;
    bool last_used_mixed = false;

    // full-covariance model (fullcov.hip); the diagonal fields above are
    // unused while full_model is set
    gmmhip::FullState *full = nullptr;
    bool full_model = false;
    gmmhip::AuxState *aux = nullptr;

    // Gaussian selection (gmmhip_set_mask): per-frame component bitmap
    // [mask_N][mask_W] in HBM; while set, every E-step call runs the GEMM
    // form with it
    uint32_t *d_mask = nullptr;
    size_t mask_cap = 0;
    uint64_t mask_N = 0;
    int mask_W = 0, mask_M = 0;

    // batched per-utterance state (batch_api.hip)
    gmmhip_batch_state *batch = nullptr;
    float *d_mu_f32 = nullptr, *d_var_f32 = nullptr;  // UBM means / variances
    size_t mu_f32_cap = 0, var_f32_cap = 0;
    double *d_gconst = nullptr;  // per-component Gaussian constants (FP64)
    size_t gconst_cap = 0;
    unsigned int *d_flag = nullptr;  // split-model overflow flag
    uint64_t mu_f32_version = ~0ull;
    int num_cus = 0;

    // device address of the no-allocation entry point's file-scope scratch
    // array on this handle's device (noalloc.hip), fetched at creation
    unsigned char *noalloc = nullptr;

    // tiles of the last call (gmmhip_tiles_used)
    int used_ll_frames = 0, used_chunk = 0, used_comp_tile = 0;

    // kernel timers (opts.profile)
    hipEvent_t ev[5] = {};
    bool ev_ok = false;
    double kt_ms[GMMHIP_KT_COUNT] = {};
    uint64_t kt_n[GMMHIP_KT_COUNT] = {};
};

namespace gmmhip_detail {

inline int fail(gmmhip_ctx *h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, expr)                                                              \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail((h), e_ == hipErrorOutOfMemory ? GMMHIP_ERR_OOM : GMMHIP_ERR_HIP, \
                        "%s failed: %s", #expr, hipGetErrorString(e_));             \
    } while (0)

// Bumped whenever a device buffer of any handle is (re)allocated or freed
// for regrowth.  A captured hipGraph holds raw device addresses, so its
// cache key includes this generation: a replay after any reallocation would
// otherwise touch freed memory.
inline std::atomic<uint64_t> g_alloc_generation{0};

inline void bump_alloc_generation() { g_alloc_generation.fetch_add(1, std::memory_order_relaxed); }

// Grow-only device buffer.
template <typename T>
inline hipError_t ensure(T **p, size_t *cap, size_t n) {
    if (*cap >= n && *p) return hipSuccess;
    bump_alloc_generation();
    if (*p) {
        hipError_t e = hipFree(*p);
        if (e != hipSuccess) return e;
        *p = nullptr;
        *cap = 0;
    }
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) *cap = n;
    return e;
}

// Rows of the mixed form's frame buffers: a multiple of the 128-frame
// log-likelihood block.
inline uint64_t pad128(uint64_t n) { return (n + 127) / 128 * 128; }

inline void apply_defaults(gmmhip_opts *o) {
    if (o->min_loglike == 0.0f) o->min_loglike = -1e20f;
    if (o->min_gamma == 0.0f) o->min_gamma = 1e-4f;
    if (o->min_var == 0.0f) o->min_var = 1e-6f;
}

struct DevGuard {
    int prev = -1;
    explicit DevGuard(int d) {
        if (hipGetDevice(&prev) == hipSuccess && prev != d) hipSetDevice(d);
        else prev = -1;
    }
    ~DevGuard() {
        if (prev >= 0) hipSetDevice(prev);
    }
};

// Device address of gmmhip_accumulate_noalloc's scratch array on device dev
// (which must be current), fetched with hipGetSymbolAddress once per device;
// nullptr when the runtime cannot give it (noalloc.hip).
unsigned char *noalloc_scratch(int dev);

// Frees the batched state of a handle (batch_api.hip).
void batch_state_free(gmmhip_batch_state *b);

// Full-covariance path (fullcov.hip): state release, per-frame
// log-likelihoods into device memory, statistics accumulation with the
// packed upper-triangle second-order statistics.
void full_state_free(gmmhip::FullState *fs);
void aux_state_free(gmmhip::AuxState *a);
// GMMHIP_ERR_ARG when a Gaussian-selection mask does not fit the handle's
// frames and model.
int mask_check(gmmhip_ctx *h);
int full_accumulate(gmmhip_ctx *h, int flags, double *n, double *F, double *S, double *tot_ll,
                    uint64_t *nframes);
// Full second-order statistics of the handle's diagonal model (GMMHIP_ACC_FULLVAR).
int diag_full_accumulate(gmmhip_ctx *h, int flags, double *n, double *F, double *S,
                         double *tot_ll, uint64_t *nframes);
// The GEMM form for the handle's model (diagonal or full): per-frame
// log-likelihoods into device memory, statistics (full_accumulate after
// gemm_operand_for_model), posteriors before the cut into host memory.
int gemm_operand_for_model(gmmhip_ctx *h);
int gemm_loglikes_device(gmmhip_ctx *h, float *d_out);
int gemm_posteriors(gmmhip_ctx *h, float *ll, float *post);

// Device copies of the active model for the per-utterance kernels: FP32
// means / variances and, with mllr, the UBM-only MLLR operands
// P[m][packed a<=b] = xi_a xi_b, Xi[m][a] (xi = [mu, 1]), the prior G (one
// per row i, packed like P), the prior k and 1/var in FP64 (batch_api.hip).  Valid until the next gmmhip_set_model.
struct ModelOperands {
    const float *mu, *var;
    const double *gconst;
    const double *P, *Xi, *Ginit, *kinit, *ivar;
};
int model_operands(gmmhip_ctx *h, bool mllr, ModelOperands *o);

// Split model of the mixed form for the feature map z = (x - zoff) * zscale
// (device arrays of D): [Mt32][KPm/16][hi, lo][64 lanes][8] 16-bit pieces of
// the log2-unit coefficients [const, nu*omega, -omega/2], formed on the
// device into d_out (Mt32 * KPm * 64 values) on the handle's stream
// (gmmhip_api.hip).  *overflow is set when a coefficient does not fit the
// 16-bit type; the call waits for the stream.
int pack_split_model(gmmhip_ctx *h, int prec, int KPm, const float *d_zoff, const float *d_zscale,
                     uint16_t *d_out, bool *overflow);

// Frames per statistics work item of an E-step over N frames: the
// opts.chunk_frames override, else enough items for ~4096 blocks, capped by
// the tuning table (8192 without an entry), a multiple of the statistics
// kernel's 64-frame step (gmmhip_api.hip).
uint64_t estep_chunk_frames(const gmmhip_ctx *h, uint64_t N, int ngroups, bool mixed);

// Log-likelihood frame tile of the FP32 form (in units of 64 frames):
// opts.ll_frames, else the tuning table's entry for the path (then the
// E-step entry), else 2; 4 only for contraction depths up to 128
// (gmmhip_api.hip).
int choose_ll_sub(const gmmhip_ctx *h, int path);
// Log-likelihood waves per block of the mixed form (4: 128-frame blocks,
// 8: 256-frame blocks, only where the depth is a multiple of 64):
// opts.ll_frames, else the tuning table's entry for (path, precision),
// else 4 (gmmhip_api.hip).
int choose_mixed_ll_nw(const gmmhip_ctx *h, int path);
// Mixed statistics waves per block: opts.comp_tile, else the tuning table's
// entry for (path, precision), else 128 components (4 waves).
int choose_mixed_st_nw(const gmmhip_ctx *h, int path);
// Most frames per statistics work item from the tuning table for (path,
// precision) and the handle's model, rounded to the 64-frame step; 0 when
// the table has no entry.
uint64_t tuned_chunk(const gmmhip_ctx *h, int path, int prec);

}  // namespace gmmhip_detail
