// gmmhip_accumulate_noalloc: the plain FP32 E-step statistics of device
// frames into device outputs with no device or host allocation, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops).
//
// Its scratch -- per-frame log-likelihoods, per-block log-likelihood sums,
// FP32 chunk partials and the running FP64 sums -- lives in the file-scope
// device array `scratch` (GMMHIP_NOALLOC_SCRATCH_BYTES, held by every device
// the library runs on until the process ends).  Host code never uses the
// array's name as an address: gmmhip_create fetches the array's device
// address for its device once (hipGetSymbolAddress with that device
// current) and the kernels take that pointer.  Every call on one device
// shares the array, so calls that overlap on one device race.
//
// The frames are processed in segments of whole chunks (chunk boundaries as
// gmmhip_accumulate places them on the same frames), each segment's chunk
// partials added in chunk order to the running FP64 sums, so n, F and S are
// bit-identical to gmmhip_accumulate's; the total log-likelihood is summed
// per segment (equal to FP64 rounding).
#include "ctx.h"

#include <mutex>

using namespace gmmhip_detail;

namespace {

// scratch layout (bytes): [red: E doubles | tot: 1 double | blockLL | LL2 | partials]
constexpr size_t kRedBytes = (size_t)4 << 20;       // E = Mp * KP <= 524288
constexpr size_t kTotBytes = 256;
constexpr size_t kBlockLLBytes = (size_t)128 << 10; // 16384 blocks of >= 64 frames
constexpr size_t kLL2Bytes = (size_t)4 << 20;       // 1,048,576 frames per segment
constexpr size_t kPartBytes = (size_t)16 << 20;     // chunk partials of one segment
constexpr size_t kScratchBytes = kRedBytes + kTotBytes + kBlockLLBytes + kLL2Bytes + kPartBytes;
static_assert(kScratchBytes == GMMHIP_NOALLOC_SCRATCH_BYTES, "scratch size of the header");
constexpr size_t kMaxE = kRedBytes / sizeof(double);
constexpr uint64_t kMaxSegFrames = kLL2Bytes / sizeof(float);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};

// out[e] = (first ? 0 : out[e]) + part[0][e] + part[1][e] + ... in chunk
// order: continuing one running sum across segments gives the bits of a
// single pass over all chunks.
__global__ __launch_bounds__(256) void reduce_continue_kernel(const float *__restrict__ part,
                                                              int nchunks, size_t E, int first,
                                                              double *__restrict__ out) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= E) return;
    double s = first ? 0.0 : out[e];
    for (int c = 0; c < nchunks; c++) s += (double)part[(size_t)c * E + e];
    out[e] = s;
}

// tot = (first ? 0 : tot) + fixed-order sum of the segment's block sums.
__global__ __launch_bounds__(256) void total_continue_kernel(const double *__restrict__ v,
                                                             uint64_t n, int first,
                                                             double *__restrict__ tot) {
    __shared__ double sh[256];
    double s = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += 256) s += v[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[0] = (first ? 0.0 : tot[0]) + sh[0];
}

// Un-centre the FP64 sums [n | F' | S'] (frames were centred by off) and
// add them into the caller's outputs: F += F' + n o, S += S' + 2 o F' + n o^2,
// in the operation order of gmmhip_accumulate's host loop, unfused.
__global__ __launch_bounds__(256) void uncentre_add_kernel(const double *__restrict__ red, int KP,
                                                           const float *__restrict__ off, int M,
                                                           int D, int flags, const double *tot,
                                                           double *__restrict__ n,
                                                           double *__restrict__ F,
                                                           double *__restrict__ S,
                                                           double *__restrict__ tot_out) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) tot_out[0] += tot[0];
    if (i >= (size_t)M * D) return;
    const int m = (int)(i / D), d = (int)(i - (size_t)m * D);
    const double *r = red + (size_t)m * KP;
    const double nm = r[0];
    if (d == 0) n[m] += nm;
    const double o = (double)off[d];
    const double f1 = r[1 + d], s1 = r[1 + D + d];
    if ((flags & GMMHIP_ACC_MEAN) && F) F[i] += f1 + nm * o;
    if ((flags & GMMHIP_ACC_VAR) && S) S[i] += s1 + 2.0 * o * f1 + nm * o * o;
}

}  // namespace

namespace gmmhip_detail {

// Device address of `scratch` on device dev (current), fetched once per
// device; nullptr if the runtime cannot give it.
unsigned char *noalloc_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_ptr[dev] = (unsigned char *)p;
    }
    return g_ptr[dev];
}

}  // namespace gmmhip_detail

extern "C" int gmmhip_accumulate_noalloc(gmmhip_t h, const float *frames, uint64_t N, int D,
                                         int flags, double *n, double *F, double *S,
                                         double *tot_ll, uint64_t *nframes) {
    if (!h) return GMMHIP_ERR_ARG;
    if (h->M <= 0) return fail(h, GMMHIP_ERR_STATE, "accumulate_noalloc: no model inserted");
    if (!frames || N == 0 || D != h->D)
        return fail(h, GMMHIP_ERR_ARG, "accumulate_noalloc: need device frames [N][%d]", h->D);
    if (!n || !tot_ll || !nframes || ((flags & GMMHIP_ACC_MEAN) && !F) ||
        ((flags & GMMHIP_ACC_VAR) && !S))
        return fail(h, GMMHIP_ERR_ARG, "accumulate_noalloc: null output for a requested statistic");
    if (flags & GMMHIP_ACC_FULLVAR)
        return fail(h, GMMHIP_ERR_ARG, "accumulate_noalloc: diagonal statistics only");
    if (h->full_model || h->d_mask)
        return fail(h, GMMHIP_ERR_ARG,
                    "accumulate_noalloc: needs a diagonal model and no Gaussian-selection mask");
    if (!h->noalloc)
        return fail(h, GMMHIP_ERR_STATE, "accumulate_noalloc: scratch address not available");
    const int KP = h->KP;
    const int ngroups = (h->Mt + 3) / 4;
    const size_t Mp = (size_t)h->Mt * 16;
    const size_t E = Mp * KP;
    if (E > kMaxE)
        return fail(h, GMMHIP_ERR_ARG,
                    "accumulate_noalloc: M x depth = %zu x %d exceeds the scratch maximum of %zu "
                    "(e.g. M <= 4096 at D <= 63, M <= 2048 at D <= 127)", Mp, KP, kMaxE);
    // chunks as gmmhip_accumulate places them on N frames; whole chunks per
    // segment, bounded by the partials and the per-frame regions
    const uint64_t cf = estep_chunk_frames(h, N, ngroups, false);
    const int ll_sub = choose_ll_sub(h, GMMHIP_TUNE_ESTEP);
    uint64_t seg_chunks = std::min<uint64_t>(kPartBytes / (E * sizeof(float)), kMaxSegFrames / cf);
    // per-block sums of one segment must fit too
    while (seg_chunks > 1 && ll_blocks(seg_chunks * cf, ll_sub) * sizeof(double) > kBlockLLBytes)
        seg_chunks--;
    if (seg_chunks < 1)
        return fail(h, GMMHIP_ERR_ARG,
                    "accumulate_noalloc: one chunk of %llu frames does not fit the scratch",
                    (unsigned long long)cf);
    const uint64_t seg_frames = seg_chunks * cf;

    unsigned char *base = h->noalloc;
    double *red = (double *)base;
    double *tot = (double *)(base + kRedBytes);
    double *blockLL = (double *)(base + kRedBytes + kTotBytes);
    float *LL2 = (float *)(base + kRedBytes + kTotBytes + kBlockLLBytes);
    float *part = (float *)(base + kRedBytes + kTotBytes + kBlockLLBytes + kLL2Bytes);

    DevGuard g(h->dev);
    hipStream_t st = h->stream;
    int first = 1;
    for (uint64_t s0 = 0; s0 < N; s0 += seg_frames, first = 0) {
        const uint64_t ns = std::min<uint64_t>(seg_frames, N - s0);
        EstepArgs a{};
        a.X = frames + s0 * (uint64_t)D;
        a.ldx = D;
        a.N = ns;
        a.D = D;
        a.KP = KP;
        a.P = h->d_P;
        a.M = h->M;
        a.Mt = h->Mt;
        a.off = h->d_off;
        a.min_ll = h->opt.min_loglike;
        a.min_gamma = h->opt.min_gamma;
        a.LL2 = LL2;
        a.LLnat = nullptr;
        a.blockLL = blockLL;
        a.ll_sub = ll_sub;
        a.ngroups = ngroups;
        a.chunk_frames = cf;
        a.nchunks = (int)((ns + cf - 1) / cf);
        a.partials = part;
        HIPCHK(h, launch_loglik_fp32(a, st));
        HIPCHK(h, launch_stats_fp32(a, st));
        hipLaunchKernelGGL(reduce_continue_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0,
                           st, part, a.nchunks, E, first, red);
        HIPCHK(h, hipGetLastError());
        hipLaunchKernelGGL(total_continue_kernel, dim3(1), dim3(256), 0, st, blockLL,
                           ll_blocks(ns, ll_sub), first, tot);
        HIPCHK(h, hipGetLastError());
    }
    const size_t MD = (size_t)h->M * D;
    hipLaunchKernelGGL(uncentre_add_kernel, dim3((unsigned)((MD + 255) / 256)), dim3(256), 0, st,
                       red, KP, h->d_off, h->M, D, flags, tot, n, F, S, tot_ll);
    HIPCHK(h, hipGetLastError());
    *nframes += N;
    return GMMHIP_OK;
}
