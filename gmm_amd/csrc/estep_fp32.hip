// FP32 E-step kernels for gfx950 (plain form).
//
// The per-frame-by-component log-likelihood is the dense contraction
//   l2[t,m] = sum_k Z[t,k] * B[k,m],  Z[t,:] = [1, x', x'^2, 1, 0...],
//   B[:,m]  = log2(e) * [c_m hi, mu'_m * ivar_m, -ivar_m / 2, c_m lo, 0...]
// with x' = x - off (mean-centred frames, mu' = mu - off), evaluated on
// v_mfma_f32_16x16x4_f32.  It restates gammasKernel
// (trainer/CU_GMMStatsEstimator_CUDAkernels.cuh:681-750), which forms
// gconst - 1/2 sum (x-mu)^2 ivar per frame and component on scalar FMAs.
//
// Three kernels run per accumulate call:
//   loglik_kernel  per-frame log-sum-exp over all components, online in
//                  registers, lanes combined with xor shuffles
//                  (logLikeKernel + checkLogLike, kernels.cuh:922-995,
//                  :96-110; sumLogLike :998-1032 becomes a fixed-order
//                  per-block FP64 sum);
//   stats_kernel   recomputes l2 for its 64 components over a frame chunk,
//                  forms posteriors g = 2^(l2 - LL2) with the 1e-4 cut
//                  (gammasNormKernel, kernels.cuh:1037-1110), and contracts
//                  g^T [1, x', x'^2] on MFMA into FP32 register accumulators
//                  (accStatsDiagKernel, kernels.cuh:582-676);
//   reduce_kernel  fixed-order FP64 sum of the FP32 chunk partials
//                  (addArrays4Kernel/addArraysKernel, kernels.cuh:1115-1154).
// Posteriors never touch HBM.
#include "estep.h"

#include <hip/hip_runtime.h>
#include <math.h>

namespace gmmhip {

typedef float f32x4 __attribute__((ext_vector_type(4)));

static constexpr float kLn2 = 0.69314718055994530942f;

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float fast_exp2(float x) {
    return __builtin_amdgcn_exp2f(x);
}

__device__ __forceinline__ float fast_log2(float x) {
    return __builtin_amdgcn_logf(x);
}

// Row k of the frame feature vector [1, x', x'^2, 1, 0...] for one frame:
// the second 1 (column 2D+1) carries the low part of the model constant.
__device__ __forceinline__ float zval(const float *__restrict__ xr,
                                      const float *__restrict__ off, int D,
                                      int k) {
    if (k == 0) return 1.0f;
    if (k <= D) return xr[k - 1] - off[k - 1];
    if (k <= 2 * D) {
        float v = xr[k - 1 - D] - off[k - 1 - D];
        return v * v;
    }
    return k == 2 * D + 1 ? 1.0f : 0.0f;
}

// ---------------------------------------------------------------------------
// Contraction index mapping shared by both kernels: MFMA step kk of lane
// group q (q = lane/16) covers k = q*KK + kk, so a lane's operands for
// consecutive steps are consecutive in memory and load as float4.  The model
// is packed as P[t][lane][kk] = B[q*KK + kk][16t + c] (c = lane%16).
//
// loglik_kernel: one block = 4 waves x SUB sub-tiles x 16 frames (SUB = 1,
// 2 or 4: the frame tile is a tuning choice).  Each wave keeps the MFMA
// A-fragments of its 16 SUB frames in registers for the whole
// component loop; component tiles of 16 are staged through LDS (register
// prefetch of the next tile overlaps the current tile's MFMAs) and shared by
// the 4 waves.  Output layout of the 16x16x4 MFMA: lane (q, c) holds rows
// 4q..4q+3 (frames) of column c (component).
template <int KP, int SUB>
__global__ __launch_bounds__(256) void loglik_kernel(EstepArgs a) {
    constexpr int KK = KP / 4;
    constexpr int KU = KK / 4;              // float4 per lane per tile
    constexpr int PSTR = KK | 4;            // padded LDS row stride (floats)
    constexpr int TILE4 = 16 * KK;          // float4 per component tile
    constexpr int PF = (TILE4 + 255) / 256; // float4 per thread per tile
    __shared__ float Pl[64 * PSTR];
    __shared__ double red[16];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, c = lane & 15;
    const uint64_t base = (uint64_t)blockIdx.x * (64 * SUB);
    const int D = a.D;

    float af[SUB][KK];
#pragma unroll
    for (int s = 0; s < SUB; s++) {
        const uint64_t fr = base + (uint64_t)(wave * SUB + s) * 16 + c;
        const bool ok = fr < a.N;
        const float *xr = a.X + (ok ? fr : 0) * a.ldx;
#pragma unroll
        for (int kk = 0; kk < KK; kk++)
            af[s][kk] = ok ? zval(xr, a.off, D, q * KK + kk) : 0.0f;
    }

    // Online log-sum-exp per (frame, lane): mx is a per-lane reference and
    // sm = sum 2^(v - mx).  A value more than 32 above the reference (rare
    // after the first tiles) moves the reference first, so every exp2
    // argument stays <= 32; the common path is sub + exp2 + add.  Padded
    // components carry l = -1e30 and vanish in exp2; a NaN frame makes sm NaN
    // and is clamped at the end.
    float mx[SUB][4], sm[SUB][4];

    const float4 *P4 = reinterpret_cast<const float4 *>(a.P);
    float4 pf[PF];
    auto prefetch = [&](int t) {
#pragma unroll
        for (int j = 0; j < PF; j++) {
            const int i = tid + 256 * j;
            if (i < TILE4) pf[j] = P4[(size_t)t * TILE4 + i];
        }
    };
    prefetch(0);
    for (int t = 0; t < a.Mt; t++) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PF; j++) {
            const int i = tid + 256 * j;
            if (i < TILE4) {
                const int ln = i / KU, u = i - ln * KU;
                *reinterpret_cast<float4 *>(&Pl[ln * PSTR + 4 * u]) = pf[j];
            }
        }
        __syncthreads();
        if (t + 1 < a.Mt) prefetch(t + 1);

        f32x4 acc[SUB];
#pragma unroll
        for (int s = 0; s < SUB; s++) acc[s] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int u = 0; u < KU; u++) {
            const float4 b = *reinterpret_cast<const float4 *>(&Pl[lane * PSTR + 4 * u]);
#pragma unroll
            for (int s = 0; s < SUB; s++) {
                acc[s] = mfma4(af[s][4 * u + 0], b.x, acc[s]);
                acc[s] = mfma4(af[s][4 * u + 1], b.y, acc[s]);
                acc[s] = mfma4(af[s][4 * u + 2], b.z, acc[s]);
                acc[s] = mfma4(af[s][4 * u + 3], b.w, acc[s]);
            }
        }

        if (t == 0) {
#pragma unroll
            for (int s = 0; s < SUB; s++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    mx[s][r] = acc[s][r];
                    sm[s][r] = 1.0f;
                }
        } else {
#pragma unroll
            for (int s = 0; s < SUB; s++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    // branch-free: a value more than 32 above the reference
                    // moves it (sm * 2^-d + 1), otherwise sm += 2^d
                    const float d = acc[s][r] - mx[s][r];
                    const bool up = d > 32.0f;
                    const float e = fast_exp2(up ? -d : d);
                    sm[s][r] = up ? fmaf(sm[s][r], e, 1.0f) : sm[s][r] + e;
                    mx[s][r] = up ? acc[s][r] : mx[s][r];
                }
        }
    }

    // combine the 16 component lanes of each frame row
#pragma unroll
    for (int s = 0; s < SUB; s++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const float m2 = __shfl_xor(mx[s][r], o);
                const float s2 = __shfl_xor(sm[s][r], o);
                const float hi = fmaxf(mx[s][r], m2);
                sm[s][r] = sm[s][r] * fast_exp2(mx[s][r] - hi) + s2 * fast_exp2(m2 - hi);
                mx[s][r] = hi;
            }
        }

    double local = 0.0;
    if (c == 0) {
#pragma unroll
        for (int s = 0; s < SUB; s++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint64_t fr = base + (uint64_t)(wave * SUB + s) * 16 + 4 * q + r;
                float l2 = mx[s][r] + fast_log2(sm[s][r]);
                float nat = l2 * kLn2;
                if (!(nat > a.min_ll) || isinf(nat)) {
                    // every component underflowed, or a NaN/Inf frame: the
                    // log-likelihood is clamped and +inf marks "no statistics"
                    nat = a.min_ll;
                    l2 = INFINITY;
                }
                if (fr < a.N) {
                    a.LL2[fr] = l2;
                    if (a.LLnat) a.LLnat[fr] = nat;
                    local += (double)nat;
                }
            }
        red[wave * 4 + q] = local;
    }
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int i = 0; i < 16; i++) s += red[i];
        a.blockLL[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// stats_kernel: block = (frame chunk, group of 64 components); wave w owns
// component tile 4*group + w and its 16 x KP statistics accumulators.  Blocks
// are numbered so that the groups of one chunk land on the same XCD (the
// dispatcher deals blocks round-robin over 8 XCDs), so a chunk's frames are
// fetched from HBM once per XCD L2.  The next 64 frames are prefetched into
// registers while the current ones are contracted.
//
// Statistics column mapping: MFMA jt of the g^T Z contraction covers column
// stats_col(c, jt) of lane c = lane%16.  Columns are [1, x', x'^2, 0...]:
// n, F', S'.
// Two waves per SIMD (<= 256 VGPRs) up to KP = 160 (D <= 79); larger
// depths keep their accumulators in registers at one wave per SIMD.
#define GMMHIP_STATS_BOUNDS __launch_bounds__(256, (KP <= 160 ? 2 : 1))

// Statistics column of MFMA jt in lane c (c = lane % 16).  When KP is a
// multiple of 64 the lane's columns come in runs of 4 spaced 64 apart, so
// each 16-byte LDS read of the g^T Z contraction covers 16 consecutive
// 16-byte pieces across the 16 lanes (distinct banks); otherwise a lane owns
// JT consecutive columns.
template <int KP>
__device__ __forceinline__ int stats_col(int c, int jt) {
    constexpr int JT = KP / 16;
    if constexpr (JT % 4 == 0) return (jt >> 2) * 64 + 4 * c + (jt & 3);
    else return c * JT + jt;
}

// Zr rows are read as 16-byte vectors: keep the tile 16-byte aligned.
template <int KP>
struct alignas(16) StatsShared {
    alignas(16) float Zr[ST_FRAMES_PER_ITER * lds_stride(KP)];
    float Ls[ST_FRAMES_PER_ITER];
    float offL[KP];
    int item;
};

// One (chunk, group) work item: all 256 threads of the block take part.
template <int KP>
__device__ __forceinline__ void stats_item(const EstepArgs &a, StatsShared<KP> &sh, int chunk,
                                           int group) {
    constexpr int KK = KP / 4;
    constexpr int KU = KK / 4;
    constexpr int JT = KP / 16;
    constexpr int S = lds_stride(KP);
    constexpr int FB = ST_FRAMES_PER_ITER;
    constexpr int NPF = KP / 8;  // >= ceil(FB*D/256) for 2D+1 <= KP
    float *Zr = sh.Zr;
    float *Ls = sh.Ls;
    const float *offL = sh.offL;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, c = lane & 15;

    const int tile = group * 4 + wave;
    const bool active = tile < a.Mt;
    const int D = a.D;
    const int FD = FB * D;

    float bf[KK];
    {
        const float4 *P4 = reinterpret_cast<const float4 *>(a.P);
#pragma unroll
        for (int u = 0; u < KU; u++) {
            const float4 v = active ? P4[((size_t)tile * 64 + lane) * KU + u]
                                    : make_float4(0.f, 0.f, 0.f, 0.f);
            bf[4 * u + 0] = v.x;
            bf[4 * u + 1] = v.y;
            bf[4 * u + 2] = v.z;
            bf[4 * u + 3] = v.w;
        }
    }

    f32x4 acc2[JT];
#pragma unroll
    for (int j = 0; j < JT; j++) acc2[j] = (f32x4){0.f, 0.f, 0.f, 0.f};

    uint64_t f0, f1;
    if (a.chunk_lo) {
        f0 = a.chunk_lo[chunk];
        f1 = a.chunk_hi[chunk];
    } else {
        f0 = (uint64_t)chunk * a.chunk_frames;
        f1 = f0 + a.chunk_frames;
        if (f1 > a.N) f1 = a.N;
    }

    // Frames are contiguous rows (ldx == D): a 64-frame step is one flat run
    // of 64*D floats.  Element i of the run belongs to frame f = i / D, dim
    // d = i - f*D; the float quotient is exact because i < 2^15 and the
    // fractional part of (i + 0.5)/D is at least 1/(2D).
    const float invD = 1.0f / (float)D;

    float pf[NPF];
    float pl = INFINITY;
#define GMMHIP_STATS_PREFETCH(B_)                                                \
    {                                                                            \
        const float *src = a.X + (B_) * (uint64_t)D;                             \
        const uint64_t nf = (f1 - (B_) < (uint64_t)FB) ? f1 - (B_) : (uint64_t)FB; \
        const int nvalid = (int)nf * D;                                          \
        _Pragma("unroll") for (int j = 0; j < NPF; j++) {                        \
            const int i = tid + 256 * j;                                         \
            pf[j] = (i < nvalid) ? src[i] : 0.0f;                                \
        }                                                                        \
        if (tid < FB) pl = ((B_) + tid < f1) ? a.LL2[(B_) + tid] : INFINITY;     \
    }
    GMMHIP_STATS_PREFETCH(f0)

    for (uint64_t base = f0; base < f1; base += FB) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NPF; j++) {
            const int i = tid + 256 * j;
            if (i < FD) {
                const int f = (int)(((float)i + 0.5f) * invD);
                const int d = i - f * D;
                float v = pf[j] - offL[d];
                // Non-finite or huge values belong to frames whose
                // log-likelihood was clamped (posterior 0); zero them so that
                // 0 * NaN cannot reach the contraction over frames.  Frames
                // past the chunk end were loaded as 0 and stay finite.
                v = (fabsf(v) <= 1e18f) ? v : 0.0f;
                Zr[f * S + 1 + d] = v;
                Zr[f * S + 1 + D + d] = v * v;
            }
        }
        if (tid < FB) {
            const float one = (base + tid < f1) ? 1.0f : 0.0f;
            Zr[tid * S] = one;
            Zr[tid * S + 2 * D + 1] = one;  // low part of the model constant
            Ls[tid] = pl;
        }
        __syncthreads();
        if (base + FB < f1) GMMHIP_STATS_PREFETCH(base + FB)
        if (!active) continue;

        f32x4 acc1[4];
#pragma unroll
        for (int s = 0; s < 4; s++) acc1[s] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int u = 0; u < KU; u++) {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const float4 z =
                    *reinterpret_cast<const float4 *>(&Zr[(16 * s + c) * S + q * KK + 4 * u]);
                acc1[s] = mfma4(z.x, bf[4 * u + 0], acc1[s]);
                acc1[s] = mfma4(z.y, bf[4 * u + 1], acc1[s]);
                acc1[s] = mfma4(z.z, bf[4 * u + 2], acc1[s]);
                acc1[s] = mfma4(z.w, bf[4 * u + 3], acc1[s]);
            }
        }

#pragma unroll
        for (int s = 0; s < 4; s++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                // g = 2^(l2 - LL2); LL2 = +inf (no statistics / padding
                // frame) and l2 = -1e30 (padding component) give 0, a NaN
                // frame fails the cut.
                const int row = 16 * s + 4 * q + r;
                float g = fast_exp2(acc1[s][r] - Ls[row]);
                g = (g > a.min_gamma) ? g : 0.0f;
                // a 4-frame x 16-component block with every posterior under
                // the cut adds exact zeros: skip its products
                if (__builtin_amdgcn_ballot_w64(g != 0.0f) == 0) continue;
                const float *zr = &Zr[row * S];
#pragma unroll
                for (int j = 0; j < JT; j++) acc2[j] = mfma4(g, zr[stats_col<KP>(c, j)], acc2[j]);
            }
        }
    }
#undef GMMHIP_STATS_PREFETCH

    if (active) {
        const size_t Mp = (size_t)a.Mt * 16;
        float *out = a.partials + ((size_t)chunk * Mp + (size_t)tile * 16) * KP;
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int j = 0; j < JT; j++)
                out[(size_t)(4 * q + r) * KP + stats_col<KP>(c, j)] = acc2[j][r];
    }
}

template <int KP>
__device__ __forceinline__ void stats_init_shared(const EstepArgs &a, StatsShared<KP> &sh) {
    for (int i = threadIdx.x; i < ST_FRAMES_PER_ITER * lds_stride(KP); i += 256) sh.Zr[i] = 0.0f;
    for (int i = threadIdx.x; i < a.D; i += 256) sh.offL[i] = a.off[i];
}

// stats_kernel: block = (frame chunk, group of 64 components); wave w owns
// component tile 4*group + w and its 16 x KP statistics accumulators.  Blocks
// are numbered so that the groups of one chunk land on the same XCD (the
// dispatcher deals blocks round-robin over 8 XCDs), so a chunk's frames are
// fetched from HBM once per XCD L2.  The next 64 frames are prefetched into
// registers while the current ones are contracted.
//
// Statistics column mapping: MFMA jt of the g^T Z contraction covers column
// stats_col(c, jt) of lane c = lane%16.  Columns are [1, x', x'^2, 0...]:
// n, F', S'.
template <int KP>
__global__ GMMHIP_STATS_BOUNDS void stats_kernel(EstepArgs a) {
    __shared__ StatsShared<KP> sh;
    const int b = blockIdx.x;
    const int xcd = b & 7, idx = b >> 3;
    const int group = idx % a.ngroups;
    const int chunk = (idx / a.ngroups) * 8 + xcd;
    if (chunk >= a.nchunks) return;
    stats_init_shared<KP>(a, sh);
    stats_item<KP>(a, sh, chunk, group);
}

// Persistent form: a fixed grid pulls (chunk, group) items, chunk-major in
// the order of the chunk table, from a device counter until they run out;
// every block reaches the exit test after each item.
template <int KP>
__global__ GMMHIP_STATS_BOUNDS void stats_queue_kernel(EstepArgs a) {
    __shared__ StatsShared<KP> sh;
    stats_init_shared<KP>(a, sh);
    const int nitems = a.nchunks * a.ngroups;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sh.item = (int)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int item = sh.item;
        if (item >= nitems) break;
        stats_item<KP>(a, sh, item / a.ngroups, item % a.ngroups);
    }
}

// ---------------------------------------------------------------------------
// Fixed-order FP64 sum of the chunk partials: out[e] = sum_c partials[c][e].
__global__ __launch_bounds__(256) void reduce_kernel(const float *__restrict__ part,
                                                     int nchunks, size_t E,
                                                     double *__restrict__ out) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= E) return;
    double s = 0.0;
    int c = 0;
    // eight chunks' loads in flight, added in chunk order
    for (; c + 8 <= nchunks; c += 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = part[(size_t)(c + j) * E + e];
#pragma unroll
        for (int j = 0; j < 8; j++) s += (double)v[j];
    }
    for (; c < nchunks; c++) s += (double)part[(size_t)c * E + e];
    out[e] = s;
}

// Fixed-order FP64 sum of the per-block log-likelihood sums (one block).
__global__ __launch_bounds__(256) void sum_blockll_kernel(const double *__restrict__ v,
                                                          uint64_t n,
                                                          double *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    uint64_t i = threadIdx.x;
    // eight loads in flight, added in index order
    for (; i + 7 * 256 < n; i += 8 * 256) {
        double t[8];
#pragma unroll
        for (int j = 0; j < 8; j++) t[j] = v[i + 256 * j];
#pragma unroll
        for (int j = 0; j < 8; j++) s += t[j];
    }
    for (; i < n; i += 256) s += v[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// ---------------------------------------------------------------------------
#define GMMHIP_KP_CASES(X) \
    X(16) X(32) X(48) X(64) X(80) X(96) X(112) X(128) X(144) X(160) X(176) \
    X(192) X(208) X(224) X(240) X(256)

// Frame tiles of 256 (SUB = 4) keep 4 KP/4 frame values per lane in
// registers; they are built for KP <= 128 only.
hipError_t launch_loglik_fp32(const EstepArgs &a, hipStream_t s) {
    const int sub = ll_sub_of(a);
    if (sub == 4 && a.KP > 128) return hipErrorInvalidValue;
    const dim3 grid((unsigned)ll_blocks(a.N, sub));
    switch (a.KP) {
#define X(K)                                                                           \
    case K:                                                                            \
        if (sub == 1)                                                                  \
            hipLaunchKernelGGL((loglik_kernel<K, 1>), grid, dim3(256), 0, s, a);       \
        else if (sub == 2)                                                             \
            hipLaunchKernelGGL((loglik_kernel<K, 2>), grid, dim3(256), 0, s, a);       \
        else                                                                           \
            hipLaunchKernelGGL((loglik_kernel<(K <= 128 ? K : 128), 4>), grid, dim3(256), 0, s, a); \
        break;
        GMMHIP_KP_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stats_fp32(const EstepArgs &a, hipStream_t s) {
    const int nchunks_pad = (a.nchunks + 7) / 8 * 8;
    const dim3 grid((unsigned)(nchunks_pad * a.ngroups));
    switch (a.KP) {
#define X(K) \
    case K: hipLaunchKernelGGL(stats_kernel<K>, grid, dim3(256), 0, s, a); break;
        GMMHIP_KP_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stats_fp32_queue(const EstepArgs &a, hipStream_t s) {
    const dim3 grid((unsigned)a.queue_blocks);
    switch (a.KP) {
#define X(K) \
    case K: hipLaunchKernelGGL(stats_queue_kernel<K>, grid, dim3(256), 0, s, a); break;
        GMMHIP_KP_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reduce_stats(const EstepArgs &a, hipStream_t s) {
    const size_t E = (size_t)a.Mt * 16 * a.KP;
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, s,
                       a.partials, a.nchunks, E, a.red);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const float *part, int nchunks, size_t E, double *out,
                                  hipStream_t s) {
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, s, part,
                       nchunks, E, out);
    return hipGetLastError();
}

hipError_t launch_sum_doubles(const double *v, uint64_t n, double *out, hipStream_t s) {
    hipLaunchKernelGGL(sum_blockll_kernel, dim3(1), dim3(256), 0, s, v, n, out);
    return hipGetLastError();
}

hipError_t launch_sum_blockll(const EstepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(sum_blockll_kernel, dim3(1), dim3(256), 0, s, a.blockLL,
                       ll_blocks(a.N, ll_sub_of(a)), a.tot);
    return hipGetLastError();
}

}  // namespace gmmhip
