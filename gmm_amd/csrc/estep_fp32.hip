// FP32 E-step kernels for gfx950 (plain form).
//
// The per-frame-by-component log-likelihood is the dense contraction
//   l2[t,m] = sum_k Z[t,k] * B[k,m],  Z[t,:] = [1, x', x'^2, 0...],
//   B[:,m]  = log2(e) * [c_m, mu'_m * ivar_m, -ivar_m / 2, 0...]
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

// Row k of the frame feature vector [1, x', x'^2, 0...] for one frame.
__device__ __forceinline__ float zval(const float *__restrict__ xr,
                                      const float *__restrict__ off, int D,
                                      int k) {
    if (k == 0) return 1.0f;
    if (k <= D) return xr[k - 1] - off[k - 1];
    if (k <= 2 * D) {
        float v = xr[k - 1 - D] - off[k - 1 - D];
        return v * v;
    }
    return 0.0f;
}

// ---------------------------------------------------------------------------
// loglik_kernel: one block = 4 waves x 2 sub-tiles x 16 frames.  Each wave
// keeps the MFMA A-fragments of its 32 frames in registers for the whole
// component loop; component tiles of 16 are staged through LDS and shared by
// the 4 waves.  Output layout of the 16x16x4 MFMA: lane (q = lane/16,
// c = lane%16) holds rows 4q..4q+3 (frames) of column c (component).
template <int KP>
__global__ __launch_bounds__(256) void loglik_kernel(EstepArgs a) {
    constexpr int KK = KP / 4;
    constexpr int SUB = 2;
    __shared__ float Pl[KK * 64];
    __shared__ double red[16];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, c = lane & 15;
    const uint64_t base = (uint64_t)blockIdx.x * LL_FRAMES_PER_BLOCK;
    const int D = a.D;

    float af[SUB][KK];
#pragma unroll
    for (int s = 0; s < SUB; s++) {
        const uint64_t fr = base + (uint64_t)(wave * SUB + s) * 16 + c;
        const bool ok = fr < a.N;
        const float *xr = a.X + (ok ? fr : 0) * a.ldx;
#pragma unroll
        for (int kk = 0; kk < KK; kk++)
            af[s][kk] = ok ? zval(xr, a.off, D, 4 * kk + q) : 0.0f;
    }

    float mx[SUB][4], sm[SUB][4];
    bool bad[SUB][4];
#pragma unroll
    for (int s = 0; s < SUB; s++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            mx[s][r] = -INFINITY;
            sm[s][r] = 0.0f;
            bad[s][r] = false;
        }

    const float4 *P4 = reinterpret_cast<const float4 *>(a.P);
    float4 *Pl4 = reinterpret_cast<float4 *>(Pl);
    for (int t = 0; t < a.Mt; t++) {
        __syncthreads();
        for (int i = tid; i < KK * 16; i += 256)
            Pl4[i] = P4[(size_t)t * KK * 16 + i];
        __syncthreads();

        f32x4 acc[SUB];
#pragma unroll
        for (int s = 0; s < SUB; s++) acc[s] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < KK; kk++) {
            const float b = Pl[kk * 64 + lane];
#pragma unroll
            for (int s = 0; s < SUB; s++) acc[s] = mfma4(af[s][kk], b, acc[s]);
        }

        const bool cvalid = (t * 16 + c) < a.M;
#pragma unroll
        for (int s = 0; s < SUB; s++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float v = cvalid ? acc[s][r] : -INFINITY;
                bad[s][r] = bad[s][r] || (v != v);
                const float hi = fmaxf(mx[s][r], v);
                const float lo = fminf(mx[s][r], v);
                const float e = (hi == -INFINITY) ? 0.0f : fast_exp2(lo - hi);
                sm[s][r] = (v > mx[s][r]) ? fmaf(sm[s][r], e, 1.0f) : sm[s][r] + e;
                mx[s][r] = hi;
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
                const int b2 = __shfl_xor((int)bad[s][r], o);
                const float hi = fmaxf(mx[s][r], m2);
                float v = 0.0f;
                if (hi != -INFINITY)
                    v = sm[s][r] * fast_exp2(mx[s][r] - hi) + s2 * fast_exp2(m2 - hi);
                sm[s][r] = v;
                mx[s][r] = hi;
                bad[s][r] = bad[s][r] || (b2 != 0);
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
                if (bad[s][r] || !(nat > a.min_ll) || isinf(nat)) {
                    nat = a.min_ll;
                    l2 = -INFINITY;
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
// fetched from HBM once per XCD L2.
template <int KP>
__global__ __launch_bounds__(256) void stats_kernel(EstepArgs a) {
    constexpr int KK = KP / 4;
    constexpr int JT = KP / 16;
    constexpr int S = lds_stride(KP);
    constexpr int FB = ST_FRAMES_PER_ITER;
    __shared__ float Zr[FB * S];
    __shared__ float Ls[FB];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, c = lane & 15;

    const int b = blockIdx.x;
    const int xcd = b & 7, idx = b >> 3;
    const int group = idx % a.ngroups;
    const int chunk = (idx / a.ngroups) * 8 + xcd;
    if (chunk >= a.nchunks) return;

    const int tile = group * 4 + wave;
    const bool active = tile < a.Mt;
    const int D = a.D;

    float bf[KK];
#pragma unroll
    for (int kk = 0; kk < KK; kk++)
        bf[kk] = active ? a.P[((size_t)tile * KK + kk) * 64 + lane] : 0.0f;

    f32x4 acc2[JT];
#pragma unroll
    for (int j = 0; j < JT; j++) acc2[j] = (f32x4){0.f, 0.f, 0.f, 0.f};

    for (int i = tid; i < FB * S; i += 256) Zr[i] = 0.0f;

    const uint64_t f0 = (uint64_t)chunk * a.chunk_frames;
    uint64_t f1 = f0 + a.chunk_frames;
    if (f1 > a.N) f1 = a.N;
    const bool cvalid = (tile * 16 + c) < a.M;

    for (uint64_t base = f0; base < f1; base += FB) {
        __syncthreads();
        for (int i = tid; i < FB * D; i += 256) {
            const int f = i / D, d = i - f * D;
            const uint64_t fr = base + f;
            const float v = (fr < f1) ? a.X[fr * a.ldx + d] - a.off[d] : 0.0f;
            Zr[f * S + 1 + d] = v;
            Zr[f * S + 1 + D + d] = v * v;
        }
        if (tid < FB) {
            const uint64_t fr = base + tid;
            const bool ok = fr < f1;
            Zr[tid * S] = ok ? 1.0f : 0.0f;
            Ls[tid] = ok ? a.LL2[fr] : -INFINITY;
        }
        __syncthreads();
        if (!active) continue;

        f32x4 acc1[4];
#pragma unroll
        for (int s = 0; s < 4; s++) acc1[s] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < KK; kk++) {
#pragma unroll
            for (int s = 0; s < 4; s++)
                acc1[s] = mfma4(Zr[(16 * s + c) * S + 4 * kk + q], bf[kk], acc1[s]);
        }

#pragma unroll
        for (int s = 0; s < 4; s++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * s + 4 * q + r;
                const float L = Ls[row];
                float g = 0.0f;
                if (cvalid && L != -INFINITY) g = fast_exp2(acc1[s][r] - L);
                g = (g > a.min_gamma) ? g : 0.0f;
#pragma unroll
                for (int j = 0; j < JT; j++)
                    acc2[j] = mfma4(g, Zr[row * S + 16 * j + c], acc2[j]);
            }
        }
    }

    if (!active) return;
    const size_t Mp = (size_t)a.Mt * 16;
    float *out = a.partials + ((size_t)chunk * Mp + (size_t)tile * 16) * KP;
#pragma unroll
    for (int j = 0; j < JT; j++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            out[(size_t)(4 * q + r) * KP + 16 * j + c] = acc2[j][r];
}

// ---------------------------------------------------------------------------
// Fixed-order FP64 sum of the chunk partials: out[e] = sum_c partials[c][e].
__global__ __launch_bounds__(256) void reduce_kernel(const float *__restrict__ part,
                                                     int nchunks, size_t E,
                                                     double *__restrict__ out) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= E) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; c++) s += (double)part[(size_t)c * E + e];
    out[e] = s;
}

// Fixed-order FP64 sum of the per-block log-likelihood sums (one block).
__global__ __launch_bounds__(256) void sum_blockll_kernel(const double *__restrict__ v,
                                                          uint64_t n,
                                                          double *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += 256) s += v[i];
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

hipError_t launch_loglik_fp32(const EstepArgs &a, hipStream_t s) {
    const dim3 grid((unsigned)ll_blocks(a.N));
    switch (a.KP) {
#define X(K) \
    case K: hipLaunchKernelGGL(loglik_kernel<K>, grid, dim3(256), 0, s, a); break;
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

hipError_t launch_reduce_stats(const EstepArgs &a, hipStream_t s) {
    const size_t E = (size_t)a.Mt * 16 * a.KP;
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, s,
                       a.partials, a.nchunks, E, a.red);
    return hipGetLastError();
}

hipError_t launch_sum_blockll(const EstepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(sum_blockll_kernel, dim3(1), dim3(256), 0, s, a.blockLL,
                       ll_blocks(a.N), a.tot);
    return hipGetLastError();
}

}  // namespace gmmhip
