// Mixed-precision E-step kernels for gfx950.
//
// Same contractions as the FP32 form (estep_fp32.hip), evaluated on
// v_mfma_f32_32x32x16_{f16,bf16} with split operands: every FP32 value v is
// carried as hi = T(v), lo = T(v - hi) and a product is formed from three
// MFMAs, hi*hi + hi*lo + lo*hi, accumulated in FP32 (lo*lo is below the
// representation error).  Frames are centred and scaled per dimension with
// their own mean and standard deviation, z = (x - o) * s, which keeps z and
// z^2 inside the 16-bit exponent range and removes the x^2 cancellation of
// offset data; the model coefficients absorb o and s.
//
// Frame features [1, z, z^2, 0...] are split once per frame set into two
// layouts (prep_frames_kernel):
//   ZA[n][k]                   row-major: the log-likelihood contraction
//                              reads 8 consecutive k of one frame per lane;
//   ZB[blk][u][ct][c][h][j]    per 32-frame block: the statistics
//                              contraction over frames reads, for column
//                              32ct+c, the 8 frames 16u + 8(j/4) + 4h + j%4
//                              that lane (h, c) holds after the first
//                              contraction.
// MFMA 32x32x16 lane layout (lane = 32h + c): A[row c][k 8h+j],
// B[k 8h+j][col c], C reg r -> row 8(r/4) + 4h + r%4, col c.
#include "estep.h"
#include "estep_mixed.h"

#include <hip/hip_runtime.h>
#include <math.h>

namespace gmmhip {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// 16-byte piece as a native vector (HIP's uint4 class defeats register
// promotion of the prefetch arrays)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

static constexpr float kLn2m = 0.69314718055994530942f;

template <typename T>
struct Vec8;
template <>
struct Vec8<_Float16> {
    typedef f16x8 type;
};
template <>
struct Vec8<__bf16> {
    typedef bf16x8 type;
};

__device__ __forceinline__ f32x16 mfma32(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float exp2f_fast(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float log2f_fast(float x) { return __builtin_amdgcn_logf(x); }

// ---------------------------------------------------------------------------
// Per-dimension sum and sum of squares in FP64, fixed order.  Thread
// (g, d) of a block walks frames start+g, start+g+G, ... (G = 256 / D).
__global__ __launch_bounds__(256) void colstats_kernel(const float *__restrict__ X, uint64_t N,
                                                       int D, uint64_t per_block,
                                                       double *__restrict__ part) {
    __shared__ double sh[2][256];
    const int t = threadIdx.x;
    const int G = 256 / D;
    const int g = t / D, d = t - g * D;
    double s1 = 0.0, s2 = 0.0;
    const uint64_t start = (uint64_t)blockIdx.x * per_block;
    uint64_t end = start + per_block;
    if (end > N) end = N;
    if (g < G) {
        for (uint64_t n = start + g; n < end; n += G) {
            const double v = X[n * D + d];
            s1 += v;
            s2 += v * v;
        }
    }
    sh[0][t] = s1;
    sh[1][t] = s2;
    __syncthreads();
    if (t < D) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < G; k++) {
            a += sh[0][k * D + t];
            b += sh[1][k * D + t];
        }
        part[(size_t)blockIdx.x * 2 * D + t] = a;
        part[(size_t)blockIdx.x * 2 * D + D + t] = b;
    }
}

// ---------------------------------------------------------------------------
// Split frame features into hi/lo pieces in both layouts.  One thread per
// (frame, k); frames past N (padding to a multiple of 32) are zero rows.
template <typename T>
__global__ __launch_bounds__(256) void prep_frames_kernel(PrepArgs p) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const int KP = p.KP;
    const uint64_t total = p.Npad * (uint64_t)KP;
    if (i >= total) return;
    const uint64_t n = i / KP;
    const int k = (int)(i - n * KP);
    const int D = p.D;
    float v = 0.0f;
    if (n < p.N) {
        if (k == 0) {
            v = 1.0f;
        } else if (k <= 2 * D) {
            const int d = (k <= D) ? k - 1 : k - 1 - D;
            const float x = p.X[n * D + d];
            float z = (x - p.off[d]) * p.scale[d];
            if (!isfinite(x)) {
                // non-finite frame: its log-likelihood is clamped, no statistics
                z = 0.0f;
                p.bad[n] = 1;
            } else if (fabsf(z) > p.zlimit) {
                // outside the 16-bit range: the call runs in FP32 instead
                z = 0.0f;
                p.bad[n] = 2;
                if (k <= D) atomicAdd(p.nrange, 1u);
            }
            v = (k <= D) ? z : z * z;
        }
    }
    const T hi = (T)v;
    const T lo = (T)(v - (float)hi);
    const uint16_t hb = __builtin_bit_cast(uint16_t, hi), lb = __builtin_bit_cast(uint16_t, lo);
    p.za_hi_raw[i] = hb;
    p.za_lo_raw[i] = lb;
    const uint64_t blk = n >> 5;
    const int f = (int)(n & 31);
    const int u = f >> 4, r16 = f & 15;
    const int h = (r16 >> 2) & 1, j = ((r16 >> 3) << 2) | (r16 & 3);
    const int ct = k >> 5, c = k & 31, CT = KP >> 5;
    const uint64_t zb = blk * 32 * (uint64_t)KP + ((((uint64_t)u * CT + ct) * 32 + c) * 2 + h) * 8 + j;
    p.zb_hi_raw[zb] = hb;
    p.zb_lo_raw[zb] = lb;
}

// ---------------------------------------------------------------------------
// Per-frame log-sum-exp: block = 4 waves x 32 frames; a wave keeps the split
// A fragments of its frames in registers and streams 32-component tiles of
// the split model through LDS.  Two accumulators (hi*hi, cross terms) keep
// two independent MFMA chains per wave.
template <int KP, typename T>
__global__ __launch_bounds__(256) void loglik_mixed_kernel(MixedArgs a) {
    typedef typename Vec8<T>::type V8;
    constexpr int KS = KP / 16;
    constexpr int TILE16 = KS * 2 * 64;      // 16-byte pieces per component tile
    constexpr int PF = TILE16 / 256;
    static_assert(TILE16 % 256 == 0, "component tile must split evenly over 256 threads");
    __shared__ u32x4 Pl[TILE16];
    __shared__ double red[8];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, c = lane & 31;
    const uint64_t base = (uint64_t)blockIdx.x * 128;

    V8 ahi[KS], alo[KS];
    {
        const uint64_t fr = base + 32 * wave + c;  // < Npad
        const u32x4 *zh = reinterpret_cast<const u32x4 *>(a.za_hi) + fr * (KP / 8);
        const u32x4 *zl = reinterpret_cast<const u32x4 *>(a.za_lo) + fr * (KP / 8);
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const u32x4 x = zh[2 * ks + h], y = zl[2 * ks + h];
            ahi[ks] = __builtin_bit_cast(V8, x);
            alo[ks] = __builtin_bit_cast(V8, y);
        }
    }

    float mx[16], sm[16];
    const u32x4 *P16 = reinterpret_cast<const u32x4 *>(a.P16);
    u32x4 pf[PF];
#define GMMHIP_PREFETCH_TILE(T_)                                        \
    _Pragma("unroll") for (int j = 0; j < PF; j++) {                    \
        pf[j] = P16[(size_t)(T_) * TILE16 + tid + 256 * j];             \
    }
    GMMHIP_PREFETCH_TILE(0)
    for (int t = 0; t < a.Mt32; t++) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PF; j++) Pl[tid + 256 * j] = pf[j];
        __syncthreads();
        if (t + 1 < a.Mt32) {
            GMMHIP_PREFETCH_TILE(t + 1)
        }

        f32x16 accA = {}, accB = {};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const u32x4 bh4 = Pl[(ks * 2 + 0) * 64 + lane];
            const u32x4 bl4 = Pl[(ks * 2 + 1) * 64 + lane];
            const V8 bh = __builtin_bit_cast(V8, bh4);
            const V8 bl = __builtin_bit_cast(V8, bl4);
            accA = mfma32(ahi[ks], bh, accA);
            accB = mfma32(ahi[ks], bl, accB);
            accB = mfma32(alo[ks], bh, accB);
        }
        // padded components (>= M) are masked per lane
        const bool cvalid = (32 * t + c) < a.M;
        if (t == 0) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                mx[r] = cvalid ? accA[r] + accB[r] : -INFINITY;
                sm[r] = cvalid ? 1.0f : 0.0f;
            }
        } else {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const float v = cvalid ? accA[r] + accB[r] : -INFINITY;
                const float d = v - mx[r];
                if (d > 32.0f) {
                    sm[r] = fmaf(sm[r], exp2f_fast(-d), 1.0f);
                    mx[r] = v;
                } else {
                    sm[r] += exp2f_fast(d);
                }
            }
        }
    }

    // combine the 32 component lanes of each frame row (xor within 32 lanes)
#pragma unroll
    for (int r = 0; r < 16; r++) {
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const float m2 = __shfl_xor(mx[r], o);
            const float s2 = __shfl_xor(sm[r], o);
            const float hi = fmaxf(mx[r], m2);
            // lanes that only saw padded components carry (-inf, 0)
            const float e1 = (mx[r] == -INFINITY) ? 0.0f : sm[r] * exp2f_fast(mx[r] - hi);
            const float e2 = (m2 == -INFINITY) ? 0.0f : s2 * exp2f_fast(m2 - hi);
            sm[r] = e1 + e2;
            mx[r] = hi;
        }
    }
    double local = 0.0;
    if (c == 0) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 8 * (r >> 2) + 4 * h + (r & 3);
            const uint64_t fr = base + 32 * wave + row;
            float l2 = mx[r] + log2f_fast(sm[r]);
            float nat = l2 * kLn2m;
            if (!(nat > a.min_ll) || isinf(nat) || (fr < a.N && a.bad[fr])) {
                nat = a.min_ll;
                l2 = INFINITY;
            }
            if (fr < a.N) {
                a.LL2[fr] = l2;
                if (a.LLnat) a.LLnat[fr] = nat;
                local += (double)nat;
            }
        }
        red[wave * 2 + h] = local;
    }
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int i = 0; i < 8; i++) s += red[i];
        a.blockLL[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Statistics: block = (frame chunk, group of 128 components); wave w owns the
// 32-component tile 4*group + w and its 32 x KP accumulators.  Per 32 frames
// the block stages the ZA rows and the ZB block of hi/lo pieces in LDS
// (register prefetch of the next step), forms posteriors from the first
// contraction, splits them and contracts them with ZB.
template <int KP, typename T>
__global__ __launch_bounds__(256, (KP <= 192 ? 2 : 1)) void stats_mixed_kernel(MixedArgs a) {
    typedef typename Vec8<T>::type V8;
    constexpr int KS = KP / 16;
    constexpr int CT = KP / 32;
    constexpr int ZAS = KP / 8 + 1;                 // ZA row stride in u32x4 (padded)
    constexpr int ZA16 = 32 * (KP / 8);             // u32x4 per hi (or lo) ZA block
    constexpr int ZB16 = 32 * (KP / 8);             // u32x4 per hi (or lo) ZB block
    constexpr int NPF = (2 * ZA16 + 2 * ZB16) / 256;  // u32x4 per thread per step
    __shared__ u32x4 ZAl[2][32 * ZAS];
    __shared__ u32x4 ZBl[2][ZB16];
    __shared__ float Ls[32];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, c = lane & 31;

    const int b = blockIdx.x;
    const int xcd = b & 7, idx = b >> 3;
    const int group = idx % a.ngroups;
    const int chunk = (idx / a.ngroups) * 8 + xcd;
    if (chunk >= a.nchunks) return;
    const int tile = group * 4 + wave;
    const bool active = tile < a.Mt32;
    const bool cvalid = (32 * tile + c) < a.M;

    V8 bhi[KS], blo[KS];
    {
        const u32x4 *P16 = reinterpret_cast<const u32x4 *>(a.P16);
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            u32x4 x = (u32x4){0u, 0u, 0u, 0u}, y = (u32x4){0u, 0u, 0u, 0u};
            if (active) {
                x = P16[((size_t)tile * KS * 2 + ks * 2 + 0) * 64 + lane];
                y = P16[((size_t)tile * KS * 2 + ks * 2 + 1) * 64 + lane];
            }
            bhi[ks] = __builtin_bit_cast(V8, x);
            blo[ks] = __builtin_bit_cast(V8, y);
        }
    }
    f32x16 acc2[CT];
#pragma unroll
    for (int j = 0; j < CT; j++) acc2[j] = (f32x16){};

    uint64_t f0, f1;
    if (a.chunk_lo) {
        f0 = a.chunk_lo[chunk];
        f1 = a.chunk_hi[chunk];
    } else {
        f0 = (uint64_t)chunk * a.chunk_frames;
        f1 = f0 + a.chunk_frames;
        if (f1 > a.N) f1 = a.N;
    }

    const u32x4 *zah = reinterpret_cast<const u32x4 *>(a.za_hi);
    const u32x4 *zal = reinterpret_cast<const u32x4 *>(a.za_lo);
    const u32x4 *zbh = reinterpret_cast<const u32x4 *>(a.zb_hi);
    const u32x4 *zbl = reinterpret_cast<const u32x4 *>(a.zb_lo);
    u32x4 pf[NPF];
    float pl = INFINITY;
    // base is a multiple of 32; the 32 ZA rows and the ZB block of a step are
    // contiguous runs of ZA16 pieces each
#define GMMHIP_PREFETCH_STEP(B_)                                                   \
    _Pragma("unroll") for (int j = 0; j < NPF; j++) {                              \
        const int i = tid + 256 * j;                                               \
        const int part = i / ZA16, off = i - part * ZA16;                          \
        const u32x4 *src = part == 0 ? zah : part == 1 ? zal : part == 2 ? zbh : zbl; \
        pf[j] = src[(B_) * (KP / 8) + off];                                        \
    }                                                                              \
    if (tid < 32) pl = ((B_) + tid < f1) ? a.LL2[(B_) + tid] : INFINITY;
    GMMHIP_PREFETCH_STEP(f0)

    for (uint64_t base = f0; base < f1; base += 32) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NPF; j++) {
            const int i = tid + 256 * j;
            const int part = i / ZA16, off = i - part * ZA16;
            if (part < 2) {
                const int row = off / (KP / 8), col = off - row * (KP / 8);
                ZAl[part][row * ZAS + col] = pf[j];
            } else {
                ZBl[part - 2][off] = pf[j];
            }
        }
        if (tid < 32) Ls[tid] = pl;
        __syncthreads();
        if (base + 32 < f1) {
            GMMHIP_PREFETCH_STEP(base + 32)
        }
        if (!active) continue;

        f32x16 accA = {}, accB = {};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const u32x4 xh = ZAl[0][c * ZAS + 2 * ks + h];
            const u32x4 xl = ZAl[1][c * ZAS + 2 * ks + h];
            const V8 zh = __builtin_bit_cast(V8, xh);
            const V8 zl = __builtin_bit_cast(V8, xl);
            accA = mfma32(zh, bhi[ks], accA);
            accB = mfma32(zh, blo[ks], accB);
            accB = mfma32(zl, bhi[ks], accB);
        }

        // posteriors, split into hi/lo A operands of the frame contraction:
        // register r = 8u + j holds frame 16u + 8(j/4) + 4h + j%4
        V8 ghi[2], glo[2];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 8 * (r >> 2) + 4 * h + (r & 3);
            float g = exp2f_fast(accA[r] + accB[r] - Ls[row]);
            g = (cvalid && g > a.min_gamma) ? g : 0.0f;
            const T gh = (T)g;
            const T gl = (T)(g - (float)gh);
            ghi[r >> 3][r & 7] = gh;
            glo[r >> 3][r & 7] = gl;
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                const u32x4 yh4 = ZBl[0][((u * CT + ct) * 32 + c) * 2 + h];
                const u32x4 yl4 = ZBl[1][((u * CT + ct) * 32 + c) * 2 + h];
                const V8 yh = __builtin_bit_cast(V8, yh4);
                const V8 yl = __builtin_bit_cast(V8, yl4);
                acc2[ct] = mfma32(ghi[u], yh, acc2[ct]);
                acc2[ct] = mfma32(ghi[u], yl, acc2[ct]);
                acc2[ct] = mfma32(glo[u], yh, acc2[ct]);
            }
        }
    }

    if (!active) return;
    const size_t Mp = (size_t)a.Mt32 * 32;
    float *out = a.partials + ((size_t)chunk * Mp + (size_t)tile * 32) * KP;
#pragma unroll
    for (int ct = 0; ct < CT; ct++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 8 * (r >> 2) + 4 * h + (r & 3);
            out[(size_t)row * KP + 32 * ct + c] = acc2[ct][r];
        }
}

// ---------------------------------------------------------------------------
#define GMMHIP_KP32_CASES(X) X(32) X(64) X(96) X(128) X(160) X(192) X(224) X(256)

hipError_t launch_colstats(const float *X, uint64_t N, int D, int blocks, uint64_t per_block,
                           double *part, hipStream_t s) {
    hipLaunchKernelGGL(colstats_kernel, dim3(blocks), dim3(256), 0, s, X, N, D, per_block, part);
    return hipGetLastError();
}

hipError_t launch_prep_frames(const PrepArgs &p, int prec, hipStream_t s) {
    const uint64_t total = p.Npad * (uint64_t)p.KP;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (prec == 2)
        hipLaunchKernelGGL(prep_frames_kernel<_Float16>, grid, dim3(256), 0, s, p);
    else
        hipLaunchKernelGGL(prep_frames_kernel<__bf16>, grid, dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_loglik_mixed(const MixedArgs &a, int prec, hipStream_t s) {
    const dim3 grid((unsigned)((a.Npad + 127) / 128));
    switch (a.KP) {
#define X(K)                                                                                   \
    case K:                                                                                    \
        if (prec == 2)                                                                         \
            hipLaunchKernelGGL((loglik_mixed_kernel<K, _Float16>), grid, dim3(256), 0, s, a); \
        else                                                                                   \
            hipLaunchKernelGGL((loglik_mixed_kernel<K, __bf16>), grid, dim3(256), 0, s, a);   \
        break;
        GMMHIP_KP32_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stats_mixed(const MixedArgs &a, int prec, hipStream_t s) {
    const int nchunks_pad = (a.nchunks + 7) / 8 * 8;
    const dim3 grid((unsigned)(nchunks_pad * a.ngroups));
    switch (a.KP) {
#define X(K)                                                                                  \
    case K:                                                                                   \
        if (prec == 2)                                                                        \
            hipLaunchKernelGGL((stats_mixed_kernel<K, _Float16>), grid, dim3(256), 0, s, a); \
        else                                                                                  \
            hipLaunchKernelGGL((stats_mixed_kernel<K, __bf16>), grid, dim3(256), 0, s, a);   \
        break;
        GMMHIP_KP32_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gmmhip
