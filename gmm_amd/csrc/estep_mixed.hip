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
// Frame features [1, z, z^2, 0...] are split once per frame set
// (prep_frames_kernel) into rows ZA[n][k] of KP columns, KP a power of two
// (zero columns past 2D+1): the log-likelihood contraction reads 8
// consecutive k of one frame per lane (the first 2D+1 rounded up to 32 of
// them), and the statistics kernel takes its frame-contraction operand from
// the same rows by transposing LDS reads.
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
// (g, d) of a block walks frames start+g, start+g+G, ... (G = 256 / D),
// eight frames a round with their eight loads in flight before the adds
// (one load at a time left the kernel latency-bound at 2 TB/s).
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
        uint64_t n = start + g;
        const uint64_t step = (uint64_t)G * D;
        for (; n + 7 * (uint64_t)G < end; n += 8 * (uint64_t)G) {
            const float *xp = X + n * D + d;
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = xp[j * step];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double w = v[j];
                s1 += w;
                s2 += w * w;
            }
        }
        for (; n < end; n += G) {
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
// Split model.  binary16 pieces round once from the FP64 coefficient (round
// to nearest even), bfloat16 pieces from its binary32 value, in integer
// arithmetic (the same bits the 16-bit casts give).
__device__ __forceinline__ uint16_t f64_to_f16_bits(double v) {
    const uint64_t x = __builtin_bit_cast(uint64_t, v);
    const uint32_t sign = (uint32_t)((x >> 48) & 0x8000u);
    const uint64_t ax = x & 0x7fffffffffffffffull;
    // normal range: rebias the exponent (1023 -> 15), round the 52-bit mantissa to 10
    const uint32_t rn = (uint32_t)((ax - (1008ull << 52) + ((1ull << 41) - 1) + ((ax >> 42) & 1)) >> 42);
    // below 2^-14: a multiple of 2^-24 (exact scaling), rounded to even by the 2^52 shift
    const double units = __builtin_bit_cast(double, ax) * 16777216.0;
    const uint32_t rs = (uint32_t)((units + 4503599627370496.0) - 4503599627370496.0);
    uint32_t r = ax < 0x3f10000000000000ull ? rs : rn;
    r = ax >= 0x40effe0000000000ull ? 0x7c00u : r;  // >= 65520 rounds to infinity
    r = ax > 0x7ff0000000000000ull ? 0x7e00u : r;   // NaN
    return (uint16_t)(sign | r);
}
__device__ __forceinline__ float f16_bits_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    const float sub = (float)m * 5.9604644775390625e-08f;  // m * 2^-24
    const uint32_t b = e == 31 ? (sign | 0x7f800000u | (m << 13)) : (sign | ((e + 112u) << 23) | (m << 13));
    return e == 0 ? (sign ? -sub : sub) : __builtin_bit_cast(float, b);
}
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
    const uint32_t x = __builtin_bit_cast(uint32_t, f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

// One wave per component slot: the wave forms the D pairs (omega, nu) and
// the linear and quadratic coefficients, lane 0 the constant (its sum over
// d in order), then the lanes split and store the KP coefficients.  FP64
// arithmetic as written: no contraction into fused multiply-adds.
template <bool FP16>
__global__ __launch_bounds__(64) void pack_split_model_kernel(SplitModelArgs a) {
#pragma clang fp contract(off)
    __shared__ double om[PREP_MAX_D], nu[PREP_MAX_D], coef[256];
    const int m = blockIdx.x, lane = threadIdx.x;
    const int D = a.D, KP = a.KP;
    const bool real = m < a.M;
    for (int d = lane; d < D; d += 64) {
        double o = 0.0, n = 0.0;
        if (real) {
            const double sd = a.scale[d];
            n = sd * ((double)a.mu[(size_t)m * D + d] - (double)a.off[d]);
            o = 1.0 / ((double)a.var[(size_t)m * D + d] * sd * sd);
        }
        om[d] = o;
        nu[d] = n;
        coef[1 + d] = o * n;
        coef[1 + D + d] = -0.5 * o;
    }
    for (int k = 2 * D + 1 + lane; k < KP; k += 64) coef[k] = 0.0;
    __syncthreads();
    if (lane == 0) {
        double cq = 0.0;
        for (int d = 0; d < D; d++) cq += om[d] * nu[d] * nu[d];
        coef[0] = real ? a.gconst[m] - 0.5 * cq : 0.0;
    }
    __syncthreads();
    const double log2e = 1.4426950408889634;
    const int t = m >> 5, c = m & 31, KS = KP / 16;
    bool bad = false;
    for (int k = lane; k < KP; k += 64) {
        const double v = coef[k] * log2e;
        uint16_t hb, lb;
        float hf;
        if constexpr (FP16) {
            hb = f64_to_f16_bits(v);
            hf = f16_bits_to_f32(hb);
            lb = f64_to_f16_bits(v - (double)hf);
        } else {
            hb = f32_to_bf16_bits((float)v);
            hf = __builtin_bit_cast(float, (uint32_t)hb << 16);
            lb = f32_to_bf16_bits((float)(v - (double)hf));
        }
        bad |= !__builtin_isfinite(hf);
        const int ks = k >> 4, hh = (k >> 3) & 1, j = k & 7;
        const size_t ih = ((((size_t)t * KS + ks) * 2 + 0) * 64 + 32 * hh + c) * 8 + j;
        a.out[ih] = hb;
        a.out[ih + 64 * 8] = lb;
    }
    if (bad) atomicOr(a.overflow, 1u);
}

hipError_t launch_pack_split_model(const SplitModelArgs &a, int prec, hipStream_t s) {
    if (a.D < 1 || a.D > PREP_MAX_D || a.KP < 2 * a.D + 1 || a.KP > 256 || a.KP % 32 ||
        a.Mt32 * 32 < a.M)
        return hipErrorInvalidValue;
    if (prec == 2)
        hipLaunchKernelGGL(pack_split_model_kernel<true>, dim3((unsigned)(a.Mt32 * 32)), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(pack_split_model_kernel<false>, dim3((unsigned)(a.Mt32 * 32)), dim3(64), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split frame features into hi/lo pieces.  One block per 32-frame block:
// the block's centred features are staged in LDS once, then every store is
// one 16-byte piece (8 features of a frame) and consecutive threads write
// consecutive pieces.
// Frames past N (padding to a multiple of 32) are zero rows.  bad[n] is
// written for every frame: 1 if the frame holds a non-finite value, else 2
// if a centred value lies outside the 16-bit range, else 0.
template <typename T>
__device__ __forceinline__ unsigned int split_bits(float v, unsigned int &lo) {
    const T hi = (T)v;
    const T l = (T)(v - (float)hi);
    lo = __builtin_bit_cast(uint16_t, l);
    return __builtin_bit_cast(uint16_t, hi);
}

template <typename T, int KP>
__global__ __launch_bounds__(256) void prep_frames_kernel(PrepArgs p) {
    __shared__ float zs[32 * PREP_MAX_D];
    __shared__ unsigned int flags[32];
    __shared__ unsigned int cnt;
    const int tid = threadIdx.x;
    const int D = p.D;
    const uint64_t blk = blockIdx.x;
    const uint64_t n0 = blk * 32;
    const int nf = (n0 >= p.N) ? 0 : (int)((p.N - n0) < 32 ? (p.N - n0) : 32);
    if (tid < 32) flags[tid] = 0u;
    if (tid == 0) cnt = 0u;
    __syncthreads();
    // the block's frames are 32 x D consecutive floats; thread (g, d) takes
    // dimension d of frames g, g + G, ... (G = 256 / D rows per pass, one
    // division per thread instead of one per element)
    const int G = 256 / D;
    const int g = tid / D, d = tid - g * D;
    if (g < G) {
        const float o = p.off[d], sc = p.scale[d];
        for (int f = g; f < nf; f += G) {
            const float x = p.X[(n0 + f) * D + d];
            float z = (x - o) * sc;
            if (!isfinite(x)) {
                // non-finite frame: its log-likelihood is clamped, no statistics
                z = 0.0f;
                atomicOr(&flags[f], 1u);
            } else if (fabsf(z) > p.zlimit) {
                // outside the 16-bit range: the call runs in FP32 instead
                z = 0.0f;
                atomicOr(&flags[f], 2u);
                atomicAdd(&cnt, 1u);
            }
            zs[f * D + d] = z;
        }
    }
    __syncthreads();
    if (tid < nf) {
        const unsigned int fl = flags[tid];
        p.bad[n0 + tid] = (fl & 1u) ? 1 : (fl ? 2 : 0);
    }
    if (tid == 0 && cnt) atomicAdd(p.nrange, cnt);

    // feature k of frame f: 1, z_1..z_D, z_1^2..z_D^2, zero padding
    auto val = [&](int f, int k) -> float {
        if (f >= nf) return 0.0f;
        if (k == 0) return 1.0f;
        if (k <= D) return zs[f * D + k - 1];
        if (k <= 2 * D) {
            const float z = zs[f * D + k - 1 - D];
            return z * z;
        }
        return 0.0f;
    };
    constexpr int U = KP / 8;
    constexpr int pieces = 32 * U;  // per part
    u32x4 *zah = reinterpret_cast<u32x4 *>(p.za_hi_raw) + blk * pieces;
    u32x4 *zal = reinterpret_cast<u32x4 *>(p.za_lo_raw) + blk * pieces;
    // ZA[n][k]: piece q of frame f holds k = 8q .. 8q+7
    for (int e = tid; e < pieces; e += 256) {
        const int f = e / U, k0 = (e - f * U) * 8;
        u32x4 H, L;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            unsigned int l0, l1;
            const unsigned int h0 = split_bits<T>(val(f, k0 + 2 * i), l0);
            const unsigned int h1 = split_bits<T>(val(f, k0 + 2 * i + 1), l1);
            H[i] = h0 | (h1 << 16);
            L[i] = l0 | (l1 << 16);
        }
        zah[e] = H;
        zal[e] = L;
    }
}

// ---------------------------------------------------------------------------
// Two FP32 values split into packed 16-bit hi and lo words (hi = T(v),
// lo = T(v - hi), round to nearest), as the A operand words of the frame
// contraction.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <typename T>
struct Vec2;
template <>
struct Vec2<_Float16> {
    typedef f16x2 type;
};
template <>
struct Vec2<__bf16> {
    typedef bf16x2 type;
};
template <typename T>
__device__ __forceinline__ void split_pair(float a, float b, unsigned int &hi, unsigned int &lo) {
    typedef typename Vec2<T>::type T2;
    const T2 h = __builtin_convertvector((f32x2){a, b}, T2);
    f32x2 r;
    if constexpr (sizeof(T) == 2 && __is_same(T, _Float16)) {
        // a - hi (exact) in one mixed-precision FMA per value, reading the
        // half straight from the packed pair: -hi * 1.0 + a
        const unsigned int hp = __builtin_bit_cast(unsigned int, h);
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r.x) : "v"(hp), "v"(a));
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
            : "=v"(r.y) : "v"(hp), "v"(b));
    } else {
        const f32x2 hb = __builtin_convertvector(h, f32x2);
        r = (f32x2){a - hb.x, b - hb.y};
    }
    const T2 l = __builtin_convertvector(r, T2);
    hi = __builtin_bit_cast(unsigned int, h);
    lo = __builtin_bit_cast(unsigned int, l);
}

// ---------------------------------------------------------------------------
// Direct-to-LDS staging (gfx950 global_load_lds_dwordx4: lane L of a wave
// writes its 16 bytes at the wave-uniform LDS base + 16 L), for depths whose
// feature row is a power-of-two number of 16-byte pieces (KP = 32, 64, 128,
// 256).  The copies of a later step run while the current one is
// contracted, with one barrier per step and no staging registers.
//
// ZA rows are stored swizzled: piece w of frame row c sits at slot
// c*U + (w ^ za_swz(c)).  za_swz is a bijection on any 16 consecutive rows
// (of the rows that share a 256-byte bank window), so the 16 lanes of a
// 16-byte read of piece w of consecutive rows hit 16 distinct bank groups;
// and for the four rows of an aligned block it differs in the bits above the
// lowest two piece bits (U >= 8), so a transposing read of four pieces
// 4i .. 4i+3 of each of four rows is conflict-free as well.
template <int U>
__host__ __device__ constexpr int za_swz(int c) {
    if constexpr (U >= 16) return ((c & 3) << 2) | ((c >> 2) & 3);
    else if constexpr (U == 8) return (((c >> 1) & 1) << 2) | ((c >> 2) & 3);
    else return (c >> 2) & 3;
}

__device__ __forceinline__ void lds_dma16(const void *src, void *lds_wave_base) {
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)lds_wave_base,
                                     16, 0, 0);
}

// The same copy through a buffer resource over [base, base + bytes): the
// base stays in scalar registers and each lane supplies a 32-bit byte
// offset, so a copy costs no 64-bit address arithmetic.  Lane L still lands
// at lds_wave_base + 16 L.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t dma_rsrc(const void *base, unsigned int bytes) {
    // dword 3 of a gfx9 buffer descriptor: 32-bit raw data format
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void lds_dma16(__amdgpu_buffer_rsrc_t r, unsigned int voff,
                                          void *lds_wave_base) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        r, (__attribute__((address_space(3))) void *)lds_wave_base, 16, voff, 0, 0, 0);
}

// Transposing LDS read (gfx950 ds_read_b64_tr_b16).  Within each group of 16
// lanes, lane j supplies the address of an 8-byte piece P_j (four 16-bit
// values); lane i of the group receives P_{4k + i/4}[i % 4] as element k,
// k = 0..3.  With lane j pointing at frame fb + j/4, features b + 4 (j % 4),
// lane i receives feature b + i of frames fb .. fb+3.
typedef short v4s16 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4s16 lds_read_tr16(const void *lds_addr) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16(
        (__attribute__((address_space(3))) v4s16 *)const_cast<void *>(lds_addr));
}

// Next (chunk, group) item of the persistent statistics grid, or -1 when
// none is left.  XCD x's list (counter a.queue[x]) holds the chunks x, x+8,
// x+16, ..., each with all its component groups in a row, so the groups
// that share a chunk's frames run under one L2; a block (on XCD blockIdx % 8)
// drains its own list first, then takes items from the others in turn.
// drained: bit x set once list x was found empty.
__device__ __forceinline__ int mixed_queue_pull(const MixedArgs &a, int xcd,
                                                unsigned int &drained) {
    for (int k = 0; k < 8; k++) {
        const int x = (xcd + k) & 7;
        if (drained & (1u << x)) continue;
        const unsigned int nch = a.nchunks > x ? (unsigned int)(a.nchunks - x + 7) / 8u : 0u;
        const unsigned int n = nch * (unsigned int)a.ngroups;
        const unsigned int j = n ? atomicAdd(a.queue + x, 1u) : n;
        if (j < n) {
            const unsigned int jc = j / (unsigned int)a.ngroups;
            return (int)((x + 8 * jc) * (unsigned int)a.ngroups + (j - jc * (unsigned int)a.ngroups));
        }
        drained |= 1u << x;
    }
    return -1;
}

// Statistics with direct-to-LDS staging.  A block of NWS waves (4 or 8: the
// component tile, 128 or 256 components per block) owns NWS component tiles
// (32 components each) and walks its frame chunk in 32-frame steps; per step
// the waves copy the step's ZA hi and lo rows (U pieces x 32 rows each, the
// first half of the waves hi, the second lo) and the last wave also the
// step's 32 per-frame log-likelihoods.  Eight waves share one copy of the
// rows between twice the components (half the copies per MFMA) at one block
// per CU (depths 64 and 128; at 256 the two waves per SIMD would spill); the
// tuning table picks per (D, M).  Iteration s contracts ZA of step s+1 with the
// split model (the first contraction, interleaved with step s's posteriors)
// and the posteriors of step s with the features of step s (the frame
// contraction), whose operand -- 8 frames of one feature per lane -- comes
// out of the ZA rows of step s through transposing LDS reads.  So ZA is
// copied two steps ahead into three rotating slots.  Each slot is a
// __shared__ variable of its own and the loop is unrolled by three so that
// every access names its slot: the compiler can then see that the copies in
// flight do not alias the reads of the iteration and does not wait for them.
//
// QUEUE: persistent form; the blocks pull (chunk, group) items through
// mixed_queue_pull until every list is empty, with a barrier between items
// (the next item's copies overwrite slots the last one read).
template <int KP, typename T, int NWS, bool QUEUE>
__global__ __launch_bounds__(64 * NWS, (NWS == 4 && KP <= 128 ? 2 : 1)) void stats_mixed_tr_kernel(MixedArgs a) {
    typedef typename Vec8<T>::type V8;
    constexpr int KS = KP / 16;
    constexpr int CT = KP / 32;
    constexpr int U = KP / 8;               // 16-byte pieces per feature row
    constexpr int PART = 32 * U;            // pieces per part (hi or lo) and step
    constexpr int HW = NWS / 2;             // waves per part
    constexpr int PW = PART / HW;           // pieces a wave copies per step
    constexpr int NDMA = PW / 64;           // copies per wave and step
    static_assert((U & (U - 1)) == 0 && PW % 64 == 0 && NDMA >= 1, "power-of-two feature rows");
    __shared__ u32x4 ZA0[2][PART], ZA1[2][PART], ZA2[2][PART];
    __shared__ float Ls0[32], Ls1[32], Ls2[32];

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    __shared__ int sh_item;
    unsigned int drained = 0;  // (thread 0) item lists found empty
    for (;;) {
    // lane-derived offsets formed anew per item: hoisted out of the item
    // loop they would stay live beside the body's 255 registers and spill
    int lane = tid & 63;
    if constexpr (QUEUE) asm volatile("" : "+v"(lane));
    const int h = lane >> 5, c = lane & 31;
    int chunk, group;
    if constexpr (QUEUE) {
        __syncthreads();
        if (tid == 0) sh_item = mixed_queue_pull(a, (int)(blockIdx.x & 7), drained);
        __syncthreads();
        const int item = sh_item;
        if (item < 0) break;
        chunk = item / a.ngroups;
        group = item - chunk * a.ngroups;
    } else {
        const int b = blockIdx.x;
        const int xcd = b & 7, idx = b >> 3;
        group = idx % a.ngroups;
        chunk = (idx / a.ngroups) * 8 + xcd;
        if (chunk >= a.nchunks) return;
    }
    const int tile = group * NWS + wave;
    const bool active = tile < a.Mt32;
    const bool cvalid = (32 * tile + c) < a.M;
    const float thr = cvalid ? a.min_gamma : INFINITY;

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

    // waves 0 .. HW-1 copy ZA hi, waves HW .. NWS-1 ZA lo, PW pieces each;
    // the last wave also the step's 32 per-frame log-likelihoods.  Copy j of
    // a wave moves piece p = q PW + 64 j + lane of its part (row p / U,
    // de-swizzled source).
    const int hl = wave / HW, q = wave % HW;
    const bool ll_wave = wave == NWS - 1;
    const u32x4 *src = reinterpret_cast<const u32x4 *>(hl ? a.za_lo : a.za_hi);
    unsigned int soff[NDMA];
#pragma unroll
    for (int j = 0; j < NDMA; j++) {
        const int p = q * PW + j * 64 + lane;
        const int row = p / U, v = p % U;
        soff[j] = (unsigned int)(row * U + (v ^ za_swz<U>(row))) * 16u;
    }
    // transposing reads of the frame contraction: this lane's two 8-byte
    // pieces per (u, ct) are frames 16u + 8r + 4h + j/4, r = 0, 1, features
    // 32 ct + 16 (lane / 16 % 2) + 4 (j % 4), j = lane % 16, at byte offset
    // (fr U + (w ^ za_swz(fr))) 16 + 8 (j % 2) of a slot's hi (or lo) array,
    // fr the frame row, w = fr's feature / 8.  The swizzle has period 16 in
    // fr, so u only adds 16 rows.
    static_assert(za_swz<U>(16) == za_swz<U>(0) && za_swz<U>(31) == za_swz<U>(15),
                  "row swizzle of period 16");
    unsigned int troff[2][CT];
    {
        const int g = lane >> 4, j = lane & 15;
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                const int fr = 8 * r + 4 * h + (j >> 2);
                const int feat = 32 * ct + 16 * (g & 1) + 4 * (j & 3);
                const int w = feat >> 3;
                troff[r][ct] = (unsigned int)((fr * U + (w ^ za_swz<U>(fr))) * 16 + (feat & 4) * 2);
            }
    }
    const unsigned int ulane = (unsigned int)lane;
    const int nsteps = (int)((f1 - f0 + 31) / 32);
#define GMMHIP_TR_DMA(K_, ZD_, LD_)                                                \
    {                                                                              \
        const __amdgpu_buffer_rsrc_t rs =                                          \
            dma_rsrc(src + (f0 + 32 * (uint64_t)(K_)) * U, PART * 16);             \
        _Pragma("unroll") for (int j = 0; j < NDMA; j++)                           \
            lds_dma16(rs, soff[j], &ZD_[hl][q * PW + j * 64]);             \
        if (ll_wave && lane < 32)                                                  \
            __builtin_amdgcn_global_load_lds(                                      \
                a.LL2 + f0 + 32 * (uint64_t)(K_) + ulane,                          \
                (__attribute__((address_space(3))) void *)&LD_[0], 4, 0, 0);       \
    }
#define GMMHIP_TR_CONTRACT1(ACC_, ZR_)                                             \
    _Pragma("unroll") for (int ks = 0; ks < KS; ks++) {                            \
        const int slot = c * U + ((2 * ks + h) ^ za_swz<U>(c));                    \
        const V8 zh = __builtin_bit_cast(V8, ZR_[0][slot]);                        \
        const V8 zl = __builtin_bit_cast(V8, ZR_[1][slot]);                        \
        ACC_ = mfma32(zh, bhi[ks], ACC_);                                          \
        ACC_ = mfma32(zh, blo[ks], ACC_);                                          \
        ACC_ = mfma32(zl, bhi[ks], ACC_);                                          \
    }
    // prologue: ZA of steps 0 and 1, LL2 of step 0
    GMMHIP_TR_DMA(0, ZA0, Ls0)
    if (nsteps > 1) {
        const __amdgpu_buffer_rsrc_t rs = dma_rsrc(src + (f0 + 32) * U, PART * 16);
#pragma unroll
        for (int j = 0; j < NDMA; j++) lds_dma16(rs, soff[j], &ZA1[hl][q * PW + j * 64]);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    f32x16 acc = {};
    if (active) {
        GMMHIP_TR_CONTRACT1(acc, ZA0)
    }

    // iteration s: slot C holds ZA / LL2 of step s (frame contraction,
    // posteriors), slot N ZA of step s+1 (first contraction) and receives
    // LL2 of step s+1, slot F receives ZA of step s+2
#define GMMHIP_TR_STEP(S_, ZC, ZN, ZF, LC, LN, ACC_, NACC_)                        \
    {                                                                              \
        const int s_ = (S_);                                                       \
        __builtin_amdgcn_s_waitcnt(0);                                             \
        if (ll_wave && s_ == nsteps - 1 && lane < 32 && f0 + 32 * (uint64_t)s_ + lane >= f1) \
            LC[lane] = INFINITY; /* frames past the chunk end: posterior 0 */      \
        __syncthreads();                                                           \
        if (s_ + 2 < nsteps) {                                                     \
            const __amdgpu_buffer_rsrc_t rs =                                      \
                dma_rsrc(src + (f0 + 32 * (uint64_t)(s_ + 2)) * U, PART * 16);     \
            _Pragma("unroll") for (int j = 0; j < NDMA; j++)                       \
                lds_dma16(rs, soff[j], &ZF[hl][q * PW + j * 64]);          \
        }                                                                          \
        if (ll_wave && lane < 32 && s_ + 1 < nsteps)                               \
            __builtin_amdgcn_global_load_lds(                                      \
                a.LL2 + f0 + 32 * (uint64_t)(s_ + 1) + ulane,                      \
                (__attribute__((address_space(3))) void *)&LN[0], 4, 0, 0);        \
        if (active) {                                                              \
            NACC_ = (f32x16){};                                                    \
            GMMHIP_TR_CONTRACT1(NACC_, ZN)                                         \
            u32x4 gh4[2], gl4[2];                                                  \
            _Pragma("unroll") for (int r = 0; r < 16; r += 2) {                    \
                const int row = 8 * (r >> 2) + 4 * h + (r & 3);                    \
                float g0 = exp2f_fast(ACC_[r] - LC[row]);                          \
                float g1 = exp2f_fast(ACC_[r + 1] - LC[row + 1]);                  \
                g0 = (g0 > thr) ? g0 : 0.0f;                                       \
                g1 = (g1 > thr) ? g1 : 0.0f;                                       \
                unsigned int wh, wl;                                               \
                split_pair<T>(g0, g1, wh, wl);                                     \
                gh4[r >> 3][(r & 7) >> 1] = wh;                                    \
                gl4[r >> 3][(r & 7) >> 1] = wl;                                    \
            }                                                                      \
            GMMHIP_TR_SCHED                                                        \
            _Pragma("unroll") for (int u = 0; u < 2; u++) {                        \
                const unsigned int any = gh4[u][0] | gh4[u][1] | gh4[u][2] | gh4[u][3]; \
                if (__builtin_amdgcn_ballot_w64(any != 0u) == 0) continue;         \
                const V8 ghi = __builtin_bit_cast(V8, gh4[u]);                     \
                const V8 glo = __builtin_bit_cast(V8, gl4[u]);                     \
                _Pragma("unroll") for (int ct = 0; ct < CT; ct++) {                \
                    const char *bh = reinterpret_cast<const char *>(&ZC[0][0]);    \
                    const char *bl = reinterpret_cast<const char *>(&ZC[1][0]);    \
                    const char *bhu = bh + u * (16 * U * 16);                      \
                    const char *blu = bl + u * (16 * U * 16);                      \
                    const v4s16 h0 = lds_read_tr16(bhu + troff[0][ct]);            \
                    const v4s16 h1 = lds_read_tr16(bhu + troff[1][ct]);            \
                    const v4s16 l0 = lds_read_tr16(blu + troff[0][ct]);            \
                    const v4s16 l1 = lds_read_tr16(blu + troff[1][ct]);            \
                    const V8 yh = __builtin_bit_cast(                              \
                        V8, __builtin_shufflevector(h0, h1, 0, 1, 2, 3, 4, 5, 6, 7)); \
                    const V8 yl = __builtin_bit_cast(                              \
                        V8, __builtin_shufflevector(l0, l1, 0, 1, 2, 3, 4, 5, 6, 7)); \
                    acc2[ct] = mfma32(ghi, yh, acc2[ct]);                          \
                    acc2[ct] = mfma32(ghi, yl, acc2[ct]);                          \
                    acc2[ct] = mfma32(glo, yh, acc2[ct]);                          \
                }                                                                  \
            }                                                                      \
        }                                                                          \
    }
    f32x16 acc_n;
    int s = 0;
    // the first contraction's LDS reads one step ahead of their MFMAs, the
    // step's posterior references first, and up to 16 VALU instructions of
    // the posteriors after each three MFMAs, so that all of them (about 100
    // per step) issue among the chain's 24 MFMAs (with 10 per group a fifth
    // were left to run after the chain: 41.7 -> 40.3 ms at D=60, M=2048,
    // N=2e7)
#define GMMHIP_TR_SCHED                                                            \
    __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);                             \
    _Pragma("unroll") for (int i = 0; i < KS; i++) {                               \
        if (i + 1 < KS) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);         \
        __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);                         \
        __builtin_amdgcn_sched_group_barrier(0x002, 16, 0);                        \
    }
    for (; s + 2 < nsteps; s += 3) {
        GMMHIP_TR_STEP(s, ZA0, ZA1, ZA2, Ls0, Ls1, acc, acc_n)
        GMMHIP_TR_STEP(s + 1, ZA1, ZA2, ZA0, Ls1, Ls2, acc_n, acc)
        GMMHIP_TR_STEP(s + 2, ZA2, ZA0, ZA1, Ls2, Ls0, acc, acc_n)
        acc = acc_n;
    }
    if (s < nsteps) {
        GMMHIP_TR_STEP(s, ZA0, ZA1, ZA2, Ls0, Ls1, acc, acc_n)
        acc = acc_n;
    }
    if (s + 1 < nsteps) GMMHIP_TR_STEP(s + 1, ZA1, ZA2, ZA0, Ls1, Ls2, acc, acc_n)
#undef GMMHIP_TR_STEP
#undef GMMHIP_TR_SCHED
#undef GMMHIP_TR_DMA
#undef GMMHIP_TR_CONTRACT1

    if (active) {
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
    if constexpr (!QUEUE) break;
    }
}

// Maximum over the 32 lanes of a half-wave (lanes 32h .. 32h+31): DPP
// within each 16-lane row, then one exchange between the two rows.
__device__ __forceinline__ float row_max32(float v) {
    v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                     0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false)));  // quad [1,0,3,2]
    v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                     0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false)));  // quad [2,3,0,1]
    v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                     0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, false))); // row half-mirror
    v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                     0, __builtin_bit_cast(int, v), 0x140, 0xF, 0xF, false))); // row mirror
    return fmaxf(v, __shfl_xor(v, 16));
}

// Headroom of the log-sum-exp reference (log2 units): values up to this
// far above a row's reference are summed without moving it.
static constexpr float LSE_HEADROOM = 64.0f;

// Last tile (padded components >= M masked per lane), then the 32 component
// lanes of each frame row combined (xor within 32 lanes), the per-frame
// log-likelihoods written (clamped as the reference does) and the block's
// FP64 sum.  base = first frame of the block (NW waves x 32 frames).
template <int NW>
__device__ __forceinline__ void ll_finish(const MixedArgs &a, const f32x16 &acc, float (&mx)[16],
                                          float (&sm)[16], uint64_t base, double *red) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, c = lane & 31;
    {
        const bool valid = (32 * (a.Mt32 - 1) + c) < a.M;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            // (a padded lane that saw no component keeps mx = -inf, sm = 0)
            const float d = acc[r] - mx[r];
            const bool up = valid && d > LSE_HEADROOM;
            const float e = valid ? exp2f_fast(up ? -d : d) : 0.0f;
            sm[r] = up ? fmaf(sm[r], e, 1.0f) : sm[r] + e;
            mx[r] = up ? acc[r] : mx[r];
        }
    }
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
        for (int i = 0; i < 2 * NW; i++) s += red[i];
        a.blockLL[blockIdx.x] = s;
    }
}

// Per-frame log-sum-exp: block = 4 waves x 32 frames; a wave keeps the split
// A fragments of its frames in registers and streams 32-component tiles of
// the split model through LDS by direct-to-LDS copies (two LDS buffers, one
// barrier per tile).  hi*hi + hi*lo + lo*hi accumulate into one chain per
// tile; the log-sum-exp of tile t runs on the VALU between tile t+1's MFMAs.
// KP: contraction depth (2D+1 rounded up to 32); KPL: the feature rows' and
// the split model's layout depth (KP rounded up to a power of two, zero
// columns past 2D+1), of which the first KP columns are read.
template <int KP, int KPL, typename T, int NW>
__global__ __launch_bounds__(64 * NW) void loglik_mixed_dma_kernel(MixedArgs a) {
    typedef typename Vec8<T>::type V8;
    static_assert(KP <= KPL && KPL % 32 == 0, "contraction within the layout");
    constexpr int KS = KP / 16;
    constexpr int TILE16 = KS * 2 * 64;      // 16-byte pieces per component tile read
    constexpr int TILEL16 = (KPL / 16) * 2 * 64;  // ... per tile of the layout
    constexpr int PF = TILE16 / (64 * NW);
    static_assert(TILE16 % (64 * NW) == 0, "component tile must split evenly over the block");
    // one __shared__ variable per buffer: the compiler then sees that a copy
    // in flight does not alias the reads of the other buffer
    __shared__ u32x4 Pl0[TILE16], Pl1[TILE16];
    __shared__ double red[2 * NW];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, c = lane & 31;
    const uint64_t base = (uint64_t)blockIdx.x * (32 * NW);

    V8 ahi[KS], alo[KS];
    {
        // rows past Npad (a multiple of 128) belong to no frame: zero operands
        const uint64_t fr = base + 32 * wave + c;
        const bool ok = fr < a.Npad;
        const u32x4 *zh = reinterpret_cast<const u32x4 *>(a.za_hi) + (ok ? fr : 0) * (KPL / 8);
        const u32x4 *zl = reinterpret_cast<const u32x4 *>(a.za_lo) + (ok ? fr : 0) * (KPL / 8);
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const u32x4 z = (u32x4){0u, 0u, 0u, 0u};
            const u32x4 x = ok ? zh[2 * ks + h] : z, y = ok ? zl[2 * ks + h] : z;
            ahi[ks] = __builtin_bit_cast(V8, x);
            alo[ks] = __builtin_bit_cast(V8, y);
        }
    }

    // mx = -inf, sm = 0: the general update below starts every lane's sum
    // at its first value (d = +inf takes the reset branch, sm = 0 * 0 + 1)
    float mx[16], sm[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        mx[r] = -INFINITY;
        sm[r] = 0.0f;
    }
    const u32x4 *P16 = reinterpret_cast<const u32x4 *>(a.P16);
    // wave w copies pieces [w PF 64, (w+1) PF 64) of each component tile
#define GMMHIP_DMA_TILE(T_, PW_)                                                     \
    _Pragma("unroll") for (int j = 0; j < PF; j++) {                                 \
        const int o = (wave * PF + j) * 64;                                          \
        lds_dma16(dma_rsrc(P16 + (size_t)(T_) * TILEL16, TILE16 * 16),               \
                  (unsigned int)(o + lane) * 16u, &PW_[o]);                          \
    }
#define GMMHIP_TILE_MFMA(ACC_)                                                       \
    _Pragma("unroll") for (int ks = 0; ks < KS; ks++) {                              \
        const V8 bh = __builtin_bit_cast(V8, bt[2 * ks]);                            \
        const V8 bl = __builtin_bit_cast(V8, bt[2 * ks + 1]);                        \
        ACC_ = mfma32(ahi[ks], bh, ACC_);                                            \
        ACC_ = mfma32(ahi[ks], bl, ACC_);                                            \
        ACC_ = mfma32(alo[ks], bh, ACC_);                                            \
    }
    u32x4 bt[2 * KS];
    // tile 0
    GMMHIP_DMA_TILE(0, Pl0)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (a.Mt32 > 1) {
        GMMHIP_DMA_TILE(1, Pl1)
    }
#pragma unroll
    for (int j = 0; j < 2 * KS; j++) bt[j] = Pl0[j * 64 + lane];
    // one FP32 accumulator takes hi*hi, hi*lo and lo*hi
    f32x16 acc = {};
    GMMHIP_TILE_MFMA(acc)
    if (a.Mt32 > 1) {
        // every tile but the last has 32 valid components: start each frame
        // row's reference at the row maximum of tile 0 (all 32 lanes agree)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            mx[r] = row_max32(acc[r]);
            sm[r] = 0.0f;
        }
    }
    // Tile t's MFMAs interleave with the log-sum-exp update of tile t-1.
    // Fast path: no value of the wave's tile lies more than LSE_HEADROOM
    // above its row reference, and sm += 2^d (sm stays below 2^(11+64)).
    // Otherwise the general update moves the reference of every value that
    // exceeded it, row register by row register, skipping the registers in
    // which no lane did: with a narrow-variance model a frame's own
    // component lies hundreds of log2 units above the first tile's maximum,
    // so about half the tiles hold such a jump, mostly in one or two of the
    // sixteen rows a lane carries.  The running sums alternate
    // between two register sets: the fast-path sums SMN = SMC + 2^d are
    // formed among the tile's MFMAs, before the ballot that decides the path
    // (after it, the adds found no MFMA left to overlap), and the general
    // update rebuilds SMN from the untouched SMC.
#define GMMHIP_LL_STEP(ACC_, NACC_, T_, PR_, PW_, SMC, SMN)                          \
    {                                                                                \
        __builtin_amdgcn_s_waitcnt(0);                                               \
        __syncthreads();                                                             \
        if ((T_) + 1 < a.Mt32) {                                                     \
            GMMHIP_DMA_TILE((T_) + 1, PW_)                                           \
        }                                                                            \
        _Pragma("unroll") for (int j = 0; j < 2 * KS; j++) bt[j] = PR_[j * 64 + lane]; \
        GMMHIP_LL_READ_ORDER                                                         \
        NACC_ = (f32x16){};                                                          \
        GMMHIP_TILE_MFMA(NACC_)                                                      \
        float d[16], e[16];                                                          \
        _Pragma("unroll") for (int r = 0; r < 16; r++) {                             \
            d[r] = ACC_[r] - mx[r];                                                  \
            e[r] = exp2f_fast(d[r]);                                                 \
            SMN[r] = SMC[r] + e[r];                                                  \
        }                                                                            \
        float dm = fmaxf(fmaxf(d[0], d[1]), d[2]);                                   \
        _Pragma("unroll") for (int r = 3; r < 16; r += 2) dm = fmaxf(fmaxf(dm, d[r]), d[r + 1 < 16 ? r + 1 : r]); \
        GMMHIP_LL_SCHED                                                              \
        if (__builtin_amdgcn_ballot_w64(dm > LSE_HEADROOM) != 0) {                   \
            _Pragma("unroll") for (int r = 0; r < 16; r++) {                         \
                const bool up = d[r] > LSE_HEADROOM;                                 \
                if (__builtin_amdgcn_ballot_w64(up) == 0) continue;                  \
                const float ee = up ? exp2f_fast(-d[r]) : e[r];                      \
                SMN[r] = up ? fmaf(SMC[r], ee, 1.0f) : SMC[r] + ee;                  \
                mx[r] = up ? ACC_[r] : mx[r];                                        \
            }                                                                        \
        }                                                                            \
    }
    // the tile's LDS reads spread through its MFMA chain, two contraction
    // steps ahead of their use, so that a wave never holds a run of reads in
    // front of MFMAs it could issue (with all 16 reads first, the LDS command
    // queue fills and in-order issue stalls the MFMAs)
#define GMMHIP_LL_READ_ORDER
#define GMMHIP_LL_SCHED                                                              \
    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);                               \
    _Pragma("unroll") for (int ks = 0; ks < KS; ks++) {                              \
        if (ks + 2 < KS) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);          \
        _Pragma("unroll") for (int i = 0; i < 3; i++) {                              \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                       \
            __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);                       \
        }                                                                            \
    }
    f32x16 acc2;
    float sm2[16];
    int t = 1;
    // tile t sits in Pl1 for odd t, Pl0 for even t
    for (; t + 1 < a.Mt32; t += 2) {
        GMMHIP_LL_STEP(acc, acc2, t, Pl1, Pl0, sm, sm2)
        GMMHIP_LL_STEP(acc2, acc, t + 1, Pl0, Pl1, sm2, sm)
    }
    if (t < a.Mt32) {
        GMMHIP_LL_STEP(acc, acc2, t, Pl1, Pl0, sm, sm2)
        acc = acc2;
#pragma unroll
        for (int r = 0; r < 16; r++) sm[r] = sm2[r];
    }
#undef GMMHIP_LL_STEP
#undef GMMHIP_LL_READ_ORDER
#undef GMMHIP_LL_SCHED
#undef GMMHIP_TILE_MFMA
    ll_finish<NW>(a, acc, mx, sm, base, red);
}

// ---------------------------------------------------------------------------
// contraction depths of the log-likelihood kernel, and the layout depths of
// the feature rows and the split model
#define GMMHIP_KP32_CASES(X) X(32) X(64) X(96) X(128) X(160) X(192) X(224) X(256)
#define GMMHIP_KPL_CASES(X) X(32) X(64) X(128) X(256)

hipError_t launch_colstats(const float *X, uint64_t N, int D, int blocks, uint64_t per_block,
                           double *part, hipStream_t s) {
    hipLaunchKernelGGL(colstats_kernel, dim3(blocks), dim3(256), 0, s, X, N, D, per_block, part);
    return hipGetLastError();
}

hipError_t launch_prep_frames(const PrepArgs &p, int prec, hipStream_t s) {
    if (p.D < 1 || p.D > PREP_MAX_D || p.KP != kp_mixed_for_dim(p.D) || p.Npad % 32)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)(p.Npad / 32));
    switch (p.KP) {
#define X(K)                                                                        \
    case K:                                                                         \
        if (prec == 2)                                                              \
            hipLaunchKernelGGL((prep_frames_kernel<_Float16, K>), grid, dim3(256), 0, s, p); \
        else                                                                        \
            hipLaunchKernelGGL((prep_frames_kernel<__bf16, K>), grid, dim3(256), 0, s, p);   \
        break;
        GMMHIP_KPL_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int mixed_ll_frames_per_block(int D, int nw) {
    return (nw == 8 && kp_mixed_contract(D) % 64 == 0) ? 256 : 128;
}

hipError_t launch_loglik_mixed(const MixedArgs &a, int prec, hipStream_t s) {
    const int KPc = kp_mixed_contract(a.D);
    if (a.D < 1 || a.D > PREP_MAX_D || a.KP != kp_mixed_for_dim(a.D)) return hipErrorInvalidValue;
    const int fpb = mixed_ll_frames_per_block(a.D, a.ll_nw);
    const dim3 grid((unsigned)((a.Npad + fpb - 1) / fpb));
    switch (KPc) {
#define X(K)                                                                                      \
    case K: {                                                                                     \
        constexpr int KL = K <= 32 ? 32 : K <= 64 ? 64 : K <= 128 ? 128 : 256;                    \
        if (fpb == 256) {                                                                         \
            if (prec == 2)                                                                        \
                hipLaunchKernelGGL((loglik_mixed_dma_kernel<K, KL, _Float16, (K % 64 == 0 ? 8 : 4)>), grid, dim3(512), 0, s, a); \
            else                                                                                  \
                hipLaunchKernelGGL((loglik_mixed_dma_kernel<K, KL, __bf16, (K % 64 == 0 ? 8 : 4)>), grid, dim3(512), 0, s, a); \
        } else if (prec == 2)                                                                     \
            hipLaunchKernelGGL((loglik_mixed_dma_kernel<K, KL, _Float16, 4>), grid, dim3(256), 0, s, a); \
        else                                                                                      \
            hipLaunchKernelGGL((loglik_mixed_dma_kernel<K, KL, __bf16, 4>), grid, dim3(256), 0, s, a);   \
        break;                                                                                    \
    }
        GMMHIP_KP32_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int mixed_stats_waves(int KP, int comp_tile) {
    // (at KP = 256 eight waves would share a 256-register budget and spill)
    return (comp_tile == 256 && (KP == 64 || KP == 128)) ? 8 : 4;
}

template <bool Q>
static hipError_t launch_stats_mixed_form(const MixedArgs &a, int prec, dim3 grid, hipStream_t s) {
    if (a.st_nw != 4 && a.st_nw != 8) return hipErrorInvalidValue;
    if (a.st_nw == 8 && a.KP != 64 && a.KP != 128) return hipErrorInvalidValue;
    if (a.ngroups != (a.Mt32 + a.st_nw - 1) / a.st_nw) return hipErrorInvalidValue;
    if (a.KP != kp_mixed_for_dim(a.D)) return hipErrorInvalidValue;
    switch (a.KP) {
#define X(K)                                                                                      \
    case K:                                                                                       \
        if (a.st_nw == 8 && (K == 64 || K == 128)) {                                              \
            if (prec == 2)                                                                        \
                hipLaunchKernelGGL((stats_mixed_tr_kernel<K, _Float16, (K == 64 || K == 128 ? 8 : 4), Q>), grid, dim3(512), 0, s, a); \
            else                                                                                  \
                hipLaunchKernelGGL((stats_mixed_tr_kernel<K, __bf16, (K == 64 || K == 128 ? 8 : 4), Q>), grid, dim3(512), 0, s, a);   \
        } else if (prec == 2)                                                                     \
            hipLaunchKernelGGL((stats_mixed_tr_kernel<K, _Float16, 4, Q>), grid, dim3(256), 0, s, a); \
        else                                                                                      \
            hipLaunchKernelGGL((stats_mixed_tr_kernel<K, __bf16, 4, Q>), grid, dim3(256), 0, s, a);   \
        break;
        GMMHIP_KPL_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stats_mixed(const MixedArgs &a, int prec, hipStream_t s) {
    const int nchunks_pad = (a.nchunks + 7) / 8 * 8;
    return launch_stats_mixed_form<false>(a, prec, dim3((unsigned)(nchunks_pad * a.ngroups)), s);
}

int mixed_stats_queue_blocks(int KP, int st_nw, int num_cus, int nitems) {
    // resident blocks: two per CU for the 4-wave kernel at depths up to 128
    const int per_cu = (st_nw == 4 && KP <= 128) ? 2 : 1;
    int n = per_cu * (num_cus > 0 ? num_cus : 256);
    if (n > nitems) n = (nitems + 7) / 8 * 8;
    return n < 8 ? 8 : n;
}

hipError_t launch_stats_mixed_queue(const MixedArgs &a, int prec, hipStream_t s) {
    if (!a.queue || a.queue_blocks < 1) return hipErrorInvalidValue;
    return launch_stats_mixed_form<true>(a, prec, dim3((unsigned)a.queue_blocks), s);
}

}  // namespace gmmhip
