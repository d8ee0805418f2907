// Dense-contraction (GEMM-form) E-step kernels on FP32 MFMA with the
// feature map formed on the fly.  The feature map of a centred frame x' is
// phi_k = xt[a_k] * xt[b_k] over the extended row xt = [1, x'_1..x'_D, 0, ..]
// (index 0 the constant, D+1 a zero): (0, 0) -> 1, (d+1, 0) -> x'_d,
// (i+1, j+1) -> x'_i x'_j, (D+1, D+1) -> 0 for padding.  Only the extended
// rows (DXP floats per frame) go through HBM, not phi (KP floats per frame:
// 1,920 against 64 at D = 60).
//
//   fc_rows_kernel   extended rows of a frame chunk, bad-frame flags
//   fc_ll_kernel     L[t][m] = phi(x'_t) . theta_m   (log-likelihoods)
//   fc_stats_kernel  part[s][m][k] = sum over the frames t of sub-chunk s
//                    of g[t][m] phi_k(x'_t)          (statistics)
//
// Both contractions run on v_mfma_f32_16x16x4f32: a block of 4 waves owns
// a 128 x 128 output tile, each wave 64 x 64 (4 x 4 MFMA tiles of 16 x 16;
// lane (q, c) supplies A[c][q] and B[q][c], result register r of lane
// (q, c) is row 4q + r, column c).
#include "gemm_path.h"

namespace gmmhip {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Extended rows of frames [n0, n0 + nc) into xt[ncp][DXP] (rows [nc, ncp)
// zero).  A frame holding a non-finite (or overflowing, |x'| > 1e18) value
// gets a zero row and bad = 1: its log-likelihood is clamped and it
// contributes no statistics, as the reference (non-finite component
// log-likelihoods -> _minLogLike -> no posterior above the cut).  One wave
// per frame.
__global__ __launch_bounds__(256) void fc_rows_kernel(const float *__restrict__ X, uint64_t n0,
                                                      int nc, int ncp, int D, int DXP,
                                                      const float *__restrict__ cen,
                                                      float *__restrict__ xt,
                                                      uint8_t *__restrict__ bad) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ncp) return;
    // lane covers elements lane, lane + 64, lane + 128 of the row (DXP <= 192)
    bool b = t >= nc;
    float y[3];
#pragma unroll
    for (int h = 0; h < 3; h++) {
        const int e = lane + 64 * h;
        y[h] = e == 0 ? 1.0f : 0.0f;
        if (t < nc && e >= 1 && e <= D) {
            const float v = X[(n0 + t) * (uint64_t)D + (e - 1)] - cen[e - 1];
            if (!(fabsf(v) <= 1e18f)) b = true;
            y[h] = v;
        }
    }
    const bool fb = __builtin_amdgcn_ballot_w64(b) != 0;
    float *row = xt + (size_t)t * DXP;
#pragma unroll
    for (int h = 0; h < 3; h++) {
        const int e = lane + 64 * h;
        if (e < DXP) row[e] = fb ? 0.0f : y[h];
    }
    if (lane == 0) bad[t] = fb ? 1 : 0;
}

namespace {
// LDS strides are chosen for the bank pattern of the operand reads (32 banks
// of 4 bytes, a wave's read served 32 lanes at a time: lanes (q, c) with
// q in {0, 1} or {2, 3}, c = 0..15):
//  * rows of 34 floats: lane (q, c) reading row c, column q hits bank
//    2c + q, all distinct;
//  * columns of 144 floats: lanes c reading 16 consecutive columns of rows
//    a_q, a_q' hit disjoint halves whenever a_q - a_q' is odd (consecutive
//    features of a row of the upper triangle), and the same address when
//    a_q = a_q'.
constexpr int KC = 32;         // contraction step staged per LDS buffer (LL kernel)
constexpr int TST = KC + 2;    // theta row stride in LDS (LL kernel)
constexpr int FS = 128 + 16;   // frame stride of the transposed rows in LDS (LL kernel)
constexpr int FC = 32;         // frames staged per step (statistics kernel)
constexpr int GST = 128 + 16;  // posterior row stride in LDS (statistics kernel)
constexpr int FCS = FC + 2;    // frame stride of the transposed rows in LDS (statistics kernel)
}  // namespace

// L[t][m] for frames t < nc of the chunk and components m < M; theta
// [M][KP] (KP a multiple of 32, padded with zeros), pairs[KP] packed
// a | b << 16.  Block (frame tile, component tile) of 128 x 128.  Out-of-range
// frames and components of a tile read the last valid row (their outputs
// are not written), so no load is guarded.
__global__ __launch_bounds__(256) void fc_ll_kernel(const float *__restrict__ xt, int nc, int D, int DXP,
                                                    const float *__restrict__ theta, int M, int KP,
                                                    const uint32_t *__restrict__ pairs,
                                                    float *__restrict__ L) {
    extern __shared__ float smem[];
    // rows 0 .. D + 1 (= DXP - 2 at most) are the only ones the pairs name
    const int XR = min(DXP, D + 2);
    float *Xs = smem;                   // [XR][FS]: element d of frame f at d * FS + f
    float *Ts = smem + XR * FS;         // [128][TST]
    int *Ps = reinterpret_cast<int *>(Ts + 128 * TST);  // [2][KC]: a * FS, b * FS of the step's pairs
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, c = lane & 15;
    const int f0 = blockIdx.x * 128, m0 = blockIdx.y * 128;
    const int wf = (wave >> 1) * 64, wm = (wave & 1) * 64;
    // the frame tile's extended rows, transposed (consecutive threads take
    // consecutive frames: conflict-free LDS writes)
    const int D4 = DXP >> 2;
    for (int e = tid; e < 128 * D4; e += 256) {
        const int f = e & 127, d4 = e >> 7;
        const float4 v =
            *reinterpret_cast<const float4 *>(xt + (size_t)min(f0 + f, nc - 1) * DXP + 4 * d4);
        float *x = Xs + 4 * d4 * FS + f;
        const float w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (4 * d4 + j < XR) x[j * FS] = w[j];
    }
    // theta staging: 128 components x 32 k per step, 4 float4 per thread;
    // the step's 32 feature pairs ride along (threads 0..31)
    const float *tp[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int idx = tid + 256 * i;
        tp[i] = theta + (size_t)min(m0 + (idx >> 3), M - 1) * KP + 4 * (idx & 7);
    }
    float4 pre[4];
    uint32_t ppre = 0;
    auto load_theta = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; i++) pre[i] = *reinterpret_cast<const float4 *>(tp[i] + k0);
        if (tid < KC) ppre = pairs[k0 + tid];
    };
    load_theta(0);
    f32x4 acc[4][4];
#pragma unroll
    for (int I = 0; I < 4; I++)
#pragma unroll
        for (int J = 0; J < 4; J++) acc[I][J] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    const float *xl = Xs + wf + c;                    // this lane's frame column
    const float *tl = Ts + (wm + c) * TST + q;        // this lane's theta element
    for (int k0 = 0; k0 < KP; k0 += KC) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int idx = tid + 256 * i;
            float *d = Ts + (idx >> 3) * TST + 4 * (idx & 7);
            *reinterpret_cast<float2 *>(d) = make_float2(pre[i].x, pre[i].y);
            *reinterpret_cast<float2 *>(d + 2) = make_float2(pre[i].z, pre[i].w);
        }
        if (tid < KC) {
            Ps[tid] = (int)(ppre & 0xffffu) * FS;
            Ps[KC + tid] = (int)(ppre >> 16) * FS;
        }
        __syncthreads();
        if (k0 + KC < KP) load_theta(k0 + KC);
#pragma unroll
        for (int s = 0; s < KC / 4; s++) {
            const float *xa = xl + Ps[4 * s + q], *xb = xl + Ps[KC + 4 * s + q];
            float av[4], bv[4];
#pragma unroll
            for (int I = 0; I < 4; I++) av[I] = xa[16 * I] * xb[16 * I];
#pragma unroll
            for (int J = 0; J < 4; J++) bv[J] = tl[16 * J * TST + 4 * s];
#pragma unroll
            for (int I = 0; I < 4; I++)
#pragma unroll
                for (int J = 0; J < 4; J++)
                    acc[I][J] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[I], bv[J], acc[I][J], 0, 0, 0);
        }
    }
#pragma unroll
    for (int I = 0; I < 4; I++)
#pragma unroll
        for (int J = 0; J < 4; J++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int f = f0 + wf + 16 * I + 4 * q + r, m = m0 + wm + 16 * J + c;
                if (f < nc && m < M) L[(size_t)f * M + m] = acc[I][J][r];
            }
}

// part[s][m][k] = sum over frames t in [s*sub, (s+1)*sub) of g[t][m] phi_k(t)
// for m < M, k < KP; g [S*sub][M], xt [S*sub][DXP], sub a multiple of FC
// (padding frames have zero posteriors and zero rows).  One block per
// (128 x 128 output tile, sub-chunk) = tiles * S blocks, launched as a 1-D
// grid of a multiple of 8 blocks: hardware block b runs on XCD b % 8, and
// the logical blocks of one XCD are a contiguous range, so the tiles of a
// sub-chunk share its posteriors and rows through that XCD's L2.
// Components past M read the last component's posteriors (their outputs
// are not written), so no load is guarded.
template <int NXI>  // float4 of the staged rows per thread: ceil(DXP / 32)
__global__ __launch_bounds__(256) void fc_stats_kernel(const float *__restrict__ xt, int DXP,
                                                       const float *__restrict__ g, int M, int KP,
                                                       const uint32_t *__restrict__ pairs, int sub,
                                                       int S, float *__restrict__ part) {
    extern __shared__ float smem[];
    float *Gs = smem;               // [FC][GST]
    float *Xs = smem + FC * GST;    // [DXP][FCS]: element d of frame f at d * FCS + f
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, c = lane & 15;
    const int mt = (M + 127) / 128, tiles = mt * ((KP + 127) / 128), nb = tiles * S;
    const int per = (nb + 7) / 8;
    const int lb = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (lb >= nb) return;
    const int s = lb / tiles, tile = lb - s * tiles;
    const int m0 = (tile % mt) * 128, k0 = (tile / mt) * 128;
    const int wm = (wave >> 1) * 64, wk = (wave & 1) * 64;
    // this lane's four features 16 J + c of the wave's 64, as LDS offsets
    // of its frame column (features past KP read the zero slot DXP - 1)
    const float *xa[4], *xb[4];
#pragma unroll
    for (int J = 0; J < 4; J++) {
        const int k = k0 + wk + 16 * J + c;
        const uint32_t p = k < KP ? pairs[k] : (uint32_t)(DXP - 1) * 0x10001u;
        xa[J] = Xs + (int)(p & 0xffffu) * FCS + q;
        xb[J] = Xs + (int)(p >> 16) * FCS + q;
    }
    // stage prefetch: FC x 128 posteriors (16 per thread), FC x DXP row
    // floats (NXI float4 per thread; consecutive threads take consecutive
    // frames, so the transposed LDS writes are conflict-free).  Offsets
    // within a sub-chunk fit 32 bits (a chunk's posteriors are < 2^31 floats).
    const float *gb = g + (size_t)s * sub * M + min(m0 + (tid & 127), M - 1);
    const float *xbase = xt + (size_t)s * sub * DXP;
    const int nx4 = FC * (DXP >> 2);
    float gp[FC * 128 / 256];
    float4 xp[NXI];
    auto load_stage = [&](int f0) {
#pragma unroll
        for (int i = 0; i < FC * 128 / 256; i++) gp[i] = gb[(f0 + (tid >> 7) + 2 * i) * M];
#pragma unroll
        for (int i = 0; i < NXI; i++) {
            const int e = min(tid + 256 * i, nx4 - 1);
            xp[i] = *reinterpret_cast<const float4 *>(xbase + (f0 + (e & (FC - 1))) * DXP + 4 * (e / FC));
        }
    };
    load_stage(0);
    f32x4 acc[4][4];
#pragma unroll
    for (int I = 0; I < 4; I++)
#pragma unroll
        for (int J = 0; J < 4; J++) acc[I][J] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    const float *gl = Gs + q * GST + wm + c;  // this lane's posterior element
    for (int f0 = 0; f0 < sub; f0 += FC) {
        __syncthreads();
        // (a row element past the stage's DXP x FC is the clamped last one,
        // rewritten with its own value)
#pragma unroll
        for (int i = 0; i < FC * 128 / 256; i++) Gs[((tid >> 7) + 2 * i) * GST + (tid & 127)] = gp[i];
#pragma unroll
        for (int i = 0; i < NXI; i++) {
            const int e = min(tid + 256 * i, nx4 - 1);
            float *x = Xs + (4 * (e / FC)) * FCS + (e & (FC - 1));
            x[0] = xp[i].x;
            x[FCS] = xp[i].y;
            x[2 * FCS] = xp[i].z;
            x[3 * FCS] = xp[i].w;
        }
        __syncthreads();
        if (f0 + FC < sub) load_stage(f0 + FC);
#pragma unroll
        for (int st = 0; st < FC / 4; st++) {
            float av[4], bv[4];
#pragma unroll
            for (int I = 0; I < 4; I++) av[I] = gl[4 * st * GST + 16 * I];
#pragma unroll
            for (int J = 0; J < 4; J++) bv[J] = xa[J][4 * st] * xb[J][4 * st];
#pragma unroll
            for (int I = 0; I < 4; I++)
#pragma unroll
                for (int J = 0; J < 4; J++)
                    acc[I][J] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[I], bv[J], acc[I][J], 0, 0, 0);
        }
    }
    float *out = part + (size_t)s * M * KP;
#pragma unroll
    for (int I = 0; I < 4; I++)
#pragma unroll
        for (int J = 0; J < 4; J++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * I + 4 * q + r, k = k0 + wk + 16 * J + c;
                if (m < M && k < KP) out[(size_t)m * KP + k] = acc[I][J][r];
            }
}

}  // namespace gmmhip

namespace gmmhip_detail {

int ext_row_floats(int D) { return (D + 2 + 3) / 4 * 4; }

std::vector<uint32_t> feature_pairs(const std::vector<int> &codes, int D, int KP) {
    const uint32_t z = (uint32_t)(D + 1);
    std::vector<uint32_t> p((size_t)KP, z | (z << 16));
    for (size_t k = 0; k < codes.size(); k++) {
        const int cd = codes[k];
        uint32_t a, b;
        if (cd == -1) {
            a = 0;
            b = 0;
        } else if (cd < 0) {
            a = (uint32_t)(-2 - cd) + 1;
            b = 0;
        } else {
            a = (uint32_t)(cd >> 16) + 1;
            b = (uint32_t)(cd & 0xffff) + 1;
        }
        p[k] = a | (b << 16);
    }
    return p;
}

int stats_subchunks(int M, int KP, int num_cus, uint64_t nc_max) {
    const int tiles = ((M + 127) / 128) * ((KP + 127) / 128);
    const int cus = num_cus > 0 ? num_cus : 256;
    int S = std::max(1, (16 * cus + tiles / 2) / tiles);
    S = std::min(S, 128);
    S = (int)std::min<uint64_t>((uint64_t)S, std::max<uint64_t>(1, nc_max / 256));
    const uint64_t part = (uint64_t)M * (uint64_t)KP * 4;
    S = (int)std::min<uint64_t>((uint64_t)S, std::max<uint64_t>(1, (512ull << 20) / part));
    return std::max(1, S);
}

hipError_t launch_fc_rows(const float *X, uint64_t n0, int nc, int ncp, int D, const float *cen,
                          float *xt, uint8_t *bad, hipStream_t s) {
    hipLaunchKernelGGL(gmmhip::fc_rows_kernel, dim3((ncp + 3) / 4), dim3(256), 0, s, X, n0, nc, ncp,
                       D, ext_row_floats(D), cen, xt, bad);
    return hipGetLastError();
}

hipError_t launch_fc_ll(const float *xt, int nc, int D, const float *theta, int M, int KP,
                        const uint32_t *pairs, float *L, hipStream_t s) {
    const int DXP = ext_row_floats(D);
    if (DXP > 192 || DXP % 4 != 0 || KP % KC != 0) return hipErrorInvalidValue;
    const size_t shm = (size_t)(std::min(DXP, D + 2) * FS + 128 * TST + 2 * KC) * sizeof(float);
    if (shm > 65536) {
        // above the default dynamic LDS limit (D > 88): raise it, once
        static bool raised = false;
        if (!raised) {
            const hipError_t e = hipFuncSetAttribute(
                reinterpret_cast<const void *>(&gmmhip::fc_ll_kernel),
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)(192 * FS + 128 * TST + 2 * KC) * 4);
            if (e != hipSuccess) return e;
            raised = true;
        }
    }
    const dim3 grid((unsigned)((nc + 127) / 128), (unsigned)((M + 127) / 128));
    hipLaunchKernelGGL(gmmhip::fc_ll_kernel, grid, dim3(256), shm, s, xt, nc, D, DXP, theta, M, KP, pairs,
                       L);
    return hipGetLastError();
}

hipError_t launch_fc_stats(const float *xt, int DXP, const float *g, int M, int KP,
                           const uint32_t *pairs, int sub, int S, float *part, hipStream_t s) {
    if (DXP > 192) return hipErrorInvalidValue;
    const size_t shm = (size_t)(FC * GST + DXP * FCS) * sizeof(float);
    if (DXP % 4 != 0 || S < 1 || sub % FC != 0 || (uint64_t)sub * (uint64_t)M >= (1ull << 31))
        return hipErrorInvalidValue;
    const int nb = ((M + 127) / 128) * ((KP + 127) / 128) * S;
    const dim3 grid((unsigned)((nb + 7) / 8 * 8));
    switch ((DXP + 31) / 32) {
#define FC_STATS(n)                                                                                    \
    case n:                                                                                            \
        hipLaunchKernelGGL(gmmhip::fc_stats_kernel<n>, grid, dim3(256), shm, s, xt, DXP, g, M, KP, pairs, \
                           sub, S, part);                                                              \
        break;
        FC_STATS(1)
        FC_STATS(2)
        FC_STATS(3)
        FC_STATS(4)
        FC_STATS(5)
        FC_STATS(6)
#undef FC_STATS
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gmmhip_detail
