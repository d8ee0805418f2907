// Batched per-utterance statistics, MAP means and MLLR transforms on gfx950.
//
// Thousands of utterances of uneven length are processed together: the
// per-frame log-likelihoods of all frames run as one launch (the E-step
// kernels), the posteriors + statistics of every (utterance chunk,
// component group) work item run in one persistent launch that pulls items
// from a device work queue (stats kernel, queue mode), and the kernels below
// turn the chunk partials into per-utterance statistics, MAP-adapted means
// (UpdateModelMAP, trainer/OL_ModelAdapt.cpp:993-1106) and MLLR row
// statistics and transforms (ComputeKiGi_MLLR / prepareKiGiInitMatrices /
// ComputeW_MLLR, OL_ModelAdapt.cpp:1417-1581).  MLLR statistics are
// contractions over components:
//   G[(u,i)][p(a,b)] = sum_m (n_um / var_mi) xi_ma xi_mb,  xi_m = [mu_m; 1]
//   k[(u,i)][a]      = sum_m (F_umi / var_mi) xi_ma
// i.e. two FP64 GEMMs against UBM-only matrices (xi xi^T packed, and xi),
// followed by one FP64 Cholesky solve per (utterance, row).
#include "batch.h"

#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

namespace gmmhip {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
__global__ void fill_u32_kernel(unsigned int *p, unsigned int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Utterance u's rows X[src[u] .. src[u+1]) to Xp[dst[u] ..), the rows up to
// dst[u+1] zero (each utterance of a mixed-precision batch starts on a
// 32-row block).  One block per utterance.
__global__ __launch_bounds__(256) void pad_rows_kernel(const float *__restrict__ X,
                                                       const uint64_t *__restrict__ src,
                                                       const uint64_t *__restrict__ dst, int D,
                                                       float *__restrict__ Xp) {
    const int u = blockIdx.x;
    const uint64_t n = (src[u + 1] - src[u]) * (uint64_t)D;
    const uint64_t np = (dst[u + 1] - dst[u]) * (uint64_t)D;
    const float *x = X + src[u] * (uint64_t)D;
    float *y = Xp + dst[u] * (uint64_t)D;
    for (uint64_t e = threadIdx.x; e < np; e += 256) y[e] = e < n ? x[e] : 0.0f;
}

// The same copy in 16-byte pieces (D a multiple of 4, both bases 16-byte
// aligned): a row of D floats is D/4 pieces.
__global__ __launch_bounds__(256) void pad_rows4_kernel(const float4 *__restrict__ X,
                                                        const uint64_t *__restrict__ src,
                                                        const uint64_t *__restrict__ dst, int D4,
                                                        float4 *__restrict__ Xp) {
    const int u = blockIdx.x;
    const uint64_t n = (src[u + 1] - src[u]) * (uint64_t)D4;
    const uint64_t np = (dst[u + 1] - dst[u]) * (uint64_t)D4;
    const float4 *x = X + src[u] * (uint64_t)D4;
    float4 *y = Xp + dst[u] * (uint64_t)D4;
    for (uint64_t e = threadIdx.x; e < np; e += 256)
        y[e] = e < n ? x[e] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

hipError_t launch_pad_rows(const float *X, const uint64_t *src, const uint64_t *dst, int U, int D,
                           float *Xp, hipStream_t s) {
    if (D % 4 == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0 &&
        (reinterpret_cast<uintptr_t>(Xp) & 15) == 0)
        hipLaunchKernelGGL(pad_rows4_kernel, dim3(U), dim3(256), 0, s,
                           reinterpret_cast<const float4 *>(X), src, dst, D / 4,
                           reinterpret_cast<float4 *>(Xp));
    else
        hipLaunchKernelGGL(pad_rows_kernel, dim3(U), dim3(256), 0, s, X, src, dst, D, Xp);
    return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int *p, unsigned int v, int n, hipStream_t s) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-utterance reduction of the chunk partials [chunk][Mp][KP] (columns
// [1, x', x'^2] of centred frames) into n[u][m] and un-centred F[u][m][d],
// FP64, chunks of an utterance summed in order, and, for the first Umeans
// utterances, the MAP means of UpdateModelMAP (trainer/OL_ModelAdapt.cpp:
// 993-1106; means only): alpha = n/(n+tau), mu' = alpha F/n + (1-alpha) mu
// for n > 0, the UBM mean otherwise.  A block of 4 waves takes 4 components
// of UB consecutive utterances (groups of UB strided by gridDim.y), wave w
// component 4 blockIdx.x + w and lane j the columns j, j + 64, ... of 0..D,
// so no thread divides an index; every lane sums the occupancy column
// itself, in the same order, so each has n without a broadcast.  The first
// chunk of each of the UB utterances is loaded before any is used: with one
// utterance per round a wave waited out one memory latency per 61 values
// (3.15 ms for 4,096 utterances x 2048 components at D=60; 2.17 ms with 16).
template <int UB>
__global__ __launch_bounds__(256) void seg_reduce_kernel(SegReduceArgs a) {
    // the wave's component in scalar registers: the UB row addresses are
    // then a scalar base plus the lane's column offset
    const int m = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (m >= a.M) return;
    const int J = a.D + 1;
    const size_t E = (size_t)a.Mp * a.KP;
    const float *pm = a.partials + (size_t)m * a.KP;
    for (int u0 = blockIdx.y * UB; u0 < a.U; u0 += gridDim.y * UB) {
        int cb[UB], ce[UB];
#pragma unroll
        for (int k = 0; k < UB; k++) {
            const bool in = u0 + k < a.U;
            cb[k] = in ? a.chunk_first[u0 + k] : 0;
            ce[k] = in ? a.chunk_first[u0 + k + 1] : 0;
        }
        for (int j = (int)(threadIdx.x & 63); j < J; j += 64) {
            float v[UB], v0[UB];
#pragma unroll
            for (int k = 0; k < UB; k++) {
                v[k] = 0.0f;
                v0[k] = 0.0f;
                if (cb[k] < ce[k]) {
                    const float *p = pm + (size_t)cb[k] * E;
                    v[k] = p[j];
                    v0[k] = p[0];
                }
            }
#pragma unroll
            for (int k = 0; k < UB; k++) {
                const int u = u0 + k;
                if (u >= a.U) break;
                double s = 0.0, n = 0.0;
                if (cb[k] < ce[k]) {
                    s += (double)v[k];
                    if (j > 0) n += (double)v0[k];
                }
                for (int c = cb[k] + 1; c < ce[k]; c++) {
                    const float *p = pm + (size_t)c * E;
                    s += (double)p[j];
                    if (j > 0) n += (double)p[0];
                }
                const uint64_t um = (uint64_t)u * a.M + m;
                if (j == 0) {
                    a.n[um] = s;
                } else {
                    // x = x' + o (FP32 form) or x = z / scale + o (mixed form):
                    // F = F' + n o or F = F' / scale + n o
                    const double f = a.scale ? s / (double)a.scale[j - 1] : s;
                    const double Fv = f + n * (double)a.off[j - 1];
                    a.F[um * a.D + (j - 1)] = Fv;
                    if (u < a.Umeans) {
                        const float ub = a.mu[(size_t)m * a.D + (j - 1)];
                        float mv = ub;
                        if (n > 0.0) {
                            const double alfa = n / (n + (double)a.tau);
                            mv = (float)(alfa * Fv / n + (1.0 - alfa) * (double)ub);
                        }
                        a.means[um * a.D + (j - 1)] = mv;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// CheckInvertibility (OL_ModelAdapt.cpp:1851-1878) per utterance: status 0, or
// 7 when fewer than D+1 components are active or the occupancy is below D+1.
__global__ __launch_bounds__(256) void mllr_check_kernel(const double *__restrict__ n, int U, int M,
                                                         int D, int *__restrict__ status) {
    __shared__ int cnt[256];
    __shared__ double occ[256];
    const int u = blockIdx.x;
    int c = 0;
    double o = 0.0;
    for (int m = threadIdx.x; m < M; m += 256) {
        const double v = n[(size_t)u * M + m];
        if (v > 0.0) {
            c++;
            o += v;
        }
    }
    cnt[threadIdx.x] = c;
    occ[threadIdx.x] = o;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            cnt[threadIdx.x] += cnt[threadIdx.x + w];
            occ[threadIdx.x] += occ[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool ok = M >= D + 1 && cnt[0] >= D + 1 && occ[0] * D >= (double)D * (D + 1);
        status[u] = ok ? 0 : 7;
    }
}

// ---------------------------------------------------------------------------
// MLLR G / k statistics of every utterance (ComputeKiGi_MLLR,
// OL_ModelAdapt.cpp:1525-1581):
//   C[(u,i)][col] = sum over components m with n[u][m] > 0 of
//                   A_u[i][m] B[m][col],
//   A_u[i][m] = n[u][m] ivar[m][i] (G) or F[u][m][i] ivar[m][i] (k).
// Block = (utterance u, 64-row tile, 128-column tile).  The block first compacts the
// components u occupies into an ordered list in LDS (a wave ballot and
// prefix count per 256 components), so the K loop runs over those only —
// a few hundred of 2048 for a 300-frame utterance — and the result depends
// on u's statistics alone.  Then a 64 x 128 tile (rows i, four waves of
// 32 x 64, 8 MFMA tiles each) accumulates on v_mfma_f64_16x16x4f64 (lane
// (q, c) gives A[c][q] and B[q][c]; result register r of lane (q, c) is row
// q + 4r, column c), K staged 16 components at a time through k-major LDS
// tiles with the next chunk prefetched into registers.  The column tiles of
// one utterance run on one XCD.
template <int KMODE>
__global__ __launch_bounds__(256) void mllr_gk_kernel(MllrGkArgs a) {
    constexpr int LDA = 80, LDB = 144;  // row strides (doubles) of the k-major tiles
    constexpr int MAXM = MLLR_MAX_COMPONENTS;
    __shared__ double At[16 * LDA];
    __shared__ double Bt[16 * LDB];
    __shared__ int act[MAXM];
    __shared__ int wcount[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, c = lane & 15;
    const int ncol = (a.Cn + 127) / 128, nrt = (a.D + 63) / 64;
    const int b = blockIdx.x, xcd = b & 7, idx = b >> 3;
    const int u = (idx / (ncol * nrt)) * 8 + xcd;
    const int rt = (idx % (ncol * nrt)) / ncol, ct = idx % ncol;
    if (u >= a.U) return;
    const int c0 = ct * 128, i0 = rt * 64;
    const double *nu = a.n + (size_t)u * a.M;

    // ordered list of the occupied components
    int K = 0;
    for (int m0 = 0; m0 < a.M; m0 += 256) {
        const int m = m0 + tid;
        const bool on = m < a.M && nu[m] > 0.0;
        const unsigned long long bal = __ballot(on);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wcount[wave] = __popcll(bal);
        __syncthreads();
        int off = K;
        for (int w = 0; w < wave; w++) off += wcount[w];
        if (on) act[off + before] = m;
        K += wcount[0] + wcount[1] + wcount[2] + wcount[3];
        __syncthreads();
    }

    // staging: A tile 16 x 64 (4 per thread), B tile 16 x 128 (8 per thread)
    const int ai = tid & 63, ak = tid >> 6;     // A: row i0 + ai, K rows ak + 4j
    const int bc = tid & 127, bk = tid >> 7;    // B: column, K rows bk + 2j
    const bool iok = i0 + ai < a.D;
    const bool cok = c0 + bc < a.Cn;
    const double *fu = a.F + (size_t)u * a.M * a.D + i0 + ai;
    const double *ip = a.ivar + i0 + ai;
    const double *bp = a.B + c0 + bc;
    double pa[4], pb[8];
#define GMMHIP_GK_LOAD(K0_)                                                      \
    _Pragma("unroll") for (int j = 0; j < 4; j++) {                              \
        const int kk = (K0_) + ak + 4 * j;                                       \
        double v = 0.0;                                                          \
        if (kk < K && iok) {                                                     \
            const int m = act[kk];                                               \
            const double x = KMODE ? fu[(size_t)m * a.D] : nu[m];                \
            v = x * ip[(size_t)m * a.D];                                         \
        }                                                                        \
        pa[j] = v;                                                               \
    }                                                                            \
    _Pragma("unroll") for (int j = 0; j < 8; j++) {                              \
        const int kk = (K0_) + bk + 2 * j;                                       \
        pb[j] = (kk < K && cok) ? bp[(size_t)act[kk] * a.ldb] : 0.0;             \
    }
    GMMHIP_GK_LOAD(0)

    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64;
    f64x4 acc[2][4];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = (f64x4){0.0, 0.0, 0.0, 0.0};

    for (int k0 = 0; k0 < K; k0 += 16) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++) At[(ak + 4 * j) * LDA + ai] = pa[j];
#pragma unroll
        for (int j = 0; j < 8; j++) Bt[(bk + 2 * j) * LDB + bc] = pb[j];
        __syncthreads();
        if (k0 + 16 < K) {
            GMMHIP_GK_LOAD(k0 + 16)
        }
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            double av[2], bv[4];
#pragma unroll
            for (int x = 0; x < 2; x++) av[x] = At[(kk + q) * LDA + wr + 16 * x + c];
#pragma unroll
            for (int y = 0; y < 4; y++) bv[y] = Bt[(kk + q) * LDB + wc + 16 * y + c];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 4; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
        }
    }
#undef GMMHIP_GK_LOAD
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = i0 + wr + 16 * x + q + 4 * r, col = c0 + wc + 16 * y + c;
                if (i < a.D && col < a.Cn) a.C[((size_t)u * a.D + i) * a.ldc + col] = acc[x][y][r];
            }
}

hipError_t launch_mllr_gk(const MllrGkArgs &a, hipStream_t s) {
    if (a.U <= 0) return hipSuccess;
    if (a.M > MLLR_MAX_COMPONENTS) return hipErrorInvalidValue;
    const int ncol = (a.Cn + 127) / 128, nrt = (a.D + 63) / 64;
    const dim3 grid((unsigned)((a.U + 7) / 8 * 8 * ncol * nrt));
    if (a.kmode)
        hipLaunchKernelGGL(mllr_gk_kernel<1>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(mllr_gk_kernel<0>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// One wave per (u, i): G (packed upper triangle + optional prior) and k are
// assembled in LDS, G = L L^T by a left-looking Cholesky (lane = row), then
// L y = k and L^T w = y.  A non-positive pivot marks the utterance's status.
__global__ __launch_bounds__(64) void mllr_solve_kernel(MllrSolveArgs a) {
    extern __shared__ double sh[];
    const int E = a.D + 1;
    const int LDS = E + 1;
    double *G = sh;               // [E][LDS]
    double *v = sh + E * LDS;     // [E]
    const uint64_t ui = blockIdx.x;
    const int u = (int)(ui / a.D), i = (int)(ui % a.D);
    const int lane = threadIdx.x;
    if (a.status[u] != 0) {
        for (int p = lane; p < E; p += 64) a.W[ui * E + p] = 0.0;
        return;
    }
    const double *gp = a.Gp + ui * a.ldg;
    const int NP = E * (E + 1) / 2;
    for (int p = lane; p < E * E; p += 64) {
        const int r = p / E, cc = p - r * E;
        const int lo = r < cc ? r : cc, hi = r < cc ? cc : r;
        // packed index of (lo, hi), lo <= hi: lo*E - lo(lo-1)/2 + (hi - lo)
        const int pk = lo * E - (lo * (lo - 1)) / 2 + (hi - lo);
        double g = gp[pk];
        if (a.Ginit) g += a.Ginit[(size_t)i * NP + pk];
        G[r * LDS + cc] = g;
    }
    for (int p = lane; p < E; p += 64) {
        double kv = a.Kp[ui * a.ldk + p];
        if (a.kinit) kv += a.kinit[(size_t)i * E + p];
        v[p] = kv;
    }
    __syncthreads();
    bool bad = false;
    for (int j = 0; j < E; j++) {
        // diagonal
        double d = G[j * LDS + j];
        for (int p = 0; p < j; p++) d -= G[j * LDS + p] * G[j * LDS + p];
        bad = bad || !(d > 0.0);
        const double ljj = sqrt(d > 0.0 ? d : 1.0);
        // column below the diagonal, one row per lane
        for (int r = j + 1 + lane; r < E; r += 64) {
            double s = G[r * LDS + j];
            for (int p = 0; p < j; p++) s -= G[r * LDS + p] * G[j * LDS + p];
            G[r * LDS + j] = s / ljj;
        }
        __syncthreads();
        if (lane == 0) G[j * LDS + j] = ljj;
        __syncthreads();
    }
    // forward substitution L y = k (serial over rows, lanes share the dot)
    for (int r = 0; r < E; r++) {
        double s = 0.0;
        for (int p = lane; p < r; p += 64) s += G[r * LDS + p] * v[p];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) v[r] = (v[r] - s) / G[r * LDS + r];
        __syncthreads();
    }
    // back substitution L^T w = y
    for (int r = E - 1; r >= 0; r--) {
        double s = 0.0;
        for (int p = r + 1 + lane; p < E; p += 64) s += G[p * LDS + r] * v[p];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) v[r] = (v[r] - s) / G[r * LDS + r];
        __syncthreads();
    }
    for (int p = lane; p < E; p += 64) a.W[ui * E + p] = v[p];
    if (bad && lane == 0) atomicOr(&a.status[u], 8);
}

// ---------------------------------------------------------------------------
// Per-segment FP64 sums: thread t of block u adds frames t, t+256, ... of
// segment u in order, then a fixed tree combines the 256 partial sums.
__global__ __launch_bounds__(256) void seg_sum_kernel(const float *__restrict__ v,
                                                      const uint64_t *__restrict__ begin,
                                                      const uint64_t *__restrict__ end,
                                                      double *__restrict__ out) {
    __shared__ double red[256];
    const int u = blockIdx.x;
    const uint64_t b = begin[u], e = end[u];
    double s = 0.0;
    for (uint64_t i = b + threadIdx.x; i < e; i += 256) s += (double)v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[u] = red[0];
}

hipError_t launch_seg_sum(const float *v, const uint64_t *begin, const uint64_t *end, int U,
                          double *out, hipStream_t s) {
    if (U <= 0) return hipSuccess;
    hipLaunchKernelGGL(seg_sum_kernel, dim3(U), dim3(256), 0, s, v, begin, end, out);
    return hipGetLastError();
}

hipError_t launch_seg_reduce(const SegReduceArgs &a, hipStream_t s) {
    if (a.U <= 0 || a.M <= 0) return hipSuccess;
    constexpr int UB = 16;
    const dim3 grid((unsigned)((a.M + 3) / 4), (unsigned)std::min((a.U + UB - 1) / UB, 65535));
    hipLaunchKernelGGL(seg_reduce_kernel<UB>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mllr_check(const double *n, int U, int M, int D, int *status, hipStream_t s) {
    hipLaunchKernelGGL(mllr_check_kernel, dim3(U), dim3(256), 0, s, n, U, M, D, status);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Blocked LDL^T solve for E = D+1 <= 64 on FP64 MFMA.  One wave per (u, i).
// W_i = G_i^-1 k_i of ComputeW_MLLR, which the reference forms by inverting
// G_i with Gauss-Jordan elimination (MATRIX::_Invert_Gen,
// Matrix4_LMa.cpp:2232); G_i is symmetric positive definite, so here it is
// factored instead and k_i substituted through the factor.
// G (symmetric, unpacked from the packed upper triangle plus the prior) is
// padded to 64 x 64 with an identity block and held in registers in the
// accumulator layout of v_mfma_f64_16x16x4f64: tile (I, J) of 16 x 16, lane
// (q, c), register r = element (16 I + q + 4 r, 16 J + c); k sits one row per
// lane.  Step s factors the columns K = 4s .. 4s+3: with P = A[:, K] (through
// LDS) and the 4 x 4 block B = A[K, K] = L_K D_K L_K^T (unit lower L_K,
// every lane repeats it), X = P L_K^-T for the rows below K (lane = row), and
//   A <- A - X D_K^-1 X^T
// on the tiles on or below the diagonal from tile column K / 16 on (rows and
// columns up to K carry X = 0, so finished columns and the block itself are
// untouched): one rank-4 MFMA per tile, 80 per system against 160 for the
// full Gauss-Jordan update.  The lanes holding the columns K keep
// L[:, K] = X D_K^-1 (the unit lower factor) for the back substitution, and
// k takes the forward substitution L z = k in the same step (z_K by L_K,
// the rows below by X D_K^-1 z_K).  After the last step u = D^-1 z, and
// L^T w = u runs block by block from the last: w_K = L_K^-T u_K, then
// u_i -= L[K, i] . w_K for every row i above K, L[K, :] staged from the
// registers that hold row block K.  A non-positive D_K entry (G not
// positive definite) marks the utterance's status.
// The solve is latency-bound (a serial chain of 4 x 4 factorisations, LDS
// round trips and dependent FP64 MFMAs per step), so the kernel is held to
// four waves per SIMD: only the ten tiles on and below the diagonal are
// live (128 VGPRs) and G is read from the packed triangle straight into
// them (8 KB of LDS per wave).  Measured on the 4,096-utterance batch
// (245,760 systems of 61): 3.07 ms, against 3.56 ms at three waves and
// 3.62 ms for a blocked Gauss-Jordan with the full 64 x 64 system (160
// MFMAs per system, two waves per SIMD).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void mllr_solve_ldl_kernel(
    MllrSolveArgs a) {
    // LDS: P (64 x 4; the back substitution's row block of L in the same
    // place), X D^-1 and X (64 x 4 each), the factored blocks (16 per step:
    // L_K below the diagonal, 1/D_K on it), k of the rows K
    __shared__ double sW[4 * 256 + 8];
    double *sP = sW;
    double *sR = sW;
    double *sX = sW + 256;
    double *sXr = sW + 512;
    double *sL = sW + 768;
    double *sK = sW + 1024;
    const int E = a.D + 1;
    const int NP = E * (E + 1) / 2;
    const uint64_t ui = blockIdx.x;
    const int u = (int)(ui / a.D), i = (int)(ui % a.D);
    const int lane = threadIdx.x, q = lane >> 4, c = lane & 15;
    if (a.status[u] != 0) {
        if (lane < E) a.W[ui * E + lane] = 0.0;
        return;
    }
    // the tiles on and below the diagonal, straight from the packed upper
    // triangle (and the prior); the tiles above it are never read
    const double *gp = a.Gp + ui * a.ldg;
    const double *gip = a.Ginit ? a.Ginit + (size_t)i * NP : nullptr;
    f64x4 A[4][4];
#pragma unroll
    for (int I = 0; I < 4; I++)
#pragma unroll
        for (int J = 0; J <= I; J++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = 16 * I + q + 4 * r, col = 16 * J + c;
                double v = row == col ? 1.0 : 0.0;
                if (row < E && col < E) {
                    const int lo = row < col ? row : col, hi = row < col ? col : row;
                    const int pk = lo * E - (lo * (lo - 1)) / 2 + (hi - lo);
                    v = gp[pk];
                    if (gip) v += gip[pk];
                }
                A[I][J][r] = v;
            }
    double rhs = 0.0;
    if (lane < E) {
        rhs = a.Kp[ui * a.ldk + lane];
        if (a.kinit) rhs += a.kinit[(size_t)i * E + lane];
    }
    bool bad = false;
    const int S = (E + 3) / 4;
#pragma unroll
    for (int s = 0; s < 16; s++) {
        if (s < S) {
            const int Ib = s >> 2, cb = 4 * (s & 3);
            const int k0 = 4 * s;
            __syncthreads();  // the previous step's reads are done
            // (rows above tile row Ib are not read: the tiles above the
            // diagonal stay dead)
            if (c >= cb && c < cb + 4) {
#pragma unroll
                for (int I = 0; I < 4; I++)
                    if (I >= Ib) {
#pragma unroll
                        for (int r = 0; r < 4; r++) sP[(16 * I + q + 4 * r) * 4 + (c - cb)] = A[I][Ib][r];
                    }
            }
            if (lane >= k0 && lane < k0 + 4) sK[lane - k0] = rhs;
            __syncthreads();
            // B = L_K D_K L_K^T (every lane)
            double l[4][4], dv[4], dinv[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // d_j = B[j][j] - sum_p L[j][p]^2 d_p, then column j of L
                double d = sP[(k0 + j) * 4 + j];
#pragma unroll
                for (int p = 0; p < j; p++) d = fma(-l[j][p] * dv[p], l[j][p], d);
                bad = bad || !(d > 0.0);
                const double dd = d > 0.0 ? d : 1.0;
                dv[j] = dd;
                // 1/d: the hardware reciprocal and two Newton steps
                double rc = __builtin_amdgcn_rcp(dd);
                rc = fma(rc, fma(-dd, rc, 1.0), rc);
                dinv[j] = fma(rc, fma(-dd, rc, 1.0), rc);
#pragma unroll
                for (int tt = j + 1; tt < 4; tt++) {
                    double v = sP[(k0 + tt) * 4 + j];
#pragma unroll
                    for (int p = 0; p < j; p++) v = fma(-l[tt][p] * dv[p], l[j][p], v);
                    l[tt][j] = v * dinv[j];
                }
            }
            // this lane's row: x = P[row] L_K^-T (rows below K; zero above)
            double x[4] = {0.0, 0.0, 0.0, 0.0};
            if (lane >= k0 + 4) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    double v = sP[lane * 4 + j];
#pragma unroll
                    for (int p = 0; p < j; p++) v = fma(-x[p], l[j][p], v);
                    x[j] = v;
                }
            }
            // X and L[row][K] = X D_K^-1 (the factor kept for the back substitution)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                sXr[lane * 4 + j] = x[j];
                sX[lane * 4 + j] = x[j] * dinv[j];
            }
            // forward substitution: z_K = L_K^-1 k_K, rows below -= L[row][K] z_K
            double z[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                double v = sK[j];
#pragma unroll
                for (int p = 0; p < j; p++) v = fma(-l[j][p], z[p], v);
                z[j] = v;
            }
            if (lane >= k0 + 4) {
#pragma unroll
                for (int j = 0; j < 4; j++) rhs = fma(-x[j] * dinv[j], z[j], rhs);
            } else if (lane >= k0) {
                // u_K = D_K^-1 z_K (selected, not indexed: no scratch array)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (lane - k0 == j) rhs = z[j] * dinv[j];
            }
            if (lane < 16) {
                double val = 0.0;
#pragma unroll
                for (int tt = 0; tt < 4; tt++)
#pragma unroll
                    for (int v = 0; v <= tt; v++)
                        if (lane == 4 * tt + v) val = tt > v ? l[tt][v] : dinv[tt];
                sL[s * 16 + lane] = val;
            }
            __syncthreads();
            // operands: A[row c][k q] = -X[16 I + c][q], B[k q][col c] = (X D^-1)[16 J + c][q]
            double aop[4], bop[4];
#pragma unroll
            for (int I = 0; I < 4; I++) {
                aop[I] = -sXr[(16 * I + c) * 4 + q];
                bop[I] = sX[(16 * I + c) * 4 + q];
            }
            // the column block K in registers becomes L[:, K] below the block:
            // row 16 I + q + 4 r lies below K = 16 Ib + 4 rr .. + 3 exactly
            // when I > Ib, or I == Ib and r > rr (q < 4), a condition known at
            // compile time, so the loads write the registers under the lane
            // mask alone
            if (c >= cb && c < cb + 4) {
                const int rr = s & 3;
#pragma unroll
                for (int I = 0; I < 4; I++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (I > Ib || (I == Ib && r > rr))
                            A[I][Ib][r] = sX[(16 * I + q + 4 * r) * 4 + (c - cb)];
            }
#pragma unroll
            for (int I = 0; I < 4; I++)
#pragma unroll
                for (int J = 0; J < 4; J++) {
                    // (tiles wholly in the identity padding add X = 0)
                    if (J < Ib || I < J || 16 * I >= E) continue;
                    A[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(aop[I], bop[J], A[I][J], 0, 0, 0);
                }
        }
    }
    // back substitution L^T w = u, blocks from the last
#pragma unroll
    for (int s = 15; s >= 0; s--) {
        if (s < S) {
            const int Ib = s >> 2, rr = s & 3;
            const int k0 = 4 * s;
            __syncthreads();
            // row 4s + q of L is register rr of tiles (Ib, J), columns 16 J + c
#pragma unroll
            for (int J = 0; J < 4; J++)
                if (J <= Ib) sR[q * 64 + 16 * J + c] = A[Ib][J][rr];
            if (lane >= k0 && lane < k0 + 4) sK[lane - k0] = rhs;
            __syncthreads();
            double w[4];
#pragma unroll
            for (int j = 3; j >= 0; j--) {
                double v = sK[j];
#pragma unroll
                for (int p = j + 1; p < 4; p++) v = fma(-sL[s * 16 + p * 4 + j], w[p], v);
                w[j] = v;
            }
            if (lane < k0) {
#pragma unroll
                for (int j = 0; j < 4; j++) rhs = fma(-sR[j * 64 + lane], w[j], rhs);
            } else if (lane < k0 + 4) {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (lane - k0 == j) rhs = w[j];
            }
        }
    }
    if (lane < E) a.W[ui * E + lane] = rhs;
    if (bad && lane == 0) atomicOr(&a.status[u], 8);
}

hipError_t launch_mllr_solve(const MllrSolveArgs &a, hipStream_t s) {
    const int E = a.D + 1;
    const dim3 grid((unsigned)((uint64_t)a.U * a.D));
    if (E <= 64) {
        hipLaunchKernelGGL(mllr_solve_ldl_kernel, grid, dim3(64), 0, s, a);
    } else {
        const size_t shm = ((size_t)E * (E + 1) + E) * sizeof(double);
        hipLaunchKernelGGL(mllr_solve_kernel, grid, dim3(64), shm, s, a);
    }
    return hipGetLastError();
}

}  // namespace gmmhip
