// Full-covariance E-step and statistics (the reference's gammasKernelFull +
// logLikeKernel + gammasNormKernel + accStatsFullKernel,
// trainer/CU_GMMStatsEstimator_CUDAkernels.cuh:836-917, :922-995,
// :1037-1110, :459-576; host side insertModel's full branch,
// trainer/OL_GMMStatsEstimator.cpp:156-229, :291-299).
//
// With centred frames x' = x - c and the feature map
//     phi(x') = [1, x'_1..x'_D, x'_i x'_j (i <= j, packed upper order)]
// of K = 1 + D + D(D+1)/2 entries, both halves of the E-step are dense
// contractions over K:
//     l[t][m]   = phi(x'_t) . theta_m
//       theta_m = [log w + gconst - mu'^T P mu' / 2, (P mu')_d, -P_ii/2 (i == j), -P_ij (i < j)]
//     stats_m   = sum_t g_tm phi(x'_t) = [n, sum g x', sum g x'_i x'_j]
// so a frame chunk runs: extended rows [1, x'] (fc_rows_kernel), L = Phi
// Theta^T with phi formed from those rows inside the MFMA kernel
// (fc_ll_kernel), one fused pass per frame computing the log-sum-exp, the
// clamp rules and the cut posteriors in place, and the statistics
// Gamma^T Phi the same way (fc_stats_kernel).  The statistics have a small
// output (K x M) and a long contraction (the chunk's frames), so they run
// over S frame sub-chunks to fill the chip; the S partials and the chunks
// are summed into FP64 accumulators in a fixed order.
#include "gemm_path.h"

namespace gmmhip {

// One wave per frame: log-sum-exp of the frame's M component
// log-likelihoods (exact, max-shifted), the clamp rules of logLikeKernel
// (LL <= min_ll or non-finite -> min_ll, posterior 0), posteriors
// g = exp(l - LL) (with cut: set to 0 when g <= min_gamma), written in place.  LLnat[n0 + t]
// (optional) receives the frame's log-likelihood; blockLL[block] the FP64
// sum of the block's four frames.
__global__ __launch_bounds__(256) void fc_lse_kernel(float *__restrict__ L, int nc, int ncp, int M,
                                                     const uint8_t *__restrict__ bad,
                                                     float min_ll, float min_gamma, int cut,
                                                     float *__restrict__ LLnat, uint64_t n0,
                                                     double *__restrict__ blockLL) {
    __shared__ double red[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.x * 4 + wave;
    double lld = 0.0;
    if (t >= nc && t < ncp) {
        // padding frames: posterior 0
        for (int m = lane; m < M; m += 64) L[(size_t)t * M + m] = 0.0f;
    }
    if (t < nc) {
        float *row = L + (size_t)t * M;
        float mx = -INFINITY;
        bool nonfinite = false;
        for (int m = lane; m < M; m += 64) {
            const float v = row[m];
            nonfinite |= !isfinite(v);
            mx = fmaxf(mx, v);
        }
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        float s = 0.0f;
        for (int m = lane; m < M; m += 64) s += expf(row[m] - mx);
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        const bool anynf = __builtin_amdgcn_ballot_w64(nonfinite) != 0;
        float ll = mx + logf(s);
        const bool clamp = bad[t] || anynf || !(ll > min_ll) || !isfinite(ll);
        if (clamp) ll = min_ll;
        for (int m = lane; m < M; m += 64) {
            float g = 0.0f;
            if (!clamp) {
                g = (M == 1) ? 1.0f : expf(row[m] - ll);
                if (cut) g = (g > min_gamma) ? g : 0.0f;
            }
            row[m] = g;
        }
        if (lane == 0 && LLnat) LLnat[n0 + t] = ll;
        lld = (double)ll;
    }
    if (lane == 0) red[wave] = lld;
    __syncthreads();
    if (threadIdx.x == 0) blockLL[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Gaussian selection: log-likelihoods of the components outside frame
// n0 + t's selection (bit m of its row of the bitmap, W 32-bit words per
// frame) become -3e38, so they vanish from the log-sum-exp and get posterior
// 0; a frame with an empty selection is clamped.  One block per frame.
__global__ __launch_bounds__(256) void fc_mask_kernel(float *__restrict__ L, int nc, int M,
                                                      const uint32_t *__restrict__ bits, int W,
                                                      uint64_t n0) {
    const int t = blockIdx.x;
    if (t >= nc) return;
    const uint32_t *row = bits + (n0 + t) * (uint64_t)W;
    float *l = L + (size_t)t * M;
    for (int m = threadIdx.x; m < M; m += 256)
        if (!((row[m >> 5] >> (m & 31)) & 1u)) l[m] = -3e38f;
}

// acc[e] += sum_s part[s][e] (FP64, s in order), one element per thread:
// the fixed-order reduction of the FP32 sub-chunk statistics.
__global__ __launch_bounds__(256) void fc_acc_kernel(const float *__restrict__ part, size_t E,
                                                     int S, int first, double *__restrict__ acc) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= E) return;
    double v = first ? 0.0 : acc[e];
    for (int s = 0; s < S; s++) v += (double)part[(size_t)s * E + e];
    acc[e] = v;
}

}  // namespace gmmhip

using namespace gmmhip_detail;

namespace gmmhip_detail {

std::vector<int> feature_codes(int D, bool full) {
    // feature order of the reference's full second-order statistics
    // (trainer/OL_GMMStatsEstimator.cpp:1358-1364) for full models
    std::vector<int> p;
    p.push_back(-1);
    for (int d = 0; d < D; d++) p.push_back(-2 - d);
    for (int i = 0; i < D; i++)
        for (int j = i; j < D; j++)
            if (full || j == i) p.push_back((i << 16) | j);
    return p;
}

int gemm_state(gmmhip_ctx *h, gmmhip::FullState **out) {
    if (!h->full) {
        h->full = new gmmhip::FullState();
    }
    *out = h->full;
    return GMMHIP_OK;
}

hipError_t launch_fc_lse(float *L, int nc, int ncp, int M, const uint8_t *bad, float min_ll,
                         float min_gamma, int cut, float *LLnat, uint64_t n0, double *blockLL,
                         hipStream_t s) {
    hipLaunchKernelGGL(gmmhip::fc_lse_kernel, dim3((ncp + 3) / 4), dim3(256), 0, s, L, nc, ncp, M,
                       bad, min_ll, min_gamma, cut, LLnat, n0, blockLL);
    return hipGetLastError();
}

hipError_t launch_fc_acc(const float *part, size_t E, int S, int first, double *acc,
                         hipStream_t s) {
    hipLaunchKernelGGL(gmmhip::fc_acc_kernel, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, s,
                       part, E, S, first, acc);
    return hipGetLastError();
}

}  // namespace gmmhip_detail

namespace {

// Frames per chunk: the extended rows and the log-likelihood rows of a chunk
// take at most ~512 MB;
// opts.chunk_frames (at least 256) overrides.
uint64_t full_chunk(const gmmhip_ctx *h) {
    uint64_t nc;
    if (h->opt.chunk_frames > 0) {
        nc = std::max<uint64_t>(256, (uint64_t)h->opt.chunk_frames);
    } else {
        const uint64_t per = 4ull * (uint64_t)(ext_row_floats(h->D) + h->M) + 1;
        nc = std::max<uint64_t>(1024, (512ull << 20) / per / 256 * 256);
    }
    return std::min<uint64_t>(nc, h->N);
}

// One pass over the resident frames: per chunk features, L = Phi Theta^T
// (Gaussian selection applied when a mask is set), fused log-sum-exp /
// posteriors, and (stats) Gamma^T Phi into the FP64 accumulators.  LLnat
// (device, optional) receives per-frame log-likelihoods; *tot the FP64
// total; post_out (host, optional, [N][M]) the posteriors before the cut.
int full_pass(gmmhip_ctx *h, bool stats, float *LLnat, double *tot, float *post_out = nullptr) {
    const int post_cut = post_out ? 0 : 1;
    gmmhip::FullState *fs = h->full;
    const int M = h->M, D = h->D, KP = fs->KP, DXP = ext_row_floats(D);
    const uint64_t nc_max = full_chunk(h);
    const int S = stats_subchunks(M, KP, h->num_cus, nc_max);
    const uint64_t ncp_max = (uint64_t)stats_sub(nc_max, S) * S;
    HIPCHK(h, ensure(&fs->d_xt, &fs->xt_cap, (size_t)ncp_max * DXP));
    HIPCHK(h, ensure(&fs->d_L, &fs->L_cap, (size_t)ncp_max * M));
    HIPCHK(h, ensure(&fs->d_bad, &fs->bad_cap, (size_t)ncp_max));
    const uint64_t nblk_total = (h->N + nc_max - 1) / nc_max * ((ncp_max + 3) / 4);
    HIPCHK(h, ensure(&h->d_blockLL, &h->blockLL_cap, (size_t)nblk_total));
    if (stats) {
        HIPCHK(h, ensure(&fs->d_part, &fs->part_cap, (size_t)S * M * KP));
        HIPCHK(h, ensure(&fs->d_acc, &fs->acc_cap, (size_t)M * KP));
    }
    uint64_t nblk = 0;
    int first = 1;
    for (uint64_t n0 = 0; n0 < h->N; n0 += nc_max) {
        const int nc = (int)std::min<uint64_t>(nc_max, h->N - n0);
        const int sub = stats_sub(nc, S);  // frames per sub-chunk
        const int ncp = sub * S;
        HIPCHK(h, launch_fc_rows(h->d_x, n0, nc, ncp, D, fs->d_cen, fs->d_xt, fs->d_bad, h->stream));
        // L [nc][M] = Phi Theta^T
        HIPCHK(h, launch_fc_ll(fs->d_xt, nc, D, fs->d_theta, M, KP, fs->d_pairs, fs->d_L, h->stream));
        if (h->d_mask) {
            hipLaunchKernelGGL(gmmhip::fc_mask_kernel, dim3(nc), dim3(256), 0, h->stream, fs->d_L,
                               nc, M, h->d_mask, h->mask_W, n0);
            HIPCHK(h, hipGetLastError());
        }
        const int blocks = (ncp + 3) / 4;
        HIPCHK(h, launch_fc_lse(fs->d_L, nc, ncp, M, fs->d_bad, h->opt.min_loglike,
                                h->opt.min_gamma, post_cut, LLnat, n0, h->d_blockLL + nblk,
                                h->stream));
        if (post_out)
            HIPCHK(h, hipMemcpyAsync(post_out + n0 * (uint64_t)M, fs->d_L, (size_t)nc * M * sizeof(float),
                                     hipMemcpyDeviceToHost, h->stream));
        nblk += blocks;
        if (stats) {
            // part[s] [M][KP] = Gamma_s^T Phi_s over frames [s sub, (s+1) sub)
            HIPCHK(h, launch_fc_stats(fs->d_xt, DXP, fs->d_L, M, KP, fs->d_pairs, sub, S,
                                      fs->d_part, h->stream));
            const size_t E = (size_t)M * KP;
            HIPCHK(h, launch_fc_acc(fs->d_part, E, S, first, fs->d_acc, h->stream));
            first = 0;
        }
    }
    HIPCHK(h, launch_sum_doubles(h->d_blockLL, nblk, h->d_tot, h->stream));
    HIPCHK(h, hipMemcpyAsync(tot, h->d_tot, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    return GMMHIP_OK;
}

}  // namespace

namespace gmmhip_detail {

void full_state_free(gmmhip::FullState *fs) {
    if (!fs) return;
    hipFree(fs->d_theta);
    hipFree(fs->d_pairs);
    hipFree(fs->d_cen);
    hipFree(fs->d_xt);
    hipFree(fs->d_L);
    hipFree(fs->d_bad);
    hipFree(fs->d_part);
    hipFree(fs->d_acc);
    delete fs;
}

int mask_check(gmmhip_ctx *h) {
    if (!h->d_mask) return GMMHIP_OK;
    if (h->mask_N != h->N || h->mask_M != h->M)
        return fail(h, GMMHIP_ERR_ARG,
                    "Gaussian-selection mask set for %llu frames x %d components, but the handle "
                    "has %llu frames x %d components", (unsigned long long)h->mask_N, h->mask_M,
                    (unsigned long long)h->N, h->M);
    return GMMHIP_OK;
}

int diag_operand(gmmhip_ctx *h, bool full_features) {
    gmmhip::FullState *fs = nullptr;
    int rc = gemm_state(h, &fs);
    if (rc) return rc;
    const int kind = full_features ? gmmhip::OP_DIAG_FULLFEAT : gmmhip::OP_DIAG;
    if (fs->operand_version == h->model_version && fs->operand_kind == kind) return GMMHIP_OK;
    const int M = h->M, D = h->D;
    const int K = full_features ? 1 + D + D * (D + 1) / 2 : 1 + 2 * D, KP = feature_pad(K);
    std::vector<float> theta((size_t)M * KP, 0.0f);
    for (int m = 0; m < M; m++) {
        float *th = theta.data() + (size_t)m * KP;
        double cq = 0.0;
        int k = 1 + D;
        for (int i = 0; i < D; i++) {
            const double iv = 1.0 / (double)h->m_var[(size_t)m * D + i];
            const double mc = (double)h->m_mu[(size_t)m * D + i] - (double)h->h_off[i];
            th[1 + i] = (float)(mc * iv);
            th[k] = (float)(-0.5 * iv);  // x'_i^2; with full features (i, j > i) stay 0
            k += full_features ? D - i : 1;
            cq += mc * mc * iv;
        }
        th[0] = (float)(h->m_gconst[m] - 0.5 * cq);
    }
    const std::vector<uint32_t> pairs = feature_pairs(feature_codes(D, full_features), D, KP);
    HIPCHK(h, ensure(&fs->d_theta, &fs->theta_cap, theta.size()));
    HIPCHK(h, ensure(&fs->d_pairs, &fs->pairs_cap, pairs.size()));
    HIPCHK(h, ensure(&fs->d_cen, &fs->cen_cap, (size_t)D));
    HIPCHK(h, hipMemcpyAsync(fs->d_theta, theta.data(), theta.size() * sizeof(float),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(fs->d_pairs, pairs.data(), pairs.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(fs->d_cen, h->h_off.data(), D * sizeof(float), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    fs->K = K;
    fs->KP = KP;
    fs->h_cen = h->h_off;
    fs->operand_version = h->model_version;
    fs->operand_kind = kind;
    return GMMHIP_OK;
}

int diag_full_accumulate(gmmhip_ctx *h, int flags, double *n, double *F, double *S,
                         double *tot_ll, uint64_t *nframes) {
    int rc = diag_operand(h, true);
    if (rc) return rc;
    return full_accumulate(h, flags, n, F, S, tot_ll, nframes);
}

int gemm_operand_for_model(gmmhip_ctx *h) {
    // a full model's operand is built by gmmhip_set_model_full; a diagonal
    // model gets its [1, x', x'^2] operand
    return h->full_model ? GMMHIP_OK : diag_operand(h, false);
}

int gemm_loglikes_device(gmmhip_ctx *h, float *d_out) {
    int rc = gemm_operand_for_model(h);
    if (rc) return rc;
    double tot = 0.0;
    rc = full_pass(h, false, d_out, &tot);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return GMMHIP_OK;
}

int gemm_posteriors(gmmhip_ctx *h, float *ll, float *post) {
    int rc = gemm_operand_for_model(h);
    if (rc) return rc;
    HIPCHK(h, ensure(&h->d_LLnat, &h->LLnat_cap, (size_t)h->N));
    double tot = 0.0;
    rc = full_pass(h, false, h->d_LLnat, &tot, post);
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(ll, h->d_LLnat, h->N * sizeof(float), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return GMMHIP_OK;
}

int full_accumulate(gmmhip_ctx *h, int flags, double *n, double *F, double *S, double *tot_ll,
                    uint64_t *nframes) {
    gmmhip::FullState *fs = h->full;
    const int M = h->M, D = h->D, KP = fs->KP;
    double tot = 0.0;
    int rc = full_pass(h, true, nullptr, &tot);
    if (rc) return rc;
    std::vector<double> acc((size_t)M * KP);
    HIPCHK(h, hipMemcpyAsync(acc.data(), fs->d_acc, acc.size() * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (std::isnan(tot))
        return fail(h, GMMHIP_ERR_NAN, "accumulateStatsMT(): overall log-likelihood = NaN!");
    // Undo the centring in FP64: x = x' + c ->
    //   F_i = F'_i + n c_i,  S_ij = S'_ij + c_i F'_j + c_j F'_i + n c_i c_j.
    const size_t DP = (size_t)D * (D + 1) / 2;
    const bool diag = fs->operand_kind == gmmhip::OP_DIAG;
    for (int m = 0; m < M && diag; m++) {
        const double *r = acc.data() + (size_t)m * KP;
        const double nm = r[0];
        n[m] += nm;
        for (int d = 0; d < D; d++) {
            const double c = fs->h_cen[d], f1 = r[1 + d];
            if (flags & GMMHIP_ACC_MEAN) F[(size_t)m * D + d] += f1 + nm * c;
            if (flags & GMMHIP_ACC_VAR) S[(size_t)m * D + d] += r[1 + D + d] + 2.0 * c * f1 + nm * c * c;
        }
    }
    for (int m = 0; m < M && !diag; m++) {
        const double *r = acc.data() + (size_t)m * KP;
        const double nm = r[0];
        const double *f1 = r + 1, *s1 = r + 1 + D;
        n[m] += nm;
        if (flags & GMMHIP_ACC_MEAN)
            for (int d = 0; d < D; d++) F[(size_t)m * D + d] += f1[d] + nm * (double)fs->h_cen[d];
        if (flags & GMMHIP_ACC_VAR) {
            size_t p = 0;
            for (int i = 0; i < D; i++) {
                const double ci = fs->h_cen[i];
                for (int j = i; j < D; j++, p++) {
                    const double cj = fs->h_cen[j];
                    S[(size_t)m * DP + p] += s1[p] + ci * f1[j] + cj * f1[i] + nm * ci * cj;
                }
            }
        }
    }
    *tot_ll += tot;
    *nframes += h->N;
    return GMMHIP_OK;
}

}  // namespace gmmhip_detail

// Full-covariance counterpart of gmmhip_set_model: the inverse covariance
// and log-determinant by an FP64 Cholesky (getCovInverse / getLogDet use
// FP32 LAPACK spotrf/spotri), then the feature-space operand theta in FP32.
int gmmhip_set_model_full(gmmhip_t h, int M, int D, const float *w, const float *mu,
                          const float *cov) {
    if (!h) return GMMHIP_ERR_ARG;
    if (M <= 0 || D <= 0 || D > 127 || !w || !mu || !cov)
        return fail(h, GMMHIP_ERR_ARG, "set_model_full: bad arguments (M=%d, D=%d)", M, D);
    for (int m = 0; m < M; m++)
        if (!(w[m] > 0.0f))
            return fail(h, GMMHIP_ERR_ARG,
                        "set_model_full: component %d has weight %g; insert only the active "
                        "(weight > 0) components in RearrangeMixtures order", m, (double)w[m]);
    gmmhip::FullState *fs = nullptr;
    int rc = gemm_state(h, &fs);
    if (rc) return rc;
    const int K = 1 + D + D * (D + 1) / 2;
    const int KP = feature_pad(K);
    // centre: the weighted mean of the component means
    std::vector<double> c(D, 0.0);
    double ws = 0.0;
    for (int m = 0; m < M; m++) ws += w[m];
    for (int m = 0; m < M; m++)
        for (int d = 0; d < D; d++) c[d] += (double)w[m] * mu[(size_t)m * D + d] / ws;
    std::vector<float> cf(D);
    for (int d = 0; d < D; d++) cf[d] = (float)c[d];

    const double minus_log_2pi = -1.837877066409345;
    std::vector<float> theta((size_t)M * KP, 0.0f);
    std::vector<double> L((size_t)D * D), Li((size_t)D * D), P((size_t)D * D), mc(D), Pm(D);
    for (int m = 0; m < M; m++) {
        const float *cm = cov + (size_t)m * D * D;
        for (int i = 0; i < D * D; i++) L[i] = cm[i];
        // lower Cholesky, row-major
        for (int j = 0; j < D; j++) {
            double s = L[(size_t)j * D + j];
            for (int k = 0; k < j; k++) s -= L[(size_t)j * D + k] * L[(size_t)j * D + k];
            if (!(s > 0.0))
                return fail(h, GMMHIP_ERR_VAR,
                            "getCovInverse(): cholesky factorization failed! (component %d)", m);
            const double ljj = std::sqrt(s);
            L[(size_t)j * D + j] = ljj;
            for (int i = j + 1; i < D; i++) {
                double t = L[(size_t)i * D + j];
                for (int k = 0; k < j; k++) t -= L[(size_t)i * D + k] * L[(size_t)j * D + k];
                L[(size_t)i * D + j] = t / ljj;
            }
        }
        double logdet = 0.0;
        for (int i = 0; i < D; i++) logdet += std::log(L[(size_t)i * D + i]);
        logdet *= 2.0;
        std::fill(Li.begin(), Li.end(), 0.0);
        for (int j = 0; j < D; j++) {
            Li[(size_t)j * D + j] = 1.0 / L[(size_t)j * D + j];
            for (int i = j + 1; i < D; i++) {
                double s = 0.0;
                for (int k = j; k < i; k++) s -= L[(size_t)i * D + k] * Li[(size_t)k * D + j];
                Li[(size_t)i * D + j] = s / L[(size_t)i * D + i];
            }
        }
        for (int i = 0; i < D; i++)
            for (int j = i; j < D; j++) {
                double s = 0.0;
                for (int k = std::max(i, j); k < D; k++) s += Li[(size_t)k * D + i] * Li[(size_t)k * D + j];
                P[(size_t)i * D + j] = P[(size_t)j * D + i] = s;
            }
        for (int d = 0; d < D; d++) mc[d] = (double)mu[(size_t)m * D + d] - (double)cf[d];
        double q = 0.0;
        for (int i = 0; i < D; i++) {
            double s = 0.0;
            for (int j = 0; j < D; j++) s += P[(size_t)i * D + j] * mc[j];
            Pm[i] = s;
            q += mc[i] * s;
        }
        float *th = theta.data() + (size_t)m * KP;
        th[0] = (float)(std::log((double)w[m]) + minus_log_2pi * D / 2.0 - logdet / 2.0 - 0.5 * q);
        for (int d = 0; d < D; d++) th[1 + d] = (float)Pm[d];
        int k = 1 + D;
        for (int i = 0; i < D; i++)
            for (int j = i; j < D; j++, k++)
                th[k] = (float)((i == j ? -0.5 : -1.0) * P[(size_t)i * D + j]);
    }
    const std::vector<uint32_t> pairs = feature_pairs(feature_codes(D, true), D, KP);
    DevGuard g(h->dev);
    HIPCHK(h, ensure(&fs->d_theta, &fs->theta_cap, theta.size()));
    HIPCHK(h, ensure(&fs->d_pairs, &fs->pairs_cap, pairs.size()));
    HIPCHK(h, ensure(&fs->d_cen, &fs->cen_cap, (size_t)D));
    HIPCHK(h, hipMemcpyAsync(fs->d_theta, theta.data(), theta.size() * sizeof(float),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(fs->d_pairs, pairs.data(), pairs.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(fs->d_cen, cf.data(), D * sizeof(float), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    fs->K = K;
    fs->KP = KP;
    fs->h_cen = cf;
    h->M = M;
    h->D = D;
    h->full_model = true;
    h->m_w.assign(w, w + M);
    h->m_mu.assign(mu, mu + (size_t)M * D);
    h->m_var.clear();
    h->model_version++;
    fs->operand_version = h->model_version;
    fs->operand_kind = gmmhip::OP_FULL;
    return GMMHIP_OK;
}

// Gaussian selection (the reference's logLikeMask in "use" mode,
// trainer/OL_GMMStatsEstimator.h:73-114, :1157-1163: getLikeAndGammas and
// accumulateStats restricted to each frame's component list).
int gmmhip_set_mask(gmmhip_t h, uint64_t N, const uint64_t *offsets, const int *idx) {
    if (!h) return GMMHIP_ERR_ARG;
    DevGuard g(h->dev);
    if (!offsets) {
        h->mask_N = 0;
        h->mask_M = 0;
        if (h->d_mask) hipFree(h->d_mask);
        h->d_mask = nullptr;
        h->mask_cap = 0;
        return GMMHIP_OK;
    }
    if (h->M <= 0) return fail(h, GMMHIP_ERR_STATE, "set_mask: no model inserted");
    if (N == 0 || (offsets[N] > 0 && !idx))
        return fail(h, GMMHIP_ERR_ARG, "set_mask: bad arguments");
    const int W = (h->M + 31) / 32;
    std::vector<uint32_t> bits((size_t)N * W, 0u);
    for (uint64_t t = 0; t < N; t++) {
        if (offsets[t + 1] < offsets[t])
            return fail(h, GMMHIP_ERR_ARG, "set_mask: offsets decrease at frame %llu",
                        (unsigned long long)t);
        for (uint64_t e = offsets[t]; e < offsets[t + 1]; e++) {
            const int m = idx[e];
            if (m < 0 || m >= h->M)
                return fail(h, GMMHIP_ERR_ARG, "set_mask: component %d of frame %llu out of range",
                            m, (unsigned long long)t);
            bits[(size_t)t * W + (m >> 5)] |= 1u << (m & 31);
        }
    }
    HIPCHK(h, ensure(&h->d_mask, &h->mask_cap, bits.size()));
    HIPCHK(h, hipMemcpyAsync(h->d_mask, bits.data(), bits.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->mask_N = N;
    h->mask_W = W;
    h->mask_M = h->M;
    return GMMHIP_OK;
}

// The reference's logLikeMask in "fill" mode (mask2fill,
// trainer/OL_GMMStatsEstimator.cpp:1168-1171): per-frame log-likelihoods and
// posteriors before the cut; a clamped frame gets min_gamma in every
// component (compGammas, :540-546; 1 for a single-component model).
int gmmhip_posteriors(gmmhip_t h, float *ll, float *post) {
    if (!h) return GMMHIP_ERR_ARG;
    if (h->M <= 0) return fail(h, GMMHIP_ERR_STATE, "no model inserted");
    if (!h->d_x || h->N == 0) return fail(h, GMMHIP_ERR_STATE, "no frames inserted");
    if (h->xD != h->D)
        return fail(h, GMMHIP_ERR_ARG, "frame dimension %d != model dimension %d", h->xD, h->D);
    if (!ll || !post) return fail(h, GMMHIP_ERR_ARG, "posteriors: null output");
    int rc = mask_check(h);
    if (rc) return rc;
    DevGuard g(h->dev);
    rc = gemm_posteriors(h, ll, post);
    if (rc) return rc;
    const int M = h->M;
    for (uint64_t t = 0; t < h->N; t++)
        if (!(ll[t] > h->opt.min_loglike))
            std::fill(post + t * (uint64_t)M, post + (t + 1) * (uint64_t)M,
                      M == 1 ? 1.0f : h->opt.min_gamma);
    return GMMHIP_OK;
}
