// Robust auxiliary statistics (the reference's accumulateAuxStats,
// trainer/OL_GMMStatsEstimator.cpp:1226-1280, driven by accumulateStats at
// :1174-1189; on its GPU: normAuxGammasKernel + accAuxStatsKernel,
// trainer/CU_GMMStatsEstimator_CUDAkernels.cuh:1157-1224, :310-453).
//
// For every consecutive frame pair (t-1, t) of the resident frames, with the
// posteriors g before the cut:
//     p_m = g_t,m g_t-1,m,   s = sum_m [p_m > cut] p_m,   h_m = p_m / s
//     aux2_m += sqrt(g_t,m)                       (g_t,m > cut)
//     aux_m  += h_m |x_t - x_t-1|,  aux3_m += h_m  (h_m > cut)
// The normalisation s needs every component of a pair, so the posteriors of
// a frame chunk are formed as a whole: the GEMM form of the E-step
// (gemm_path.h) gives L = Theta^T phi and the uncut posteriors in place; one
// wave per frame then forms the pair weights H = [h | sqrt-cut g] (2M per
// frame) and the pair features psi = [1, |x_t - x_t-1|]; the sums H^T psi
// run on the statistics kernel of the GEMM form (features psi_d = psi[d] *
// psi[0]) over frame sub-chunks, into FP64 in a fixed order.  The
// posteriors of a chunk's last frame are carried to the next chunk, so
// pairs never break at chunk boundaries.
#include "gemm_path.h"

namespace gmmhip {

struct AuxState {
    uint64_t version = ~0ull;    // model_version of the diagonal GEMM operand
    int K = 0, KP = 0;
    float *d_theta = nullptr;    // [M][KP] diagonal models: [const, mu' ivar, -ivar/2]
    size_t theta_cap = 0;
    uint32_t *d_pairs = nullptr; // [KP] feature pairs of the diagonal operand
    size_t pairs_cap = 0;
    uint32_t *d_psipairs = nullptr;  // [DP1] (d, 0): psi_d = psi[d] * psi[0]
    size_t psipairs_cap = 0;
    float *d_cen = nullptr;      // [D]
    size_t cen_cap = 0;
    float *d_H = nullptr;        // [ncp][2M] pair weights | sqrt-cut posteriors
    size_t H_cap = 0;
    float *d_psi = nullptr;      // [ncp][DP1] [1, |x_t - x_t-1|]
    size_t psi_cap = 0;
    float *d_carry = nullptr;    // [M] posteriors of the previous chunk's last frame
    size_t carry_cap = 0;
    float *d_part = nullptr;     // [S][2M][DP1]
    size_t part_cap = 0;
    double *d_acc = nullptr;     // [2M][DP1]
    size_t acc_cap = 0;
};

// One wave per frame t = n0 + i of the chunk: pair weights against frame
// t-1 (row i-1, or the carried row for i = 0; none for t = 0), the
// sqrt-cut posteriors, and psi.  A pair with s = 0 (no product above the
// cut, e.g. a clamped or non-finite frame on either side) has zero pair
// weights and zero differences in psi, so non-finite frame values never
// reach the GEMM.
__global__ __launch_bounds__(256) void aux_pair_kernel(const float *__restrict__ G,
                                                       const float *__restrict__ carry,
                                                       const float *__restrict__ X, uint64_t n0,
                                                       int nc, int ncp, int M, int D, int DP1,
                                                       float min_gamma, float *__restrict__ H,
                                                       float *__restrict__ psi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wave;
    if (i >= ncp) return;
    float *hrow = H + (size_t)i * 2 * M;
    float *prow = psi + (size_t)i * DP1;
    const bool has_pair = i < nc && (n0 + i) > 0;
    if (!has_pair) {
        // padding rows, and the first frame of the frame set: the reference
        // starts the auxiliary statistics (aux2 included) at a run's second
        // frame
        for (int m = lane; m < 2 * M; m += 64) hrow[m] = 0.0f;
        for (int k = lane; k < DP1; k += 64) prow[k] = 0.0f;
        return;
    }
    const float *g2 = G + (size_t)i * M;
    const float *g1 = (i == 0) ? carry : G + (size_t)(i - 1) * M;
    double s = 0.0;
    for (int m = lane; m < M; m += 64) {
        const float p = g2[m] * g1[m];
        if (p > min_gamma) s += (double)p;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const double gs = s > 0.0 ? 1.0 / s : 0.0;
    for (int m = lane; m < M; m += 64) {
        const float a = g2[m];
        const double hv = gs * (double)a * (double)g1[m];
        hrow[m] = hv > (double)min_gamma ? (float)hv : 0.0f;
        hrow[M + m] = a > min_gamma ? sqrtf(a) : 0.0f;
    }
    const float *x2 = X + (n0 + i) * (uint64_t)D;
    const float *x1 = x2 - D;
    // column 0 carries aux3 (pair weights) and aux2 (this frame's sqrt-cut
    // posteriors, counted whatever the pair's weights); the differences are
    // only formed when some pair weight is non-zero
    for (int k = lane; k < DP1; k += 64) {
        float v = 0.0f;
        if (k == 0) v = 1.0f;
        else if (k <= D && s > 0.0) v = fabsf(x2[k - 1] - x1[k - 1]);
        prow[k] = v;
    }
}

}  // namespace gmmhip

using namespace gmmhip_detail;

namespace {

// The GEMM-form operand of a diagonal model (built from the handle's host
// copy, FP64, rounded once): theta_m = [log w + gconst - sum mu'^2 ivar / 2,
// mu' ivar, -ivar / 2] over features [1, x', x'^2], x' = x - c.
int aux_diag_operand(gmmhip_ctx *h, gmmhip::AuxState *a) {
    if (a->version == h->model_version) return GMMHIP_OK;
    const int M = h->M, D = h->D;
    const int K = 1 + 2 * D, KP = feature_pad(K);
    std::vector<float> theta((size_t)M * KP, 0.0f);
    for (int m = 0; m < M; m++) {
        float *th = theta.data() + (size_t)m * KP;
        double cq = 0.0;
        for (int d = 0; d < D; d++) {
            const double iv = 1.0 / (double)h->m_var[(size_t)m * D + d];
            const double mc = (double)h->m_mu[(size_t)m * D + d] - (double)h->h_off[d];
            th[1 + d] = (float)(mc * iv);
            th[1 + D + d] = (float)(-0.5 * iv);
            cq += mc * mc * iv;
        }
        th[0] = (float)(h->m_gconst[m] - 0.5 * cq);
    }
    const std::vector<uint32_t> pairs = feature_pairs(feature_codes(D, false), D, KP);
    HIPCHK(h, ensure(&a->d_theta, &a->theta_cap, theta.size()));
    HIPCHK(h, ensure(&a->d_pairs, &a->pairs_cap, pairs.size()));
    HIPCHK(h, ensure(&a->d_cen, &a->cen_cap, (size_t)D));
    HIPCHK(h, hipMemcpyAsync(a->d_theta, theta.data(), theta.size() * sizeof(float),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(a->d_pairs, pairs.data(), pairs.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(a->d_cen, h->h_off.data(), D * sizeof(float), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    a->K = K;
    a->KP = KP;
    a->version = h->model_version;
    return GMMHIP_OK;
}

}  // namespace

namespace gmmhip_detail {

void aux_state_free(gmmhip::AuxState *a) {
    if (!a) return;
    hipFree(a->d_theta);
    hipFree(a->d_pairs);
    hipFree(a->d_psipairs);
    hipFree(a->d_cen);
    hipFree(a->d_H);
    hipFree(a->d_psi);
    hipFree(a->d_carry);
    hipFree(a->d_part);
    hipFree(a->d_acc);
    delete a;
}

}  // namespace gmmhip_detail

int gmmhip_accumulate_aux(gmmhip_t h, double *aux) {
    if (!h) return GMMHIP_ERR_ARG;
    if (!aux) return fail(h, GMMHIP_ERR_ARG, "accumulate_aux: null output");
    if (h->M <= 0) return fail(h, GMMHIP_ERR_STATE, "no model inserted");
    if (!h->d_x || h->N == 0) return fail(h, GMMHIP_ERR_STATE, "no frames inserted");
    if (h->xD != h->D)
        return fail(h, GMMHIP_ERR_ARG, "frame dimension %d != model dimension %d", h->xD, h->D);
    DevGuard g(h->dev);
    gmmhip::FullState *fs = nullptr;
    int rc = gemm_state(h, &fs);
    if (rc) return rc;
    if (!h->aux) h->aux = new gmmhip::AuxState();
    gmmhip::AuxState *a = h->aux;
    const int M = h->M, D = h->D, DP1 = (D + 1 + 7) / 8 * 8;
    int KP;
    const float *theta, *cen;
    const uint32_t *pairs;
    if (h->full_model) {
        KP = fs->KP, theta = fs->d_theta, cen = fs->d_cen, pairs = fs->d_pairs;
    } else {
        rc = aux_diag_operand(h, a);
        if (rc) return rc;
        KP = a->KP, theta = a->d_theta, cen = a->d_cen, pairs = a->d_pairs;
    }
    if (!a->d_psipairs || a->psipairs_cap < (size_t)DP1) {
        std::vector<uint32_t> pp(DP1);
        for (int d = 0; d < DP1; d++) pp[d] = (uint32_t)d;  // (d, 0)
        HIPCHK(h, ensure(&a->d_psipairs, &a->psipairs_cap, (size_t)DP1));
        HIPCHK(h, hipMemcpyAsync(a->d_psipairs, pp.data(), pp.size() * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, h->stream));
    }
    const int DXP = ext_row_floats(D);
    // chunk: extended rows, L, H and psi rows of a chunk within ~768 MB
    uint64_t nc_max;
    if (h->opt.chunk_frames > 0) {
        nc_max = std::max<uint64_t>(256, (uint64_t)h->opt.chunk_frames);
    } else {
        const uint64_t per = 4ull * (uint64_t)(DXP + 3 * M + DP1) + 1;
        nc_max = std::max<uint64_t>(1024, (768ull << 20) / per / 256 * 256);
    }
    nc_max = std::min<uint64_t>(nc_max, h->N);
    const int S = stats_subchunks(2 * M, DP1, h->num_cus, nc_max);
    const uint64_t ncp_max = (uint64_t)stats_sub(nc_max, S) * S;
    HIPCHK(h, ensure(&fs->d_xt, &fs->xt_cap, (size_t)ncp_max * DXP));
    HIPCHK(h, ensure(&fs->d_L, &fs->L_cap, (size_t)ncp_max * M));
    HIPCHK(h, ensure(&fs->d_bad, &fs->bad_cap, (size_t)ncp_max));
    HIPCHK(h, ensure(&a->d_H, &a->H_cap, (size_t)ncp_max * 2 * M));
    HIPCHK(h, ensure(&a->d_psi, &a->psi_cap, (size_t)ncp_max * DP1));
    HIPCHK(h, ensure(&a->d_carry, &a->carry_cap, (size_t)M));
    HIPCHK(h, ensure(&a->d_part, &a->part_cap, (size_t)S * 2 * M * DP1));
    HIPCHK(h, ensure(&a->d_acc, &a->acc_cap, (size_t)2 * M * DP1));
    HIPCHK(h, ensure(&h->d_blockLL, &h->blockLL_cap, (size_t)(ncp_max + 3) / 4));
    int first = 1;
    for (uint64_t n0 = 0; n0 < h->N; n0 += nc_max) {
        const int nc = (int)std::min<uint64_t>(nc_max, h->N - n0);
        const int sub = stats_sub(nc, S);
        const int ncp = sub * S;
        HIPCHK(h, launch_fc_rows(h->d_x, n0, nc, ncp, D, cen, fs->d_xt, fs->d_bad, h->stream));
        HIPCHK(h, launch_fc_ll(fs->d_xt, nc, D, theta, M, KP, pairs, fs->d_L, h->stream));
        HIPCHK(h, launch_fc_lse(fs->d_L, nc, ncp, M, fs->d_bad, h->opt.min_loglike,
                                h->opt.min_gamma, 0, nullptr, n0, h->d_blockLL, h->stream));
        hipLaunchKernelGGL(gmmhip::aux_pair_kernel, dim3((ncp + 3) / 4), dim3(256), 0, h->stream,
                           fs->d_L, a->d_carry, h->d_x, n0, nc, ncp, M, D, DP1, h->opt.min_gamma,
                           a->d_H, a->d_psi);
        HIPCHK(h, hipGetLastError());
        // part[s] [2M][DP1] = H_s^T Psi_s over frames [s sub, (s+1) sub)
        HIPCHK(h, launch_fc_stats(a->d_psi, DP1, a->d_H, 2 * M, DP1, a->d_psipairs, sub, S,
                                  a->d_part, h->stream));
        HIPCHK(h, launch_fc_acc(a->d_part, (size_t)2 * M * DP1, S, first, a->d_acc, h->stream));
        first = 0;
        // the next chunk's first pair needs this chunk's last posteriors
        HIPCHK(h, hipMemcpyAsync(a->d_carry, fs->d_L + (size_t)(nc - 1) * M, M * sizeof(float),
                                 hipMemcpyDeviceToDevice, h->stream));
    }
    std::vector<double> acc((size_t)2 * M * DP1);
    HIPCHK(h, hipMemcpyAsync(acc.data(), a->d_acc, acc.size() * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int m = 0; m < M; m++) {
        const double *r = acc.data() + (size_t)m * DP1;
        double *o = aux + (size_t)m * (D + 2);
        for (int d = 0; d < D; d++) o[d] += r[1 + d];
        o[D] += acc[(size_t)(M + m) * DP1];
        o[D + 1] += r[0];
    }
    return GMMHIP_OK;
}
