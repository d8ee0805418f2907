// gmmhip_batch_adapt: batched per-utterance statistics, MAP means and MLLR
// transforms (include/gmmhip.h).
//
// Pipeline on the handle's stream:
//   loglik_kernel over all frames of the batch (per-frame LL is independent
//     of the utterance),
//   stats_queue_kernel: a persistent grid pulls (chunk, component group)
//     items from a device counter; chunks never straddle utterances,
//   seg_reduce_kernel: per-utterance FP64 sums of the chunk partials and
//     the MAP means, and for MLLR mllr_check / two FP64 GEMMs (G, k)
//     against UBM-only matrices with the A operand formed from n, F and
//     1/var while staged / mllr_solve.
// With opts.precision fp16x3 / bf16x3 the E-step part runs the mixed form:
// every utterance is copied to a 32-row-aligned place (pad_rows_kernel),
// prep_frames splits the frames centred and scaled by the UBM's global mean
// and deviation (so each utterance's result does not depend on the batch),
// then loglik_mixed and the statistics over the utterances' chunks in the
// persistent form of stats_mixed_tr, whose blocks pull (chunk, group) items
// from per-XCD device counters.  A batch with a frame value outside the
// 16-bit range, or a model whose coefficients overflow it, runs in FP32.
// With cfg->graph the whole pipeline is captured once into a hipGraph and
// replayed while the batch shape and buffers stay the same.
#include "batch.h"
#include "ctx.h"
#include "estep_mixed.h"

using namespace gmmhip_detail;

struct gmmhip_batch_state {
    uint64_t *d_lo = nullptr, *d_hi = nullptr;
    size_t lo_cap = 0, hi_cap = 0;
    int *d_first = nullptr;
    size_t first_cap = 0;
    unsigned int *d_queue = nullptr;
    double *d_n = nullptr, *d_F = nullptr;
    size_t n_cap = 0, F_cap = 0;
    float *d_means = nullptr;
    size_t means_cap = 0;
    double *d_G = nullptr, *d_K = nullptr, *d_W = nullptr;
    size_t G_cap = 0, K_cap = 0, W_cap = 0;
    int *d_status = nullptr;
    size_t status_cap = 0;
    // UBM-only MLLR operands: P[m][packed (a<=b)] = xi_a xi_b, Xi[m][a], prior,
    // 1/var in FP64
    double *d_P = nullptr, *d_Xi = nullptr, *d_Ginit = nullptr, *d_kinit = nullptr;
    double *d_ivar = nullptr;
    size_t P_cap = 0, Xi_cap = 0, Ginit_cap = 0, kinit_cap = 0, ivar_cap = 0;
    uint64_t mllr_model = ~0ull;
    // mixed form: padded frames, split features, range flags, UBM-based
    // centring / scaling and the split model for them
    float *d_xpad = nullptr;
    size_t xpad_cap = 0;
    uint16_t *d_zah = nullptr, *d_zal = nullptr;
    size_t zah_cap = 0, zal_cap = 0;
    uint8_t *d_bad = nullptr;
    size_t bad_cap = 0;
    unsigned int *d_nrange = nullptr;
    uint64_t *d_src = nullptr, *d_dst = nullptr;
    size_t src_cap = 0, dst_cap = 0;
    float *d_zoff = nullptr, *d_zscale = nullptr;
    size_t zoff_cap = 0, zscale_cap = 0;
    uint16_t *d_P16 = nullptr;
    size_t P16_cap = 0;
    uint64_t p16_model = ~0ull;
    int p16_prec = -1;
    bool p16_ok = false;
    // graph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key;
    uint64_t captures = 0;
};

namespace gmmhip_detail {
void batch_state_free(gmmhip_batch_state *b) {
    if (!b) return;
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (void *p : {(void *)b->d_lo, (void *)b->d_hi, (void *)b->d_first, (void *)b->d_queue,
                    (void *)b->d_n, (void *)b->d_F, (void *)b->d_means, (void *)b->d_G,
                    (void *)b->d_K, (void *)b->d_W, (void *)b->d_status, (void *)b->d_P,
                    (void *)b->d_Xi, (void *)b->d_Ginit, (void *)b->d_kinit, (void *)b->d_ivar,
                    (void *)b->d_xpad, (void *)b->d_zah, (void *)b->d_zal, (void *)b->d_bad,
                    (void *)b->d_nrange, (void *)b->d_src,
                    (void *)b->d_dst, (void *)b->d_zoff, (void *)b->d_zscale, (void *)b->d_P16})
        if (p) hipFree(p);
    delete b;
}
}  // namespace gmmhip_detail

namespace {

// UBM-only MLLR operands (FP64, host-built once per model).
int build_mllr_operands(gmmhip_ctx *h, gmmhip_batch_state *b) {
    if (b->mllr_model == h->model_version) return GMMHIP_OK;
    const int M = h->M, D = h->D, E = D + 1, NP = E * (E + 1) / 2;
    std::vector<double> P((size_t)M * NP), Xi((size_t)M * E), iv((size_t)M * D);
    for (size_t j = 0; j < iv.size(); j++) iv[j] = 1.0 / (double)h->m_var[j];
    std::vector<double> Gi((size_t)D * E * E, 0.0), ki((size_t)D * E, 0.0);
    std::vector<double> xi(E);
    const double minv = h->opt.min_var;
    for (int m = 0; m < M; m++) {
        for (int a = 0; a < D; a++) xi[a] = h->m_mu[(size_t)m * D + a];
        xi[D] = 1.0;
        for (int a = 0, pk = 0; a < E; a++)
            for (int c = a; c < E; c++, pk++) P[(size_t)m * NP + pk] = xi[a] * xi[c];
        for (int a = 0; a < E; a++) Xi[(size_t)m * E + a] = xi[a];
        // prepareKiGiInitMatrices (OL_ModelAdapt.cpp:1484-1523): one virtual
        // frame per component at its mean, with its variance
        const float *var = &h->m_var[(size_t)m * D];
        const float *mean = &h->m_mu[(size_t)m * D];
        for (int i = 0; i < D; i++) {
            const double cG = (var[i] < minv) ? 1.0 / minv : 1.0 / var[i];
            const double ck = mean[i] * cG;
            double *G = &Gi[(size_t)i * E * E], *k = &ki[(size_t)i * E];
            for (int a = 0; a < D; a++) {
                k[a] += ck * mean[a];
                G[(size_t)D * E + a] += cG * mean[a];
                G[(size_t)a * E + D] += cG * mean[a];
                for (int c = 0; c < D; c++) G[(size_t)a * E + c] += cG * mean[a] * mean[c];
                G[(size_t)a * E + a] += cG * var[a];
            }
            G[(size_t)E * E - 1] += cG;
            k[D] += ck;
        }
    }
    // the prior of row i, packed like the per-utterance G (upper triangle)
    std::vector<double> Gip((size_t)D * NP);
    for (int i = 0; i < D; i++)
        for (int a = 0, pk = 0; a < E; a++)
            for (int c = a; c < E; c++, pk++) Gip[(size_t)i * NP + pk] = Gi[((size_t)i * E + a) * E + c];
    HIPCHK(h, ensure(&b->d_P, &b->P_cap, P.size()));
    HIPCHK(h, ensure(&b->d_Xi, &b->Xi_cap, Xi.size()));
    HIPCHK(h, ensure(&b->d_Ginit, &b->Ginit_cap, Gip.size()));
    HIPCHK(h, ensure(&b->d_kinit, &b->kinit_cap, ki.size()));
    HIPCHK(h, ensure(&b->d_ivar, &b->ivar_cap, iv.size()));
    HIPCHK(h, hipMemcpyAsync(b->d_ivar, iv.data(), iv.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(b->d_P, P.data(), P.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(b->d_Xi, Xi.data(), Xi.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(b->d_Ginit, Gip.data(), Gip.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(b->d_kinit, ki.data(), ki.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    b->mllr_model = h->model_version;
    return GMMHIP_OK;
}

}  // namespace

namespace gmmhip_detail {

int model_operands(gmmhip_ctx *h, bool mllr, ModelOperands *o) {
    DevGuard g(h->dev);
    if (!h->batch) h->batch = new gmmhip_batch_state();
    gmmhip_batch_state *b = h->batch;
    const size_t MD = (size_t)h->M * h->D;
    if (!h->d_mu_f32 || h->mu_f32_version != h->model_version) {
        HIPCHK(h, ensure(&h->d_mu_f32, &h->mu_f32_cap, MD));
        HIPCHK(h, ensure(&h->d_var_f32, &h->var_f32_cap, MD));
        HIPCHK(h, ensure(&h->d_gconst, &h->gconst_cap, (size_t)h->M));
        HIPCHK(h, hipMemcpyAsync(h->d_mu_f32, h->m_mu.data(), MD * 4, hipMemcpyHostToDevice,
                                 h->stream));
        HIPCHK(h, hipMemcpyAsync(h->d_var_f32, h->m_var.data(), MD * 4, hipMemcpyHostToDevice,
                                 h->stream));
        HIPCHK(h, hipMemcpyAsync(h->d_gconst, h->m_gconst.data(), (size_t)h->M * 8,
                                 hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        h->mu_f32_version = h->model_version;
    }
    o->mu = h->d_mu_f32;
    o->var = h->d_var_f32;
    o->gconst = h->d_gconst;
    o->P = o->Xi = o->Ginit = o->kinit = o->ivar = nullptr;
    if (mllr) {
        const int rc = build_mllr_operands(h, b);
        if (rc) return rc;
        o->P = b->d_P;
        o->Xi = b->d_Xi;
        o->Ginit = b->d_Ginit;
        o->kinit = b->d_kinit;
        o->ivar = b->d_ivar;
    }
    return GMMHIP_OK;
}

}  // namespace gmmhip_detail

namespace {

struct BatchPlan {
    int U = 0, M = 0, D = 0, KP = 0, Mt = 0, ngroups = 0, nchunks = 0;
    uint64_t N = 0, cf = 0;
    // mixed form: split precision, padded rows (Np, Npad), KP / tiles of the split model
    bool mixed = false;
    int prec = 0, KPm = 0, Mt32 = 0;
    uint64_t Np = 0, Npad = 0;
    bool want_means = false, want_mllr = false, prior = false, per_utt = false;
    int tau = 0;
    int ll_sub = 2;
    int ll_nw = 4;  // mixed log-likelihood waves per block
    int st_nw = 4;  // mixed statistics waves per block (32 components each)
    std::vector<uint64_t> lo, hi;
    std::vector<int> first;
};

// Mixed form operands of the handle's model: centring o = sum_m w_m mu_m and
// scaling s = 1 / sqrt(sum_m w_m (var_m + (mu_m - o)^2)) (the UBM's global
// mean and deviation, the same for every batch) and the split model for
// them; *ok is false when a coefficient overflows the 16-bit type.
int batch_mixed_operands(gmmhip_ctx *h, gmmhip_batch_state *b, int prec, int KPm, bool *ok) {
    if (b->p16_model == h->model_version && b->p16_prec == prec) {
        *ok = b->p16_ok;
        return GMMHIP_OK;
    }
    const int M = h->M, D = h->D;
    std::vector<float> zoff(D, 0.0f), zscale(D, 1.0f);
    double wsum = 0.0;
    for (int m = 0; m < M; m++) wsum += h->m_w[m];
    if (!(wsum > 0.0)) wsum = 1.0;
    for (int d = 0; d < D; d++) {
        double o = 0.0;
        for (int m = 0; m < M; m++) o += h->m_w[m] * (double)h->m_mu[(size_t)m * D + d];
        o /= wsum;
        double v = 0.0;
        for (int m = 0; m < M; m++) {
            const double dm = (double)h->m_mu[(size_t)m * D + d] - o;
            v += h->m_w[m] * ((double)h->m_var[(size_t)m * D + d] + dm * dm);
        }
        v /= wsum;
        zoff[d] = (float)o;
        if (std::isfinite(v) && v > 0.0) zscale[d] = (float)(1.0 / std::sqrt(v));
    }
    HIPCHK(h, ensure(&b->d_zoff, &b->zoff_cap, (size_t)D));
    HIPCHK(h, ensure(&b->d_zscale, &b->zscale_cap, (size_t)D));
    HIPCHK(h, ensure(&b->d_P16, &b->P16_cap, (size_t)((M + 31) / 32) * KPm * 64));
    HIPCHK(h, hipMemcpyAsync(b->d_zoff, zoff.data(), D * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(b->d_zscale, zscale.data(), D * 4, hipMemcpyHostToDevice, h->stream));
    bool overflow = false;
    int rc = pack_split_model(h, prec, KPm, b->d_zoff, b->d_zscale, b->d_P16, &overflow);
    if (rc) return rc;
    b->p16_model = h->model_version;
    b->p16_prec = prec;
    b->p16_ok = !overflow;
    *ok = b->p16_ok;
    return GMMHIP_OK;
}

// Enqueue the device pipeline (no host synchronisation; graph-capturable:
// kernels only, the queue counter is reset by a kernel).
int enqueue_pipeline(gmmhip_ctx *h, gmmhip_batch_state *b, const BatchPlan &p) {
    hipStream_t s = h->stream;
    EstepArgs a{};
    a.X = h->d_x;
    a.ldx = p.D;
    a.N = p.N;
    a.D = p.D;
    a.KP = p.KP;
    a.P = h->d_P;
    a.M = p.M;
    a.Mt = p.Mt;
    a.off = h->d_off;
    a.min_ll = h->opt.min_loglike;
    a.min_gamma = h->opt.min_gamma;
    a.LL2 = h->d_LL2;
    a.LLnat = nullptr;
    a.blockLL = h->d_blockLL;
    a.ngroups = p.ngroups;
    a.partials = h->d_part;
    a.chunk_frames = p.cf;
    a.ll_sub = p.ll_sub;
    const int blocks_cap = std::max(1, h->num_cus) * 2;
    if (p.mixed) {
        MixedArgs m{};
        m.N = p.Np;
        m.Npad = p.Npad;
        m.D = p.D;
        m.KP = p.KPm;
        m.za_hi = b->d_zah;
        m.za_lo = b->d_zal;
        m.bad = b->d_bad;
        m.P16 = b->d_P16;
        m.M = p.M;
        m.Mt32 = p.Mt32;
        m.min_ll = h->opt.min_loglike;
        m.min_gamma = h->opt.min_gamma;
        m.LL2 = h->d_LL2;
        m.LLnat = nullptr;
        m.blockLL = h->d_blockLL;
        m.ll_nw = p.ll_nw;
        HIPCHK(h, launch_loglik_mixed(m, p.prec, s));
        m.chunk_frames = p.cf;
        m.nchunks = p.nchunks;
        m.st_nw = p.st_nw;
        m.ngroups = (p.Mt32 + p.st_nw - 1) / p.st_nw;
        m.partials = h->d_part;
        m.chunk_lo = b->d_lo;
        m.chunk_hi = b->d_hi;
        // persistent grid over the (chunk, group) items; counters reset by a
        // kernel so the pipeline stays capturable
        m.queue = b->d_queue;
        m.queue_blocks =
            mixed_stats_queue_blocks(p.KPm, p.st_nw, h->num_cus, p.nchunks * m.ngroups);
        HIPCHK(h, launch_fill_u32(b->d_queue, 0u, 8, s));
        HIPCHK(h, launch_stats_mixed_queue(m, p.prec, s));
    } else if (!p.per_utt) {
        HIPCHK(h, launch_loglik_fp32(a, s));
        a.nchunks = p.nchunks;
        a.chunk_lo = b->d_lo;
        a.chunk_hi = b->d_hi;
        a.queue = b->d_queue;
        a.queue_blocks = std::min(p.nchunks * p.ngroups, blocks_cap);
        HIPCHK(h, launch_fill_u32(b->d_queue, 0u, 1, s));
        HIPCHK(h, launch_stats_fp32_queue(a, s));
    } else {
        // baseline: the same kernels, one launch set per utterance
        for (int u = 0; u < p.U; u++) {
            const int c0 = p.first[u], c1 = p.first[u + 1];
            if (c1 == c0) continue;
            const uint64_t f0 = p.lo[c0], f1 = p.hi[c1 - 1];
            EstepArgs au = a;
            au.X = h->d_x + f0 * p.D;
            au.N = f1 - f0;
            au.LL2 = h->d_LL2 + f0;
            HIPCHK(h, launch_loglik_fp32(au, s));
            au.X = h->d_x;
            au.N = p.N;
            au.LL2 = h->d_LL2;
            au.nchunks = c1 - c0;
            au.chunk_lo = b->d_lo + c0;
            au.chunk_hi = b->d_hi + c0;
            au.partials = h->d_part + (size_t)c0 * p.Mt * 16 * p.KP;
            au.queue = b->d_queue + 1 + (u % 255);
            au.queue_blocks = std::min((c1 - c0) * p.ngroups, blocks_cap);
            HIPCHK(h, launch_fill_u32(au.queue, 0u, 1, s));
            HIPCHK(h, launch_stats_fp32_queue(au, s));
        }
    }
    SegReduceArgs r{};
    r.partials = h->d_part;
    r.chunk_first = b->d_first;
    r.off = p.mixed ? b->d_zoff : h->d_off;
    r.scale = p.mixed ? b->d_zscale : nullptr;
    r.U = p.U;
    r.M = p.M;
    r.Mp = p.mixed ? p.Mt32 * 32 : p.Mt * 16;
    r.D = p.D;
    r.KP = p.mixed ? p.KPm : p.KP;
    r.n = b->d_n;
    r.F = b->d_F;
    r.mu = h->d_mu_f32;
    r.tau = p.tau;
    r.Umeans = p.want_means ? p.U : 0;
    r.means = b->d_means;
    HIPCHK(h, launch_seg_reduce(r, s));
    if (p.want_mllr) {
        const int E = p.D + 1, NP = E * (E + 1) / 2;
        const int R = p.U * p.D;
        HIPCHK(h, launch_mllr_check(b->d_n, p.U, p.M, p.D, b->d_status, s));
        MllrGkArgs g{};
        g.n = b->d_n;
        g.F = b->d_F;
        g.ivar = b->d_ivar;
        g.U = p.U;
        g.M = p.M;
        g.D = p.D;
        g.B = b->d_P;
        g.ldb = NP;
        g.Cn = NP;
        g.C = b->d_G;
        g.ldc = NP;
        g.kmode = 0;
        HIPCHK(h, launch_mllr_gk(g, s));
        g.B = b->d_Xi;
        g.ldb = E;
        g.Cn = E;
        g.C = b->d_K;
        g.ldc = E;
        g.kmode = 1;
        HIPCHK(h, launch_mllr_gk(g, s));
        MllrSolveArgs m{};
        m.Gp = b->d_G;
        m.Kp = b->d_K;
        m.Ginit = p.prior ? b->d_Ginit : nullptr;
        m.kinit = p.prior ? b->d_kinit : nullptr;
        m.U = p.U;
        m.D = p.D;
        m.ldg = NP;
        m.ldk = E;
        m.status = b->d_status;
        m.W = b->d_W;
        HIPCHK(h, launch_mllr_solve(m, s));
    }
    return GMMHIP_OK;
}

}  // namespace

extern "C" int gmmhip_batch_adapt(gmmhip_t h, const float *frames, int frames_on_device,
                                  const uint64_t *offsets, int U, int D,
                                  const gmmhip_batch_cfg *cfg, gmmhip_batch_out *out) {
    if (!h) return GMMHIP_ERR_ARG;
    if (!frames || !offsets || U <= 0 || !cfg || !out)
        return fail(h, GMMHIP_ERR_ARG, "batch_adapt: null argument");
    if (h->M <= 0) return fail(h, GMMHIP_ERR_STATE, "no model inserted");
    if (h->full_model)
        return fail(h, GMMHIP_ERR_ARG, "batch_adapt: needs a diagonal model (full covariance inserted)");
    if (D != h->D) return fail(h, GMMHIP_ERR_ARG, "batch_adapt: D=%d but the model has D=%d", D, h->D);
    if (offsets[0] != 0) return fail(h, GMMHIP_ERR_ARG, "batch_adapt: offsets[0] must be 0");
    for (int u = 0; u < U; u++)
        if (offsets[u + 1] < offsets[u])
            return fail(h, GMMHIP_ERR_ARG, "batch_adapt: offsets must be non-decreasing");
    if (out->mllr_W && !out->mllr_status)
        return fail(h, GMMHIP_ERR_ARG, "batch_adapt: mllr_W needs mllr_status");
    if (out->mllr_W && h->M > MLLR_MAX_COMPONENTS)
        return fail(h, GMMHIP_ERR_ARG, "batch_adapt: MLLR supports at most %d components",
                    MLLR_MAX_COMPONENTS);
    const uint64_t N = offsets[U];
    if (N == 0) return fail(h, GMMHIP_ERR_ARG, "batch_adapt: no frames");
    DevGuard g(h->dev);
    if (!h->batch) h->batch = new gmmhip_batch_state();
    gmmhip_batch_state *b = h->batch;
    if (h->num_cus == 0) {
        int n = 0;
        HIPCHK(h, hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, h->dev));
        h->num_cus = n > 0 ? n : 256;
    }
    int rc;
    // frames become the handle's frames (borrowed device memory or a copy)
    if (frames_on_device)
        rc = gmmhip_set_frames_device(h, frames, N, D, D);
    else
        rc = gmmhip_set_frames_contig(h, frames, N, D, D);
    if (rc) return rc;

    BatchPlan p;
    p.U = U;
    p.M = h->M;
    p.D = D;
    p.KP = h->KP;
    p.Mt = h->Mt;
    p.ngroups = (h->Mt + 3) / 4;
    p.N = N;
    // the table's entry for the handle's precision, else the FP32 entry
    uint64_t tcf = tuned_chunk(h, GMMHIP_TUNE_BATCH, h->opt.precision);
    if (tcf == 0) tcf = tuned_chunk(h, GMMHIP_TUNE_BATCH, GMMHIP_PREC_FP32);
    p.cf = cfg->chunk_frames > 0 ? (uint64_t)(cfg->chunk_frames + 63) / 64 * 64 : tcf;
    if (p.cf == 0) p.cf = 1024;
    p.ll_sub = choose_ll_sub(h, GMMHIP_TUNE_BATCH);
    p.ll_nw = choose_mixed_ll_nw(h, GMMHIP_TUNE_BATCH);
    p.st_nw = choose_mixed_st_nw(h, GMMHIP_TUNE_BATCH);
    h->used_ll_frames = 64 * p.ll_sub;
    h->used_chunk = (int)p.cf;
    p.want_means = out->means != nullptr;
    p.want_mllr = out->mllr_W != nullptr;
    p.prior = cfg->mllr_prior != 0;
    p.per_utt = cfg->mode == GMMHIP_BATCH_PER_UTTERANCE;
    p.tau = cfg->tau;
    // mixed form (queue mode): utterances on 32-row blocks, split features
    // of the padded rows; a frame out of the 16-bit range sends the batch to
    // FP32
    std::vector<uint64_t> dst(U + 1, 0);
    const int prec = h->opt.precision;
    if (prec != GMMHIP_PREC_FP32 && !p.per_utt) {
        p.KPm = kp_mixed_for_dim(D);
        p.Mt32 = (h->M + 31) / 32;
        bool ok = false;
        if ((rc = batch_mixed_operands(h, b, prec, p.KPm, &ok))) return rc;
        if (ok) {
            for (int u = 0; u < U; u++)
                dst[u + 1] = dst[u] + (offsets[u + 1] - offsets[u] + 31) / 32 * 32;
            p.Np = dst[U];
            p.Npad = pad128(p.Np);
            HIPCHK(h, ensure(&b->d_src, &b->src_cap, (size_t)U + 1));
            HIPCHK(h, ensure(&b->d_dst, &b->dst_cap, (size_t)U + 1));
            HIPCHK(h, ensure(&b->d_xpad, &b->xpad_cap, (size_t)p.Np * D));
            const size_t zn = (size_t)p.Npad * p.KPm;
            HIPCHK(h, ensure(&b->d_zah, &b->zah_cap, zn));
            HIPCHK(h, ensure(&b->d_zal, &b->zal_cap, zn));
            HIPCHK(h, ensure(&b->d_bad, &b->bad_cap, (size_t)p.Npad));
            if (!b->d_nrange) HIPCHK(h, hipMalloc((void **)&b->d_nrange, sizeof(unsigned int)));
            HIPCHK(h, hipMemcpyAsync(b->d_src, offsets, (U + 1) * 8, hipMemcpyHostToDevice, h->stream));
            HIPCHK(h, hipMemcpyAsync(b->d_dst, dst.data(), (U + 1) * 8, hipMemcpyHostToDevice,
                                     h->stream));
            HIPCHK(h, launch_pad_rows(h->d_x, b->d_src, b->d_dst, U, D, b->d_xpad, h->stream));
            HIPCHK(h, launch_fill_u32(b->d_nrange, 0u, 1, h->stream));
            PrepArgs pa{};
            pa.X = b->d_xpad;
            pa.N = p.Np;
            pa.Npad = p.Npad;
            pa.D = D;
            pa.KP = p.KPm;
            pa.off = b->d_zoff;
            pa.scale = b->d_zscale;
            pa.za_hi_raw = b->d_zah;
            pa.za_lo_raw = b->d_zal;
            pa.bad = b->d_bad;
            pa.nrange = b->d_nrange;
            pa.zlimit = prec == GMMHIP_PREC_FP16X3 ? 240.0f : INFINITY;
            HIPCHK(h, launch_prep_frames(pa, prec, h->stream));
            unsigned int nrange = 0;
            HIPCHK(h, hipMemcpyAsync(&nrange, b->d_nrange, sizeof(nrange), hipMemcpyDeviceToHost,
                                     h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            p.mixed = nrange == 0;
            p.prec = prec;
        }
    }
    h->used_comp_tile = p.mixed ? 32 * p.st_nw : 0;
    p.first.assign(U + 1, 0);
    for (int u = 0; u < U; u++) {
        p.first[u] = (int)p.lo.size();
        // chunks in the rows the kernels see: padded rows in the mixed form
        const uint64_t base = p.mixed ? dst[u] : offsets[u];
        const uint64_t end = base + (offsets[u + 1] - offsets[u]);
        for (uint64_t s0 = base; s0 < end; s0 += p.cf) {
            p.lo.push_back(s0);
            p.hi.push_back(std::min<uint64_t>(end, s0 + p.cf));
        }
    }
    p.first[U] = (int)p.lo.size();
    p.nchunks = (int)p.lo.size();
    const size_t M = p.M, E = D + 1, NP = E * (E + 1) / 2;
    const size_t Mp = p.mixed ? (size_t)p.Mt32 * 32 : (size_t)p.Mt * 16;
    const size_t KPx = p.mixed ? (size_t)p.KPm : (size_t)p.KP;
    const uint64_t Nrows = p.mixed ? p.Npad : N;

    // buffers
    HIPCHK(h, ensure(&h->d_LL2, &h->LL2_cap, Nrows));
    HIPCHK(h, ensure(&h->d_blockLL, &h->blockLL_cap, ll_blocks(Nrows, 1) + (size_t)U));
    HIPCHK(h, ensure(&h->d_part, &h->part_cap, (size_t)p.nchunks * Mp * KPx));
    HIPCHK(h, ensure(&b->d_lo, &b->lo_cap, (size_t)p.nchunks));
    HIPCHK(h, ensure(&b->d_hi, &b->hi_cap, (size_t)p.nchunks));
    HIPCHK(h, ensure(&b->d_first, &b->first_cap, (size_t)U + 1));
    if (!b->d_queue) HIPCHK(h, hipMalloc((void **)&b->d_queue, 256 * sizeof(unsigned int)));
    HIPCHK(h, ensure(&b->d_n, &b->n_cap, (size_t)U * M));
    HIPCHK(h, ensure(&b->d_F, &b->F_cap, (size_t)U * M * D));
    if (p.want_means || p.want_mllr) {
        ModelOperands mo;
        if ((rc = model_operands(h, p.want_mllr, &mo))) return rc;
    }
    if (p.want_means) HIPCHK(h, ensure(&b->d_means, &b->means_cap, (size_t)U * M * D));
    if (p.want_mllr) {
        HIPCHK(h, ensure(&b->d_G, &b->G_cap, (size_t)U * D * NP));
        HIPCHK(h, ensure(&b->d_K, &b->K_cap, (size_t)U * D * E));
        HIPCHK(h, ensure(&b->d_W, &b->W_cap, (size_t)U * D * E));
        HIPCHK(h, ensure(&b->d_status, &b->status_cap, (size_t)U));
    }
    HIPCHK(h, hipMemcpyAsync(b->d_lo, p.lo.data(), p.lo.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(b->d_hi, p.hi.data(), p.hi.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(b->d_first, p.first.data(), p.first.size() * 4, hipMemcpyHostToDevice,
                             h->stream));

    // run, directly or through a cached graph
    if (cfg->graph) {
        std::vector<uint64_t> key = {N,
                                     (uint64_t)U,
                                     (uint64_t)(uintptr_t)h->d_x,
                                     h->model_version,
                                     p.cf,
                                     (uint64_t)p.ll_sub,
                                     (uint64_t)p.ll_nw,
                                     (uint64_t)p.st_nw,
                                     (uint64_t)p.want_means,
                                     (uint64_t)p.want_mllr,
                                     (uint64_t)p.prior,
                                     (uint64_t)p.per_utt,
                                     (uint64_t)p.tau,
                                     (uint64_t)p.mixed,
                                     (uint64_t)p.prec,
                                     (uint64_t)(uintptr_t)b->d_zah,
                                     (uint64_t)(uintptr_t)b->d_P16,
                                     (uint64_t)(uintptr_t)h->d_part,
                                     (uint64_t)(uintptr_t)b->d_n,
                                     (uint64_t)(uintptr_t)b->d_G,
                                     // every buffer the pipeline captures
                                     // (handle-wide scratch included) stays
                                     // where it was only while no device
                                     // buffer has been reallocated since
                                     g_alloc_generation.load(std::memory_order_relaxed)};
        key.insert(key.end(), offsets, offsets + U + 1);
        if (!b->exec || key != b->key) {
            if (b->exec) HIPCHK(h, hipGraphExecDestroy(b->exec));
            if (b->graph) HIPCHK(h, hipGraphDestroy(b->graph));
            b->exec = nullptr;
            b->graph = nullptr;
            HIPCHK(h, hipStreamSynchronize(h->stream));
            HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_pipeline(h, b, p);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(h->stream, &graph);
            if (rc) {
                if (graph) hipGraphDestroy(graph);
                return rc;
            }
            HIPCHK(h, e);
            b->graph = graph;
            HIPCHK(h, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->key = key;
            b->captures++;
        }
        HIPCHK(h, hipGraphLaunch(b->exec, h->stream));
    } else {
        if ((rc = enqueue_pipeline(h, b, p))) return rc;
    }

    // outputs
    const hipMemcpyKind kind = cfg->outputs_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (out->occ) HIPCHK(h, hipMemcpyAsync(out->occ, b->d_n, (size_t)U * M * 8, kind, h->stream));
    if (out->first)
        HIPCHK(h, hipMemcpyAsync(out->first, b->d_F, (size_t)U * M * D * 8, kind, h->stream));
    if (out->means)
        HIPCHK(h, hipMemcpyAsync(out->means, b->d_means, (size_t)U * M * D * 4, kind, h->stream));
    if (out->mllr_W) {
        HIPCHK(h, hipMemcpyAsync(out->mllr_W, b->d_W, (size_t)U * D * E * 8, kind, h->stream));
        if (cfg->outputs_on_device) {
            HIPCHK(h, hipMemcpyAsync(out->mllr_status, b->d_status, (size_t)U * 4, kind, h->stream));
        } else {
            HIPCHK(h, hipMemcpyAsync(out->mllr_status, b->d_status, (size_t)U * 4, kind, h->stream));
        }
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (out->mllr_W && !cfg->outputs_on_device)
        for (int u = 0; u < U; u++)
            if (out->mllr_status[u]) out->mllr_status[u] = GMMHIP_ERR_MLLR;
    return GMMHIP_OK;
}

extern "C" int gmmhip_batch_graph_captures(gmmhip_t h, uint64_t *captures) {
    if (!h || !captures) return GMMHIP_ERR_ARG;
    *captures = h->batch ? h->batch->captures : 0;
    return GMMHIP_OK;
}
