// C ABI of the gfx950 diagonal-GMM accumulator (include/gmmhip.h).
//
// One handle = one device + one HIP stream + device buffers it owns.  The
// model is preprocessed on the host in FP64 (insertModel,
// trainer/OL_GMMStatsEstimator.cpp:234-313), mean-centred and packed into
// the MFMA B-operand layout; frames stay resident in HBM row-major, so the
// reference's host transpose into 8-frame tiles (alignVectors,
// trainer/OL_GMMStatsEstimator_CUDA.cpp:292-328) has no counterpart here.
#include "batch.h"
#include "ctx.h"
#include "tune.h"

namespace gmmhip_detail {

// Frames per FP32 partial accumulator: enough chunks to give >= ~4096
// blocks, at most 8192 frames per FP32 accumulator, a multiple of the
// statistics kernel's frame step.
// The upper bound comes from the tuning table when it has an entry for the
// precision in use.
uint64_t estep_chunk_frames(const gmmhip_ctx *h, uint64_t N, int ngroups, bool mixed) {
    if (h->opt.chunk_frames > 0) {
        uint64_t cf = (uint64_t)h->opt.chunk_frames;
        return (cf + ST_FRAMES_PER_ITER - 1) / ST_FRAMES_PER_ITER * ST_FRAMES_PER_ITER;
    }
    uint64_t cap = tuned_chunk(h, GMMHIP_TUNE_ESTEP, mixed ? h->opt.precision : GMMHIP_PREC_FP32);
    if (cap == 0) cap = 8192;
    const uint64_t target_chunks = std::max<uint64_t>(1, 4096 / std::max(1, ngroups));
    uint64_t cf = (N + target_chunks - 1) / target_chunks;
    cf = (cf + ST_FRAMES_PER_ITER - 1) / ST_FRAMES_PER_ITER * ST_FRAMES_PER_ITER;
    cf = std::max<uint64_t>(cf, 128);
    cf = std::min<uint64_t>(cf, cap);
    return cf;
}

}  // namespace gmmhip_detail

namespace {
using namespace gmmhip_detail;

uint64_t pick_chunk_frames(const gmmhip_ctx *h, int ngroups, bool mixed) {
    return estep_chunk_frames(h, h->N, ngroups, mixed);
}

int check_ready(gmmhip_ctx *h) {
    if (!h) return GMMHIP_ERR_ARG;
    if (h->M <= 0) return fail(h, GMMHIP_ERR_STATE, "no model inserted");
    if (h->full_model && h->opt.precision != GMMHIP_PREC_FP32)
        return fail(h, GMMHIP_ERR_ARG, "full-covariance models run in FP32 only");
    if (!h->d_x || h->N == 0) return fail(h, GMMHIP_ERR_STATE, "no frames inserted");
    if (h->xD != h->D)
        return fail(h, GMMHIP_ERR_ARG, "frame dimension %d != model dimension %d", h->xD, h->D);
    return GMMHIP_OK;
}

EstepArgs make_args(gmmhip_ctx *h) {
    EstepArgs a{};
    a.X = h->d_x;
    a.ldx = h->ld;
    a.N = h->N;
    a.D = h->D;
    a.KP = h->KP;
    a.P = h->d_P;
    a.M = h->M;
    a.Mt = h->Mt;
    a.off = h->d_off;
    a.min_ll = h->opt.min_loglike;
    a.min_gamma = h->opt.min_gamma;
    return a;
}

int ensure_events(gmmhip_ctx *h) {
    if (h->ev_ok) return GMMHIP_OK;
    for (auto &e : h->ev) HIPCHK(h, hipEventCreate(&e));
    h->ev_ok = true;
    return GMMHIP_OK;
}

void add_time(gmmhip_ctx *h, int cls, int e0, int e1) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, h->ev[e0], h->ev[e1]) == hipSuccess) {
        h->kt_ms[cls] += ms;
        h->kt_n[cls] += 1;
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision form: per-frame-set features and the split model.

// Frames per mixed-precision feature row block (the log-likelihood kernel
// works on 128 frames per block).

}  // namespace

namespace gmmhip_detail {

int choose_ll_sub(const gmmhip_ctx *h, int path) {
    int f = h->opt.ll_frames;
    if (f <= 0) {
        gmmhip_tune::Entry e;
        if (gmmhip_tune::lookup(path, GMMHIP_PREC_FP32, h->D, h->M, &e) ||
            gmmhip_tune::lookup(GMMHIP_TUNE_ESTEP, GMMHIP_PREC_FP32, h->D, h->M, &e))
            f = e.ll_frames;
        else
            f = 128;
    }
    int sub = f >= 256 ? 4 : f >= 128 ? 2 : 1;
    if (sub == 4 && h->KP > 128) sub = 2;
    return sub;
}

int choose_mixed_st_nw(const gmmhip_ctx *h, int path) {
    int ct = h->opt.comp_tile;
    if (ct <= 0) {
        gmmhip_tune::Entry e;
        ct = gmmhip_tune::lookup(path, h->opt.precision, h->D, h->M, &e) ? e.comp_tile : 128;
    }
    return mixed_stats_waves(kp_mixed_for_dim(h->D), ct);
}

int choose_mixed_ll_nw(const gmmhip_ctx *h, int path) {
    int f = h->opt.ll_frames;
    if (f <= 0) {
        gmmhip_tune::Entry e;
        f = gmmhip_tune::lookup(path, h->opt.precision, h->D, h->M, &e) ? e.ll_frames : 128;
    }
    return (f >= 256 && kp_mixed_contract(h->D) % 64 == 0) ? 8 : 4;
}

uint64_t tuned_chunk(const gmmhip_ctx *h, int path, int prec) {
    gmmhip_tune::Entry e;
    if (!gmmhip_tune::lookup(path, prec, h->D, h->M, &e)) return 0;
    const uint64_t cf = (uint64_t)e.chunk_frames;
    return (cf + ST_FRAMES_PER_ITER - 1) / ST_FRAMES_PER_ITER * ST_FRAMES_PER_ITER;
}

int pack_split_model(gmmhip_ctx *h, int prec, int KPm, const float *d_zoff, const float *d_zscale,
                     uint16_t *d_out, bool *overflow) {
    DevGuard g(h->dev);
    ModelOperands mo;
    int rc = model_operands(h, false, &mo);
    if (rc) return rc;
    if (!h->d_flag) HIPCHK(h, hipMalloc((void **)&h->d_flag, sizeof(unsigned int)));
    HIPCHK(h, launch_fill_u32(h->d_flag, 0u, 1, h->stream));
    SplitModelArgs a{};
    a.mu = mo.mu;
    a.var = mo.var;
    a.gconst = mo.gconst;
    a.off = d_zoff;
    a.scale = d_zscale;
    a.M = h->M;
    a.D = h->D;
    a.KP = KPm;
    a.Mt32 = (h->M + 31) / 32;
    a.out = d_out;
    a.overflow = h->d_flag;
    HIPCHK(h, launch_pack_split_model(a, prec, h->stream));
    unsigned int flag = 0;
    HIPCHK(h, hipMemcpyAsync(&flag, h->d_flag, sizeof(flag), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *overflow = flag != 0;
    return GMMHIP_OK;
}

}  // namespace gmmhip_detail

namespace {

// Prepares the features and the split model of the current frames and model
// for opts.precision.  *use_mixed stays false (the call then runs the FP32
// form, on the GPU) when the precision is FP32, when a frame value lies
// outside the FP16 range after centring and scaling, or when a model
// coefficient overflows FP16.
int prepare_mixed(gmmhip_ctx *h, bool *use_mixed) {
    *use_mixed = false;
    const int prec = h->opt.precision;
    if (prec != GMMHIP_PREC_BF16X3 && prec != GMMHIP_PREC_FP16X3) return GMMHIP_OK;
    const int D = h->D, KPm = kp_mixed_for_dim(D);
    const uint64_t N = h->N;
    if (h->z_prec != prec || h->z_version != h->frames_version || h->KPm != KPm || h->z_N != N) {
        const uint64_t Npad = pad128(N);
        // per-dimension mean and standard deviation of the frames (FP64)
        const int blocks = (int)std::min<uint64_t>(1024, (N + 255) / 256);
        const uint64_t per_block = (N + blocks - 1) / blocks;
        HIPCHK(h, ensure(&h->d_colpart, &h->colpart_cap, (size_t)blocks * 2 * D));
        HIPCHK(h, launch_colstats(h->d_x, N, D, blocks, per_block, h->d_colpart, h->stream));
        std::vector<double> part((size_t)blocks * 2 * D);
        HIPCHK(h, hipMemcpyAsync(part.data(), h->d_colpart, part.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        h->h_zoff.assign(D, 0.0f);
        h->h_zscale.assign(D, 1.0f);
        for (int d = 0; d < D; d++) {
            double s1 = 0.0, s2 = 0.0;
            for (int b = 0; b < blocks; b++) {
                s1 += part[(size_t)b * 2 * D + d];
                s2 += part[(size_t)b * 2 * D + D + d];
            }
            const double mean = s1 / (double)N, var = s2 / (double)N - mean * mean;
            if (std::isfinite(mean)) h->h_zoff[d] = (float)mean;
            if (std::isfinite(var) && var > 0.0) h->h_zscale[d] = (float)(1.0 / std::sqrt(var));
        }
        HIPCHK(h, ensure(&h->d_zoff, &h->zos_cap, (size_t)D));
        HIPCHK(h, ensure(&h->d_zscale, &h->zscale_cap, (size_t)D));
        HIPCHK(h, hipMemcpyAsync(h->d_zoff, h->h_zoff.data(), D * sizeof(float),
                                 hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->d_zscale, h->h_zscale.data(), D * sizeof(float),
                                 hipMemcpyHostToDevice, h->stream));
        const size_t zn = (size_t)Npad * KPm;
        if (h->z_cap < zn) {
            bump_alloc_generation();
            hipFree(h->d_za_hi);
            hipFree(h->d_za_lo);
            h->d_za_hi = h->d_za_lo = nullptr;
            h->z_cap = 0;
            HIPCHK(h, hipMalloc((void **)&h->d_za_hi, zn * 2));
            HIPCHK(h, hipMalloc((void **)&h->d_za_lo, zn * 2));
            h->z_cap = zn;
        }
        HIPCHK(h, ensure(&h->d_bad, &h->bad_cap, (size_t)N));
        if (!h->d_nrange) HIPCHK(h, hipMalloc((void **)&h->d_nrange, sizeof(unsigned int)));
        HIPCHK(h, hipMemsetAsync(h->d_nrange, 0, sizeof(unsigned int), h->stream));
        PrepArgs p{};
        p.X = h->d_x;
        p.N = N;
        p.Npad = Npad;
        p.D = D;
        p.KP = KPm;
        p.off = h->d_zoff;
        p.scale = h->d_zscale;
        p.za_hi_raw = h->d_za_hi;
        p.za_lo_raw = h->d_za_lo;
        p.bad = h->d_bad;
        p.nrange = h->d_nrange;
        p.zlimit = (prec == GMMHIP_PREC_FP16X3) ? 240.0f : INFINITY;
        HIPCHK(h, launch_prep_frames(p, prec, h->stream));
        unsigned int nrange = 0;
        HIPCHK(h, hipMemcpyAsync(&nrange, h->d_nrange, sizeof(unsigned int), hipMemcpyDeviceToHost,
                                 h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        h->z_range_fallback = nrange > 0;
        h->z_prec = prec;
        h->z_version = h->frames_version;
        h->KPm = KPm;
        h->z_N = N;
        h->z_Npad = Npad;
    }
    if (h->z_range_fallback) return GMMHIP_OK;
    if (h->p16_model != h->model_version || h->p16_z != h->z_version || h->p16_prec != prec) {
        bool overflow = false;
        uint16_t *p16 = static_cast<uint16_t *>(h->d_P16);
        HIPCHK(h, ensure(&p16, &h->P16_cap, (size_t)((h->M + 31) / 32) * KPm * 64));
        h->d_P16 = p16;
        const int rc = pack_split_model(h, prec, KPm, h->d_zoff, h->d_zscale, p16, &overflow);
        if (rc) return rc;
        h->p16_model = h->model_version;
        h->p16_z = h->z_version;
        h->p16_prec = prec;
        h->p16_overflow = overflow;
    }
    *use_mixed = !h->p16_overflow;
    return GMMHIP_OK;
}

MixedArgs make_mixed_args(gmmhip_ctx *h) {
    MixedArgs m{};
    m.N = h->N;
    m.Npad = h->z_Npad;
    m.D = h->D;
    m.KP = h->KPm;
    m.za_hi = h->d_za_hi;
    m.za_lo = h->d_za_lo;
    m.bad = h->d_bad;
    m.P16 = h->d_P16;
    m.M = h->M;
    m.Mt32 = (h->M + 31) / 32;
    m.min_ll = h->opt.min_loglike;
    m.min_gamma = h->opt.min_gamma;
    m.ll_nw = choose_mixed_ll_nw(h, GMMHIP_TUNE_ESTEP);
    m.st_nw = choose_mixed_st_nw(h, GMMHIP_TUNE_ESTEP);
    return m;
}

}  // namespace

extern "C" {

void gmmhip_default_opts(gmmhip_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    apply_defaults(o);
    o->precision = GMMHIP_PREC_FP32;
}

int gmmhip_device_count(int *count) {
    if (!count) return GMMHIP_ERR_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return GMMHIP_OK;
}

int gmmhip_create(int device, const gmmhip_opts *opts, gmmhip_t *out) {
    if (!out) return GMMHIP_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return GMMHIP_ERR_NODEV;
    if (device < 0 || device >= n) return GMMHIP_ERR_ARG;
    gmmhip_ctx *h = new gmmhip_ctx();
    h->dev = device;
    if (opts) h->opt = *opts;
    apply_defaults(&h->opt);
    DevGuard g(device);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return GMMHIP_ERR_HIP;
    }
    if (hipMalloc((void **)&h->d_tot, sizeof(double)) != hipSuccess) {
        hipStreamDestroy(h->stream);
        delete h;
        return GMMHIP_ERR_OOM;
    }
    // the no-allocation entry point's scratch on this device (its address is
    // fetched once per device, with the device current)
    h->noalloc = noalloc_scratch(device);
    *out = h;
    return GMMHIP_OK;
}

int gmmhip_destroy(gmmhip_t h) {
    if (!h) return GMMHIP_OK;
    DevGuard g(h->dev);
    if (h->stream) hipStreamSynchronize(h->stream);
    hipFree(h->d_P);
    hipFree(h->d_off);
    hipFree(h->d_xown);
    hipFree(h->d_LL2);
    hipFree(h->d_LLnat);
    hipFree(h->d_blockLL);
    hipFree(h->d_part);
    hipFree(h->d_red);
    hipFree(h->d_tot);
    hipFree(h->d_za_hi);
    hipFree(h->d_za_lo);
    hipFree(h->d_bad);
    hipFree(h->d_nrange);
    hipFree(h->d_colpart);
    hipFree(h->d_zoff);
    hipFree(h->d_zscale);
    hipFree(h->d_P16);
    batch_state_free(h->batch);
    aux_state_free(h->aux);
    full_state_free(h->full);
    hipFree(h->d_mask);
    hipFree(h->d_mu_f32);
    hipFree(h->d_var_f32);
    hipFree(h->d_gconst);
    hipFree(h->d_flag);
    if (h->h_res) hipHostFree(h->h_res);
    if (h->ev_ok)
        for (auto &e : h->ev) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return GMMHIP_OK;
}

int gmmhip_set_opts(gmmhip_t h, const gmmhip_opts *opts) {
    if (!h || !opts) return GMMHIP_ERR_ARG;
    h->opt = *opts;
    apply_defaults(&h->opt);
    return GMMHIP_OK;
}

const char *gmmhip_last_error(gmmhip_t h) { return h ? h->err.c_str() : "null handle"; }

int gmmhip_kernel_times(gmmhip_t h, double *ms, uint64_t *count, int n) {
    if (!h || !ms || !count || n < 0) return GMMHIP_ERR_ARG;
    for (int i = 0; i < n && i < GMMHIP_KT_COUNT; i++) {
        ms[i] = h->kt_ms[i];
        count[i] = h->kt_n[i];
    }
    return GMMHIP_OK;
}

int gmmhip_reset_kernel_times(gmmhip_t h) {
    if (!h) return GMMHIP_ERR_ARG;
    for (int i = 0; i < GMMHIP_KT_COUNT; i++) {
        h->kt_ms[i] = 0.0;
        h->kt_n[i] = 0;
    }
    return GMMHIP_OK;
}

void *gmmhip_stream(gmmhip_t h) { return h ? (void *)h->stream : nullptr; }

int gmmhip_sync(gmmhip_t h) {
    if (!h) return GMMHIP_ERR_ARG;
    DevGuard g(h->dev);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return GMMHIP_OK;
}

int gmmhip_set_model(gmmhip_t h, int M, int D, const float *w, const float *mu,
                     const float *var) {
    if (!h) return GMMHIP_ERR_ARG;
    if (M <= 0 || D <= 0 || !w || !mu || !var)
        return fail(h, GMMHIP_ERR_ARG, "set_model: bad arguments (M=%d, D=%d)", M, D);
    if (kp_for_dim(D) > KP_MAX)
        return fail(h, GMMHIP_ERR_ARG, "set_model: D=%d exceeds the supported maximum 127", D);
    for (int m = 0; m < M; m++)
        if (!(w[m] > 0.0f))
            return fail(h, GMMHIP_ERR_ARG,
                        "set_model: component %d has weight %g; insert only the active "
                        "(weight > 0) components in RearrangeMixtures order", m, (double)w[m]);
    // Per-component constants in FP64 (insertModel); the reference throws on
    // var < min_var (OL_GMMStatsEstimator.cpp:279-285).
    const double minus_log_2pi = -1.837877066409345;
    std::vector<double> gconst(M);
    for (int m = 0; m < M; m++) {
        double slv = 0.0;
        for (int d = 0; d < D; d++) {
            const float v = var[(size_t)m * D + d];
            if (!(v >= h->opt.min_var))
                return fail(h, GMMHIP_ERR_VAR,
                            "insertModel(): model contains very low variances [%g] (component %d, dim %d)",
                            (double)v, m, d);
            slv += std::log((double)v);
        }
        gconst[m] = minus_log_2pi * D / 2.0 - slv / 2.0 + std::log((double)w[m]);
    }
    // Centring offset: the weighted mean of the component means.
    std::vector<double> offd(D, 0.0);
    double wsum = 0.0;
    for (int m = 0; m < M; m++) wsum += w[m];
    for (int m = 0; m < M; m++)
        for (int d = 0; d < D; d++) offd[d] += (double)w[m] * mu[(size_t)m * D + d] / wsum;
    std::vector<float> off(D);
    for (int d = 0; d < D; d++) off[d] = (float)offd[d];

    const int KP = kp_for_dim(D);
    const int KK = KP / 4;
    const int Mt = (M + 15) / 16;
    const double log2e = 1.4426950408889634;
    // MFMA B-operand packing P[t][lane][kk] = B[q*KK + kk][16t + c] with
    // lane = 16q + c (the contraction index mapping of the E-step kernels).
    std::vector<float> P((size_t)Mt * 64 * KK, 0.0f);
    auto pidx = [&](int m, int k) {
        const int t = m / 16, c = m % 16, qq = k / KK, kk = k % KK;
        return ((size_t)t * 64 + qq * 16 + c) * KK + kk;
    };
    // Padding components (M..16*Mt-1): constant -1e30 (log2 units), so their
    // log-likelihood vanishes in every exp2 without a lane mask.
    for (int m = M; m < Mt * 16; m++) P[pidx(m, 0)] = -1e30f;
    // Coefficients (log2 units) b_d = mu'_d ivar_d, a_d = -ivar_d / 2 are
    // rounded to FP32; the constant is then chosen in FP64 so that the
    // rounded coefficients give the exact log-likelihood at the component's
    // own mean, c = l(mu') - sum_d (b_d mu'_d + a_d mu'_d^2), and it is
    // carried as hi + lo: hi in column 0, lo in column 2D+1 (a padding
    // column, KP >= 2D+2 always, whose frame feature is 1).  Without this the
    // rounding of the coefficients and of the constant (|c| up to hundreds)
    // shifts every posterior of a component by the same ~1e-5 relative.
    for (int m = 0; m < M; m++) {
        double cq = 0.0, cadj = 0.0;
        for (int d = 0; d < D; d++) {
            const double iv = 1.0 / (double)var[(size_t)m * D + d];
            const double mc = (double)mu[(size_t)m * D + d] - (double)off[d];
            const double b = mc * iv * log2e, a = -0.5 * iv * log2e;
            const float bf = (float)b, af = (float)a;
            P[pidx(m, 1 + d)] = bf;
            P[pidx(m, 1 + D + d)] = af;
            cq += mc * mc * iv;
            cadj += (b - (double)bf) * mc + (a - (double)af) * mc * mc;
        }
        const double c = (gconst[m] - 0.5 * cq) * log2e + cadj;
        const float chi = (float)c;
        P[pidx(m, 0)] = chi;
        P[pidx(m, 2 * D + 1)] = (float)(c - (double)chi);
    }
    DevGuard g(h->dev);
    HIPCHK(h, ensure(&h->d_P, &h->P_cap, P.size()));
    HIPCHK(h, ensure(&h->d_off, &h->off_cap, (size_t)D));
    HIPCHK(h, hipMemcpyAsync(h->d_P, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d_off, off.data(), D * sizeof(float), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->M = M;
    h->D = D;
    h->KP = KP;
    h->Mt = Mt;
    h->h_off = off;
    h->m_w.assign(w, w + M);
    h->m_mu.assign(mu, mu + (size_t)M * D);
    h->m_var.assign(var, var + (size_t)M * D);
    h->m_gconst = gconst;
    h->full_model = false;
    h->model_version++;
    return GMMHIP_OK;
}

int gmmhip_set_frames_contig(gmmhip_t h, const float *x, uint64_t N, int D, uint64_t ld) {
    if (!h) return GMMHIP_ERR_ARG;
    if (!x || N == 0 || D <= 0 || ld < (uint64_t)D)
        return fail(h, GMMHIP_ERR_ARG, "set_frames: bad arguments");
    DevGuard g(h->dev);
    HIPCHK(h, ensure(&h->d_xown, &h->xown_cap, (size_t)N * D));
    if (ld == (uint64_t)D) {
        HIPCHK(h, hipMemcpy(h->d_xown, x, (size_t)N * D * sizeof(float), hipMemcpyHostToDevice));
    } else {
        HIPCHK(h, hipMemcpy2D(h->d_xown, D * sizeof(float), x, ld * sizeof(float),
                              D * sizeof(float), N, hipMemcpyHostToDevice));
    }
    h->d_x = h->d_xown;
    h->N = N;
    h->ld = D;
    h->xD = D;
    h->frames_version++;
    return GMMHIP_OK;
}

int gmmhip_set_frames(gmmhip_t h, const float *const *rows, uint64_t N, int D) {
    if (!h) return GMMHIP_ERR_ARG;
    if (!rows || N == 0 || D <= 0) return fail(h, GMMHIP_ERR_ARG, "set_frames: bad arguments");
    bool contig = true;
    for (uint64_t i = 1; i < N && contig; i++) contig = (rows[i] == rows[0] + i * (uint64_t)D);
    if (contig) return gmmhip_set_frames_contig(h, rows[0], N, D, D);
    DevGuard g(h->dev);
    HIPCHK(h, ensure(&h->d_xown, &h->xown_cap, (size_t)N * D));
    // Pack row blocks through a bounded host buffer.
    const uint64_t blk = std::max<uint64_t>(1, (64ull << 20) / (sizeof(float) * D));
    std::vector<float> buf((size_t)std::min<uint64_t>(blk, N) * D);
    for (uint64_t s = 0; s < N; s += blk) {
        const uint64_t e = std::min<uint64_t>(N, s + blk);
        for (uint64_t i = s; i < e; i++)
            std::memcpy(&buf[(size_t)(i - s) * D], rows[i], sizeof(float) * D);
        HIPCHK(h, hipMemcpy(h->d_xown + (size_t)s * D, buf.data(), (size_t)(e - s) * D * sizeof(float),
                            hipMemcpyHostToDevice));
    }
    h->d_x = h->d_xown;
    h->N = N;
    h->ld = D;
    h->xD = D;
    h->frames_version++;
    return GMMHIP_OK;
}

int gmmhip_set_frames_device(gmmhip_t h, const float *dx, uint64_t N, int D, uint64_t ld) {
    if (!h) return GMMHIP_ERR_ARG;
    if (!dx || N == 0 || D <= 0 || ld < (uint64_t)D)
        return fail(h, GMMHIP_ERR_ARG, "set_frames_device: bad arguments");
    if (ld != (uint64_t)D) {
        // The kernels stream frames as contiguous rows: compact a strided
        // view once into the handle's own buffer.
        DevGuard g(h->dev);
        HIPCHK(h, ensure(&h->d_xown, &h->xown_cap, (size_t)N * D));
        HIPCHK(h, hipMemcpy2DAsync(h->d_xown, D * sizeof(float), dx, ld * sizeof(float),
                                   D * sizeof(float), N, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        dx = h->d_xown;
        ld = D;
    }
    h->d_x = dx;
    h->N = N;
    h->ld = ld;
    h->xD = D;
    h->frames_version++;
    return GMMHIP_OK;
}

int gmmhip_loglikes_device(gmmhip_t h, float *d_out) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!d_out) return fail(h, GMMHIP_ERR_ARG, "loglikes_device: null output");
    DevGuard g(h->dev);
    if ((rc = mask_check(h))) return rc;
    if (h->full_model || h->d_mask) return gemm_loglikes_device(h, d_out);
    // padded: the mixed statistics kernel stages 32-frame steps of LL2
    HIPCHK(h, ensure(&h->d_LL2, &h->LL2_cap, pad128(h->N)));
    HIPCHK(h, ensure(&h->d_blockLL, &h->blockLL_cap, ll_blocks(h->N, 1)));
    bool mixed = false;
    if ((rc = prepare_mixed(h, &mixed))) return rc;
    h->last_used_mixed = mixed;
    if (mixed) {
        MixedArgs m = make_mixed_args(h);
        m.LL2 = h->d_LL2;
        m.LLnat = d_out;
        m.blockLL = h->d_blockLL;
        h->used_ll_frames = mixed_ll_frames_per_block(m.D, m.ll_nw);
        HIPCHK(h, launch_loglik_mixed(m, h->opt.precision, h->stream));
        return GMMHIP_OK;
    }
    EstepArgs a = make_args(h);
    a.LL2 = h->d_LL2;
    a.LLnat = d_out;
    a.blockLL = h->d_blockLL;
    a.ll_sub = choose_ll_sub(h, GMMHIP_TUNE_ESTEP);
    h->used_ll_frames = 64 * a.ll_sub;
    HIPCHK(h, launch_loglik_fp32(a, h->stream));
    return GMMHIP_OK;
}

int gmmhip_loglikes(gmmhip_t h, float *out) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!out) return fail(h, GMMHIP_ERR_ARG, "loglikes: null output");
    DevGuard g(h->dev);
    HIPCHK(h, ensure(&h->d_LLnat, &h->LLnat_cap, h->N));
    const bool prof = h->opt.profile != 0;
    if (prof) {
        if ((rc = ensure_events(h))) return rc;
        HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
    }
    rc = gmmhip_loglikes_device(h, h->d_LLnat);
    if (rc) return rc;
    if (prof) HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
    HIPCHK(h, hipMemcpyAsync(out, h->d_LLnat, h->N * sizeof(float), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (prof) add_time(h, GMMHIP_KT_LOGLIK, 0, 1);
    // NaN check of compVecsLogLikeMT (OL_GMMStatsEstimator_CUDA.cpp:383-387);
    // NaN frames are already clamped to min_loglike on the device.
    return GMMHIP_OK;
}

int gmmhip_accumulate(gmmhip_t h, int flags, double *n, double *F, double *S, double *tot_ll,
                      uint64_t *nframes) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!n || !tot_ll || !nframes || ((flags & GMMHIP_ACC_MEAN) && !F) ||
        ((flags & GMMHIP_ACC_VAR) && !S))
        return fail(h, GMMHIP_ERR_ARG, "accumulate: null output for a requested statistic");
    DevGuard g(h->dev);
    if ((rc = mask_check(h))) return rc;
    if (h->full_model) return full_accumulate(h, flags, n, F, S, tot_ll, nframes);
    if ((flags & GMMHIP_ACC_FULLVAR) && (flags & GMMHIP_ACC_VAR))
        return diag_full_accumulate(h, flags, n, F, S, tot_ll, nframes);
    if (h->d_mask) {
        // Gaussian selection: the GEMM form with the diagonal operand
        if ((rc = gemm_operand_for_model(h))) return rc;
        return full_accumulate(h, flags, n, F, S, tot_ll, nframes);
    }
    // padded: the mixed statistics kernel stages 32-frame steps of LL2
    HIPCHK(h, ensure(&h->d_LL2, &h->LL2_cap, pad128(h->N)));
    HIPCHK(h, ensure(&h->d_blockLL, &h->blockLL_cap, ll_blocks(h->N, 1)));
    bool mixed = false;
    if ((rc = prepare_mixed(h, &mixed))) return rc;
    h->last_used_mixed = mixed;

    const int D = h->D;
    const int KPx = mixed ? h->KPm : h->KP;
    const int tile = mixed ? 32 : 16;
    const int Mtx = (h->M + tile - 1) / tile;
    const int st_nw = mixed ? choose_mixed_st_nw(h, GMMHIP_TUNE_ESTEP) : 4;
    const int ngroups = (Mtx + st_nw - 1) / st_nw;
    const uint64_t chunk_frames = pick_chunk_frames(h, ngroups, mixed);
    const int ll_sub =
        mixed ? mixed_ll_frames_per_block(h->D, choose_mixed_ll_nw(h, GMMHIP_TUNE_ESTEP)) / 64
              : choose_ll_sub(h, GMMHIP_TUNE_ESTEP);
    h->used_ll_frames = 64 * ll_sub;
    h->used_chunk = (int)chunk_frames;
    h->used_comp_tile = mixed ? 32 * st_nw : 0;
    const int nchunks = (int)((h->N + chunk_frames - 1) / chunk_frames);
    const size_t Mp = (size_t)Mtx * tile;
    const size_t E = Mp * KPx;
    HIPCHK(h, ensure(&h->d_part, &h->part_cap, (size_t)nchunks * E));
    HIPCHK(h, ensure(&h->d_red, &h->red_cap, E));
    if (h->h_res_cap < E + 1) {
        if (h->h_res) HIPCHK(h, hipHostFree(h->h_res));
        h->h_res = nullptr;
        h->h_res_cap = 0;
        HIPCHK(h, hipHostMalloc((void **)&h->h_res, (E + 1) * sizeof(double)));
        h->h_res_cap = E + 1;
    }

    const bool prof = h->opt.profile != 0;
    if (prof) {
        if ((rc = ensure_events(h))) return rc;
        HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
    }
    if (mixed) {
        MixedArgs m = make_mixed_args(h);
        m.LL2 = h->d_LL2;
        m.LLnat = nullptr;
        m.blockLL = h->d_blockLL;
        m.chunk_frames = chunk_frames;
        m.nchunks = nchunks;
        m.ngroups = ngroups;
        m.partials = h->d_part;
        HIPCHK(h, launch_loglik_mixed(m, h->opt.precision, h->stream));
        if (prof) HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
        HIPCHK(h, launch_stats_mixed(m, h->opt.precision, h->stream));
    } else {
        EstepArgs a = make_args(h);
        a.LL2 = h->d_LL2;
        a.LLnat = nullptr;
        a.blockLL = h->d_blockLL;
        a.ngroups = ngroups;
        a.chunk_frames = chunk_frames;
        a.nchunks = nchunks;
        a.partials = h->d_part;
        a.ll_sub = ll_sub;
        HIPCHK(h, launch_loglik_fp32(a, h->stream));
        if (prof) HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
        HIPCHK(h, launch_stats_fp32(a, h->stream));
    }
    if (prof) HIPCHK(h, hipEventRecord(h->ev[2], h->stream));
    HIPCHK(h, launch_reduce_partials(h->d_part, nchunks, E, h->d_red, h->stream));
    HIPCHK(h, launch_sum_doubles(h->d_blockLL, ll_blocks(h->N, ll_sub), h->d_tot, h->stream));
    if (prof) HIPCHK(h, hipEventRecord(h->ev[3], h->stream));
    HIPCHK(h, hipMemcpyAsync(h->h_res, h->d_red, E * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipMemcpyAsync(h->h_res + E, h->d_tot, sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (prof) {
        add_time(h, GMMHIP_KT_LOGLIK, 0, 1);
        add_time(h, GMMHIP_KT_STATS, 1, 2);
        add_time(h, GMMHIP_KT_REDUCE, 2, 3);
    }
    const double tot = h->h_res[E];
    if (std::isnan(tot))
        return fail(h, GMMHIP_ERR_NAN, "accumulateStatsMT(): overall log-likelihood = NaN!");

    // Undo the centring (and scaling) of the features, in FP64:
    // x = z/s + o  ->  F = F'/s + n o,  S = S'/s^2 + 2 o F'/s + n o^2.
    for (int m = 0; m < h->M; m++) {
        const double *r = h->h_res + (size_t)m * KPx;
        const double nm = r[0];
        n[m] += nm;
        for (int d = 0; d < D; d++) {
            const double o = mixed ? (double)h->h_zoff[d] : (double)h->h_off[d];
            const double sc = mixed ? 1.0 / (double)h->h_zscale[d] : 1.0;
            const double f1 = r[1 + d] * sc, s1 = r[1 + D + d] * sc * sc;
            if (flags & GMMHIP_ACC_MEAN) F[(size_t)m * D + d] += f1 + nm * o;
            if (flags & GMMHIP_ACC_VAR) S[(size_t)m * D + d] += s1 + 2.0 * o * f1 + nm * o * o;
        }
    }
    *tot_ll += tot;
    *nframes += h->N;
    return GMMHIP_OK;
}

int gmmhip_tune_load(const char *path) { return gmmhip_tune::load(path); }

int gmmhip_tune_lookup(int path, int precision, int D, int M, int *ll_frames, int *chunk_frames,
                       int *comp_tile) {
    if (D <= 0 || M <= 0) return GMMHIP_ERR_ARG;
    gmmhip_tune::Entry e;
    if (!gmmhip_tune::lookup(path, precision, D, M, &e)) return GMMHIP_ERR_STATE;
    if (ll_frames) *ll_frames = e.ll_frames;
    if (chunk_frames) *chunk_frames = e.chunk_frames;
    if (comp_tile) *comp_tile = e.comp_tile;
    return GMMHIP_OK;
}

int gmmhip_tiles_used(gmmhip_t h, int *ll_frames, int *chunk_frames, int *comp_tile) {
    if (!h) return GMMHIP_ERR_ARG;
    if (ll_frames) *ll_frames = h->used_ll_frames;
    if (chunk_frames) *chunk_frames = h->used_chunk;
    if (comp_tile) *comp_tile = h->used_comp_tile;
    return GMMHIP_OK;
}

}  // extern "C"
