// C ABI of the host layer (include/gmmhost.h).
#include "../../../include/gmmhost.h"

#include <cstring>
#include <memory>
#include <string>

#include "feature_io.h"
#include "gmm_model.h"
#include "model_adapt.h"

struct gmmhost_model {
    gmm::GMModel m;
};

struct gmmhost_features {
    gmm::Features f;
};

namespace {
thread_local std::string g_err;
int fail(const std::string &msg, int code = 1) {
    g_err = msg;
    return code;
}

// Statistics wrapper for the update rules (GMMStats<double> view).
gmm::GMMStats<double> make_stats(int M, int D, const double *n, const double *F, const double *S,
                                 uint64_t tot, bool full = false, const double *aux = nullptr) {
    gmm::GMMStats<double> st(D, M, F != nullptr, S != nullptr, full, aux != nullptr);
    st.reset();
    std::memcpy(st.probData(), n, sizeof(double) * M);
    if (F) std::memcpy(st.meanData(), F, sizeof(double) * M * D);
    if (S) std::memcpy(st.varData(), S, sizeof(double) * M * st.getVarSize());
    if (aux) std::memcpy(st.auxData(), aux, sizeof(double) * M * (D + 2));
    st._totalAccSamples = tot;
    return st;
}

// CModelAdapt exposes the statistics it works on through Stats().
struct RuleHost : gmm::CModelAdapt {};
}  // namespace

extern "C" {

const char *gmmhost_last_error(void) { return g_err.c_str(); }

static int feat_rc(int rc, const std::string &err) {
    if (rc == 0) return 0;
    return fail(err, rc == -1 ? GMMHOST_FEAT_MISSING : 1);
}

int gmmhost_features_load(const char *path, int type, uint32_t dwnsmp, gmmhost_features_t *out) {
    if (!path || !out) return fail("gmmhost_features_load: null argument");
    try {
        std::unique_ptr<gmmhost_features> h(new gmmhost_features);
        std::string err;
        const int rc = gmm::LoadFeatures(path, type, h->f, &err);
        if (rc) return feat_rc(rc, err);
        if (dwnsmp > 1 && h->f.N > 0) {
            gmm::Features k;
            k.D = h->f.D;
            k.samplePeriod = h->f.samplePeriod;
            for (uint32_t t = 0; t < h->f.N; t += dwnsmp) {
                k.data.insert(k.data.end(), h->f.data.begin() + (size_t)t * k.D,
                              h->f.data.begin() + (size_t)(t + 1) * k.D);
                k.N++;
            }
            h->f = std::move(k);
        }
        *out = h.release();
        return 0;
    } catch (std::exception &e) {
        return fail(e.what());
    }
}

int gmmhost_features_load_ranges(const char *path, int type, const uint32_t *ranges, int nranges,
                                 uint32_t sample_period, gmmhost_features_t *out) {
    if (!path || !out || nranges < 0 || (nranges > 0 && !ranges))
        return fail("gmmhost_features_load_ranges: bad argument");
    try {
        gmm::element_ranges r;
        for (int k = 0; k < nranges; k++) r.push_back({ranges[2 * k], ranges[2 * k + 1]});
        std::unique_ptr<gmmhost_features> h(new gmmhost_features);
        std::string err;
        const int rc = gmm::LoadFeatureRanges(path, type, r, sample_period, h->f, &err);
        if (rc) return feat_rc(rc, err);
        *out = h.release();
        return 0;
    } catch (std::exception &e) {
        return fail(e.what());
    }
}

int gmmhost_features_dims(gmmhost_features_t f, uint32_t *N, uint32_t *D, uint32_t *sample_period) {
    if (!f) return fail("gmmhost_features_dims: null handle");
    if (N) *N = f->f.N;
    if (D) *D = f->f.D;
    if (sample_period) *sample_period = f->f.samplePeriod;
    return 0;
}

int gmmhost_features_get(gmmhost_features_t f, float *x) {
    if (!f || (!x && f->f.N)) return fail("gmmhost_features_get: null argument");
    if (f->f.N) std::memcpy(x, f->f.data.data(), sizeof(float) * f->f.data.size());
    return 0;
}

void gmmhost_features_free(gmmhost_features_t f) { delete f; }

int gmmhost_features_save(const char *path, int type, const float *x, uint32_t N, uint32_t D,
                          uint32_t sample_period) {
    if (!path || (!x && N && D)) return fail("gmmhost_features_save: null argument");
    int rc;
    switch (type) {
    case 0: rc = gmm::SaveSVES(path, x, N, D); break;
    case 1: rc = gmm::SaveHTK(path, x, N, D, sample_period); break;
    case 2: rc = gmm::SaveRaw(path, x, N, D); break;
    default: return fail("gmmhost_features_save: unknown type");
    }
    return rc ? fail(std::string("cannot write ") + path) : 0;
}

int gmmhost_model_load(const char *path, int txt, gmmhost_model_t *out) {
    if (!path || !out) return fail("null argument");
    gmmhost_model *h = new gmmhost_model();
    if (h->m.Load(path, txt != 0) != 0) {
        const std::string e = h->m.LastError();
        delete h;
        return fail(e);
    }
    *out = h;
    return 0;
}

int gmmhost_model_save(gmmhost_model_t m, const char *path, int txt) {
    if (!m || !path) return fail("null argument");
    if (m->m.Save(path, txt != 0) != 0) return fail(m->m.LastError());
    return 0;
}

int gmmhost_model_create(int M, int D, const float *w, const float *mu, const float *var,
                         gmmhost_model_t *out) {
    if (M <= 0 || D <= 0 || !w || !mu || !var || !out) return fail("bad arguments");
    gmmhost_model *h = new gmmhost_model();
    h->m = gmm::GMModel(D, M);
    for (int k = 0; k < M; k++) h->m.SetMixParam(k, mu + (size_t)k * D, var + (size_t)k * D, w[k]);
    *out = h;
    return 0;
}

int gmmhost_model_create_full(int M, int D, const float *w, const float *mu, const float *cov,
                              gmmhost_model_t *out) {
    if (M <= 0 || D <= 0 || !w || !mu || !cov || !out) return fail("bad arguments");
    gmmhost_model *h = new gmmhost_model();
    h->m = gmm::GMModel(D, M, true);
    for (int k = 0; k < M; k++)
        h->m.SetMixParam(k, mu + (size_t)k * D, cov + (size_t)k * D * D, w[k]);
    *out = h;
    return 0;
}

int gmmhost_model_full(gmmhost_model_t m, int *full) {
    if (!m || !full) return fail("null argument");
    *full = m->m.GetFullCovStatus() ? 1 : 0;
    return 0;
}

int gmmhost_model_dims(gmmhost_model_t m, int *M, int *D) {
    if (!m || !M || !D) return fail("null argument");
    *M = m->m.GetNumberOfMixtures();
    *D = m->m.GetDimension();
    return 0;
}

int gmmhost_model_get(gmmhost_model_t m, float *w, float *mu, float *var) {
    if (!m) return fail("null argument");
    if (w) std::memcpy(w, m->m.weights().data(), sizeof(float) * m->m.weights().size());
    if (mu) std::memcpy(mu, m->m.means().data(), sizeof(float) * m->m.means().size());
    if (var) std::memcpy(var, m->m.vars().data(), sizeof(float) * m->m.vars().size());
    return 0;
}

int gmmhost_model_rearrange(gmmhost_model_t m, int *nzero) {
    if (!m) return fail("null argument");
    const unsigned nz = m->m.RearrangeMixtures();
    if (nzero) *nzero = (int)nz;
    return 0;
}

void gmmhost_model_free(gmmhost_model_t m) { delete m; }

int gmmhost_update_em(gmmhost_model_t m, const double *n, const double *F, const double *S,
                      uint64_t tot_samples, float min_var, int upweight, int upmean, int upvar,
                      int *left) {
    if (!m || !n) return fail("null argument");
    try {
        const int M = m->m.GetNumberOfMixtures(), D = m->m.GetDimension();
        gmm::GMMStats<double> st = make_stats(M, D, n, F, S, tot_samples, m->m.GetFullCovStatus());
        RuleHost ad;
        gmm::CModelAdapt::OPTIONS o;
        o._min_var = min_var;
        ad.SetOptions(o);
        const unsigned l = ad.UpdateModelEM(&m->m, st, upweight != 0, upmean != 0, upvar != 0);
        if (left) *left = (int)l;
    } catch (std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int gmmhost_update_em_robust(gmmhost_model_t m, const double *n, const double *F, const double *S,
                             const double *aux, uint64_t tot_samples, float min_var, int upweight,
                             int upmean, int upvar, int *left) {
    if (!m || !n || !F || !S || !aux) return fail("null argument");
    try {
        const int M = m->m.GetNumberOfMixtures(), D = m->m.GetDimension();
        gmm::GMMStats<double> st =
            make_stats(M, D, n, F, S, tot_samples, m->m.GetFullCovStatus(), aux);
        RuleHost ad;
        gmm::CModelAdapt::OPTIONS o;
        o._min_var = min_var;
        o._robustCov = true;
        ad.SetOptions(o);
        const unsigned l = ad.UpdateModelEM(&m->m, st, upweight != 0, upmean != 0, upvar != 0);
        if (left) *left = (int)l;
    } catch (std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int gmmhost_make_robust_stats(int M, int D, int full, const double *n, const double *F, double *S,
                              const double *aux, double *nind) {
    if (M <= 0 || D <= 0 || !n || !F || !S || !aux) return fail("bad argument");
    try {
        gmm::GMMStats<double> st = make_stats(M, D, n, F, S, 0, full != 0, aux);
        const std::vector<double> ni = gmm::CModelAdapt::MakeRobustStats(st);
        std::memcpy(S, st.varData(), sizeof(double) * M * st.getVarSize());
        if (nind) std::memcpy(nind, ni.data(), sizeof(double) * M);
    } catch (std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int gmmhost_update_map(gmmhost_model_t m, gmmhost_model_t ubm, const double *n, const double *F,
                       const double *S, uint64_t tot_samples, int tau, float min_var, int upweight,
                       int upmean, int upvar) {
    if (!m || !ubm || !n) return fail("null argument");
    try {
        const int M = m->m.GetNumberOfMixtures(), D = m->m.GetDimension();
        RuleHost ad;
        gmm::CModelAdapt::OPTIONS o;
        o._min_var = min_var;
        o._tau = tau;
        o._AdaptWeightBool = upweight != 0;
        o._AdaptMeanBool = upmean != 0;
        o._AdaptVarBool = upvar != 0;
        ad.SetOptions(o);
        ad.Stats() = make_stats(M, D, n, F, S, tot_samples, m->m.GetFullCovStatus());
        ad.UpdateModelMAP(&m->m, &ubm->m);
    } catch (std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int gmmhost_mllr_transform(gmmhost_model_t m, gmmhost_model_t ubm, const double *n,
                           const double *F, float min_var, double *W) {
    if (!m || !n || !F || !W) return fail("null argument");
    try {
        const int M = m->m.GetNumberOfMixtures(), D = m->m.GetDimension(), E = D + 1;
        // CheckInvertibility (OL_ModelAdapt.cpp:1851-1878)
        int active = 0;
        double occ = 0.0;
        for (int k = 0; k < M; k++)
            if (n[k] > 0) {
                active++;
                occ += n[k];
            }
        if (M < D + 1 || active < D + 1 || occ * D < (double)D * (D + 1))
            return fail("CheckInvertibility(): MLLR is under-determined for these statistics", 7);
        RuleHost ad;
        ad.Stats() = make_stats(M, D, n, F, nullptr, 0);
        std::vector<double> Gi, ki;
        if (ubm) gmm::CModelAdapt::prepareKiGiInitMatrices(ubm->m, min_var, Gi, ki);
        std::vector<double> G((size_t)E * E), k(E);
        ad.SetNumMix(M);
        for (int i = 0; i < D; i++) {
            ad.ComputeKiGi(m->m, i, Gi.empty() ? nullptr : &Gi[(size_t)i * E * E],
                           ki.empty() ? nullptr : &ki[(size_t)i * E], G.data(), k.data());
            gmm::spd_solve(E, G.data(), k.data(), W + (size_t)i * E);
        }
    } catch (std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int gmmhost_fmllr_transform(gmmhost_model_t m, const double *n, const double *F, const double *S,
                            uint64_t tot_samples, int iters, double *W) {
    if (!m || !n || !F || !S || !W) return fail("null argument");
    try {
        const int M = m->m.GetNumberOfMixtures(), D = m->m.GetDimension();
        gmm::GMMStats<double> st = make_stats(M, D, n, F, S, tot_samples, true);
        const gmm::MLLRTransform T = gmm::CModelAdapt::EstimateW_fMLLR(m->m, st, iters);
        std::memcpy(W, T.W.data(), sizeof(double) * T.W.size());
    } catch (std::exception &e) {
        return fail(e.what());
    }
    return 0;
}

int gmmhost_mllr_apply(gmmhost_model_t m, const double *W) {
    if (!m || !W) return fail("null argument");
    gmm::MLLRTransform T;
    T.D = m->m.GetDimension();
    T.W.assign(W, W + (size_t)T.D * (T.D + 1));
    m->m = *gmm::CModelAdapt::AdaptModelMLLR(m->m, T);
    return 0;
}

}  // extern "C"
