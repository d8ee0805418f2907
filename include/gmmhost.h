/*
 * gmmhost.h -- C ABI of the host layer around the gfx950 accumulator: the
 * `.gmm` model file, the RearrangeMixtures component order and the EM / MAP /
 * MLLR update rules that turn statistics into models, and the feature files
 * the trainer reads.  No device work; these entry points run on any host.  Reference paths are relative to
 * GMM/lib_bmssr/ of GMMTeam/GMM.
 */
#ifndef GMMHOST_H
#define GMMHOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gmmhost_model *gmmhost_model_t;

/* GMModel::Load (model/OL_GMModel.cpp:1048, ReadTXT :627-753, ReadBIN
 * :757-853).  Returns 0 or 1 (message: gmmhost_last_error()). */
int gmmhost_model_load(const char *path, int txt, gmmhost_model_t *out);
/* GMModel::Save (WriteTXT :475-547, WriteBIN :551-624); zero-weight
 * components are skipped. */
int gmmhost_model_save(gmmhost_model_t m, const char *path, int txt);
/* A model from arrays w[M], mu[M*D], var[M*D] (GMModel(dim, nummix) +
 * SetMixParam, :120-156). */
int gmmhost_model_create(int M, int D, const float *w, const float *mu, const float *var,
                         gmmhost_model_t *out);
/* A full-covariance model from w[M], mu[M*D], cov[M*D*D] (GMModel(dim,
 * nummix, full_cov = true) + SetMixParam). */
int gmmhost_model_create_full(int M, int D, const float *w, const float *mu, const float *cov,
                              gmmhost_model_t *out);
int gmmhost_model_dims(gmmhost_model_t m, int *M, int *D);
/* *full = 1 for a full-covariance model (GetFullCovStatus, OL_GMModel.h:100). */
int gmmhost_model_full(gmmhost_model_t m, int *full);
/* var receives M*D variances, or M*D*D for a full-covariance model. */
int gmmhost_model_get(gmmhost_model_t m, float *w, float *mu, float *var);
/* RearrangeMixtures (:857-885): zero-weight components to the end,
 * reversed; *nzero receives their count. */
int gmmhost_model_rearrange(gmmhost_model_t m, int *nzero);
void gmmhost_model_free(gmmhost_model_t m);

/* UpdateModelEM (trainer/OL_ModelAdapt.cpp:823-949) on model m with
 * statistics n[M], F[M*D], S[M*D] over tot_samples frames (a full-covariance
 * model takes S[M*D(D+1)/2], the packed upper triangle of sum g x x^T);
 * returns the reference's count of components with non-zero weight on entry
 * through *left. */
int gmmhost_update_em(gmmhost_model_t m, const double *n, const double *F, const double *S,
                      uint64_t tot_samples, float min_var, int upweight, int upmean, int upvar,
                      int *left);
/* UpdateModelEM with _robustCov (OL_ModelAdapt.cpp:848-857, :880): the
 * statistics are made robust (MakeRobustStats) from the auxiliary statistics
 * aux[M*(D+2)] (per component [aux (D) | aux2 | aux3], as gmmhip_accumulate_aux
 * returns them) and the independent-frame counts replace the occupancies in
 * the minimum-data test. */
int gmmhost_update_em_robust(gmmhost_model_t m, const double *n, const double *F, const double *S,
                             const double *aux, uint64_t tot_samples, float min_var, int upweight,
                             int upmean, int upvar, int *left);
/* MakeRobustStats (OL_ModelAdapt.cpp:468-588) on statistics of M components:
 * S (M*D diagonal, or M*D(D+1)/2 packed upper when full) is rewritten in
 * place; nind[M] (may be NULL) receives GetNIndFrames' counts.  Fails when a
 * global variance is zero. */
int gmmhost_make_robust_stats(int M, int D, int full, const double *n, const double *F, double *S,
                              const double *aux, double *nind);
/* UpdateModelMAP (OL_ModelAdapt.cpp:993-1106): m starts as (a copy of) the
 * UBM ubm and is adapted with relevance factor tau (full-covariance models:
 * packed S as for gmmhost_update_em). */
int gmmhost_update_map(gmmhost_model_t m, gmmhost_model_t ubm, const double *n, const double *F,
                       const double *S, uint64_t tot_samples, int tau, float min_var,
                       int upweight, int upmean, int upvar);
/* Global MLLR transform W[D*(D+1)] (row-major [A|b]) of model m from
 * statistics n[M], F[M*D] (ComputeW_MLLR + ComputeKiGi_MLLR, :1417-1581;
 * a full-covariance model contributes the diagonal of its covariances),
 * optionally with the UBM prior of prepareKiGiInitMatrices (:1484-1523,
 * ubm may be NULL).  Returns 7 (GMMHIP_ERR_MLLR) when CheckInvertibility
 * (:1851-1878) rejects the statistics. */
int gmmhost_mllr_transform(gmmhost_model_t m, gmmhost_model_t ubm, const double *n,
                           const double *F, float min_var, double *W);
/* fMLLR transform W[D*(D+1)] = [A | b] (ComputeW_fMLLR + EstimateAb_fmllr +
 * ComputeKiGi_fMLLR, OL_ModelAdapt.cpp:1614-1840) of model m (its diagonal
 * variances) from statistics n[M], F[M*D] and full second-order S[M*D(D+1)/2]
 * (packed upper, as gmmhip_accumulate with GMMHIP_ACC_FULLVAR gives them)
 * over tot_samples frames, iters row-update sweeps (the reference: 20). */
int gmmhost_fmllr_transform(gmmhost_model_t m, const double *n, const double *F, const double *S,
                            uint64_t tot_samples, int iters, double *W);
/* AdaptModelMLLR (:1583-1612): means of m replaced by A mu + b. */
int gmmhost_mllr_apply(gmmhost_model_t m, const double *W);

/* Feature files (param/OL_Param.h:59-127).  type: 0 SVES, 1 HTK, 2 RAW
 * (--data-T).  Loaders return 0, GMMHOST_FEAT_MISSING when the file cannot
 * be opened (the reference returns -1 and its callers skip the file) or 1
 * for a malformed / truncated file or bad interval (the reference throws);
 * an empty file loads as N = D = 0. */
#define GMMHOST_FEAT_MISSING 2
typedef struct gmmhost_features *gmmhost_features_t;
/* Param::Load(filename, loadType, dwnsmp) (OL_Param.cpp:222-244; SVES
 * :246-354, RAW :358-453, HTK :477-576): every dwnsmp-th frame from the
 * first (dwnsmp 0 or 1: all). */
int gmmhost_features_load(const char *path, int type, uint32_t dwnsmp, gmmhost_features_t *out);
/* Param::Load with frame intervals (OL_Param.cpp:581-1037): nranges
 * intervals [ranges[2k], ranges[2k+1]) in sample-period units, ascending;
 * the period is the HTK header's, or sample_period for SVES and RAW.  The
 * reference's HTK variant (:897-1000) is followed for every type (its SVES
 * and RAW variants count each interval twice). */
int gmmhost_features_load_ranges(const char *path, int type, const uint32_t *ranges, int nranges,
                                 uint32_t sample_period, gmmhost_features_t *out);
/* Frames, dimension and (HTK) sample period of loaded features. */
int gmmhost_features_dims(gmmhost_features_t f, uint32_t *N, uint32_t *D, uint32_t *sample_period);
/* Copy the N x D frames (row-major) into x. */
int gmmhost_features_get(gmmhost_features_t f, float *x);
void gmmhost_features_free(gmmhost_features_t f);
/* Write N x D frames: SVES (Param::Save, OL_Param.cpp:177-220, confidence
 * factors 1), HTK (the layout LoadHTK reads, parmKind USER) or RAW. */
int gmmhost_features_save(const char *path, int type, const float *x, uint32_t N, uint32_t D,
                          uint32_t sample_period);

const char *gmmhost_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GMMHOST_H */
