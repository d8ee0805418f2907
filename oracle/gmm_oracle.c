/*
 * gmm_oracle.c -- FP64 CPU restatement of the reference's diagonal-GMM EM
 * hot path.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP
 * library is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product never links it.
 *
 * Every function names the reference file:line it restates.  Paths are
 * relative to GMM/lib_bmssr/ of the reference tree.
 *
 * Arithmetic: all sums, logs and exps are in double.  Inputs that the
 * reference stores as float (frames, model parameters) are taken as float;
 * where the reference rounds an intermediate to float before using it again
 * (the M-step mean, MAP mean), this restatement rounds at the same place so
 * that updated models agree bit-for-bit in structure and to FP64 rounding in
 * value.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_prepare_model   -- closed form vs numpy/scipy normal log-density;
 *                          example/data.logL golden vector.
 *   orc_frame_loglik    -- example/data.logL (N=10000, D=20, M=4) to print
 *                          precision; scipy logsumexp brute force; M=1
 *                          Gaussian closed form; equal components -> 1/M;
 *                          translation invariance.
 *   orc_accumulate      -- brute force numpy on tiny inputs (posterior cut
 *                          planted both sides of 1e-4); sum of n == N when
 *                          no posterior is cut; permutation invariance.
 *   orc_update_em       -- example/myModel.gmm reproduced by the from-scratch
 *                          split driver (tests/_scratch_driver.py).
 *   orc_update_map      -- tau=0 equals ML means; n=0 keeps the UBM; alpha
 *                          closed form on a one-component case.
 *   orc_mllr_init/kg    -- G symmetric positive definite; brute-force outer
 *                          product sums in numpy; identity transform when the
 *                          statistics equal the UBM's own moments.
 *   orc_gj_invert       -- A * inv(A) == I against numpy.linalg.inv.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Reference thresholds: OL_GMMStatsEstimator.cpp:82-83 (_minLogLike=-1e20f,
 * _minGamma=1e-4f), OL_GMMStatsEstimator.h:283 (LPDIF=15),
 * OL_GMMStatsEstimator.h:274 (MINUS_LOG_2PI). */
#define ORC_MINUS_LOG_2PI (-1.837877066409345)
static const double ORC_LPDIF = 15.0;

double orc_min_loglike(void) { return (double)(-1e20f); }
double orc_min_gamma(void) { return (double)(1e-4f); }

/* RearrangeMixtures, OL_GMModel.cpp:857-885: components with weight > 0 keep
 * their order at the front; the others fill the tail from the back, so their
 * relative order is reversed.  perm[new] = old.  Returns the number of
 * zero-weight components. */
int orc_rearrange_perm(int M, const float *w, int *perm)
{
    int front = 0, back = M - 1;
    for (int m = 0; m < M; m++) {
        if (w[m] > 0.0f)
            perm[front++] = m;
        else
            perm[back--] = m;
    }
    return M - front;
}

/* insertModel (diagonal), OL_GMMStatsEstimator.cpp:234-313 together with the
 * GPU packing constant gconst = logGconst + logw,
 * OL_GMMStatsEstimator_CUDA.cpp:195-203.  Inputs are the M active
 * (weight > 0) components.  Outputs: gconst[m] and ivar[m*D+d].
 * Returns 0, or -(1+m) when component m holds a variance below min_var
 * (the reference throws there). */
int orc_prepare_model(int M, int D, const float *w, const float *mu,
                      const float *var, double min_var, double *gconst,
                      double *ivar)
{
    (void)mu;
    for (int m = 0; m < M; m++) {
        double logw = (w[m] > 0.0f) ? log((double)w[m]) : orc_min_loglike();
        double sum_logvar = 0.0;
        for (int d = 0; d < D; d++) {
            double v = var[(size_t)m * D + d];
            if (v < min_var)
                return -(1 + m);
            ivar[(size_t)m * D + d] = 1.0 / v;
            sum_logvar += log(v);
        }
        gconst[m] = ORC_MINUS_LOG_2PI * D / 2.0 - sum_logvar / 2.0 + logw;
    }
    return 0;
}

/* Component log-likelihood incl. log-weight: computeMixLogLike
 * (OL_GMMStatsEstimator_SSE.cpp:206-253, non-finite -> _minLogLike) plus
 * _logWeights (OL_GMMStatsEstimator.cpp:488). */
static double comp_ll(const float *x, int D, const float *mu_m,
                      const double *ivar_m, double gconst_m)
{
    double q = 0.0;
    for (int d = 0; d < D; d++) {
        double t = (double)x[d] - (double)mu_m[d];
        q += t * t * ivar_m[d];
    }
    double l = gconst_m - 0.5 * q;
    if (!isfinite(l))
        return orc_min_loglike();
    return l;
}

/* addLog with the LPDIF early-out, OL_GMMStatsEstimator.cpp:1450-1458. */
static double add_log_lpdif(double p1, double p2)
{
    if (p1 - ORC_LPDIF > p2)
        return p1;
    if (p1 > p2)
        return p1 + log(1.0 + exp(p2 - p1));
    return p2 + log(1.0 + exp(p1 - p2));
}

/* Per-frame log-likelihood and posteriors, getLikeAndGammas + compGammas
 * (OL_GMMStatsEstimator.cpp:444-558) with the GPU's LSE
 * (CU_GMMStatsEstimator_CUDAkernels.cuh:922-995, exact pairwise log-add, no
 * early-out) when lpdif == 0, or the CPU's sequential addLog with the
 * LPDIF=15 early-out when lpdif != 0.
 * LL[n] <= _minLogLike means every component underflowed: LL is clamped to
 * _minLogLike and the frame contributes no statistics.
 * gamma (optional, N*M) receives the posteriors before the 1e-4 cut. */
void orc_frame_loglik(int N, int D, int M, const float *x, const float *mu,
                      const double *ivar, const double *gconst, int lpdif,
                      double *LL, double *gamma)
{
    const double minLL = orc_min_loglike();
    double *l = (double *)malloc(sizeof(double) * (M > 0 ? M : 1));
    for (int n = 0; n < N; n++) {
        const float *xn = x + (size_t)n * D;
        double mx = minLL;
        for (int m = 0; m < M; m++) {
            l[m] = comp_ll(xn, D, mu + (size_t)m * D, ivar + (size_t)m * D,
                           gconst[m]);
            if (l[m] > mx)
                mx = l[m];
        }
        double ll;
        if (mx <= minLL) {
            ll = minLL;
        } else if (lpdif) {
            ll = minLL;
            for (int m = 0; m < M; m++)
                ll = add_log_lpdif(ll, l[m]);
        } else {
            double s = 0.0;
            for (int m = 0; m < M; m++)
                s += exp(l[m] - mx);
            ll = mx + log(s);
        }
        if (!isfinite(ll) || ll < minLL)
            ll = minLL;
        LL[n] = ll;
        if (gamma) {
            for (int m = 0; m < M; m++) {
                double g;
                if (ll <= minLL)
                    g = 0.0;
                else if (M == 1)
                    g = 1.0;
                else
                    g = (l[m] > minLL) ? exp(l[m] - ll) : 0.0;
                gamma[(size_t)n * M + m] = g;
            }
        }
    }
    free(l);
}

/* Statistics accumulation, accumulateStats over frames and per vector
 * (OL_GMMStatsEstimator.cpp:1105-1193, :1334-1378): for every component with
 * posterior > _minGamma, n += g, F += g*x, S += g*x*x.  tot_ll receives the
 * sum of per-frame log-likelihoods (including _minLogLike for frames whose
 * components all underflowed, as the reference does).  Outputs are
 * overwritten, not accumulated into. */
void orc_accumulate(int N, int D, int M, const float *x, const float *mu,
                    const double *ivar, const double *gconst, int lpdif,
                    double min_gamma, double *n_out, double *F, double *S,
                    double *tot_ll)
{
    double *g = (double *)malloc(sizeof(double) * (M > 0 ? M : 1));
    double LL;
    memset(n_out, 0, sizeof(double) * M);
    memset(F, 0, sizeof(double) * (size_t)M * D);
    memset(S, 0, sizeof(double) * (size_t)M * D);
    double tot = 0.0;
    for (int n = 0; n < N; n++) {
        const float *xn = x + (size_t)n * D;
        orc_frame_loglik(1, D, M, xn, mu, ivar, gconst, lpdif, &LL, g);
        tot += LL;
        for (int m = 0; m < M; m++) {
            if (!(g[m] > min_gamma))
                continue;
            n_out[m] += g[m];
            double *Fm = F + (size_t)m * D, *Sm = S + (size_t)m * D;
            for (int d = 0; d < D; d++) {
                double xv = xn[d];
                Fm[d] += g[m] * xv;
                Sm[d] += g[m] * xv * xv;
            }
        }
    }
    *tot_ll = tot;
    free(g);
}

/* Posteriors near the _minGamma cut (OL_GMMStatsEstimator.cpp:1350): for
 * every frame and component whose FP64 posterior lies in [lo, hi] (a band
 * around the cut), cnt[m] += 1 and the bounds of what one such frame can
 * move the statistics when an implementation lands on the other side of the
 * cut: gx[m*D+d] += |x_d|, gx2[m*D+d] += x_d^2.  Outputs are overwritten.
 * Used by the parity tests to give each component an allowance for cut
 * flips of exactly the frames that can flip; pinned by a numpy brute force
 * over the posteriors of orc_frame_loglik. */
void orc_near_cut(int N, int D, int M, const float *x, const float *mu,
                  const double *ivar, const double *gconst, double lo,
                  double hi, double *cnt, double *gx, double *gx2)
{
    double *g = (double *)malloc(sizeof(double) * (M > 0 ? M : 1));
    double LL;
    memset(cnt, 0, sizeof(double) * M);
    memset(gx, 0, sizeof(double) * (size_t)M * D);
    memset(gx2, 0, sizeof(double) * (size_t)M * D);
    for (int n = 0; n < N; n++) {
        const float *xn = x + (size_t)n * D;
        orc_frame_loglik(1, D, M, xn, mu, ivar, gconst, 0, &LL, g);
        for (int m = 0; m < M; m++) {
            if (!(g[m] >= lo && g[m] <= hi))
                continue;
            cnt[m] += 1.0;
            double *a = gx + (size_t)m * D, *b = gx2 + (size_t)m * D;
            for (int d = 0; d < D; d++) {
                double xv = xn[d];
                a[d] += fabs(xv);
                b[d] += xv * xv;
            }
        }
    }
    free(g);
}

/* Posterior-weighted sums of the posteriors that pass the cut, each taken
 * with the weight gamma (1 - gamma): for every frame and component with
 * gamma > min_gamma (FP64 posteriors as in orc_frame_loglik),
 * tn[m] += gamma (1 - gamma), tx[m*D+d] += gamma (1 - gamma) |x_d|,
 * tx2[m*D+d] += gamma (1 - gamma) x_d^2.  Outputs are overwritten.
 * If every component log-likelihood of a frame carries an absolute error of
 * at most E, a posterior gamma_m = exp(l_m - LL) moves by at most
 * 2 E gamma_m (1 - gamma_m); the parity tests turn these sums into the
 * statistics' allowance for that.  Pinned by a numpy brute force over the
 * posteriors. */
void orc_tail_mass(int N, int D, int M, const float *x, const float *mu,
                   const double *ivar, const double *gconst, double min_gamma,
                   double *tn, double *tx, double *tx2)
{
    double *g = (double *)malloc(sizeof(double) * (M > 0 ? M : 1));
    double LL;
    memset(tn, 0, sizeof(double) * M);
    memset(tx, 0, sizeof(double) * (size_t)M * D);
    memset(tx2, 0, sizeof(double) * (size_t)M * D);
    for (int n = 0; n < N; n++) {
        const float *xn = x + (size_t)n * D;
        orc_frame_loglik(1, D, M, xn, mu, ivar, gconst, 0, &LL, g);
        for (int m = 0; m < M; m++) {
            if (!(g[m] > min_gamma))
                continue;
            const double v = g[m] * (1.0 - g[m]);
            tn[m] += v;
            double *a = tx + (size_t)m * D, *b = tx2 + (size_t)m * D;
            for (int d = 0; d < D; d++) {
                double xv = xn[d];
                a[d] += v * fabs(xv);
                b[d] += v * xv * xv;
            }
        }
    }
    free(g);
}

/* GMMMixture::Reset for a diagonal mixture, OL_GMModel.cpp:76-86: mean = 0
 * and the variance bytes set to 0x01 (memset(Var, 1, ...)), i.e. every
 * variance becomes the float with bit pattern 0x01010101. */
static void mixture_reset(float *mu_m, float *var_m, int D)
{
    memset(mu_m, 0, sizeof(float) * D);
    memset(var_m, 1, sizeof(float) * D);
}

/* UpdateModelEM (diagonal, non-robust), OL_ModelAdapt.cpp:823-949.
 * w/mu/var are updated in place for the first M components (model order),
 * n/F/S are the statistics, tot_samples = GMMStats::_totalAccSamples.
 * Returns the number of components left with non-zero weight. */
static int update_em_core(int M, int D, float *w, float *mu, float *var,
                          const double *n, const double *F, const double *S,
                          double tot_samples, double min_var, int upweight,
                          int upmean, int upvar, const double *nind);

int orc_update_em(int M, int D, float *w, float *mu, float *var,
                  const double *n, const double *F, const double *S,
                  double tot_samples, double min_var, int upweight,
                  int upmean, int upvar)
{
    return update_em_core(M, D, w, mu, var, n, F, S, tot_samples, min_var, upweight, upmean,
                          upvar, NULL);
}

/* The EM update with the minimum-data test on nind[j] (the robust
 * independent-frame counts) when nind is given, else on n[j]. */
static int update_em_core(int M, int D, float *w, float *mu, float *var,
                          const double *n, const double *F, const double *S,
                          double tot_samples, double min_var, int upweight,
                          int upmean, int upvar, const double *nind)
{
    const double MIN_IND_SAMPLES = 2.0; /* OL_ModelAdapt.h:54 */
    int nzero = 0, wchange = 0;
    double gcorr = 0.0;
    float *mean = (float *)malloc(sizeof(float) * (D > 0 ? D : 1));
    for (int j = 0; j < M; j++) {
        if (w[j] == 0.0f) {
            nzero++;
            continue;
        }
        double np = n[j];
        if ((nind ? nind[j] : np) < MIN_IND_SAMPLES) {
            mixture_reset(mu + (size_t)j * D, var + (size_t)j * D, D);
            w[j] = 0.0f;
            wchange = 1;
            continue;
        }
        if (upweight) {
            w[j] = (float)(np / tot_samples);
            wchange = 1;
        }
        gcorr += w[j];
        if (upmean || upvar) {
            for (int k = 0; k < D; k++)
                mean[k] = (float)(F[(size_t)j * D + k] / np);
            if (upvar) {
                for (int k = 0; k < D; k++) {
                    float v = (float)(S[(size_t)j * D + k] / np -
                                      pow((double)mean[k], 2));
                    if (v < (float)min_var)
                        v = (float)min_var;
                    var[(size_t)j * D + k] = v;
                }
            }
            if (upmean)
                memcpy(mu + (size_t)j * D, mean, sizeof(float) * D);
        }
    }
    if (wchange && gcorr > 0.0)
        for (int j = 0; j < M; j++)
            w[j] = (float)(w[j] / gcorr);
    free(mean);
    return M - nzero;
}

/* UpdateModelMAP for tau > 0 (diagonal, non-robust),
 * OL_ModelAdapt.cpp:993-1106.  The model arrays start as a copy of the UBM
 * and are updated in place; ubm_* are the prior.  tau is an int in the
 * reference (OL_ModelAdapt.h:75). */
void orc_update_map(int M, int D, float *w, float *mu, float *var,
                    const float *ubm_w, const float *ubm_mu,
                    const float *ubm_var, const double *n, const double *F,
                    const double *S, double tot_samples, int tau,
                    double min_var, int upweight, int upmean, int upvar)
{
    double mixcorr = 0.0;
    float *amean = (float *)malloc(sizeof(float) * (D > 0 ? D : 1));
    for (int j = 0; j < M; j++) {
        if (ubm_w[j] == 0.0f)
            continue;
        double np = n[j];
        if (!(np > 0.0))
            continue;
        const float *um = ubm_mu + (size_t)j * D, *uv = ubm_var + (size_t)j * D;
        double alfa = np / (np + tau);
        if (upweight && tot_samples > 0) {
            float aw = (float)(alfa * (np / tot_samples) + (1 - alfa) * ubm_w[j]);
            w[j] = aw;
            mixcorr += aw;
        }
        if (upmean || upvar) {
            for (int k = 0; k < D; k++)
                amean[k] = (float)(alfa * F[(size_t)j * D + k] / np +
                                   (1 - alfa) * um[k]);
            if (upvar) {
                for (int k = 0; k < D; k++) {
                    float av = (float)(alfa * S[(size_t)j * D + k] / np +
                                       (1 - alfa) * (uv[k] + um[k] * um[k]) -
                                       amean[k] * amean[k]);
                    if (av < (float)min_var)
                        av = (float)min_var;
                    var[(size_t)j * D + k] = av;
                }
            }
            memcpy(mu + (size_t)j * D, amean, sizeof(float) * D);
        }
    }
    if (upweight && mixcorr != 0.0)
        for (int j = 0; j < M; j++)
            w[j] = w[j] / (float)mixcorr;
    free(amean);
}

/* prepareKiGiInitMatrices, OL_ModelAdapt.cpp:1484-1523: the UBM prior for the
 * MLLR row statistics, as if every UBM component contributed one frame at
 * its own mean with its own variance.  G is D blocks of (D+1)^2 (row-major),
 * k is D blocks of D+1. */
void orc_mllr_init(int M, int D, const float *ubm_mu, const float *ubm_var,
                   double min_var, double *G, double *k)
{
    const int E = D + 1;
    memset(G, 0, sizeof(double) * (size_t)D * E * E);
    memset(k, 0, sizeof(double) * (size_t)D * E);
    for (int m = 0; m < M; m++) {
        const float *var = ubm_var + (size_t)m * D, *mean = ubm_mu + (size_t)m * D;
        for (int i = 0; i < D; i++) {
            double cG = (var[i] < min_var) ? 1.0 / min_var : 1.0 / var[i];
            double ck = mean[i] * cG;
            double *Gi = G + (size_t)i * E * E, *ki = k + (size_t)i * E;
            for (int d = 0; d < D; d++) {
                ki[d] += ck * mean[d];
                Gi[(size_t)D * E + d] += cG * mean[d];
                Gi[(size_t)d * E + D] += cG * mean[d];
                for (int dd = 0; dd < D; dd++)
                    Gi[(size_t)d * E + dd] += cG * mean[d] * mean[dd];
                Gi[(size_t)d * E + d] += cG * var[d];
            }
            Gi[(size_t)E * E - 1] += cG;
            ki[D] += ck;
        }
    }
}

/* ComputeKiGi_MLLR, OL_ModelAdapt.cpp:1525-1581, for one row i:
 * G_i = Ginit_i + sum_{n_m>0} (n_m/var_mi) xi xi^T,
 * k_i = kinit_i + sum_{n_m>0} (F_mi/var_mi) xi, xi = [mu_m; 1].
 * Ginit/kinit may be NULL (zero prior).  Returns the total occupancy. */
double orc_mllr_kg(int M, int D, const float *mu, const float *var,
                   const double *n, const double *F, const double *Ginit,
                   const double *kinit, int i, double *G, double *k)
{
    const int E = D + 1;
    if (Ginit && kinit) {
        memcpy(G, Ginit, sizeof(double) * (size_t)E * E);
        memcpy(k, kinit, sizeof(double) * E);
    } else {
        memset(G, 0, sizeof(double) * (size_t)E * E);
        memset(k, 0, sizeof(double) * E);
    }
    double tot = 0.0;
    for (int m = 0; m < M; m++) {
        double np = n[m];
        if (!(np > 0.0))
            continue;
        tot += np;
        const float *mm = mu + (size_t)m * D;
        double vi = var[(size_t)m * D + i];
        double cm = F[(size_t)m * D + i] / vi;
        double cv = np / vi;
        for (int a = 0; a < D; a++) {
            k[a] += cm * mm[a];
            for (int b = a; b < D; b++) {
                G[(size_t)a * E + b] += cv * mm[a] * mm[b];
                G[(size_t)b * E + a] = G[(size_t)a * E + b];
            }
            G[(size_t)a * E + D] += cv * mm[a];
            G[(size_t)D * E + a] = G[(size_t)a * E + D];
        }
        k[D] += cm;
        G[(size_t)D * E + D] += cv;
    }
    return tot;
}

/* MATRIX::_Invert_Gen, Matrix4_LMa.cpp:2232-2290: in-place Gauss-Jordan
 * elimination without pivoting, FP64.  A is n x n row-major. */
void orc_gj_invert(int n, double *A)
{
    for (int k = 0; k < n; k++) {
        double p = A[(size_t)k * n + k];
        for (int i = 0; i < n; i++) {
            if (i == k)
                continue;
            double aik = A[(size_t)i * n + k];
            for (int j = 0; j < n; j++)
                if (j != k)
                    A[(size_t)i * n + j] -= A[(size_t)k * n + j] * aik / p;
        }
        for (int j = 0; j < n; j++)
            if (j != k)
                A[(size_t)k * n + j] = -A[(size_t)k * n + j] / p;
        for (int i = 0; i < n; i++)
            if (i != k)
                A[(size_t)i * n + k] = A[(size_t)i * n + k] / p;
        A[(size_t)k * n + k] = 1.0 / p;
    }
}

/* ComputeW_MLLR row solve, OL_ModelAdapt.cpp:1448-1461: w_i = inv(G_i) k_i
 * with the Gauss-Jordan inverse above.  G is overwritten. */
void orc_mllr_solve_row(int E, double *G, const double *k, double *w)
{
    orc_gj_invert(E, G);
    for (int a = 0; a < E; a++) {
        double s = 0.0;
        for (int b = 0; b < E; b++)
            s += G[(size_t)a * E + b] * k[b];
        w[a] = s;
    }
}

/* CheckInvertibility, OL_ModelAdapt.cpp:1851-1878.  Returns 0 when MLLR is
 * well-posed, 1 (too few components), 2 (too few active components) or
 * 3 (too little occupancy). */
int orc_mllr_check(int M, int D, const double *n)
{
    if (M < D + 1)
        return 1;
    int active = 0;
    double occ = 0.0;
    for (int m = 0; m < M; m++)
        if (n[m] > 0) {
            active++;
            occ += n[m];
        }
    if (active < D + 1)
        return 2;
    if (occ * D < (double)D * (D + 1))
        return 3;
    return 0;
}

/* AdaptModelMLLR, OL_ModelAdapt.cpp:1583-1612: mu' = A mu + b with
 * W = [A | b] (D x (D+1), row-major), FP64 accumulate, float store. */
void orc_mllr_apply(int M, int D, const double *W, const float *mu_in,
                    float *mu_out)
{
    const int E = D + 1;
    for (int m = 0; m < M; m++) {
        const float *src = mu_in + (size_t)m * D;
        for (int a = 0; a < D; a++) {
            double v = 0.0;
            for (int b = 0; b < D; b++)
                v += W[(size_t)a * E + b] * src[b];
            mu_out[(size_t)m * D + a] = (float)(v + W[(size_t)a * E + D]);
        }
    }
}

/* ------------------------------------------------------------------------
 * Full covariance.  Paths relative to GMM/lib_bmssr/.
 *
 * How each function is pinned (tests/test_oracle_fullcov.py):
 *   orc_prepare_model_full / orc_frame_loglik_full -- example/data.logL
 *       through example/myModel.gmm written as full matrices with zero
 *       off-diagonals; scipy multivariate_normal + logsumexp brute force;
 *       affine invariance (x -> Ax + b shifts every log-likelihood by
 *       -log|det A|).
 *   orc_accumulate_full -- numpy posterior-weighted outer products on tiny
 *       inputs; the diagonal of S equals orc_accumulate's S on a diagonal
 *       model.
 *   orc_update_em_full -- one component with all-ones posteriors gives the
 *       sample mean and covariance; example/myModelFC.gmm reproduced by the
 *       from-scratch driver (to the level stated in the test).
 *   orc_update_map_full -- diagonal models written as full matrices give
 *       orc_update_map's result; tau -> infinity keeps the UBM; a tiny tau
 *       gives orc_update_em_full's means and covariances; one component
 *       against the closed-form interpolation of the moments.
 *   orc_eig_leading -- numpy.linalg.eigh on random SPD matrices.
 * ------------------------------------------------------------------------ */

/* In-place lower Cholesky factor of a row-major SPD matrix in double: the
 * FP64 counterpart of spotrf('L') in getCovInverse / getLogDet
 * (trainer/OL_GMMStatsEstimator.cpp:156-229).  Returns 0, or 1 when a pivot
 * is not positive (the reference throws there). */
static int chol_lower(int D, double *A)
{
    for (int j = 0; j < D; j++) {
        double s = A[(size_t)j * D + j];
        for (int k = 0; k < j; k++)
            s -= A[(size_t)j * D + k] * A[(size_t)j * D + k];
        if (!(s > 0.0))
            return 1;
        const double ljj = sqrt(s);
        A[(size_t)j * D + j] = ljj;
        for (int i = j + 1; i < D; i++) {
            double t = A[(size_t)i * D + j];
            for (int k = 0; k < j; k++)
                t -= A[(size_t)i * D + k] * A[(size_t)j * D + k];
            A[(size_t)i * D + j] = t / ljj;
        }
        for (int k = j + 1; k < D; k++)
            A[(size_t)j * D + k] = 0.0;
    }
    return 0;
}

/* insertModel, full-covariance branch (OL_GMMStatsEstimator.cpp:291-299):
 * icov = cov^-1 (getCovInverse, spotrf + spotri, :156-189), logGconst =
 * MINUS_LOG_2PI*D/2 - logdet/2 with logdet = 2 sum log L_ii (getLogDet,
 * :194-229), plus log w.  prec[m*D*D] is the full symmetric inverse.
 * Returns 0, or -(1+m) when component m is not positive definite. */
int orc_prepare_model_full(int M, int D, const float *w, const float *cov,
                           double *prec, double *gconst)
{
    double *L = (double *)malloc(sizeof(double) * (size_t)D * D);
    double *Li = (double *)malloc(sizeof(double) * (size_t)D * D);
    for (int m = 0; m < M; m++) {
        const float *c = cov + (size_t)m * D * D;
        for (int i = 0; i < D * D; i++)
            L[i] = c[i];
        if (chol_lower(D, L)) {
            free(L);
            free(Li);
            return -(1 + m);
        }
        double logdet = 0.0;
        for (int i = 0; i < D; i++)
            logdet += log(L[(size_t)i * D + i]);
        logdet *= 2.0;
        /* Li = L^-1 (lower), by forward substitution on the identity */
        memset(Li, 0, sizeof(double) * (size_t)D * D);
        for (int j = 0; j < D; j++) {
            Li[(size_t)j * D + j] = 1.0 / L[(size_t)j * D + j];
            for (int i = j + 1; i < D; i++) {
                double s = 0.0;
                for (int k = j; k < i; k++)
                    s -= L[(size_t)i * D + k] * Li[(size_t)k * D + j];
                Li[(size_t)i * D + j] = s / L[(size_t)i * D + i];
            }
        }
        /* prec = Li^T Li */
        double *P = prec + (size_t)m * D * D;
        for (int i = 0; i < D; i++)
            for (int j = i; j < D; j++) {
                double s = 0.0;
                for (int k = (i > j ? i : j); k < D; k++)
                    s += Li[(size_t)k * D + i] * Li[(size_t)k * D + j];
                P[(size_t)i * D + j] = s;
                P[(size_t)j * D + i] = s;
            }
        double logw = (w[m] > 0.0f) ? log((double)w[m]) : orc_min_loglike();
        gconst[m] = ORC_MINUS_LOG_2PI * D / 2.0 - logdet / 2.0 + logw;
    }
    free(L);
    free(Li);
    return 0;
}

/* computeMixLogLikeFullCov (OL_GMMStatsEstimator.cpp:593-621, SSE form
 * OL_GMMStatsEstimator_SSE.cpp:258-317): gconst - 0.5 (x-mu)^T icov (x-mu),
 * non-finite -> _minLogLike; log-weight included in gconst. */
static double comp_ll_full(const float *x, int D, const float *mu_m,
                           const double *P, double gconst_m)
{
    double q = 0.0;
    for (int i = 0; i < D; i++) {
        const double ti = (double)x[i] - (double)mu_m[i];
        double r = 0.0;
        for (int j = 0; j < D; j++)
            r += P[(size_t)i * D + j] * ((double)x[j] - (double)mu_m[j]);
        q += ti * r;
    }
    double l = gconst_m - 0.5 * q;
    if (!isfinite(l))
        return orc_min_loglike();
    return l;
}

/* Per-frame log-likelihood and posteriors with full covariances: the same
 * combination rules as orc_frame_loglik (exact log-sum-exp, the GPU path's
 * semantics; clamp at _minLogLike; gamma before the 1e-4 cut). */
void orc_frame_loglik_full(int N, int D, int M, const float *x, const float *mu,
                           const double *prec, const double *gconst, double *LL,
                           double *gamma)
{
    const double minLL = orc_min_loglike();
    double *l = (double *)malloc(sizeof(double) * (M > 0 ? M : 1));
    for (int n = 0; n < N; n++) {
        const float *xn = x + (size_t)n * D;
        double mx = minLL;
        for (int m = 0; m < M; m++) {
            l[m] = comp_ll_full(xn, D, mu + (size_t)m * D, prec + (size_t)m * D * D,
                                gconst[m]);
            if (l[m] > mx)
                mx = l[m];
        }
        double ll;
        if (mx <= minLL) {
            ll = minLL;
        } else {
            double s = 0.0;
            for (int m = 0; m < M; m++)
                s += exp(l[m] - mx);
            ll = mx + log(s);
        }
        if (!isfinite(ll) || ll < minLL)
            ll = minLL;
        LL[n] = ll;
        if (gamma) {
            for (int m = 0; m < M; m++) {
                double g;
                if (ll <= minLL)
                    g = 0.0;
                else if (M == 1)
                    g = 1.0;
                else
                    g = (l[m] > minLL) ? exp(l[m] - ll) : 0.0;
                gamma[(size_t)n * M + m] = g;
            }
        }
    }
    free(l);
}

/* Packed upper-triangle index of (k, kk), kk >= k, in the reference's full
 * second-order statistics (OL_GMMStatsEstimator.cpp:1358-1364: shift =
 * k(k+1)/2, index k*D + kk - shift). */
static size_t packed_upper(int D, int k, int kk)
{
    return (size_t)k * D + kk - (size_t)k * (k + 1) / 2;
}

/* accumulateStats with _fFullVar (OL_GMMStatsEstimator.cpp:1105-1193,
 * :1334-1378): for posterior > min_gamma, n += g, F += g x, S[k<=kk] +=
 * g x_k x_kk (packed upper, D(D+1)/2 per component).  Outputs overwritten. */
void orc_accumulate_full(int N, int D, int M, const float *x, const float *mu,
                         const double *prec, const double *gconst, double min_gamma,
                         double *n_out, double *F, double *S, double *tot_ll)
{
    const size_t DP = (size_t)D * (D + 1) / 2;
    double *g = (double *)malloc(sizeof(double) * (M > 0 ? M : 1));
    double LL;
    memset(n_out, 0, sizeof(double) * M);
    memset(F, 0, sizeof(double) * (size_t)M * D);
    memset(S, 0, sizeof(double) * (size_t)M * DP);
    double tot = 0.0;
    for (int n = 0; n < N; n++) {
        const float *xn = x + (size_t)n * D;
        orc_frame_loglik_full(1, D, M, xn, mu, prec, gconst, &LL, g);
        tot += LL;
        for (int m = 0; m < M; m++) {
            if (!(g[m] > min_gamma))
                continue;
            n_out[m] += g[m];
            double *Fm = F + (size_t)m * D, *Sm = S + (size_t)m * DP;
            for (int k = 0; k < D; k++) {
                const double xk = xn[k];
                Fm[k] += g[m] * xk;
                for (int kk = k; kk < D; kk++)
                    Sm[packed_upper(D, k, kk)] += g[m] * xk * (double)xn[kk];
            }
        }
    }
    *tot_ll = tot;
    free(g);
}

/* UpdateModelEM, full-covariance branch, non-robust (OL_ModelAdapt.cpp:
 * 823-949): w = n/N_tot; mean = (float)(F/n); cov[k][kk] = cov[kk][k] =
 * (float)(S/n - mean_k * mean_kk) with the product of the float means taken
 * in float (:917-923); no variance floor.  Components with n < 2 are reset
 * (GMMMixture::Reset, OL_GMModel.cpp:73-83: zero mean, identity covariance)
 * and get weight 0; weights are renormalised over the survivors. */
static int update_em_full_core(int M, int D, float *w, float *mu, float *cov,
                               const double *n, const double *F, const double *S,
                               double tot_samples, const double *nind);

int orc_update_em_full(int M, int D, float *w, float *mu, float *cov,
                       const double *n, const double *F, const double *S,
                       double tot_samples)
{
    return update_em_full_core(M, D, w, mu, cov, n, F, S, tot_samples, NULL);
}

static int update_em_full_core(int M, int D, float *w, float *mu, float *cov,
                               const double *n, const double *F, const double *S,
                               double tot_samples, const double *nind)
{
    const double MIN_IND_SAMPLES = 2.0; /* OL_ModelAdapt.h:54 */
    const size_t DP = (size_t)D * (D + 1) / 2;
    int nzero = 0, wchange = 0;
    double gcorr = 0.0;
    float *mean = (float *)malloc(sizeof(float) * (D > 0 ? D : 1));
    for (int j = 0; j < M; j++) {
        float *c = cov + (size_t)j * D * D;
        if (w[j] == 0.0f) {
            nzero++;
            continue;
        }
        const double np = n[j];
        if ((nind ? nind[j] : np) < MIN_IND_SAMPLES) {
            memset(mu + (size_t)j * D, 0, sizeof(float) * D);
            memset(c, 0, sizeof(float) * (size_t)D * D);
            for (int d = 0; d < D; d++)
                c[(size_t)d * D + d] = 1.0f;
            w[j] = 0.0f;
            wchange = 1;
            continue;
        }
        w[j] = (float)(np / tot_samples);
        wchange = 1;
        gcorr += w[j];
        for (int k = 0; k < D; k++)
            mean[k] = (float)(F[(size_t)j * D + k] / np);
        const double *Sj = S + (size_t)j * DP;
        for (int k = 0; k < D; k++)
            for (int kk = k; kk < D; kk++) {
                const float mm = mean[k] * mean[kk];
                const float v = (float)(Sj[packed_upper(D, k, kk)] / np - (double)mm);
                c[(size_t)k * D + kk] = v;
                c[(size_t)kk * D + k] = v;
            }
        memcpy(mu + (size_t)j * D, mean, sizeof(float) * D);
    }
    if (wchange && gcorr > 0.0)
        for (int j = 0; j < M; j++)
            w[j] = (float)(w[j] / gcorr);
    free(mean);
    return M - nzero;
}

/* UpdateModelMAP for tau > 0, full-covariance branch, non-robust
 * (OL_ModelAdapt.cpp:993-1106, :1079-1085): as orc_update_map, with
 * cov[k][kk] = cov[kk][k] = (float)(alfa S_kk'/n + (1-alfa)(ubm_cov[k][kk] +
 * ubm_mu_k ubm_mu_kk) - mu_k mu_kk), the products of float means taken in
 * float, no variance floor.  S is packed upper per component. */
void orc_update_map_full(int M, int D, float *w, float *mu, float *cov,
                         const float *ubm_w, const float *ubm_mu, const float *ubm_cov,
                         const double *n, const double *F, const double *S,
                         double tot_samples, int tau, int upweight, int upmean, int upvar)
{
    const size_t DP = (size_t)D * (D + 1) / 2;
    double mixcorr = 0.0;
    float *amean = (float *)malloc(sizeof(float) * (D > 0 ? D : 1));
    for (int j = 0; j < M; j++) {
        if (ubm_w[j] == 0.0f)
            continue;
        const double np = n[j];
        if (!(np > 0.0))
            continue;
        const float *um = ubm_mu + (size_t)j * D, *uc = ubm_cov + (size_t)j * D * D;
        const double alfa = np / (np + tau);
        if (upweight && tot_samples > 0) {
            const float aw = (float)(alfa * (np / tot_samples) + (1 - alfa) * ubm_w[j]);
            w[j] = aw;
            mixcorr += aw;
        }
        if (upmean || upvar) {
            for (int k = 0; k < D; k++)
                amean[k] = (float)(alfa * F[(size_t)j * D + k] / np + (1 - alfa) * um[k]);
            if (upvar) {
                float *c = cov + (size_t)j * D * D;
                const double *Sj = S + (size_t)j * DP;
                for (int k = 0; k < D; k++)
                    for (int kk = k; kk < D; kk++) {
                        /* the UBM moment is summed in float, as the reference's
                         * float operands make it */
                        const float ubm_m2 = uc[(size_t)k * D + kk] + um[k] * um[kk];
                        const float mm = amean[k] * amean[kk];
                        const float v = (float)(alfa * Sj[packed_upper(D, k, kk)] / np +
                                                (1 - alfa) * (double)ubm_m2 - (double)mm);
                        c[(size_t)k * D + kk] = v;
                        c[(size_t)kk * D + k] = v;
                    }
            }
            memcpy(mu + (size_t)j * D, amean, sizeof(float) * D);
        }
    }
    if (upweight && mixcorr != 0.0)
        for (int j = 0; j < M; j++)
            w[j] = w[j] / (float)mixcorr;
    free(amean);
}

/* Leading eigenpair of a symmetric matrix (the ssyevr('V','I','U', il = iu
 * = D) call of getEigValVecs, OL_ModelAdapt.cpp:72-160, used by the
 * full-covariance splitGaussian, :614-660), by cyclic Jacobi rotations in
 * double.  The eigenvector has unit length; its sign, which LAPACK leaves
 * open, is fixed so that the entry of largest magnitude (lowest index on a
 * tie) is negative -- the choice that reproduces example/myModelFC.gmm. */
void orc_eig_leading(int D, const float *cov, double *lam, double *v)
{
    double *A = (double *)malloc(sizeof(double) * (size_t)D * D);
    double *V = (double *)malloc(sizeof(double) * (size_t)D * D);
    for (int i = 0; i < D * D; i++)
        A[i] = cov[i];
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++)
            V[(size_t)i * D + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < D; p++)
            for (int q = p + 1; q < D; q++)
                off += A[(size_t)p * D + q] * A[(size_t)p * D + q];
        if (off < 1e-30)
            break;
        for (int p = 0; p < D; p++)
            for (int q = p + 1; q < D; q++) {
                const double apq = A[(size_t)p * D + q];
                if (fabs(apq) < 1e-300)
                    continue;
                const double app = A[(size_t)p * D + p], aqq = A[(size_t)q * D + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < D; k++) {
                    const double akp = A[(size_t)k * D + p], akq = A[(size_t)k * D + q];
                    A[(size_t)k * D + p] = cs * akp - sn * akq;
                    A[(size_t)k * D + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < D; k++) {
                    const double apk = A[(size_t)p * D + k], aqk = A[(size_t)q * D + k];
                    A[(size_t)p * D + k] = cs * apk - sn * aqk;
                    A[(size_t)q * D + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < D; k++) {
                    const double vkp = V[(size_t)k * D + p], vkq = V[(size_t)k * D + q];
                    V[(size_t)k * D + p] = cs * vkp - sn * vkq;
                    V[(size_t)k * D + q] = sn * vkp + cs * vkq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < D; i++)
        if (A[(size_t)i * D + i] > A[(size_t)best * D + best])
            best = i;
    *lam = A[(size_t)best * D + best];
    int imax = 0;
    for (int i = 1; i < D; i++)
        if (fabs(V[(size_t)i * D + best]) > fabs(V[(size_t)imax * D + best]))
            imax = i;
    const double sgn = (V[(size_t)imax * D + best] > 0.0) ? -1.0 : 1.0;
    for (int i = 0; i < D; i++)
        v[i] = sgn * V[(size_t)i * D + best];
    free(A);
    free(V);
}

/* ------------------------------------------------------------------------
 * Robust ("aux") statistics and the robust covariance estimate.  Paths
 * relative to GMM/lib_bmssr/.
 *
 * How each function is pinned (tests/test_oracle_robust.py):
 *   orc_accumulate_aux -- numpy brute force over every consecutive frame
 *       pair on tiny inputs (with a planted pair-normalisation and a planted
 *       cut, both caught); M = 1 closed form (aux = sum |x_t - x_{t-1}|,
 *       aux2 = aux3 = N - 1); time reversal leaves aux and aux3 unchanged.
 *   orc_nind_frames -- i.i.d. Gaussian frames give about N independent
 *       frames (the 1.12 ~ 2/sqrt(pi) normalisation of the mean absolute
 *       difference), a slow random walk far fewer; brute force of the
 *       formula's pieces (global deviation, aux2 share).
 *   orc_make_robust_stats -- a large independent-frame count leaves the
 *       statistics almost unchanged (multipliers -> 1); diagonal models
 *       written as full give the same diagonal; multipliers against their
 *       closed forms at chosen counts.
 * ------------------------------------------------------------------------ */

/* accumulateAuxStats without Gaussian pruning
 * (trainer/OL_GMMStatsEstimator.cpp:1226-1280), driven as accumulateStats
 * drives it (:1174-1189) over ONE contiguous run of frames (one thread):
 * for every consecutive pair (t-1, t), t >= 1, with posteriors before the
 * cut g (gamma[N][M]):
 *   p_m = g_t,m g_t-1,m;  s = sum_m [p_m > cut] p_m;  h_m = p_m / s (0 if s = 0)
 *   aux2_m += sqrt(g_t,m)            if g_t,m > cut
 *   aux_m[k] += h_m |x_t[k] - x_t-1[k]|, aux3_m += h_m   if h_m > cut
 * aux[M][D+2] = [aux_m (D) | aux2_m | aux3_m] (the layout of the
 * reference's GPU aux buffer, trainer/CU_GMMStatsEstimator_CUDA.cu:862),
 * overwritten.  The frame difference is taken in float, as the reference's
 * TV = float. */
void orc_accumulate_aux(int N, int D, int M, const float *x, const double *gamma,
                        double min_gamma, double *aux)
{
    const int W = D + 2;
    memset(aux, 0, sizeof(double) * (size_t)M * W);
    for (int t = 1; t < N; t++) {
        const double *g2 = gamma + (size_t)t * M, *g1 = gamma + (size_t)(t - 1) * M;
        const float *x2 = x + (size_t)t * D, *x1 = x + (size_t)(t - 1) * D;
        double gs = 0.0;
        for (int m = 0; m < M; m++) {
            const double p = g2[m] * g1[m];
            if (p > min_gamma)
                gs += p;
        }
        if (gs > 0.0)
            gs = 1.0 / gs;
        for (int m = 0; m < M; m++) {
            double *a = aux + (size_t)m * W;
            const double h = gs * g2[m] * g1[m];
            if (g2[m] > min_gamma)
                a[D] += sqrt(g2[m]);
            if (h <= min_gamma)
                continue;
            for (int k = 0; k < D; k++) {
                const float ad = fabsf(x2[k] - x1[k]);
                a[k] += h * (double)ad;
            }
            a[D + 1] += h;
        }
    }
}

/* Index of S_dd in a component's second-order statistics: d for diagonal
 * statistics, the packed upper position for full ones. */
static size_t var_diag_index(int D, int full, int d)
{
    return full ? packed_upper(D, d, d) : (size_t)d;
}

/* GetNIndFrames (trainer/OL_ModelAdapt.cpp:468-535): the effective number of
 * independent frames per component,
 *   MII_m  = (1/D) sum_d aux_m[d] / aux3_m / std_d   (std: the global
 *            deviation from the summed statistics)
 *   nind_m = 1 + (aux2_m N_glob / sum_m aux2_m - 1) (MII_m / 1.12)^3,
 * 0 for a component with n = 0 or aux3 = 0.  Returns 0, or -1 when a global
 * deviation is 0 (the reference throws). */
int orc_nind_frames(int M, int D, int full, const double *n, const double *F, const double *S,
                    const double *aux, double *nind)
{
    const int W = D + 2;
    const size_t VS = full ? (size_t)D * (D + 1) / 2 : (size_t)D;
    double *gstd = (double *)calloc((size_t)D, sizeof(double));
    double *gmean = (double *)calloc((size_t)D, sizeof(double));
    double gprob = 0.0, sumNg = 0.0;
    for (int m = 0; m < M; m++) {
        gprob += n[m];
        for (int d = 0; d < D; d++) {
            gmean[d] += F[(size_t)m * D + d];
            gstd[d] += S[(size_t)m * VS + var_diag_index(D, full, d)];
        }
        nind[m] = 0.0;
        sumNg += aux[(size_t)m * W + D];
    }
    for (int d = 0; d < D; d++) {
        gmean[d] /= gprob;
        gstd[d] = sqrt(gstd[d] / gprob - gmean[d] * gmean[d]);
        if (gstd[d] == 0.0) {
            free(gstd);
            free(gmean);
            return -1;
        }
    }
    for (int m = 0; m < M; m++) {
        const double *a = aux + (size_t)m * W;
        if (n[m] == 0.0 || a[D + 1] == 0.0) {
            nind[m] = 0.0;
            continue;
        }
        double v = 0.0;
        for (int d = 0; d < D; d++)
            v += a[d] / a[D + 1] / gstd[d];
        v /= D;
        v = pow(v / 1.12, 3);
        nind[m] = 1 + (a[D] * gprob / sumNg - 1) * v;
    }
    free(gstd);
    free(gmean);
    return 0;
}

/* getDiagCovMultiplier / getOffDiagCovMultiplier (OL_ModelAdapt.cpp:537-548),
 * rounded to float as the reference returns them. */
float orc_diag_cov_multiplier(double nind)
{
    if (nind < 2)
        return (float)pow(pow(1.0 + 1.0 / (2.0 - 1.25), 3.5), 3 - nind);
    return (float)pow(1.0 + 1.0 / (nind - 1.25), 3.5);
}

float orc_offdiag_cov_multiplier(int D, double nind)
{
    if (nind < 2)
        return 0.0f;
    return (float)(1 - pow(((double)D / ((double)D - 1 + nind)), 1.4));
}

/* MakeRobustStats (OL_ModelAdapt.cpp:550-588): S is replaced, per component,
 * by  dc S + (1 - dc) F_k F_kk / n  with dc = diag multiplier x nind/(nind-1)
 * (float) on the diagonal and the off-diagonal multiplier odc on the
 * off-diagonal entries of full statistics; (1 - dc) is a float difference,
 * as in the reference.  nind receives the independent-frame counts.
 * Returns orc_nind_frames' status. */
int orc_make_robust_stats(int M, int D, int full, const double *n, const double *F, double *S,
                          const double *aux, double *nind)
{
    const int rc = orc_nind_frames(M, D, full, n, F, S, aux, nind);
    if (rc)
        return rc;
    const size_t VS = full ? (size_t)D * (D + 1) / 2 : (size_t)D;
    for (int j = 0; j < M; j++) {
        float dc = orc_diag_cov_multiplier(nind[j]);
        dc = (float)((double)dc * (nind[j] / (nind[j] - 1)));
        const double *Fj = F + (size_t)j * D;
        double *Sj = S + (size_t)j * VS;
        if (full) {
            const float odc = orc_offdiag_cov_multiplier(D, nind[j]);
            for (int k = 0; k < D; k++)
                for (int kk = k; kk < D; kk++) {
                    const size_t p = packed_upper(D, k, kk);
                    if (kk == k)
                        Sj[p] = dc * Sj[p] + (double)(1.0f - dc) * pow(Fj[k], 2.0) / n[j];
                    else
                        Sj[p] = odc * Sj[p] + (double)(1.0f - odc) * Fj[k] * Fj[kk] / n[j];
                }
        } else {
            for (int k = 0; k < D; k++)
                Sj[k] = dc * Sj[k] + (double)(1.0f - dc) * pow(Fj[k], 2.0) / n[j];
        }
    }
    return 0;
}

/* UpdateModelEM with _robustCov (OL_ModelAdapt.cpp:848-857, :880): S (a
 * copy) made robust by orc_make_robust_stats, then the EM update with the
 * minimum-data test on the independent-frame counts.  full selects the
 * full-covariance form (cov [M][D][D], S packed upper).  Returns the number
 * of components left, or -1 when a global variance is 0. */
int orc_update_em_robust(int M, int D, int full, float *w, float *mu, float *var,
                         const double *n, const double *F, const double *S,
                         const double *aux, double tot_samples, double min_var)
{
    const size_t VS = full ? (size_t)D * (D + 1) / 2 : (size_t)D;
    double *Sr = (double *)malloc(sizeof(double) * (size_t)M * VS);
    double *nind = (double *)malloc(sizeof(double) * (size_t)M);
    memcpy(Sr, S, sizeof(double) * (size_t)M * VS);
    int left = -1;
    if (orc_make_robust_stats(M, D, full, n, F, Sr, aux, nind) == 0)
        left = full ? update_em_full_core(M, D, w, mu, var, n, F, Sr, tot_samples, nind)
                    : update_em_core(M, D, w, mu, var, n, F, Sr, tot_samples, min_var, 1, 1, 1,
                                     nind);
    free(Sr);
    free(nind);
    return left;
}

/* ------------------------------------------------------------------------
 * fMLLR (constrained MLLR: one affine feature transform).  Paths relative
 * to GMM/lib_bmssr/.
 *
 * How it is pinned (tests/test_oracle_fmllr.py): G_i / k_i against numpy
 * sums over the statistics; frames drawn from the model give a transform
 * close to the identity; frames pushed through a planted affine map give
 * that map's inverse; each row update does not decrease the row's
 * auxiliary function; the determinant against numpy.
 * ------------------------------------------------------------------------ */

/* MATRIX::_Det_Gen (matrices/Matrix4_LMa.cpp): Gaussian elimination with
 * partial pivoting (row of the largest magnitude in the column), FP64. */
double orc_det(int n, const double *A_in)
{
    double *A = (double *)malloc(sizeof(double) * (size_t)n * n);
    memcpy(A, A_in, sizeof(double) * (size_t)n * n);
    double det = 1.0;
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(A[(size_t)i * n + k]) > fabs(A[(size_t)p * n + k]))
                p = i;
        if (A[(size_t)p * n + k] == 0.0) {
            free(A);
            return 0.0;
        }
        if (p != k) {
            for (int j = 0; j < n; j++) {
                const double t = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)p * n + j];
                A[(size_t)p * n + j] = t;
            }
            det = -det;
        }
        det *= A[(size_t)k * n + k];
        for (int i = k + 1; i < n; i++) {
            const double piv = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            for (int j = k + 1; j < n; j++)
                A[(size_t)i * n + j] -= piv * A[(size_t)k * n + j];
        }
    }
    free(A);
    return det;
}

/* ComputeKiGi_fMLLR (trainer/OL_ModelAdapt.cpp:1793-1840) for row i, from
 * full second-order statistics S (packed upper) of a model with diagonal
 * variances var (GetMixtureVarDiag): over components with n > 0,
 *   G_i += (1/var_mi) [[S_m, F_m], [F_m^T, n_m]],  k_i += (mu_mi/var_mi) [F_m, n_m],
 * with 1/var_mi and mu_mi/var_mi rounded to float as the reference computes
 * them from its float model.  G (E x E, E = D+1) and k (E) are overwritten. */
void orc_fmllr_kg(int M, int D, const float *mu, const float *var, const double *n,
                  const double *F, const double *S, int i, double *G, double *k)
{
    const int E = D + 1;
    const size_t DP = (size_t)D * (D + 1) / 2;
    memset(G, 0, sizeof(double) * (size_t)E * E);
    memset(k, 0, sizeof(double) * E);
    for (int m = 0; m < M; m++) {
        const double np = n[m];
        if (!(np > 0))
            continue;
        const float vm = var[(size_t)m * D + i], mm = mu[(size_t)m * D + i];
        const double cm = (double)(mm / vm), cv = (double)(1.0f / vm);
        const double *Fm = F + (size_t)m * D, *Sm = S + (size_t)m * DP;
        for (int a = 0; a < D; a++) {
            k[a] += cm * Fm[a];
            for (int b = a; b < D; b++) {
                G[(size_t)a * E + b] += cv * Sm[packed_upper(D, a, b)];
                G[(size_t)b * E + a] = G[(size_t)a * E + b];
            }
            G[(size_t)a * E + D] += cv * Fm[a];
            G[(size_t)D * E + a] = G[(size_t)a * E + D];
        }
        k[D] += cm * np;
        G[(size_t)D * E + D] += cv * np;
    }
}

/* ComputeW_fMLLR + EstimateAb_fmllr (trainer/OL_ModelAdapt.cpp:1614-1791):
 * A = I, b = 0; per iteration the cofactors of the current A (columns of
 * inv(A) det(A), fixed for the whole sweep), then for every row i the
 * roots f of  beta f^2 - (p Ginv k) f - p Ginv p^T = 0  (beta = the frame
 * count, rounded to float; p = [cofactor column i, 0]), the candidates
 * w = (p / f + k) Ginv, and the one with the larger
 * Q = sum w (k - G w / 2) becomes row i of [A | b].  Inverses by
 * orc_gj_invert (the reference's Gauss-Jordan), the determinant by
 * orc_det.  W (D x (D+1), row-major) is written; returns 0, or -1 on a
 * negative discriminant / zero root (the reference throws). */
int orc_fmllr(int M, int D, const float *mu, const float *var, const double *n, const double *F,
              const double *S, double tot_samples, int iters, double *W)
{
    const int E = D + 1;
    double *G = (double *)malloc(sizeof(double) * (size_t)D * E * E);
    double *Gi = (double *)malloc(sizeof(double) * (size_t)D * E * E);
    double *K = (double *)malloc(sizeof(double) * (size_t)D * E);
    double *A = (double *)calloc((size_t)D * D, sizeof(double));
    double *b = (double *)calloc((size_t)D, sizeof(double));
    double *cof = (double *)malloc(sizeof(double) * (size_t)D * D);
    double *p = (double *)malloc(sizeof(double) * E), *u = (double *)malloc(sizeof(double) * E);
    double *w[2], *gw = (double *)malloc(sizeof(double) * E);
    double *tv = (double *)malloc(sizeof(double) * E);
    w[0] = (double *)malloc(sizeof(double) * E);
    w[1] = (double *)malloc(sizeof(double) * E);
    int rc = 0;
    for (int i = 0; i < D; i++) {
        orc_fmllr_kg(M, D, mu, var, n, F, S, i, G + (size_t)i * E * E, K + (size_t)i * E);
        memcpy(Gi + (size_t)i * E * E, G + (size_t)i * E * E, sizeof(double) * (size_t)E * E);
        orc_gj_invert(E, Gi + (size_t)i * E * E);
    }
    for (int d = 0; d < D; d++)
        A[(size_t)d * D + d] = 1.0;
    const double beta = (double)(float)tot_samples;
    for (int it = 0; it < iters && rc == 0; it++) {
        const double det = orc_det(D, A);
        memcpy(cof, A, sizeof(double) * (size_t)D * D);
        orc_gj_invert(D, cof);
        for (int e = 0; e < D * D; e++)
            cof[e] *= det;
        for (int i = 0; i < D && rc == 0; i++) {
            const double *Gr = G + (size_t)i * E * E, *Gir = Gi + (size_t)i * E * E;
            const double *kr = K + (size_t)i * E;
            for (int a = 0; a < D; a++)
                p[a] = cof[(size_t)a * D + i];
            p[D] = 0.0;
            /* u = p^T Ginv */
            for (int c = 0; c < E; c++) {
                double s = 0.0;
                for (int a = 0; a < E; a++)
                    s += p[a] * Gir[(size_t)a * E + c];
                u[c] = s;
            }
            double bb = 0.0, cc = 0.0;
            for (int c = 0; c < E; c++) {
                bb += u[c] * kr[c];
                cc += u[c] * p[c];
            }
            bb = -bb;
            cc = -cc;
            const double sq = sqrt(pow(bb, 2) - 4 * beta * cc);
            if (sq < 0 || isnan(sq)) {
                rc = -1;
                break;
            }
            const double f[2] = {(-bb + sq) / (2 * beta), (-bb - sq) / (2 * beta)};
            double Q[2];
            for (int j = 0; j < 2; j++) {
                if (f[j] == 0) {
                    rc = -1;
                    break;
                }
                for (int c = 0; c < E; c++)
                    tv[c] = p[c] / f[j] + kr[c];
                for (int c = 0; c < E; c++) {
                    double s = 0.0;
                    for (int a = 0; a < E; a++)
                        s += tv[a] * Gir[(size_t)a * E + c];
                    w[j][c] = s;
                }
                for (int a = 0; a < E; a++) {
                    double s = 0.0;
                    for (int c = 0; c < E; c++)
                        s += Gr[(size_t)a * E + c] * w[j][c];
                    gw[a] = s;
                }
                Q[j] = 0.0;
                for (int c = 0; c < E; c++)
                    Q[j] += w[j][c] * (kr[c] - 0.5 * gw[c]);
            }
            if (rc)
                break;
            const int idx = (Q[0] > Q[1]) ? 0 : 1;
            for (int c = 0; c < D; c++)
                A[(size_t)i * D + c] = w[idx][c];
            b[i] = w[idx][D];
        }
    }
    for (int i = 0; i < D; i++) {
        for (int c = 0; c < D; c++)
            W[(size_t)i * E + c] = A[(size_t)i * D + c];
        W[(size_t)i * E + D] = b[i];
    }
    free(G); free(Gi); free(K); free(A); free(b); free(cof); free(p); free(u); free(gw); free(tv);
    free(w[0]); free(w[1]);
    return rc;
}

/* ------------------------------------------------------------------------
 * Gaussian selection (logLikeMask in "use" mode).
 *
 * Pinned (tests/test_oracle_mask.py): a mask holding every component gives
 * orc_frame_loglik / orc_accumulate exactly; scipy logsumexp and numpy sums
 * over random subsets; empty selections clamp; a selection of the
 * components above the cut (estimateMaskIdxs) changes the statistics by no
 * more than the dropped posterior mass.
 * ------------------------------------------------------------------------ */

/* getLikeAndGammas with mixIdxs (trainer/OL_GMMStatsEstimator.cpp:444-512)
 * + compGammas (:515-550) + accumulateStats with mixIdxs (:1334-1378), diagonal
 * model, exact log-sum-exp: frame t's log-likelihood over components
 * idx[off[t] .. off[t+1]) only, posteriors exp(l - LL) for those, the
 * statistics of those with posterior > min_gamma.  An empty selection (or
 * all components underflowing) clamps the frame to _minLogLike with no
 * statistics.  LL[N] and (optional) n/F/S (overwritten) and *tot. */
void orc_accumulate_masked(int N, int D, int M, const float *x, const float *mu,
                           const double *ivar, const double *gconst, const uint64_t *off,
                           const int *idx, double min_gamma, double *LL, double *n_out,
                           double *F, double *S, double *tot_ll)
{
    const double minLL = orc_min_loglike();
    double *l = (double *)malloc(sizeof(double) * (M > 0 ? M : 1));
    if (n_out) {
        memset(n_out, 0, sizeof(double) * M);
        memset(F, 0, sizeof(double) * (size_t)M * D);
        memset(S, 0, sizeof(double) * (size_t)M * D);
    }
    double tot = 0.0;
    for (int t = 0; t < N; t++) {
        const float *xt = x + (size_t)t * D;
        const uint64_t b = off[t], e = off[t + 1];
        double mx = minLL;
        for (uint64_t j = b; j < e; j++) {
            const int m = idx[j];
            l[m] = comp_ll(xt, D, mu + (size_t)m * D, ivar + (size_t)m * D, gconst[m]);
            if (l[m] > mx)
                mx = l[m];
        }
        double ll = minLL;
        if (mx > minLL) {
            double s = 0.0;
            for (uint64_t j = b; j < e; j++)
                s += exp(l[idx[j]] - mx);
            ll = mx + log(s);
            if (!isfinite(ll) || ll < minLL)
                ll = minLL;
        }
        LL[t] = ll;
        tot += ll;
        if (!n_out || ll <= minLL)
            continue;
        for (uint64_t j = b; j < e; j++) {
            const int m = idx[j];
            const double g = (M == 1) ? 1.0 : exp(l[m] - ll);
            if (!(g > min_gamma))
                continue;
            n_out[m] += g;
            for (int d = 0; d < D; d++) {
                const double xv = xt[d];
                F[(size_t)m * D + d] += g * xv;
                S[(size_t)m * D + d] += g * xv * xv;
            }
        }
    }
    if (tot_ll)
        *tot_ll = tot;
    free(l);
}
