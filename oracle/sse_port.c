/*
 * sse_port.c -- restatement of the reference's multi-threaded SSE CPU
 * accumulator, used ONLY as the timed CPU baseline beside the GPU numbers
 * (bench.py's cpu_baseline leg) and checked against gmm_oracle.c (with the
 * same LPDIF rule) at 1, 4 and 16 threads in
 * tests/test_oracle.py::test_sse_port_matches_oracle_at_every_thread_count.
 * It is test infrastructure, never part of the product.
 *
 * Follows (paths relative to GMM/lib_bmssr/):
 *   trainer/OL_GMMStatsEstimator.cpp:234-313   model preprocessing (float)
 *   trainer/OL_GMMStatsEstimator_SSE.cpp:206-253 4-wide (x-mu)^2*ivar in float
 *   trainer/OL_GMMStatsEstimator.cpp:444-558   log-likelihood + posteriors
 *   trainer/OL_GMMStatsEstimator.cpp:1450-1458 addLog with LPDIF = 15
 *   trainer/OL_GMMStatsEstimator.cpp:1334-1378 accumulate g > 1e-4 into double
 *   trainer/OL_GMMStatsEstimator.cpp:884-971   contiguous thread split
 *                                              (>= 100 frames per thread),
 *                                              per-thread stats summed in
 *                                              thread order
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

#define MIN_FRAMES_PER_THREAD 100u

typedef struct {
    int D, Da, M;
    const float *mean_a; /* M x Da, zero padded */
    const float *ivar_a; /* M x Da, zero padded */
    const float *logw;   /* M */
    const float *lgc;    /* M, logGconst */
} sse_model;

typedef struct {
    const sse_model *mdl;
    const float *x;
    size_t n0, n1;
    double *n, *F, *S, tot;
} sse_job;

static const float MIN_LL = -1e20f;
static const float MIN_G = 1e-4f;
static const float LPDIF = 15.0f;

static float mix_ll(const sse_model *md, int m, const float *xa)
{
    const __m128 *mu = (const __m128 *)(md->mean_a + (size_t)m * md->Da);
    const __m128 *iv = (const __m128 *)(md->ivar_a + (size_t)m * md->Da);
    const __m128 *xv = (const __m128 *)xa;
    __m128 acc = _mm_setzero_ps();
    for (int b = 0; b < md->Da / 4; b++) {
        __m128 t = _mm_sub_ps(xv[b], mu[b]);
        acc = _mm_add_ps(acc, _mm_mul_ps(_mm_mul_ps(t, t), iv[b]));
    }
    __m128 h = _mm_add_ps(acc, _mm_movehl_ps(acc, acc));
    h = _mm_add_ss(h, _mm_shuffle_ps(h, h, 1));
    float q;
    _mm_store_ss(&q, h);
    float ll = md->lgc[m] - 0.5f * q;
    if (!isfinite(ll))
        return MIN_LL;
    return ll;
}

static void add_log(float *p1, float p2)
{
    if (*p1 - LPDIF > p2)
        return;
    if (*p1 > p2)
        *p1 = *p1 + logf(1.0f + expf(p2 - *p1));
    else
        *p1 = p2 + logf(1.0f + expf(*p1 - p2));
}

static void *sse_worker(void *arg)
{
    sse_job *jb = (sse_job *)arg;
    const sse_model *md = jb->mdl;
    const int D = md->D, M = md->M;
    float *xa = (float *)aligned_alloc(16, sizeof(float) * md->Da);
    float *g = (float *)malloc(sizeof(float) * (M > 0 ? M : 1));
    memset(xa, 0, sizeof(float) * md->Da);
    double tot = 0.0;
    for (size_t n = jb->n0; n < jb->n1; n++) {
        const float *xn = jb->x + n * (size_t)D;
        memcpy(xa, xn, sizeof(float) * D);
        float mx = MIN_LL, sum = MIN_LL;
        for (int m = 0; m < M; m++) {
            if (md->logw[m] == MIN_LL) {
                g[m] = MIN_LL;
                continue;
            }
            g[m] = mix_ll(md, m, xa) + md->logw[m];
            if (mx < g[m])
                mx = g[m];
            add_log(&sum, g[m]);
        }
        float ll = (mx == MIN_LL) ? MIN_LL : sum;
        if (M == 1) {
            g[0] = 1.0f;
        } else if (ll == MIN_LL) {
            for (int m = 0; m < M; m++)
                g[m] = MIN_G;
        } else {
            for (int m = 0; m < M; m++)
                g[m] = (float)(g[m] > MIN_LL) * expf(g[m] - ll);
        }
        tot += ll;
        for (int m = 0; m < M; m++) {
            if (g[m] <= MIN_G)
                continue;
            jb->n[m] += (double)g[m];
            double *Fm = jb->F + (size_t)m * D, *Sm = jb->S + (size_t)m * D;
            for (int d = 0; d < D; d++) {
                Fm[d] += (double)(g[m] * xn[d]);
                Sm[d] += (double)(g[m] * xn[d] * xn[d]);
            }
        }
    }
    jb->tot = tot;
    free(xa);
    free(g);
    return NULL;
}

/* Full accumulate call: model preprocessing + threaded accumulation.
 * Inputs are the M active components.  n/F/S are overwritten.  Returns the
 * number of threads used, or -1 when a variance is below 1e-6. */
int sse_accumulate_mt(int N, int D, int M, const float *x, const float *w,
                      const float *mu, const float *var, int nthreads,
                      double *n_out, double *F, double *S, double *tot_ll)
{
    const int Da = (D + 3) / 4 * 4;
    float *mean_a = (float *)aligned_alloc(16, sizeof(float) * (size_t)M * Da + 16);
    float *ivar_a = (float *)aligned_alloc(16, sizeof(float) * (size_t)M * Da + 16);
    float *logw = (float *)malloc(sizeof(float) * (M > 0 ? M : 1));
    float *lgc = (float *)malloc(sizeof(float) * (M > 0 ? M : 1));
    memset(mean_a, 0, sizeof(float) * (size_t)M * Da);
    memset(ivar_a, 0, sizeof(float) * (size_t)M * Da);
    for (int m = 0; m < M; m++) {
        logw[m] = (w[m] > 0) ? logf(w[m]) : MIN_LL;
        float acc = 0.0f;
        for (int d = 0; d < D; d++) {
            float v = var[(size_t)m * D + d];
            if (v < 1e-6f) {
                free(mean_a); free(ivar_a); free(logw); free(lgc);
                return -1;
            }
            ivar_a[(size_t)m * Da + d] = 1.0f / v;
            mean_a[(size_t)m * Da + d] = mu[(size_t)m * D + d];
            acc += logf(v);
        }
        lgc[m] = (float)(-1.837877066409345 * D / 2 - acc / 2);
    }
    sse_model md = {D, Da, M, mean_a, ivar_a, logw, lgc};

    unsigned int nthr = (unsigned int)N / MIN_FRAMES_PER_THREAD;
    if (nthr > (unsigned int)nthreads)
        nthr = (unsigned int)nthreads;
    if (nthr < 1)
        nthr = 1;
    sse_job *jobs = (sse_job *)calloc(nthr, sizeof(sse_job));
    pthread_t *th = (pthread_t *)calloc(nthr, sizeof(pthread_t));
    size_t part = (size_t)N / nthr, shift = 0;
    for (unsigned int t = 0; t < nthr; t++) {
        jobs[t].mdl = &md;
        jobs[t].x = x;
        jobs[t].n0 = shift;
        jobs[t].n1 = (t + 1 == nthr) ? (size_t)N : shift + part;
        shift += part;
        jobs[t].n = (double *)calloc(M, sizeof(double));
        jobs[t].F = (double *)calloc((size_t)M * D, sizeof(double));
        jobs[t].S = (double *)calloc((size_t)M * D, sizeof(double));
        if (nthr > 1)
            pthread_create(&th[t], NULL, sse_worker, &jobs[t]);
    }
    if (nthr == 1)
        sse_worker(&jobs[0]);
    memset(n_out, 0, sizeof(double) * M);
    memset(F, 0, sizeof(double) * (size_t)M * D);
    memset(S, 0, sizeof(double) * (size_t)M * D);
    double tot = 0.0;
    for (unsigned int t = 0; t < nthr; t++) {
        if (nthr > 1)
            pthread_join(th[t], NULL);
        for (int m = 0; m < M; m++)
            n_out[m] += jobs[t].n[m];
        for (size_t i = 0; i < (size_t)M * D; i++) {
            F[i] += jobs[t].F[i];
            S[i] += jobs[t].S[i];
        }
        tot += jobs[t].tot;
        free(jobs[t].n);
        free(jobs[t].F);
        free(jobs[t].S);
    }
    *tot_ll = tot;
    free(jobs);
    free(th);
    free(mean_a);
    free(ivar_a);
    free(logw);
    free(lgc);
    return (int)nthr;
}
