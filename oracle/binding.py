"""ctypes binding of the CPU oracle (oracle/gmm_oracle.c, oracle/sse_port.c).

Test infrastructure only: this is the checker the HIP library is compared
against, never the thing measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "libgmm_oracle.so")

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    if a is None:
        return None
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        _lib = ctypes.CDLL(ORACLE_SO)
        _lib.orc_min_loglike.restype = ctypes.c_double
        _lib.orc_min_gamma.restype = ctypes.c_double
        _lib.orc_mllr_kg.restype = ctypes.c_double
        _lib.orc_det.restype = ctypes.c_double
        _lib.orc_diag_cov_multiplier.restype = ctypes.c_float
        _lib.orc_offdiag_cov_multiplier.restype = ctypes.c_float
    return _lib


MIN_LOGLIKE = float(np.float32(-1e20))
MIN_GAMMA = float(np.float32(1e-4))


def rearrange_perm(w):
    w = _f32(w)
    perm = np.zeros(len(w), dtype=np.int32)
    nz = lib().orc_rearrange_perm(len(w), _p(w), _p(perm))
    return perm, nz


def prepare_model(w, mu, var, min_var=1e-6):
    w, mu, var = _f32(w), _f32(mu), _f32(var)
    M, D = mu.shape
    gconst = np.zeros(M)
    ivar = np.zeros((M, D))
    # the reference compares float variances with a float _min_var
    # (OL_GMMStatsEstimator.h:62, OL_GMMStatsEstimator.cpp:279)
    rc = lib().orc_prepare_model(M, D, _p(w), _p(mu), _p(var),
                                 ctypes.c_double(float(np.float32(min_var))), _p(gconst),
                                 _p(ivar))
    if rc != 0:
        raise ValueError("variance below min_var in component %d" % (-rc - 1))
    return gconst, ivar


def frame_loglik(x, w, mu, var, lpdif=False, want_gamma=False):
    x, mu = _f32(x), _f32(mu)
    N, D = x.shape
    M = mu.shape[0]
    gconst, ivar = prepare_model(w, mu, var)
    LL = np.zeros(N)
    gamma = np.zeros((N, M)) if want_gamma else None
    lib().orc_frame_loglik(N, D, M, _p(x), _p(mu), _p(ivar), _p(gconst),
                           int(lpdif), _p(LL), _p(gamma))
    return (LL, gamma) if want_gamma else LL


def accumulate(x, w, mu, var, lpdif=False, min_gamma=MIN_GAMMA):
    x, mu = _f32(x), _f32(mu)
    N, D = x.shape
    M = mu.shape[0]
    gconst, ivar = prepare_model(w, mu, var)
    n = np.zeros(M)
    F = np.zeros((M, D))
    S = np.zeros((M, D))
    tot = ctypes.c_double(0.0)
    lib().orc_accumulate(N, D, M, _p(x), _p(mu), _p(ivar), _p(gconst),
                         int(lpdif), ctypes.c_double(min_gamma), _p(n), _p(F),
                         _p(S), ctypes.byref(tot))
    return n, F, S, tot.value


def near_cut(x, w, mu, var, band, min_gamma=MIN_GAMMA):
    """Per component: the number of frames whose FP64 posterior lies within a
    relative band of the cut, [min_gamma(1-band), min_gamma(1+band)], and the
    sums of |x| and x^2 over them (what such frames move the statistics by
    when they land on the other side of the cut, divided by the posterior)."""
    x, mu = _f32(x), _f32(mu)
    N, D = x.shape
    M = mu.shape[0]
    gconst, ivar = prepare_model(w, mu, var)
    cnt = np.zeros(M)
    gx = np.zeros((M, D))
    gx2 = np.zeros((M, D))
    lib().orc_near_cut(N, D, M, _p(x), _p(mu), _p(ivar), _p(gconst),
                       ctypes.c_double(min_gamma * (1.0 - band)),
                       ctypes.c_double(min_gamma * (1.0 + band)), _p(cnt), _p(gx), _p(gx2))
    return cnt, gx, gx2


def tail_mass(x, w, mu, var, min_gamma=MIN_GAMMA):
    """Per component: the sums of g(1-g), g(1-g)|x| and g(1-g) x^2 over the
    posteriors g that pass the cut (what an absolute error E in the
    component log-likelihoods can move the statistics by, divided by 2E)."""
    x, mu = _f32(x), _f32(mu)
    N, D = x.shape
    M = mu.shape[0]
    gconst, ivar = prepare_model(w, mu, var)
    tn = np.zeros(M)
    tx = np.zeros((M, D))
    tx2 = np.zeros((M, D))
    lib().orc_tail_mass(N, D, M, _p(x), _p(mu), _p(ivar), _p(gconst), ctypes.c_double(min_gamma),
                        _p(tn), _p(tx), _p(tx2))
    return tn, tx, tx2


def update_em(w, mu, var, n, F, S, tot_samples, min_var=1e-6,
              upweight=True, upmean=True, upvar=True):
    w, mu, var = _f32(w).copy(), _f32(mu).copy(), _f32(var).copy()
    M, D = mu.shape
    n, F, S = _f64(n), _f64(F), _f64(S)
    left = lib().orc_update_em(M, D, _p(w), _p(mu), _p(var), _p(n), _p(F), _p(S),
                               ctypes.c_double(tot_samples),
                               ctypes.c_double(min_var), int(upweight),
                               int(upmean), int(upvar))
    return w, mu, var, left


def update_map(ubm_w, ubm_mu, ubm_var, n, F, S, tot_samples, tau,
               min_var=1e-6, upweight=False, upmean=True, upvar=False):
    ubm_w, ubm_mu, ubm_var = _f32(ubm_w), _f32(ubm_mu), _f32(ubm_var)
    w, mu, var = ubm_w.copy(), ubm_mu.copy(), ubm_var.copy()
    M, D = mu.shape
    n, F, S = _f64(n), _f64(F), _f64(S)
    lib().orc_update_map(M, D, _p(w), _p(mu), _p(var), _p(ubm_w), _p(ubm_mu),
                         _p(ubm_var), _p(n), _p(F), _p(S),
                         ctypes.c_double(tot_samples), int(tau),
                         ctypes.c_double(min_var), int(upweight), int(upmean),
                         int(upvar))
    return w, mu, var


def mllr_init(ubm_mu, ubm_var, min_var=1e-6):
    ubm_mu, ubm_var = _f32(ubm_mu), _f32(ubm_var)
    M, D = ubm_mu.shape
    G = np.zeros((D, D + 1, D + 1))
    k = np.zeros((D, D + 1))
    lib().orc_mllr_init(M, D, _p(ubm_mu), _p(ubm_var), ctypes.c_double(min_var),
                        _p(G), _p(k))
    return G, k


def mllr_kg(mu, var, n, F, i, Ginit=None, kinit=None):
    mu, var = _f32(mu), _f32(var)
    M, D = mu.shape
    n, F = _f64(n), _f64(F)
    G = np.zeros((D + 1, D + 1))
    k = np.zeros(D + 1)
    gi = _f64(Ginit) if Ginit is not None else None
    ki = _f64(kinit) if kinit is not None else None
    tot = lib().orc_mllr_kg(M, D, _p(mu), _p(var), _p(n), _p(F), _p(gi), _p(ki),
                            int(i), _p(G), _p(k))
    return G, k, tot


def gj_invert(A):
    A = _f64(A).copy()
    lib().orc_gj_invert(A.shape[0], _p(A))
    return A


def mllr_solve_row(G, k):
    G = _f64(G).copy()
    k = _f64(k)
    w = np.zeros(len(k))
    lib().orc_mllr_solve_row(len(k), _p(G), _p(k), _p(w))
    return w


def mllr_check(M, D, n):
    n = _f64(n)
    return lib().orc_mllr_check(M, D, _p(n))


def mllr_transform(mu, var, n, F, ubm_mu=None, ubm_var=None):
    """Global MLLR W (D x (D+1)) from statistics, ComputeW_MLLR
    (OL_ModelAdapt.cpp:1417-1482), optional UBM prior."""
    mu = _f32(mu)
    M, D = mu.shape
    Gi = ki = None
    if ubm_mu is not None:
        Gi, ki = mllr_init(ubm_mu, ubm_var)
    W = np.zeros((D, D + 1))
    for i in range(D):
        G, k, _ = mllr_kg(mu, var, n, F, i,
                          None if Gi is None else Gi[i],
                          None if ki is None else ki[i])
        W[i] = mllr_solve_row(G, k)
    return W


def mllr_apply(W, mu):
    W = _f64(W)
    mu = _f32(mu)
    M, D = mu.shape
    out = np.zeros_like(mu)
    lib().orc_mllr_apply(M, D, _p(W), _p(mu), _p(out))
    return out


def sse_accumulate(x, w, mu, var, nthreads):
    x, w, mu, var = _f32(x), _f32(w), _f32(mu), _f32(var)
    N, D = x.shape
    M = mu.shape[0]
    n = np.zeros(M)
    F = np.zeros((M, D))
    S = np.zeros((M, D))
    tot = ctypes.c_double(0.0)
    used = lib().sse_accumulate_mt(N, D, M, _p(x), _p(w), _p(mu), _p(var),
                                   int(nthreads), _p(n), _p(F), _p(S),
                                   ctypes.byref(tot))
    if used < 0:
        raise ValueError("variance below 1e-6")
    return n, F, S, tot.value, used


# ---------------------------------------------------------------- full covariance

def prepare_model_full(w, cov):
    """(prec[M,D,D], gconst[M]) of a full-covariance model; raises on a
    matrix that is not positive definite."""
    w, cov = _f32(w), _f32(cov)
    M, D = cov.shape[0], cov.shape[1]
    prec = np.zeros((M, D, D))
    gconst = np.zeros(M)
    rc = lib().orc_prepare_model_full(M, D, _p(w), _p(cov), _p(prec), _p(gconst))
    if rc != 0:
        raise ValueError("covariance of component %d is not positive definite" % (-rc - 1))
    return prec, gconst


def frame_loglik_full(x, w, mu, cov, want_gamma=False):
    x, mu = _f32(x), _f32(mu)
    N, D = x.shape
    M = mu.shape[0]
    prec, gconst = prepare_model_full(w, cov)
    LL = np.zeros(N)
    gamma = np.zeros((N, M)) if want_gamma else None
    lib().orc_frame_loglik_full(N, D, M, _p(x), _p(mu), _p(prec), _p(gconst),
                                _p(LL), _p(gamma))
    return (LL, gamma) if want_gamma else LL


def accumulate_full(x, w, mu, cov, min_gamma=MIN_GAMMA):
    """(n[M], F[M,D], S[M,D(D+1)/2] packed upper, tot_ll)."""
    x, mu = _f32(x), _f32(mu)
    N, D = x.shape
    M = mu.shape[0]
    prec, gconst = prepare_model_full(w, cov)
    n = np.zeros(M)
    F = np.zeros((M, D))
    S = np.zeros((M, D * (D + 1) // 2))
    tot = ctypes.c_double(0.0)
    lib().orc_accumulate_full(N, D, M, _p(x), _p(mu), _p(prec), _p(gconst),
                              ctypes.c_double(min_gamma), _p(n), _p(F), _p(S),
                              ctypes.byref(tot))
    return n, F, S, tot.value


def update_em_full(w, mu, cov, n, F, S, tot_samples):
    w, mu, cov = _f32(w).copy(), _f32(mu).copy(), _f32(cov).copy()
    M, D = mu.shape
    n, F, S = _f64(n), _f64(F), _f64(S)
    left = lib().orc_update_em_full(M, D, _p(w), _p(mu), _p(cov), _p(n), _p(F), _p(S),
                                    ctypes.c_double(tot_samples))
    return w, mu, cov, left


def update_map_full(ubm_w, ubm_mu, ubm_cov, n, F, S, tot_samples, tau,
                    upweight=False, upmean=True, upvar=False):
    ubm_w, ubm_mu, ubm_cov = _f32(ubm_w), _f32(ubm_mu), _f32(ubm_cov)
    w, mu, cov = ubm_w.copy(), ubm_mu.copy(), ubm_cov.copy()
    M, D = mu.shape
    n, F, S = _f64(n), _f64(F), _f64(S)
    lib().orc_update_map_full(M, D, _p(w), _p(mu), _p(cov), _p(ubm_w), _p(ubm_mu), _p(ubm_cov),
                              _p(n), _p(F), _p(S), ctypes.c_double(tot_samples), int(tau),
                              int(upweight), int(upmean), int(upvar))
    return w, mu, cov


def eig_leading(cov):
    cov = _f32(cov)
    D = cov.shape[0]
    lam = ctypes.c_double(0.0)
    v = np.zeros(D)
    lib().orc_eig_leading(D, _p(cov), ctypes.byref(lam), _p(v))
    return lam.value, v


def packed_upper_index(D):
    """(rows, cols) of the packed upper-triangle order of the full
    second-order statistics."""
    r, c = np.triu_indices(D)
    return r, c


def accumulate_aux(x, gamma, min_gamma=MIN_GAMMA):
    """Robust auxiliary statistics [M, D+2] = [aux (D) | aux2 | aux3] over
    every consecutive frame pair, from the posteriors before the cut."""
    x = _f32(x)
    gamma = _f64(gamma)
    N, D = x.shape
    M = gamma.shape[1]
    aux = np.zeros((M, D + 2))
    lib().orc_accumulate_aux(N, D, M, _p(x), _p(gamma), ctypes.c_double(min_gamma), _p(aux))
    return aux


def nind_frames(n, F, S, aux, full=False):
    n, F, S, aux = _f64(n), _f64(F), _f64(S), _f64(aux)
    M, D = F.shape
    nind = np.zeros(M)
    rc = lib().orc_nind_frames(M, D, int(full), _p(n), _p(F), _p(S), _p(aux), _p(nind))
    if rc:
        raise ValueError("global variance equal zero")
    return nind


def make_robust_stats(n, F, S, aux, full=False):
    """(S', nind): MakeRobustStats applied to a copy of S."""
    n, F, aux = _f64(n), _f64(F), _f64(aux)
    S = _f64(S).copy()
    M, D = F.shape
    nind = np.zeros(M)
    rc = lib().orc_make_robust_stats(M, D, int(full), _p(n), _p(F), _p(S), _p(aux), _p(nind))
    if rc:
        raise ValueError("global variance equal zero")
    return S, nind


def diag_cov_multiplier(nind):
    return lib().orc_diag_cov_multiplier(ctypes.c_double(nind))


def offdiag_cov_multiplier(D, nind):
    return lib().orc_offdiag_cov_multiplier(int(D), ctypes.c_double(nind))


def update_em_robust(w, mu, var, n, F, S, aux, tot_samples, min_var=1e-6, full=False):
    """Robust EM update (weights, means, variances / covariances)."""
    w, mu, var = _f32(w).copy(), _f32(mu).copy(), _f32(var).copy()
    M, D = mu.shape
    n, F, S, aux = _f64(n), _f64(F), _f64(S), _f64(aux)
    left = lib().orc_update_em_robust(M, D, int(full), _p(w), _p(mu), _p(var), _p(n), _p(F),
                                      _p(S), _p(aux), ctypes.c_double(tot_samples),
                                      ctypes.c_double(min_var))
    if left < 0:
        raise ValueError("global variance equal zero")
    return w, mu, var, left


def det(A):
    A = _f64(A)
    return lib().orc_det(A.shape[0], _p(A))


def fmllr_kg(mu, var, n, F, S, i):
    """(G_i [(D+1)^2], k_i [D+1]) of fMLLR row i from full (packed upper)
    statistics and the diagonal variances."""
    mu, var = _f32(mu), _f32(var)
    M, D = mu.shape
    n, F, S = _f64(n), _f64(F), _f64(S)
    G = np.zeros((D + 1, D + 1))
    k = np.zeros(D + 1)
    lib().orc_fmllr_kg(M, D, _p(mu), _p(var), _p(n), _p(F), _p(S), int(i), _p(G), _p(k))
    return G, k


def fmllr(mu, var, n, F, S, tot_samples, iters=20):
    """fMLLR transform W = [A | b] (D x (D+1))."""
    mu, var = _f32(mu), _f32(var)
    M, D = mu.shape
    n, F, S = _f64(n), _f64(F), _f64(S)
    W = np.zeros((D, D + 1))
    rc = lib().orc_fmllr(M, D, _p(mu), _p(var), _p(n), _p(F), _p(S),
                         ctypes.c_double(tot_samples), int(iters), _p(W))
    if rc:
        raise ValueError("fMLLR: negative discriminant or zero root")
    return W


def accumulate_masked(x, w, mu, var, lists, min_gamma=MIN_GAMMA):
    """Gaussian selection: (LL[N], n, F, S, tot) with frame t restricted to
    the components lists[t]."""
    x, mu = _f32(x), _f32(mu)
    N, D = x.shape
    M = mu.shape[0]
    gconst, ivar = prepare_model(w, mu, var)
    off = np.zeros(N + 1, np.uint64)
    off[1:] = np.cumsum([len(l) for l in lists])
    idx = np.ascontiguousarray(np.concatenate([np.asarray(l, np.int32) for l in lists])
                               if off[-1] else np.zeros(1, np.int32), np.int32)
    LL = np.zeros(N)
    n = np.zeros(M)
    F = np.zeros((M, D))
    S = np.zeros((M, D))
    tot = ctypes.c_double(0.0)
    lib().orc_accumulate_masked(N, D, M, _p(x), _p(mu), _p(ivar), _p(gconst),
                                off.ctypes.data_as(ctypes.c_void_p),
                                idx.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(min_gamma),
                                _p(LL), _p(n), _p(F), _p(S), ctypes.byref(tot))
    return LL, n, F, S, tot.value
