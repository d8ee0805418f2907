"""From-scratch split training (trainGMM --type 0), restated on top of the
oracle to reproduce the reference's committed example/myModel.gmm.  Test
infrastructure: it exists to pin the oracle's E-step, statistics and EM
update against the reference's own output.

Follows OL_ModelAdapt.cpp:673-821 (TrainFromScratch), :590-660
(splitMultipleGaussians / splitGaussian, diagonal branch), :965-991
(GetGaussStats: highest weight, ties to the higher index; zero slots listed
from the highest index), OL_GMModel.cpp:857-885 (RearrangeMixtures inside
every insertModel) and :76-86 (GMMMixture::Reset).  With robust=True the
statistics carry the auxiliary frame-pair sums (accumulateAuxStats,
OL_GMMStatsEstimator.cpp:1226-1280) and every EM update is the robust one
(OL_ModelAdapt.cpp:848-857: MakeRobustStats, independent-frame test).
"""
import numpy as np

import _oracle as orc


def _rearrange(w, mu, var):
    perm, nz = orc.rearrange_perm(w)
    return w[perm].copy(), mu[perm].copy(), var[perm].copy(), nz


def _reset_var(D):
    return np.frombuffer(b"\x01" * (4 * D), dtype=np.float32).copy()


def _aux(x, w, mu, var, act, nummix, full):
    D = x.shape[1]
    if full:
        _, g = orc.frame_loglik_full(x, w[act], mu[act], var[act], want_gamma=True)
    else:
        _, g = orc.frame_loglik(x, w[act], mu[act], var[act], want_gamma=True)
    aux = np.zeros((nummix, D + 2))
    aux[act] = orc.accumulate_aux(x, g)
    return aux


def _split_diag(w, mu, var, m2, mn):
    """splitGaussian, diagonal branch (OL_ModelAdapt.cpp:614-660)."""
    var[mn] = var[m2]
    half = np.float32(w[m2] / np.float32(2.0))
    w[mn] = half
    w[m2] = half
    shift = (np.sqrt(var[m2]).astype(np.float64) * 0.5).astype(np.float32)
    mu[mn] = mu[m2] + shift
    mu[m2] = mu[m2] - shift


def _split_multiple(w, mu, var, split_x, split_th):
    """splitMultipleGaussians + GetGaussStats (OL_ModelAdapt.cpp:590-611,
    :951-990), written as the list the reference builds: scanning from the
    last component down, the heaviest is the first strict maximum met, empty
    slots are listed in scan order, and each non-empty component is inserted
    before the first listed one that is not heavier."""
    M = len(w)
    imax, wm = 0, np.float32(0.0)
    zero, order = [], []
    for m in range(M - 1, -1, -1):
        if wm < w[m]:
            wm, imax = w[m], m
        if w[m] == 0:
            zero.append(m)
            continue
        if split_x > 0:
            k = 0
            while k < len(order) and w[order[k]] > w[m]:
                k += 1
            order.insert(k, m)
    nonzero = M - len(zero)
    nsplit = int(np.ceil(np.float32(split_x) * np.float32(nonzero)))
    if nsplit < 2:
        _split_diag(w, mu, var, imax, zero[0])
        return
    th = np.float32(split_th) * w[imax]
    for n, m in enumerate(order):
        if w[m] < th or n >= len(zero) or n >= nsplit:
            break
        _split_diag(w, mu, var, m, zero[n])


def train_from_scratch(x, nummix, it_em, robust=False, split_x=0.0, split_th=0.0, ll_th=0.0,
                       history=None):
    """split_x / split_th / ll_th: trainGMM --split-X / --split-th / --ll-Th
    (TrainFromScratch, OL_ModelAdapt.cpp:738-797, with the mean
    log-likelihood of each round's last EM iteration computed at every
    verbosity).  history, if a list, receives (round, mean LL, non-zero
    components) per round."""
    x = np.ascontiguousarray(x, np.float32)
    N, D = x.shape
    # global statistics with a single all-ones posterior
    n = np.array([float(N)])
    F = x.astype(np.float64).sum(0)[None]
    S = (x.astype(np.float64) * x.astype(np.float64)).sum(0)[None]
    if robust:
        w0, mu0, var0, left = orc.update_em_robust(
            np.ones(1), np.zeros((1, D)), _reset_var(D)[None], n, F, S,
            orc.accumulate_aux(x, np.ones((N, 1))), N)
    else:
        w0, mu0, var0, left = orc.update_em(np.ones(1), np.zeros((1, D)),
                                            _reset_var(D)[None], n, F, S, N)
    w = np.zeros(nummix, np.float32)
    mu = np.zeros((nummix, D), np.float32)
    var = np.tile(_reset_var(D), (nummix, 1))
    w[0], mu[0], var[0] = w0[0], mu0[0], var0[0]
    if ll_th > 0 and it_em < 2:
        it_em = 2
    lmin = -np.finfo(np.float32).max
    ll = [lmin, lmin, lmin]
    m2 = m3 = None
    nonzero = 1
    for it in range(nummix if nummix >= 2 else 0):
        for _ in range(it_em):
            w, mu, var, _ = _rearrange(w, mu, var)
            act = w > 0
            na, Fa, Sa, tot = orc.accumulate(x, w[act], mu[act], var[act])
            ll[0] = np.float32(tot / N)
            n = np.zeros(nummix); F = np.zeros((nummix, D)); S = np.zeros((nummix, D))
            n[act], F[act], S[act] = na, Fa, Sa
            if nonzero > 1 and robust:
                aux = _aux(x, w, mu, var, act, nummix, False)
                w, mu, var, nonzero = orc.update_em_robust(w, mu, var, n, F, S, aux, N)
            elif nonzero > 1:
                w, mu, var, nonzero = orc.update_em(w, mu, var, n, F, S, N)
            if nonzero < 2:
                break
        if history is not None:
            history.append((it, float(ll[0]), int(nonzero)))
        if nonzero == 0 or nonzero == nummix:
            break
        if ll_th > 0:
            mx = max(ll[1], ll[2])
            if it > 2 and ll[0] - mx < np.float32(ll_th):
                if mx > ll[0]:
                    w, mu, var = m2 if mx == ll[1] else m3
                break
            m3 = m2
            m2 = (w.copy(), mu.copy(), var.copy())
            ll[2] = ll[1]
            ll[1] = ll[0]
        if it < nummix - 1:
            _split_multiple(w, mu, var, split_x, split_th)
            nonzero += 1
    w, mu, var, _ = _rearrange(w, mu, var)
    return w, mu, var


def _rearrange_full(w, mu, cov):
    perm, nz = orc.rearrange_perm(w)
    return w[perm].copy(), mu[perm].copy(), cov[perm].copy(), nz


def train_from_scratch_full(x, nummix, it_em, robust=False):
    """trainGMM --type 0 --full-cov: the same driver with the full-covariance
    statistics / update, GMMMixture::Reset's identity covariance for empty
    slots (OL_GMModel.cpp:73-83) and the full-covariance splitGaussian
    (OL_ModelAdapt.cpp:630-646): the new component copies the covariance and
    the two means move by -+ sqrt(lambda) v along the leading eigenpair."""
    x = np.ascontiguousarray(x, np.float32)
    N, D = x.shape
    xd = x.astype(np.float64)
    r, c = np.triu_indices(D)
    n = np.array([float(N)])
    F = xd.sum(0)[None]
    S = (xd.T @ xd)[r, c][None]
    eye = np.eye(D, dtype=np.float32)
    if robust:
        w0, mu0, cov0, left = orc.update_em_robust(
            np.ones(1), np.zeros((1, D)), eye[None], n, F, S,
            orc.accumulate_aux(x, np.ones((N, 1))), N, full=True)
    else:
        w0, mu0, cov0, left = orc.update_em_full(np.ones(1), np.zeros((1, D)), eye[None],
                                                 n, F, S, N)
    w = np.zeros(nummix, np.float32)
    mu = np.zeros((nummix, D), np.float32)
    cov = np.tile(eye, (nummix, 1, 1))
    w[0], mu[0], cov[0] = w0[0], mu0[0], cov0[0]
    nonzero = 1
    for it in range(nummix):
        for _ in range(it_em):
            w, mu, cov, _ = _rearrange_full(w, mu, cov)
            act = w > 0
            na, Fa, Sa, _tot = orc.accumulate_full(x, w[act], mu[act], cov[act])
            n = np.zeros(nummix)
            F = np.zeros((nummix, D))
            S = np.zeros((nummix, D * (D + 1) // 2))
            n[act], F[act], S[act] = na, Fa, Sa
            if nonzero > 1 and robust:
                aux = _aux(x, w, mu, cov, act, nummix, True)
                w, mu, cov, nonzero = orc.update_em_robust(w, mu, cov, n, F, S, aux, N,
                                                           full=True)
            elif nonzero > 1:
                w, mu, cov, nonzero = orc.update_em_full(w, mu, cov, n, F, S, N)
            if nonzero < 2:
                break
        if nonzero == 0 or nonzero == nummix:
            break
        if it < nummix - 1:
            imax, wm = 0, 0.0
            for m in range(nummix - 1, -1, -1):
                if wm < w[m]:
                    wm, imax = w[m], m
            znew = max(m for m in range(nummix) if w[m] == 0)
            cov[znew] = cov[imax]
            half = np.float32(w[imax] / np.float32(2.0))
            w[znew] = half
            w[imax] = half
            lam, v = orc.eig_leading(cov[imax])
            shift = (np.sqrt(lam) * v).astype(np.float32)
            mu[znew] = mu[imax] + shift
            mu[imax] = mu[imax] - shift
            nonzero += 1
    w, mu, cov, _ = _rearrange_full(w, mu, cov)
    return w, mu, cov
