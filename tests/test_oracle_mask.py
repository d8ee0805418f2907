"""Pins the oracle's Gaussian selection (orc_accumulate_masked) before it
checks the HIP masked path: the full selection is the unmasked E-step,
random subsets against scipy/numpy, empty selections clamp, and a selection
of the components above the cut moves the statistics by no more than the
posterior mass it drops."""
import numpy as np
import pytest
from scipy.special import logsumexp

import _oracle as orc


def case(seed, N=300, D=4, M=8):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 2.0, (M, D)).astype(np.float32)
    x = rng.normal(0, 2.5, (N, D)).astype(np.float32)
    return rng, w, mu, var, x


def test_full_selection_is_the_unmasked_estep():
    _, w, mu, var, x = case(1)
    LL, n, F, S, tot = orc.accumulate_masked(x, w, mu, var, [range(8)] * len(x))
    assert np.array_equal(LL, orc.frame_loglik(x, w, mu, var))
    rn, rF, rS, rtot = orc.accumulate(x, w, mu, var)
    assert np.array_equal(n, rn) and np.array_equal(F, rF) and np.array_equal(S, rS)
    assert tot == rtot


def test_random_subsets_brute_force():
    rng, w, mu, var, x = case(2, N=60)
    M = len(w)
    lists = [np.sort(rng.choice(M, size=rng.integers(1, M + 1), replace=False)) for _ in x]
    LL, n, F, S, tot = orc.accumulate_masked(x, w, mu, var, lists)
    # per-component log-likelihoods from the unmasked oracle's posteriors
    LLu, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    l = np.log(np.maximum(g, 1e-300)) + LLu[:, None]
    xd = x.astype(np.float64)
    rn = np.zeros(M)
    rF = np.zeros(F.shape)
    for t, sel in enumerate(lists):
        ref = logsumexp(l[t, sel])
        assert LL[t] == pytest.approx(ref, rel=1e-12, abs=1e-9)
        gs = np.exp(l[t, sel] - ref)
        keep = gs > orc.MIN_GAMMA
        rn[sel[keep]] += gs[keep]
        rF[sel[keep]] += gs[keep][:, None] * xd[t]
    assert np.allclose(n, rn, rtol=1e-10, atol=1e-12)
    assert np.allclose(F, rF, rtol=1e-10, atol=1e-10)
    # a planted mistake: posteriors normalised over all components
    assert not np.allclose(n, np.where(g > orc.MIN_GAMMA, g, 0).sum(0), rtol=1e-3)


def test_empty_selection_clamps():
    _, w, mu, var, x = case(3, N=10)
    lists = [[] if t % 3 == 0 else [0, 2] for t in range(10)]
    LL, n, F, S, tot = orc.accumulate_masked(x, w, mu, var, lists)
    assert np.all(LL[::3] == orc.MIN_LOGLIKE)
    keep = [t for t in range(10) if t % 3]
    LL2, n2, F2, S2, _ = orc.accumulate_masked(x[keep], w, mu, var, [[0, 2]] * len(keep))
    assert np.allclose(n, n2, rtol=1e-12) and np.allclose(F, F2, rtol=1e-12)


def test_selection_above_the_cut():
    """estimateMaskIdxs (OL_GMMStatsEstimator.cpp:383-394): the components
    whose posterior exceeds the cut.  Restricting to them renormalises each
    frame's posteriors by the dropped mass (each dropped posterior <= cut)."""
    _, w, mu, var, x = case(4, N=500, M=16)
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    lists = [np.nonzero(row > orc.MIN_GAMMA)[0] for row in g]
    LL, n, F, S, tot = orc.accumulate_masked(x, w, mu, var, lists)
    rn, rF, rS, rtot = orc.accumulate(x, w, mu, var)
    dropped = np.array([row[row <= orc.MIN_GAMMA].sum() for row in g])
    assert np.all(dropped <= 16 * orc.MIN_GAMMA)
    assert np.abs(n - rn).max() <= 2 * dropped.sum() + 1e-9
    LLu = orc.frame_loglik(x, w, mu, var)
    assert np.all(LL <= LLu + 1e-12) and np.all(LLu - LL <= 16 * orc.MIN_GAMMA * 1.01)
