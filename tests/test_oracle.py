"""Pins the CPU oracle (oracle/gmm_oracle.c) before it is trusted as the
checker: against the reference's committed example fixtures and against what
the mathematics fixes (closed forms, brute force, invariants)."""
import numpy as np
import pytest
from scipy.special import logsumexp
from scipy.stats import multivariate_normal

import _oracle as orc
import _scratch_driver as sd


def rand_model(rng, M, D, mean_sd=2.0):
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, mean_sd, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 3.0, (M, D)).astype(np.float32)
    return w, mu, var


def np_loglik_matrix(x, w, mu, var):
    """l[t,m] = log w_m + log N(x_t; mu_m, diag var_m), textbook FP64."""
    x = x.astype(np.float64)
    mu = mu.astype(np.float64)
    var = var.astype(np.float64)
    D = x.shape[1]
    q = (((x[:, None, :] - mu[None]) ** 2) / var[None]).sum(-1)
    return (np.log(w.astype(np.float64))[None] - 0.5 * D * np.log(2 * np.pi)
            - 0.5 * np.log(var).sum(-1)[None] - 0.5 * q)


# ---------------------------------------------------------------- fixtures
def test_data_logl_golden(golden):
    """example/data.logL (getLogLike.bat:9 on myModel.gmm) is reproduced to
    its 6-significant-digit print precision."""
    for lp in (0, 1):
        LL = orc.frame_loglik(golden["x"], golden["w"], golden["mu"], golden["var"], lpdif=lp)
        ref = golden["ll"]
        # 6 significant digits: half a unit in the 6th digit, plus the
        # reference's own FP32 evaluation error (~1e-6 relative)
        tol = 0.5 * 10.0 ** (np.floor(np.log10(np.abs(ref))) - 5) + 1e-6 * np.abs(ref)
        assert np.all(np.abs(LL - ref) <= tol), np.abs(LL - ref).max()


def test_data_logl_detects_planted_mistakes(golden):
    """A dropped log-weight or a wrong 2*pi term fails the golden check."""
    x, w, mu, var, ref = golden["x"], golden["w"], golden["mu"], golden["var"], golden["ll"]
    LL_now = orc.frame_loglik(x, np.ones_like(w), mu, var)  # log w dropped
    assert np.abs(LL_now - ref).max() > 0.5
    LL_sw = orc.frame_loglik(x, w[::-1].copy(), mu, var)  # weights misindexed
    assert np.abs(LL_sw - ref).max() > 1e-2


def test_mymodel_reproduced(golden):
    """example/myModel.gmm (trainModel.bat:4: --type 0 --nummix 4 --it-EM 8)
    is reproduced by the split driver on top of the oracle's E-step,
    statistics and EM update, to the file's 7-digit print precision."""
    w, mu, var = sd.train_from_scratch(golden["x"], 4, 8)
    rw = golden["w"]
    assert np.abs(w - rw).max() / rw.max() < 5e-6
    assert np.abs(mu - golden["mu"]).max() < 2e-6
    assert (np.abs(var - golden["var"]) / golden["var"]).max() < 2e-6


# ---------------------------------------------------------------- closed forms
def test_single_gaussian_closed_form():
    rng = np.random.default_rng(1)
    D = 7
    w, mu, var = rand_model(rng, 1, D)
    x = rng.normal(0, 2, (50, D)).astype(np.float32)
    LL, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    ref = multivariate_normal(mu[0].astype(np.float64), np.diag(var[0].astype(np.float64))).logpdf(
        x.astype(np.float64)) + np.log(np.float64(w[0]))
    assert np.allclose(LL, ref, rtol=0, atol=1e-10)
    assert np.all(g == 1.0)


def test_logsumexp_brute_force():
    rng = np.random.default_rng(2)
    for M, D in [(2, 3), (5, 11), (8, 1), (33, 6)]:
        w, mu, var = rand_model(rng, M, D)
        x = rng.normal(0, 3, (64, D)).astype(np.float32)
        LL, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
        l = np_loglik_matrix(x, w, mu, var)
        ref = logsumexp(l, axis=1)
        assert np.allclose(LL, ref, rtol=1e-13, atol=1e-10)
        assert np.allclose(g, np.exp(l - ref[:, None]), rtol=1e-12, atol=1e-300)
        assert np.allclose(g.sum(1), 1.0, atol=1e-12)


def test_equal_components_uniform_posterior():
    rng = np.random.default_rng(3)
    M, D = 6, 4
    mu = np.tile(rng.normal(0, 1, (1, D)), (M, 1)).astype(np.float32)
    var = np.tile(rng.uniform(0.5, 2, (1, D)), (M, 1)).astype(np.float32)
    w = np.full(M, 1.0 / M, np.float32)
    x = rng.normal(0, 1, (20, D)).astype(np.float32)
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    assert np.allclose(g, 1.0 / M, rtol=1e-6)


def test_translation_invariance():
    rng = np.random.default_rng(4)
    M, D = 5, 6
    w, mu, var = rand_model(rng, M, D)
    # values on a 2^-10 grid so that the shifted values are exact in float
    x = (np.round(rng.normal(0, 2, (40, D)) * 1024) / 1024).astype(np.float32)
    mu = (np.round(mu.astype(np.float64) * 1024) / 1024).astype(np.float32)
    sh = np.float32(4.0)
    a = orc.frame_loglik(x, w, mu, var)
    b = orc.frame_loglik(x + sh, w, mu + sh, var)
    assert np.allclose(a, b, rtol=0, atol=1e-9)


def test_lpdif_early_out_differs_only_by_skipped_terms():
    """The CPU addLog skips terms more than 15 below the running sum; the
    result is never above the exact LSE and below it by < M*exp(-15)."""
    rng = np.random.default_rng(5)
    M, D = 64, 8
    w, mu, var = rand_model(rng, M, D, mean_sd=1.0)
    x = rng.normal(0, 1, (200, D)).astype(np.float32)
    exact = orc.frame_loglik(x, w, mu, var, lpdif=0)
    cpu = orc.frame_loglik(x, w, mu, var, lpdif=1)
    d = exact - cpu
    assert np.all(d >= -1e-12)
    assert np.all(d <= M * np.exp(-15.0) * 1.01)


# ---------------------------------------------------------------- statistics
def test_accumulate_brute_force_with_cut():
    rng = np.random.default_rng(6)
    M, D, N = 7, 5, 300
    w, mu, var = rand_model(rng, M, D)
    x = rng.normal(0, 2.5, (N, D)).astype(np.float32)
    n, F, S, tot = orc.accumulate(x, w, mu, var)
    l = np_loglik_matrix(x, w, mu, var)
    LL = logsumexp(l, axis=1)
    g = np.exp(l - LL[:, None])
    assert (g <= orc.MIN_GAMMA).any() and (g > orc.MIN_GAMMA).any()
    g = np.where(g > orc.MIN_GAMMA, g, 0.0)
    xd = x.astype(np.float64)
    assert np.allclose(n, g.sum(0), rtol=1e-12)
    assert np.allclose(F, g.T @ xd, rtol=1e-11, atol=1e-11)
    assert np.allclose(S, g.T @ (xd * xd), rtol=1e-11, atol=1e-11)
    assert tot == pytest.approx(LL.sum(), rel=1e-13)


def test_posterior_cut_boundary_planted():
    """Two components placed so that one frame's posterior sits just above
    and another just below 1e-4: only the first is accumulated."""
    D = 1
    w = np.array([0.5, 0.5], np.float32)
    mu = np.array([[0.0], [1.0]], np.float32)
    var = np.array([[1.0], [1.0]], np.float32)
    # g_2 / g_1 = exp(x - 0.5); g_2 = 1e-4 at r = 1e-4/(1-1e-4)
    r = 1e-4 / (1 - 1e-4)
    xc = 0.5 + np.log(r)
    x = np.array([[xc + 1e-3], [xc - 1e-3]], np.float32)
    n, F, S, _ = orc.accumulate(x, w, mu, var)
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    assert g[0, 1] > 1e-4 > g[1, 1]
    assert n[1] == pytest.approx(g[0, 1], rel=1e-12)


def test_sum_of_occupancies_is_n_without_cut():
    rng = np.random.default_rng(7)
    M, D, N = 4, 3, 500
    w, mu, var = rand_model(rng, M, D, mean_sd=0.2)
    x = rng.normal(0, 0.5, (N, D)).astype(np.float32)
    n, _, _, _ = orc.accumulate(x, w, mu, var, min_gamma=0.0)
    assert n.sum() == pytest.approx(N, rel=1e-12)


def test_frame_permutation_invariance():
    rng = np.random.default_rng(8)
    M, D, N = 9, 4, 400
    w, mu, var = rand_model(rng, M, D)
    x = rng.normal(0, 2, (N, D)).astype(np.float32)
    a = orc.accumulate(x, w, mu, var)
    b = orc.accumulate(x[rng.permutation(N)], w, mu, var)
    for u, v in zip(a, b):
        assert np.allclose(u, v, rtol=1e-12, atol=1e-12)


def test_nan_frame_contributes_nothing():
    rng = np.random.default_rng(9)
    M, D = 3, 4
    w, mu, var = rand_model(rng, M, D)
    x = rng.normal(0, 1, (10, D)).astype(np.float32)
    x[3, 2] = np.nan
    n, F, S, tot = orc.accumulate(x, w, mu, var)
    n2, F2, S2, tot2 = orc.accumulate(np.delete(x, 3, 0), w, mu, var)
    assert np.allclose(n, n2) and np.allclose(F, F2) and np.allclose(S, S2)
    assert tot == pytest.approx(tot2 + orc.MIN_LOGLIKE, rel=1e-15)


def test_low_variance_rejected():
    w = np.array([1.0], np.float32)
    mu = np.zeros((1, 2), np.float32)
    var = np.array([[1.0, 1e-7]], np.float32)
    with pytest.raises(ValueError):
        orc.prepare_model(w, mu, var)


def test_rearrange_perm_reverses_zero_weight_tail():
    perm, nz = orc.rearrange_perm(np.array([0.2, 0, 0.3, 0, 0, 0.5], np.float32))
    assert nz == 3
    assert list(perm) == [0, 2, 5, 4, 3, 1]


# ---------------------------------------------------------------- M-step
def test_update_em_closed_form():
    rng = np.random.default_rng(10)
    M, D = 4, 3
    w = np.full(M, 0.25, np.float32)
    mu = np.zeros((M, D), np.float32)
    var = np.ones((M, D), np.float32)
    n = np.array([10.0, 20.0, 1.5, 30.0])
    F = rng.normal(0, 1, (M, D)) * n[:, None]
    S = (F / n[:, None]) ** 2 * n[:, None] + n[:, None] * 0.7
    w2, mu2, var2, left = orc.update_em(w, mu, var, n, F, S, tot_samples=61.5)
    # the reference counts only components that were already at zero weight
    # on entry (OL_ModelAdapt.cpp:948): the one dropped here still counts
    assert left == 4
    assert w2[2] == 0.0  # n < 2 -> dropped
    keep = [0, 1, 3]
    assert np.allclose(w2[keep], n[keep] / n[keep].sum(), rtol=1e-6)
    assert np.allclose(mu2[keep], (F / n[:, None])[keep], rtol=1e-6)
    assert np.allclose(var2[keep], 0.7, rtol=1e-5)
    assert np.all(mu2[2] == 0.0)


def test_map_tau0_limit_and_no_data():
    rng = np.random.default_rng(11)
    M, D = 3, 2
    w, mu, var = rand_model(rng, M, D)
    n = np.array([5.0, 0.0, 12.0])
    F = rng.normal(0, 1, (M, D)) * n[:, None]
    S = np.abs(F) + 1
    _, m_map, _ = orc.update_map(w, mu, var, n, F, S, 17.0, tau=16)
    alfa = n / (n + 16)
    for j in (0, 2):
        assert np.allclose(m_map[j], alfa[j] * F[j] / n[j] + (1 - alfa[j]) * mu[j], rtol=1e-6)
    assert np.all(m_map[1] == mu[1])  # no data -> UBM mean
    _, m_big, _ = orc.update_map(w, mu, var, n * 1e9, F * 1e9, S * 1e9, 17e9, tau=16)
    assert np.allclose(m_big[[0, 2]], (F / np.maximum(n, 1)[:, None])[[0, 2]], rtol=1e-6)


# ---------------------------------------------------------------- MLLR
def test_gauss_jordan_inverse():
    rng = np.random.default_rng(12)
    for E in (2, 5, 13, 61):
        A = rng.normal(0, 1, (E, E))
        A = A @ A.T + E * np.eye(E)
        inv = orc.gj_invert(A)
        assert np.allclose(inv @ A, np.eye(E), atol=1e-10)


def test_mllr_kg_brute_force_and_spd():
    rng = np.random.default_rng(13)
    M, D = 12, 4
    w, mu, var = rand_model(rng, M, D)
    n = rng.uniform(0, 10, M)
    n[3] = 0.0
    F = rng.normal(0, 1, (M, D)) * n[:, None]
    for i in range(D):
        G, k, tot = orc.mllr_kg(mu, var, n, F, i)
        xi = np.concatenate([mu.astype(np.float64), np.ones((M, 1))], 1)
        Gr = sum(n[m] / np.float64(var[m, i]) * np.outer(xi[m], xi[m]) for m in range(M) if n[m] > 0)
        kr = sum(F[m, i] / np.float64(var[m, i]) * xi[m] for m in range(M) if n[m] > 0)
        assert np.allclose(G, Gr, rtol=1e-12)
        assert np.allclose(k, kr, rtol=1e-12)
        assert np.allclose(G, G.T)
        assert np.all(np.linalg.eigvalsh(G) > 0)
        assert tot == pytest.approx(n.sum())


def test_mllr_init_is_virtual_frame_per_component():
    rng = np.random.default_rng(14)
    M, D = 6, 3
    _, mu, var = rand_model(rng, M, D)
    G, k = orc.mllr_init(mu, var)
    for i in range(D):
        xi = np.concatenate([mu.astype(np.float64), np.ones((M, 1))], 1)
        Gr = sum((np.outer(xi[m], xi[m]) + np.diag(np.append(var[m].astype(np.float64), 0)))
                 / np.float64(var[m, i]) for m in range(M))
        kr = sum(mu[m, i] / np.float64(var[m, i]) * xi[m] for m in range(M))
        assert np.allclose(G[i], Gr, rtol=1e-12)
        assert np.allclose(k[i], kr, rtol=1e-12)


def test_mllr_identity_when_stats_match_model():
    """F_m = n_m * mu_m for every component gives W = [I | 0]."""
    rng = np.random.default_rng(15)
    M, D = 20, 4
    _, mu, var = rand_model(rng, M, D)
    n = rng.uniform(1, 10, M)
    F = n[:, None] * mu.astype(np.float64)
    assert orc.mllr_check(M, D, n) == 0
    W = orc.mllr_transform(mu, var, n, F)
    assert np.allclose(W, np.hstack([np.eye(D), np.zeros((D, 1))]), atol=1e-9)
    out = orc.mllr_apply(W, mu)
    assert np.allclose(out, mu, atol=1e-5)


def test_mllr_recovers_affine_transform():
    rng = np.random.default_rng(16)
    M, D = 30, 3
    _, mu, var = rand_model(rng, M, D)
    A = np.eye(D) + rng.normal(0, 0.1, (D, D))
    b = rng.normal(0, 0.5, D)
    n = rng.uniform(1, 10, M)
    F = n[:, None] * (mu.astype(np.float64) @ A.T + b)
    W = orc.mllr_transform(mu, var, n, F)
    assert np.allclose(W[:, :D], A, atol=1e-9)
    assert np.allclose(W[:, D], b, atol=1e-9)


def test_mllr_check_codes():
    assert orc.mllr_check(3, 4, np.ones(3)) == 1
    n = np.zeros(10)
    n[:4] = 100
    assert orc.mllr_check(10, 4, n) == 2
    n = np.full(10, 0.1)
    assert orc.mllr_check(10, 4, n) == 3


# ---------------------------------------------------------------- near-cut counts
def test_near_cut_matches_brute_force():
    """orc_near_cut: per-component count of frames whose posterior lies in a
    band around the 1e-4 cut, and the sums of |x| and x^2 over them, against
    numpy posteriors (logsumexp) on a case with many posteriors near 1e-4."""
    rng = np.random.default_rng(71)
    w, mu, var = rand_model(rng, 9, 4, mean_sd=1.5)
    x = rng.normal(0, 2.5, (4000, 4)).astype(np.float32)
    l = np_loglik_matrix(x, w, mu, var)
    g = np.exp(l - logsumexp(l, axis=1)[:, None])
    for band in (0.05, 0.3):
        lo, hi = orc.MIN_GAMMA * (1 - band), orc.MIN_GAMMA * (1 + band)
        sel = (g >= lo) & (g <= hi)
        cnt, gx, gx2 = orc.near_cut(x, w, mu, var, band)
        assert sel.sum() > 20  # the case does exercise the band
        assert np.array_equal(cnt, sel.sum(0).astype(np.float64))
        xa = np.abs(x.astype(np.float64))
        assert np.allclose(gx, sel.T.astype(np.float64) @ xa, rtol=1e-12)
        assert np.allclose(gx2, sel.T.astype(np.float64) @ (xa * xa), rtol=1e-12)
    # a planted mistake (band taken on the log-likelihood instead of the
    # posterior) selects different frames
    sel_wrong = (np.abs(l - np.log(orc.MIN_GAMMA)) <= 0.05)
    assert not np.array_equal(sel_wrong.sum(0), orc.near_cut(x, w, mu, var, 0.05)[0])


def test_tail_mass_matches_brute_force():
    """orc_tail_mass: per-component sums of g(1-g), g(1-g)|x|, g(1-g)x^2
    over the posteriors above the cut, against numpy posteriors; and the
    bound it stands for: with every component log-likelihood of a frame
    moved by at most E, each posterior moves by at most 2 E g (1 - g)."""
    rng = np.random.default_rng(72)
    w, mu, var = rand_model(rng, 11, 5, mean_sd=1.0)
    x = rng.normal(0, 2.0, (3000, 5)).astype(np.float32)
    l = np_loglik_matrix(x, w, mu, var)
    g = np.exp(l - logsumexp(l, axis=1)[:, None])
    xa = np.abs(x.astype(np.float64))
    gs = np.where(g > orc.MIN_GAMMA, g * (1 - g), 0.0)
    tn, tx, tx2 = orc.tail_mass(x, w, mu, var)
    assert np.allclose(tn, gs.sum(0), rtol=1e-10)
    assert np.allclose(tx, gs.T @ xa, rtol=1e-10)
    assert np.allclose(tx2, gs.T @ (xa * xa), rtol=1e-10)
    E = 1e-3
    for _ in range(3):
        lp = l + rng.uniform(-E, E, l.shape)
        gp = np.exp(lp - logsumexp(lp, axis=1)[:, None])
        assert np.all(np.abs(gp - g) <= 2 * E * g * (1 - g) * (1 + 1e-3) + 1e-15)
    # a planted mistake (weight g instead of g(1-g)) gives other sums
    assert not np.allclose(tn, np.where(g > orc.MIN_GAMMA, g, 0.0).sum(0))


# ---------------------------------------------------------------- CPU baseline port
def test_sse_port_matches_oracle_at_every_thread_count():
    """oracle/sse_port.c (the timed CPU baseline: the reference's SSE path
    with float per-frame arithmetic, addLog with LPDIF = 15, g > 1e-4 into
    double, contiguous thread split) against the FP64 oracle with the same
    LPDIF rule, at 1, 4 and 16 threads; the thread split changes only the
    order of the final double sums."""
    import _check
    rng = np.random.default_rng(73)
    M, D, N = 40, 13, 6000
    w, mu, var = rand_model(rng, M, D)
    k = rng.choice(M, size=N, p=w.astype(np.float64) / w.sum())
    x = (mu[k] + np.sqrt(var[k]) * rng.standard_normal((N, D))).astype(np.float32)
    ref = orc.accumulate(x, w, mu, var, lpdif=True)
    near = orc.near_cut(x, w, mu, var, 1e-2)
    outs = []
    for t in (1, 4, 16):
        n, F, S, tot, used = orc.sse_accumulate(x, w, mu, var, t)
        assert used == t
        _check.check_stats((n, F, S, tot), ref, x, w, mu, var, rtol=1e-5, tot_rtol=1e-6,
                           near=near, band=1e-2, what="threads=%d " % t)
        outs.append((n, F, S, tot))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.allclose(a, b, rtol=1e-12, atol=1e-9)
    # the LPDIF early-out is what the port implements: a frame whose second
    # component is 14 log-units below the first gets its addLog term, one 16
    # below does not (the oracle agrees on both)
    w2 = np.array([0.5, 0.5], np.float32)
    mu2 = np.zeros((2, 1), np.float32)
    for gap, keep in ((14.0, True), (16.0, False)):
        var2 = np.array([[1.0], [1.0]], np.float32)
        mu2[1, 0] = np.sqrt(2 * gap)
        xx = np.zeros((1, 1), np.float32)
        t_port = orc.sse_accumulate(xx, w2, mu2, var2, 1)[3]
        t_orc = orc.accumulate(xx, w2, mu2, var2, lpdif=True)[3]
        t_exact = orc.accumulate(xx, w2, mu2, var2, lpdif=False)[3]
        assert t_port == pytest.approx(t_orc, abs=2e-6)
        assert (abs(t_orc - t_exact) < 1e-12) == keep
