"""Pins the oracle's robust statistics (oracle/gmm_oracle.c:
orc_accumulate_aux, orc_nind_frames, orc_make_robust_stats) before they
check the HIP path and the host's robust EM/MAP: brute force over frame
pairs, closed forms, and what the statistic is for (the number of
independent frames)."""
import numpy as np
import pytest

import _oracle as orc

CUT = orc.MIN_GAMMA


def brute_aux(x, g, plant_norm=False, plant_cut=False):
    N, D = x.shape
    M = g.shape[1]
    aux = np.zeros((M, D + 2))
    for t in range(1, N):
        p = g[t] * g[t - 1]
        s = p[p > CUT].sum()
        if plant_norm:
            s = p.sum()
        h = p / s if s > 0 else np.zeros(M)
        ad = np.abs(x[t] - x[t - 1]).astype(np.float64)
        for m in range(M):
            if g[t, m] > CUT:
                aux[m, D] += np.sqrt(g[t, m])
            if h[m] <= CUT and not plant_cut:
                continue
            aux[m, :D] += h[m] * ad
            aux[m, D + 1] += h[m]
    return aux


def diag_case(seed, N=60, D=3, M=4):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 1.5, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 2.0, (M, D)).astype(np.float32)
    x = rng.normal(0, 2.0, (N, D)).astype(np.float32)
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    return x, w, mu, var, g


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_aux_brute_force(seed):
    x, _, _, _, g = diag_case(seed)
    aux = orc.accumulate_aux(x, g)
    ref = brute_aux(x, g)
    assert np.allclose(aux, ref, rtol=1e-13, atol=1e-13)
    # planted mistakes: normalising by the uncut sum, or dropping the cut
    assert not np.allclose(brute_aux(x, g, plant_norm=True), aux, rtol=1e-9, atol=1e-12)
    assert not np.allclose(brute_aux(x, g, plant_cut=True), aux, rtol=1e-9, atol=1e-12)


def test_aux_single_component_closed_form():
    rng = np.random.default_rng(4)
    N, D = 50, 5
    x = rng.normal(0, 1, (N, D)).astype(np.float32)
    aux = orc.accumulate_aux(x, np.ones((N, 1)))
    d = np.abs(np.diff(x, axis=0)).astype(np.float64)
    assert np.allclose(aux[0, :D], d.sum(0), rtol=1e-13)
    assert aux[0, D] == N - 1 and aux[0, D + 1] == N - 1


def test_aux_time_reversal():
    """Pairs are symmetric in time: aux and aux3 do not change when the
    frames run backwards (aux2 counts frames 1..N-1, so it may)."""
    x, _, _, _, g = diag_case(5, N=80)
    a = orc.accumulate_aux(x, g)
    b = orc.accumulate_aux(x[::-1].copy(), g[::-1].copy())
    D = x.shape[1]
    assert np.allclose(a[:, :D], b[:, :D], rtol=1e-12, atol=1e-13)
    assert np.allclose(a[:, D + 1], b[:, D + 1], rtol=1e-12, atol=1e-13)


def _single_stats(x):
    xd = x.astype(np.float64)
    N, D = x.shape
    n = np.array([float(N)])
    F = xd.sum(0)[None]
    S = (xd * xd).sum(0)[None]
    aux = orc.accumulate_aux(x, np.ones((N, 1)))
    return n, F, S, aux


def test_nind_iid_frames_are_independent():
    """For i.i.d. Gaussian frames E|x_t - x_{t-1}| = 2 sigma / sqrt(pi) =
    1.128 sigma, which the 1.12 normalisation maps to ~1: every frame counts
    as independent.  A slow random walk (strongly correlated frames) counts
    as far fewer."""
    rng = np.random.default_rng(7)
    N, D = 20000, 4
    x = rng.normal(3.0, 2.0, (N, D)).astype(np.float32)
    nind = orc.nind_frames(*_single_stats(x))
    assert nind[0] == pytest.approx(N * (2 / np.sqrt(np.pi) / 1.12) ** 3, rel=0.03)
    assert 0.9 * N < nind[0] < 1.1 * N
    walk = np.cumsum(rng.normal(0, 0.02, (N, D)), axis=0).astype(np.float32)
    nwalk = orc.nind_frames(*_single_stats(walk))
    assert nwalk[0] < 0.01 * N


def test_nind_pieces_brute_force():
    x, w, mu, var, g = diag_case(8, N=400, D=3, M=3)
    gc = np.where(g > CUT, g, 0.0)
    xd = x.astype(np.float64)
    n, F, S = gc.sum(0), gc.T @ xd, gc.T @ (xd * xd)
    aux = orc.accumulate_aux(x, g)
    nind = orc.nind_frames(n, F, S, aux)
    gm = F.sum(0) / n.sum()
    gsd = np.sqrt(S.sum(0) / n.sum() - gm * gm)
    D = x.shape[1]
    mii = (aux[:, :D] / aux[:, D + 1:D + 2] / gsd).mean(1)
    ref = 1 + (aux[:, D] * n.sum() / aux[:, D].sum() - 1) * (mii / 1.12) ** 3
    assert np.allclose(nind, ref, rtol=1e-12)


def test_multipliers_closed_forms():
    assert orc.diag_cov_multiplier(11.25) == np.float32((1 + 1 / 10.0) ** 3.5)
    assert orc.diag_cov_multiplier(1.0) == np.float32(((1 + 1 / 0.75) ** 3.5) ** 2)
    assert orc.offdiag_cov_multiplier(5, 1.5) == 0.0
    assert orc.offdiag_cov_multiplier(5, 6.0) == np.float32(1 - (5 / 10.0) ** 1.4)
    # many independent frames: both multipliers -> 1 (the plain estimate)
    assert abs(orc.diag_cov_multiplier(1e7) - 1) < 1e-6
    assert abs(orc.offdiag_cov_multiplier(20, 1e7) - 1) < 1e-5


def test_make_robust_stats_limits_and_diag_as_full():
    rng = np.random.default_rng(9)
    N, D = 30000, 3
    x = rng.normal(0, 1.5, (N, D)).astype(np.float32)
    n, F, S, aux = _single_stats(x)
    S2, nind = orc.make_robust_stats(n, F, S, aux)
    # ~N independent frames: the covariance n S - F^2 is kept to ~1/N
    cov = S / n[:, None] - (F / n[:, None]) ** 2
    cov2 = S2 / n[:, None] - (F / n[:, None]) ** 2
    assert np.allclose(cov2, cov, rtol=5e-4)
    assert np.all(cov2 > cov)  # the robust estimate inflates the variance
    # diagonal statistics written as full: the same diagonal
    xd = x.astype(np.float64)
    r, c = np.triu_indices(D)
    Sf = (xd.T @ xd)[r, c][None]
    Sf[:, r == c] = S  # the same diagonal sums
    Sf2, nind_f = orc.make_robust_stats(n, F, Sf, aux, full=True)
    assert np.array_equal(nind_f, nind)
    assert np.array_equal(Sf2[:, r == c], S2)
    # few independent frames: off-diagonal moments shrink towards F_k F_l / n
    walk = np.cumsum(rng.normal(0, 0.05, (3000, D)), axis=0).astype(np.float32)
    n, F, _, aux = _single_stats(walk)
    wd = walk.astype(np.float64)
    Sw = (wd.T @ wd)[r, c][None]
    Sw2, nw = orc.make_robust_stats(n, F, Sw, aux, full=True)
    odc = orc.offdiag_cov_multiplier(D, nw[0])
    off = r != c
    ref = odc * Sw[0, off] + np.float64(np.float32(1) - np.float32(odc)) * (F[0, r] * F[0, c])[off] / n[0]
    assert np.allclose(Sw2[0, off], ref, rtol=1e-14)


def test_update_em_robust_limits():
    """Many independent frames: the robust EM update is the plain one up to
    the ~1/N inflation of the variances.  Strongly correlated frames: a
    component whose occupancy passes the minimum-data test but whose
    independent-frame count does not is reset."""
    x, w, mu, var, g = diag_case(31, N=20000, D=3, M=2)
    gc = np.where(g > CUT, g, 0.0)
    xd = x.astype(np.float64)
    n, F, S = gc.sum(0), gc.T @ xd, gc.T @ (xd * xd)
    aux = orc.accumulate_aux(x, g)
    pw, pmu, pvar, _ = orc.update_em(w, mu, var, n, F, S, x.shape[0])
    rw, rmu, rvar, left = orc.update_em_robust(w, mu, var, n, F, S, aux, x.shape[0])
    assert left == 2
    assert np.array_equal(rw, pw) and np.array_equal(rmu, pmu)
    assert np.all(rvar >= pvar) and np.allclose(rvar, pvar, rtol=2e-3)
    # correlated frames: one component sees a slow drift only
    rng = np.random.default_rng(32)
    N, D = 400, 2
    drift = (5.0 + np.cumsum(rng.normal(0, 0.01, (N, D)), axis=0)).astype(np.float32)
    rest = rng.normal(-5.0, 1.0, (N, D)).astype(np.float32)
    x = np.concatenate([drift, rest])
    w = np.array([0.5, 0.5], np.float32)
    mu = np.array([[5, 5], [-5, -5]], np.float32)
    var = np.ones((2, D), np.float32)
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    gc = np.where(g > CUT, g, 0.0)
    xd = x.astype(np.float64)
    n, F, S = gc.sum(0), gc.T @ xd, gc.T @ (xd * xd)
    aux = orc.accumulate_aux(x, g)
    nind = orc.nind_frames(n, F, S, aux)
    # (the mean absolute difference is measured in units of the global
    # deviation, which the two distant clusters make large: nind[1] stays
    # small too, but above the minimum)
    assert n[0] > 300 and nind[0] < 2 < nind[1]
    rw, rmu, rvar, left = orc.update_em_robust(w, mu, var, n, F, S, aux, len(x))
    assert left == 2 and rw[0] == 0 and rw[1] == 1
