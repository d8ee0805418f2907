"""Parity of the HIP robust auxiliary statistics (gmmhip_accumulate_aux
through the C ABI) with the FP64 oracle (orc_accumulate_aux over the
oracle's posteriors), for diagonal and full-covariance models.

Tolerances (FP32 GEMM-form log-likelihoods and posteriors, FP64 sums):
  aux_m[d]   |d| <= 2e-4 aux_m[d] + 5e-3 max|x_t - x_t-1| (cut allowance)
  aux2_m     |d| <= 2e-4 aux2_m + 0.05   (a posterior at the 1e-4 cut adds 0.01)
  aux3_m     |d| <= 2e-4 aux3_m + 5e-3
"""
import numpy as np
import pytest

import _oracle as orc
import gmm_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def acc():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    a = gmm_amd.Accumulator(0)
    yield a
    a.close()


def markov_frames(seed, N, D, M, stay=0.9, spread=2.5):
    """Frames from a mixture whose component persists over time (stay
    probability) with AR(1) noise: consecutive frames are correlated, as in
    speech features."""
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, spread, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 1.5, (M, D)).astype(np.float32)
    k = np.empty(N, np.int64)
    k[0] = rng.integers(M)
    jumps = rng.random(N) > stay
    pw = w.astype(np.float64)
    draws = rng.choice(M, size=N, p=pw / pw.sum())
    for t in range(1, N):
        k[t] = draws[t] if jumps[t] else k[t - 1]
    e = np.zeros((N, D))
    z = rng.normal(size=(N, D))
    for t in range(1, N):
        e[t] = 0.7 * e[t - 1] + np.sqrt(1 - 0.49) * z[t]
    x = (mu[k] + np.sqrt(var[k]) * e).astype(np.float32)
    return x, w, mu, var


def check_aux(got, ref, x):
    D = x.shape[1]
    dmax = float(np.abs(np.diff(x[np.all(np.isfinite(x), 1)], axis=0)).max())
    err = np.abs(got - ref)
    assert np.all(err[:, :D] <= 2e-4 * np.abs(ref[:, :D]) + 5e-3 * dmax), err[:, :D].max()
    assert np.all(err[:, D] <= 2e-4 * ref[:, D] + 0.05), err[:, D].max()
    assert np.all(err[:, D + 1] <= 2e-4 * ref[:, D + 1] + 5e-3), err[:, D + 1].max()


@pytest.mark.parametrize("N,D,M", [(2, 3, 2), (500, 5, 4), (5000, 20, 32), (20000, 39, 64),
                                   (3000, 60, 128)])
def test_aux_diag_matches_oracle(acc, N, D, M):
    x, w, mu, var = markov_frames(100 + N + D + M, N, D, M)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    got = acc.accumulate_aux()
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    ref = orc.accumulate_aux(x, g)
    check_aux(got, ref, x)


def test_aux_single_component(acc):
    rng = np.random.default_rng(3)
    N, D = 3000, 7
    x = np.cumsum(rng.normal(0, 0.3, (N, D)), axis=0).astype(np.float32)
    acc.set_model(np.ones(1, np.float32), x.mean(0)[None], x.var(0)[None] + 1.0)
    acc.set_frames(x)
    got = acc.accumulate_aux()
    d = np.abs(np.diff(x, axis=0)).astype(np.float64).sum(0)
    assert np.allclose(got[0, :D], d, rtol=1e-6)
    assert got[0, D] == N - 1 and got[0, D + 1] == N - 1


def test_aux_chunking_and_accumulate_into(acc):
    """Pairs do not break at chunk boundaries; repeated calls are
    bit-identical; the into form adds."""
    x, w, mu, var = markov_frames(7, 6000, 12, 16)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    a = acc.accumulate_aux()
    b = acc.accumulate_aux()
    assert np.array_equal(a, b)
    twice = acc.accumulate_aux(into=a.copy())
    assert np.allclose(twice, 2 * a, rtol=1e-12)
    acc.set_opts(chunk_frames=700)
    c = acc.accumulate_aux()
    acc.set_opts(chunk_frames=0)
    assert np.allclose(c, a, rtol=1e-5, atol=1e-4)


def test_aux_full_covariance_matches_oracle(acc):
    rng = np.random.default_rng(11)
    N, D, M = 4000, 8, 6
    x, w, mu, _ = markov_frames(12, N, D, M)
    A = rng.normal(size=(M, D, D)) / np.sqrt(D)
    cov = (A @ A.transpose(0, 2, 1) + np.diag(rng.uniform(0.3, 1.5, D))).astype(np.float32)
    cov = (0.5 * (cov + cov.transpose(0, 2, 1))).astype(np.float32)
    acc.set_model_full(w, mu, cov)
    acc.set_frames(x)
    got = acc.accumulate_aux()
    _, g = orc.frame_loglik_full(x, w, mu, cov, want_gamma=True)
    check_aux(got, orc.accumulate_aux(x, g), x)


def test_aux_nan_frames(acc):
    """A non-finite frame has no posteriors: the pairs on both sides of it
    contribute nothing, and its values never reach the sums."""
    x, w, mu, var = markov_frames(13, 800, 6, 5)
    x[[0, 100, 101, 500]] = np.nan
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    got = acc.accumulate_aux()
    assert np.all(np.isfinite(got))
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    check_aux(got, orc.accumulate_aux(x, g), x)


def test_aux_errors(acc):
    x, w, mu, var = markov_frames(14, 100, 4, 3)
    acc.set_model(w, mu, var)
    acc.set_frames(x[:, :3].copy())
    with pytest.raises(gmm_amd.GmmHipError):
        acc.accumulate_aux()
