"""Parity of the HIP full-covariance E-step (gmmhip_set_model_full +
gmmhip_accumulate / gmmhip_loglikes through the C ABI) with the FP64 oracle.

Tolerances (FP32 form; the log-likelihood and statistics contractions run
over K = 1 + D + D(D+1)/2 features, up to 1891 at D = 60, so the per-frame
log-likelihood carries a larger absolute FP32 error than the diagonal form's):
  per-frame log-likelihood   |dLL| <= 1e-5 |LL| + 1e-4   (measured: <= 5e-6 relative at D = 127)
  total log-likelihood       relative 2e-6
  n_m                        |dn| <= 1e-4 n + 3e-4 (cut allowance)
  F_md                       |dF| <= 1e-4 sqrt(n S_dd) + 3e-4 max|x|
  S_m(k,l)                   |dS| <= 1e-4 sqrt(S_kk S_ll) + 3e-4 max x^2
"""
import os

import numpy as np
import pytest

import _formats as fm
import _oracle as orc
import gmm_amd

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def acc():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    a = gmm_amd.Accumulator(0)
    yield a
    a.close()


def full_case(seed, N, D, M, spread=2.0):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, spread, (M, D)).astype(np.float32)
    A = rng.normal(size=(M, D, D)) / np.sqrt(D)
    cov = (A @ A.transpose(0, 2, 1) + np.diag(rng.uniform(0.3, 1.5, D))).astype(np.float32)
    cov = (0.5 * (cov + cov.transpose(0, 2, 1))).astype(np.float32)
    p = w.astype(np.float64)
    k = rng.choice(M, size=N, p=p / p.sum())
    L = np.linalg.cholesky(cov.astype(np.float64))
    x = (mu[k] + np.einsum("nij,nj->ni", L[k], rng.normal(size=(N, D)))).astype(np.float32)
    return x, w, mu, cov


def check_full(acc, x, w, mu, cov, label=""):
    N, D = x.shape
    acc.set_model_full(w, mu, cov)
    acc.set_frames(x)
    ll = acc.loglikes().astype(np.float64)
    n, F, S, tot, nf = acc.accumulate()
    rll = orc.frame_loglik_full(x, w, mu, cov)
    rn, rF, rS, rtot = orc.accumulate_full(x, w, mu, cov)
    r, c = np.triu_indices(D)
    dd = np.where(r == c)[0]
    rSd = np.maximum(rS[:, dd], 0)  # S_kk per component
    xm = float(np.abs(x).max())
    err_ll = np.max(np.abs(ll - rll) / (np.abs(rll) + 1))
    print("%s N=%d D=%d M=%d: max rel dLL %.2e, tot rel %.2e, max dn %.2e" % (
        label, N, D, len(w), err_ll, abs(tot - rtot) / abs(rtot), np.abs(n - rn).max()))
    assert nf == N
    assert np.all(np.abs(ll - rll) <= 1e-5 * np.abs(rll) + 1e-4), err_ll
    assert tot == pytest.approx(rtot, rel=2e-6)
    assert np.all(np.abs(n - rn) <= 1e-4 * rn + 3e-4)
    assert np.all(np.abs(F - rF) <= 1e-4 * np.sqrt(rn[:, None] * rSd) + 3e-4 * xm)
    scale = np.sqrt(rSd[:, r] * rSd[:, c])
    assert np.all(np.abs(S - rS) <= 1e-4 * scale + 3e-4 * xm * xm)
    return ll, (n, F, S, tot)


def test_diagonal_model_as_full_reproduces_data_logl(acc, golden):
    """example/myModel.gmm written as full matrices: the per-frame
    log-likelihoods match example/data.logL and the diagonal form's."""
    w, mu, var, x, ref = golden["w"], golden["mu"], golden["var"], golden["x"], golden["ll"]
    D = mu.shape[1]
    cov = np.zeros((len(w), D, D), np.float32)
    cov[:, np.arange(D), np.arange(D)] = var
    acc.set_model_full(w, mu, cov)
    acc.set_frames(x)
    ll = acc.loglikes().astype(np.float64)
    tol = 0.5 * 10.0 ** (np.floor(np.log10(np.abs(ref))) - 5) + 2e-5 * np.abs(ref)
    assert np.all(np.abs(ll - ref) <= tol), np.abs(ll - ref).max()
    acc.set_model(w, mu, var)
    lld = acc.loglikes().astype(np.float64)
    assert np.abs(ll - lld).max() <= 1e-4 * np.abs(lld).max()


def test_mymodelfc_on_data_bin(acc, golden):
    """The reference's example/myModelFC.gmm on example/data.bin (the
    getLogLike.bat:8 command), against the oracle."""
    w, mu, cov, _ = fm.read_gmm_txt_full(os.path.join(HERE, "golden", "myModelFC.gmm"))
    check_full(acc, golden["x"], w, mu, cov, "myModelFC")


@pytest.mark.parametrize("N,D,M", [(1, 3, 1), (77, 5, 3), (1000, 20, 4), (3000, 39, 64),
                                   (2000, 60, 33), (500, 127, 5)])
def test_full_shapes(acc, N, D, M):
    x, w, mu, cov = full_case(900 + N + D + M, N, D, M)
    check_full(acc, x, w, mu, cov)


def test_full_nan_frames(acc):
    x, w, mu, cov = full_case(31, 600, 8, 5)
    x[[3, 400]] = np.nan
    acc.set_model_full(w, mu, cov)
    acc.set_frames(x)
    ll = acc.loglikes()
    assert np.all(ll[[3, 400]] == np.float32(-1e20))
    n = acc.accumulate()[0]
    rn = orc.accumulate_full(np.delete(x, [3, 400], 0), w, mu, cov)[0]
    assert np.all(np.abs(n - rn) <= 1e-4 * rn + 3e-4)


def test_full_determinism_chunking_and_accumulate_into(acc):
    x, w, mu, cov = full_case(37, 5000, 12, 20)
    acc.set_model_full(w, mu, cov)
    acc.set_frames(x)
    a = acc.accumulate()
    b = acc.accumulate()
    for u, v in zip(a, b):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    twice = acc.accumulate(into=tuple(np.copy(t) if isinstance(t, np.ndarray) else t for t in a))
    assert np.allclose(twice[0], 2 * a[0], rtol=1e-12) and twice[4] == 2 * a[4]
    acc.set_opts(chunk_frames=1000)
    c = acc.accumulate()
    acc.set_opts(chunk_frames=0)
    assert np.allclose(c[0], a[0], rtol=1e-5, atol=1e-4)
    assert np.allclose(c[2], a[2], rtol=1e-4, atol=1e-2)
    # the GEMM kernel (and so the FP32 summation order over the features) is
    # chosen per chunk shape: per-frame log-likelihoods may differ in the
    # last bits between chunk sizes, never between runs of one size
    assert c[3] == pytest.approx(a[3], rel=1e-8)


def test_full_errors(acc):
    x, w, mu, cov = full_case(41, 100, 4, 3)
    bad = cov.copy()
    bad[1] = -np.eye(4, dtype=np.float32)
    with pytest.raises(gmm_amd.GmmHipError):
        acc.set_model_full(w, mu, bad)
    acc.set_model_full(w, mu, cov)
    acc.set_frames(x)
    acc.set_opts(precision=gmm_amd.PREC_FP16X3)
    with pytest.raises(gmm_amd.GmmHipError):
        acc.accumulate()
    acc.set_opts(precision=gmm_amd.PREC_FP32)
    acc.accumulate()
    # a diagonal model inserted afterwards replaces the full one
    acc.set_model(w, mu, np.ones((3, 4), np.float32))
    n, F, S, tot, nf = acc.accumulate()
    assert S.shape == (3, 4)


@pytest.mark.parametrize("N,D,M", [(3000, 6, 16), (2000, 39, 64)])
def test_full_second_order_statistics_of_a_diagonal_model(acc, N, D, M):
    """GMMHIP_ACC_FULLVAR (setAccFlags' varFull, what fMLLR accumulates):
    the posteriors of a diagonal model, the packed upper sum g x x^T."""
    rng = np.random.default_rng(N + D)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 1.5, (M, D)).astype(np.float32)
    x = rng.normal(0, 2.5, (N, D)).astype(np.float32)
    acc.set_opts(precision=gmm_amd.PREC_FP32)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    n, F, S, tot, nf = acc.accumulate(gmm_amd.ACC_MEAN | gmm_amd.ACC_VAR | gmm_amd.ACC_FULLVAR)
    cov = np.zeros((M, D, D), np.float32)
    cov[:, np.arange(D), np.arange(D)] = var
    rn, rF, rS, rtot = orc.accumulate_full(x, w, mu, cov)
    assert nf == N and tot == pytest.approx(rtot, rel=2e-6)
    r, c = np.triu_indices(D)
    rSd = np.maximum(rS[:, r == c], 0)
    xm = float(np.abs(x).max())
    assert np.all(np.abs(n - rn) <= 1e-4 * rn + 3e-4)
    assert np.all(np.abs(F - rF) <= 1e-4 * np.sqrt(rn[:, None] * rSd) + 3e-4 * xm)
    assert np.all(np.abs(S - rS) <= 1e-4 * np.sqrt(rSd[:, r] * rSd[:, c]) + 3e-4 * xm * xm)
    # the diagonal of the full statistics matches the diagonal path
    n2, F2, S2, _, _ = acc.accumulate()
    assert np.allclose(S[:, r == c], S2, rtol=2e-4, atol=3e-4 * xm * xm)
    # mixed precision: the GEMM path still runs in FP32
    acc.set_opts(precision=gmm_amd.PREC_FP16X3)
    n3, _, S3, _, _ = acc.accumulate(gmm_amd.ACC_MEAN | gmm_amd.ACC_VAR | gmm_amd.ACC_FULLVAR)
    acc.set_opts(precision=gmm_amd.PREC_FP32)
    assert np.array_equal(S3, S) and np.array_equal(n3, n)
