"""Parity of the HIP E-step (through the C ABI) with the FP64 oracle.

Tolerances (plain FP32 form):
  per-frame log-likelihood   |dLL| <= 1e-5 |LL|
  total log-likelihood       relative 1e-6
  n_m                        |dn| <= 1e-5 n + cut allowance of m
  F_md                       |dF| <= 1e-5 sqrt(n S) + cut allowance of m
  S_md                       |dS| <= 1e-5 S + cut allowance of m
The cut allowance of component m covers only the frames whose FP64
posterior for m lies within 1 % of the 1e-4 threshold (counted by the
oracle): each may land on the other side of the cut and moves the
statistics by at most its posterior times (1, |x|, x^2) (tests/_check.py).
Component indexing must be identical (statistics of component m land in
slot m).
"""
import ctypes

import numpy as np
import pytest

import _check
import _oracle as orc
import gmm_amd
from gmm_amd import synth

pytestmark = pytest.mark.gpu

RTOL = 1e-5


@pytest.fixture(scope="module")
def acc():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    a = gmm_amd.Accumulator(0)
    yield a
    a.close()


def rand_case(seed, N, D, M, mean_sd=2.0, offset=0.0):
    rng = np.random.default_rng(seed)
    w64 = rng.dirichlet(np.ones(M))
    mu = (rng.normal(0, mean_sd, (M, D)) + offset).astype(np.float32)
    var = rng.uniform(0.3, 3.0, (M, D)).astype(np.float32)
    x = synth.sample_frames_np(w64, mu, var, N, seed + 1)
    return x, w64.astype(np.float32), mu, var


def check_ll(ll_gpu, ll_ref, rtol=RTOL):
    err = np.abs(ll_gpu.astype(np.float64) - ll_ref)
    lim = rtol * np.abs(ll_ref) + 1e-6
    bad = np.nonzero(err > lim)[0]
    assert bad.size == 0, "LL mismatch at %s: gpu %s ref %s" % (
        bad[:5], ll_gpu[bad[:5]], ll_ref[bad[:5]])


def check_stats(got, ref, x, model, rtol=RTOL, tot_rtol=1e-6):
    w, mu, var = model
    return _check.check_stats(got, ref, x, w, mu, var, rtol=rtol, tot_rtol=tot_rtol,
                              band=_check.BAND["fp32"])


def run_gpu(acc, x, w, mu, var):
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    ll = acc.loglikes()
    st = acc.accumulate()
    return ll, st


# ----------------------------------------------------------------- fixtures
def test_golden_example_loglikes(acc, golden):
    """example/data.logL reproduced by the GPU path (N=10000, D=20, M=4)."""
    x, w, mu, var, ref = golden["x"], golden["w"], golden["mu"], golden["var"], golden["ll"]
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    ll = acc.loglikes().astype(np.float64)
    tol = 0.5 * 10.0 ** (np.floor(np.log10(np.abs(ref))) - 5) + 2e-6 * np.abs(ref)
    assert np.all(np.abs(ll - ref) <= tol), np.abs(ll - ref).max()
    check_ll(ll, orc.frame_loglik(x, w, mu, var))
    st = acc.accumulate()
    check_stats(st, orc.accumulate(x, w, mu, var), x, (w, mu, var))
    assert st[4] == x.shape[0]


# ----------------------------------------------------------------- configs
def test_plain_em_small_config(acc):
    """D=39, M=64, N=100,000 (the reference's CPU-runnable config), model
    initialised from 64 random frames with the global variance."""
    cfg = synth.CONFIGS["plain_em_small"]
    w0, mu0, var0 = synth.true_mixture(**cfg)
    x = synth.sample_frames_np(w0, mu0, var0, cfg["N"], cfg["seed"])
    w, mu, var = synth.init_model_from_frames(x, cfg["M"], cfg["seed"])
    ll, st = run_gpu(acc, x, w, mu, var)
    check_ll(ll, orc.frame_loglik(x, w, mu, var))
    check_stats(st, orc.accumulate(x, w, mu, var), x, (w, mu, var))


def test_plain_em_small_after_one_update(acc):
    """Second EM iteration (a separated model) -- many posteriors cut."""
    cfg = synth.CONFIGS["plain_em_small"]
    w0, mu0, var0 = synth.true_mixture(**cfg)
    x = synth.sample_frames_np(w0, mu0, var0, 30_000, cfg["seed"] + 7)
    w, mu, var = synth.init_model_from_frames(x, cfg["M"], cfg["seed"])
    n, F, S, _ = orc.accumulate(x, w, mu, var)
    w, mu, var, _ = orc.update_em(w, mu, var, n, F, S, x.shape[0])
    keep = w > 0
    w, mu, var = w[keep], mu[keep], var[keep]
    ll, st = run_gpu(acc, x, w, mu, var)
    check_ll(ll, orc.frame_loglik(x, w, mu, var))
    check_stats(st, orc.accumulate(x, w, mu, var), x, (w, mu, var))


@pytest.mark.parametrize("N,D,M", [
    (1, 3, 1), (2, 1, 2), (17, 5, 3), (129, 20, 17), (1000, 60, 33),
    (4999, 13, 100), (300, 127, 16), (2050, 39, 64), (777, 60, 513), (1500, 60, 2048),
])
def test_shapes(acc, N, D, M):
    x, w, mu, var = rand_case(1000 + N + D + M, N, D, M)
    ll, st = run_gpu(acc, x, w, mu, var)
    check_ll(ll, orc.frame_loglik(x, w, mu, var))
    check_stats(st, orc.accumulate(x, w, mu, var), x, (w, mu, var))
    assert st[4] == N


def test_offset_frames(acc):
    """Frame offsets ~N(5,1) per dim (x^2 cancellation stress)."""
    rng = np.random.default_rng(3)
    N, D, M = 20000, 60, 128
    x, w, mu, var = rand_case(3, N, D, M, offset=rng.normal(5, 1, (1, D)))
    ll, st = run_gpu(acc, x, w, mu, var)
    check_ll(ll, orc.frame_loglik(x, w, mu, var))
    check_stats(st, orc.accumulate(x, w, mu, var), x, (w, mu, var))


def test_lse_large_jumps_between_component_tiles(acc):
    """The first component tiles sit far from the data (log-likelihoods
    ~ -1e4) and later tiles close to it: the running log-sum-exp must move
    its reference by far more than the exp2 range."""
    rng = np.random.default_rng(21)
    N, D, M = 3000, 20, 80
    x, w, mu, var = rand_case(21, N, D, M, mean_sd=1.0)
    mu[:32] += 60.0
    w = w.copy()
    w[:32] = 1e-30
    w /= w.sum()
    ll, st = run_gpu(acc, x, w, mu, var)
    ref_ll = orc.frame_loglik(x, w, mu, var)
    check_ll(ll, ref_ll)
    check_stats(st, orc.accumulate(x, w, mu, var), x, (w, mu, var))
    # reversed order: near components first
    ll2, _ = run_gpu(acc, x, w[::-1].copy(), mu[::-1].copy(), var[::-1].copy())
    check_ll(ll2, ref_ll)


def test_single_component_posterior_is_one(acc):
    x, w, mu, var = rand_case(5, 3000, 12, 1)
    ll, st = run_gpu(acc, x, w, mu, var)
    assert st[0][0] == pytest.approx(3000.0, rel=1e-6)
    assert np.allclose(st[1][0], x.astype(np.float64).sum(0), rtol=1e-5,
                       atol=1e-5 * np.abs(x).sum(0).max())


# ----------------------------------------------------------------- semantics
def test_nan_frames_contribute_nothing(acc):
    x, w, mu, var = rand_case(11, 500, 8, 5)
    x[[3, 100, 499]] = np.nan
    ll, st = run_gpu(acc, x, w, mu, var)
    assert np.all(ll[[3, 100, 499]] == np.float32(-1e20))
    ref = orc.accumulate(x, w, mu, var)
    xs = np.delete(x, [3, 100, 499], 0)
    check_stats((st[0], st[1], st[2], st[3] - 3 * orc.MIN_LOGLIKE),
                (ref[0], ref[1], ref[2], ref[3] - 3 * orc.MIN_LOGLIKE), xs, (w, mu, var))


def test_low_variance_rejected(acc):
    w = np.array([1.0], np.float32)
    mu = np.zeros((1, 2), np.float32)
    var = np.array([[1.0, 1e-7]], np.float32)
    with pytest.raises(gmm_amd.GmmHipError) as ei:
        acc.set_model(w, mu, var)
    assert ei.value.code == gmm_amd.ERR_VAR


def test_zero_weight_component_rejected(acc):
    w = np.array([1.0, 0.0], np.float32)
    with pytest.raises(gmm_amd.GmmHipError) as ei:
        acc.set_model(w, np.zeros((2, 2), np.float32), np.ones((2, 2), np.float32))
    assert ei.value.code == gmm_amd.ERR_ARG


def test_row_pointer_frames_match_contiguous(acc):
    x, w, mu, var = rand_case(12, 700, 10, 9)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    a = acc.accumulate()
    rows = [np.array(r) for r in x[::-1]][::-1]  # separate allocations
    acc.set_frames_rows(rows)
    b = acc.accumulate()
    for u, v in zip(a, b):
        assert np.array_equal(np.asarray(u), np.asarray(v))


def test_accumulate_into_and_determinism(acc):
    x, w, mu, var = rand_case(13, 4000, 16, 40)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    a = acc.accumulate()
    b = acc.accumulate()
    for u, v in zip(a, b):
        assert np.array_equal(np.asarray(u), np.asarray(v))  # bitwise repeatable
    c = acc.accumulate(into=(a[0].copy(), a[1].copy(), a[2].copy(), a[3], a[4]))
    assert np.array_equal(c[0], a[0] + b[0])
    assert c[4] == 2 * a[4]


def test_flags_export_only_requested(acc):
    x, w, mu, var = rand_case(14, 1000, 6, 7)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    n = np.zeros(7)
    tot = ctypes.c_double(0)
    nf = ctypes.c_uint64(0)
    rc = gmm_amd.lib().gmmhip_accumulate(acc.handle, 0, gmm_amd._dp(n), None, None,
                                         ctypes.byref(tot), ctypes.byref(nf))
    assert rc == 0 and nf.value == 1000
    assert n.sum() > 0


def test_component_indexing(acc):
    """Well-separated components: statistics of component m land in slot m
    and follow a permutation of the model."""
    rng = np.random.default_rng(15)
    M, D = 37, 4
    mu = (np.arange(M)[:, None] * 10.0 + np.zeros((1, D))).astype(np.float32)
    var = np.ones((M, D), np.float32)
    w64 = rng.dirichlet(np.ones(M))
    w = w64.astype(np.float32)
    x = synth.sample_frames_np(w64, mu, var, 5000, 16)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    n1 = acc.accumulate()[0]
    perm = rng.permutation(M)
    acc.set_model(w[perm], mu[perm], var[perm])
    n2 = acc.accumulate()[0]
    assert np.allclose(n2, n1[perm], rtol=1e-6)
    counts = np.bincount(np.argmin(np.abs(x[:, :1] - mu[:, 0][None]), 1), minlength=M)
    assert np.allclose(n1, counts, atol=1e-3)


def test_chunk_size_invariance(acc):
    """Different FP32 partial sizes agree to FP32 accumulation accuracy."""
    x, w, mu, var = rand_case(17, 60000, 24, 48)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    ref = orc.accumulate(x, w, mu, var)
    for cf in (64, 1000, 8192, 60000):
        acc.set_opts(chunk_frames=cf)
        check_stats(acc.accumulate(), ref, x, (w, mu, var), rtol=2e-5)
    acc.set_opts(chunk_frames=0)


def test_large_n_properties(acc):
    """At 2M frames (oracle too slow): sum of occupancies, total LL equals the
    sum of per-frame LLs, and the statistics of the two halves add up."""
    N, D, M = 2_000_000, 60, 64
    x, w, mu, var = rand_case(18, N, D, M, mean_sd=0.5)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    ll = acc.loglikes()
    st = acc.accumulate()
    assert st[3] == pytest.approx(ll.astype(np.float64).sum(), rel=1e-9)
    assert st[0].sum() == pytest.approx(N, rel=2e-3)  # only cut posteriors missing
    acc.set_frames(x[: N // 2])
    a = acc.accumulate()
    acc.set_frames(x[N // 2:])
    b = acc.accumulate()
    assert np.allclose(a[0] + b[0], st[0], rtol=1e-5)
    assert np.allclose(a[2] + b[2], st[2], rtol=1e-5)
    assert a[3] + b[3] == pytest.approx(st[3], rel=1e-9)
