"""Mixed-precision E-step (split fp16 / bf16 operands, FP32 accumulate, FP64
reduction) against the FP64 oracle.

Stated tolerances (per-frame log-likelihood relative / total
log-likelihood relative / statistics rtol on n, sqrt(nS)-scaled F, S):
  fp16x3   1e-5 / 1e-6 / 2e-4
  bf16x3   2e-4 / 2e-5 / 2e-3
Posteriors that cross the 1e-4 cut get a per-component allowance for the
frames whose FP64 posterior lies within 1 % (fp16x3) or 10 % (bf16x3) of the
cut (tests/_check.py).
"""
import numpy as np
import pytest

import _check
import _oracle as orc
import gmm_amd
from gmm_amd import synth

pytestmark = pytest.mark.gpu

TOL = {gmm_amd.PREC_FP16X3: (1e-5, 1e-6, 2e-4), gmm_amd.PREC_BF16X3: (2e-4, 2e-5, 2e-3)}
PRECS = [gmm_amd.PREC_FP16X3, gmm_amd.PREC_BF16X3]
NAME = {gmm_amd.PREC_FP16X3: "fp16x3", gmm_amd.PREC_BF16X3: "bf16x3"}
BAND = {p: _check.BAND[NAME[p]] for p in PRECS}


@pytest.fixture(scope="module")
def acc():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    a = gmm_amd.Accumulator(0)
    yield a
    a.close()


def case(seed, N, D, M, offset=0.0, mean_sd=2.0):
    rng = np.random.default_rng(seed)
    w64 = rng.dirichlet(np.ones(M))
    mu = (rng.normal(0, mean_sd, (M, D)) + offset).astype(np.float32)
    var = rng.uniform(0.3, 3.0, (M, D)).astype(np.float32)
    x = synth.sample_frames_np(w64, mu, var, N, seed + 1)
    return x, w64.astype(np.float32), mu, var


def check(acc, prec, x, w, mu, var, ref_ll=None, ref_st=None):
    lt, tt, st = TOL[prec]
    acc.set_opts(precision=prec)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    ll = acc.loglikes().astype(np.float64)
    got = acc.accumulate()
    acc.set_opts(precision=gmm_amd.PREC_FP32)
    rll = orc.frame_loglik(x, w, mu, var) if ref_ll is None else ref_ll
    rn, rF, rS, rtot = orc.accumulate(x, w, mu, var) if ref_st is None else ref_st
    assert np.all(np.abs(ll - rll) <= lt * np.abs(rll) + 1e-6), np.max(np.abs(ll - rll) / np.abs(rll))
    used = _check.check_stats(got, (rn, rF, rS, rtot), x, w, mu, var, rtol=st, tot_rtol=tt,
                              band=BAND[prec])
    print("%s N=%d D=%d M=%d: max rel LL %.3g, allowance used %.3f" % (
        NAME[prec], x.shape[0], x.shape[1], mu.shape[0],
        np.max(np.abs(ll - rll) / np.abs(rll)), used))
    return ll, got


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("N,D,M", [(1, 3, 1), (77, 5, 3), (1000, 20, 33), (5000, 60, 100),
                                   (3000, 39, 64), (400, 127, 40), (2049, 60, 257),
                                   (600, 70, 20), (500, 90, 9), (300, 100, 12)])
def test_mixed_shapes(acc, prec, N, D, M):
    """Contraction depths 2D+1 rounded up to 32 = 32, 64, 128, 256 (feature
    rows of that depth) and 96, 160, 192, 224 (feature rows padded with zero
    columns to 128 or 256, of which the log-likelihood kernel reads only the
    contraction depth; the statistics kernel reads the whole row)."""
    x, w, mu, var = case(500 + N + D + M, N, D, M)
    check(acc, prec, x, w, mu, var)


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_offset_frames(acc, prec):
    """Frame offsets ~N(5,1) per dim: the mixed form centres and scales the
    frames, so x^2 does not cancel."""
    rng = np.random.default_rng(7)
    x, w, mu, var = case(7, 20000, 60, 128, offset=rng.normal(5, 1, (1, 60)))
    check(acc, prec, x, w, mu, var)


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_nan_frames(acc, prec):
    x, w, mu, var = case(9, 600, 8, 40)
    x[[5, 300]] = np.nan
    acc.set_opts(precision=prec)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    ll = acc.loglikes()
    n = acc.accumulate()[0]
    acc.set_opts(precision=gmm_amd.PREC_FP32)
    assert np.all(ll[[5, 300]] == np.float32(-1e20))
    xs = np.delete(x, [5, 300], 0)
    rn, rF, _, _ = orc.accumulate(xs, w, mu, var)
    _check.check_stats((n, rF, None, None), (rn, rF, None, None), xs, w, mu, var,
                       rtol=TOL[prec][2], band=BAND[prec])


def test_fp16_out_of_range_runs_fp32(acc):
    """A frame 1e4 standard deviations out does not fit FP16 after scaling:
    the call runs the FP32 form and matches it bit for bit."""
    x, w, mu, var = case(11, 3000, 10, 20)
    x[17, 3] = 1e6
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    ref = acc.accumulate()
    acc.set_opts(precision=gmm_amd.PREC_FP16X3)
    got = acc.accumulate()
    acc.set_opts(precision=gmm_amd.PREC_FP32)
    for u, v in zip(ref, got):
        assert np.array_equal(np.asarray(u), np.asarray(v))


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_determinism_and_chunking(acc, prec):
    x, w, mu, var = case(13, 40000, 24, 96)
    acc.set_opts(precision=prec)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    a = acc.accumulate()
    b = acc.accumulate()
    for u, v in zip(a, b):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    acc.set_opts(precision=prec, chunk_frames=1024)
    c = acc.accumulate()
    acc.set_opts(precision=gmm_amd.PREC_FP32, chunk_frames=0)
    assert np.allclose(c[0], a[0], rtol=1e-5, atol=1e-4)
    assert c[3] == pytest.approx(a[3], rel=1e-12)


def test_mixed_error_report_ubm_like(acc):
    """Errors of both mixed forms and the FP32 form against FP64 on a
    UBM-like case (D=60, M=512, N=20,000, the mixed workload's mixture with
    frame offsets): per-frame LL, total LL, occupancies and M-step means,
    each within the form's stated tolerance."""
    cfg = synth.CONFIGS["mixed_estep"]
    w0, mu0, var0 = synth.true_mixture(**dict(cfg, M=512))
    x = synth.sample_frames_np(w0, mu0, var0, 20000, 99)
    w, mu, var = synth.init_model_from_frames(x, 512, 99)
    rll = orc.frame_loglik(x, w, mu, var)
    rst = orc.accumulate(x, w, mu, var)
    rn = rst[0]
    # (per-frame LL rel, total LL rel, occupancy rel and mean abs for
    # components with n >= 100, where a posterior crossing the cut moves
    # the mean by <= 1e-4 |x - mu| / 100); the per-component check below is
    # the exact one
    lim = {"fp32": (1e-5, 1e-6, 2e-5, 5e-5), "fp16x3": (1e-5, 1e-6, 2e-4, 2e-4),
           "bf16x3": (2e-4, 2e-5, 2e-3, 2e-3)}
    big = rn >= 100
    for prec, name in [(gmm_amd.PREC_FP32, "fp32"), (gmm_amd.PREC_FP16X3, "fp16x3"),
                       (gmm_amd.PREC_BF16X3, "bf16x3")]:
        acc.set_opts(precision=prec)
        acc.set_model(w, mu, var)
        acc.set_frames(x)
        ll = acc.loglikes().astype(np.float64)
        n, F, S, tot, _ = acc.accumulate()
        e_ll = np.max(np.abs(ll - rll) / np.abs(rll))
        e_tot = abs(tot - rst[3]) / abs(rst[3])
        e_n = np.max(np.abs(n - rn)[big] / rn[big])
        e_mu = np.max(np.abs(F / np.maximum(n, 1e-30)[:, None] -
                             rst[1] / np.maximum(rn, 1e-30)[:, None])[big])
        print("%s: max rel LL %.3g, rel totLL %.3g, max rel n %.3g, max |dmu| %.3g" % (
            name, e_ll, e_tot, e_n, e_mu))
        L = lim[name]
        assert e_ll <= L[0] and e_tot <= L[1] and e_n <= L[2] and e_mu <= L[3], name
        _check.check_stats((n, F, S, tot), rst, x, w, mu, var, rtol=max(L[2], 1e-5),
                           tot_rtol=L[1], band=_check.BAND[name], what=name + " ")
    acc.set_opts(precision=gmm_amd.PREC_FP32)


def test_split_model_overflow_runs_fp32(acc):
    """A model coefficient beyond the FP16 range (a variance of 2e-6 in one
    dimension: its quadratic coefficient is ~1e6 after the frames' scaling)
    is flagged by the split-model packing on the GPU and sends the fp16x3 call
    to the FP32 kernels: log-likelihoods and statistics equal the FP32 form's
    bit for bit and no mixed statistics kernel ran.  bf16x3, whose range holds
    the coefficient, stays mixed."""
    x, w, mu, var = case(31, 3000, 12, 20)
    var = var.copy()
    var[3, 5] = 2e-6  # just above the 1e-6 variance floor
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    acc.set_opts(precision=gmm_amd.PREC_FP32)
    ll32 = acc.loglikes()
    r32 = acc.accumulate()
    try:
        acc.set_opts(precision=gmm_amd.PREC_FP16X3)
        ll16 = acc.loglikes()
        r16 = acc.accumulate()
        tiles16 = acc.tiles_used()
        acc.set_opts(precision=gmm_amd.PREC_BF16X3)
        llb = acc.loglikes().astype(np.float64)
        acc.accumulate()
        tilesb = acc.tiles_used()
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32)
    assert np.array_equal(ll16, ll32)
    for a, b in zip(r16, r32):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert tiles16[2] == 0, tiles16
    assert tilesb[2] != 0, tilesb
    rll = ll32.astype(np.float64)
    assert np.all(np.abs(llb - rll) <= TOL[gmm_amd.PREC_BF16X3][0] * np.abs(rll) + 1e-6)
