"""Mixed-precision E-step (split fp16 / bf16 operands, FP32 accumulate, FP64
reduction) against the FP64 oracle.

Stated tolerances (per-frame log-likelihood relative / total
log-likelihood relative / statistics rtol on n, sqrt(nS)-scaled F, S):
  fp16x3   1e-5 / 1e-6 / 2e-4
  bf16x3   2e-4 / 2e-5 / 2e-3
The statistics tolerance also absorbs posteriors that cross the 1e-4 cut.
"""
import numpy as np
import pytest

import _oracle as orc
import gmm_amd
from gmm_amd import synth

pytestmark = pytest.mark.gpu

TOL = {gmm_amd.PREC_FP16X3: (1e-5, 1e-6, 2e-4), gmm_amd.PREC_BF16X3: (2e-4, 2e-5, 2e-3)}
PRECS = [gmm_amd.PREC_FP16X3, gmm_amd.PREC_BF16X3]


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
    assert got[3] == pytest.approx(rtot, rel=tt)
    n, F, S = got[:3]
    xm = float(np.abs(x).max())
    assert np.all(np.abs(n - rn) <= st * rn + 3e-4), np.max(np.abs(n - rn) / np.maximum(rn, 1e-30))
    scaleF = np.sqrt(np.maximum(rn[:, None] * rS, 0))
    assert np.all(np.abs(F - rF) <= st * scaleF + 3e-4 * xm)
    assert np.all(np.abs(S - rS) <= st * rS + 3e-4 * xm * xm)
    return ll, got


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("N,D,M", [(1, 3, 1), (77, 5, 3), (1000, 20, 33), (5000, 60, 100),
                                   (3000, 39, 64), (400, 127, 40), (2049, 60, 257),
                                   (600, 70, 20), (500, 90, 9), (300, 100, 12)])
def test_mixed_shapes(acc, prec, N, D, M):
    """Depths KP = 32, 64, 128, 256 (statistics from the ZA rows by
    transposing LDS reads) and 96, 160, 192, 224 (register-staged statistics
    kernel with the ZB copy)."""
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
    rn = orc.accumulate(np.delete(x, [5, 300], 0), w, mu, var)[0]
    assert np.all(np.abs(n - rn) <= TOL[prec][2] * rn + 3e-4)


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
    UBM-like case (D=60, M=512, N=20,000; the numbers are printed)."""
    cfg = synth.CONFIGS["mixed_estep"]
    w0, mu0, var0 = synth.true_mixture(**dict(cfg, M=512))
    x = synth.sample_frames_np(w0, mu0, var0, 20000, 99)
    w, mu, var = synth.init_model_from_frames(x, 512, 99)
    rll = orc.frame_loglik(x, w, mu, var)
    rst = orc.accumulate(x, w, mu, var)
    for prec, name in [(gmm_amd.PREC_FP32, "fp32"), (gmm_amd.PREC_FP16X3, "fp16x3"),
                       (gmm_amd.PREC_BF16X3, "bf16x3")]:
        acc.set_opts(precision=prec)
        acc.set_model(w, mu, var)
        acc.set_frames(x)
        ll = acc.loglikes().astype(np.float64)
        n, F, S, tot, _ = acc.accumulate()
        print("%s: max rel LL %.3g, rel totLL %.3g, max rel n %.3g, max |dmu| %.3g" % (
            name, np.max(np.abs(ll - rll) / np.abs(rll)), abs(tot - rst[3]) / abs(rst[3]),
            np.max(np.abs(n - rst[0]) / np.maximum(rst[0], 1)),
            np.max(np.abs(F / np.maximum(n, 1e-30)[:, None] -
                          rst[1] / np.maximum(rst[0], 1e-30)[:, None])[rst[0] > 10])))
    acc.set_opts(precision=gmm_amd.PREC_FP32)
