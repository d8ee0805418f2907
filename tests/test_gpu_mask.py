"""Gaussian selection (gmmhip_set_mask) and posterior export
(gmmhip_posteriors) through the C ABI, against the oracle.

Tolerances (FP32 GEMM form, as the full-covariance path):
  per-frame log-likelihood   |dLL| <= 1e-5 |LL| + 1e-4
  n_m                        |dn| <= 1e-4 n + 3e-4 (cut allowance)
  F_md                       |dF| <= 1e-4 |F| + 3e-4 max|x| (+ n-scaled)
  posteriors                 |dg| <= 2e-4
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


def case(seed, N, D, M):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 2.0, (M, D)).astype(np.float32)
    x = rng.normal(0, 2.5, (N, D)).astype(np.float32)
    return rng, w, mu, var, x


def check_stats(got, ref, x):
    n, F, S = got
    rn, rF, rS = ref
    xm = float(np.abs(x).max())
    assert np.all(np.abs(n - rn) <= 1e-4 * rn + 3e-4), np.abs(n - rn).max()
    assert np.all(np.abs(F - rF) <= 1e-4 * np.abs(rF) + 3e-4 * xm + 1e-4 * rn[:, None] * xm)
    assert np.all(np.abs(S - rS) <= 1e-4 * np.abs(rS) + 3e-4 * xm * xm + 1e-4 * rn[:, None] * xm * xm)


@pytest.mark.parametrize("N,D,M", [(500, 5, 8), (4000, 20, 64), (2000, 60, 256)])
def test_masked_estep_matches_oracle(acc, N, D, M):
    rng, w, mu, var, x = case(N + D + M, N, D, M)
    lists = [np.sort(rng.choice(M, size=rng.integers(0, min(M, 12) + 1), replace=False))
             for _ in range(N)]
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    acc.set_mask(lists)
    ll = acc.loglikes().astype(np.float64)
    n, F, S, tot, nf = acc.accumulate()
    RLL, rn, rF, rS, rtot = orc.accumulate_masked(x, w, mu, var, lists)
    assert np.all(np.abs(ll - RLL) <= 1e-5 * np.abs(RLL) + 1e-4)
    assert tot == pytest.approx(rtot, rel=2e-6)
    check_stats((n, F, S), (rn, rF, rS), x)
    acc.set_mask(None)


def test_full_selection_and_removal(acc):
    rng, w, mu, var, x = case(5, 3000, 12, 32)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    plain = acc.accumulate()
    ll_plain = acc.loglikes()
    acc.set_mask([range(32)] * len(x))
    masked = acc.accumulate()
    ll_masked = acc.loglikes()
    assert np.allclose(ll_masked, ll_plain, rtol=1e-5, atol=1e-4)
    check_stats(masked[:3], plain[:3], x)
    acc.set_mask(None)
    again = acc.accumulate()
    for u, v in zip(plain, again):
        assert np.array_equal(np.asarray(u), np.asarray(v))


def test_mask_must_fit_frames_and_model(acc):
    rng, w, mu, var, x = case(6, 100, 4, 6)
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    acc.set_mask([[0, 1]] * 100)
    acc.set_frames(x[:50].copy())
    with pytest.raises(gmm_amd.GmmHipError):
        acc.accumulate()
    with pytest.raises(gmm_amd.GmmHipError):
        acc.set_mask([[7]] * 50)  # component out of range
    acc.set_mask(None)
    acc.accumulate()


def test_posteriors_match_oracle(acc):
    rng, w, mu, var, x = case(7, 2000, 10, 24)
    x[[5, 77]] = np.nan
    acc.set_model(w, mu, var)
    acc.set_frames(x)
    ll, post = acc.posteriors()
    RLL, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    ok = np.isfinite(x).all(1)
    assert np.all(np.abs(ll[ok] - RLL[ok]) <= 1e-5 * np.abs(RLL[ok]) + 1e-4)
    assert np.abs(post[ok] - g[ok]).max() <= 2e-4
    # clamped frames: min_gamma everywhere, as compGammas leaves them
    assert np.all(ll[~ok] == np.float32(-1e20))
    assert np.all(post[~ok] == np.float32(1e-4))
    # estimateMaskIdxs from the export, then the masked E-step
    lists = [np.nonzero(row > 1e-4)[0] for row in post[ok]]
    acc.set_frames(np.ascontiguousarray(x[ok]))
    acc.set_mask(lists)
    got = acc.accumulate()
    _, rn, rF, rS, _ = orc.accumulate_masked(x[ok], w, mu, var, lists)
    check_stats(got[:3], (rn, rF, rS), x[ok])
    acc.set_mask(None)


def test_full_covariance_model_with_mask(acc):
    rng = np.random.default_rng(8)
    N, D, M = 1500, 6, 10
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    A = rng.normal(size=(M, D, D)) / np.sqrt(D)
    cov = (A @ A.transpose(0, 2, 1) + np.eye(D)).astype(np.float32)
    cov = (0.5 * (cov + cov.transpose(0, 2, 1))).astype(np.float32)
    x = rng.normal(0, 2.5, (N, D)).astype(np.float32)
    acc.set_model_full(w, mu, cov)
    acc.set_frames(x)
    plain = acc.accumulate()
    acc.set_mask([range(M)] * N)
    masked = acc.accumulate()
    acc.set_mask(None)
    assert np.allclose(masked[0], plain[0], rtol=1e-6, atol=1e-6)
    assert masked[3] == pytest.approx(plain[3], rel=1e-9)


def test_cpp_estimator_mask_api(tmp_path, golden):
    """The C++ drop-in's mask2fill / estimateMaskIdxs / mask2use on the
    reference's example model and data: the filled posteriors against the
    oracle, then the masked statistics against the oracle's masked E-step
    on the same selections."""
    import os
    import subprocess
    import _formats as fm
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "mask_api"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "gmm_amd", "csrc", "host"),
                    os.path.join(root, "tests", "cpp", "mask_api.cpp"), "-o", str(exe),
                    "-L", os.path.join(root, "gmm_amd", "_build"), "-lgmmhip", "-pthread",
                    "-Wl,-rpath," + os.path.join(root, "gmm_amd", "_build"),
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    x = golden["x"][:3000].copy()
    fm.write_raw(tmp_path / "x.raw", x)
    out = subprocess.run([str(exe), os.path.join(root, "tests", "golden", "myModel.gmm"),
                          str(tmp_path / "x.raw")], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    N, M = int(lines[0].split()[1]), int(lines[0].split()[3])
    rows = np.array([[float(v) for v in ln.split()[1:]] for ln in lines[1:1 + N]])
    ll, post = rows[:, 0], rows[:, 1:]
    w, mu, var = golden["w"], golden["mu"], golden["var"]
    RLL, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    assert np.all(np.abs(ll - RLL) <= 1e-5 * np.abs(RLL) + 1e-4)
    assert np.abs(post - g).max() <= 2e-4
    stats = np.array([[float(v) for v in ln.split()[1:]] for ln in lines[1 + N:1 + N + M]])
    lists = [np.nonzero(row > np.float32(1e-4))[0] for row in post.astype(np.float32)]
    _, rn, rF, rS, rtot = orc.accumulate_masked(x, w, mu, var, lists)
    check_stats((stats[:, 0], stats[:, 1:], rS), (rn, rF, rS), x)
