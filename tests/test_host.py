"""CPU tests of the host layer (include/gmmhost.h): the `.gmm` reader and
writers against an independent parser and the reference's own model file,
RearrangeMixtures, and the EM / MAP / MLLR update rules against the oracle."""
import os

import numpy as np
import pytest

import _formats as fm
import _oracle as orc
from gmm_amd import host

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def test_read_reference_model_txt(golden):
    m = host.Model.load(os.path.join(GOLD, "myModel.gmm"), True)
    w, mu, var = m.get()
    assert np.array_equal(w, golden["w"])
    assert np.array_equal(mu, golden["mu"])
    assert np.array_equal(var, golden["var"])


@pytest.mark.parametrize("txt", [True, False])
def test_round_trip(tmp_path, golden, txt):
    m = host.Model.load(os.path.join(GOLD, "myModel.gmm"), True)
    p = tmp_path / ("m.gmm" if txt else "m.bin")
    m.save(p, txt)
    r = host.Model.load(p, txt)
    a, b = m.get(), r.get()
    for u, v in zip(a, b):
        if txt:  # %15e keeps 7 significant digits
            assert np.allclose(u, v, rtol=6e-7, atol=0)
        else:
            assert np.array_equal(u, v)
    if txt:
        w2, mu2, var2, _ = fm.read_gmm_txt(p)
        assert np.allclose(w2, golden["w"], rtol=6e-7)


def test_bin_layout(tmp_path):
    w = np.array([0.5, 0.0, 0.5], np.float32)
    mu = np.arange(6, dtype=np.float32).reshape(3, 2)
    var = np.ones((3, 2), np.float32) * 2
    p = tmp_path / "m.bin"
    host.Model.create(w, mu, var).save(p, False)
    raw = p.read_bytes()
    assert raw[:9] == b"GMM-DB-VN"
    assert np.frombuffer(raw[9:11], "<i2")[0] == 2
    assert np.frombuffer(raw[19:21], "<i2")[0] == 2  # zero-weight component skipped
    body = np.frombuffer(raw[21:], "<f4")
    assert np.array_equal(body[:2], [0.5, 0.5])
    assert np.array_equal(body[2:6], [0, 1, 2, 2])
    assert np.array_equal(body[6:10], [4, 5, 2, 2])


def test_rearrange_matches_oracle():
    w = np.array([0.2, 0, 0.3, 0, 0, 0.5], np.float32)
    mu = np.arange(12, dtype=np.float32).reshape(6, 2)
    m = host.Model.create(w, mu, np.ones((6, 2), np.float32))
    assert m.rearrange() == 3
    perm, _ = orc.rearrange_perm(w)
    w2, mu2, _ = m.get()
    assert np.array_equal(w2, w[perm])
    assert np.array_equal(mu2, mu[perm])


def _stats(seed, M, D):
    rng = np.random.default_rng(seed)
    n = rng.uniform(0, 50, M)
    n[1] = 1.5
    F = rng.normal(0, 2, (M, D)) * n[:, None]
    S = (F / n[:, None]) ** 2 * n[:, None] + rng.uniform(0.5, 2, (M, D)) * n[:, None]
    return n, F, S


def test_update_em_matches_oracle():
    rng = np.random.default_rng(1)
    M, D = 9, 5
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    w[4] = 0.0
    mu = rng.normal(0, 1, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2, (M, D)).astype(np.float32)
    n, F, S = _stats(2, M, D)
    n[4] = 0
    m = host.Model.create(w, mu, var)
    left = m.update_em(n, F, S, 300)
    rw, rmu, rvar, rleft = orc.update_em(w, mu, var, n, F, S, 300)
    hw, hmu, hvar = m.get()
    assert left == rleft
    assert np.array_equal(hw, rw) and np.array_equal(hmu, rmu) and np.array_equal(hvar, rvar)


@pytest.mark.parametrize("flags", [(0, 1, 0), (1, 1, 1), (1, 0, 0)])
def test_update_map_matches_oracle(flags):
    rng = np.random.default_rng(3)
    M, D = 7, 4
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 1, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2, (M, D)).astype(np.float32)
    n, F, S = _stats(4, M, D)
    n[2] = 0.0
    ubm = host.Model.create(w, mu, var)
    m = host.Model.create(w, mu, var)
    m.update_map(ubm, n, F, S, 250, 16, upweight=flags[0], upmean=flags[1], upvar=flags[2])
    rw, rmu, rvar = orc.update_map(w, mu, var, n, F, S, 250, 16, upweight=flags[0],
                                   upmean=flags[1], upvar=flags[2])
    hw, hmu, hvar = m.get()
    assert np.array_equal(hw, rw) and np.array_equal(hmu, rmu) and np.array_equal(hvar, rvar)


@pytest.mark.parametrize("prior", [False, True])
def test_mllr_transform_matches_oracle(prior):
    rng = np.random.default_rng(5)
    M, D = 40, 6
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2, (M, D)).astype(np.float32)
    A = np.eye(D) + rng.normal(0, 0.1, (D, D))
    b = rng.normal(0, 0.5, D)
    n = rng.uniform(2, 20, M)
    F = n[:, None] * (mu.astype(np.float64) @ A.T + b) + rng.normal(0, 0.1, (M, D))
    m = host.Model.create(w, mu, var)
    ubm = m if prior else None
    W = m.mllr_transform(n, F, ubm=ubm)
    R = orc.mllr_transform(mu, var, n, F, ubm_mu=mu if prior else None,
                           ubm_var=var if prior else None)
    assert np.allclose(W, R, rtol=1e-9, atol=1e-9)
    m.mllr_apply(W)
    assert np.array_equal(m.get()[1], orc.mllr_apply(W, mu))


def test_mllr_underdetermined_rejected():
    M, D = 5, 6
    m = host.Model.create(np.full(M, 0.2, np.float32), np.zeros((M, D), np.float32),
                          np.ones((M, D), np.float32))
    with pytest.raises(host.HostError):
        m.mllr_transform(np.ones(M), np.zeros((M, D)))


def test_missing_file_reports_error(tmp_path):
    with pytest.raises(host.HostError):
        host.Model.load(tmp_path / "nope.gmm", True)


# ---------------------------------------------------------------- full covariance
def test_read_reference_fullcov_model(tmp_path):
    """example/myModelFC.gmm through the library's reader equals the
    independent parser; TXT and BIN (negative dimension) round trips."""
    path = os.path.join(GOLD, "myModelFC.gmm")
    m = host.Model.load(path, True)
    assert m.full()
    w, mu, cov = m.get()
    rw, rmu, rcov, _ = fm.read_gmm_txt_full(path)
    assert np.array_equal(w, rw) and np.array_equal(mu, rmu) and np.array_equal(cov, rcov)
    for txt in (True, False):
        p = tmp_path / ("fc.gmm" if txt else "fc.bin")
        m.save(p, txt)
        r = host.Model.load(p, txt)
        assert r.full()
        for u, v in zip((w, mu, cov), r.get()):
            if txt:
                assert np.allclose(u, v, rtol=6e-7, atol=0)
            else:
                assert np.array_equal(u, v)
    with open(tmp_path / "fc.bin", "rb") as f:
        hdr = f.read(11)
    assert hdr[:9] == b"GMM-DB-VN" and np.frombuffer(hdr[9:11], "<i2")[0] == -20
    w2, mu2, cov2, _ = fm.read_gmm_txt_full(tmp_path / "fc.gmm")
    assert np.allclose(cov2, rcov, rtol=6e-7, atol=0)


def test_update_em_full_matches_oracle():
    rng = np.random.default_rng(17)
    M, D, N = 3, 4, 400
    w = np.array([0.3, 0.3, 0.4], np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    A = rng.normal(size=(M, D, D))
    cov = (A @ A.transpose(0, 2, 1) / D + np.eye(D)).astype(np.float32)
    x = rng.normal(0, 2, (N, D)).astype(np.float32)
    n, F, S, _ = orc.accumulate_full(x, w, mu, cov)
    m = host.Model.create_full(w, mu, cov)
    left = m.update_em(n, F, S, N)
    gw, gmu, gcov = m.get()
    ow, omu, ocov, oleft = orc.update_em_full(w, mu, cov, n, F, S, N)
    assert left == oleft
    assert np.array_equal(gw, ow) and np.array_equal(gmu, omu)
    assert np.array_equal(gcov, ocov)


@pytest.mark.parametrize("flags", [(True, True, True), (False, True, False), (False, False, True)])
def test_update_map_full_matches_oracle(flags):
    """UpdateModelMAP's full-covariance branch (OL_ModelAdapt.cpp:1079-1085)
    against the oracle, bit for bit."""
    rng = np.random.default_rng(19)
    M, D, N = 3, 5, 400
    w = np.array([0.3, 0.3, 0.4], np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    A = rng.normal(size=(M, D, D))
    cov = (A @ A.transpose(0, 2, 1) / D + np.eye(D)).astype(np.float32)
    cov = 0.5 * (cov + cov.transpose(0, 2, 1))
    x = (rng.normal(0, 2, (N, D)) + 0.3).astype(np.float32)
    n, F, S, _ = orc.accumulate_full(x, w, mu, cov)
    ubm = host.Model.create_full(w, mu, cov)
    m = host.Model.create_full(w, mu, cov)
    m.update_map(ubm, n, F, S, N, 16, upweight=flags[0], upmean=flags[1], upvar=flags[2])
    hw, hmu, hcov = m.get()
    rw, rmu, rcov = orc.update_map_full(w, mu, cov, n, F, S, N, 16, upweight=flags[0],
                                        upmean=flags[1], upvar=flags[2])
    assert np.array_equal(hw, rw) and np.array_equal(hmu, rmu) and np.array_equal(hcov, rcov)


def test_mllr_transform_full_model_uses_diagonal():
    """A full-covariance model enters the MLLR statistics through the
    diagonal of its covariances (GetMixtureVarDiag, OL_ModelAdapt.cpp:1502,
    :1561): same transform as the diagonal model of those variances."""
    rng = np.random.default_rng(6)
    M, D = 30, 5
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    A = rng.normal(size=(M, D, D))
    cov = (A @ A.transpose(0, 2, 1) / D + np.eye(D)).astype(np.float32)
    var = np.ascontiguousarray(cov[:, np.arange(D), np.arange(D)])
    n = rng.uniform(2, 20, M)
    F = n[:, None] * (mu.astype(np.float64) + 0.3) + rng.normal(0, 0.1, (M, D))
    mf = host.Model.create_full(w, mu, cov)
    md = host.Model.create(w, mu, var)
    Wf = mf.mllr_transform(n, F, ubm=mf)
    Wd = md.mllr_transform(n, F, ubm=md)
    assert np.array_equal(Wf, Wd)
    R = orc.mllr_transform(mu, var, n, F, ubm_mu=mu, ubm_var=var)
    assert np.allclose(Wf, R, rtol=1e-9, atol=1e-9)


def _robust_case(seed, M=3, D=4, N=3000, full=False):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 1.0, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2.0, (M, D)).astype(np.float32)
    x = (np.cumsum(rng.normal(0, 0.4, (N, D)), axis=0) * 0.2 + rng.normal(0, 1, (N, D))).astype(np.float32)
    if full:
        A = rng.normal(size=(M, D, D))
        cov = (A @ A.transpose(0, 2, 1) / D + np.eye(D)).astype(np.float32)
        cov = 0.5 * (cov + cov.transpose(0, 2, 1))
        n, F, S, _ = orc.accumulate_full(x, w, mu, cov)
        _, g = orc.frame_loglik_full(x, w, mu, cov, want_gamma=True)
        return x, w, mu, cov, n, F, S, orc.accumulate_aux(x, g)
    n, F, S, _ = orc.accumulate(x, w, mu, var)
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    return x, w, mu, var, n, F, S, orc.accumulate_aux(x, g)


@pytest.mark.parametrize("full", [False, True])
def test_make_robust_stats_matches_oracle(full):
    x, w, mu, v, n, F, S, aux = _robust_case(23, full=full)
    hS, hn = host.make_robust_stats(n, F, S, aux, full=full)
    rS, rn = orc.make_robust_stats(n, F, S, aux, full=full)
    assert np.array_equal(hn, rn)
    assert np.array_equal(hS, rS)


@pytest.mark.parametrize("full", [False, True])
def test_update_em_robust_matches_oracle(full):
    x, w, mu, v, n, F, S, aux = _robust_case(24, full=full)
    m = host.Model.create_full(w, mu, v) if full else host.Model.create(w, mu, v)
    left = m.update_em_robust(n, F, S, aux, x.shape[0])
    ow, omu, ov, oleft = orc.update_em_robust(w, mu, v, n, F, S, aux, x.shape[0], full=full)
    gw, gmu, gv = m.get()
    assert left == oleft
    assert np.array_equal(gw, ow) and np.array_equal(gmu, omu) and np.array_equal(gv, ov)


@pytest.mark.parametrize("iters", [1, 20])
def test_fmllr_transform_matches_oracle(iters):
    rng = np.random.default_rng(29)
    M, D, N = 32, 5, 6000
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 1.5, (M, D)).astype(np.float32)
    p = w.astype(np.float64)
    k = rng.choice(M, size=N, p=p / p.sum())
    A0 = np.eye(D) + rng.normal(0, 0.1, (D, D))
    y = mu[k] + np.sqrt(var[k]) * rng.normal(size=(N, D))
    x = ((y - 0.2) @ np.linalg.inv(A0).T).astype(np.float32)
    cov = np.zeros((M, D, D), np.float32)
    cov[:, np.arange(D), np.arange(D)] = var
    n, F, S, _ = orc.accumulate_full(x, w, mu, cov)
    m = host.Model.create(w, mu, var)
    W = m.fmllr_transform(n, F, S, N, iters=iters)
    R = orc.fmllr(mu, var, n, F, S, N, iters=iters)
    assert np.allclose(W, R, rtol=1e-9, atol=1e-9)
    # the same transform from a full-covariance model with those variances
    mf = host.Model.create_full(w, mu, cov)
    assert np.allclose(mf.fmllr_transform(n, F, S, N, iters=iters), W, rtol=1e-12, atol=1e-12)
