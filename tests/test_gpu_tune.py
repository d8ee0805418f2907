"""Tile choices (log-likelihood frame tile, statistics work-item size) do
not change results: per-frame log-likelihoods are bit-identical for every
frame tile (each frame's log-sum-exp runs over the same component order);
with the work-item size fixed the statistics are bit-identical too and the
total log-likelihood differs only in the FP64 grouping of block sums
(1e-12 relative).  The tuning table's entry is what a handle uses when the
options leave the tiles to it."""
import numpy as np
import pytest

import _oracle as orc
import gmm_amd
from gmm_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def acc():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    a = gmm_amd.Accumulator(0)
    yield a
    a.close()


@pytest.mark.parametrize("D,M,N", [(20, 100, 5000), (60, 300, 20000), (39, 64, 1000)])
def test_ll_frame_tiles_identical(acc, D, M, N):
    rng = np.random.default_rng(D + M)
    w = rng.dirichlet(np.ones(M))
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2, (M, D)).astype(np.float32)
    x = synth.sample_frames_np(w, mu, var, N, 3)
    acc.set_model(w.astype(np.float32), mu, var)
    acc.set_frames(x)
    res = {}
    for ll in (64, 128, 256):
        acc.set_opts(ll_frames=ll, chunk_frames=1024)
        lls = acc.loglikes()
        st = acc.accumulate()
        assert acc.tiles_used() == (ll, 1024, 0)
        res[ll] = (lls, st)
    acc.set_opts(ll_frames=0, chunk_frames=0)
    for ll in (64, 256):
        assert np.array_equal(res[ll][0], res[128][0])
        for a, b in zip(res[ll][1][:3], res[128][1][:3]):
            assert np.array_equal(a, b)
        assert res[ll][1][3] == pytest.approx(res[128][1][3], rel=1e-12)
    ref = orc.frame_loglik(x, w.astype(np.float32), mu, var)
    assert np.allclose(res[64][0], ref, rtol=1e-5, atol=1e-5)


def test_table_entry_is_used(acc, tmp_path):
    D, M, N = 24, 80, 30000
    p = tmp_path / "t.tune"
    p.write_text("estep fp32 %d %d 256 2048 1.0\nbatch fp32 %d %d 64 256 1.0\n" % (D, M, D, M))
    rng = np.random.default_rng(1)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2, (M, D)).astype(np.float32)
    x = synth.sample_frames_np(w.astype(np.float64), mu, var, N, 4)
    try:
        assert gmm_amd.tune_load(str(p)) == 2
        acc.set_opts(ll_frames=0, chunk_frames=0)
        acc.set_model(w, mu, var)
        acc.set_frames(x)
        acc.accumulate()
        # the table's value bounds the work item; N/64 groups-worth of
        # chunks would be smaller, so the bound binds only for large N
        ll, cf, ct = acc.tiles_used()
        assert ll == 256 and cf <= 2048 and ct == 0
        off = np.array([0, 1000, 5000, N], np.uint64)
        acc.batch_adapt(x, off, mllr=False)
        assert acc.tiles_used() == (64, 256, 0)
    finally:
        gmm_amd.tune_load(None)


@pytest.mark.parametrize("prec", [gmm_amd.PREC_FP16X3, gmm_amd.PREC_BF16X3])
@pytest.mark.parametrize("D,M,N", [(60, 300, 20000 + 77), (20, 70, 3001)])
def test_mixed_frame_tiles_identical(acc, prec, D, M, N):
    """The mixed form's log-likelihood frame tile (128 or 256 frames per
    block; 256 only where the depth is a multiple of 64, else the handle
    keeps 128): per-frame log-likelihoods and statistics bit-identical, the
    total to the FP64 grouping of block sums; the batched path the same."""
    rng = np.random.default_rng(D + M + prec)
    w = rng.dirichlet(np.ones(M))
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2, (M, D)).astype(np.float32)
    x = synth.sample_frames_np(w, mu, var, N, 5)
    acc.set_model(w.astype(np.float32), mu, var)
    acc.set_frames(x)
    kpm = (2 * D + 1 + 31) // 32 * 32
    res = {}
    try:
        for ll in (128, 256):
            acc.set_opts(precision=prec, ll_frames=ll, chunk_frames=1024, comp_tile=128)
            lls = acc.loglikes()
            st = acc.accumulate()
            used = acc.tiles_used()
            assert used == ((ll if kpm % 64 == 0 else 128), 1024, 128), used
            off = np.array([0, 500, 3000, N], np.uint64)
            bt = acc.batch_adapt(x, off, mllr=False)
            res[ll] = (lls, st, bt)
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32, ll_frames=0, chunk_frames=0, comp_tile=0)
    assert np.array_equal(res[128][0], res[256][0])
    for a, b in zip(res[128][1][:3], res[256][1][:3]):
        assert np.array_equal(a, b)
    assert res[128][1][3] == pytest.approx(res[256][1][3], rel=1e-12)
    for k in res[128][2]:
        assert np.array_equal(res[128][2][k], res[256][2][k]), k


@pytest.mark.parametrize("prec", [gmm_amd.PREC_FP16X3, gmm_amd.PREC_BF16X3])
@pytest.mark.parametrize("D,M,N", [(60, 300, 20000 + 77), (20, 70, 3001), (39, 64, 1000),
                                   (60, 2048, 4096 + 5)])
def test_mixed_comp_tiles_identical(acc, prec, D, M, N):
    """The mixed statistics component tile (128 or 256 components per block;
    256 only for feature rows of 64 and 128 columns, else the handle keeps
    128): every
    component's statistics come from the same wave-level contraction over
    the same frame steps, so n, F, S, the per-frame log-likelihoods and the
    batched path's outputs are bit-identical.  M = 300 and 70 leave waves of
    the last 256-component group without a component tile."""
    rng = np.random.default_rng(D + 7 * M + prec)
    w = rng.dirichlet(np.ones(M))
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2, (M, D)).astype(np.float32)
    x = synth.sample_frames_np(w, mu, var, N, 6)
    acc.set_model(w.astype(np.float32), mu, var)
    acc.set_frames(x)
    kpc = (2 * D + 1 + 31) // 32 * 32
    kpl = 32 if kpc <= 32 else 64 if kpc <= 64 else 128 if kpc <= 128 else 256  # row layout
    res = {}
    try:
        for ct in (128, 256):
            acc.set_opts(precision=prec, ll_frames=128, chunk_frames=1024, comp_tile=ct)
            st = acc.accumulate()
            used = acc.tiles_used()
            assert used == (128, 1024, ct if kpl in (64, 128) else 128), used
            off = np.array([0, 300, N // 2, N], np.uint64)
            bt = acc.batch_adapt(x, off, mllr=False)
            assert acc.tiles_used()[2] == (ct if kpl in (64, 128) else 128)
            res[ct] = (st, bt)
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32, ll_frames=0, chunk_frames=0, comp_tile=0)
    for a, b in zip(res[128][0], res[256][0]):
        assert np.array_equal(a, b)
    for k in res[128][1]:
        assert np.array_equal(res[128][1][k], res[256][1][k]), k
    # and the statistics are the oracle's, to the form's stated total
    # log-likelihood tolerance (fp16x3 1e-6, bf16x3 2e-5 relative)
    n_ref, F_ref, _, tot_ref = orc.accumulate(x, w.astype(np.float32), mu, var)
    tot_rel = 1e-6 if prec == gmm_amd.PREC_FP16X3 else 2e-5
    assert res[256][0][3] == pytest.approx(tot_ref, rel=tot_rel)
    assert np.allclose(res[256][0][0], n_ref, rtol=2e-4, atol=2e-3)
