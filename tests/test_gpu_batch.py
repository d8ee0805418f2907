"""Batched per-utterance statistics, MAP means and MLLR transforms
(gmmhip_batch_adapt) against the oracle applied utterance by utterance.

Tolerances: statistics as the E-step (1e-5 relative + the per-component
allowance for posteriors within 1 % of the 1e-4 cut, tests/_check.py); MAP
means 1e-5 relative; MLLR transforms: exact FP64 algebra on the GPU's own
statistics (1e-8), and 1e-3 against the oracle's FP64 statistics (the
transform amplifies the FP32 statistics' 1e-5 by the conditioning of G).
The work-queue launch, the per-utterance launches and hipGraph replays give
bit-identical results.
"""
import numpy as np
import pytest

import _check
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


def make_batch(seed, U, M, D, lo=30, hi=400, extra=()):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M))
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2.0, (M, D)).astype(np.float32)
    lens = list(rng.integers(lo, hi, U - len(extra))) + list(extra)
    xs = [synth.sample_frames_np(w, mu + rng.normal(0, 0.3, (M, D)).astype(np.float32), var,
                                 int(n), seed + 10 + u) if n > 0 else np.zeros((0, D), np.float32)
          for u, n in enumerate(lens)]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    x = np.concatenate(xs).astype(np.float32)
    return x, off, w.astype(np.float32), mu, var


def per_utt_oracle(x, off, w, mu, var):
    out = []
    for u in range(len(off) - 1):
        xu = x[int(off[u]):int(off[u + 1])]
        if len(xu) == 0:
            M, D = mu.shape
            out.append((np.zeros(M), np.zeros((M, D))))
        else:
            n, F, S, _ = orc.accumulate(xu, w, mu, var)
            out.append((n, F))
    return out


def test_batch_stats_map_mllr(acc):
    x, off, w, mu, var = make_batch(1, 24, 48, 6, extra=(0, 4))
    acc.set_model(w, mu, var)
    res = acc.batch_adapt(x, off, tau=16, prior=False)
    ref = per_utt_oracle(x, off, w, mu, var)
    xm = float(np.abs(x).max())
    for u, (rn, rF) in enumerate(ref):
        n, F = res["occ"][u], res["first"][u]
        xu = x[int(off[u]):int(off[u + 1])]
        _check.check_stats((n, F, None, None), (rn, rF, None, None), xu, w, mu, var,
                           band=_check.BAND["fp32"], what="utt %d " % u)
        # MAP means from the GPU's statistics and from the oracle's
        _, rmu, _ = orc.update_map(w, mu, var, rn, rF, rF ** 2, 1, 16)
        assert np.allclose(res["means"][u], rmu, rtol=1e-5, atol=1e-5 * xm), u
    # MLLR: under-determined utterances are flagged
    lens = np.diff(off.astype(np.int64))
    for u in range(len(lens)):
        n = res["occ"][u]
        expect_ok = orc.mllr_check(len(w), mu.shape[1], n) == 0
        assert (res["status"][u] == 0) == expect_ok, (u, lens[u])
        if expect_ok:
            Wg = orc.mllr_transform(mu, var, n, res["first"][u])
            assert np.allclose(res["W"][u], Wg, rtol=1e-8, atol=1e-8), u
            Wo = orc.mllr_transform(mu, var, ref[u][0], ref[u][1])
            assert np.allclose(res["W"][u], Wo, rtol=1e-3, atol=1e-3), u
        else:
            assert np.all(res["W"][u] == 0)


def test_batch_mllr_prior(acc):
    x, off, w, mu, var = make_batch(2, 10, 40, 5, lo=40, hi=200)
    acc.set_model(w, mu, var)
    res = acc.batch_adapt(x, off, tau=16, prior=True, means=False)
    for u in range(10):
        assert res["status"][u] == 0
        Wg = orc.mllr_transform(mu, var, res["occ"][u], res["first"][u], ubm_mu=mu, ubm_var=var)
        assert np.allclose(res["W"][u], Wg, rtol=1e-8, atol=1e-8), u


def test_batch_modes_and_graph_bit_identical(acc):
    x, off, w, mu, var = make_batch(3, 40, 100, 12, lo=20, hi=700)
    acc.set_model(w, mu, var)
    a = acc.batch_adapt(x, off, mode=gmm_amd.BATCH_QUEUE)
    b = acc.batch_adapt(x, off, mode=gmm_amd.BATCH_PER_UTTERANCE)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # graph: capture once, replay several times with eager work in between
    for rep in range(3):
        g = acc.batch_adapt(x, off, graph=True)
        for k in a:
            assert np.array_equal(a[k], g[k]), (rep, k)
        e = acc.batch_adapt(x, off, graph=False)
        for k in a:
            assert np.array_equal(a[k], e[k]), (rep, k)
    # a different batch shape rebuilds the graph
    x2, off2, _, _, _ = make_batch(4, 17, 100, 12)
    g2 = acc.batch_adapt(x2, off2, graph=True)
    e2 = acc.batch_adapt(x2, off2, graph=False)
    for k in g2:
        assert np.array_equal(g2[k], e2[k]), k


def test_batch_chunk_size_invariance(acc):
    x, off, w, mu, var = make_batch(5, 16, 64, 8, lo=100, hi=3000)
    acc.set_model(w, mu, var)
    a = acc.batch_adapt(x, off, chunk_frames=256, mllr=False)
    b = acc.batch_adapt(x, off, chunk_frames=4096, mllr=False)
    assert np.allclose(a["occ"], b["occ"], rtol=2e-5, atol=1e-4)
    assert np.allclose(a["means"], b["means"], rtol=2e-5, atol=1e-4)


def test_batch_matches_single_accumulate(acc):
    """One utterance through the batch path equals the plain accumulate."""
    x, off, w, mu, var = make_batch(6, 1, 64, 10, lo=5000, hi=5001)
    acc.set_model(w, mu, var)
    res = acc.batch_adapt(x, off, mllr=False, chunk_frames=8192)
    acc.set_opts(chunk_frames=8192)
    acc.set_frames(x)
    n, F, S, tot, _ = acc.accumulate()
    acc.set_opts(chunk_frames=0)
    assert np.allclose(res["occ"][0], n, rtol=1e-12, atol=1e-12)
    assert np.allclose(res["first"][0], F, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("D,M", [(70, 160), (20, 600)])
def test_batch_mllr_wide_and_many_components(acc, D, M):
    """D = 70: two 64-row tiles in the G/k kernel and the E > 64 solve;
    M = 600: occupied-component lists of several hundred entries."""
    x, off, w, mu, var = make_batch(7 + D, 5, M, D, lo=1500, hi=2500)
    acc.set_model(w, mu, var)
    res = acc.batch_adapt(x, off, tau=16, prior=True, means=False)
    ref = per_utt_oracle(x, off, w, mu, var)
    for u in range(len(off) - 1):
        assert res["status"][u] == 0, u
        Wg = orc.mllr_transform(mu, var, res["occ"][u], res["first"][u], ubm_mu=mu, ubm_var=var)
        assert np.allclose(res["W"][u], Wg, rtol=1e-8, atol=1e-8), u
        Wo = orc.mllr_transform(mu, var, ref[u][0], ref[u][1], ubm_mu=mu, ubm_var=var)
        assert np.allclose(res["W"][u], Wo, rtol=1e-3, atol=1e-3), u


@pytest.mark.parametrize("prec,tol", [(gmm_amd.PREC_FP16X3, 2e-4), (gmm_amd.PREC_BF16X3, 2e-3)])
@pytest.mark.parametrize("D", [12, 30, 45, 120])
def test_batch_mixed_precision(acc, prec, tol, D):
    """The mixed form in the batched path: per-utterance statistics, MAP
    means and MLLR transforms against the oracle (mixed-precision
    tolerances, as the E-step's), graph replays bit-identical to eager runs,
    and each utterance's result independent of its batch-mates (UBM-based
    centring and scaling, chunks relative to the utterance start).  Feature
    rows of 32, 64, 128 and 256 columns (contraction depths 32, 64, 96 and
    256), every one through the persistent work-queue form of the statistics
    kernel."""
    x, off, w, mu, var = make_batch(11, 30, 64, D, lo=20, hi=600, extra=(0, 5))
    acc.set_model(w, mu, var)
    ref = per_utt_oracle(x, off, w, mu, var)
    acc.set_opts(precision=prec)
    try:
        res = acc.batch_adapt(x, off, tau=16, prior=True)
        graphs = [acc.batch_adapt(x, off, tau=16, prior=True, graph=True) for _ in range(3)]
        # a sub-batch: utterances 3..9 alone
        sub = off[3:11] - off[3]
        res_sub = acc.batch_adapt(x[int(off[3]):int(off[10])], sub, tau=16, prior=True)
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32)
    xm = float(np.abs(x).max())
    band = _check.BAND["fp16x3" if prec == gmm_amd.PREC_FP16X3 else "bf16x3"]
    for u, (rn, rF) in enumerate(ref):
        n, F = res["occ"][u], res["first"][u]
        xu = x[int(off[u]):int(off[u + 1])]
        _check.check_stats((n, F, None, None), (rn, rF, None, None), xu, w, mu, var, rtol=tol,
                           band=band, what="utt %d " % u)
        _, rmu, _ = orc.update_map(w, mu, var, n, F, F ** 2, 1, 16)
        assert np.allclose(res["means"][u], rmu, rtol=1e-5, atol=1e-5 * xm), u
        if res["status"][u] == 0:
            Wg = orc.mllr_transform(mu, var, n, F, ubm_mu=mu, ubm_var=var)
            assert np.allclose(res["W"][u], Wg, rtol=1e-8, atol=1e-8), u
    for g in graphs:
        for k in res:
            assert np.array_equal(res[k], g[k]), k
    for k in res:
        assert np.array_equal(res_sub[k], res[k][3:10]), k


def test_batch_mixed_out_of_range_runs_fp32(acc):
    """A frame value outside the FP16 range after centring and scaling sends
    the whole batch to the FP32 kernels: the results equal the FP32 run."""
    x, off, w, mu, var = make_batch(12, 8, 32, 6, lo=50, hi=300)
    x[int(off[2]) + 3, 1] = 5e4  # |z| >> 240
    acc.set_model(w, mu, var)
    fp32 = acc.batch_adapt(x, off, tau=16)
    acc.set_opts(precision=gmm_amd.PREC_FP16X3)
    try:
        mixed = acc.batch_adapt(x, off, tau=16)
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32)
    for k in fp32:
        assert np.array_equal(fp32[k], mixed[k]), k


def test_batch_graph_recaptured_after_reallocation(acc):
    """A cached batch graph holds raw device addresses: after other calls on
    the same handle grow handle-wide buffers (per-frame log-likelihoods of a
    larger frame set whose N is not a multiple of 128, the auxiliary
    statistics' per-block sums), the next graph call must capture anew, and
    every result must equal the eager run."""
    x, off, w, mu, var = make_batch(21, 20, 64, 10, lo=20, hi=300)
    acc.set_model(w, mu, var)
    eager = acc.batch_adapt(x, off, graph=False)
    c0 = acc.graph_captures()
    g1 = acc.batch_adapt(x, off, graph=True)
    assert acc.graph_captures() == c0 + 1
    g1b = acc.batch_adapt(x, off, graph=True)  # same shape and buffers: replay
    assert acc.graph_captures() == c0 + 1
    y = synth.sample_frames_np(w.astype(np.float64), mu, var, 3 * int(off[-1]) + 77, 22)
    assert y.shape[0] % 128 != 0
    acc.set_frames(y)
    acc.loglikes()
    acc.accumulate_aux()
    g2 = acc.batch_adapt(x, off, graph=True)
    assert acc.graph_captures() == c0 + 2
    g2b = acc.batch_adapt(x, off, graph=True)
    assert acc.graph_captures() == c0 + 2
    for k in eager:
        for g in (g1, g1b, g2, g2b):
            assert np.array_equal(eager[k], g[k]), k


def test_batch_mixed_model_overflow_runs_fp32(acc):
    """A UBM coefficient beyond the FP16 range (a variance of 2e-6) is flagged
    by the split-model packing on the GPU and sends the fp16x3 batch to the
    FP32 kernels: the results equal the FP32 run bit for bit."""
    x, off, w, mu, var = make_batch(14, 8, 32, 6, lo=50, hi=300)
    var = var.copy()
    var[7, 2] = 2e-6  # just above the 1e-6 variance floor
    acc.set_model(w, mu, var)
    fp32 = acc.batch_adapt(x, off, tau=16)
    acc.set_opts(precision=gmm_amd.PREC_FP16X3)
    try:
        mixed = acc.batch_adapt(x, off, tau=16)
        tiles = acc.tiles_used()
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32)
    assert tiles[2] == 0, tiles
    for k in fp32:
        assert np.array_equal(fp32[k], mixed[k]), k


@pytest.mark.parametrize("prec", [gmm_amd.PREC_FP32, gmm_amd.PREC_FP16X3])
def test_batch_device_frames_any_alignment(acc, prec):
    """Device-resident frames at a 16-byte aligned address and at one float
    past it: the row padding copies 16-byte pieces in the first case and
    single floats in the second, and every output is the same bit for bit."""
    import torch
    x, off, w, mu, var = make_batch(31, 24, 64, 12, lo=20, hi=500, extra=(0, 7))
    N, D = x.shape
    M = mu.shape[0]
    U = len(off) - 1
    acc.set_model(w, mu, var)
    buf = torch.zeros(N * D + 8, dtype=torch.float32, device="cuda")
    base = buf.data_ptr()
    assert base % 16 == 0
    xt = torch.from_numpy(x.reshape(-1))

    def run(shift):
        buf.zero_()
        buf[shift:shift + N * D] = xt.to("cuda")
        o = {"means": torch.empty((U, M, D), dtype=torch.float32, device="cuda"),
             "occ": torch.empty((U, M), dtype=torch.float64, device="cuda"),
             "first": torch.empty((U, M, D), dtype=torch.float64, device="cuda"),
             "W": torch.empty((U, D, D + 1), dtype=torch.float64, device="cuda"),
             "status": torch.empty(U, dtype=torch.int32, device="cuda")}
        acc.batch_adapt_device(base + 4 * shift, off, D, {k: v.data_ptr() for k, v in o.items()},
                               tau=16, prior=True)
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in o.items()}

    acc.set_opts(precision=prec)
    try:
        aligned, shifted = run(0), run(1)
        host = acc.batch_adapt(x, off, tau=16, prior=True)
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32)
    for k in aligned:
        assert np.array_equal(aligned[k], shifted[k]), k
        assert np.array_equal(aligned[k], host[k]), k
