"""Serving loop (gmmhip_server_*): continuous batching of score / MAP / MLLR
requests over a resident UBM.

Stated tolerance: none -- a request's results (log-likelihood sum, per-frame
log-likelihoods, occupancies, MAP means, MLLR transform, status) are
bit-identical whether it runs alone or mixed into batches with others, and
bit-identical to gmmhip_batch_adapt on the same utterance with the same
chunk size (whose agreement with the oracle test_gpu_batch.py checks).  The
per-frame log-likelihoods are also checked against the oracle directly
(1e-5 relative, the plain FP32 bar).
"""
import threading

import numpy as np
import pytest

import _oracle as orc
import gmm_amd
from gmm_amd import REQ_MAP, REQ_MLLR, REQ_SCORE, synth

pytestmark = pytest.mark.gpu


def make_ubm(seed, M, D):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2.0, (M, D)).astype(np.float32)
    return w, mu, var


def make_requests(seed, w, mu, var, n, lo=20, hi=1500):
    rng = np.random.default_rng(seed)
    M, D = mu.shape
    kinds = rng.choice([REQ_SCORE, REQ_MAP, REQ_MLLR], size=n, p=[0.5, 0.3, 0.2])
    lens = rng.integers(lo, hi, n)
    out = []
    for i in range(n):
        shift = rng.normal(0, 0.3, (M, D)).astype(np.float32)
        x = synth.sample_frames_np(w.astype(np.float64), mu + shift, var, int(lens[i]),
                                   seed * 1000 + i)
        out.append((int(kinds[i]), x))
    return out


def results(p):
    return (p.status, p.ll_sum, None if p.frame_ll is None else p.frame_ll.copy(),
            None if p.means is None else p.means.copy(),
            None if p.W is None else p.W.copy(), None if p.occ is None else p.occ.copy())


def same(a, b):
    assert a[0] == b[0]
    assert a[1] == b[1]
    for x, y in zip(a[2:], b[2:]):
        assert (x is None) == (y is None)
        if x is not None:
            assert np.array_equal(x, y)


@pytest.fixture(scope="module")
def ubm():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    return make_ubm(5, 96, 12)


def test_server_alone_vs_mix_bit_identical(ubm):
    w, mu, var = ubm
    reqs = make_requests(11, w, mu, var, 48)
    srv = gmm_amd.Server(w, mu, var, slots=2, tau=16, chunk_frames=256)
    try:
        alone = []
        for kind, x in reqs:
            p = srv.submit(kind, x, frame_ll=True, occ=True)
            assert srv.wait(p, 60000) in (0, gmm_amd.ERR_MLLR)
            alone.append(results(p))
        st = srv.stats(reset=True)
        assert st["batches"] == len(reqs) and st["completed"] == len(reqs)
        # all at once: batched together, several batches in flight
        ps = [srv.submit(kind, x, frame_ll=True, occ=True) for kind, x in reqs]
        srv.drain()
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["batches"] < len(reqs)
        assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
        for a, p in zip(alone, ps):
            same(a, results(p))
    finally:
        srv.close()


def test_server_matches_batch_adapt_and_oracle(ubm):
    w, mu, var = ubm
    M, D = mu.shape
    reqs = make_requests(12, w, mu, var, 12, lo=30, hi=900)
    srv = gmm_amd.Server(w, mu, var, tau=16, chunk_frames=256, mllr_prior=True)
    acc = gmm_amd.Accumulator(0)
    try:
        ps = [srv.submit(kind, x, frame_ll=True, occ=True) for kind, x in reqs]
        srv.drain()
        acc.set_model(w, mu, var)
        for (kind, x), p in zip(reqs, ps):
            # per-frame log-likelihood against the oracle
            ref = orc.frame_loglik(x, w, mu, var)
            assert np.allclose(p.frame_ll, ref, rtol=1e-5, atol=1e-5)
            assert abs(p.ll_sum - ref.sum()) <= 1e-5 * abs(ref.sum())
            assert p.ll_sum == float(np.sum(p.frame_ll.astype(np.float64))) or \
                abs(p.ll_sum - np.sum(p.frame_ll.astype(np.float64))) <= 1e-9 * abs(p.ll_sum)
            if kind == REQ_SCORE:
                continue
            off = np.array([0, len(x)], np.uint64)
            res = acc.batch_adapt(x, off, tau=16, prior=True, chunk_frames=256)
            assert np.array_equal(p.occ, res["occ"][0])
            if kind == REQ_MAP:
                assert np.array_equal(p.means, res["means"][0])
            else:
                assert np.array_equal(p.W, res["W"][0])
                assert (p.status == 0) == (res["status"][0] == 0)
    finally:
        srv.close()
        acc.close()


def test_server_budget_admission(ubm):
    """A small slot budget splits the load into many batches; a request that
    can never fit fails alone with ERR_OOM; the rest are unaffected."""
    w, mu, var = ubm
    reqs = make_requests(13, w, mu, var, 20, lo=50, hi=400)
    big = synth.sample_frames_np(w.astype(np.float64), mu, var, 20000, 99)
    srv_ref = gmm_amd.Server(w, mu, var, chunk_frames=256)
    srv = gmm_amd.Server(w, mu, var, chunk_frames=256, slot_bytes=256 << 10, slots=3)
    try:
        ref = [srv_ref.submit(k, x, occ=True) for k, x in reqs]
        srv_ref.drain()
        ps = [srv.submit(k, x, occ=True) for k, x in reqs[:10]]
        pb = srv.submit(REQ_MAP, big)
        ps += [srv.submit(k, x, occ=True) for k, x in reqs[10:]]
        srv.drain()
        assert pb.status == gmm_amd.ERR_OOM
        assert "budget" in srv.last_error()
        st = srv.stats()
        assert st["failed"] == 1 and st["completed"] == len(reqs)
        assert st["batches"] >= 4
        for a, b in zip(ref, ps):
            same(results(a), results(b))
    finally:
        srv.close()
        srv_ref.close()


def test_server_threads_and_shutdown(ubm):
    """Submissions from several threads; closing the server finishes every
    queued request."""
    w, mu, var = ubm
    reqs = make_requests(14, w, mu, var, 64, lo=20, hi=300)
    srv = gmm_amd.Server(w, mu, var, chunk_frames=256, max_batch=8)
    ps = [None] * len(reqs)

    def worker(k):
        for i in range(k, len(reqs), 4):
            ps[i] = srv.submit(reqs[i][0], reqs[i][1])

    th = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.close()
    assert all(p.status in (0, gmm_amd.ERR_MLLR) for p in ps)
    ref = gmm_amd.Server(w, mu, var, chunk_frames=256)
    try:
        for (k, x), p in zip(reqs, ps):
            q = ref.submit(k, x)
            ref.wait(q)
            same(results(p), results(q))
    finally:
        ref.close()


def test_server_rejects_bad_requests(ubm):
    w, mu, var = ubm
    srv = gmm_amd.Server(w, mu, var)
    try:
        r = gmm_amd.Request()
        r.kind = REQ_MAP
        x = np.zeros((10, mu.shape[1]), np.float32)
        r.frames = x.ctypes.data
        r.n_frames = 10
        import ctypes
        assert gmm_amd.lib().gmmhip_server_submit(srv._h, ctypes.byref(r)) == gmm_amd.ERR_ARG
        r.kind = REQ_SCORE
        r.n_frames = 0
        assert gmm_amd.lib().gmmhip_server_submit(srv._h, ctypes.byref(r)) == gmm_amd.ERR_ARG
    finally:
        srv.close()


# Mixed forms: per-frame log-likelihood rtol / statistics rtol (n, F scaled
# by sqrt(n S)), as the mixed E-step states; MAP means are checked through
# the oracle's MAP update of the server's own statistics (exact algebra,
# 1e-6) and the occupancies against the oracle's.
MIXED_TOL = {gmm_amd.PREC_FP16X3: (1e-5, 2e-4), gmm_amd.PREC_BF16X3: (2e-4, 2e-3)}


@pytest.mark.parametrize("prec", [gmm_amd.PREC_FP16X3, gmm_amd.PREC_BF16X3])
def test_server_mixed_alone_vs_mix_and_oracle(ubm, prec):
    w, mu, var = ubm
    check_mixed_server(w, mu, var, prec, make_requests(21, w, mu, var, 40, lo=20, hi=1200))


def test_server_mixed_padded_rows():
    """D = 40: the split feature rows hold 128 columns, of which the
    log-likelihood kernel contracts 96 (2D+1 = 81 rounded up to 32) and the
    statistics kernel reads all; alone = in the mix, and the oracle."""
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    w, mu, var = make_ubm(6, 64, 40)
    check_mixed_server(w, mu, var, gmm_amd.PREC_FP16X3,
                       make_requests(23, w, mu, var, 16, lo=20, hi=800))


def check_mixed_server(w, mu, var, prec, reqs):
    lt, st = MIXED_TOL[prec]
    srv = gmm_amd.Server(w, mu, var, tau=16, chunk_frames=256, precision=prec)
    try:
        assert srv.stats()["precision"] == prec
        alone = []
        for kind, x in reqs:
            p = srv.submit(kind, x, frame_ll=True, occ=True)
            assert srv.wait(p, 60000) in (0, gmm_amd.ERR_MLLR)
            alone.append(results(p))
        ps = [srv.submit(kind, x, frame_ll=True, occ=True) for kind, x in reqs]
        srv.drain()
        s = srv.stats()
        assert s["fp32_fallbacks"] == 0 and s["failed"] == 0
        assert s["batches"] < 2 * len(reqs)
        for a, p in zip(alone, ps):
            same(a, results(p))
        for (kind, x), p in zip(reqs, ps):
            ref = orc.frame_loglik(x, w, mu, var)
            assert np.all(np.abs(p.frame_ll - ref) <= lt * np.abs(ref) + 1e-6)
            assert abs(p.ll_sum - ref.sum()) <= lt * np.abs(ref).sum()
            if kind == REQ_SCORE:
                continue
            rn, rF, rS, _ = orc.accumulate(x, w, mu, var)
            assert np.all(np.abs(p.occ - rn) <= st * rn + 3e-4)
            if kind == REQ_MAP:
                # the server's MAP from its own statistics: F is not returned,
                # so check mu' against the oracle's MAP of the oracle's
                # statistics within the statistics tolerance
                _, rmu, _ = orc.update_map(w, mu, var, rn, rF, rS, 1, 16)
                xm = float(np.abs(x).max())
                assert np.allclose(p.means, rmu, rtol=st, atol=st * xm)
    finally:
        srv.close()


def test_server_fp16_out_of_range_request_runs_fp32(ubm):
    """A request with a value far outside the FP16 range after the UBM's
    centring and scaling runs in an FP32 batch: its results equal the FP32
    server's bit for bit, the others stay in the FP16 form."""
    w, mu, var = ubm
    reqs = make_requests(22, w, mu, var, 6, lo=50, hi=300)
    odd = reqs[1][1].copy()
    odd[7, 2] = 1e6
    reqs[1] = (REQ_MAP, odd)
    srv = gmm_amd.Server(w, mu, var, chunk_frames=256, precision=gmm_amd.PREC_FP16X3)
    ref = gmm_amd.Server(w, mu, var, chunk_frames=256)
    try:
        ps = [srv.submit(k, x, occ=True) for k, x in reqs]
        srv.drain()
        assert srv.stats()["fp32_fallbacks"] == 1
        q = ref.submit(REQ_MAP, odd, occ=True)
        ref.wait(q)
        same(results(ps[1]), results(q))
        q0 = ref.submit(reqs[0][0], reqs[0][1], occ=True)
        ref.wait(q0)
        assert ps[0].ll_sum != q0.ll_sum  # the others ran in the FP16 form
    finally:
        srv.close()
        ref.close()
