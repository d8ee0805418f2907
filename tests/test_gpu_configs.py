"""The benchmark configurations on their own workloads, against the oracle.

  plain_em_small  D=39, M=64, N=100,000: 10 EM iterations (--type 2 --tau 0
                  --mean --var --weight) through CModelAdapt (trainGMM) and
                  through the C ABI step by step, against the oracle's EM
                  loop.
  ubm_plain       D=60, M=512: the workload's own mixture and model
                  initialisation, N=20,000 frames against the oracle; the
                  full N=1e7 set through size-independent properties.
  mixed_estep     D=60, M=2048, frame offsets ~N(5,1): fp16x3 and bf16x3
                  against the oracle at N=4,000; the full N=2e7 set through
                  properties (sum of occupancies with the cut disabled, total
                  LL = sum of per-frame LL, halves add up, first-order sums).
  batch_map       D=60, M=2048, utterance lengths lognormal(300, 0.8) in
                  [50, 3000]: 64 utterances against the per-utterance oracle
                  (FP32 and fp16x3); the full 4,096-utterance batch: queue,
                  graph replay and per-utterance launches bit-identical.
  serve           D=60, M=2048, lognormal(300, 1.0) in [20, 6000], 60/30/10
                  score/MAP/MLLR: serveLoad at 5,000 req/s (p99 <= 20 ms,
                  alone-vs-mix bit-identical), and requests against the
                  oracle.

Tolerances are the stated ones: FP32 1e-5 relative on per-frame LL and
statistics, 1e-6 on the total; fp16x3 1e-5 / 1e-6 / 2e-4; bf16x3
2e-4 / 2e-5 / 2e-3; statistics with the per-component cut allowance of
tests/_check.py.
"""
import json
import os
import subprocess
import time

import numpy as np
import pytest

import _check
import _formats as fm
import _oracle as orc
import gmm_amd
from gmm_amd import host, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(os.path.dirname(HERE), "gmm_amd", "_build")
TOL = {"fp32": (1e-5, 1e-6, 1e-5), "fp16x3": (1e-5, 1e-6, 2e-4), "bf16x3": (2e-4, 2e-5, 2e-3)}
PREC = {"fp32": gmm_amd.PREC_FP32, "fp16x3": gmm_amd.PREC_FP16X3,
        "bf16x3": gmm_amd.PREC_BF16X3}


@pytest.fixture(scope="module")
def acc():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    a = gmm_amd.Accumulator(0)
    yield a
    a.close()


def say(*a):
    print(*a, flush=True)


def check_estep(acc, name, x, w, mu, var, what=""):
    """Per-frame LL, total LL and statistics of one form against the oracle."""
    lt, tt, st = TOL[name]
    acc.set_opts(precision=PREC[name])
    try:
        acc.set_model(w, mu, var)
        acc.set_frames(x)
        ll = acc.loglikes().astype(np.float64)
        got = acc.accumulate()
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32)
    rll = orc.frame_loglik(x, w, mu, var)
    e = np.abs(ll - rll) / np.abs(rll)
    assert np.all(np.abs(ll - rll) <= lt * np.abs(rll) + 1e-6), (what, name, e.max())
    ref = orc.accumulate(x, w, mu, var)
    used = _check.check_stats(got, ref, x, w, mu, var, rtol=st, tot_rtol=tt,
                              band=_check.BAND[name], what="%s %s " % (what, name))
    say("%s %s: max rel LL %.3g, rel tot %.3g, allowance used %.3f" % (
        what, name, e.max(), abs(got[3] - ref[3]) / abs(ref[3]), used))
    assert got[4] == x.shape[0]
    return got


# ------------------------------------------------------------ plain_em_small
def _active(w, mu, var):
    """RearrangeMixtures order (OL_GMModel.cpp:857-885) and the number of
    components with weight > 0 (they come first)."""
    perm, nz = orc.rearrange_perm(w)
    return w[perm], mu[perm], var[perm], len(w) - nz


def _pad(M, n, F, S):
    Ma, D = F.shape
    out = (np.zeros(M), np.zeros((M, D)), np.zeros((M, D)))
    out[0][:Ma], out[1][:Ma], out[2][:Ma] = n, F, S
    return out


def test_plain_em_small_10_iterations(acc, tmp_path):
    """Every iteration's statistics against the oracle from the same model
    (FP32 bar), the host EM rule against the oracle's from the same
    statistics, CModelAdapt (trainGMM) bit-identical to the step-by-step
    run, and the 10-iteration trajectory against the oracle's own.
    Components whose occupancy falls under 2 are dropped on the way and
    move to the end (RearrangeMixtures), as in the reference."""
    cfg = synth.CONFIGS["plain_em_small"]
    w0, mu0, var0 = synth.true_mixture(**cfg)
    x = synth.sample_frames_np(w0, mu0, var0, cfg["N"], cfg["seed"])
    w, mu, var = synth.init_model_from_frames(x, cfg["M"], cfg["seed"])
    N, M = x.shape[0], cfg["M"]
    # trainGMM: the reference's CLI over CModelAdapt, binary model files
    init_p, data_p, out_p = tmp_path / "init.gmm", tmp_path / "data.raw", tmp_path / "em.gmm"
    host.Model.create(w, mu, var).save(init_p, False)
    fm.write_raw(data_p, x)
    r = subprocess.run([os.path.join(BIN, "trainGMM"), "-i", str(data_p), "--data-T", "2",
                        "--in-UBM", str(init_p), "--type", "2", "--it-EM", "10", "--tau", "0",
                        "--mean", "--var", "--weight", "-o", str(out_p)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    tw, tmu, tvar = host.Model.load(out_p, False).get()
    # the same iterations step by step through the C ABI and the host update
    # rule, each compared with the oracle from the same starting model
    model = host.Model.create(w, mu, var)
    ow, omu, ovar = w.copy(), mu.copy(), var.copy()
    worst, dropped = 0.0, 0
    for it in range(10):
        model.rearrange()
        gw, gmu, gvar = model.get()
        Ma = int((gw > 0).sum())
        aw, amu, avar = gw[:Ma], gmu[:Ma], gvar[:Ma]
        acc.set_model(aw, amu, avar)
        acc.set_frames(x)
        n, F, S, tot, nf = acc.accumulate()
        ref = orc.accumulate(x, aw, amu, avar)
        used = _check.check_stats((n, F, S, tot), ref, x, aw, amu, avar,
                                  band=_check.BAND["fp32"], what="iteration %d " % it)
        worst = max(worst, used)
        # one M-step from the same statistics: the host rule and the oracle's
        nP, FP, SP = _pad(M, n, F, S)
        model.update_em(nP, FP, SP, nf)
        hw, hmu, hvar = model.get()
        qw, qmu, qvar, _ = orc.update_em(gw, gmu, gvar, nP, FP, SP, nf)
        assert np.array_equal(hw > 0, qw > 0)
        k = hw > 0
        assert np.allclose(hw[k], qw[k], rtol=1e-6, atol=0)
        assert np.allclose(hmu[k], qmu[k], rtol=1e-6, atol=1e-6)
        assert np.allclose(hvar[k], qvar[k], rtol=1e-6, atol=0)
        # the oracle's own trajectory
        ow, omu, ovar, Mo = _active(ow, omu, ovar)
        on, oF, oS, _ = orc.accumulate(x, ow[:Mo], omu[:Mo], ovar[:Mo])
        ow, omu, ovar, _ = orc.update_em(ow, omu, ovar, *_pad(M, on, oF, oS), N)
        dropped = int((hw == 0).sum())
    gw, gmu, gvar = model.get()
    k = gw > 0
    # CModelAdapt (trainGMM) runs exactly these kernels and update rules; the
    # saved model skips zero-weight components
    assert np.array_equal(tw, gw[k]) and np.array_equal(tmu, gmu[k])
    assert np.array_equal(tvar, gvar[k])
    # after 10 iterations, against the oracle's trajectory (same components
    # dropped, same order).  Each iteration above meets the parity bar from
    # the same model; EM then carries each iteration's differences into the
    # next, so the trajectories drift apart by more than one iteration's
    # difference: parameters within 1e-4, the data log-likelihood within 1e-6.
    assert np.array_equal(ow > 0, k)
    Ma = int(k.sum())
    sd = np.sqrt(ovar[k])
    e_w = np.abs(gw[k] - ow[k]) / ow[k]
    e_mu = np.max(np.abs(gmu[k] - omu[k]) / sd, axis=1)
    e_var = np.max(np.abs(gvar[k] - ovar[k]) / ovar[k], axis=1)
    occ = ow[k] * N
    ll_g = orc.frame_loglik(x, gw[:Ma], gmu[:Ma], gvar[:Ma]).mean()
    ll_o = orc.frame_loglik(x, ow[:Ma], omu[:Ma], ovar[:Ma]).mean()
    say("10 EM iterations (%d components dropped): per-iteration allowance used %.3f; "
        "trajectory rel err w %.3g (n=%.0f), mu/sd %.3g (n=%.0f), var %.3g (n=%.0f); "
        "mean LL %.8f vs %.8f" % (dropped, worst, e_w.max(), occ[e_w.argmax()], e_mu.max(),
                                  occ[e_mu.argmax()], e_var.max(), occ[e_var.argmax()],
                                  ll_g, ll_o))
    assert e_w.max() <= 1e-4 and e_mu.max() <= 1e-4 and e_var.max() <= 1e-4
    assert abs(ll_g - ll_o) <= 1e-6 * abs(ll_o)


# ----------------------------------------------------------------- ubm_plain
def test_ubm_plain_against_oracle(acc):
    cfg = synth.CONFIGS["ubm_plain"]
    w0, mu0, var0 = synth.true_mixture(**cfg)
    x = synth.sample_frames_np(w0, mu0, var0, 20000, cfg["seed"])
    w, mu, var = synth.init_model_from_frames(x, cfg["M"], cfg["seed"])
    n, F, S, _, nf = check_estep(acc, "fp32", x, w, mu, var, "ubm_plain init")
    # after one EM update (a separated model: many posteriors cut)
    w1, mu1, var1, _ = orc.update_em(w, mu, var, n, F, S, nf)
    keep = w1 > 0
    check_estep(acc, "fp32", x, w1[keep], mu1[keep], var1[keep], "ubm_plain EM1")


def _device_frames(cfg, N):
    import torch
    w0, mu0, var0 = synth.true_mixture(**cfg)
    x = synth.sample_frames_torch(w0, mu0, var0, N, cfg["seed"])
    w, mu, var = synth.init_model_from_frames(x, cfg["M"], cfg["seed"])
    torch.cuda.synchronize()
    return x, w, mu, var


def _column_sums(x, chunk=1 << 20):
    import torch
    s = torch.zeros(x.shape[1], dtype=torch.float64, device=x.device)
    a = torch.zeros(x.shape[1], dtype=torch.float64, device=x.device)
    for i in range(0, x.shape[0], chunk):
        c = x[i:i + chunk].double()
        s += c.sum(0)
        a += c.abs().sum(0)
    return s.cpu().numpy(), a.cpu().numpy()


def _full_n_properties(acc, x, w, mu, var, name):
    """Size-independent checks of one form on a frame set resident in HBM."""
    N, D = x.shape
    ptr = x.data_ptr()
    acc.set_opts(precision=PREC[name])
    try:
        acc.set_model(w, mu, var)
        acc.set_frames_device(ptr, N, D)
        t0 = time.perf_counter()
        n, F, S, tot, nf = acc.accumulate()
        ll = acc.loglikes()
        say("%s N=%d: accumulate + loglikes %.2f s" % (name, N, time.perf_counter() - t0))
        assert nf == N
        # total LL = sum of the per-frame LLs (both FP64 sums of the same values)
        assert tot == pytest.approx(ll.astype(np.float64).sum(), rel=1e-9)
        assert np.all(np.isfinite(ll)) and np.all(ll > -1e19)
        # with the cut disabled, the posteriors of a frame sum to one: sum of
        # occupancies = N and sum over components of F = column sums of x
        acc.set_opts(precision=PREC[name], min_gamma=1e-30)
        n0, F0, S0, tot0, _ = acc.accumulate()
        acc.set_opts(precision=PREC[name], min_gamma=0.0)
        cs, ca = _column_sums(x)
        e_n = abs(n0.sum() - N) / N
        e_F = np.max(np.abs(F0.sum(0) - cs) / ca)
        say("%s: cut disabled: |sum n - N|/N %.3g, max |sum F - sum x|/sum|x| %.3g" % (
            name, e_n, e_F))
        assert e_n <= 1e-5 and e_F <= 1e-5
        assert tot0 == tot  # the cut does not touch the log-likelihoods
        # the cut only removes mass: no component gains
        assert np.all(n <= n0 * (1 + 1e-5) + 1e-9)
        # the mass of the posteriors within the form's band around the cut,
        # per component (cut at the band's two edges): the most by which two
        # runs that round differently can differ through cut flips
        band = _check.BAND[name]
        acc.set_opts(precision=PREC[name], min_gamma=float(np.float32(1e-4 * (1 - band))))
        lo = acc.accumulate()
        acc.set_opts(precision=PREC[name], min_gamma=float(np.float32(1e-4 * (1 + band))))
        hi = acc.accumulate()
        acc.set_opts(precision=PREC[name], min_gamma=0.0)
        dn = np.maximum(lo[0] - hi[0], 0.0)
        dS = np.maximum(lo[2] - hi[2], 0.0)
        # halves of the frame set add up to the whole (the mixed forms centre
        # and scale each frame set by its own statistics, so the halves round
        # differently from the whole: the form's tolerance applies)
        st = TOL[name][2]
        h = N // 2
        acc.set_frames_device(ptr, h, D)
        a = acc.accumulate()
        acc.set_frames_device(ptr + h * D * 4, N - h, D)
        b = acc.accumulate()
        assert a[3] + b[3] == pytest.approx(tot, rel=TOL[name][1])
        scaleF = np.sqrt(np.maximum(n[:, None] * S, 0.0))
        e = [np.abs(a[0] + b[0] - n) - (st * n + dn),
             np.abs(a[1] + b[1] - F) - (st * scaleF + np.sqrt(dn[:, None] * dS)),
             np.abs(a[2] + b[2] - S) - (st * S + dS)]
        say("%s: halves vs whole, worst excess over the tolerance n %.3g F %.3g S %.3g; "
            "components with near-cut mass %d" % (name, e[0].max(), e[1].max(), e[2].max(),
                                                   int((dn > 0).sum())))
        assert all(np.all(v <= 1e-9) for v in e)
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32, min_gamma=0.0)


def test_ubm_plain_full_n_properties(acc):
    cfg = synth.CONFIGS["ubm_plain"]
    x, w, mu, var = _device_frames(cfg, cfg["N"])
    _full_n_properties(acc, x, w, mu, var, "fp32")
    del x


# --------------------------------------------------------------- mixed_estep
@pytest.mark.parametrize("name", ["fp16x3", "bf16x3"])
def test_mixed_estep_against_oracle(acc, name):
    """Mixed forms at the headline shape (D=60, M=2048, offset frames); the
    model is the workload's initialisation (2048 frames of the mixture as
    means, global variance), drawn from frames other than those scored."""
    cfg = synth.CONFIGS["mixed_estep"]
    w0, mu0, var0 = synth.true_mixture(**cfg)
    xi = synth.sample_frames_np(w0, mu0, var0, 20000, cfg["seed"] + 11)
    w, mu, var = synth.init_model_from_frames(xi, cfg["M"], cfg["seed"])
    x = synth.sample_frames_np(w0, mu0, var0, 4000, cfg["seed"])
    check_estep(acc, name, x, w, mu, var, "mixed_estep")
    check_estep(acc, "fp32", x, w, mu, var, "mixed_estep")


@pytest.mark.parametrize("name", ["fp16x3", "bf16x3"])
def test_mixed_estep_full_n_properties(acc, name):
    cfg = synth.CONFIGS["mixed_estep"]
    x, w, mu, var = _device_frames(cfg, cfg["N"])
    _full_n_properties(acc, x, w, mu, var, name)
    del x


# ----------------------------------------------------------------- batch_map
@pytest.fixture(scope="module")
def batch_workload():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    cfg = dict(synth.CONFIGS["batch_map"])
    w, mu, var = synth.true_mixture(**cfg)
    lens = synth.utterance_lengths(**cfg)
    x, off = synth.sample_utterances_torch(w, mu, var, lens, cfg["shift_sd"], cfg["seed"])
    yield cfg, w.astype(np.float32), mu, var, x, off
    del x


@pytest.mark.parametrize("name", ["fp32", "fp16x3"])
def test_batch_map_against_oracle(acc, batch_workload, name):
    """64 utterances of the workload against the oracle applied utterance by
    utterance: occupancies and first-order statistics, MAP means (tau = 16)
    from them, MLLR status; 8 MLLR transforms against the oracle's solve of
    the GPU's statistics (exact FP64 algebra)."""
    cfg, w, mu, var, x, off = batch_workload
    U = 64
    xs = x[:int(off[U])].cpu().numpy()
    offs = off[:U + 1].copy()
    acc.set_model(w, mu, var)
    acc.set_opts(precision=PREC[name])
    try:
        res = acc.batch_adapt(xs, offs, tau=16, prior=True)
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32)
    lt, tt, st = TOL[name]
    worst = 0.0
    for u in range(U):
        xu = xs[int(offs[u]):int(offs[u + 1])]
        rn, rF, rS, _ = orc.accumulate(xu, w, mu, var)
        n, F = res["occ"][u], res["first"][u]
        worst = max(worst, _check.check_stats((n, F, None, None), (rn, rF, None, None), xu, w,
                                              mu, var, rtol=st, band=_check.BAND[name],
                                              what="utt %d " % u))
        _, rmu, _ = orc.update_map(w, mu, var, n, F, F ** 2, 1, 16)
        assert np.allclose(res["means"][u], rmu, rtol=1e-6, atol=1e-6), u
        assert (res["status"][u] == 0) == (orc.mllr_check(len(w), mu.shape[1], n) == 0), u
    for u in range(8):
        if res["status"][u] == 0:
            Wg = orc.mllr_transform(mu, var, res["occ"][u], res["first"][u], ubm_mu=mu,
                                    ubm_var=var)
            assert np.allclose(res["W"][u], Wg, rtol=1e-8, atol=1e-8), u
    say("batch_map %s: 64 utterances, allowance used %.3f, MLLR ok %d" % (
        name, worst, int((res["status"] == 0).sum())))


@pytest.mark.parametrize("name", ["fp32", "fp16x3"])
def test_batch_map_full_batch_modes(acc, batch_workload, name):
    """The whole 4,096-utterance batch: the graph-captured queue launch,
    its replays and the eager queue launch are bit-identical (and, in FP32,
    so are the per-utterance launches); the first 64 utterances alone give
    the same results as inside the full batch."""
    import torch
    cfg, w, mu, var, x, off = batch_workload
    U, D, M = cfg["U"], cfg["D"], cfg["M"]
    acc.set_model(w, mu, var)

    def run(mode, graph, xx=x, oo=off):
        u = len(oo) - 1
        o = {"means": torch.empty((u, M, D), dtype=torch.float32, device="cuda"),
             "occ": torch.empty((u, M), dtype=torch.float64, device="cuda"),
             "first": torch.empty((u, M, D), dtype=torch.float64, device="cuda"),
             "W": torch.empty((u, D, D + 1), dtype=torch.float64, device="cuda"),
             "status": torch.empty(u, dtype=torch.int32, device="cuda")}
        acc.batch_adapt_device(xx.data_ptr(), oo, D, {k: v.data_ptr() for k, v in o.items()},
                               tau=16, prior=True, mode=mode, graph=graph)
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in o.items()}

    acc.set_opts(precision=PREC[name])
    try:
        eager = run(gmm_amd.BATCH_QUEUE, False)
        graphs = [run(gmm_amd.BATCH_QUEUE, True) for _ in range(3)]
        per_utt = run(gmm_amd.BATCH_PER_UTTERANCE, False) if name == "fp32" else None
        sub = run(gmm_amd.BATCH_QUEUE, False, x[:int(off[64])], off[:65].copy())
    finally:
        acc.set_opts(precision=gmm_amd.PREC_FP32)
    for k in eager:
        for g in graphs:
            assert np.array_equal(eager[k], g[k]), k
        if per_utt is not None:
            assert np.array_equal(eager[k], per_utt[k]), k
        assert np.array_equal(sub[k], eager[k][:64]), k
    # every utterance's occupancies sum to (at most) its frame count
    lens = np.diff(off.astype(np.int64))
    tot = eager["occ"].sum(1)
    assert np.all(tot <= lens * (1 + 1e-5)) and np.all(tot >= lens * 0.98)
    say("batch_map %s: %d utterances, %d frames, MLLR under-determined %d" % (
        name, U, int(off[-1]), int((eager["status"] != 0).sum())))


# --------------------------------------------------------------------- serve
@pytest.mark.parametrize("prec", ["fp32", "fp16x3"])
def test_serve_load_p99(prec):
    """serveLoad: Poisson arrivals at 5,000 req/s for 2 s on the D=60,
    M=2048 UBM (60/30/10 score/MAP/MLLR, lengths lognormal(300, 1.0) in
    [20, 6000]); latency from the scheduled arrival.  p99 <= 20 ms, no
    failures, and the first 64 requests re-run alone are bit-identical to
    their results in the mix."""
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    cmd = [os.path.join(BIN, "serveLoad"), "--rates", "5000", "--seconds", "2",
           "--precision", prec, "--verify", "64"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    row = [json.loads(ln) for ln in r.stdout.splitlines() if ln.startswith("{")][0]
    say("serve %s @5000: p50 %.3f p99 %.3f max %.3f ms, %d requests, %.0f req/s" % (
        prec, row["p50_ms"], row["p99_ms"], row["max_ms"], row["requests"],
        row["throughput_rps"]))
    assert row["M"] == 2048 and row["D"] == 60
    assert row["failed"] == 0
    assert row["verify_checked"] == 64 and row["verify_identical"] == 64
    assert row["p99_ms"] <= 20.0
    assert row["throughput_rps"] >= 0.9 * 5000


@pytest.mark.parametrize("prec", ["fp32", "fp16x3"])
def test_serve_requests_against_oracle(prec):
    """Requests of the serving mix on the D=60, M=2048 UBM: every result
    bit-identical alone and in the mix, and equal to gmmhip_batch_adapt on
    the same utterance; per-frame log-likelihoods and occupancies against
    the oracle, MAP means against the MAP rule on the request's statistics."""
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")
    cfg = dict(synth.CONFIGS["batch_map"])
    w, mu, var = synth.true_mixture(**cfg)
    w = w.astype(np.float32)
    M, D = mu.shape
    rng = np.random.default_rng(55)
    kinds = rng.choice([gmm_amd.REQ_SCORE, gmm_amd.REQ_MAP, gmm_amd.REQ_MLLR], size=24,
                       p=[0.6, 0.3, 0.1])
    kinds[:3] = [gmm_amd.REQ_SCORE, gmm_amd.REQ_MAP, gmm_amd.REQ_MLLR]
    lens = np.clip(np.round(np.exp(rng.normal(np.log(300), 1.0, 24))), 20, 6000).astype(int)
    reqs = []
    for i in range(24):
        shift = rng.normal(0, 0.3, (M, D)).astype(np.float32)
        reqs.append((int(kinds[i]), synth.sample_frames_np(w.astype(np.float64), mu + shift, var,
                                                           int(lens[i]), 900 + i)))
    lt, tt, st = TOL[prec]
    srv = gmm_amd.Server(w, mu, var, tau=16, chunk_frames=512, precision=PREC[prec])
    acc = gmm_amd.Accumulator(0, precision=PREC[prec])
    try:
        alone = []
        for kind, x in reqs:
            p = srv.submit(kind, x, frame_ll=True, occ=True)
            assert srv.wait(p, 60000) in (0, gmm_amd.ERR_MLLR)
            alone.append(p)
        mixed = [srv.submit(kind, x, frame_ll=True, occ=True) for kind, x in reqs]
        srv.drain()
        acc.set_model(w, mu, var)
        for i, ((kind, x), a, b) in enumerate(zip(reqs, alone, mixed)):
            assert a.status == b.status and a.ll_sum == b.ll_sum, i
            assert np.array_equal(a.frame_ll, b.frame_ll), i
            for u, v in ((a.means, b.means), (a.W, b.W), (a.occ, b.occ)):
                assert (u is None) == (v is None)
                if u is not None:
                    assert np.array_equal(u, v), i
            if i >= 6:
                continue
            rll = orc.frame_loglik(x, w, mu, var)
            assert np.all(np.abs(b.frame_ll - rll) <= lt * np.abs(rll) + 1e-6), i
            assert abs(b.ll_sum - rll.sum()) <= tt * abs(rll.sum()) * 10, i
            if kind == gmm_amd.REQ_SCORE:
                continue
            rn, rF, _, _ = orc.accumulate(x, w, mu, var)
            off = np.array([0, len(x)], np.uint64)
            res = acc.batch_adapt(x, off, tau=16, chunk_frames=512)
            assert np.array_equal(b.occ, res["occ"][0]), i
            _check.check_stats((res["occ"][0], res["first"][0], None, None),
                               (rn, rF, None, None), x, w, mu, var, rtol=st,
                               band=_check.BAND[prec], what="request %d " % i)
            if kind == gmm_amd.REQ_MAP:
                assert np.array_equal(b.means, res["means"][0]), i
                _, rmu, _ = orc.update_map(w, mu, var, res["occ"][0], res["first"][0],
                                           res["first"][0] ** 2, 1, 16)
                assert np.allclose(b.means, rmu, rtol=1e-6, atol=1e-6), i
            else:
                assert np.array_equal(b.W, res["W"][0]), i
    finally:
        srv.close()
        acc.close()
