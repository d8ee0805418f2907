#!/usr/bin/env python3
"""E-step throughput of the gfx950 diagonal-GMM accumulator.

A step is one E-step pass (posteriors + zeroth/first/second-order
statistics + total log-likelihood, the reference's accumulateStatsMT) over
the whole resident batch of synthetic frames.  Default workload: the
mixed-precision E-step configuration, D=60, M=2048, N=20,000,000 frames
from a seeded 2048-component mixture with frame offsets ~N(5, 1) per dim;
the model is initialised from 2048 random frames with the global variance.
The headline is the mixed-precision form (split fp16 operands, FP32
accumulate, FP64 reduction), timed beside the plain FP32 form on the same
frames (`fp32_compare`), with each form's error against an FP64 reference
of the same E-step over all frames (`accuracy_vs_fp64`).

Other workloads: --workload ubm_plain (D=60, M=512, N=1e7), plain_em_small,
batch_map (batched MAP + MLLR of 4,096 utterances), serve (serving loop),
fullcov_estep (full-covariance E-step, D=60, M=512, N=1e6).

Prints ONE JSON line (rank 0).  Frames are generated directly in HBM, so the
timed region starts with inputs resident.  Beside the headline the line
carries: the first E-step on a new frame set (the mixed form's
once-per-frame-set preparation included), errors against FP64 (total LL,
per-frame LL over all frames, posteriors of a frame sample, occupancies,
M-step means and variances), the serving loop at 5,000 req/s on the
2048-component UBM (p50/p99 latency, FP32 and fp16x3), and the restated
CPU path on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3e12, "bf16": 2.5e15, "fp16": 2.5e15}
HBM_PEAK = 8.0e12


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(x_sample, w, mu, var, budget_s=12.0, max_frames=10_000_000):
    """Restated SSE CPU path (oracle/sse_port.c) on the host cores over a
    bounded prefix of the same frames: one thread per core this process may
    run on, at most 16 (the host-core share of one GPU on a GPU box)."""
    from oracle import binding as orc
    visible = len(os.sched_getaffinity(0))
    threads = max(1, min(16, visible))
    M = mu.shape[0]
    n0 = min(4000, x_sample.shape[0])
    t0 = time.perf_counter()
    orc.sse_accumulate(x_sample[:n0], w, mu, var, threads)
    dt = time.perf_counter() - t0
    n = int(min(max_frames, x_sample.shape[0], max(n0, n0 * budget_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    _, _, _, _, used = orc.sse_accumulate(x_sample[:n], w, mu, var, threads)
    dt = time.perf_counter() - t0
    return {"value": n * M / dt, "unit": "frame*comp/s", "cores": int(used),
            "kind": "port", "cpu_model": cpu_model(), "cores_visible": visible,
            "sample": "first %d frames of the workload, %.1f s, reference SSE CPU path restated "
                      "(oracle/sse_port.c), %d threads on %s (%d cores visible to the "
                      "process)" % (n, dt, used, cpu_model(), visible)}


def fp64_reference(x, w, mu, var, min_gamma=1e-4, chunk=262144, frame_ll=None):
    """FP64 E-step over all frames on the GPU (torch float64; a checker,
    not the product): total log-likelihood and n, F, S with the 1e-4
    posterior cut, for the error report; per-frame log-likelihoods into
    frame_ll (float64 tensor [N]) when given."""
    import torch
    dev = x.device
    M, D = mu.shape
    mu64 = torch.as_tensor(mu, dtype=torch.float64, device=dev)
    var64 = torch.as_tensor(var, dtype=torch.float64, device=dev)
    ivar = 1.0 / var64
    gconst = (torch.log(torch.as_tensor(w, dtype=torch.float64, device=dev))
              - 0.5 * (D * np.log(2 * np.pi) + torch.log(var64).sum(1))
              - 0.5 * (mu64 * mu64 * ivar).sum(1))
    B1 = (mu64 * ivar).t().contiguous()
    B2 = (-0.5 * ivar).t().contiguous()
    tot = torch.zeros((), dtype=torch.float64, device=dev)
    n = torch.zeros(M, dtype=torch.float64, device=dev)
    F = torch.zeros((M, D), dtype=torch.float64, device=dev)
    S = torch.zeros((M, D), dtype=torch.float64, device=dev)
    for a in range(0, x.shape[0], chunk):
        xc = x[a:a + chunk].double()
        x2 = xc * xc
        L = torch.mm(xc, B1) + torch.mm(x2, B2) + gconst
        lse = torch.logsumexp(L, dim=1)
        if frame_ll is not None:
            frame_ll[a:a + chunk] = lse
        tot += lse.sum()
        g = torch.exp(L - lse[:, None])
        g = torch.where(g > min_gamma, g, torch.zeros_like(g))
        gt = g.t().contiguous()
        n += g.sum(0)
        F += torch.mm(gt, xc)
        S += torch.mm(gt, x2)
        del L, g, gt
    return float(tot), n.cpu().numpy(), F.cpu().numpy(), S.cpu().numpy()


def accuracy(res, ref):
    """Errors of one form's statistics against the FP64 reference: total
    log-likelihood (relative), occupancies n (max abs; n is the sum of the
    posteriors), and the M-step means F/n and variances S/n - mean^2 (max
    abs over components with n >= 1)."""
    n, F, S, tot = res[0], res[1], res[2], res[3]
    rt, rn, rF, rS = ref
    ok = rn >= 1.0
    mu, rmu = F[ok] / n[ok, None], rF[ok] / rn[ok, None]
    v, rv = S[ok] / n[ok, None] - mu * mu, rS[ok] / rn[ok, None] - rmu * rmu
    return {"tot_loglike_rel_err": abs(tot - rt) / abs(rt),
            "occupancy_max_abs_err": float(np.max(np.abs(n - rn))),
            "occupancy_max_rel_err": float(np.max(np.abs(n[ok] - rn[ok]) / rn[ok])),
            "mean_max_abs_err": float(np.max(np.abs(mu - rmu))),
            "var_max_abs_err": float(np.max(np.abs(v - rv))),
            "var_max_rel_err": float(np.max(np.abs(v - rv) / rv))}


def frame_ll_error(acc, x, ref_ll):
    """Max relative error of the per-frame log-likelihoods (the handle's
    current form, written to device memory) against FP64, over all frames."""
    import torch
    got = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    acc.loglikes_device(got.data_ptr())
    torch.cuda.synchronize()
    e = 0.0
    for a in range(0, x.shape[0], 1 << 22):
        r = ref_ll[a:a + (1 << 22)]
        e = max(e, float(((got[a:a + (1 << 22)].double() - r).abs() / r.abs()).max()))
    return e


def posterior_error(acc, x, w, mu, var, nsample=2048):
    """Posteriors of a sample of frames (every N/nsample-th) through the
    handle's current form, against FP64: each sampled frame is one
    utterance of the batched path with the cut disabled, so its occupancies
    are its posteriors.  The batched mixed form centres and scales by the
    model's global mean and deviation instead of the frame set's."""
    import torch
    N, D = x.shape
    M = mu.shape[0]
    idx = torch.arange(0, N, max(1, N // nsample), device=x.device)[:nsample]
    xs = x[idx].contiguous()
    U = xs.shape[0]
    off = np.arange(U + 1, dtype=np.uint64)
    occ = torch.empty((U, M), dtype=torch.float64, device=x.device)
    acc.set_opts(min_gamma=1e-30)
    try:
        acc.batch_adapt_device(xs.data_ptr(), off, D, {"occ": occ.data_ptr()}, tau=16)
    finally:
        acc.set_opts(min_gamma=0.0)
    torch.cuda.synchronize()
    x64 = xs.double()
    mu64 = torch.as_tensor(mu, dtype=torch.float64, device=x.device)
    var64 = torch.as_tensor(var, dtype=torch.float64, device=x.device)
    L = (torch.log(torch.as_tensor(w, dtype=torch.float64, device=x.device))
         - 0.5 * (D * np.log(2 * np.pi) + torch.log(var64).sum(1))
         - 0.5 * (((x64[:, None, :] - mu64[None]) ** 2) / var64[None]).sum(-1))
    g = torch.softmax(L, dim=1)
    d = (occ - g).abs()
    big = g > 1e-3
    return {"frames": int(U), "max_abs_err": float(d.max()),
            "max_rel_err_above_1e-3": float((d[big] / g[big]).max()) if bool(big.any()) else 0.0}


def serve_leg(rate=5000, seconds=2.0, precisions=("fp32", "fp16x3")):
    """The serving loop at one arrival rate (serveLoad, a child process that
    links the same library): p50/p99 latency from the scheduled arrival,
    throughput and the alone-vs-mix check, per precision."""
    import subprocess
    exe = os.path.join(ROOT, "gmm_amd", "_build", "serveLoad")
    out = {}
    for prec in precisions:
        cmd = [exe, "--rates", str(rate), "--seconds", str(seconds), "--precision", prec]
        print("bench: " + " ".join(cmd), file=sys.stderr, flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            out[prec] = {"error": "serveLoad rc %d" % r.returncode}
            continue
        row = [json.loads(ln) for ln in r.stdout.splitlines() if ln.startswith("{")][0]
        out[prec] = {k: row[k] for k in ("rate", "requests", "failed", "throughput_rps",
                                         "p50_ms", "p90_ms", "p99_ms", "max_ms", "map_p99_ms",
                                         "mllr_p99_ms", "verify_checked", "verify_identical")}
    return out


def traffic_of(kernel, workload):
    """HBM bytes per launch of a kernel on a workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/traffic.json (scaled
    by the copy-kernel calibration stated there), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        t = json.load(open(p))
        return t["kernels"][workload][kernel]["bytes_per_launch"]
    except (KeyError, ValueError):
        return None


def bench_batch(args):
    """Batched MAP (tau=16) + global MLLR statistics and transforms for 4,096
    utterances of uneven length against a 2048-component D=60 UBM: the whole
    batch in one pipeline captured in a hipGraph (E-step in --precision; the
    FP32 form's work-queue launch reported beside it) vs one launch set per
    utterance.  Frames and outputs stay in HBM."""
    import torch
    import gmm_amd
    from gmm_amd import synth
    cfg = dict(synth.CONFIGS["batch_map"])
    D, M, U = cfg["D"], cfg["M"], cfg["U"]
    w, mu, var = synth.true_mixture(**cfg)
    lens = synth.utterance_lengths(**cfg)
    x, off = synth.sample_utterances_torch(w, mu, var, lens, cfg["shift_sd"], cfg["seed"])
    N = int(off[-1])
    acc = gmm_amd.Accumulator(0)
    acc.set_model(w.astype(np.float32), mu, var)
    dev = "cuda"
    means = torch.empty((U, M, D), dtype=torch.float32, device=dev)
    occ = torch.empty((U, M), dtype=torch.float64, device=dev)
    W = torch.empty((U, D, D + 1), dtype=torch.float64, device=dev)
    status = torch.empty(U, dtype=torch.int32, device=dev)
    outs = {"means": means.data_ptr(), "occ": occ.data_ptr(), "W": W.data_ptr(),
            "status": status.data_ptr()}
    torch.cuda.synchronize()

    def timed(mode, graph, steps):
        acc.batch_adapt_device(x.data_ptr(), off, D, outs, tau=16, prior=True, mode=mode,
                               graph=graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            acc.batch_adapt_device(x.data_ptr(), off, D, outs, tau=16, prior=True, mode=mode,
                                   graph=graph)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    precs = {"fp32": gmm_amd.PREC_FP32, "fp16x3": gmm_amd.PREC_FP16X3,
             "bf16x3": gmm_amd.PREC_BF16X3}
    acc.set_opts(precision=precs[args.precision])
    t_q = timed(gmm_amd.BATCH_QUEUE, True, args.steps)
    st = status.cpu().numpy()
    means_q = means.cpu().numpy()
    # the FP32 queue form on the same batch, and the per-utterance launches
    acc.set_opts(precision=gmm_amd.PREC_FP32)
    t_q32 = timed(gmm_amd.BATCH_QUEUE, True, args.steps) if args.precision != "fp32" else t_q
    means_32 = means.cpu().numpy()
    t_u = timed(gmm_amd.BATCH_PER_UTTERANCE, False, max(1, args.steps // 4))
    line = {
        "metric": "batched MAP+MLLR utterances/s (persistent queue + hipGraph)",
        "value": U / t_q, "unit": "utterances/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": t_q * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "%s E-step, fp64 MLLR statistics and solve" % args.precision,
        "data": "synthetic (seeded UBM, per-utterance shifted mixtures, frames in HBM)",
        "config": {"workload": "batch_map", "D": D, "M": M, "U": U, "frames": N, "tau": 16,
                   "mllr_prior": True},
        "fp32_queue": {"ms_per_batch": t_q32 * 1e3, "utterances_per_s": U / t_q32,
                       "map_means_max_abs_diff_vs_this": float(np.abs(means_q - means_32).max())},
        "per_utterance_launches": {"ms_per_batch": t_u * 1e3, "utterances_per_s": U / t_u,
                                   "precision": "fp32"},
        "speedup_vs_per_utterance": t_u / t_q,
        "frame_comp_per_s": N * M / t_q,
        "mllr_underdetermined": int((st != 0).sum()),
    }
    acc.close()
    print(json.dumps(line), flush=True)


def bench_copy(args):
    """A plain 256 MiB device-to-device copy (torch, the runtime's copy
    kernel): the known-size launch that scales rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
    import torch
    n = 64 * 1024 * 1024
    x = torch.randn(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"metric": "device copy GB/s", "value": 2 * n * 4 * args.steps / el / 1e9,
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
                      "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
                      "data": "synthetic", "config": {"workload": "copy", "bytes": n * 4}}),
          flush=True)


def bench_fullcov(args):
    """Full-covariance E-step (the reference's gammasKernelFull +
    accStatsFullKernel path): D=60, M=512, N=1,000,000 frames drawn from a
    seeded 512-component full-covariance mixture; the model is initialised
    from 512 random frames with the global covariance (the first EM
    iteration).  One step = gmmhip_accumulate over the resident frames:
    extended rows, the two FP32 MFMA contractions with the feature map
    formed in LDS (log-likelihoods, gamma^T Phi), the fused
    log-sum-exp/posterior pass and the FP64 chunk reduction.
    Algorithmic work 4K FLOP per frame x component with K = 1 + D + D(D+1)/2
    features."""
    import torch
    import gmm_amd
    D, M = 60, 512
    N = args.frames or 1_000_000
    seed = 20261208
    rng = np.random.default_rng(seed)
    wt = rng.dirichlet(np.full(M, 0.5))
    mut = rng.normal(0, 2.0, (M, D))
    A = rng.normal(size=(M, D, D)) / np.sqrt(D)
    covt = A @ A.transpose(0, 2, 1) + np.diag(rng.uniform(0.3, 1.5, D))
    Lt = np.linalg.cholesky(covt)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.empty((N, D), dtype=torch.float32, device="cuda")
    Lg = torch.as_tensor(Lt, dtype=torch.float32, device="cuda")
    mug = torch.as_tensor(mut, dtype=torch.float32, device="cuda")
    wg = torch.as_tensor(wt, dtype=torch.float64, device="cuda")
    for a in range(0, N, 250_000):
        b = min(N, a + 250_000)
        k = torch.multinomial(wg, b - a, replacement=True, generator=g)
        z = torch.randn((b - a, D, 1), generator=g, device="cuda")
        x[a:b] = mug[k] + torch.bmm(Lg[k].contiguous(), z)[:, :, 0]
    torch.cuda.synchronize()
    idx = np.sort(rng.choice(N, size=M, replace=False))
    mu = x[torch.as_tensor(idx, device="cuda")].cpu().numpy()
    xd = x.double()
    cg = torch.cov(xd.t()).cpu().numpy().astype(np.float32)
    cov = np.tile(0.5 * (cg + cg.T), (M, 1, 1)).astype(np.float32)
    w = np.full(M, 1.0 / M, np.float32)
    del xd
    acc = gmm_amd.Accumulator(0)
    acc.set_model_full(w, mu, cov)
    acc.set_frames_device(x.data_ptr(), N, D)
    for _ in range(args.warmup):
        acc.accumulate()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = acc.accumulate()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    K = 1 + D + D * (D + 1) // 2
    flop = 4.0 * K * N * M
    step = el / args.steps
    line = {
        "metric": "full-covariance E-step frame*comp/s (fp32)",
        "value": N * M * args.steps / el, "unit": "frame*comp/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded full-covariance mixture, frames generated in HBM)",
        "config": {"workload": "fullcov_estep", "D": D, "M": M, "N": N, "K": K,
                   "tot_loglike_per_frame": res[3] / N},
        "roofline": {"bound": "mfma", "achieved": flop / step / 1e12, "peak": PEAK["fp32"] / 1e12,
                     "unit": "TFLOP/s", "frac": flop / step / PEAK["fp32"], "traffic": None,
                     "kernel": "whole E-step (extended rows + fc_ll + log-sum-exp + fc_stats + "
                               "FP64 reduction)", "avg_launch_ms": step * 1e3},
    }
    if not args.no_cpu_baseline:
        from oracle import binding as orc
        xs = x[:64].cpu().numpy()
        t0 = time.perf_counter()
        orc.accumulate_full(xs, w, mu, cov)
        dt = time.perf_counter() - t0
        n = int(min(N, max(64, 64 * 10.0 / max(dt, 1e-3))))
        xs = x[:n].cpu().numpy()
        t0 = time.perf_counter()
        orc.accumulate_full(xs, w, mu, cov)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": n * M / dt, "unit": "frame*comp/s", "cores": 1, "kind": "port",
            "sample": "first %d frames of the workload, %.1f s, the reference's full-covariance "
                      "CPU E-step restated in FP64 (oracle/gmm_oracle.c, orc_accumulate_full), "
                      "1 thread" % (n, dt)}
    acc.close()
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """Serving loop: UBM D=60, M=2048 resident in HBM; open-loop Poisson
    arrivals (60% score / 30% MAP / 10% MLLR, lengths lognormal(median 300,
    sigma 1.0) clipped to [20, 6000] frames) at each rate; latency from the
    scheduled arrival to completion.  Runs the serveLoad tool (C++, links
    libgmmhip.so) as a child process, one JSON line per rate, and reports the
    p99 at 5000 req/s as the value."""
    import subprocess
    exe = os.path.join(ROOT, "gmm_amd", "_build", "serveLoad")
    if not os.path.exists(exe):
        raise SystemExit("serveLoad not built: run `python -m gmm_amd.build`")
    rows = []
    precisions = ["fp32"] if args.precision == "fp32" else ["fp32", args.precision]
    for prec in precisions:
        cmd = [exe, "--rates", args.rates, "--seconds", str(args.serve_seconds),
               "--precision", prec]
        print("bench: " + " ".join(cmd), file=sys.stderr, flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        sys.stderr.write(r.stderr)
        if r.returncode != 0:
            raise SystemExit("serveLoad failed (%d): %s" % (r.returncode, r.stdout[-2000:]))
        rows += [json.loads(ln) for ln in r.stdout.splitlines() if ln.startswith("{")]
    head = [r for r in rows if r["precision"] == "fp32" and r["rate"] == 5000] or rows[:1]
    h = head[0]
    line = {
        "metric": "serving p99 latency at %d req/s (score/MAP/MLLR mix, M=2048 UBM)" % h["rate"],
        "value": h["p99_ms"], "unit": "ms", "n_gpus": 1, "steps": len(rows),
        "warmup": 1, "ms_per_step": h["p50_ms"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": h["precision"],
        "data": "synthetic (seeded UBM, per-utterance shifted mixtures, Poisson arrivals)",
        "config": {"workload": "serve", "D": h["D"], "M": h["M"], "rates": args.rates,
                   "seconds_per_rate": args.serve_seconds, "mix": "0.6/0.3/0.1",
                   "lengths": "lognormal(300, 1.0) in [20, 6000]"},
        "rates": rows,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mixed_estep")
    ap.add_argument("--frames", type=int, default=0, help="override N")
    ap.add_argument("--dim", type=int, default=0,
                    help="override D of an E-step workload (the headline is D=60)")
    ap.add_argument("--precision", default="fp16x3", choices=["fp32", "fp16x3", "bf16x3"])
    ap.add_argument("--no-compare", action="store_true", help="skip the FP32 comparison leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 reference")
    ap.add_argument("--rates", default="2000,5000,10000", help="serve: arrival rates (req/s)")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="serve: seconds per rate")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving leg of the E-step workloads")
    ap.add_argument("--serve-seconds-headline", type=float, default=2.0,
                    help="seconds of the serving leg (5000 req/s) in the E-step line")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: the E-step of this build runs on a single "
              "MI355X (frames would shard N/K per GPU with one all-reduce of the statistics "
              "per E-step, not built here). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch_map":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "fullcov_estep":
        return bench_fullcov(args)

    import torch
    import gmm_amd
    from gmm_amd import synth

    cfg = dict(synth.CONFIGS[args.workload])
    if args.frames:
        cfg["N"] = args.frames
    if args.dim:
        cfg["D"] = args.dim
    D, M, N = cfg["D"], cfg["M"], cfg["N"]
    w0, mu0, var0 = synth.true_mixture(**cfg)
    x = synth.sample_frames_torch(w0, mu0, var0, N, cfg["seed"])
    w, mu, var = synth.init_model_from_frames(x, M, cfg["seed"])
    torch.cuda.synchronize()

    precs = {"fp32": gmm_amd.PREC_FP32, "fp16x3": gmm_amd.PREC_FP16X3,
             "bf16x3": gmm_amd.PREC_BF16X3}
    acc = gmm_amd.Accumulator(0, profile=1)
    acc.set_model(w, mu, var)
    acc.set_frames_device(x.data_ptr(), N, D)

    def run(precision):
        """K timed E-steps (after W untimed ones) in one precision."""
        acc.set_opts(profile=1, precision=precs[precision])
        for _ in range(args.warmup):
            acc.accumulate()
        acc.reset_kernel_times()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = acc.accumulate()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        return res, el, acc.kernel_times()

    compare = None
    if args.precision != "fp32" and not args.no_compare:
        r32, el32, kt32 = run("fp32")
        compare = {"precision": "fp32", "value": N * M * args.steps / el32,
                   "ms_per_step": el32 * 1e3 / args.steps,
                   "stats_kernel_ms": kt32["stats"][0] / max(kt32["stats"][1], 1),
                   "tot_loglike": r32[3]}
    res, el, kt = run(args.precision)
    ref = None
    ll_err = {}
    if not args.no_accuracy:
        ref_ll = torch.empty(N, dtype=torch.float64, device=x.device)
        ref = fp64_reference(x, w, mu, var, frame_ll=ref_ll)
        for prec in ([args.precision] + (["fp32"] if compare is not None else [])):
            acc.set_opts(profile=1, precision=precs[prec])
            acc.set_frames_device(x.data_ptr(), N, D)  # the posterior sample replaces them
            ll_err[prec] = {"frame_loglike_max_rel_err": frame_ll_error(acc, x, ref_ll),
                            "posteriors": posterior_error(acc, x, w, mu, var)}
        del ref_ll
        acc.set_frames_device(x.data_ptr(), N, D)
    value = N * M * args.steps / el

    # the first E-step on a new frame set: the mixed form's once-per-frame-set
    # work (per-dimension frame statistics for the centring and the split
    # features) is included, the timed steps above reuse it
    first = {}
    for prec in ([args.precision] + (["fp32"] if compare is not None else [])):
        acc.set_opts(profile=1, precision=precs[prec])
        acc.set_frames_device(x.data_ptr(), N, D)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        acc.accumulate()
        torch.cuda.synchronize()
        first[prec] = (time.perf_counter() - t0) * 1e3
    acc.set_opts(profile=1, precision=precs[args.precision])

    flop_fc = 8 * D + 4
    st_ms, st_n = kt["stats"]
    ll_ms, ll_n = kt["loglik"]
    rd_ms, rd_n = kt["reduce"]
    st_avg = st_ms / max(st_n, 1) / 1e3
    achieved = flop_fc * N * M / st_avg
    peak = PEAK["fp32"] if args.precision == "fp32" else PEAK["fp16"]
    kname = "stats_kernel" if args.precision == "fp32" else "stats_mixed_tr_kernel"
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": achieved / peak,
                # measured for the workload's own N and D only
                "traffic": None if (args.frames or args.dim) else traffic_of(kname, args.workload),
                "kernel": kname,
                "avg_launch_ms": st_avg * 1e3,
                "pipeline_ms": {"loglik": ll_ms / max(ll_n, 1), "stats": st_ms / max(st_n, 1),
                                "reduce": rd_ms / max(rd_n, 1)},
                "pipeline_frac": flop_fc * N * M / ((ll_ms + st_ms + rd_ms) / max(st_n, 1) / 1e3)
                / peak}
    if args.precision != "fp32":
        roofline["executed_products"] = 3

    line = {
        "metric": "E-step frame*comp/s (%s%s)" % (
            args.precision, "" if compare is None else " vs fp32"),
        "value": value,
        "unit": "frame*comp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "fp32", "fp16x3": "fp16x3 (split fp16, fp32 accumulate)",
                  "bf16x3": "bf16x3 (split bf16, fp32 accumulate)"}[args.precision],
        "data": "synthetic (seeded true mixture, frames generated in HBM)",
        "config": {"workload": args.workload, "D": D, "M": M, "N": N,
                   "precision": args.precision, "tot_loglike_per_frame": res[3] / N},
        "roofline": roofline,
    }
    if compare is not None:
        compare["speedup"] = value / compare["value"]
        compare["tot_loglike_rel_diff"] = abs(res[3] - compare["tot_loglike"]) / abs(
            compare["tot_loglike"])
        line["fp32_compare"] = compare
    line["first_call"] = {"ms": {k: v for k, v in first.items()},
                          "note": "one E-step on a frame set just handed over (frames "
                                  "resident), once-per-frame-set preparation included"}
    if compare is not None:
        line["first_call"]["speedup_vs_fp32"] = first["fp32"] / first[args.precision]
    if ref is not None:
        line["accuracy_vs_fp64"] = {args.precision: accuracy(res, ref)}
        if compare is not None:
            line["accuracy_vs_fp64"]["fp32"] = accuracy(r32, ref)
        for prec, e in ll_err.items():
            line["accuracy_vs_fp64"][prec].update(e)
    if not args.no_serve:
        line["serving"] = serve_leg(rate=5000, seconds=args.serve_seconds_headline)
    if not args.no_cpu_baseline:
        ns = min(N, 10_000_000)
        xs = x[:ns].cpu().numpy()
        line["cpu_baseline"] = cpu_baseline(xs, w, mu, var)
    acc.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
