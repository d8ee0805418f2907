#!/usr/bin/env python3
"""E-step throughput of the gfx950 diagonal-GMM accumulator.

A step is one E-step pass (posteriors + zeroth/first/second-order
statistics + total log-likelihood, the reference's accumulateStatsMT) over
the whole resident batch of synthetic frames.  Default workload: UBM-scale
plain EM, D=60, M=512, N=10,000,000 frames (seeded true mixture), model
initialised from 512 random frames with the global variance.  The headline
is the mixed-precision form (split fp16 operands, FP32 accumulate; total
log-likelihood within 1e-6 relative of FP64), timed beside the plain FP32
form on the same frames (`fp32_compare`).

Prints ONE JSON line (rank 0).  Frames are generated directly in HBM, so the
timed region starts with inputs resident.
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


def cpu_baseline(x_sample, w, mu, var, budget_s=12.0, max_frames=10_000_000):
    """Restated SSE CPU path (oracle/sse_port.c) on the host cores over a
    bounded prefix of the same frames."""
    from oracle import binding as orc
    threads = max(1, min(16, os.cpu_count() or 1))
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
            "kind": "port",
            "sample": "first %d frames of the workload, %.1f s, reference SSE CPU path restated "
                      "(oracle/sse_port.c), %d threads" % (n, dt, used)}


def bench_batch(args):
    """Batched MAP (tau=16) + global MLLR statistics and transforms for 4,096
    utterances of uneven length against a 2048-component D=60 UBM: one
    persistent work-queue launch captured in a hipGraph vs one launch set per
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

    t_q = timed(gmm_amd.BATCH_QUEUE, True, args.steps)
    t_u = timed(gmm_amd.BATCH_PER_UTTERANCE, False, max(1, args.steps // 4))
    st = status.cpu().numpy()
    line = {
        "metric": "batched MAP+MLLR utterances/s (persistent queue + hipGraph)",
        "value": U / t_q, "unit": "utterances/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": t_q * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 E-step, fp64 MLLR statistics and solve",
        "data": "synthetic (seeded UBM, per-utterance shifted mixtures, frames in HBM)",
        "config": {"workload": "batch_map", "D": D, "M": M, "U": U, "frames": N, "tau": 16,
                   "mllr_prior": True},
        "per_utterance_launches": {"ms_per_batch": t_u * 1e3, "utterances_per_s": U / t_u},
        "speedup_vs_per_utterance": t_u / t_q,
        "frame_comp_per_s": N * M / t_q,
        "mllr_underdetermined": int((st != 0).sum()),
    }
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
    ap.add_argument("--workload", default="ubm_plain")
    ap.add_argument("--frames", type=int, default=0, help="override N")
    ap.add_argument("--precision", default="fp16x3", choices=["fp32", "fp16x3", "bf16x3"])
    ap.add_argument("--no-compare", action="store_true", help="skip the FP32 comparison leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--rates", default="2000,5000,10000", help="serve: arrival rates (req/s)")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="serve: seconds per rate")
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

    import torch
    import gmm_amd
    from gmm_amd import synth

    cfg = dict(synth.CONFIGS[args.workload])
    if args.frames:
        cfg["N"] = args.frames
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
    value = N * M * args.steps / el

    flop_fc = 8 * D + 4
    st_ms, st_n = kt["stats"]
    ll_ms, ll_n = kt["loglik"]
    rd_ms, rd_n = kt["reduce"]
    st_avg = st_ms / max(st_n, 1) / 1e3
    achieved = flop_fc * N * M / st_avg
    peak = PEAK["fp32"] if args.precision == "fp32" else PEAK["fp16"]
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "kernel": "stats_kernel" if args.precision == "fp32" else "stats_mixed_kernel",
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
    if not args.no_cpu_baseline:
        ns = min(N, 10_000_000)
        xs = x[:ns].cpu().numpy()
        line["cpu_baseline"] = cpu_baseline(xs, w, mu, var)
    acc.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
