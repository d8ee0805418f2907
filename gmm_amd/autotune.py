"""Tile autotuner for the E-step kernels.

For each (precision, D, M) it times the E-step (accumulate: log-likelihood,
posteriors, statistics, FP64 reduction) over resident synthetic frames for
every candidate (log-likelihood frame tile, statistics component tile,
frames per statistics work item), and for the batched path
(gmmhip_batch_adapt) the same; the fastest choice of each is written to the
tuning table that libgmmhip.so reads (gmm_amd/tuning/gfx950.tune, or
$GMMHIP_TUNE_FILE).  Entries for other shapes in the table are kept.

    python -m gmm_amd.autotune --D 60 --M 512,2048 --prec fp32,fp16x3 \\
        --frames 4000000 [--batch] [--out PATH]

Frame tiles: the FP32 log-likelihood kernel's 64 / 128 / 256 frames per
block; the mixed form's 128 / 256 frames per block (four or eight waves of 32
frames sharing each model tile staged in LDS; 256 only where the depth
2D+1 rounded up to 32 is a multiple of 64).  Component tiles: the mixed
statistics kernel's 128 / 256 components per block (four or eight waves of 32
components sharing each staged frame step; 256 where the split feature
rows hold 64 or 128 columns);
the FP32 form has one (written as 128).  The work-item size is swept for
every form, in the E-step and (--batch) in the batched path, per precision.
Frames come from a seeded true mixture (means ~ N(0, 2^2), variances
~ U[0.3, 3], weights ~ Dirichlet(0.5)) generated in HBM.
"""
import argparse
import os
import sys
import time

import numpy as np

PREC = {"fp32": 0, "bf16x3": 1, "fp16x3": 2}
PREC_NAME = {v: k for k, v in PREC.items()}
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "gfx950.tune")

LL_FRAMES = [64, 128, 256]
LL_FRAMES_MIXED = [128, 256]


def frame_tiles(prec, D):
    """Candidate log-likelihood frame tiles of a form at dimension D."""
    if prec == PREC["fp32"]:
        return [ll for ll in LL_FRAMES if not (ll == 256 and 2 * D + 1 > 128)]
    kpm = (2 * D + 1 + 31) // 32 * 32
    return [ll for ll in LL_FRAMES_MIXED if not (ll == 256 and kpm % 64)]


def comp_tiles(prec, D):
    """Candidate statistics component tiles of a form at dimension D (256 where
    the split feature rows, 2D+1 rounded up to 32 and then to a power of two,
    hold 64 or 128 columns)."""
    kpm = (2 * D + 1 + 31) // 32 * 32
    if prec == PREC["fp32"] or not 32 < kpm <= 128:
        return [128]
    return [128, 256]


CHUNKS = [1024, 2048, 4096, 8192, 16384]
BATCH_CHUNKS = [256, 512, 1024, 2048]


def read_table(path):
    rows = {}
    if os.path.exists(path):
        for ln in open(path):
            t = ln.split("#", 1)[0].split()
            if len(t) >= 6:
                rows[(t[0], t[1], int(t[2]), int(t[3]))] = (
                    int(t[4]), int(t[5]), float(t[6]) if len(t) > 6 else 0.0,
                    int(t[7]) if len(t) > 7 else 128)
    return rows


def write_table(path, rows):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        f.write("# gmmhip tuning table (gfx950), written by python -m gmm_amd.autotune\n")
        f.write("# path prec D M ll_frames chunk_frames ms comp_tile\n")
        for k in sorted(rows):
            ll, cf, ms, ct = rows[k]
            f.write("%s %s %d %d %d %d %.3f %d\n" % (k[0], k[1], k[2], k[3], ll, cf, ms, ct))
    os.replace(tmp, path)


def _time(fn, reps):
    fn()  # warm-up (allocations, graph capture)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)) * 1e3


def sweep_estep(acc, prec, D, M, reps, log):
    """Returns (ll_frames, chunk_frames, ms, comp_tile) of the fastest E-step."""
    import gmm_amd
    best = None
    for ll in frame_tiles(prec, D):
        for ct in comp_tiles(prec, D):
            for cf in CHUNKS:
                acc.set_opts(precision=prec, ll_frames=ll, chunk_frames=cf, comp_tile=ct)
                ms = _time(acc.accumulate, reps)
                log("  estep %s D=%d M=%d ll=%d comp=%d chunk=%d: %.3f ms" % (
                    PREC_NAME[prec], D, M, ll, ct, cf, ms))
                if best is None or ms < best[2]:
                    best = (ll, cf, ms, ct)
    acc.set_opts(precision=gmm_amd.PREC_FP32, ll_frames=0, chunk_frames=0, comp_tile=0)
    return best


def sweep_batch(acc, prec, D, M, U, reps, seed, log):
    """Returns (ll_frames, chunk_frames, ms, comp_tile) of the fastest batched MAP+MLLR
    pass in one precision over U utterances (lengths lognormal(median 300,
    sigma 0.8) in [50, 3000])."""
    import torch
    import gmm_amd
    from gmm_amd import synth
    cfg = dict(synth.CONFIGS["batch_map"], D=D, M=M, U=U, seed=seed)
    w, mu, var = synth.true_mixture(**cfg)
    lens = synth.utterance_lengths(**cfg)
    x, off = synth.sample_utterances_torch(w, mu, var, lens, 0.3, seed)
    acc.set_model(w.astype(np.float32), mu, var)
    dev = "cuda"
    outs = {"means": torch.empty((U, M, D), dtype=torch.float32, device=dev).data_ptr(),
            "W": torch.empty((U, D, D + 1), dtype=torch.float64, device=dev).data_ptr(),
            "status": torch.empty(U, dtype=torch.int32, device=dev).data_ptr()}
    best = None
    for ll in frame_tiles(prec, D):
        for ct in comp_tiles(prec, D):
            acc.set_opts(precision=prec, ll_frames=ll, comp_tile=ct)
            for cf in BATCH_CHUNKS:
                def run():
                    acc.batch_adapt_device(x.data_ptr(), off, D, outs, tau=16, prior=True,
                                           chunk_frames=cf, graph=True)
                    torch.cuda.synchronize()
                ms = _time(run, reps)
                log("  batch %s D=%d M=%d U=%d ll=%d comp=%d chunk=%d: %.3f ms" % (
                    PREC_NAME[prec], D, M, U, ll, ct, cf, ms))
                if best is None or ms < best[2]:
                    best = (ll, cf, ms, ct)
    acc.set_opts(precision=PREC["fp32"], ll_frames=0, comp_tile=0)
    return best


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--D", default="60", help="comma-separated dimensions")
    ap.add_argument("--M", default="512,2048", help="comma-separated component counts")
    ap.add_argument("--prec", default="fp32,fp16x3,bf16x3")
    ap.add_argument("--frames", type=int, default=4_000_000)
    ap.add_argument("--batch", action="store_true",
                    help="also tune the batched path (every precision in --prec)")
    ap.add_argument("--utterances", type=int, default=2048)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=20261209)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    args = ap.parse_args(argv)

    import torch
    import gmm_amd
    from gmm_amd import synth

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    rows = read_table(args.out)
    acc = gmm_amd.Accumulator(0)
    for D in [int(v) for v in args.D.split(",")]:
        for M in [int(v) for v in args.M.split(",")]:
            cfg = dict(synth.CONFIGS["ubm_plain"], D=D, M=M, seed=args.seed + M)
            w, mu, var = synth.true_mixture(**cfg)
            x = synth.sample_frames_torch(w, mu, var, args.frames, args.seed + D)
            torch.cuda.synchronize()
            acc.set_model(w.astype(np.float32), mu, var)
            acc.set_frames_device(x.data_ptr(), args.frames, D)
            for pname in args.prec.split(","):
                prec = PREC[pname]
                ll, cf, ms, ct = sweep_estep(acc, prec, D, M, args.reps, log)
                rows[("estep", pname, D, M)] = (ll, cf, ms, ct)
                log("estep %s D=%d M=%d -> ll_frames=%d comp_tile=%d chunk=%d (%.3f ms)" % (
                    pname, D, M, ll, ct, cf, ms))
            del x
            torch.cuda.empty_cache()
            if args.batch:
                for pname in args.prec.split(","):
                    ll, cf, ms, ct = sweep_batch(acc, PREC[pname], D, M, args.utterances,
                                                 args.reps, args.seed, log)
                    rows[("batch", pname, D, M)] = (ll, cf, ms, ct)
                    log("batch %s D=%d M=%d -> ll_frames=%d comp_tile=%d chunk=%d (%.3f ms)" % (
                        pname, D, M, ll, ct, cf, ms))
            write_table(args.out, rows)
    acc.close()
    print(args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
