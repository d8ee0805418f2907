"""Seeded synthetic workloads of the benchmark configurations.

Frames are drawn from a seeded "true" diagonal mixture; the model under test
is initialised from random frames with the global variance and uniform
weights (the "random initialisation then EM" recipe of GMM/README.md:94-101).

  plain_em_small : D=39, M=64,   N=100,000    means~N(0,3^2), var~U[0.5,2],
                                              w~Dirichlet(1)
  ubm_plain      : D=60, M=512,  N=10,000,000 means~N(0,2^2), var~U[0.3,3],
                                              w~Dirichlet(0.5)
  mixed_estep    : D=60, M=2048, N=20,000,000 as ubm_plain, frame offsets
                                              ~N(5,1) per dim
"""
import numpy as np

CONFIGS = {
    "plain_em_small": dict(D=39, M=64, N=100_000, mean_sd=3.0, var_lo=0.5, var_hi=2.0,
                           alpha=1.0, offset=False, seed=20261204),
    "ubm_plain": dict(D=60, M=512, N=10_000_000, mean_sd=2.0, var_lo=0.3, var_hi=3.0,
                      alpha=0.5, offset=False, seed=20261205),
    "mixed_estep": dict(D=60, M=2048, N=20_000_000, mean_sd=2.0, var_lo=0.3, var_hi=3.0,
                        alpha=0.5, offset=True, seed=20261206),
    # batched MAP + MLLR: 4,096 utterances, lengths lognormal(ln 300, 0.8)
    # clipped to [50, 3000], each from the UBM mixture with its means shifted
    # by N(0, 0.3^2)
    "batch_map": dict(D=60, M=2048, U=4096, mean_sd=2.0, var_lo=0.3, var_hi=3.0, alpha=0.5,
                      offset=False, len_median=300, len_sigma=0.8, len_lo=50, len_hi=3000,
                      shift_sd=0.3, seed=20261207),
}


def utterance_lengths(U, len_median, len_sigma, len_lo, len_hi, seed, **_):
    rng = np.random.default_rng(seed + 3)
    n = np.exp(rng.normal(np.log(len_median), len_sigma, U))
    return np.clip(np.round(n), len_lo, len_hi).astype(np.int64)


def sample_utterances_torch(w, mu, var, lens, shift_sd, seed, device="cuda", group=64):
    """Concatenated frames of len(lens) utterances; utterance u is drawn from
    the mixture with every mean shifted by its own N(0, shift_sd^2) offset."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    M, D = mu.shape
    wt = torch.as_tensor(w, dtype=torch.float64, device=device)
    mut = torch.as_tensor(mu, dtype=torch.float32, device=device)
    sdt = torch.as_tensor(np.sqrt(var), dtype=torch.float32, device=device)
    N = int(np.sum(lens))
    x = torch.empty((N, D), dtype=torch.float32, device=device)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    for u0 in range(0, len(lens), group):
        u1 = min(len(lens), u0 + group)
        shift = torch.randn((u1 - u0, M, D), generator=g, device=device) * shift_sd
        for u in range(u0, u1):
            n = int(lens[u])
            k = torch.multinomial(wt, n, replacement=True, generator=g)
            z = torch.randn((n, D), generator=g, device=device)
            x[off[u]:off[u + 1]] = mut[k] + shift[u - u0, k] + sdt[k] * z
    return x, off.astype(np.uint64)


def true_mixture(D, M, mean_sd, var_lo, var_hi, alpha, offset, seed, **_):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.full(M, alpha)).astype(np.float64)
    w = np.maximum(w, 1e-12)
    w /= w.sum()
    mu = rng.normal(0.0, mean_sd, size=(M, D))
    if offset:
        mu = mu + rng.normal(5.0, 1.0, size=(1, D))
    var = rng.uniform(var_lo, var_hi, size=(M, D))
    return w, mu.astype(np.float32), var.astype(np.float32)


def sample_frames_np(w, mu, var, N, seed):
    """numpy sampler (CPU, small N)."""
    rng = np.random.default_rng(seed)
    M, D = mu.shape
    k = rng.choice(M, size=N, p=w)
    x = mu[k] + np.sqrt(var[k]) * rng.standard_normal((N, D)).astype(np.float32)
    return np.ascontiguousarray(x, np.float32)


def sample_frames_torch(w, mu, var, N, seed, device="cuda", chunk=2_000_000):
    """torch sampler that writes the frames directly into device memory."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    M, D = mu.shape
    wt = torch.as_tensor(w, dtype=torch.float64, device=device)
    mut = torch.as_tensor(mu, dtype=torch.float32, device=device)
    sdt = torch.as_tensor(np.sqrt(var), dtype=torch.float32, device=device)
    x = torch.empty((N, D), dtype=torch.float32, device=device)
    for s in range(0, N, chunk):
        e = min(N, s + chunk)
        k = torch.multinomial(wt, e - s, replacement=True, generator=g)
        z = torch.randn((e - s, D), generator=g, device=device, dtype=torch.float32)
        x[s:e] = mut[k] + sdt[k] * z
    return x


def init_model_from_frames(x, M, seed):
    """M random frames as means, the global variance for every component,
    uniform weights.  x: numpy [N, D] or torch tensor."""
    rng = np.random.default_rng(seed + 1)
    N = x.shape[0]
    idx = np.sort(rng.choice(N, size=M, replace=False))
    try:
        import torch
        if isinstance(x, torch.Tensor):
            mu = x[torch.as_tensor(idx, device=x.device)].double().cpu().numpy()
            xd = x.double()
            gm = xd.mean(0)
            gv = (xd * xd).mean(0) - gm * gm
            gv = gv.cpu().numpy()
        else:
            raise ImportError
    except ImportError:
        mu = np.asarray(x[idx], np.float64)
        xd = np.asarray(x, np.float64)
        gm = xd.mean(0)
        gv = (xd * xd).mean(0) - gm * gm
    var = np.tile(gv, (M, 1))
    w = np.full(M, 1.0 / M)
    return (w.astype(np.float32), mu.astype(np.float32), var.astype(np.float32))
