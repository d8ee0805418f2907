"""gmmhip_accumulate_noalloc (the entry point for callers that cannot
allocate) against gmmhip_accumulate on the same device frames: n, F and S
bit-identical, the total log-likelihood to FP64 rounding (1e-12 relative),
over two calls in a row on one handle, the second with a smaller input and
then with another model, with frame sets long enough to take several
scratch segments.  Sizes beyond the scratch maximum are rejected."""
import ctypes

import numpy as np
import pytest

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


def noalloc(acc, x, out=None, flags=gmm_amd.ACC_MEAN | gmm_amd.ACC_VAR):
    import torch
    N, D = x.shape
    M = acc.M
    if out is None:
        out = torch.zeros(M + 2 * M * D + 1, dtype=torch.float64, device="cuda")
    n, F, S, tot = out[:M], out[M:M + M * D], out[M + M * D:M + 2 * M * D], out[-1:]
    nf = ctypes.c_uint64(0)
    rc = gmm_amd.lib().gmmhip_accumulate_noalloc(acc.handle, x.data_ptr(), N, D, flags,
                                                 n.data_ptr(), F.data_ptr(), S.data_ptr(),
                                                 tot.data_ptr(), ctypes.byref(nf))
    if rc:
        raise gmm_amd.GmmHipError(rc, gmm_amd.lib().gmmhip_last_error(acc.handle).decode())
    acc.sync()
    return (n.cpu().numpy(), F.cpu().numpy().reshape(M, D), S.cpu().numpy().reshape(M, D),
            float(tot.item()), nf.value), out


def reference(acc, x):
    N, D = x.shape
    acc.set_frames_device(x.data_ptr(), N, D)
    return acc.accumulate()


def same(got, ref, what):
    for i, k in enumerate("nFS"):
        assert np.array_equal(got[i], ref[i]), (what, k, np.abs(got[i] - ref[i]).max())
    assert got[3] == pytest.approx(ref[3], rel=1e-12), what
    assert got[4] == ref[4]


def frames(cfg, N, seed):
    w0, mu0, var0 = synth.true_mixture(**cfg)
    return synth.sample_frames_torch(w0, mu0, var0, N, seed)


def test_noalloc_matches_accumulate(acc):
    import torch
    cfg = dict(synth.CONFIGS["mixed_estep"])
    x1 = frames(cfg, 400_000, 31)
    w, mu, var = synth.init_model_from_frames(x1, cfg["M"], cfg["seed"])
    acc.set_model(w, mu, var)
    print("noalloc: D=60 M=2048 N=400000", flush=True)
    ref1 = reference(acc, x1)
    got1, _ = noalloc(acc, x1)
    same(got1, ref1, "first call")
    # second call, smaller input, same handle and model
    x2 = frames(cfg, 50_000 + 37, 32)
    print("noalloc: D=60 M=2048 N=50037", flush=True)
    ref2 = reference(acc, x2)
    got2, out2 = noalloc(acc, x2)
    same(got2, ref2, "second call")
    # adding into the same outputs doubles them exactly
    got2b, _ = noalloc(acc, x2, out=out2)
    assert np.array_equal(got2b[0], 2 * got2[0])
    # another model and dimension
    c1 = dict(synth.CONFIGS["plain_em_small"])
    x3 = frames(c1, 1_300_000, 33)
    w3, mu3, var3 = synth.init_model_from_frames(x3, c1["M"], c1["seed"])
    acc.set_model(w3, mu3, var3)
    print("noalloc: D=39 M=64 N=1300000", flush=True)
    ref3 = reference(acc, x3)
    got3, _ = noalloc(acc, x3)
    same(got3, ref3, "third call")
    del x1, x2, x3
    torch.cuda.synchronize()


def test_noalloc_rejects_sizes_beyond_the_scratch(acc):
    import torch
    rng = np.random.default_rng(5)
    M, D = 4096, 127
    acc.set_model(np.full(M, 1.0 / M, np.float32), rng.normal(0, 1, (M, D)).astype(np.float32),
                  np.ones((M, D), np.float32))
    x = torch.zeros((256, D), dtype=torch.float32, device="cuda")
    with pytest.raises(gmm_amd.GmmHipError) as ei:
        noalloc(acc, x)
    assert ei.value.code == gmm_amd.ERR_ARG
    assert "scratch maximum" in str(ei.value)
