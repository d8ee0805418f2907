"""The drop-in tools on the GPU: trainGMM / getLogLike (C++ over the C ABI)
reproduce the reference's example outputs and the oracle's EM / MAP / MLLR
pipelines."""
import os
import subprocess

import numpy as np
import pytest

import _formats as fm
import _oracle as orc
import gmm_amd
from gmm_amd import host, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
BIN = os.path.join(os.path.dirname(HERE), "gmm_amd", "_build")


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if gmm_amd.device_count() < 1:
        pytest.skip("no GPU")


def run(tool, *args):
    r = subprocess.run([os.path.join(BIN, tool)] + [str(a) for a in args], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_getloglike_for_each_reproduces_example(tmp_path, golden):
    """getLogLike.bat:9 -- per-frame log-likelihoods of data.bin under
    myModel.gmm, TXT saveMat output."""
    run("getLogLike", "-i", os.path.join(GOLD, "data.bin"), "--data-T", 2, "--in-GMM",
        os.path.join(GOLD, "myModel.gmm"), "-L", "--for-each", "-o", tmp_path, "-T")
    out = fm.read_savemat_txt(tmp_path / "data.logL")
    assert out.shape == (1, 10000)
    ref = golden["ll"]
    # both sides are printed with 6 significant digits: one unit in the last
    # digit apart at most, plus the FP32 evaluation difference
    tol = 1.0 * 10.0 ** (np.floor(np.log10(np.abs(ref))) - 5) + 2e-6 * np.abs(ref)
    assert np.all(np.abs(out[0] - ref) <= tol)
    assert np.mean(out[0] == ref) > 0.9


def test_getloglike_mean(tmp_path, golden):
    outp = run("getLogLike", "-i", os.path.join(GOLD, "data.bin"), "--data-T", 2, "--in-GMM",
               os.path.join(GOLD, "myModel.gmm"), "-L")
    name, val = outp.strip().split("\t")
    assert name.endswith("data.bin")
    assert float(val) == pytest.approx(golden["ll"].mean(), rel=1e-5)


def test_train_from_scratch_reproduces_example(tmp_path, golden):
    """trainModel.bat:4 -- --type 0 --nummix 4 --it-EM 8 on data.bin gives
    myModel.gmm (reference SSE path, FP32 per-frame arithmetic)."""
    out = tmp_path / "myModel.gmm"
    run("trainGMM", "-i", os.path.join(GOLD, "data.bin"), "--data-T", 2, "--type", 0,
        "--nummix", 4, "--it-EM", 8, "-o", out, "-T")
    w, mu, var, _ = fm.read_gmm_txt(out)
    assert np.abs(w - golden["w"]).max() / golden["w"].max() < 5e-5
    assert np.abs(mu - golden["mu"]).max() < 5e-5
    assert (np.abs(var - golden["var"]) / golden["var"]).max() < 5e-5


def test_train_from_scratch_full_cov(tmp_path, golden):
    """trainModel.bat:6 -- --full-cov: the same run with full covariances
    gives the components of myModelFC.gmm in the same order, to the level at
    which the FP64 oracle's own split driver reproduces that file, and agrees
    more closely with that driver."""
    import _scratch_driver as sd
    out = tmp_path / "myModelFC.gmm"
    run("trainGMM", "-i", os.path.join(GOLD, "data.bin"), "--data-T", 2, "--type", 0,
        "--nummix", 4, "--it-EM", 8, "-o", out, "-T", "--full-cov")
    w, mu, cov, _ = fm.read_gmm_txt_full(out)
    rw, rmu, rcov, _ = fm.read_gmm_txt_full(os.path.join(GOLD, "myModelFC.gmm"))
    assert np.abs(w - rw).max() < 3e-4
    assert np.abs(mu - rmu).max() < 6e-3
    assert np.abs(cov - rcov).max() < 4e-3
    ow, omu, ocov = sd.train_from_scratch_full(golden["x"], 4, 8)
    assert np.abs(w - ow).max() < 2e-5
    assert np.abs(mu - omu).max() < 2e-4
    assert np.abs(cov - ocov).max() < 2e-4


def test_getloglike_full_cov_model(golden):
    """getLogLike.bat:8 -- mean log-likelihood of data.bin under
    myModelFC.gmm, against the oracle."""
    path = os.path.join(GOLD, "myModelFC.gmm")
    outp = run("getLogLike", "-i", os.path.join(GOLD, "data.bin"), "--data-T", 2, "--in-GMM",
               path, "-L")
    name, val = outp.strip().split("\t")
    w, mu, cov, _ = fm.read_gmm_txt_full(path)
    ref = orc.frame_loglik_full(golden["x"], w, mu, cov).mean()
    assert float(val) == pytest.approx(ref, rel=1e-5)


def _ubm_case(tmp_path, seed=41, M=64, D=12, N=20000):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M))
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.5, 2, (M, D)).astype(np.float32)
    # adaptation data from a shifted mixture
    x = synth.sample_frames_np(w, mu + rng.normal(0, 0.3, (M, D)).astype(np.float32), var, N,
                               seed + 1)
    ubm_p = tmp_path / "ubm.gmm"
    host.Model.create(w.astype(np.float32), mu, var).save(ubm_p, True)
    # the file keeps 7 digits: read it back for the oracle
    uw, umu, uvar, _ = fm.read_gmm_txt(ubm_p)
    data_p = tmp_path / "data.raw"
    fm.write_raw(data_p, x)
    return x, uw, umu, uvar, ubm_p, data_p


def test_train_em_from_ubm(tmp_path):
    x, w, mu, var, ubm_p, data_p = _ubm_case(tmp_path)
    out = tmp_path / "em.gmm"
    run("trainGMM", "-i", data_p, "--data-T", 2, "--in-UBM", ubm_p, "-L", "--type", 2,
        "--it-EM", 3, "--tau", 0, "--mean", "--var", "--weight", "-o", out, "-T")
    for _ in range(3):
        n, F, S, _ = orc.accumulate(x, w, mu, var)
        w, mu, var, _ = orc.update_em(w, mu, var, n, F, S, x.shape[0])
    gw, gmu, gvar, _ = fm.read_gmm_txt(out)
    assert np.allclose(gw, w, rtol=1e-4, atol=1e-7)
    assert np.allclose(gmu, mu, rtol=1e-4, atol=1e-4)
    assert np.allclose(gvar, var, rtol=1e-4, atol=1e-6)


def test_train_map_from_ubm(tmp_path):
    x, w, mu, var, ubm_p, data_p = _ubm_case(tmp_path, seed=43)
    out = tmp_path / "map.gmm"
    run("trainGMM", "-i", data_p, "--data-T", 2, "--in-UBM", ubm_p, "-L", "--type", 2,
        "--it-MAP", 1, "--tau", 16, "--mean", "-o", out, "-T")
    n, F, S, _ = orc.accumulate(x, w, mu, var)
    rw, rmu, rvar = orc.update_map(w, mu, var, n, F, S, x.shape[0], 16)
    gw, gmu, gvar, _ = fm.read_gmm_txt(out)
    assert np.allclose(gmu, rmu, rtol=1e-5, atol=1e-5)
    assert np.array_equal(gvar, var) and np.array_equal(gw, w)


def test_train_mllr_from_ubm(tmp_path):
    x, w, mu, var, ubm_p, data_p = _ubm_case(tmp_path, seed=45, M=64, D=8, N=30000)
    out = tmp_path / "mllr.gmm"
    run("trainGMM", "-i", data_p, "--data-T", 2, "--in-UBM", ubm_p, "-L", "--type", 1,
        "--apply-trn", "--save-trn", "-o", out, "-T")
    n, F, S, _ = orc.accumulate(x, w, mu, var)
    W = orc.mllr_transform(mu, var, n, F)
    lines = (tmp_path / "mllr.mat").read_text().split("\n")
    assert lines[0].strip() == "MLLR"
    Wg = np.array([[float(v) for v in ln.split()] for ln in lines[2:2 + 8]])
    assert np.allclose(Wg, W, rtol=1e-4, atol=1e-4)
    gw, gmu, gvar, _ = fm.read_gmm_txt(out)
    assert np.allclose(gmu, orc.mllr_apply(W, mu), rtol=1e-4, atol=1e-4)


def test_train_map_full_cov_ubm(tmp_path, golden):
    """MAP (weights, means and full covariances, tau = 16, two iterations)
    of data.bin from the reference's full-covariance myModelFC.gmm, against
    the oracle's full-covariance E-step and MAP rule applied the same way."""
    ubm_p = os.path.join(GOLD, "myModelFC.gmm")
    out = tmp_path / "mapfc.gmm"
    run("trainGMM", "-i", os.path.join(GOLD, "data.bin"), "--data-T", 2, "--in-UBM", ubm_p, "-L",
        "--type", 2, "--it-MAP", 2, "--tau", 16, "--mean", "--var", "--weight", "-o", out, "-T")
    uw, umu, ucov, _ = fm.read_gmm_txt_full(ubm_p)
    x = golden["x"]
    w, mu, cov = uw, umu, ucov
    for _ in range(2):
        n, F, S, _ = orc.accumulate_full(x, w, mu, cov)
        w, mu, cov = orc.update_map_full(uw, umu, ucov, n, F, S, x.shape[0], 16, upweight=True,
                                         upmean=True, upvar=True)
    gw, gmu, gcov, _ = fm.read_gmm_txt_full(out)
    assert np.allclose(gw, w, rtol=1e-5, atol=1e-7)
    assert np.abs(gmu - mu).max() < 1e-4 * max(1.0, np.abs(mu).max())
    assert np.abs(gcov - cov).max() < 1e-4 * max(1.0, np.abs(cov).max())


def test_train_mllr_full_cov_ubm(tmp_path):
    """Global MLLR from a full-covariance UBM: full-covariance posteriors,
    the diagonal of the covariances in the row statistics."""
    rng = np.random.default_rng(47)
    M, D, N = 40, 6, 30000
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    A = rng.normal(size=(M, D, D))
    cov = (0.3 * A @ A.transpose(0, 2, 1) / D + 0.5 * np.eye(D)).astype(np.float32)
    cov = 0.5 * (cov + cov.transpose(0, 2, 1))
    ubm_p = tmp_path / "ubmfc.gmm"
    host.Model.create_full(w, mu, cov).save(ubm_p, True)
    uw, umu, ucov, _ = fm.read_gmm_txt_full(ubm_p)
    k = rng.choice(M, size=N, p=uw.astype(np.float64) / uw.sum())
    L = np.linalg.cholesky(ucov.astype(np.float64))
    x = (1.1 * umu[k] + 0.4 + np.einsum("nij,nj->ni", L[k], rng.normal(size=(N, D)))).astype(np.float32)
    data_p = tmp_path / "data.raw"
    fm.write_raw(data_p, x)
    out = tmp_path / "mllrfc.gmm"
    run("trainGMM", "-i", data_p, "--data-T", 2, "--in-UBM", ubm_p, "-L", "--type", 1,
        "--apply-trn", "--save-trn", "-o", out, "-T")
    n, F, _, _ = orc.accumulate_full(x, uw, umu, ucov)
    var = np.ascontiguousarray(ucov[:, np.arange(D), np.arange(D)])
    W = orc.mllr_transform(umu, var, n, F)
    lines = (tmp_path / "mllrfc.mat").read_text().split("\n")
    Wg = np.array([[float(v) for v in ln.split()] for ln in lines[2:2 + D]])
    assert np.allclose(Wg, W, rtol=1e-4, atol=1e-4)
    gw, gmu, gcov, _ = fm.read_gmm_txt_full(out)
    assert np.allclose(gmu, orc.mllr_apply(W, umu), rtol=1e-4, atol=1e-4)
    assert np.array_equal(gcov, ucov)


@pytest.mark.parametrize("full", [False, True])
def test_train_from_scratch_robust(tmp_path, golden, full):
    """--robust (-R): from-scratch training on data.bin with the robust
    covariance estimate, against the oracle's split driver run the same way
    (auxiliary statistics from the oracle's posteriors, MakeRobustStats, the
    independent-frame test)."""
    import _scratch_driver as sd
    out = tmp_path / "robust.gmm"
    args = ["-i", os.path.join(GOLD, "data.bin"), "--data-T", 2, "--type", 0, "--nummix", 4,
            "--it-EM", 8, "-o", out, "-T", "--robust"]
    run("trainGMM", *(args + (["--full-cov"] if full else [])))
    if full:
        w, mu, cov, _ = fm.read_gmm_txt_full(out)
        ow, omu, ocov = sd.train_from_scratch_full(golden["x"], 4, 8, robust=True)
        assert np.abs(w - ow).max() < 2e-5
        assert np.abs(mu - omu).max() < 2e-4
        assert np.abs(cov - ocov).max() < 2e-4
    else:
        w, mu, var, _ = fm.read_gmm_txt(out)
        ow, omu, ovar = sd.train_from_scratch(golden["x"], 4, 8, robust=True)
        assert np.abs(w - ow).max() < 2e-5
        assert np.abs(mu - omu).max() < 5e-5
        assert (np.abs(var - ovar) / ovar).max() < 5e-5
        # the robust estimate differs from the plain one
        pw, pmu, pvar = sd.train_from_scratch(golden["x"], 4, 8)
        assert np.abs(var - pvar).max() > 1e-3


def test_train_map_robust(tmp_path):
    """MAP of means and variances with --robust: the auxiliary statistics
    of the adaptation frames make the second-order statistics robust before
    the MAP rule (OL_ModelAdapt.cpp:1037-1040)."""
    x, w, mu, var, ubm_p, data_p = _ubm_case(tmp_path, seed=49, M=16, D=6, N=8000)
    out = tmp_path / "maprobust.gmm"
    run("trainGMM", "-i", data_p, "--data-T", 2, "--in-UBM", ubm_p, "-L", "--type", 2,
        "--it-MAP", 1, "--tau", 16, "--mean", "--var", "--robust", "-o", out, "-T")
    n, F, S, _ = orc.accumulate(x, w, mu, var)
    _, g = orc.frame_loglik(x, w, mu, var, want_gamma=True)
    Sr, _ = orc.make_robust_stats(n, F, S, orc.accumulate_aux(x, g))
    rw, rmu, rvar = orc.update_map(w, mu, var, n, F, Sr, x.shape[0], 16, upmean=True, upvar=True)
    gw, gmu, gvar, _ = fm.read_gmm_txt(out)
    assert np.allclose(gmu, rmu, rtol=1e-5, atol=1e-5)
    assert np.allclose(gvar, rvar, rtol=1e-4, atol=1e-6)


def test_train_fmllr_from_ubm(tmp_path):
    """--type 3 (fMLLR): the saved feature transform against the oracle's
    fMLLR from the oracle's full second-order statistics under the UBM."""
    x, w, mu, var, ubm_p, data_p = _ubm_case(tmp_path, seed=51, M=32, D=6, N=20000)
    out = tmp_path / "fmllr.gmm"
    run("trainGMM", "-i", data_p, "--data-T", 2, "--in-UBM", ubm_p, "-L", "--type", 3,
        "-o", out, "-T")
    D = mu.shape[1]
    cov = np.zeros((len(w), D, D), np.float32)
    cov[:, np.arange(D), np.arange(D)] = var
    n, F, S, _ = orc.accumulate_full(x, w, mu, cov)
    W = orc.fmllr(mu, var, n, F, S, x.shape[0])
    lines = (tmp_path / "fmllr.mat").read_text().split("\n")
    assert lines[0].strip() == "fMLLR"
    Wg = np.array([[float(v) for v in ln.split()] for ln in lines[2:2 + D]])
    assert np.allclose(Wg, W, rtol=1e-4, atol=1e-4)
    assert not out.exists()  # fMLLR saves the transform only


@pytest.mark.parametrize("buff", ["0", "0.0002"])
def test_file_by_file_matches_resident(tmp_path, golden, buff):
    """--file-by-file (-f): data.bin split into 7 RAW files, read file by
    file (one file per group, or groups of --prm-buffer GB) at every EM
    iteration, gives the model of the resident run; --dwnsmp 2 keeps every
    second frame of every file."""
    x = golden["x"]
    parts = np.array_split(x, 7)
    files = []
    for i, p in enumerate(parts):
        fn = tmp_path / ("part%d.raw" % i)
        fm.write_raw(fn, p)
        files.append(fn)
    a_out, b_out = tmp_path / "a.gmm", tmp_path / "b.gmm"
    common = ["--data-T", 2, "--type", 0, "--nummix", 4, "--it-EM", 8, "-T"]
    run("trainGMM", *sum([["-i", f] for f in files], []), *common, "-o", a_out, "-f",
        "--prm-buffer", buff)
    run("trainGMM", "-i", os.path.join(GOLD, "data.bin"), *common, "-o", b_out)
    # FP32 partial sums group the frames differently (per file vs per
    # resident chunk): the models agree to a few units in the 6th digit
    wa, ma, va, _ = fm.read_gmm_txt(a_out)
    wb, mb, vb, _ = fm.read_gmm_txt(b_out)
    print("max rel diff w %.2e mu %.2e var %.2e" % (np.abs(wa / wb - 1).max(),
          (np.abs(ma - mb) / (np.abs(mb) + 1e-3)).max(), np.abs(va / vb - 1).max()))
    assert np.allclose(wa, wb, rtol=2e-5) and np.allclose(ma, mb, rtol=2e-5, atol=1e-5)
    assert np.allclose(va, vb, rtol=2e-5)
    # down-sampling per file, against the resident run on the kept frames
    c_out, d_out = tmp_path / "c.gmm", tmp_path / "d.gmm"
    run("trainGMM", *sum([["-i", f] for f in files], []), *common, "-o", c_out, "-f",
        "--prm-buffer", buff, "--dwnsmp", 2)
    kept = np.concatenate([p[::2] for p in parts])
    fm.write_raw(tmp_path / "kept.raw", kept)
    run("trainGMM", "-i", tmp_path / "kept.raw", *common, "-o", d_out)
    wc, mc, vc, _ = fm.read_gmm_txt(c_out)
    wd, md, vd, _ = fm.read_gmm_txt(d_out)
    assert np.allclose(wc, wd, rtol=2e-5) and np.allclose(mc, md, rtol=2e-5, atol=1e-5)
    assert np.allclose(vc, vd, rtol=2e-5)


def _clustered(seed=5, M=12, D=6, N=12000):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M) * 3)
    mu = rng.normal(0, 4, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 1.0, (M, D)).astype(np.float32)
    return synth.sample_frames_np(w, mu, var, N, 9)


def _cmp_models(path, ref, what):
    w, mu, var, _ = fm.read_gmm_txt(path)
    # the model file holds the components of non-zero weight only
    # (GMModel::WriteTXT, model/OL_GMModel.cpp:497-505)
    keep = ref[0] > 0
    rw, rmu, rvar = ref[0][keep], ref[1][keep], ref[2][keep]
    assert w.shape == rw.shape, (what, w.shape, rw.shape)
    dw = np.abs(w - rw).max() / rw.max()
    dmu = np.abs(mu - rmu).max() / np.abs(rmu).max()
    dvar = (np.abs(var - rvar) / rvar).max()
    print("%s: max diff w %.2e mu %.2e var %.2e" % (what, dw, dmu, dvar))
    # FP32 E-step against the FP64 oracle over ~30 EM iterations of
    # well-separated clusters (the file keeps 7 digits)
    assert dw < 1e-4 and dmu < 1e-4 and dvar < 1e-4, what


@pytest.mark.parametrize("split_x,split_th", [(0.5, 0.2), (1.0, 0.0)])
def test_train_from_scratch_split_x(tmp_path, split_x, split_th):
    """--split-X / --split-th (splitMultipleGaussians, OL_ModelAdapt.cpp:
    590-611): several components split per round, in GetGaussStats' weight
    order, against the oracle's split driver."""
    import _scratch_driver as sd
    x = _clustered()
    fm.write_raw(tmp_path / "x.raw", x)
    out = tmp_path / "m.gmm"
    run("trainGMM", "-i", tmp_path / "x.raw", "--data-T", 2, "--type", 0, "--nummix", 16,
        "--it-EM", 4, "--split-X", split_x, "--split-th", split_th, "-o", out, "-T")
    hist = []
    ref = sd.train_from_scratch(x, 16, 4, split_x=split_x, split_th=split_th, history=hist)
    # more than one split per round happened
    assert len(hist) < 16 and hist[-1][2] == 16, hist
    _cmp_models(out, ref, "split-X %g" % split_x)


def test_train_from_scratch_ll_threshold_and_backups(tmp_path):
    """--ll-Th (OL_ModelAdapt.cpp:781-797) stops the splitting once a round
    gains less than the threshold over the better of the two rounds before;
    --tmp-N writes OUT.itNNNNN.GNNNNN.gmm.bckup every N rounds.  The oracle's
    driver gives per-round gains of >= 0.17 up to round 13 and 4e-4 at round
    14, so a threshold of 0.05 stops at round 14 (15 components) with a
    wide margin on both sides."""
    import _scratch_driver as sd
    x = _clustered()
    fm.write_raw(tmp_path / "x.raw", x)
    out = tmp_path / "m.gmm"
    run("trainGMM", "-i", tmp_path / "x.raw", "--data-T", 2, "--type", 0, "--nummix", 32,
        "--it-EM", 3, "--ll-Th", 0.05, "--tmp-N", 5, "-o", out, "-T")
    hist = []
    ref = sd.train_from_scratch(x, 32, 3, ll_th=0.05, history=hist)
    assert hist[-1][0] == 14 and (ref[0] > 0).sum() == 15, hist
    _cmp_models(out, ref, "ll-Th")
    for it, _, nz in hist:
        name = "%s.it%05d.G%05d.gmm.bckup" % (out, it, nz)
        assert os.path.exists(name) == ((it + 1) % 5 == 0 and it < 14), name


def _mlf_case(tmp_path, kind):
    """Three feature files (two under a sub-directory), an MLF labelling
    their frames with 'sil', 'a' and 'b', and a classes file."""
    x = _clustered(N=1500)
    parts = {"utt1": x[:400], "sub/utt2": x[400:1000], "sub/utt3": x[1000:]}
    data = tmp_path / "data"
    (data / "sub").mkdir(parents=True)
    ext = ".htk" if kind == host.HTK else ".prm"
    for name, p in parts.items():
        host.save_features(data / (name + ext), kind, p, sample_period=100000)
    # label times in 100 ns: frame k starts at k * 100000
    labels = {
        "utt1": [(0, 50, "sil"), (50, 230, "a"), (230, 400, "b")],
        "sub/utt2": [(0, 120, "b"), (120, 130, "sil"), (130, 600, "a")],
        "sub/utt3": [(0, 90, "a"), (90, 500, "sil")],
    }
    lines = ["#!MLF!#"]
    for name, labs in labels.items():
        lines.append('"*/%s.lab"' % name)
        lines += ["%d %d %s" % (b * 100000, e * 100000, lab) for b, e, lab in labs]
        lines.append(".")
    (tmp_path / "labels.mlf").write_text("\n".join(lines) + "\n")
    (tmp_path / "classes.txt").write_text("speech a b\nsilence sil\nnothing zz\n")
    # the frames of each class, files in path order, intervals in time order
    want = {}
    for cls, elems in (("speech", ("a", "b")), ("silence", ("sil",))):
        rows = []
        for name in sorted(parts, key=lambda n: str(data / (n + ext))):
            for b, e, lab in labels[name]:
                if lab in elems:
                    rows.append(parts[name][b:e])
        want[cls] = np.concatenate(rows)
    return data, ext, want


@pytest.mark.parametrize("kind", [host.HTK, host.SVES])
@pytest.mark.parametrize("by_file", [False, True])
def test_mlf_mode_trains_one_model_per_class(tmp_path, kind, by_file):
    """MLF mode (projects/Model_ADAPT/Model_ADAPT.cpp:97-240, :478-513;
    tools/OL_MLFParser.cpp:57-163): one model per class from the labelled
    intervals of its elements, resident or read file by file at every
    accumulation through the file-to-ranges accumulateStatsMT
    (OL_GMMStatsEstimator.cpp:712-780).  Each model is the one trainGMM
    trains from a RAW file of exactly that class's frames, byte for byte;
    a class without labels gets no model."""
    data, ext, want = _mlf_case(tmp_path, kind)
    out = tmp_path / "models"
    out.mkdir()
    common = ["--type", 0, "--nummix", 4, "--it-EM", 5, "-T"]
    args = ["-S", tmp_path / "classes.txt", "-i", tmp_path / "labels.mlf", "-e", ext, "-I", data,
            "--data-T", kind, "-D", out, "-x", ".gmm"] + common
    if kind != host.HTK:
        args += ["--sample-period", 100000]
    if by_file:
        args += ["-f"]
    run("trainGMM", *args)
    assert sorted(p.name for p in out.iterdir()) == ["silence.gmm", "speech.gmm"]
    for cls, x in want.items():
        fm.write_raw(tmp_path / (cls + ".raw"), x)
        ref = tmp_path / (cls + "_ref.gmm")
        run("trainGMM", "-i", tmp_path / (cls + ".raw"), "--data-T", 2, "-o", ref, *common)
        assert (out / (cls + ".gmm")).read_bytes() == ref.read_bytes(), cls


def test_class_mode_trains_from_whole_files(tmp_path):
    """CLASS mode: 'CLASS file1 file2 ...' trains each class from whole
    files dir/fileK.ext."""
    data, ext, _ = _mlf_case(tmp_path, host.HTK)
    (tmp_path / "cls.txt").write_text("one utt1\ntwo sub/utt2 sub/utt3\n")
    out = tmp_path / "models"
    out.mkdir()
    common = ["--type", 0, "--nummix", 2, "--it-EM", 4, "-T"]
    run("trainGMM", "-S", tmp_path / "cls.txt", "-e", ext, "-I", data, "--data-T", 1, "-D", out,
        *common)
    x = _clustered(N=1500)
    for cls, rows in (("one", x[:400]), ("two", x[400:])):
        fm.write_raw(tmp_path / (cls + ".raw"), rows)
        ref = tmp_path / (cls + "_ref.gmm")
        run("trainGMM", "-i", tmp_path / (cls + ".raw"), "--data-T", 2, "-o", ref, *common)
        assert (out / (cls + ".gmm")).read_bytes() == ref.read_bytes(), cls
