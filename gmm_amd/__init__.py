"""gmm_amd -- MI355X (gfx950) diagonal-GMM EM accumulator.

The product is the native library gmm_amd/_build/libgmmhip.so (HIP kernels,
C ABI declared in include/gmmhip.h, C++ host layer).  This module is a thin
ctypes binding of that C ABI used by the tests and bench.py; it mirrors the
reference estimator's calls (insertModel, accumulateStatsMT,
compVecsLogLikeMT of trainer/OL_GMMStatsEstimator_CUDA.h:43-56).

There is no CPU fallback: if the library is missing or no GPU is present,
the calls raise.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "_build", "libgmmhip.so")

EXPORTS = [
    "gmmhip_default_opts", "gmmhip_device_count", "gmmhip_create",
    "gmmhip_destroy", "gmmhip_set_opts", "gmmhip_last_error",
    "gmmhip_set_model", "gmmhip_set_model_full", "gmmhip_set_frames", "gmmhip_set_frames_contig",
    "gmmhip_set_frames_device", "gmmhip_accumulate", "gmmhip_accumulate_aux", "gmmhip_set_mask",
    "gmmhip_posteriors", "gmmhip_loglikes",
    "gmmhip_loglikes_device", "gmmhip_stream", "gmmhip_sync",
    "gmmhip_kernel_times", "gmmhip_reset_kernel_times", "gmmhip_batch_adapt",
    "gmmhip_batch_graph_captures", "gmmhip_accumulate_noalloc",
    "gmmhip_server_create", "gmmhip_server_submit", "gmmhip_server_wait",
    "gmmhip_server_drain", "gmmhip_server_stats_get", "gmmhip_server_last_error",
    "gmmhip_server_destroy", "gmmhip_tune_load", "gmmhip_tune_lookup", "gmmhip_tiles_used",
]
TUNE_ESTEP, TUNE_BATCH = 0, 1
BATCH_QUEUE, BATCH_PER_UTTERANCE = 0, 1
KT_LOGLIK, KT_STATS, KT_REDUCE, KT_COUNT = 0, 1, 2, 3
# host-layer entry points (include/gmmhost.h), bound in gmm_amd.host
EXTRA_EXPORTS = [
    "gmmhost_model_load", "gmmhost_model_save", "gmmhost_model_create",
    "gmmhost_model_create_full", "gmmhost_model_full", "gmmhost_model_dims", "gmmhost_model_get", "gmmhost_model_rearrange",
    "gmmhost_model_free", "gmmhost_update_em", "gmmhost_update_em_robust",
    "gmmhost_make_robust_stats", "gmmhost_update_map",
    "gmmhost_mllr_transform", "gmmhost_fmllr_transform", "gmmhost_mllr_apply",
    "gmmhost_last_error", "gmmhost_features_load", "gmmhost_features_load_ranges",
    "gmmhost_features_dims", "gmmhost_features_get", "gmmhost_features_free",
    "gmmhost_features_save",
]

(OK, ERR_HIP, ERR_OOM, ERR_NAN, ERR_ARG, ERR_VAR, ERR_STATE, ERR_MLLR, ERR_NODEV, ERR_TIMEOUT,
 ERR_CLOSED) = range(11)
REQ_SCORE, REQ_MAP, REQ_MLLR = 0, 1, 2
REQ_PENDING = -1
PREC_FP32, PREC_BF16X3, PREC_FP16X3 = 0, 1, 2
ACC_MEAN, ACC_VAR, ACC_FULLVAR = 1, 2, 4


class GmmHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gmmhip error %d: %s" % (code, msg))
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [("min_loglike", ctypes.c_float), ("min_gamma", ctypes.c_float),
                ("min_var", ctypes.c_float), ("precision", ctypes.c_int),
                ("chunk_frames", ctypes.c_int), ("verbosity", ctypes.c_int),
                ("profile", ctypes.c_int), ("ll_frames", ctypes.c_int),
                ("comp_tile", ctypes.c_int)]


class BatchCfg(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_int), ("mllr_prior", ctypes.c_int), ("mode", ctypes.c_int),
                ("chunk_frames", ctypes.c_int), ("graph", ctypes.c_int),
                ("outputs_on_device", ctypes.c_int)]


class BatchOut(ctypes.Structure):
    _fields_ = [("means", ctypes.c_void_p), ("occ", ctypes.c_void_p), ("first", ctypes.c_void_p),
                ("mllr_W", ctypes.c_void_p), ("mllr_status", ctypes.c_void_p)]


class ServerCfg(ctypes.Structure):
    _fields_ = [("opts", Opts), ("slots", ctypes.c_int), ("slot_bytes", ctypes.c_uint64),
                ("max_batch", ctypes.c_int), ("chunk_frames", ctypes.c_int),
                ("tau", ctypes.c_int), ("mllr_prior", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("frames", ctypes.c_void_p),
                ("n_frames", ctypes.c_uint64), ("frame_ll", ctypes.c_void_p),
                ("means", ctypes.c_void_p), ("occ", ctypes.c_void_p),
                ("mllr_W", ctypes.c_void_p), ("ll_sum", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("status", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("submitted", "completed", "failed", "batches", "frames", "fp32_fallbacks")] + \
               [(n, ctypes.c_double) for n in
                ("p50_ms", "p90_ms", "p99_ms", "p999_ms", "max_ms", "mean_ms",
                 "mean_batch_requests", "mean_batch_gpu_ms", "wall_s")] + \
               [("precision", ctypes.c_int)]


_lib = None


def lib():
    """Load libgmmhip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libgmmhip.so not built: run `python -m gmm_amd.build` "
                              "(expected at %s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        c_u64, c_dp, c_fp, c_vp = (ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_float), ctypes.c_void_p)
        L.gmmhip_last_error.restype = ctypes.c_char_p
        L.gmmhip_last_error.argtypes = [c_vp]
        L.gmmhip_create.argtypes = [ctypes.c_int, ctypes.POINTER(Opts), ctypes.POINTER(c_vp)]
        L.gmmhip_destroy.argtypes = [c_vp]
        L.gmmhip_set_opts.argtypes = [c_vp, ctypes.POINTER(Opts)]
        L.gmmhip_set_model.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, c_fp, c_fp, c_fp]
        L.gmmhip_set_frames.argtypes = [c_vp, ctypes.POINTER(c_fp), c_u64, ctypes.c_int]
        L.gmmhip_set_frames_contig.argtypes = [c_vp, c_fp, c_u64, ctypes.c_int, c_u64]
        L.gmmhip_set_frames_device.argtypes = [c_vp, c_vp, c_u64, ctypes.c_int, c_u64]
        L.gmmhip_accumulate.argtypes = [c_vp, ctypes.c_int, c_dp, c_dp, c_dp, c_dp,
                                        ctypes.POINTER(c_u64)]
        L.gmmhip_accumulate_aux.argtypes = [c_vp, c_dp]
        L.gmmhip_set_mask.argtypes = [c_vp, c_u64, ctypes.POINTER(c_u64),
                                      ctypes.POINTER(ctypes.c_int)]
        L.gmmhip_posteriors.argtypes = [c_vp, c_fp, c_fp]
        L.gmmhip_loglikes.argtypes = [c_vp, c_fp]
        L.gmmhip_loglikes_device.argtypes = [c_vp, c_vp]
        L.gmmhip_stream.restype = c_vp
        L.gmmhip_stream.argtypes = [c_vp]
        L.gmmhip_sync.argtypes = [c_vp]
        L.gmmhip_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.gmmhip_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.gmmhip_default_opts.restype = None
        L.gmmhip_kernel_times.argtypes = [c_vp, c_dp, ctypes.POINTER(c_u64), ctypes.c_int]
        L.gmmhip_batch_adapt.argtypes = [c_vp, c_vp, ctypes.c_int, ctypes.POINTER(c_u64),
                                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(BatchCfg),
                                         ctypes.POINTER(BatchOut)]
        L.gmmhip_reset_kernel_times.argtypes = [c_vp]
        L.gmmhip_batch_graph_captures.argtypes = [c_vp, ctypes.POINTER(c_u64)]
        L.gmmhip_accumulate_noalloc.argtypes = [c_vp, c_vp, c_u64, ctypes.c_int, ctypes.c_int,
                                                c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_u64)]
        L.gmmhip_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerCfg), ctypes.c_int,
                                           ctypes.c_int, c_fp, c_fp, c_fp, ctypes.POINTER(c_vp)]
        L.gmmhip_server_submit.argtypes = [c_vp, ctypes.POINTER(Request)]
        L.gmmhip_server_wait.argtypes = [c_vp, ctypes.POINTER(Request), ctypes.c_double]
        L.gmmhip_server_drain.argtypes = [c_vp]
        L.gmmhip_server_stats_get.argtypes = [c_vp, ctypes.POINTER(ServerStats), ctypes.c_int]
        L.gmmhip_server_last_error.restype = ctypes.c_char_p
        L.gmmhip_server_last_error.argtypes = [c_vp]
        L.gmmhip_server_destroy.argtypes = [c_vp]
        L.gmmhip_server_destroy.restype = None
        L.gmmhip_tune_load.argtypes = [ctypes.c_char_p]
        L.gmmhip_tune_lookup.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)] * 3
        L.gmmhip_tiles_used.argtypes = [c_vp] + [ctypes.POINTER(ctypes.c_int)] * 3
        _lib = L
    return _lib


def default_opts(**kw):
    o = Opts()
    lib().gmmhip_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def tune_load(path=None):
    """(Re)load the tuning table; returns the number of entries or -1."""
    return lib().gmmhip_tune_load(path.encode() if path else None)


def tune_lookup(path, precision, D, M):
    """(ll_frames, chunk_frames, comp_tile) the tuning table picks, or None."""
    a, b, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    rc = lib().gmmhip_tune_lookup(path, precision, D, M, ctypes.byref(a), ctypes.byref(b),
                                  ctypes.byref(c))
    return (a.value, b.value, c.value) if rc == 0 else None


def device_count():
    n = ctypes.c_int(0)
    lib().gmmhip_device_count(ctypes.byref(n))
    return n.value


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Accumulator:
    """One handle of the C ABI (one GPU, one stream)."""

    def __init__(self, device=0, **opts):
        self._h = ctypes.c_void_p()
        self._opts = default_opts(**opts)
        self._check(lib().gmmhip_create(device, ctypes.byref(self._opts), ctypes.byref(self._h)))
        self.M = self.D = 0
        self.N = 0
        self._keep = None

    def _check(self, rc):
        if rc != OK:
            msg = lib().gmmhip_last_error(self._h).decode() if self._h else ""
            raise GmmHipError(rc, msg)

    def close(self):
        if self._h:
            lib().gmmhip_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_opts(self, **kw):
        for k, v in kw.items():
            setattr(self._opts, k, v)
        self._check(lib().gmmhip_set_opts(self._h, ctypes.byref(self._opts)))

    def set_model(self, w, mu, var):
        w = np.ascontiguousarray(w, np.float32)
        mu = np.ascontiguousarray(mu, np.float32)
        var = np.ascontiguousarray(var, np.float32)
        M, D = mu.shape
        self._check(lib().gmmhip_set_model(self._h, M, D, _fp(w), _fp(mu), _fp(var)))
        self.M, self.D = M, D
        self.full = False

    def set_model_full(self, w, mu, cov):
        """Full-covariance model: cov[M, D, D] symmetric positive definite.
        accumulate() then returns S[M, D(D+1)/2], the packed upper triangle
        of sum g x x^T (row k holds columns k..D-1)."""
        w = np.ascontiguousarray(w, np.float32)
        mu = np.ascontiguousarray(mu, np.float32)
        cov = np.ascontiguousarray(cov, np.float32)
        M, D = mu.shape
        self._check(lib().gmmhip_set_model_full(self._h, M, D, _fp(w), _fp(mu), _fp(cov)))
        self.M, self.D = M, D
        self.full = True

    def set_frames(self, x):
        x = np.ascontiguousarray(x, np.float32)
        N, D = x.shape
        self._check(lib().gmmhip_set_frames_contig(self._h, _fp(x), N, D, D))
        self.N = N

    def set_frames_rows(self, rows):
        """rows: list of float32 1-D arrays (the reference's float** frames)."""
        arrs = [np.ascontiguousarray(r, np.float32) for r in rows]
        ptrs = (ctypes.POINTER(ctypes.c_float) * len(arrs))(*[_fp(a) for a in arrs])
        self._check(lib().gmmhip_set_frames(self._h, ptrs, len(arrs), arrs[0].shape[0]))
        self.N = len(arrs)

    def set_frames_device(self, ptr, N, D, ld=None):
        self._check(lib().gmmhip_set_frames_device(self._h, ctypes.c_void_p(ptr), N, D,
                                                   ld if ld is not None else D))
        self.N = N

    def accumulate(self, flags=ACC_MEAN | ACC_VAR, into=None):
        """Returns (n[M], F[M,D], S, tot_ll, nframes), S[M,D] for a diagonal
        model, S[M,D(D+1)/2] (packed upper) for a full one or with
        ACC_FULLVAR; adds into `into` if given."""
        if into is None:
            n = np.zeros(self.M)
            F = np.zeros((self.M, self.D))
            packed = getattr(self, "full", False) or (flags & ACC_FULLVAR)
            S = np.zeros((self.M, self.D * (self.D + 1) // 2 if packed else self.D))
            tot = ctypes.c_double(0.0)
            nf = ctypes.c_uint64(0)
        else:
            n, F, S, tot0, nf0 = into
            tot = ctypes.c_double(tot0)
            nf = ctypes.c_uint64(nf0)
        self._check(lib().gmmhip_accumulate(self._h, flags, _dp(n), _dp(F), _dp(S),
                                            ctypes.byref(tot), ctypes.byref(nf)))
        return n, F, S, tot.value, nf.value

    def accumulate_aux(self, into=None):
        """Robust auxiliary statistics [M, D+2] = [aux (D) | aux2 | aux3] of
        the resident frames; adds into `into` if given."""
        aux = np.zeros((self.M, self.D + 2)) if into is None else into
        self._check(lib().gmmhip_accumulate_aux(self._h, _dp(aux)))
        return aux

    def set_mask(self, lists):
        """Gaussian selection: lists[t] = component indices of frame t (None
        removes the mask)."""
        if lists is None:
            self._check(lib().gmmhip_set_mask(self._h, 0, None, None))
            return
        off = np.zeros(len(lists) + 1, np.uint64)
        off[1:] = np.cumsum([len(l) for l in lists])
        idx = np.ascontiguousarray(np.concatenate([np.asarray(l, np.int32) for l in lists])
                                   if off[-1] else np.zeros(1, np.int32), np.int32)
        self._check(lib().gmmhip_set_mask(
            self._h, len(lists), off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def posteriors(self):
        """(ll[N], post[N, M]): per-frame log-likelihoods and posteriors
        before the cut."""
        ll = np.zeros(self.N, np.float32)
        post = np.zeros((self.N, self.M), np.float32)
        self._check(lib().gmmhip_posteriors(self._h, _fp(ll), _fp(post)))
        return ll, post

    def loglikes(self):
        out = np.zeros(self.N, np.float32)
        self._check(lib().gmmhip_loglikes(self._h, _fp(out)))
        return out

    def loglikes_device(self, ptr):
        self._check(lib().gmmhip_loglikes_device(self._h, ctypes.c_void_p(ptr)))

    def kernel_times(self):
        """{class: (total_ms, launches)} of a handle created with profile=1."""
        ms = np.zeros(KT_COUNT)
        cnt = np.zeros(KT_COUNT, np.uint64)
        self._check(lib().gmmhip_kernel_times(
            self._h, _dp(ms), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), KT_COUNT))
        return {"loglik": (ms[0], int(cnt[0])), "stats": (ms[1], int(cnt[1])),
                "reduce": (ms[2], int(cnt[2]))}

    def reset_kernel_times(self):
        self._check(lib().gmmhip_reset_kernel_times(self._h))

    def batch_adapt(self, x, offsets, tau=16, means=True, stats=True, mllr=True, prior=False,
                    mode=BATCH_QUEUE, chunk_frames=0, graph=False):
        """Batched per-utterance statistics / MAP means / MLLR transforms of
        host frames x[offsets[-1], D] split by offsets[U+1].  Returns a dict
        of numpy arrays."""
        x = np.ascontiguousarray(x, np.float32)
        off = np.ascontiguousarray(offsets, np.uint64)
        U = len(off) - 1
        M, D = self.M, self.D
        res = {}
        o = BatchOut()
        if means:
            res["means"] = np.zeros((U, M, D), np.float32)
            o.means = res["means"].ctypes.data
        if stats:
            res["occ"] = np.zeros((U, M))
            res["first"] = np.zeros((U, M, D))
            o.occ = res["occ"].ctypes.data
            o.first = res["first"].ctypes.data
        if mllr:
            res["W"] = np.zeros((U, D, D + 1))
            res["status"] = np.zeros(U, np.int32)
            o.mllr_W = res["W"].ctypes.data
            o.mllr_status = res["status"].ctypes.data
        cfg = BatchCfg(tau, int(prior), mode, chunk_frames, int(graph), 0)
        self._check(lib().gmmhip_batch_adapt(
            self._h, x.ctypes.data, 0, off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), U, D,
            ctypes.byref(cfg), ctypes.byref(o)))
        self.N = int(off[-1])
        return res

    def batch_adapt_device(self, x_ptr, offsets, D, out_ptrs, tau=16, prior=False,
                           mode=BATCH_QUEUE, chunk_frames=0, graph=False):
        """Device-resident variant: frames and outputs are device pointers
        (out_ptrs: dict with optional means/occ/first/W/status)."""
        off = np.ascontiguousarray(offsets, np.uint64)
        U = len(off) - 1
        o = BatchOut(out_ptrs.get("means"), out_ptrs.get("occ"), out_ptrs.get("first"),
                     out_ptrs.get("W"), out_ptrs.get("status"))
        cfg = BatchCfg(tau, int(prior), mode, chunk_frames, int(graph), 1)
        self._check(lib().gmmhip_batch_adapt(
            self._h, ctypes.c_void_p(x_ptr), 1, off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            U, D, ctypes.byref(cfg), ctypes.byref(o)))
        self.N = int(off[-1])

    def graph_captures(self):
        """How many times this handle captured the batch pipeline into a hipGraph."""
        c = ctypes.c_uint64(0)
        self._check(lib().gmmhip_batch_graph_captures(self._h, ctypes.byref(c)))
        return c.value

    def tiles_used(self):
        """(ll_frames, chunk_frames, comp_tile) of the last accumulate / loglikes /
        batch call (comp_tile 0 when no mixed statistics kernel ran)."""
        a, b, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        self._check(lib().gmmhip_tiles_used(self._h, ctypes.byref(a), ctypes.byref(b),
                                            ctypes.byref(c)))
        return a.value, b.value, c.value

    def stream(self):
        return lib().gmmhip_stream(self._h)

    def sync(self):
        self._check(lib().gmmhip_sync(self._h))


class PendingRequest:
    """A submitted server request: the ctypes struct plus the numpy buffers
    it points to (kept alive until the server has finished it)."""

    def __init__(self, kind, frames, M, D, frame_ll=False, occ=False):
        self.frames = np.ascontiguousarray(frames, np.float32)
        assert self.frames.ndim == 2 and self.frames.shape[1] == D
        self.kind = kind
        r = Request()
        r.kind = kind
        r.frames = self.frames.ctypes.data
        r.n_frames = self.frames.shape[0]
        self.frame_ll = self.means = self.occ = self.W = None
        if frame_ll:
            self.frame_ll = np.zeros(self.frames.shape[0], np.float32)
            r.frame_ll = self.frame_ll.ctypes.data
        if kind == REQ_MAP:
            self.means = np.zeros((M, D), np.float32)
            r.means = self.means.ctypes.data
        if kind == REQ_MLLR:
            self.W = np.zeros((D, D + 1))
            r.mllr_W = self.W.ctypes.data
        if occ and kind != REQ_SCORE:
            self.occ = np.zeros(M)
            r.occ = self.occ.ctypes.data
        self.req = r

    @property
    def status(self):
        return self.req.status

    @property
    def ll_sum(self):
        return self.req.ll_sum

    @property
    def latency_ms(self):
        return self.req.latency_ms


class Server:
    """Continuous-batching request loop over a resident UBM
    (gmmhip_server_* of include/gmmhip.h)."""

    def __init__(self, w, mu, var, device=0, slots=2, slot_bytes=0, max_batch=0,
                 chunk_frames=0, tau=16, mllr_prior=False, **opts):
        w = np.ascontiguousarray(w, np.float32)
        mu = np.ascontiguousarray(mu, np.float32)
        var = np.ascontiguousarray(var, np.float32)
        self.M, self.D = mu.shape
        cfg = ServerCfg()
        cfg.opts = default_opts(**opts)
        cfg.slots, cfg.slot_bytes, cfg.max_batch = slots, slot_bytes, max_batch
        cfg.chunk_frames, cfg.tau, cfg.mllr_prior = chunk_frames, tau, int(mllr_prior)
        h = ctypes.c_void_p()
        rc = lib().gmmhip_server_create(device, ctypes.byref(cfg), self.M, self.D, _fp(w),
                                        _fp(mu), _fp(var), ctypes.byref(h))
        if rc:
            raise GmmHipError(rc, "gmmhip_server_create failed")
        self._h = h

    def submit(self, kind, frames, frame_ll=False, occ=False):
        p = PendingRequest(kind, frames, self.M, self.D, frame_ll, occ)
        rc = lib().gmmhip_server_submit(self._h, ctypes.byref(p.req))
        if rc:
            raise GmmHipError(rc, "gmmhip_server_submit failed")
        return p

    def wait(self, p, timeout_ms=-1.0):
        return lib().gmmhip_server_wait(self._h, ctypes.byref(p.req), timeout_ms)

    def drain(self):
        lib().gmmhip_server_drain(self._h)

    def stats(self, reset=False):
        st = ServerStats()
        lib().gmmhip_server_stats_get(self._h, ctypes.byref(st), int(reset))
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def last_error(self):
        return lib().gmmhip_server_last_error(self._h).decode()

    def close(self):
        if getattr(self, "_h", None):
            lib().gmmhip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
