"""ctypes binding of the host-layer C ABI (include/gmmhost.h): the `.gmm`
model file, RearrangeMixtures, the EM / MAP / MLLR update rules and the
SVES / HTK / RAW feature files."""
import ctypes

import numpy as np

from . import lib

HOST_EXPORTS = [
    "gmmhost_model_load", "gmmhost_model_save", "gmmhost_model_create",
    "gmmhost_model_create_full", "gmmhost_model_full",
    "gmmhost_model_dims", "gmmhost_model_get", "gmmhost_model_rearrange",
    "gmmhost_model_free", "gmmhost_update_em", "gmmhost_update_map",
    "gmmhost_mllr_transform", "gmmhost_mllr_apply", "gmmhost_last_error",
    "gmmhost_fmllr_transform", "gmmhost_update_em_robust", "gmmhost_make_robust_stats",
    "gmmhost_features_load", "gmmhost_features_load_ranges", "gmmhost_features_dims",
    "gmmhost_features_get", "gmmhost_features_free", "gmmhost_features_save",
]

SVES, HTK, RAW = 0, 1, 2
FEAT_MISSING = 2

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_configured = False


def _L():
    global _configured
    L = lib()
    if not _configured:
        vp = ctypes.c_void_p
        L.gmmhost_last_error.restype = ctypes.c_char_p
        L.gmmhost_model_load.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(vp)]
        L.gmmhost_model_save.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.gmmhost_model_create.argtypes = [ctypes.c_int, ctypes.c_int, _fp, _fp, _fp,
                                           ctypes.POINTER(vp)]
        L.gmmhost_model_create_full.argtypes = [ctypes.c_int, ctypes.c_int, _fp, _fp, _fp,
                                                ctypes.POINTER(vp)]
        L.gmmhost_model_full.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.gmmhost_model_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]
        L.gmmhost_model_get.argtypes = [vp, _fp, _fp, _fp]
        L.gmmhost_model_rearrange.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.gmmhost_model_free.argtypes = [vp]
        L.gmmhost_model_free.restype = None
        L.gmmhost_update_em.argtypes = [vp, _dp, _dp, _dp, ctypes.c_uint64, ctypes.c_float,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int)]
        L.gmmhost_update_map.argtypes = [vp, vp, _dp, _dp, _dp, ctypes.c_uint64, ctypes.c_int,
                                         ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int]
        L.gmmhost_mllr_transform.argtypes = [vp, vp, _dp, _dp, ctypes.c_float, _dp]
        L.gmmhost_mllr_apply.argtypes = [vp, _dp]
        L.gmmhost_fmllr_transform.argtypes = [vp, _dp, _dp, _dp, ctypes.c_uint64, ctypes.c_int,
                                              _dp]
        L.gmmhost_update_em_robust.argtypes = [vp, _dp, _dp, _dp, _dp, ctypes.c_uint64,
                                               ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.gmmhost_make_robust_stats.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                                _dp, _dp, _dp]
        u32 = ctypes.c_uint32
        L.gmmhost_features_load.argtypes = [ctypes.c_char_p, ctypes.c_int, u32, ctypes.POINTER(vp)]
        L.gmmhost_features_load_ranges.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                                   ctypes.POINTER(u32), ctypes.c_int, u32,
                                                   ctypes.POINTER(vp)]
        L.gmmhost_features_dims.argtypes = [vp, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                            ctypes.POINTER(u32)]
        L.gmmhost_features_get.argtypes = [vp, _fp]
        L.gmmhost_features_free.argtypes = [vp]
        L.gmmhost_features_free.restype = None
        L.gmmhost_features_save.argtypes = [ctypes.c_char_p, ctypes.c_int, _fp, u32, u32, u32]
        _configured = True
    return L


class HostError(RuntimeError):
    def __init__(self, msg, code=1):
        super().__init__(msg)
        self.code = code


def _chk(rc):
    if rc != 0:
        raise HostError("gmmhost error %d: %s" % (rc, _L().gmmhost_last_error().decode()), rc)


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _d(a):
    return None if a is None else np.ascontiguousarray(a, np.float64)


def _pd(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Model:
    """GMModel handle (diagonal or full covariance)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def load(cls, path, txt):
        h = ctypes.c_void_p()
        _chk(_L().gmmhost_model_load(str(path).encode(), int(txt), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, w, mu, var):
        w, mu, var = _f(w), _f(mu), _f(var)
        h = ctypes.c_void_p()
        _chk(_L().gmmhost_model_create(mu.shape[0], mu.shape[1], w.ctypes.data_as(_fp),
                                       mu.ctypes.data_as(_fp), var.ctypes.data_as(_fp),
                                       ctypes.byref(h)))
        return cls(h)

    @classmethod
    def create_full(cls, w, mu, cov):
        w, mu, cov = _f(w), _f(mu), _f(cov)
        h = ctypes.c_void_p()
        _chk(_L().gmmhost_model_create_full(mu.shape[0], mu.shape[1], w.ctypes.data_as(_fp),
                                            mu.ctypes.data_as(_fp), cov.ctypes.data_as(_fp),
                                            ctypes.byref(h)))
        return cls(h)

    def full(self):
        f = ctypes.c_int()
        _chk(_L().gmmhost_model_full(self._h, ctypes.byref(f)))
        return bool(f.value)

    def __del__(self):
        if getattr(self, "_h", None):
            _L().gmmhost_model_free(self._h)
            self._h = None

    def save(self, path, txt):
        _chk(_L().gmmhost_model_save(self._h, str(path).encode(), int(txt)))

    def dims(self):
        M, D = ctypes.c_int(), ctypes.c_int()
        _chk(_L().gmmhost_model_dims(self._h, ctypes.byref(M), ctypes.byref(D)))
        return M.value, D.value

    def get(self):
        """(w[M], mu[M,D], var[M,D]) or, for a full model, (w, mu, cov[M,D,D])."""
        M, D = self.dims()
        w = np.zeros(M, np.float32)
        mu = np.zeros((M, D), np.float32)
        var = np.zeros((M, D, D) if self.full() else (M, D), np.float32)
        _chk(_L().gmmhost_model_get(self._h, w.ctypes.data_as(_fp), mu.ctypes.data_as(_fp),
                                    var.ctypes.data_as(_fp)))
        return w, mu, var

    def rearrange(self):
        nz = ctypes.c_int()
        _chk(_L().gmmhost_model_rearrange(self._h, ctypes.byref(nz)))
        return nz.value

    def update_em(self, n, F, S, tot, min_var=1e-6, upweight=True, upmean=True, upvar=True):
        n, F, S = _d(n), _d(F), _d(S)
        left = ctypes.c_int()
        _chk(_L().gmmhost_update_em(self._h, _pd(n), _pd(F), _pd(S), int(tot), min_var,
                                    int(upweight), int(upmean), int(upvar), ctypes.byref(left)))
        return left.value

    def update_em_robust(self, n, F, S, aux, tot, min_var=1e-6, upweight=True, upmean=True,
                         upvar=True):
        n, F, S, aux = _d(n), _d(F), _d(S), _d(aux)
        left = ctypes.c_int()
        _chk(_L().gmmhost_update_em_robust(self._h, _pd(n), _pd(F), _pd(S), _pd(aux), int(tot),
                                           min_var, int(upweight), int(upmean), int(upvar),
                                           ctypes.byref(left)))
        return left.value

    def update_map(self, ubm, n, F, S, tot, tau, min_var=1e-6, upweight=False, upmean=True,
                   upvar=False):
        n, F, S = _d(n), _d(F), _d(S)
        _chk(_L().gmmhost_update_map(self._h, ubm._h, _pd(n), _pd(F), _pd(S), int(tot),
                                     int(tau), min_var, int(upweight), int(upmean), int(upvar)))

    def mllr_transform(self, n, F, ubm=None, min_var=1e-6):
        M, D = self.dims()
        n, F = _d(n), _d(F)
        W = np.zeros((D, D + 1))
        _chk(_L().gmmhost_mllr_transform(self._h, ubm._h if ubm is not None else None, _pd(n),
                                         _pd(F), min_var, _pd(W)))
        return W

    def fmllr_transform(self, n, F, S, tot, iters=20):
        M, D = self.dims()
        n, F, S = _d(n), _d(F), _d(S)
        W = np.zeros((D, D + 1))
        _chk(_L().gmmhost_fmllr_transform(self._h, _pd(n), _pd(F), _pd(S), int(tot), int(iters),
                                          _pd(W)))
        return W

    def mllr_apply(self, W):
        W = _d(W)
        _chk(_L().gmmhost_mllr_apply(self._h, _pd(W)))


def make_robust_stats(n, F, S, aux, full=False):
    """(S', nind): MakeRobustStats on a copy of S."""
    n, F, aux = _d(n), _d(F), _d(aux)
    S = _d(S).copy()
    M, D = F.shape
    nind = np.zeros(M)
    _chk(_L().gmmhost_make_robust_stats(M, D, int(full), _pd(n), _pd(F), _pd(S), _pd(aux),
                                        _pd(nind)))
    return S, nind


def _take_features(h):
    u32 = ctypes.c_uint32
    N, D, P = u32(), u32(), u32()
    try:
        _chk(_L().gmmhost_features_dims(h, ctypes.byref(N), ctypes.byref(D), ctypes.byref(P)))
        x = np.zeros((N.value, D.value), np.float32)
        _chk(_L().gmmhost_features_get(h, x.ctypes.data_as(_fp)))
    finally:
        _L().gmmhost_features_free(h)
    return x, P.value


def load_features(path, kind, dwnsmp=1):
    """(frames [N, D] float32, sample period) of a SVES / HTK / RAW file,
    every dwnsmp-th frame (Param::Load).  HostError with code FEAT_MISSING
    when the file cannot be opened."""
    h = ctypes.c_void_p()
    _chk(_L().gmmhost_features_load(str(path).encode(), int(kind), int(dwnsmp), ctypes.byref(h)))
    return _take_features(h)


def load_feature_ranges(path, kind, ranges, sample_period=0):
    """Frames of the intervals [x0, x1) (sample-period units) of one file,
    (Param::Load with frame intervals)."""
    r = np.ascontiguousarray(np.asarray(ranges, np.uint32).reshape(-1, 2))
    h = ctypes.c_void_p()
    _chk(_L().gmmhost_features_load_ranges(str(path).encode(), int(kind),
                                           r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                           r.shape[0], int(sample_period), ctypes.byref(h)))
    return _take_features(h)


def save_features(path, kind, x, sample_period=100000):
    x = _f(x)
    if x.ndim != 2:
        raise ValueError("frames must be [N, D]")
    _chk(_L().gmmhost_features_save(str(path).encode(), int(kind), x.ctypes.data_as(_fp),
                                    x.shape[0], x.shape[1], int(sample_period)))
