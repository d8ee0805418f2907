"""Independent numpy readers for the reference's file formats, used by the
tests to read the golden fixtures and to check the library's own readers.

RAW features: int32 N, int32 D, float32[N*D]   (OL_Param.cpp:358-453)
GMM text:     OL_GMModel.cpp:475-547 (write) / :627-753 (read)
saveMat text: "x y" header then rows           (IOMats.h:168-232)
"""
import numpy as np


def read_raw(path):
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(8), dtype="<i4")
        n, d = int(hdr[0]), int(hdr[1])
        data = np.frombuffer(f.read(), dtype="<f4")
    if n < 1:
        n = data.size // d
    return data[: n * d].reshape(n, d).copy()


def write_raw(path, x):
    x = np.ascontiguousarray(x, dtype="<f4")
    with open(path, "wb") as f:
        f.write(np.array([x.shape[0], x.shape[1]], dtype="<i4").tobytes())
        f.write(x.tobytes())


def read_gmm_txt(path):
    """Returns (w[M], mu[M,D], var[M,D], norm[D]) for a diagonal model."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    pos = 0
    D = int(lines[pos].split()[0]); pos += 1
    norm = np.array([float(lines[pos + i].split()[0]) for i in range(D)],
                    dtype=np.float32)
    pos += D
    M = int(lines[pos].split()[0]); pos += 1
    w = np.zeros(M, np.float32)
    mu = np.zeros((M, D), np.float32)
    var = np.zeros((M, D), np.float32)
    for m in range(M):
        tok = lines[pos].split(); pos += 1
        if len(tok) > 1 and tok[1] == "F":
            raise ValueError("full-covariance model")
        w[m] = float(tok[0])
        pos += 1  # empty line
        mu[m] = [float(lines[pos + i].split()[0]) for i in range(D)]
        pos += D
        pos += 1  # empty line
        var[m] = [float(lines[pos + i].split()[0]) for i in range(D)]
        pos += D
        pos += 1  # empty line
    return w, mu, var, norm


def read_gmm_txt_full(path):
    """Returns (w[M], mu[M,D], cov[M,D,D], norm[D]) for a full-covariance
    model: weight line "w F", then the mean, then D rows of 1 + D numbers
    (the diagonal element, then the row), OL_GMModel.cpp:511-533 / :700-740."""
    toks = open(path).read().split()
    p = 0
    D = int(toks[p]); p += 1
    norm = np.array([float(t) for t in toks[p:p + D]], np.float32); p += D
    M = int(toks[p]); p += 1
    w = np.zeros(M, np.float32)
    mu = np.zeros((M, D), np.float32)
    cov = np.zeros((M, D, D), np.float32)
    for m in range(M):
        w[m] = float(toks[p]); p += 1
        if toks[p] != "F":
            raise ValueError("diagonal model")
        p += 1
        mu[m] = [float(t) for t in toks[p:p + D]]; p += D
        for i in range(D):
            p += 1  # the diagonal element, repeated
            cov[m, i] = [float(t) for t in toks[p:p + D]]; p += D
    return w, mu, cov, norm


def read_savemat_txt(path):
    with open(path) as f:
        toks = f.read().split()
    x, y = int(toks[0]), int(toks[1])
    vals = np.array([float(t) for t in toks[2: 2 + x * y]])
    return vals.reshape(x, y)
