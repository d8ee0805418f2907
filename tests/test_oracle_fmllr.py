"""Pins the oracle's fMLLR (oracle/gmm_oracle.c: orc_det, orc_fmllr_kg,
orc_fmllr) before it checks the host's ComputeW_fMLLR: numpy for the
determinant and the row statistics, and the mathematics of constrained MLLR
for the transform (stationarity of the objective, identity on the model's
own frames, recovery of a planted affine map)."""
import numpy as np
import pytest

import _oracle as orc


def model_and_frames(seed, M=64, D=4, N=40000):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(M)).astype(np.float32)
    mu = rng.normal(0, 2, (M, D)).astype(np.float32)
    var = rng.uniform(0.3, 1.5, (M, D)).astype(np.float32)
    p = w.astype(np.float64)
    k = rng.choice(M, size=N, p=p / p.sum())
    y = (mu[k] + np.sqrt(var[k]) * rng.normal(size=(N, D))).astype(np.float32)
    return rng, w, mu, var, y


def diag_cov(var):
    M, D = var.shape
    c = np.zeros((M, D, D), np.float32)
    c[:, np.arange(D), np.arange(D)] = var
    return c


def test_det_matches_numpy():
    rng = np.random.default_rng(0)
    for n in (1, 3, 7, 20):
        A = rng.normal(size=(n, n))
        assert orc.det(A) == pytest.approx(np.linalg.det(A), rel=1e-10)
    P = np.eye(4)[[1, 0, 3, 2]]  # two swaps: det = +1
    assert orc.det(P) == 1.0


def test_row_statistics_brute_force():
    rng, w, mu, var, y = model_and_frames(1, M=5, D=3, N=300)
    n, F, S, _ = orc.accumulate_full(y, w, mu, diag_cov(var))
    r, c = np.triu_indices(3)
    for i in range(3):
        G, k = orc.fmllr_kg(mu, var, n, F, S, i)
        Gr = np.zeros((4, 4))
        kr = np.zeros(4)
        for m in range(5):
            Sm = np.zeros((3, 3))
            Sm[r, c] = S[m]
            Sm[c, r] = S[m]
            blk = np.block([[Sm, F[m][:, None]], [F[m][None], np.array([[n[m]]])]])
            Gr += blk * np.float64(np.float32(1) / var[m, i])  # 1/var in float, as the reference
            kr += np.float64(np.float32(mu[m, i] / var[m, i])) * np.append(F[m], n[m])
        assert np.allclose(G, Gr, rtol=1e-12)
        assert np.allclose(k, kr, rtol=1e-12)


def test_identity_on_model_frames():
    _, w, mu, var, y = model_and_frames(2)
    n, F, S, _ = orc.accumulate_full(y, w, mu, diag_cov(var))
    W = orc.fmllr(mu, var, n, F, S, y.shape[0])
    D = mu.shape[1]
    assert np.abs(W[:, :D] - np.eye(D)).max() < 0.01
    assert np.abs(W[:, D]).max() < 0.02


def test_stationary_point_of_the_objective():
    """At convergence every row satisfies the stationarity condition of
    beta log|det A| + sum_i (w_i k_i - w_i G_i w_i^T / 2):
    beta p_i / (p_i . w_i) + k_i - G_i w_i = 0 with p_i the cofactors of A."""
    rng, w, mu, var, y = model_and_frames(3)
    D = mu.shape[1]
    A0 = np.eye(D) + rng.normal(0, 0.1, (D, D))
    x = ((y - 0.3) @ np.linalg.inv(A0).T).astype(np.float32)
    n, F, S, _ = orc.accumulate_full(x, w, mu, diag_cov(var))
    N = x.shape[0]
    W = orc.fmllr(mu, var, n, F, S, N, iters=60)
    A = W[:, :D]
    cof = np.linalg.inv(A).T * np.linalg.det(A)
    for i in range(D):
        G, k = orc.fmllr_kg(mu, var, n, F, S, i)
        p = np.append(cof[i], 0.0)
        g = N * p / (p @ W[i]) + k - G @ W[i]
        assert np.abs(g).max() < 1e-6 * np.abs(k).max()
    # the frames were transformed: the stationary point is not the identity
    assert np.abs(W[:, :D] - np.eye(D)).max() > 0.05


def test_iterated_fmllr_recovers_a_planted_map():
    """Frames pushed through the inverse of an affine map: re-estimating
    fMLLR on the transformed frames converges to the map."""
    rng, w, mu, var, y = model_and_frames(4)
    D = mu.shape[1]
    A0 = np.eye(D) + rng.normal(0, 0.15, (D, D))
    b0 = rng.normal(0, 0.5, D)
    x = ((y - b0) @ np.linalg.inv(A0).T).astype(np.float32)
    A, b = np.eye(D), np.zeros(D)
    for _ in range(8):
        xt = (x @ A.T + b).astype(np.float32)
        n, F, S, _ = orc.accumulate_full(xt, w, mu, diag_cov(var))
        W = orc.fmllr(mu, var, n, F, S, x.shape[0])
        A, b = W[:, :D] @ A, W[:, :D] @ b + W[:, D]
    assert np.abs(A - A0).max() < 0.02
    assert np.abs(b - b0).max() < 0.03
