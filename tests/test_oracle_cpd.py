"""The FP64 CP-ALS oracle against closed forms (CPU only)."""

import numpy as np

from oracle import cpd


def _planted(rng, dims, R):
    # well-separated (normal) factors: ALS converges in a few dozen sweeps;
    # strongly correlated positive factors show the slow "swamp" behaviour
    fac = [rng.normal(size=(d, R)) for d in dims]
    lam = rng.uniform(1.0, 2.0, R)
    return cpd.reconstruct(lam, fac), fac, lam


def test_recovers_planted_low_rank_tensor():
    rng = np.random.default_rng(0)
    X, fac, lam = _planted(rng, (9, 8, 7), 3)
    init = [rng.normal(size=(d, 3)) for d in X.shape]
    _, _, fits = cpd.cp_als(X, init, 60)
    assert fits[-1] > 0.9999, fits[-5:]


def test_residual_never_increases():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(7, 6, 5, 4))
    init = [rng.normal(size=(d, 4)) for d in X.shape]
    _, _, fits = cpd.cp_als(X, init, 15)
    assert all(b >= a - 1e-9 for a, b in zip(fits, fits[1:])), fits


def test_fit_matches_explicit_reconstruction_and_mttkrp():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(5, 4, 3))
    A = [rng.normal(size=(d, 2)) for d in X.shape]
    lam = np.array([1.5, 0.5])
    Xh = np.zeros_like(X)
    for r in range(2):
        Xh += lam[r] * np.einsum("i,j,k->ijk", A[0][:, r], A[1][:, r], A[2][:, r])
    assert np.allclose(cpd.reconstruct(lam, A), Xh)
    M = cpd.mttkrp_mode(X, A, 1)
    ref = np.zeros((4, 2))
    for i in range(5):
        for j in range(4):
            for k in range(3):
                ref[j] += X[i, j, k] * A[0][i] * A[2][k]
    assert np.allclose(M, ref)


def test_sparse_als_matches_dense_als():
    """The COO form of the oracle follows the dense one sweep by sweep
    (factors, lambda and the fit from its formula), on a 4-mode tensor with
    empty slices and repeated coordinates summed."""
    rng = np.random.default_rng(21)
    dims = (9, 8, 7, 6)
    n = 700
    coords = np.stack([rng.integers(0, d, n) for d in dims])
    vals = rng.uniform(-1, 1, n)
    X = np.zeros(dims)
    np.add.at(X, tuple(coords), vals)
    R = 5
    init = [rng.uniform(0, 1, (d, R)) for d in dims]
    Ad, ld, fd = cpd.cp_als(X, init, 4)
    As, ls, fs = cpd.cp_als_sparse(coords, vals, dims, init, 4)
    for a, b in zip(Ad, As):
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12)
    assert np.allclose(ld, ls, rtol=1e-10)
    assert np.allclose(fd, fs, rtol=1e-9, atol=1e-12)
