"""ORACLE - test infrastructure only.

FP64 numpy CP-ALS on a small dense tensor, the checker of the library's
device CP-ALS (the role splatt_cpd_als plays in the reference, ds.h:395;
splatt's own implementation is not available, so this follows the textbook
ALS update, Kolda & Bader 2009, sec. 3.4, with the library's conventions):

  for m in 0..N-1:  M = X_(m) (Khatri-Rao of the other factors)   (MTTKRP)
                    V = Hadamard product of the other factors' Grams
                    A_m = M V^-1
                    lambda = column 2-norms on the first sweep,
                             max(1, max |a|) per column afterwards
                    A_m /= lambda
  fit = 1 - ||X - [lambda; A_1..A_N]|| / ||X||

Pinned by tests/test_oracle_cpd.py: exact recovery of a planted rank-R
tensor, a residual that never increases across sweeps (the ALS property),
and the fit formula against the explicitly reconstructed tensor.
cp_als_sparse is the same update on a COO tensor (for sizes the dense form
cannot hold), pinned by agreeing with cp_als sweep by sweep. Only tests/
may import it.
"""

from __future__ import annotations

import string

import numpy as np


def mttkrp_mode(X, factors, m):
    """M[i_m, r] = sum over the other indices of X * prod_{k != m} A_k[i_k, r]."""
    N = X.ndim
    idx = string.ascii_lowercase[:N]
    ops = [X]
    subs = [idx]
    for k in range(N):
        if k != m:
            ops.append(factors[k])
            subs.append(idx[k] + "z")
    return np.einsum(",".join(subs) + "->" + idx[m] + "z", *ops, optimize=True)


def reconstruct(lam, factors):
    N = len(factors)
    idx = string.ascii_lowercase[:N]
    subs = ",".join(i + "z" for i in idx) + ",z->" + idx
    return np.einsum(subs, *factors, lam, optimize=True)


def fit_of(X, lam, factors):
    return 1.0 - np.linalg.norm(X - reconstruct(lam, factors)) / np.linalg.norm(X)


def cp_als(X, init, sweeps, first_sweep=0):
    """Returns (factors, lambda, fits per sweep)."""
    A = [np.array(f, np.float64) for f in init]
    N = X.ndim
    R = A[0].shape[1]
    G = [a.T @ a for a in A]
    lam = np.ones(R)
    fits = []
    for it in range(sweeps):
        use_max = (first_sweep + it) > 0
        for m in range(N):
            M = mttkrp_mode(X, A, m)
            V = np.ones((R, R))
            for k in range(N):
                if k != m:
                    V *= G[k]
            A[m] = np.linalg.solve(V, M.T).T
            if use_max:
                lam = np.maximum(np.abs(A[m]).max(0), 1.0)
            else:
                lam = np.linalg.norm(A[m], axis=0)
                lam[lam == 0] = 1.0
            A[m] = A[m] / lam
            G[m] = A[m].T @ A[m]
        fits.append(fit_of(X, lam, A))
    return A, lam, fits


def mttkrp_coo(coords, vals, factors, m, rows, chunk=1 << 18):
    """FP64 MTTKRP from COO (coords N x nnz): M[i_m, r] = sum over nonzeros
    of v * prod_{k != m} A_k[i_k, r], accumulated in chunks of nonzeros."""
    R = factors[0].shape[1]
    M = np.zeros((rows, R))
    n = vals.shape[0]
    for a in range(0, n, chunk):
        e = min(n, a + chunk)
        w = np.asarray(vals[a:e], np.float64)[:, None] * np.ones((1, R))
        for k, A in enumerate(factors):
            if k != m:
                w *= A[coords[k, a:e]]
        np.add.at(M, coords[m, a:e], w)
    return M


def cp_als_sparse(coords, vals, dims, init, sweeps, first_sweep=0):
    """cp_als on a sparse tensor given as COO; the fit is
    1 - sqrt(||X||^2 + lambda^T (*_k G_k) lambda - 2 <X, model>) / ||X||,
    all in FP64. Returns (factors, lambda, fits per sweep)."""
    A = [np.array(f, np.float64) for f in init]
    N = len(dims)
    R = A[0].shape[1]
    G = [a.T @ a for a in A]
    lam = np.ones(R)
    v = np.asarray(vals, np.float64)
    # ||X||^2 over distinct coordinates (repeated ones summed first)
    order = np.lexsort(coords[::-1])
    sc = coords[:, order]
    start = np.ones(order.size, bool)
    start[1:] = np.any(sc[:, 1:] != sc[:, :-1], axis=0)
    xu = np.add.reduceat(v[order], np.flatnonzero(start)) if order.size else np.zeros(0)
    x2 = float(xu @ xu)
    fits = []
    for it in range(sweeps):
        use_max = (first_sweep + it) > 0
        for m in range(N):
            M = mttkrp_coo(coords, v, A, m, dims[m])
            V = np.ones((R, R))
            for k in range(N):
                if k != m:
                    V *= G[k]
            A[m] = np.linalg.solve(V, M.T).T
            if use_max:
                lam = np.maximum(np.abs(A[m]).max(0), 1.0)
            else:
                lam = np.linalg.norm(A[m], axis=0)
                lam[lam == 0] = 1.0
            A[m] = A[m] / lam
            G[m] = A[m].T @ A[m]
            if m == N - 1:
                inner = float(np.sum(M * A[m] * lam))
        had = np.ones((R, R))
        for k in range(N):
            had *= G[k]
        resid = max(0.0, x2 + float(lam @ had @ lam) - 2.0 * inner)
        fits.append(1.0 - np.sqrt(resid) / np.sqrt(x2))
    return A, lam, fits
