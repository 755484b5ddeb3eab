"""Shared test inputs (random COO tensors, factors) and tolerances."""

import numpy as np

import tentaco_amd as tt

# FP32 GPU results against the FP32/FP64 CPU restatement of the reference
# (relative L2, the tolerance the north star sets for floating point).
REL_L2_TOL = 1e-5


def random_coo(rng, dims, n):
    coords = np.stack([rng.integers(0, d, n) for d in dims]).astype(np.int32)
    vals = rng.uniform(-1, 1, n).astype(np.float32)
    return coords, vals


def random_csf(rng, dims, n):
    coords, vals = random_coo(rng, dims, n)
    return tt.Csf.from_coo(dims, coords, vals)


def factors_for(rng, B, R, lo=-1.0, hi=1.0):
    return [rng.uniform(lo, hi, (B.dims[m], R)).astype(np.float32) for m in range(1, B.order)]


def skewed_coo(rng, dims, n, hot_slice=0, hot_frac=0.6, empty_every=3):
    """COO where one slice holds hot_frac of the nonzeros and every
    empty_every-th slice is empty (long runs spanning workgroups, empty
    slices inside the partition brackets)."""
    coords = np.stack([rng.integers(0, d, n) for d in dims]).astype(np.int64)
    hot = rng.random(n) < hot_frac
    coords[0, hot] = hot_slice
    cold = ~hot
    allowed = np.array([i for i in range(dims[0]) if i % empty_every != 1 or i == hot_slice])
    coords[0, cold] = allowed[rng.integers(0, allowed.shape[0], int(cold.sum()))]
    vals = rng.uniform(-1, 1, n).astype(np.float32)
    return coords.astype(np.int32), vals
