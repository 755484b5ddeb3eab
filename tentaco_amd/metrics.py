"""Algorithmic (compulsory) byte counts and the benchmark workloads.

alg_bytes_mttkrp / alg_bytes_ttv count every input array read once and the
output written once, as fixed for the roofline figure:

  MTTKRP_N:  4(F0+1) + sum_{m=1..N-1} 4 F_m + sum_{m=2..N-1} 4 (F_{m-1}+1)
             + 4 nnz + sum_{m=1..N-1} 4 R dim_m + 4 R N0
  TTV:       4(N0+1) + 4 F1 + 4 (F1+1) + 8 nnz + 4 dim_last + 4 F_{N-2}

(F_m = nodes at level m, F0 = N0 for the dense root). The gather model
(every factor row fetched once per use) is reported separately as
"effective" bytes; it is not the roofline figure.
"""

from __future__ import annotations

import numpy as np

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published peak


def level_nodes(B):
    return [B.nodes(m) for m in range(B.order)]


def alg_bytes_mttkrp(B, R: int) -> int:
    F = level_nodes(B)
    N = B.order
    n0 = F[0]
    b = 4 * (n0 + 1)
    b += sum(4 * F[m] for m in range(1, N))
    b += sum(4 * (F[m - 1] + 1) for m in range(2, N))
    b += 4 * B.nnz
    b += sum(4 * R * B.dims[m] for m in range(1, N))
    b += 4 * R * B.dims[0]
    return int(b)


def gather_bytes_mttkrp(B, R: int) -> int:
    """Effective bytes when each factor row is fetched once per use in the
    factorised kernels (leaf factor per nonzero, level-m factor per level-m
    node), plus the index/value streams and the output."""
    F = level_nodes(B)
    N = B.order
    b = alg_bytes_mttkrp(B, R) - sum(4 * R * B.dims[m] for m in range(1, N))
    b += sum(4 * R * F[m] for m in range(1, N))
    return int(b)


def alg_bytes_ttv(B) -> int:
    F = level_nodes(B)
    N = B.order
    b = 4 * (F[0] + 1)
    if N >= 3:
        b += sum(4 * F[m] for m in range(1, N - 1)) + sum(4 * (F[m - 1] + 1) for m in range(2, N - 1))
    b += 4 * (F[N - 2] + 1)  # pos of the leaf level
    b += 8 * B.nnz           # crd + vals of the leaves
    b += 4 * B.dims[N - 1]   # c
    b += 4 * F[N - 2]        # output
    return int(b)


# ---------------------------------------------------------------------------
# benchmark workloads (seed = 1000 + workload index)
def workload_mttkrp3_uniform(nnz=500_000, dim=4096, seed=1000):
    """MTTKRP-3 mode-0 on 4096^3 with 500,000 i.i.d. uniform coordinates,
    duplicates summed, lexicographic CSF; values U[-1,1)."""
    import tentaco_amd as tt

    rng = np.random.default_rng(seed)
    coords = rng.integers(0, dim, (3, nnz), dtype=np.int64).astype(np.int32)
    vals = rng.uniform(-1, 1, nnz).astype(np.float32)
    return tt.Csf.from_coo((dim, dim, dim), coords, vals)


def workload_zipf(order, dim, nnz, seed, slice_s=1.1, leaf_sigma=1.0):
    """Zipf-slice / Zipf(1.3)-fan-out / log-normal(1.5, sigma)-fiber tensor."""
    import tentaco_amd as tt

    return tt.Csf.synth(order, (dim,) * order, nnz, slice_s=slice_s, fanout_s=1.3,
                        mid_mu=1.0, mid_sigma=1.0, leaf_mu=1.5, leaf_sigma=leaf_sigma,
                        leaf_cap=10000, seed=seed)


def factors(B, R, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1, 1, (B.dims[m], R)).astype(np.float32) for m in range(1, B.order)]


def workload_batch(ntensors=64, seed=1003, nnz_lo=1e4, nnz_hi=5e6, dim_lo=1e3, dim_hi=1e6):
    """Uneven batch: order U{3,4,5}, per-mode dims log-U[1e3,1e6], nnz
    log-U[1e4,5e6], Zipf slice exponent U[0.8,1.5]; returns the tensors (one
    CSF each, natural mode order)."""
    import tentaco_amd as tt

    rng = np.random.default_rng(seed)
    out = []
    for t in range(ntensors):
        order = int(rng.integers(3, 6))
        dims = [int(round(10 ** rng.uniform(np.log10(dim_lo), np.log10(dim_hi)))) for _ in range(order)]
        nnz = int(round(10 ** rng.uniform(np.log10(nnz_lo), np.log10(nnz_hi))))
        s = float(rng.uniform(0.8, 1.5))
        out.append(tt.Csf.synth(order, dims, nnz, slice_s=s, fanout_s=1.3, mid_mu=1.0,
                                mid_sigma=1.0, leaf_mu=1.5, leaf_sigma=1.0, leaf_cap=10000,
                                seed=seed * 1000 + t))
    return out
