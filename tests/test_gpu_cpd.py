"""Device CP-ALS (batched MTTKRP + Hadamard/Cholesky solve + normalisation +
Grams) against the FP64 CP-ALS oracle on the same initial factors."""

import numpy as np
import pytest

import tentaco_amd as tt
from oracle import cpd as ocpd

pytestmark = pytest.mark.gpu


def _tensor(rng, dims, density=0.6):
    n = int(np.prod(dims) * density)
    coords = np.stack([rng.integers(0, d, n) for d in dims]).astype(np.int32)
    vals = rng.uniform(-1, 1, n).astype(np.float32)
    B = tt.Csf.from_coo(dims, coords, vals)
    # the dense tensor from the CSF's own (FP32-summed) values: walk every
    # leaf's path of node indices from the root
    Xc = np.zeros(dims, np.float32)
    paths = [np.arange(B.nodes(0))]
    for m in range(1, B.order):
        pos = np.asarray(B.pos[m])
        parent = np.repeat(np.arange(B.nodes(m - 1)), np.diff(pos))
        paths = [p[parent] for p in paths] + [np.arange(B.nodes(m))]
    coord = [np.asarray(B.crd[m])[paths[m]] for m in range(B.order)]
    Xc[tuple(coord)] = np.asarray(B.vals)
    return B, Xc.astype(np.float64)


def _setup(seed, shapes, R, density=0.6):
    rng = np.random.default_rng(seed)
    hosts, dense, handles = [], [], []
    for dims in shapes:
        B, X = _tensor(rng, dims, density)
        hosts.append(B)
        dense.append(X)
        handles.append([tt.DeviceCsf(Bm) for Bm in B.mode_csfs()])
    c = tt.Cpd(handles, R, seed=seed)
    inits = []
    for t, dims in enumerate(shapes):
        f = [rng.normal(size=(d, R)).astype(np.float32) for d in dims]
        for m, a in enumerate(f):
            c.set_factor(t, m, a)
        inits.append(f)
    return c, dense, inits, handles


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


SHAPES = [(12, 10, 8), (8, 7, 6, 5), (6, 5, 5, 4, 4)]


@pytest.mark.parametrize("use_graphs", [False, True])
def test_cpd_matches_oracle(gpu, use_graphs):
    R = 16
    c, dense, inits, _ = _setup(3, SHAPES, R)
    fits1 = c.iterate(1, use_graphs=use_graphs)
    for t, X in enumerate(dense):
        A, lam, fits = ocpd.cp_als(X, inits[t], 1)
        for m in range(X.ndim):
            a, l = c.factor(t, m)
            assert _rel(a, A[m]) <= 1e-4, (t, m, _rel(a, A[m]))
        assert _rel(l, lam) <= 1e-4
        assert fits1[t] == pytest.approx(fits[-1], abs=1e-4)
    fits5 = c.iterate(4, use_graphs=use_graphs)
    for t, X in enumerate(dense):
        A, lam, fits = ocpd.cp_als(X, inits[t], 5)
        for m in range(X.ndim):
            a, _ = c.factor(t, m)
            assert _rel(a, A[m]) <= 2e-3, (t, m, _rel(a, A[m]))
        assert fits5[t] == pytest.approx(fits[-1], abs=1e-3)
        assert fits5[t] >= fits1[t] - 1e-4  # ALS does not lose fit
    c.close()


@pytest.mark.parametrize("R,shapes,sweeps", [
    # a 5000-row mode: several 4096-row update blocks and every wave's tiles
    (32, [(5000, 72, 68), (96, 70, 40, 36)], 5),
    (64, [(4200, 72, 70)], 1),
])
def test_cpd_matches_oracle_ranks(gpu, R, shapes, sweeps):
    """The R = 32 update on the matrix cores and the R = 64 update kernel,
    against the FP64 oracle from the same initial factors."""
    c, dense, inits, _ = _setup(11 + R, shapes, R, density=0.02)
    fits1 = c.iterate(1, use_graphs=True)
    for t, X in enumerate(dense):
        A, lam, fits = ocpd.cp_als(X, inits[t], 1)
        for m in range(X.ndim):
            a, l = c.factor(t, m)
            assert _rel(a, A[m]) <= 1e-4, (R, t, m, _rel(a, A[m]))
        assert _rel(l, lam) <= 1e-4, (R, t)
        assert fits1[t] == pytest.approx(fits[-1], abs=1e-4)
    if sweeps > 1:
        fitsn = c.iterate(sweeps - 1, use_graphs=True)
        for t, X in enumerate(dense):
            A, lam, fits = ocpd.cp_als(X, inits[t], sweeps)
            for m in range(X.ndim):
                a, _ = c.factor(t, m)
                assert _rel(a, A[m]) <= 2e-3, (R, t, m, _rel(a, A[m]))
            assert fitsn[t] == pytest.approx(fits[-1], abs=1e-3)
    c.close()


def test_cpd_recovers_planted_tensor(gpu):
    rng = np.random.default_rng(5)
    R = 16
    dims = (24, 20, 18)
    fac = [rng.normal(size=(d, R)) for d in dims]
    X = ocpd.reconstruct(rng.uniform(1, 2, R), fac).astype(np.float32)
    coords = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")).reshape(3, -1)
    B = tt.Csf.from_coo(dims, coords.astype(np.int32), X.reshape(-1))
    handles = [[tt.DeviceCsf(Bm) for Bm in B.mode_csfs()]]
    c = tt.Cpd(handles, R, seed=1)
    # an initialisation that leaves the ALS "swamp" within 20 sweeps (others,
    # e.g. the next draws of rng, sit near a fit of 0.81 for over 100)
    r2 = np.random.default_rng(1000)
    init = [r2.normal(size=(d, R)).astype(np.float32) for d in dims]
    for m, a in enumerate(init):
        c.set_factor(0, m, a)
    fits = [c.iterate(1)[0] for _ in range(40)]
    _, _, ref = ocpd.cp_als(X.astype(np.float64), init, 40)
    # the device follows the FP64 trajectory sweep by sweep
    for k in (4, 9, 19):
        assert fits[k] == pytest.approx(ref[k], abs=2e-3), (k, fits[k], ref[k])
    # near a fit of 1, ||X||^2 + ||model||^2 - 2 <X, model> cancels: FP32 Grams
    # leave ~1e-6 relative noise in the residual, ~1e-3 in the fit
    assert all(b >= a - 5e-3 for a, b in zip(fits, fits[1:])), fits
    assert fits[-1] > 0.99, fits[-5:]
    c.close()


def test_cpd_rejects_bad_input(gpu):
    rng = np.random.default_rng(6)
    B, _ = _tensor(rng, (6, 5, 4))
    hs = [tt.DeviceCsf(Bm) for Bm in B.mode_csfs()]
    with pytest.raises(tt.TentacoError):
        tt.Cpd([hs], 24)  # R outside {16, 32, 64}
    with pytest.raises(tt.TentacoError):
        tt.Cpd([hs], 16, level_modes=[[[1, 0, 2], [1, 0, 2], [2, 0, 1]]])  # mode 0 not at its root


def _coo_of(B):
    paths = [np.arange(B.nodes(0))]
    for m in range(1, B.order):
        pos = np.asarray(B.pos[m])
        parent = np.repeat(np.arange(B.nodes(m - 1)), np.diff(pos))
        paths = [p[parent] for p in paths] + [np.arange(B.nodes(m))]
    return (np.stack([np.asarray(B.crd[m])[paths[m]] for m in range(B.order)]),
            np.asarray(B.vals, np.float64))


def test_cpd_sparse_high_order_large_modes(gpu):
    """5-mode tensors with modes of 2-6e4 rows and few nonzeros (the shape
    of the batch's sparse order-5 tensors, where FP32 products of four
    factors' scales are at risk), from the device's own seeded factors,
    against the sparse FP64 oracle: fits and factors after 1 and 5 sweeps."""
    dims_list = [(60_000, 50_000, 40_000, 30_000, 20_000), (45_000, 3_000, 52_000, 800, 38_000)]
    tensors = [tt.Csf.synth(5, d, 15_000, slice_s=1.2, seed=31 + k) for k, d in enumerate(dims_list)]
    c = tt.Cpd([[tt.DeviceCsf(Bm) for Bm in B.mode_csfs()] for B in tensors], 32, seed=1005)
    inits = [[c.factor(t, m)[0].astype(np.float64) for m in range(5)] for t in range(len(tensors))]
    f1 = c.iterate(1, use_graphs=True)
    f5 = c.iterate(4, use_graphs=True)
    for t, B in enumerate(tensors):
        co, v = _coo_of(B)
        A1, _, fits1 = ocpd.cp_als_sparse(co, v, B.dims, inits[t], 1)
        A5, lam, fits5 = ocpd.cp_als_sparse(co, v, B.dims, inits[t], 5)
        assert 0.0 <= f1[t] < 1.0 and f1[t] == pytest.approx(fits1[0], abs=1e-4), (t, f1[t], fits1)
        assert f5[t] == pytest.approx(fits5[-1], abs=1e-3), (t, f5[t], fits5)
        for m in range(5):
            a, l = c.factor(t, m)
            assert _rel(a, A5[m]) <= 2e-3, (t, m, _rel(a, A5[m]))
        assert _rel(l, lam) <= 2e-3
    c.close()
