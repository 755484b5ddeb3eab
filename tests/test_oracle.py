"""Pin the oracle before trusting it (CPU only).

The reference ships no golden vectors or fixtures for this path (its only
check, mttkrp_csf_gpu.h:42, is an all-ones structural KAT). So every oracle
function is pinned here against what the mathematics fixes:
  * a dense FP64 brute force (numpy einsum, independent of the loop nests),
  * closed-form known answers (all-ones -> slice nonzero counts, one-hot
    factors select sub-tensors, e_k picks single entries for TTV),
  * linearity in B and in each factor,
  * agreement of the three reference schedules (taco / unscheduled /
    workspace), which differ only in association,
  * the committed golden fixtures in tests/golden (FP64 brute force),
  * planted mistakes (a transposed factor, a dropped nonzero, a shifted
    coordinate) that must fail the checks.
"""

import json
import os

import numpy as np
import pytest

import oracle
import tentaco_amd as tt
from helpers import random_coo

HERE = os.path.dirname(os.path.abspath(__file__))


def _tiny(order, seed, n=80, lo=3, hi=7):
    rng = np.random.default_rng(seed)
    dims = tuple(int(x) for x in rng.integers(lo, hi, order))
    coords, vals = random_coo(rng, dims, n)
    return rng, dims, coords, vals


@pytest.mark.parametrize("order", [2, 3, 4, 5])
@pytest.mark.parametrize("schedule", ["taco", "unscheduled", "workspace"])
def test_oracle_matches_dense_bruteforce(order, schedule):
    rng, dims, coords, vals = _tiny(order, 10 + order)
    B = oracle.csf_from_coo(dims, coords, vals)
    R = 6
    facs = [rng.uniform(-1, 1, (dims[m], R)).astype(np.float32) for m in range(1, order)]
    ref = oracle.mttkrp_dense(oracle.dense_from_csf(B), facs)
    a64 = oracle.mttkrp(B, facs, schedule, fp64=True)
    assert oracle.rel_l2(a64, ref) < 1e-14
    a32 = oracle.mttkrp(B, facs, schedule, fp64=False)
    assert oracle.rel_l2(a32, ref) < 1e-6


@pytest.mark.parametrize("order", [3, 4, 5])
def test_all_ones_kat(order):
    # mttkrp_csf_gpu.h:42 sets B = C = D = 1: A[i, :] = nnz(slice i)
    rng, dims, coords, vals = _tiny(order, 20 + order, n=200)
    B = oracle.csf_from_coo(dims, coords, np.ones_like(vals))
    B.vals = np.ones_like(B.vals)  # duplicates were summed: reset to one per stored entry
    facs = [np.ones((dims[m], 4), np.float32) for m in range(1, order)]
    for fp64 in (True, False):
        A = oracle.mttkrp(B, facs, "taco", fp64=fp64)
        counts = np.zeros(dims[0])
        coords_u, _ = oracle.coo_from_csf(B)
        np.add.at(counts, coords_u[0], 1)
        assert np.array_equal(A, np.repeat(counts[:, None], 4, axis=1))


def test_one_hot_factors_select_subtensor():
    rng, dims, coords, vals = _tiny(3, 31, n=150)
    B = oracle.csf_from_coo(dims, coords, vals)
    T = oracle.dense_from_csf(B)
    k0, l0 = 1, 2
    C = np.zeros((dims[1], 1), np.float32)
    D = np.zeros((dims[2], 1), np.float32)
    C[k0, 0] = 1
    D[l0, 0] = 1
    A = oracle.mttkrp(B, [C, D], "taco", fp64=True)
    assert np.array_equal(A[:, 0], T[:, k0, l0])


@pytest.mark.parametrize("order", [3, 4])
def test_linearity(order):
    rng, dims, coords, vals = _tiny(order, 40 + order, n=120)
    B1 = oracle.csf_from_coo(dims, coords, vals)
    B2 = oracle.csf_from_coo(dims, coords, 2.0 * vals)  # powers of two scale exactly
    facs = [rng.uniform(-1, 1, (dims[m], 5)).astype(np.float32) for m in range(1, order)]
    a1 = oracle.mttkrp(B1, facs, "taco")
    a2 = oracle.mttkrp(B2, facs, "taco")
    assert np.allclose(a2, 2.0 * a1, rtol=1e-12, atol=1e-12)
    f2 = [f.copy() for f in facs]
    f2[0] = (f2[0] * 4).astype(np.float32)
    a3 = oracle.mttkrp(B1, f2, "taco")
    assert np.allclose(a3, 4 * a1, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_ttv_oracle(order):
    rng, dims, coords, vals = _tiny(order, 50 + order, n=100)
    B = oracle.csf_from_coo(dims, coords, vals)
    c = rng.uniform(-1, 1, dims[-1]).astype(np.float32)
    ref = oracle.ttv_dense_fibers(B, c)
    for schedule in ("taco", "unscheduled"):
        assert oracle.rel_l2(oracle.ttv(B, c, schedule, True), ref) < 1e-14
        assert oracle.rel_l2(oracle.ttv(B, c, schedule, False), ref) < 1e-6
    # KAT: all ones -> fiber lengths; e_k -> the single entries B(f, k)
    ones = oracle.ttv(oracle.csf_from_coo(dims, coords, np.ones_like(vals)), np.ones(dims[-1], np.float32))
    lens = np.diff(np.asarray(B.pos[order - 1]))
    Bu = oracle.csf_from_coo(dims, coords, np.ones_like(vals))
    Bu.vals = np.ones_like(Bu.vals)
    assert np.array_equal(oracle.ttv(Bu, np.ones(dims[-1], np.float32)), lens)
    del ones
    k = 1
    e = np.zeros(dims[-1], np.float32)
    e[k] = 1
    y = oracle.ttv(B, e, "taco", True)
    exp = np.zeros_like(y)
    pl = np.asarray(B.pos[order - 1])
    cl = np.asarray(B.crd[order - 1])
    for f in range(y.shape[0]):
        for q in range(pl[f], pl[f + 1]):
            if cl[q] == k:
                exp[f] = B.vals[q]
    assert np.array_equal(y, exp)


def test_planted_mistakes_are_caught():
    """A plausible bug in an MTTKRP must fail the brute-force comparison."""
    rng, dims, coords, vals = _tiny(3, 61, n=150)
    B = oracle.csf_from_coo(dims, coords, vals)
    facs = [rng.uniform(-1, 1, (dims[m], 4)).astype(np.float32) for m in (1, 2)]
    ref = oracle.mttkrp_dense(oracle.dense_from_csf(B), facs)
    # swapped factors (C <-> D on a non-cubic tensor is caught by shape; use a
    # cubic tensor)
    cub = (5, 5, 5)
    coords5 = np.stack([rng.integers(0, 5, 150) for _ in range(3)]).astype(np.int32)
    B5 = oracle.csf_from_coo(cub, coords5, vals)
    f5 = [rng.uniform(-1, 1, (5, 4)).astype(np.float32) for _ in range(2)]
    good = oracle.mttkrp(B5, f5, "taco")
    swapped = oracle.mttkrp(B5, [f5[1], f5[0]], "taco")
    assert oracle.rel_l2(swapped, oracle.mttkrp_dense(oracle.dense_from_csf(B5), f5)) > 1e-3
    assert oracle.rel_l2(good, oracle.mttkrp_dense(oracle.dense_from_csf(B5), f5)) < 1e-12
    # dropped nonzero
    Bd = oracle.csf_from_coo(dims, coords, vals)
    Bd.vals = Bd.vals.copy()
    Bd.vals[3] = 0
    assert oracle.rel_l2(oracle.mttkrp(Bd, facs, "taco"), ref) > 1e-4
    # off-by-one coordinate
    Bo = oracle.csf_from_coo(dims, coords, vals)
    Bo.crd[2] = (np.asarray(Bo.crd[2]) + 1) % dims[2]
    assert oracle.rel_l2(oracle.mttkrp(Bo, facs, "taco"), ref) > 1e-3


def test_golden_fixtures():
    """FP64 brute-force vectors committed under tests/golden (made by
    tests/golden/make_golden.py with numpy only)."""
    path = os.path.join(HERE, "golden", "mttkrp_ttv_small.json")
    with open(path) as f:
        cases = json.load(f)
    assert len(cases) >= 4
    for case in cases:
        dims = tuple(case["dims"])
        coords = np.array(case["coords"], np.int32)
        vals = np.array(case["vals"], np.float32)
        B = oracle.csf_from_coo(dims, coords, vals)
        facs = [np.array(f, np.float32) for f in case["factors"]]
        A = oracle.mttkrp(B, facs, "taco", fp64=True)
        assert np.allclose(A, np.array(case["mttkrp"]), rtol=1e-12, atol=1e-12)
        c = np.array(case["c"], np.float32)
        y = oracle.ttv(B, c, "taco", fp64=True)
        assert np.allclose(y, np.array(case["ttv"]), rtol=1e-12, atol=1e-12)
        # the product's builder reproduces the golden CSF structure bit-exactly
        P = tt.Csf.from_coo(dims, coords, vals)
        for m in range(len(dims)):
            assert np.array_equal(P.pos[m], np.array(case["pos"][m], np.int32))
            assert np.array_equal(P.crd[m], np.array(case["crd"][m], np.int32))
        assert np.array_equal(P.vals, np.array(case["csf_vals"], np.float32))


def test_reference_search_is_ambiguous_on_empty_runs():
    """taco_binarySearchBefore (gpu_library.h:185-205) returns an arbitrary
    slot among equal pos values; the product's upper-bound search returns the
    last one, i.e. the node that actually contains the target."""
    pos = np.array([0, 3, 3, 3, 3, 7, 9], np.int32)  # nodes 1..3 empty
    amb = oracle.taco_search_before(pos, 0, 6, 3)
    assert pos[amb] == 3 and amb != 4  # any match, here not the owner
    assert oracle.upper_bound_minus_one(pos, 0, 6, 3) == 4
    for key in range(0, 9):
        x = oracle.upper_bound_minus_one(pos, 0, 6, key)
        assert pos[x] <= key < pos[x + 1]
