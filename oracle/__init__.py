"""ORACLE - test infrastructure only.

CPU restatement of zhang677/tenTACO's CSF MTTKRP/TTV path, used as the
checker of the tentaco HIP library. Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import this package, and only to
check results or to time the CPU baseline; the product (``tentaco_amd``) never
imports it.

Contents
  * ``liboracle.so`` (``cpu_taco.cpp``): the reference CPU TACO loop nests
    (mttkrp{3,4,5}_cpu_taco / _unscheduled / _workspace, ttv_csf_cpu_taco /
    _unscheduled), run either in the reference's typing (64-bit indices, FP64)
    or in the GPU's typing (int32, FP32).
  * ``csf_from_coo``: an independent numpy COO -> CSF builder (lexsort, stable,
    duplicates summed in input order) that pins the product's builder
    bit-exactly.
  * dense brute force (``dense_from_csf``, ``mttkrp_dense``, ``ttv_dense``) for
    tiny tensors: pins the loop nests themselves.
  * ``taco_search_before``: the reference's partition search
    (gpu_library.h:185-205) in Python, for the partition tests.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64
from types import SimpleNamespace
from typing import List, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

SCHEDULES = {"taco": 0, "unscheduled": 1, "workspace": 2}


class _Taco(ctypes.Structure):
    # taco_tensor_t (taco_tensor_t.h:6-15 of the reference)
    _fields_ = [
        ("order", c_int32),
        ("dimensions", POINTER(c_int32)),
        ("csize", c_int32),
        ("mode_ordering", POINTER(c_int32)),
        ("mode_types", POINTER(c_int)),
        ("indices", POINTER(POINTER(POINTER(c_int32)))),
        ("vals", POINTER(c_float)),
        ("vals_size", c_int32),
    ]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = POINTER(_Taco)
        L.oracle_mttkrp_f64.argtypes = [c_int, c_int, POINTER(c_double), c_int64, c_int, P, POINTER(P)]
        L.oracle_mttkrp_f32.argtypes = [c_int, c_int, POINTER(c_float), c_int64, c_int, P, POINTER(P)]
        L.oracle_ttv_f64.argtypes = [c_int, c_int, POINTER(c_double), P, P]
        L.oracle_ttv_f32.argtypes = [c_int, c_int, POINTER(c_float), P, P]
        L.oracle_time_mttkrp.argtypes = [c_int, c_int, c_int, c_int64, c_int, P, POINTER(P), c_int,
                                         c_int64, POINTER(c_double), POINTER(c_double)]
        L.oracle_time_ttv.argtypes = [c_int, c_int, P, P, c_int, c_int64, POINTER(c_double),
                                      POINTER(c_double)]
        L.oracle_num_threads.restype = c_int
        L.oracle_set_num_threads.argtypes = [c_int]
        L.oracle_set_num_threads.restype = None
        L.oracle_search_before.argtypes = [POINTER(c_int32), c_int32, c_int32, c_int32]
        L.oracle_search_before.restype = c_int32
        _lib = L
    return _lib


def _taco_csf(B):
    dims = np.ascontiguousarray(B.dims, dtype=np.int32)
    keep = [dims]
    lvl = (POINTER(c_int32) * 2 * B.order)()
    ptrs = (POINTER(POINTER(c_int32)) * B.order)()
    for m in range(B.order):
        p = np.ascontiguousarray(B.pos[m], dtype=np.int32)
        c = np.ascontiguousarray(B.crd[m], dtype=np.int32)
        keep += [p, c]
        lvl[m][0] = p.ctypes.data_as(POINTER(c_int32))
        lvl[m][1] = c.ctypes.data_as(POINTER(c_int32))
        ptrs[m] = ctypes.cast(lvl[m], POINTER(POINTER(c_int32)))
    vals = np.ascontiguousarray(B.vals, dtype=np.float32)
    keep += [vals, lvl, ptrs]
    t = _Taco()
    t.order = B.order
    t.dimensions = dims.ctypes.data_as(POINTER(c_int32))
    t.indices = ctypes.cast(ptrs, POINTER(POINTER(POINTER(c_int32))))
    t.vals = vals.ctypes.data_as(POINTER(c_float))
    t.vals_size = vals.shape[0]
    return t, keep


def _taco_dense(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    dims = np.ascontiguousarray(a.shape, dtype=np.int32)
    t = _Taco()
    t.order = a.ndim
    t.dimensions = dims.ctypes.data_as(POINTER(c_int32))
    t.vals = a.ctypes.data_as(POINTER(c_float))
    return t, [a, dims]


def _factor_array(factors):
    fs = [_taco_dense(f) for f in factors]
    arr = (POINTER(_Taco) * len(fs))(*[ctypes.pointer(t) for t, _ in fs])
    return arr, fs


def mttkrp(B, factors: Sequence[np.ndarray], schedule: str = "taco", fp64: bool = True,
           out_rows: int | None = None) -> np.ndarray:
    """Reference CPU MTTKRP (mttkrp{3,4,5}_csf_cpu_taco*.h) on a CSF ``B``."""
    L = load()
    R = int(factors[0].shape[1])
    rows = B.dims[0] if out_rows is None else out_rows
    tB, kB = _taco_csf(B)
    arr, kf = _factor_array(factors)
    if fp64:
        A = np.empty((rows, R), np.float64)
        rc = L.oracle_mttkrp_f64(B.order, SCHEDULES[schedule], A.ctypes.data_as(POINTER(c_double)),
                                 rows, R, ctypes.byref(tB), arr)
    else:
        A = np.empty((rows, R), np.float32)
        rc = L.oracle_mttkrp_f32(B.order, SCHEDULES[schedule], A.ctypes.data_as(POINTER(c_float)),
                                 rows, R, ctypes.byref(tB), arr)
    if rc:
        raise RuntimeError(f"oracle mttkrp failed ({rc})")
    del kB, kf
    return A


def ttv(B, c: np.ndarray, schedule: str = "taco", fp64: bool = True) -> np.ndarray:
    """Reference CPU TTV (ttv_csf_cpu_taco.h:24-68 / _unscheduled)."""
    L = load()
    nf = int(B.crd[B.order - 2].shape[0])
    tB, kB = _taco_csf(B)
    tc, kc = _taco_dense(np.ascontiguousarray(c, np.float32))
    sched = 0 if schedule == "taco" else 1
    if fp64:
        y = np.empty(nf, np.float64)
        rc = L.oracle_ttv_f64(B.order, sched, y.ctypes.data_as(POINTER(c_double)), ctypes.byref(tB),
                              ctypes.byref(tc))
    else:
        y = np.empty(nf, np.float32)
        rc = L.oracle_ttv_f32(B.order, sched, y.ctypes.data_as(POINTER(c_float)), ctypes.byref(tB),
                              ctypes.byref(tc))
    if rc:
        raise RuntimeError(f"oracle ttv failed ({rc})")
    del kB, kc
    return y


def time_mttkrp(B, factors, schedule="taco", fp64=True, trials=25, flush_doubles=0):
    """(min_ms, median_ms) of the reference CPU nest, each trial after a cache
    flush (TIME_COLD / RUN, timers.h:112-130)."""
    L = load()
    R = int(factors[0].shape[1])
    tB, kB = _taco_csf(B)
    arr, kf = _factor_array(factors)
    mn, md = c_double(), c_double()
    rc = L.oracle_time_mttkrp(B.order, SCHEDULES[schedule], int(fp64), B.dims[0], R,
                              ctypes.byref(tB), arr, trials, flush_doubles, ctypes.byref(mn),
                              ctypes.byref(md))
    if rc:
        raise RuntimeError("oracle timing failed")
    del kB, kf
    return mn.value, md.value


def time_ttv(B, c, fp64=True, trials=25, flush_doubles=0):
    L = load()
    tB, kB = _taco_csf(B)
    tc, kc = _taco_dense(np.ascontiguousarray(c, np.float32))
    mn, md = c_double(), c_double()
    rc = L.oracle_time_ttv(B.order, int(fp64), ctypes.byref(tB), ctypes.byref(tc), trials,
                           flush_doubles, ctypes.byref(mn), ctypes.byref(md))
    if rc:
        raise RuntimeError("oracle timing failed")
    del kB, kc
    return mn.value, md.value


def num_threads() -> int:
    return int(load().oracle_num_threads())


def set_num_threads(n: int) -> None:
    """OpenMP threads of the timed CPU nests (later calls use n)."""
    load().oracle_set_num_threads(int(n))


# ---------------------------------------------------------------------------
def csf_from_coo(dims: Sequence[int], coords: np.ndarray, vals: np.ndarray):
    """Independent COO -> CSF (dense root), duplicates summed in input order
    in FP32. Mirrors what splatt_csf_load + compress_top_level produce for a
    lexicographic mode order (ds.h:351-384, 426-436)."""
    dims = tuple(int(d) for d in dims)
    order = len(dims)
    coords = np.asarray(coords, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float32)
    n = vals.shape[0]
    # stable lexicographic order: np.lexsort sorts by the last key first
    perm = np.lexsort(coords[::-1]) if n else np.zeros(0, np.int64)
    sc = coords[:, perm]
    sv = vals[perm]
    if n:
        newnz = np.ones(n, bool)
        newnz[1:] = np.any(sc[:, 1:] != sc[:, :-1], axis=0)
    else:
        newnz = np.zeros(0, bool)
    starts = np.flatnonzero(newnz)
    uvals = sv[starts].copy()
    # sum duplicates sequentially in input (stable) order, FP32
    ends = np.append(starts[1:], n)
    for g in np.flatnonzero(ends - starts > 1):
        acc = np.float32(0.0)
        for q in range(starts[g], ends[g]):
            acc = np.float32(acc + sv[q])
        uvals[g] = acc
    uc = sc[:, starts]
    nu = uc.shape[1]
    pos: List[np.ndarray] = [np.array([0, dims[0]], np.int32)]
    crd: List[np.ndarray] = [np.arange(dims[0], dtype=np.int32)]
    # node at level m starts where the prefix (levels 0..m) changes
    prev_nodes_start = None
    for m in range(1, order):
        if nu:
            new = np.ones(nu, bool)
            new[1:] = np.any(uc[: m + 1, 1:] != uc[: m + 1, :-1], axis=0)
        else:
            new = np.zeros(0, bool)
        idx = np.flatnonzero(new)
        crd.append(uc[m, idx].astype(np.int32))
        if m == 1:
            counts = np.bincount(uc[0, idx], minlength=dims[0]) if nu else np.zeros(dims[0], np.int64)
            p = np.zeros(dims[0] + 1, np.int64)
            p[1:] = np.cumsum(counts)
            pos.append(p.astype(np.int32))
        else:
            # parent of each level-m node: last level-(m-1) node starting at or before it
            parent_starts = prev_nodes_start
            parent_of = np.searchsorted(parent_starts, idx, side="right") - 1
            counts = np.bincount(parent_of, minlength=parent_starts.shape[0])
            p = np.zeros(parent_starts.shape[0] + 1, np.int64)
            p[1:] = np.cumsum(counts)
            pos.append(p.astype(np.int32))
        prev_nodes_start = idx
    return SimpleNamespace(order=order, dims=dims, pos=pos, crd=crd, vals=uvals.astype(np.float32))


def coo_from_csf(B):
    """CSF -> (coords [order x nnz], vals): the inverse walk."""
    order = B.order
    nnz = int(B.vals.shape[0])
    coords = np.zeros((order, nnz), np.int64)
    # node index of every nonzero at each level, top-down
    node = np.arange(nnz)
    for m in range(order - 1, -1, -1):
        coords[m] = np.asarray(B.crd[m])[node] if m > 0 else np.asarray(B.crd[0])[node]
        if m > 0:
            pm = np.asarray(B.pos[m], np.int64)
            node = np.searchsorted(pm, node, side="right") - 1
    return coords, np.asarray(B.vals)


def dense_from_csf(B) -> np.ndarray:
    """Dense FP64 tensor of a (tiny) CSF."""
    coords, vals = coo_from_csf(B)
    T = np.zeros(B.dims, np.float64)
    np.add.at(T, tuple(coords), vals.astype(np.float64))
    return T


def mttkrp_dense(T: np.ndarray, factors: Sequence[np.ndarray]) -> np.ndarray:
    """Brute-force FP64 mode-0 MTTKRP of a dense tensor:
    A(i,r) = sum_{k,l,...} T(i,k,l,...) C(k,r) D(l,r) ..."""
    order = T.ndim
    letters = "abcdefgh"
    sub = letters[:order]
    terms = [sub] + [f"{letters[m]}z" for m in range(1, order)]
    expr = ",".join(terms) + "->az"
    return np.einsum(expr, T, *[np.asarray(f, np.float64) for f in factors])


def ttv_dense_fibers(B, c: np.ndarray) -> np.ndarray:
    """Brute-force FP64 TTV per level N-2 fiber from the COO form."""
    coords, vals = coo_from_csf(B)
    order = B.order
    nnz = vals.shape[0]
    # fiber index of every nonzero = its level order-2 node
    pl = np.asarray(B.pos[order - 1], np.int64)
    fib = np.searchsorted(pl, np.arange(nnz), side="right") - 1
    y = np.zeros(int(B.crd[order - 2].shape[0]), np.float64)
    np.add.at(y, fib, vals.astype(np.float64) * np.asarray(c, np.float64)[coords[order - 1]])
    return y


def taco_search_before(a: np.ndarray, lo: int, hi: int, target: int) -> int:
    """taco_binarySearchBefore (gpu_library.h:185-205), as the C oracle runs it."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    return int(load().oracle_search_before(a.ctypes.data_as(POINTER(c_int32)), lo, hi, target))


def upper_bound_minus_one(a: np.ndarray, lo: int, hi: int, key: int) -> int:
    """Largest x in [lo, hi] with a[x] <= key, by linear scan (the partition
    semantics the product's search_le implements)."""
    x = lo
    for i in range(lo, hi + 1):
        if a[i] <= key:
            x = i
    return x


def rel_l2(got: np.ndarray, ref: np.ndarray) -> float:
    """||got - ref|| / max(||ref||, tiny) in FP64."""
    g = np.asarray(got, np.float64).ravel()
    r = np.asarray(ref, np.float64).ravel()
    den = max(float(np.linalg.norm(r)), 1e-300)
    return float(np.linalg.norm(g - r)) / den
