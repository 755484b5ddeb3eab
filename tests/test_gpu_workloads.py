"""Parity at the benchmark workloads' full sizes: the 4096^3 / 500k-coordinate
MTTKRP-3 (R = 32) with the variant the bench times and with the fixed default,
the 100k^3 / 20M-nonzero TTV, and the 5-mode 50k^5 / 30M-nonzero MTTKRP
(R = 32), each against the CPU oracle on the same seeded inputs (the bench
itself only reports agreement between GPU paths)."""

import numpy as np
import pytest

import oracle
import tentaco_amd as tt
from helpers import REL_L2_TOL
from tentaco_amd import metrics

pytestmark = pytest.mark.gpu


def test_mttkrp3_4096cube_500k_r32_full_size(gpu):
    B = metrics.workload_mttkrp3_uniform()
    R = 32
    facs = metrics.factors(B, R, 1001)
    ref64 = oracle.mttkrp(B, facs, "taco", fp64=True)
    ref32 = oracle.mttkrp(B, facs, "taco", fp64=False)
    H = tt.DeviceCsf(B)
    fd = [tt.DeviceArray.from_numpy(f) for f in facs]
    A = tt.DeviceArray.empty((B.dims[0], R))
    names = tt.variants()
    for v in (names.index("sr_b256"), names.index("sr_b128"), tt.default_variant(),
              names.index("ep_b256_g16")):
        A.upload(np.full((B.dims[0], R), 3.0, np.float32))
        H.mttkrp(fd, A, R, v)
        got = A.numpy()
        assert oracle.rel_l2(got, ref64) <= REL_L2_TOL, names[v]
        assert oracle.rel_l2(got, ref32) <= REL_L2_TOL, names[v]
    H.close()


def test_ttv_100kcube_20m_zipf_full_size(gpu):
    B = metrics.workload_zipf(3, 100_000, 20_000_000, seed=1001)
    c = np.random.default_rng(1011).uniform(-1, 1, 100_000).astype(np.float32)
    ref = oracle.ttv(B, c, "taco", fp64=True)
    H = tt.DeviceCsf(B)
    y = tt.DeviceArray.from_numpy(np.full(B.nodes(1), np.nan, np.float32))
    H.ttv(tt.DeviceArray.from_numpy(c), y)
    assert oracle.rel_l2(y.numpy(), ref) <= REL_L2_TOL
    H.close()


def test_mttkrp5_50k_30m_zipf_r32_full_size(gpu):
    B = metrics.workload_zipf(5, 50_000, 30_000_000, seed=1007)
    R = 32
    facs = metrics.factors(B, R, 1035)
    ref = oracle.mttkrp(B, facs, "taco", fp64=True)
    H = tt.DeviceCsf(B)
    fd = [tt.DeviceArray.from_numpy(f) for f in facs]
    A = tt.DeviceArray.empty((B.dims[0], R))
    names = tt.variants()
    for v in (names.index("ep_b512_g64"), names.index("es_b256_g16")):
        H.mttkrp(fd, A, R, v)
        assert oracle.rel_l2(A.numpy(), ref) <= REL_L2_TOL, names[v]
    H.close()
