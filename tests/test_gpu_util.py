"""Device copy (the counter-calibration kernel) and the per-launch kernel
timer."""

import numpy as np
import pytest

import oracle
import tentaco_amd as tt
from helpers import REL_L2_TOL, factors_for, random_csf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [16, 4096, 1000, (1 << 20) + 12])
def test_dev_copy_exact(gpu, nbytes):
    rng = np.random.default_rng(nbytes)
    n = (nbytes + 3) // 4
    a = rng.uniform(-1, 1, n).astype(np.float32)
    src = tt.DeviceArray.from_numpy(a)
    dst = tt.DeviceArray.from_numpy(np.full(n, 7.0, np.float32))
    tt.dev_copy(dst, src, nbytes)
    tt.synchronize()
    got = dst.numpy().view(np.uint8)[:nbytes]
    assert np.array_equal(got, a.view(np.uint8)[:nbytes])
    # bytes past the copied range are untouched
    rest = dst.numpy().view(np.uint8)[nbytes:]
    assert np.array_equal(rest, np.full(n, 7.0, np.float32).view(np.uint8)[nbytes:])


def test_dev_copy_rejects_misaligned_and_overlap(gpu):
    src = tt.DeviceArray.empty((1024,))
    with pytest.raises(tt.TentacoError):
        tt._check(tt.lib().tentaco_dev_copy(src.ptr.value + 4, src.ptr, 64, None))
    with pytest.raises(tt.TentacoError):
        tt._check(tt.lib().tentaco_dev_copy(src.ptr.value + 64, src.ptr, 256, None))


def test_kernel_timer_counts_main_kernels(gpu):
    rng = np.random.default_rng(5)
    B = random_csf(rng, (40, 30, 20), 3000)
    R = 32
    facs = factors_for(rng, B, R)
    H = tt.DeviceCsf(B)
    fd = [tt.DeviceArray.from_numpy(f) for f in facs]
    A = tt.DeviceArray.empty((B.dims[0], R))
    names = tt.variants()
    ref = oracle.mttkrp(B, facs, "taco", fp64=True)
    for vname, kname in (("sr_b256", "slice_kernel"), ("ep_b256_g16", "edge_kernel"),
                         ("es_b512_g4", "edge_kernel")):
        v = names.index(vname)
        tt.ktimer_start()
        for _ in range(5):
            H.mttkrp(fd, A, R, v)
        total, n, k = tt.ktimer_stop()
        assert n == 5 and k == kname and total > 0, (vname, n, k, total)
        assert oracle.rel_l2(A.numpy(), ref) <= REL_L2_TOL
    # off: nothing recorded
    H.mttkrp(fd, A, R, names.index("sr_b256"))
    tt.ktimer_start()
    total, n, k = tt.ktimer_stop()
    assert n == 0 and total == 0.0
    # TTV's main kernel
    c = np.random.default_rng(7).uniform(-1, 1, B.dims[2]).astype(np.float32)
    cd = tt.DeviceArray.from_numpy(c)
    y = tt.DeviceArray.empty((B.nodes(1),))
    tt.ktimer_start()
    H.ttv(cd, y)
    total, n, k = tt.ktimer_stop()
    assert n == 1 and k == "ttv_unit_kernel"
    H.close()
