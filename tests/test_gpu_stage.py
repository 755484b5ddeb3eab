"""Stream-ordered host staging: pinned host arrays, asynchronous CSF upload
and host-operand MTTKRP on caller streams, several handles in flight."""

import numpy as np
import pytest

import oracle
import tentaco_amd as tt
from helpers import REL_L2_TOL, random_csf

pytestmark = pytest.mark.gpu


def test_pinned_array_roundtrip(gpu):
    a = tt.pinned_empty((1000, 32))
    assert isinstance(a, tt.PinnedArray) and a.shape == (1000, 32) and a.dtype == np.float32
    a[...] = np.arange(32000, dtype=np.float32).reshape(1000, 32)
    d = tt.DeviceArray.from_numpy(np.ascontiguousarray(a))
    assert np.array_equal(d.numpy(), a)
    v = a[10:20]
    del a
    assert v[0, 0] == 320.0  # a view keeps the pinned memory alive


@pytest.mark.parametrize("pinned", [True, False])
def test_async_upload_and_host_mttkrp(gpu, pinned):
    rng = np.random.default_rng(4)
    s1, s2 = tt.Stream(), tt.Stream()
    cases = []
    for order, dims, R in ((3, (50, 40, 30), 32), (4, (20, 18, 16, 14), 16), (5, (9, 8, 7, 6, 5), 64)):
        B = random_csf(rng, dims, 4000)
        facs = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in dims[1:]]
        if pinned:
            facs = [tt.pinned_copy(f) for f in facs]
        cases.append((B, facs, R))
    handles = [tt.DeviceCsf.upload_async(B, s1 if k % 2 == 0 else s2) for k, (B, _, _) in enumerate(cases)]
    outs = []
    for k, ((B, facs, R), H) in enumerate(zip(cases, handles)):
        st = s1 if k % 2 == 0 else s2
        A = tt.pinned_empty((B.dims[0], R)) if pinned else np.empty((B.dims[0], R), np.float32)
        A[...] = 7.0  # poison: the call must overwrite every row
        H.mttkrp_host_async(facs, A, R, st, variant=tt.variants().index("ep_b256_g16"))
        outs.append(A)
    s1.synchronize()
    s2.synchronize()
    for (B, facs, R), A in zip(cases, outs):
        ref = oracle.mttkrp(B, [np.asarray(f) for f in facs], "taco", fp64=True)
        assert oracle.rel_l2(np.asarray(A), ref) <= REL_L2_TOL
    # a second round on the same handles reuses their device buffers
    for k, ((B, facs, R), H, A) in enumerate(zip(cases, handles, outs)):
        A[...] = -3.0
        H.mttkrp_host_async(facs, A, R, s1)
    s1.synchronize()
    for (B, facs, R), A in zip(cases, outs):
        ref = oracle.mttkrp(B, [np.asarray(f) for f in facs], "taco", fp64=True)
        assert oracle.rel_l2(np.asarray(A), ref) <= REL_L2_TOL
    for H in handles:
        H.close()


def test_host_async_rejects_bad_operands(gpu):
    rng = np.random.default_rng(5)
    B = random_csf(rng, (30, 20, 10), 500)
    st = tt.Stream()
    H = tt.DeviceCsf.upload_async(B, st)
    st.synchronize()
    facs = [rng.uniform(-1, 1, (d, 16)).astype(np.float32) for d in (20, 10)]
    with pytest.raises(tt.TentacoError):
        H.mttkrp_host_async(facs, np.empty((30, 32), np.float32), 16, st)  # R mismatch
    with pytest.raises(tt.TentacoError):
        H.mttkrp_host_async([facs[0], facs[1][:5]], np.empty((30, 16), np.float32), 16, st)
    H.close()
