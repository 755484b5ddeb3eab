"""Batched persistent MTTKRP (every mode of several uneven tensors in one
launch) against per-task launches and the oracle, including hipGraph
replays checked against the eager result."""

import numpy as np
import pytest

import oracle
import tentaco_amd as tt
from helpers import REL_L2_TOL

pytestmark = pytest.mark.gpu


def _tensors(seed, n=5):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n):
        order = int(rng.integers(3, 6))
        dims = [int(10 ** rng.uniform(1.5, 3.2)) for _ in range(order)]
        nnz = int(10 ** rng.uniform(3, 4.7))
        out.append(tt.Csf.synth(order, dims, nnz, slice_s=float(rng.uniform(0.8, 1.5)),
                                seed=seed * 100 + t))
    return out


def _setup(tensors, R, seed):
    rng = np.random.default_rng(seed)
    tasks, hosts, facs_h, facs_d, outs, rows = [], [], [], [], [], []
    for B in tensors:
        F = [rng.uniform(-1, 1, (d, R)).astype(np.float32) for d in B.dims]
        Fd = [tt.DeviceArray.from_numpy(f) for f in F]
        for m, Bm in enumerate(B.mode_csfs()):
            rest = [k for k in range(B.order) if k != m]
            tasks.append(tt.DeviceCsf(Bm))
            hosts.append(Bm)
            facs_h.append([F[k] for k in rest])
            facs_d.append([Fd[k] for k in rest])
            outs.append(tt.DeviceArray.empty((B.dims[m], R)))
            rows.append(B.dims[m])
    return tasks, hosts, facs_h, facs_d, outs, rows


def test_batch_matches_oracle_every_config(gpu):
    R = 32
    tensors = _tensors(3)
    tasks, hosts, facs_h, facs_d, outs, rows = _setup(tensors, R, 4)
    refs = [oracle.mttkrp(B, f, "taco", fp64=True) for B, f in zip(hosts, facs_h)]
    bt = tt.Batch(tasks, R)
    bt.set_operands(facs_d, outs, rows)
    configs = [(sc, f, b, g) for sc in (0, 1) for f in (0, 1)
               for b in (128, 256, 512) for g in (4, 16, 64)]
    for sched, fam, blk, grp in configs:
        bt.configure(fam, blk, grp)
        bt.set_schedule(sched)
        assert bt.schedule == sched
        for o in outs:  # poison: every run must overwrite
            o.upload(np.full(o.shape, 9.0, np.float32))
        bt.run()
        for t, (o, ref) in enumerate(zip(outs, refs)):
            e = oracle.rel_l2(o.numpy(), ref)
            assert e <= REL_L2_TOL, (sched, fam, blk, grp, t, e)


def test_batch_graph_replays_match_eager(gpu):
    R = 32
    tensors = _tensors(5, n=4)
    tasks, hosts, facs_h, facs_d, outs, rows = _setup(tensors, R, 6)
    bt = tt.Batch(tasks, R)
    bt.set_operands(facs_d, outs, rows)
    bt.configure(1, 256, 16)
    bt.run()
    tt.synchronize()
    eager = [o.numpy() for o in outs]
    bt.capture()
    scratch = tt.DeviceArray.empty((1 << 20,))
    for rep in range(4):
        # the program's own eager work between replays
        scratch.upload(np.full((1 << 20,), float(rep), np.float32))
        H = tasks[0]
        H.mttkrp(facs_d[0], outs[0], R, tt.VARIANT_DEFAULT)
        tt.synchronize()
        bt.launch()
        tt.synchronize()
        for t, o in enumerate(outs):
            e = oracle.rel_l2(o.numpy(), eager[t])
            assert e <= REL_L2_TOL, (rep, t, e)


def test_batch_autotune_picks_a_valid_config(gpu):
    R = 32
    tensors = _tensors(7, n=3)
    tasks, hosts, facs_h, facs_d, outs, rows = _setup(tensors, R, 8)
    bt = tt.Batch(tasks, R)
    bt.set_operands(facs_d, outs, rows)
    (fam, blk, grp), ms = bt.autotune(reps=2)
    assert fam in (0, 1) and blk in (128, 256, 512) and grp in (4, 8, 16, 32, 64) and ms > 0
    bt.run()
    for B, f, o in zip(hosts, facs_h, outs):
        assert oracle.rel_l2(o.numpy(), oracle.mttkrp(B, f, "taco", fp64=True)) <= REL_L2_TOL


@pytest.mark.parametrize("R", [16, 64])
def test_batch_other_ranks_eager_and_graph(gpu, R):
    """R = 16 and 64 (four and sixteen lanes per factor row) through both
    edge families, eager and replayed, against the oracle; every output is
    poisoned before each run."""
    tensors = _tensors(9 + R, n=4)
    tasks, hosts, facs_h, facs_d, outs, rows = _setup(tensors, R, 10 + R)
    refs = [oracle.mttkrp(B, f, "taco", fp64=True) for B, f in zip(hosts, facs_h)]
    bt = tt.Batch(tasks, R)
    bt.set_operands(facs_d, outs, rows)
    for fam, blk, grp in ((0, 256, 8), (1, 128, 64)):
        bt.configure(fam, blk, grp)
        for o in outs:
            o.upload(np.full(o.shape, 9.0, np.float32))
        bt.run()
        for t, (o, ref) in enumerate(zip(outs, refs)):
            assert oracle.rel_l2(o.numpy(), ref) <= REL_L2_TOL, (R, fam, t)
        bt.capture()
        for rep in range(2):
            for o in outs:
                o.upload(np.full(o.shape, -4.0, np.float32))
            bt.launch()
            tt.synchronize()
            for t, (o, ref) in enumerate(zip(outs, refs)):
                assert oracle.rel_l2(o.numpy(), ref) <= REL_L2_TOL, (R, fam, rep, t)
    bt.close()
