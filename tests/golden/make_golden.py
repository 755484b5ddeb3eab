"""Writes mttkrp_ttv_small.json: tiny CSF MTTKRP/TTV cases whose expected
outputs come from pure-Python / numpy brute force only (dense FP64 einsum for
MTTKRP, per-fiber loops for TTV, dict-based CSF construction), independent of
both the oracle's loop nests and the product. Run: python make_golden.py"""

import json
import os

import numpy as np


def build_csf(dims, coords, vals):
    # sorted unique coordinates, duplicates summed in input order (FP32)
    acc = {}
    order_seen = []
    for q in range(len(vals)):
        key = tuple(int(coords[m][q]) for m in range(len(dims)))
        if key not in acc:
            acc[key] = np.float32(0.0)
            order_seen.append(key)
        acc[key] = np.float32(acc[key] + np.float32(vals[q]))
    keys = sorted(acc)
    N = len(dims)
    pos = [[0, dims[0]]]
    crd = [list(range(dims[0]))]
    prev_level_nodes = [(i,) for i in range(dims[0])]
    for m in range(1, N):
        nodes = sorted({k[: m + 1] for k in keys})
        index = {p: j for j, p in enumerate(prev_level_nodes)}
        counts = [0] * len(prev_level_nodes)
        for nd in nodes:
            counts[index[nd[:m]]] += 1
        p = [0]
        for cnt in counts:
            p.append(p[-1] + cnt)
        pos.append(p)
        crd.append([nd[m] for nd in nodes])
        prev_level_nodes = nodes
    return pos, crd, [float(acc[k]) for k in keys], keys


def main():
    rng = np.random.default_rng(2024)
    cases = []
    for order, n, R in ((2, 12, 3), (3, 30, 4), (3, 40, 2), (4, 50, 3), (5, 60, 2)):
        dims = [int(x) for x in rng.integers(3, 6, order)]
        coords = [[int(v) for v in rng.integers(0, d, n)] for d in dims]
        vals = [float(np.float32(v)) for v in rng.uniform(-1, 1, n)]
        pos, crd, cvals, keys = build_csf(dims, coords, vals)
        facs = [rng.uniform(-1, 1, (dims[m], R)).astype(np.float32) for m in range(1, order)]
        T = np.zeros(dims)
        for k, v in zip(keys, cvals):
            T[k] += v
        letters = "abcdefgh"
        expr = ",".join([letters[:order]] + [f"{letters[m]}z" for m in range(1, order)]) + "->az"
        A = np.einsum(expr, T, *[f.astype(np.float64) for f in facs])
        c = rng.uniform(-1, 1, dims[-1]).astype(np.float32)
        # TTV per level order-2 node, in node order
        # (for order 2 the fibers are the dense root's slices, empty ones too)
        fibers = ([(i,) for i in range(dims[0])] if order == 2
                  else sorted({k[:-1] for k in keys}))
        y = []
        for fb in fibers:
            y.append(sum(float(v) * float(c[k[-1]]) for k, v in zip(keys, cvals) if k[:-1] == fb))
        cases.append({
            "dims": dims, "coords": coords, "vals": vals,
            "factors": [f.tolist() for f in facs], "mttkrp": A.tolist(),
            "c": c.tolist(), "ttv": y,
            "pos": pos, "crd": crd, "csf_vals": cvals,
        })
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mttkrp_ttv_small.json")
    with open(out, "w") as f:
        json.dump(cases, f)
    print("wrote", out, len(cases), "cases")


if __name__ == "__main__":
    main()
