"""FROSTT .tns reader: bit-exact against the oracle's COO -> CSF builder on
the same parsed nonzeros, the shortest-mode-first ordering, and the error
paths (CPU only)."""

import numpy as np
import pytest

import oracle
import tentaco_amd as tt


def _write(path, coords, vals, comments=True, style=0):
    """coords: order x n, 0-based; written 1-based."""
    lines = []
    if comments:
        lines += ["# FROSTT-style header comment", ""]
    for q in range(vals.size):
        idx = [str(int(c) + 1) for c in coords[:, q]]
        v = repr(float(vals[q])) if style == 0 else f"{float(vals[q]):.6e}"
        sep = " " if q % 3 else "\t"
        lines.append(sep.join(idx + [v]) + ("  " if q % 5 == 0 else ""))
        if comments and q == vals.size // 2:
            lines.append("   # a comment in the middle")
            lines.append("")
    path.write_text("\n".join(lines) + ("\n" if style == 0 else ""))


def _coo(rng, dims, n):
    coords = np.stack([rng.integers(0, d, n) for d in dims]).astype(np.int32)
    coords[:, n // 2:n // 2 + 5] = coords[:, :5]  # duplicates
    vals = rng.uniform(-1, 1, n).astype(np.float32)
    return coords, vals


def _same(a, b):
    assert a.order == b.order and tuple(a.dims) == tuple(b.dims)
    for m in range(a.order):
        assert np.array_equal(np.asarray(a.pos[m]), np.asarray(b.pos[m])), m
        assert np.array_equal(np.asarray(a.crd[m]), np.asarray(b.crd[m])), m
    assert np.array_equal(np.asarray(a.vals), np.asarray(b.vals))


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_tns_matches_oracle_builder(tmp_path, order):
    rng = np.random.default_rng(order)
    dims = tuple(int(x) for x in rng.integers(3, 40, order))
    coords, vals = _coo(rng, dims, 3000)
    f = tmp_path / "t.tns"
    _write(f, coords, vals, style=order % 2)
    # parse values as the reader does: decimal text -> double -> float
    text_vals = np.array([float(repr(float(v))) if order % 2 == 0 else float(f"{float(v):.6e}")
                          for v in vals]).astype(np.float32)
    ident = list(range(order))
    B, used = tt.Csf.from_tns(str(f), dims=dims, mode_order=ident)
    assert used == ident
    _same(B, oracle.csf_from_coo(dims, coords, text_vals))


def test_tns_shortest_mode_first_and_inferred_dims(tmp_path):
    rng = np.random.default_rng(7)
    dims = (50, 7, 20, 7)
    coords, vals = _coo(rng, dims, 2000)
    coords[:, 0] = [d - 1 for d in dims]  # the largest index of every mode occurs
    f = tmp_path / "t.tns"
    _write(f, coords, vals)
    B, used = tt.Csf.from_tns(str(f))
    assert used == [1, 3, 2, 0]  # ascending dimension, ties by mode number
    perm_dims = tuple(dims[m] for m in used)
    assert tuple(B.dims) == perm_dims
    _same(B, oracle.csf_from_coo(perm_dims, coords[used], vals))


def test_tns_large_file_parsed_in_chunks(tmp_path):
    """More than one parse chunk (over 1 MiB of text): same result as the
    oracle, so chunk boundaries lose or duplicate nothing."""
    rng = np.random.default_rng(8)
    dims = (300, 200, 100)
    coords, vals = _coo(rng, dims, 120_000)
    f = tmp_path / "big.tns"
    _write(f, coords, vals, comments=False)
    assert f.stat().st_size > 3 << 20
    B, _ = tt.Csf.from_tns(str(f), dims=dims, mode_order=[0, 1, 2])
    _same(B, oracle.csf_from_coo(dims, coords, vals))


@pytest.mark.parametrize("body,needle", [
    ("1 2 3 1.0\n1 2 0.5\n", "line 2"),            # too few columns
    ("1 2 3 1.0\n1 2 3 4 0.5\n", "line 2"),        # too many columns
    ("1 0 3 1.0\n", "line 1"),                     # index 0 (1-based format)
    ("1 -2 3 1.0\n", "line 1"),                    # negative index
    ("1 2 3 abc\n", "line 1"),                     # bad value
    ("# only comments\n\n", "no nonzeros"),
    ("1 2 3 1.0\n4 5 6\n7 8 9 1.0\n", "line 2: expected a value"),  # not taken from line 3
])
def test_tns_rejects_malformed(tmp_path, body, needle):
    f = tmp_path / "bad.tns"
    f.write_text(body)
    with pytest.raises(tt.TentacoError) as e:
        tt.Csf.from_tns(str(f))
    assert e.value.code == tt.EINVAL and needle in str(e.value)


def test_tns_rejects_unsupported_and_missing(tmp_path):
    f = tmp_path / "o6.tns"
    f.write_text("1 1 1 1 1 1 1.0\n")
    with pytest.raises(tt.TentacoError) as e:
        tt.Csf.from_tns(str(f))
    assert e.value.code == tt.EUNSUPPORTED
    with pytest.raises(tt.TentacoError) as e:
        tt.Csf.from_tns(str(tmp_path / "missing.tns"))
    assert e.value.code == tt.EINVAL
    g = tmp_path / "d.tns"
    g.write_text("5 1 1.0\n")
    with pytest.raises(tt.TentacoError):
        tt.Csf.from_tns(str(g), dims=(4, 4))  # index beyond the given dimension
    with pytest.raises(tt.TentacoError):
        tt.Csf.from_tns(str(g), mode_order=[0, 0])


@pytest.mark.parametrize("dims,n", [
    ((1_000_000, 900_000, 800_000, 700_000, 600_000), 300_000),  # 100 key bits: per-mode digits
    ((20_000, 30_000, 40_000), 400_000),                          # 45 bits: packed 64-bit keys
])
def test_coo_builder_radix_paths_match_oracle(dims, n):
    """Both sort paths of the library's COO -> CSF builder (packed keys up to
    64 bits, per-mode digits beyond), each over several thread chunks, are
    bit-exact against the oracle's builder, duplicates included."""
    rng = np.random.default_rng(len(dims))
    coords = np.stack([rng.integers(0, d, n) for d in dims]).astype(np.int32)
    coords[:, n // 3:n // 3 + 300] = coords[:, :300]
    vals = rng.uniform(-1, 1, n).astype(np.float32)
    _same(tt.Csf.from_coo(dims, coords, vals), oracle.csf_from_coo(dims, coords, vals))
