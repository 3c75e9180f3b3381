"""Vector file readers and the canonical-output contract checks (CPU).

The contract checks are pinned on outputs whose answer is known without the
checker: the oracle's canonical ids of a small join (must pass), and the same
array with one planted defect per check (must fail exactly that check)."""
import gzip
import hashlib
import struct

import numpy as np
import pytest

from tensorjoin_variants_amd import vecio


def test_fvecs_roundtrip_and_subset(tmp_path):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((37, 13)).astype(np.float32)
    p = tmp_path / "a.fvecs"
    vecio.write_fvecs(p, x)
    assert p.stat().st_size == 37 * (13 + 1) * 4
    # the record layout, written independently of write_fvecs
    raw = np.fromfile(p, dtype="<i4").reshape(37, 14)
    assert np.all(raw[:, 0] == 13)
    assert np.array_equal(raw[:, 1:].view("<f4"), x)
    assert np.array_equal(vecio.read_fvecs(p), x)
    rows = np.array([5, 0, 36, 5])
    assert np.array_equal(vecio.read_fvecs(p, rows=rows), x[rows])
    assert np.array_equal(vecio.read_vectors(p), x)
    with pytest.raises(ValueError):
        vecio.read_fvecs(p, dim=12)
    with pytest.raises(ValueError):
        vecio.read_fvecs(p, n_rows=36)


def test_fvecs_malformed(tmp_path):
    p = tmp_path / "bad.fvecs"
    rec = np.zeros((3, 5), dtype="<i4")
    rec[:, 0] = 4
    rec[1, 0] = 3  # inconsistent header
    rec.tofile(p)
    with pytest.raises(ValueError, match="inconsistent"):
        vecio.read_fvecs(p)
    q = tmp_path / "trunc.fvecs"
    np.concatenate([np.array([4], "<i4"), np.zeros(3, "<i4")]).tofile(q)  # 16 of 20 bytes
    with pytest.raises(ValueError, match="whole number"):
        vecio.read_fvecs(q)
    r = tmp_path / "nan.fvecs"
    vecio.write_fvecs(r, np.array([[1.0, np.nan]], np.float32))
    with pytest.raises(ValueError, match="non-finite"):
        vecio.read_fvecs(r)


def test_bvecs_ivecs(tmp_path):
    rng = np.random.default_rng(2)
    b = rng.integers(0, 256, (9, 7)).astype(np.uint8)
    with open(tmp_path / "a.bvecs", "wb") as f:
        for row in b:
            f.write(struct.pack("<i", 7) + row.tobytes())
    assert np.array_equal(vecio.read_bvecs(tmp_path / "a.bvecs"), b.astype(np.float32))
    iv = rng.integers(-2 ** 24, 2 ** 24, (4, 3)).astype(np.int32)
    with open(tmp_path / "a.ivecs", "wb") as f:
        for row in iv:
            f.write(struct.pack("<i", 3) + row.astype("<i4").tobytes())
    assert np.array_equal(vecio.read_ivecs(tmp_path / "a.ivecs"), iv.astype(np.float32))
    with open(tmp_path / "big.ivecs", "wb") as f:
        f.write(struct.pack("<ii", 1, 2 ** 24 + 1))
    with pytest.raises(ValueError, match="2\\^24"):
        vecio.read_ivecs(tmp_path / "big.ivecs")


@pytest.mark.parametrize("gz", [False, True])
def test_idx_images(tmp_path, gz):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (11, 28, 28)).astype(np.uint8)
    blob = struct.pack(">BBBB", 0, 0, 0x08, 3) + struct.pack(">III", 11, 28, 28) + img.tobytes()
    p = tmp_path / ("train-images-idx3-ubyte" + (".gz" if gz else ""))
    p.write_bytes(gzip.compress(blob) if gz else blob)
    x = vecio.read_vectors(p)
    assert x.shape == (11, 784) and x.dtype == np.float32
    assert np.array_equal(x, img.reshape(11, 784).astype(np.float32))
    assert np.array_equal(vecio.read_idx(p, rows=[3, 1]), x[[3, 1]])


def test_idx_types_and_errors(tmp_path):
    v = np.array([[1.5, -2.25], [3.0, 0.5]], dtype=">f4")
    p = tmp_path / "f.idx"
    p.write_bytes(struct.pack(">BBBB", 0, 0, 0x0D, 2) + struct.pack(">II", 2, 2) + v.tobytes())
    assert np.array_equal(vecio.read_idx(p), v.astype(np.float32))
    q = tmp_path / "d.idx"
    q.write_bytes(struct.pack(">BBBB", 0, 0, 0x0E, 2) + struct.pack(">II", 1, 1) + np.array([0.1], ">f8").tobytes())
    with pytest.raises(ValueError, match="binary32"):
        vecio.read_idx(q)
    r = tmp_path / "short.idx"
    r.write_bytes(struct.pack(">BBBB", 0, 0, 0x08, 2) + struct.pack(">II", 2, 3) + b"\x00" * 5)
    with pytest.raises(ValueError, match="payload"):
        vecio.read_idx(r)
    s = tmp_path / "bad.idx"
    s.write_bytes(b"\x01\x00\x08\x01" + struct.pack(">I", 1) + b"\x00")
    with pytest.raises(ValueError, match="not an IDX"):
        vecio.read_idx(s)


def test_sha256_u64_is_little_endian_bytes():
    ids = np.array([0, 1, 2 ** 40 + 3], dtype=np.uint64)
    raw = b"".join(int(v).to_bytes(8, "little") for v in ids)
    assert vecio.sha256_u64(ids) == hashlib.sha256(raw).hexdigest()
    # empty array: hash of no bytes
    assert vecio.sha256_u64(np.zeros(0, np.uint64)) == hashlib.sha256(b"").hexdigest()


def _brute_canonical(oracle, x, T):
    """Directed ids by a plain double loop over the oracle's scalar predicate."""
    n = x.shape[0]
    ids = [i * n + j for i in range(n) for j in range(n) if i == j or oracle.d2_seq_py(x[i], x[j]) <= T]
    return np.array(sorted(ids), dtype=np.uint64)


def test_validate_canonical_pass_and_planted_defects(oracle):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((40, 5)).astype(np.float32)
    T = 4.0
    ids = oracle.canonical_u64(oracle.join(x, None, T), 40)
    brute = _brute_canonical(oracle, x, T)
    assert np.array_equal(ids, brute)
    rep = vecio.validate_canonical(ids, 40, expected_count=brute.size, expected_hash=vecio.sha256_u64(brute))
    assert rep["exact_contract_pass"] and rep["self_pairs"] == 40 and rep["symmetry_missing_pairs"] == 0

    def fails(a, key, value):
        r = vecio.validate_canonical(a, 40)
        assert r[key] == value, (key, r)
        assert not r["exact_contract_pass"]

    k = int(np.flatnonzero((ids // 40) != (ids % 40))[0])  # an off-diagonal id
    fails(np.delete(ids, k), "symmetry_missing_pairs", 1)
    fails(np.insert(ids, k, ids[k]), "duplicate_pairs", 1)
    fails(ids[::-1].copy(), "sorted_output", False)
    fails(np.append(ids, np.uint64(40 * 40)), "invalid_pairs", 1)
    diag = int(np.flatnonzero((ids // 40) == (ids % 40))[0])
    fails(np.delete(ids, diag), "self_pairs", 39)
    r = vecio.validate_canonical(ids, 40, expected_count=ids.size + 1)
    assert not r["exact_contract_pass"]
    r = vecio.validate_canonical(ids, 40, expected_hash="0" * 64)
    assert not r["exact_contract_pass"]


def test_published_thresholds_are_exact_squares():
    for name, c in vecio.PUBLISHED.items():
        assert c["eps"] * c["eps"] == c["T"], name           # T = eps^2 exactly (dyadic eps)
        assert np.float32(c["T"]) == c["T"], name            # representable in binary32
        assert c["count"] >= vecio.PUBLISHED_ROWS and (c["count"] - vecio.PUBLISHED_ROWS) % 2 == 0
    assert vecio.expected("original")["expected_count"] == 3926078


def test_gdsjoin_input_checks():
    """GDS-Join's flat DATASET layout and DTYPE rules (no GPU needed)."""
    import numpy as np
    from tensorjoin_variants_amd.gds import _points
    x = _points(np.arange(12, dtype=np.float64), 3, "double")
    assert x.shape == (4, 3) and x.dtype == np.float32
    with pytest.raises(ValueError):
        _points(np.arange(10), 3, "float")           # not a multiple of NDIM
    with pytest.raises(ValueError):
        _points(np.full(3, 0.1), 3, "double")        # not exact in float32
    with pytest.raises(ValueError):
        _points(np.zeros(3), 3, "half")
