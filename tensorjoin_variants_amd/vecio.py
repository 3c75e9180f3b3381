"""Vector files and the reference's output contract.

Readers for the vector formats the reference's data preparation reads --
``.fvecs`` (CEJ/src/prepare_g2_cifar.py:53-69, memory-mapped row subsets as
in CEJ/src/prepare_g4b_public_data.py:88-97), ``.bvecs``/``.ivecs`` (the same
record layout with uint8 / int32 payloads) and IDX image files, gzip or raw
(CEJ/src/prepare_g4b_public_data.py:100-115) -- and the checks of its
canonical output: the SHA-256 of the little-endian u64 id array
(CEJ/src/run_g3a_analytic_certificate.py:45-46) and ``validate_canonical``
(CEJ/src/g2b_public_common.py:106-143).

``PUBLISHED`` lists the reference's published CIFAR-60K (GIST-512) results
(threshold, directed pair count incl. self pairs, SHA-256 of the sorted ids;
tensorjoin_20260908_output_audit/inherited/precision_routing_20260908_threshold_sweep/
artifacts/reference_manifest.json).  The dataset itself is not part of the
reference archive; with the file at hand::

    x = read_fvecs("cifar60k_base.fvecs")
    ids, _ = join_reference_compat(x, T=PUBLISHED["original"]["T"])
    validate_canonical(ids, x.shape[0], **expected("original"))["exact_contract_pass"]
"""
from __future__ import annotations

import gzip
import hashlib
import struct
from pathlib import Path

import numpy as np

# name -> eps, T = eps^2 (exact), directed count incl. the n self pairs, sha256
PUBLISHED = {
    "k4": {"eps": 0.5107421875, "T": 0.26085758209228516, "count": 297318,
           "sha256": "0b1503dbe8814b114412ba0d04f6d10d958ae15ab7817d3a8dab4413f8f7579a"},
    "k16": {"eps": 0.56884765625, "T": 0.32358765602111816, "count": 1016858,
            "sha256": "b65ded9f0b0892bf0c056a679273dea504eaa5f2bee6f1b960c8de2d1248aa85"},
    "k64": {"eps": 0.628173828125, "T": 0.39460235834121704, "count": 3862998,
            "sha256": "4863a23e3edeb18d5e72d89fb0e25c0dbaea17762bbc98a7159803be72407a80"},
    "original": {"eps": 0.62890625, "T": 0.3955230712890625, "count": 3926078,
                 "sha256": "13cae87e6a0f75a13f3e9a7881b33c42ab77680f6d13e99dcc00fa2de5963495"},
    "k256": {"eps": 0.69189453125, "T": 0.4787180423736572, "count": 15429428,
             "sha256": "49160b0fece513842a1194c0aaca5b2d06de6de417533c401ac0c383b2e70611"},
    "k1024": {"eps": 0.76171875, "T": 0.5802154541015625, "count": 61458806,
              "sha256": "1aa5d88ffa1f87e75ad08af781637cb073d2c862378f6b4dcf84f2a56c3d6fe3"},
}
PUBLISHED_ROWS, PUBLISHED_DIM = 60_000, 512


def expected(name: str) -> dict:
    """Keyword arguments of validate_canonical for a published result."""
    c = PUBLISHED[name]
    return {"expected_count": c["count"], "expected_hash": c["sha256"]}


# ------------------------------------------------------------------ readers
_VEC_DTYPES = {".fvecs": "<f4", ".ivecs": "<i4", ".bvecs": "u1"}


def _read_vecs(path, payload: str, rows=None, dim: int | None = None, n_rows: int | None = None) -> np.ndarray:
    path = Path(path)
    size = path.stat().st_size
    if size == 0:
        return np.zeros((0, dim or 0), np.float32)
    d = int(np.fromfile(path, dtype="<i4", count=1)[0])
    if d <= 0:
        raise ValueError(f"{path}: bad dimension header {d}")
    if dim is not None and d != dim:
        raise ValueError(f"{path}: dimension {d}, expected {dim}")
    esize = np.dtype(payload).itemsize
    rec = 4 + d * esize
    if size % rec:
        raise ValueError(f"{path}: {size} bytes is not a whole number of {rec}-byte records")
    n = size // rec
    if n_rows is not None and n != n_rows:
        raise ValueError(f"{path}: {n} rows, expected {n_rows}")
    raw = np.memmap(path, dtype="u1", mode="r", shape=(n, rec))
    heads = raw[:, :4].copy().view("<i4").ravel()
    if not np.all(heads == d):
        bad = np.flatnonzero(heads != d)[:8]
        raise ValueError(f"{path}: inconsistent dimension headers at rows {bad.tolist()}")
    body = raw[:, 4:] if rows is None else raw[np.asarray(rows, dtype=np.int64), 4:]
    out = np.ascontiguousarray(body).view(payload).reshape(-1, d)
    if payload == "<i4" and out.size and np.abs(out.astype(np.int64)).max() > 2 ** 24:
        raise ValueError(f"{path}: int32 values beyond 2^24 are not exact in binary32")
    x = out.astype(np.float32, copy=payload != "<f4")
    if payload == "<f4" and not np.isfinite(x).all():
        raise ValueError(f"{path}: non-finite coordinates")
    return np.ascontiguousarray(x)


def read_fvecs(path, rows=None, dim: int | None = None, n_rows: int | None = None) -> np.ndarray:
    """float32 [n, d] from an .fvecs file (each record: int32 d, then d float32).
    ``rows`` selects a subset (memory-mapped; the file is never loaded whole)."""
    return _read_vecs(path, "<f4", rows, dim, n_rows)


def read_bvecs(path, rows=None, dim: int | None = None, n_rows: int | None = None) -> np.ndarray:
    """float32 [n, d] from a .bvecs file (int32 d, then d uint8)."""
    return _read_vecs(path, "u1", rows, dim, n_rows)


def read_ivecs(path, rows=None, dim: int | None = None, n_rows: int | None = None) -> np.ndarray:
    """float32 [n, d] from an .ivecs file (int32 d, then d int32; values must be
    exactly representable in binary32, |v| <= 2^24)."""
    return _read_vecs(path, "<i4", rows, dim, n_rows)


def write_fvecs(path, x) -> None:
    """Write float32 [n, d] as .fvecs."""
    x = np.ascontiguousarray(x, dtype="<f4")
    n, d = x.shape
    rec = np.empty((n, d + 1), dtype="<f4")
    rec[:, 0] = np.array([d], dtype="<i4").view("<f4")[0]
    rec[:, 1:] = x
    rec.tofile(path)


# IDX (big-endian): magic = 0x0000, type code, number of dimensions; then
# int32 sizes; then the payload.
_IDX_TYPES = {0x08: ">u1", 0x09: ">i1", 0x0B: ">i2", 0x0C: ">i4", 0x0D: ">f4", 0x0E: ">f8"}


def read_idx(path, rows=None) -> np.ndarray:
    """float32 [n, prod(dims[1:])] from an IDX file (raw or .gz), e.g.
    train-images-idx3-ubyte: 60000 x 784.  Integer payloads beyond 2^24 and
    float64 payloads that do not round-trip through binary32 are rejected."""
    path = Path(path)
    opener = gzip.open if path.suffix == ".gz" else open
    with opener(path, "rb") as f:
        head = f.read(4)
        if len(head) != 4 or head[0] != 0 or head[1] != 0:
            raise ValueError(f"{path}: not an IDX file")
        code, nd = head[2], head[3]
        if code not in _IDX_TYPES or nd < 1:
            raise ValueError(f"{path}: unsupported IDX type 0x{code:02x} / {nd} dimensions")
        dims = struct.unpack(">" + "I" * nd, f.read(4 * nd))
        dt = np.dtype(_IDX_TYPES[code])
        count = int(np.prod(dims, dtype=np.int64))
        payload = f.read()
    if len(payload) != count * dt.itemsize:
        raise ValueError(f"{path}: payload {len(payload)} bytes, header says {count * dt.itemsize}")
    a = np.frombuffer(payload, dtype=dt).reshape(dims[0], -1)
    if rows is not None:
        a = a[np.asarray(rows, dtype=np.int64)]
    x = a.astype(np.float32)
    if dt.kind in "iu" and dt.itemsize >= 4 and a.size and np.abs(a.astype(np.int64)).max() > 2 ** 24:
        raise ValueError(f"{path}: integer values beyond 2^24 are not exact in binary32")
    if dt.kind == "f" and dt.itemsize == 8 and not np.array_equal(x.astype(np.float64), a.astype(np.float64)):
        raise ValueError(f"{path}: float64 values that binary32 does not hold exactly")
    return np.ascontiguousarray(x)


def read_vectors(path, rows=None) -> np.ndarray:
    """Dispatch on the file name: .fvecs/.bvecs/.ivecs, IDX (*-ubyte[.gz],
    *.idx[.gz]) or .npy (float32, loaded without pickle)."""
    p = Path(path)
    name = p.name[:-3] if p.name.endswith(".gz") else p.name
    suf = Path(name).suffix
    if suf in _VEC_DTYPES and not p.name.endswith(".gz"):
        return {".fvecs": read_fvecs, ".bvecs": read_bvecs, ".ivecs": read_ivecs}[suf](p, rows)
    if "idx" in name or name.endswith("ubyte"):
        return read_idx(p, rows)
    if suf == ".npy":
        x = np.load(p, allow_pickle=False, mmap_mode="r")
        if x.ndim != 2 or x.dtype != np.float32:
            raise ValueError(f"{p}: expected a 2-D float32 array, got {x.shape} {x.dtype}")
        return np.ascontiguousarray(x if rows is None else x[np.asarray(rows, dtype=np.int64)])
    raise ValueError(f"{p}: unknown vector file type")


# --------------------------------------------------------- output contract
def sha256_u64(ids) -> str:
    """SHA-256 of the ids as little-endian uint64 bytes."""
    return hashlib.sha256(np.ascontiguousarray(np.asarray(ids, dtype="<u8")).tobytes()).hexdigest()


def validate_canonical(ids, n: int, expected_count: int | None = None, expected_hash: str | None = None) -> dict:
    """Checks of a canonical directed id array (i*n + j, both directions and
    the n self pairs, sorted): sorted, ids < n^2, no duplicates, all n self
    pairs, every reverse pair present; with expected count/hash, the exact
    contract.  Same report keys as CEJ/src/g2b_public_common.py:106-143."""
    ids = np.ascontiguousarray(np.asarray(ids, dtype=np.uint64))
    N = np.uint64(n)
    srt = bool(ids.size < 2 or np.all(ids[1:] >= ids[:-1]))
    invalid = int(np.count_nonzero(ids >= N * N))
    dups = int(np.count_nonzero(ids[1:] == ids[:-1])) if ids.size > 1 else 0
    r = ids // N if n else ids
    c = ids - r * N if n else ids
    selfp = int(np.count_nonzero(r == c))
    rev = c * N + r
    missing = int(np.setdiff1d(rev, ids, assume_unique=srt and dups == 0).size)
    h = sha256_u64(ids)
    structural = srt and invalid == 0 and dups == 0 and selfp == n and missing == 0
    exact = structural and (expected_count is None or ids.size == int(expected_count)) and \
        (expected_hash is None or h == expected_hash)
    return {
        "canonical_pair_count": int(ids.size),
        "canonical_raw_u64_sha256": h,
        "expected_pair_count": None if expected_count is None else int(expected_count),
        "expected_raw_u64_sha256": expected_hash,
        "sorted_output": srt,
        "invalid_pairs": invalid,
        "duplicate_pairs": dups,
        "self_pairs": selfp,
        "symmetry_missing_pairs": missing,
        "exact_contract_pass": bool(exact),
    }
