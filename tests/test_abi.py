"""C ABI of libtjoin.so without a GPU: the library loads, exports every entry
point declared in include/tjoin.h, plans lay out workspaces, and bad
arguments return error codes instead of aborting."""
import ctypes as C
import json

import pytest

from tensorjoin_variants_amd import _lib


@pytest.fixture(scope="module")
def L():
    return _lib.load()


def test_exports_every_declared_symbol(L):
    declared = _lib.declared_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
    assert L.tj_version() >= 10000


def _plan(L, nR, nS, d, method, prec, out=1024, amb=1024, flags=0):
    p = C.c_void_p()
    rc = L.tj_plan_create(C.byref(p), nR, nS, d, method, prec, out, amb, flags)
    return rc, p


@pytest.mark.parametrize("prec,d,dp", [("fp16", 3, 16), ("fp16", 128, 128), ("fp16", 130, 256), ("bf16", 32, 32),
                                       ("fp8", 3, 64), ("fp8", 512, 512), ("fp16x3", 256, 256), ("fp8x3", 100, 128)])
def test_plan_layout(L, prec, d, dp):
    rc, p = _plan(L, 1000, 777, d, 0, _lib.PREC[prec])
    assert rc == 0, L.tj_last_error_string()
    nbytes = C.c_size_t()
    assert L.tj_plan_workspace_bytes(p, C.byref(nbytes)) == 0
    buf = C.create_string_buffer(1024)
    assert L.tj_plan_describe(p, buf, 1024) == 0
    desc = json.loads(buf.value.decode())
    assert desc["dp"] == dp and desc["nR"] == 1000 and desc["nS"] == 777 and desc["self"] == 0
    assert desc["jbits"] == 10 and desc["keybits"] == 20
    assert desc["panels"] * desc["rows_per_wg"] >= 1000
    assert desc["col_splits"] * desc["cols_per_split"] >= 777
    es = 1 if prec.startswith("fp8") else 2
    nt = 2 if prec.endswith("x3") else 1
    assert nbytes.value >= (1000 + 777) * (dp * es * nt + 32)
    assert L.tj_plan_destroy(p) == 0


def test_self_join_plan(L):
    rc, p = _plan(L, 8192, 0, 32, 0, _lib.PREC["fp16"])
    assert rc == 0
    buf = C.create_string_buffer(1024)
    L.tj_plan_describe(p, buf, 1024)
    desc = json.loads(buf.value.decode())
    assert desc["self"] == 1 and desc["nS"] == 8192
    # the triangle's columns are split for balance: whole 32-column blocks,
    # at least 1024 columns a split, covering every column
    cps, cs = desc["cols_per_split"], desc["col_splits"]
    assert cps % 32 == 0 and cps >= 1024 and cs * cps >= 8192 and (cs - 1) * cps < 8192
    L.tj_plan_destroy(p)


def _desc(L, p):
    buf = C.create_string_buffer(1024)
    assert L.tj_plan_describe(p, buf, 1024) == 0
    return json.loads(buf.value.decode())


def test_plan_set_tile(L):
    rc, p = _plan(L, 100000, 262144, 128, 0, _lib.PREC["fp16"])
    assert rc == 0
    d0 = _desc(L, p)
    assert d0["row_blocks"] == 4 and d0["rows_per_wg"] == 512
    nbytes0 = C.c_size_t()
    L.tj_plan_workspace_bytes(p, C.byref(nbytes0))
    # two row blocks per wave, 5000-column splits (rounded up to 5024)
    assert L.tj_plan_set_tile(p, 2, 5000) == 0
    d1 = _desc(L, p)
    assert d1["row_blocks"] == 2 and d1["rows_per_wg"] == 256 and d1["panels"] == (100000 + 255) // 256
    assert d1["cols_per_split"] == 5024 and d1["col_splits"] == (262144 + 5023) // 5024
    nbytes1 = C.c_size_t()
    L.tj_plan_workspace_bytes(p, C.byref(nbytes1))
    assert nbytes1.value == nbytes0.value  # the workspace does not depend on the tile
    # a split wider than S is all of S; back to the defaults
    assert L.tj_plan_set_tile(p, 0, 1 << 30) == 0
    assert _desc(L, p)["col_splits"] == 1
    assert L.tj_plan_set_tile(p, 0, 0) == 0
    assert _desc(L, p) == d0
    # blockings that are not built, negative values, non-certified plans
    assert L.tj_plan_set_tile(p, 1, 0) == -3
    assert L.tj_plan_set_tile(p, 3, 0) == -3
    assert L.tj_plan_set_tile(p, -1, 0) == -1
    assert L.tj_plan_set_tile(p, 0, -5) == -1
    assert L.tj_plan_set_tile(None, 0, 0) == -1
    assert _desc(L, p) == d0  # a rejected tile leaves the plan as it was
    L.tj_plan_destroy(p)
    rc, q = _plan(L, 1000, 1000, 8, 1, 0)  # plain FP64
    assert rc == 0 and L.tj_plan_set_tile(q, 0, 0) == -1
    L.tj_plan_destroy(q)
    rc, q = _plan(L, 1000, 1000, 128, 0, _lib.PREC["fp16x3"])  # a one-block body has no alternative
    assert rc == 0 and _desc(L, q)["row_blocks"] == 1
    assert L.tj_plan_set_tile(q, 1, 0) == 0 and L.tj_plan_set_tile(q, 2, 0) == -3
    L.tj_plan_destroy(q)


@pytest.mark.parametrize("args,code", [
    ((10, 10, 0, 0, 1), -1),        # d = 0
    ((10, 10, 8, 7, 1), -1),        # unknown method
    ((10, 10, 8, 0, 9), -1),        # unknown precision
    ((10, 10, 600, 0, 1), -3),      # d beyond the widest FP16 screen
    ((10, 10, 300, 0, 2), -3),      # fp16 hi/lo split beyond d = 256
    ((-1, 10, 8, 0, 1), -1),        # negative size
])
def test_plan_rejects_bad_arguments(L, args, code):
    rc, p = _plan(L, *args)
    assert rc == code
    assert len(L.tj_last_error_string()) > 0


def test_capacity_bounds(L):
    rc, _ = _plan(L, 10, 10, 8, 0, 1, out=1 << 33)
    assert rc == -1


def test_sort_temp_query(L):
    need = C.c_size_t()
    assert L.tj_sort_keys(None, C.byref(need), None, None, 1 << 20, 40, None) == 0
    assert need.value >= 8 << 20
    small = C.c_size_t(16)
    assert L.tj_sort_keys(C.c_void_p(1), C.byref(small), None, None, 1 << 20, 40, None) == -4


def test_join_rejects_bad_threshold(L):
    rc, p = _plan(L, 10, 10, 8, 0, 1)
    assert rc == 0
    ws = C.c_void_p(1)
    for T in (float("nan"), -1.0, float("inf")):
        assert L.tj_join(p, C.c_void_p(1), C.c_void_p(1), T, ws, C.c_void_p(1), 0, C.c_void_p(1), None) == -1
    L.tj_plan_destroy(p)


def test_stats_decoding():
    import numpy as np
    raw = np.zeros(24, np.uint64)
    raw[0] = 5
    raw[5] = np.array([0.25]).view(np.uint64)[0]
    raw[12] = np.array([-3]).view(np.uint64)[0]
    s = _lib.decode_stats(raw)
    assert s["out"] == 5 and s["max_err"] == 0.25 and s["scale_exp"] == -3


def _batch(L, n, d, prec, out=1024, cand=1024):
    J = len(n)
    b = C.c_void_p()
    rc = L.tj_batch_create(C.byref(b), J, (C.c_int64 * J)(*n), (C.c_int * J)(*d), (C.c_int * J)(*prec),
                           out, cand, 0)
    return rc, b


@pytest.mark.parametrize("n,d,prec,code", [
    ([], [], [], -1),                       # no joins
    ([10, 20], [16, 0], [1, 1], -1),        # zero width
    ([10, -1], [16, 16], [1, 1], -1),       # negative size
    ([10, 20], [16, 300], [1, 1], -3),      # beyond the batch screen's widest body
    ([10, 20], [16, 64], [1, 2], -3),       # hi/lo split not served by the batch screen
    ([10, 20], [16, 64], [1, 3], -3),       # bf16 not served by the batch screen
])
def test_batch_rejects_bad_arguments(L, n, d, prec, code):
    if not n:
        b = C.c_void_p()
        rc = L.tj_batch_create(C.byref(b), 0, None, None, None, 16, 16, 0)
    else:
        rc, b = _batch(L, n, d, prec)
    assert rc == code
    assert len(L.tj_last_error_string()) > 0


def test_batch_capacity_bounds(L):
    rc, _ = _batch(L, [10], [16], [1], out=1 << 33)
    assert rc == -1


@pytest.mark.parametrize("nR,nS,d,method,prec,chunk,code", [
    (-1, 10, 4, 0, 1, 0, -1),      # negative rows
    (10, 10, 0, 0, 1, 0, -1),      # d = 0
    (10, 10, 4, 9, 1, 0, -1),      # unknown method
    (10, 10, 4, 0, 42, 0, -1),     # unknown precision
    (10, 10, 8, 2, 0, 0, -3),      # grid beyond d = 3
    (10, 10, 4, 0, 1, -5, -1),     # negative chunk rows
    (1 << 31, 10, 4, 0, 1, 0, -3),  # beyond 2^31 rows
])
def test_host_rejects_bad_arguments(L, nR, nS, d, method, prec, chunk, code):
    """tj_host_create validates before it touches the device."""
    h = C.c_void_p()
    assert L.tj_host_create(C.byref(h), nR, nS, d, method, prec, 0, chunk, 0) == code
    assert not h.value


def test_host_join_rejects_null_handle(L):
    res, cnt = C.c_void_p(), C.c_int64()
    assert L.tj_host_join(None, None, None, 1.0, 0, C.byref(res), C.byref(cnt), None) == -1
    assert L.tj_host_destroy(None) == 0
    L.tj_host_free(None)


@pytest.mark.parametrize("n,d,prec,q,m", [(1000, 128, "fp16", 1000 * 128 * 2, 48000), (10, 3, "fp8", 10 * 64, 480),
                                          (7, 100, "bf16x3", 7 * 128 * 2, 336), (0, 16, "fp16", 0, 0)])
def test_rowmeta_sizes(L, n, d, prec, q, m):
    qb, mb = C.c_size_t(), C.c_size_t()
    assert L.tj_rowmeta_sizes(n, d, _lib.PREC[prec], C.byref(qb), C.byref(mb)) == 0
    assert (qb.value, mb.value) == (q, m)
    assert L.tj_rowmeta_sizes(n, 600, _lib.PREC["fp16"], C.byref(qb), C.byref(mb)) == -3
    assert L.tj_rowmeta_sizes(n, d, 0, C.byref(qb), C.byref(mb)) == -1


def test_stage_calls_reject_bad_arguments(L):
    """Argument checks run before anything touches the device."""
    one = C.c_void_p(1)
    # tj_screen: unknown precision, NULL params, capacity beyond 2^32
    assert L.tj_screen(9, 8, one, None, one, 10, one, None, one, 10, one, one, 10, one, one, 10, 0, one, None) == -1
    assert L.tj_screen(1, 8, one, None, one, 10, one, None, one, 10, None, one, 10, one, one, 10, 0, one, None) == -1
    assert L.tj_screen(1, 8, one, None, one, 10, one, None, one, 10, one, one, 1 << 33, one, one, 10, 0, one,
                       None) == -1
    # hi/lo precision without lo codes
    assert L.tj_screen(2, 8, one, None, one, 10, one, None, one, 10, one, one, 10, one, one, 10, 0, one, None) == -1
    # tj_certify: NULL stats, R x S without meta_S
    assert L.tj_certify(one, one, 8, 10, one, one, 10, one, one, one, one, 10, 0, None, None) == -1
    assert L.tj_certify(one, one, 8, 10, one, one, 10, one, None, one, one, 10, 0, one, None) == -1
    # tj_refine_fp64: NULL count, negative T, jbits out of range
    assert L.tj_refine_fp64(one, None, 8, 6, 1.0, one, None, 10, one, 10, one, None) == -1
    assert L.tj_refine_fp64(one, None, 8, 6, -1.0, one, one, 10, one, 10, one, None) == -1
    assert L.tj_refine_fp64(one, None, 8, 40, 1.0, one, one, 10, one, 10, one, None) == -1
