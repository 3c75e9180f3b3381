"""ctypes binding of libtjoin.so (the C ABI declared in include/tjoin.h).

The library is required: there is no CPU fallback.  Importing the package
works without a GPU (the symbols can be inspected), but every join call goes
through the HIP library and raises if it is missing or fails.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libtjoin.so"
HEADER = PKG.parent / "include" / "tjoin.h"

TJ_OK = 0
ERRORS = {-1: "TJ_ERR_INVALID", -2: "TJ_ERR_HIP", -3: "TJ_ERR_UNSUPPORTED", -4: "TJ_ERR_WORKSPACE"}

PREC = {"none": 0, "fp16": 1, "fp16x3": 2, "bf16": 3, "bf16x3": 4, "fp8": 5, "fp8x3": 6}
METHOD = {"certified": 0, "plain_fp64": 1, "grid": 2, "plain_fp32": 3}
OUT = {"int32": 0, "int64": 1}
OUT_U64_CANONICAL = 2
NOALLOC_SCRATCH_BYTES = 64 << 20
FLAG_AUDIT = 1
FLAG_TIMING = 2
FLAG_PRUNE = 4

STATS_SLOTS = 24
STAT_NAMES = {
    0: "out", 1: "ambiguous", 2: "cert_accept", 3: "fp64_accept", 4: "candidates",
    5: "max_err", 6: "max_ratio", 7: "violations", 8: "work", 9: "invalid",
    10: "grid_candidates", 11: "pairs", 12: "scale_exp", 13: "viol_key", 14: "viol_d2", 15: "viol_est",
    16: "fp32_accept", 17: "fp32_reject", 18: "fp64_refined", 19: "pruned", 20: "pruned_work",
}
DOUBLE_SLOTS = {5, 6, 14}


class TJError(RuntimeError):
    pass


_lib = None


def load():
    """Load libtjoin.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise TJError(f"{LIB_PATH} is missing: run tensorjoin_variants_amd.build.build() "
                      "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    _init_torch_device()  # before the library (and its HIP runtime) is loaded
    L = C.CDLL(str(LIB_PATH))
    vp, i64, i32, f64, sz = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_size_t
    sig = {
        "tj_version": (i32, []),
        "tj_last_error_string": (C.c_char_p, []),
        "tj_device_info": (i32, [C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "tj_plan_create": (i32, [C.POINTER(vp), i64, i64, i32, i32, i32, i64, i64, C.c_uint]),
        "tj_plan_destroy": (i32, [vp]),
        "tj_plan_set_tile": (i32, [vp, i32, i64]),
        "tj_plan_workspace_bytes": (i32, [vp, C.POINTER(sz)]),
        "tj_plan_describe": (i32, [vp, C.c_char_p, sz]),
        "tj_plan_stage_times": (i32, [vp, C.POINTER(C.c_float), i32]),
        "tj_join": (i32, [vp, vp, vp, f64, vp, vp, i32, vp, vp]),
        "tj_scale_scan": (i32, [vp, i64, i32, vp, vp, vp]),
        "tj_params": (i32, [vp, f64, i32, i32, vp, vp, vp]),
        "tj_rowmeta": (i32, [vp, i64, i32, i32, vp, vp, vp, vp, vp, vp]),
        "tj_refine_fp64": (i32, [vp, vp, i32, i32, f64, vp, vp, i64, vp, i64, vp, vp]),
        "tj_rowmeta_sizes": (i32, [i64, i32, i32, C.POINTER(sz), C.POINTER(sz)]),
        "tj_screen": (i32, [i32, i32, vp, vp, vp, i64, vp, vp, vp, i64, vp, vp, i64, vp, vp, i64, C.c_uint, vp, vp]),
        "tj_certify": (i32, [vp, vp, i32, i64, vp, vp, i64, vp, vp, vp, vp, i64, C.c_uint, vp, vp]),
        "tj_plain_fp64": (i32, [vp, i64, vp, i64, i32, i32, f64, vp, i64, vp, vp]),
        "tj_sort_keys": (i32, [vp, C.POINTER(sz), vp, vp, i64, i32, vp]),
        "tj_grid_workspace_bytes": (i32, [i64, i64, i64, C.POINTER(sz)]),
        "tj_grid_join": (i32, [vp, i64, vp, i64, i32, f64, vp, sz, i64, vp, i32, vp, vp]),
        "tj_grid_neighbors": (i32, [vp, i64, i32, f64, vp, sz, i64, vp, vp, vp, vp]),
        "tj_unpack_pairs": (i32, [vp, vp, i64, i32, vp, i32, vp]),
        "tj_mfma_probe": (i32, [i32, vp, vp, vp, vp, vp, i32, i32, vp, vp]),
        "tj_canonical_u64": (i32, [vp, C.POINTER(sz), vp, vp, i64, i64, vp, vp]),
        "tj_join_noalloc": (i32, [vp, i64, vp, i64, i32, i32, i32, f64, i64, i64, C.c_uint, vp, i32, vp, vp]),
        "tj_plan_capture": (i32, [vp, vp, vp, f64, vp, vp, i32, vp]),
        "tj_plan_replay": (i32, [vp, vp]),
        "tj_batch_create": (i32, [C.POINTER(vp), i32, C.POINTER(i64), C.POINTER(i32), C.POINTER(i32), i64, i64,
                                  C.c_uint]),
        "tj_batch_destroy": (i32, [vp]),
        "tj_batch_workspace_bytes": (i32, [vp, C.POINTER(sz)]),
        "tj_batch_describe": (i32, [vp, C.c_char_p, sz]),
        "tj_batch_stage_times": (i32, [vp, C.POINTER(C.c_float), i32]),
        "tj_batch_join": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "tj_batch_capture": (i32, [vp, vp, vp, vp, vp, vp, vp]),
        "tj_batch_replay": (i32, [vp, vp]),
        "tj_host_create": (i32, [C.POINTER(vp), i64, i64, i32, i32, i32, C.c_uint, i64, i32]),
        "tj_host_destroy": (i32, [vp]),
        "tj_host_join": (i32, [vp, vp, vp, f64, i32, C.POINTER(vp), C.POINTER(i64), vp]),
        "tj_host_free": (None, [vp]),
        "tj_host_times": (i32, [vp, C.POINTER(C.c_double), i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _init_torch_device() -> None:
    """Initialise PyTorch's GPU runtime before libtjoin.so is loaded.

    The device buffers come from PyTorch, whose wheel carries its own HIP and
    HSA runtime beside the system ROCm one libtjoin.so links.  Measured on the
    box: with PyTorch's runtime initialised first, both work in one process;
    with libtjoin.so loaded first, whichever runtime opens the device second
    finds no GPU ("No HIP GPUs are available" from PyTorch after a batched
    join as a process's first call; "no ROCm-capable device" from the
    library's hipMalloc when PyTorch initialised after the load).  Without a
    GPU (the CPU tests only inspect the symbols) this is a no-op."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def check(rc: int, what: str = "") -> None:
    if rc != TJ_OK:
        msg = load().tj_last_error_string().decode(errors="replace")
        raise TJError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")


def declared_symbols() -> list[str]:
    """Entry points declared in include/tjoin.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(tj_[a-z0-9_]+)\s*\(", text, re.M)))


def decode_stats(raw) -> dict:
    """uint64 stats block -> dict (binary64 slots decoded)."""
    import numpy as np
    raw = np.asarray(raw, dtype=np.uint64)
    out = {}
    for slot, name in STAT_NAMES.items():
        v = raw[slot]
        if slot in DOUBLE_SLOTS:
            out[name] = float(np.array([v], dtype=np.uint64).view(np.float64)[0])
        elif slot == 12:
            out[name] = int(np.array([v], dtype=np.uint64).view(np.int64)[0])
        elif slot == 15:
            out[name] = [float(f) for f in np.array([v], dtype=np.uint64).view(np.float32)]
        else:
            out[name] = int(v)
    return out
