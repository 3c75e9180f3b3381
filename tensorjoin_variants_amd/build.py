"""Build libtjoin.so (gfx950) in-tree with hipcc.

Every HIP source under csrc/ is compiled to an object in csrc/_build/ (in
parallel) and linked into tensorjoin_variants_amd/libtjoin.so.  Flags:
  -ffp-contract=off   binary64 refinement/brute force must not fuse a*b+c
  -packed-fp32-ops    off: no v_pk_{add,mul,fma}_f32 from the compiler
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = CSRC / "_build"
LIB = PKG / "libtjoin.so"
ARCH = os.environ.get("TJ_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]


def _compile(src: Path) -> tuple[Path, str]:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "'-packed-fp32-ops' is not a recognized feature for this target"
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{msg}")
    return obj, msg


def build(force: bool = True, jobs: int | None = None, verbose: bool = False) -> Path:
    """Compile every csrc/*.hip and link libtjoin.so.  Returns the library path."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    headers = list(CSRC.glob("*.h")) + [PKG.parent / "include" / "tjoin.h"]
    newest_header = max(h.stat().st_mtime for h in headers)
    todo = []
    for s in sources:
        obj = BUILD / (s.stem + ".o")
        if force or not obj.exists() or obj.stat().st_mtime < max(s.stat().st_mtime, newest_header):
            todo.append(s)
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 4, 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, msg in ex.map(_compile, todo):
            if verbose and msg.strip():
                print(msg, file=sys.stderr)
    objs = [str(BUILD / (s.stem + ".o")) for s in sources]
    tmp = LIB.with_suffix(".so.tmp")
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose=True))
