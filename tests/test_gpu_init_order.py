"""First calls in a fresh process.

The device buffers come from PyTorch, whose wheel carries its own HIP/HSA
runtime beside the system ROCm one the library links.  A process whose first
GPU call is a library call (a batched join, which creates its plan through
the library before any tensor exists) must leave PyTorch able to use the
GPU, and get the right pairs.  Each case runs in one child process (the
parent has long initialised PyTorch, so it cannot show the order).
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import sys
sys.path.insert(0, {root!r})
import numpy as np
import tensorjoin_variants_amd as tj
from oracle import oracle as O
rng = np.random.default_rng(3)
X = rng.random((300, 8)).astype(np.float32)
Y = rng.random((200, 3)).astype(np.float32)
got = tj.batched_join([X, Y], T=[0.05, 0.01], precision="fp16")
assert np.array_equal(got[0].astype(np.int64), O.join(X, None, 0.05))
assert np.array_equal(got[1].astype(np.int64), O.join(Y, None, 0.01))
import torch
z = torch.ones(8, device="cuda") * 2
torch.cuda.synchronize()
assert float(z.sum()) == 16.0
print("first-call batch ok", len(got[0]), len(got[1]))
"""


def test_batched_join_as_first_call():
    print("[case] fresh process: batched_join before any torch call", file=sys.stderr, flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=str(ROOT))], cwd=str(ROOT), capture_output=True,
                       text=True, timeout=240, env=dict(os.environ))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "first-call batch ok" in r.stdout
