"""The roctx phase ranges (PRANDTL_ROCTX=1) change nothing but the trace: a run with them enabled,
in a child process (the library reads the variable once), gives the state of the same run in
this process bit for bit.  The run covers the ranges of an RHS evaluation, a run of RK steps and
the stage hooks."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_RUN = r"""
import hashlib, sys
import numpy as np, torch
sys.path.insert(0, {root!r})
from prandtl_amd import capi, cases
mesh = cases.khi_mesh(n=8, order=3)
u0 = cases.khi_state(mesh)
op = capi.Operator(mesh, stage_hooks=capi.HOOK_FILTER | capi.HOOK_LIMITER)
u = torch.from_numpy(u0).cuda()
du = torch.empty_like(u)
op.mult(u, du)
op.rk_run(capi.SSPRK3, u, 0.0, 1e-3, 6)
torch.cuda.synchronize()
out = np.concatenate([u.cpu().numpy().ravel(), du.cpu().numpy().ravel()])
op.close()
DIGEST = hashlib.sha256(out.tobytes()).hexdigest()
if __name__ == "__main__":
    print("digest", DIGEST, flush=True)
"""


def _digest_here():
    ns = {"__name__": "roctx_run"}
    exec(compile(_RUN.format(root=ROOT), "<roctx-run>", "exec"), ns)
    return ns["DIGEST"]


def test_roctx_ranges_change_nothing(device):
    print("[gpu-case] KHI 8^2 p=3 with hooks: Mult + 6 SSP-RK3 steps, PRANDTL_ROCTX=1 in a child process",
          flush=True)
    assert os.environ.get("PRANDTL_ROCTX", "0") in ("", "0"), "run this test with the ranges off in the parent"
    env = dict(os.environ, PRANDTL_ROCTX="1")
    r = subprocess.run([sys.executable, "-c", _RUN.format(root=ROOT)], env=env, capture_output=True, text=True,
                       timeout=150)
    assert r.returncode == 0, r.stderr[-2000:]
    child = [line.split()[1] for line in r.stdout.splitlines() if line.startswith("digest ")]
    assert len(child) == 1, r.stdout[-2000:]
    assert child[0] == _digest_here()
