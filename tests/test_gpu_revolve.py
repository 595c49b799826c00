"""Binomial (revolve) checkpointing of the adjoint's reverse sweep against the segment schedule.

The reverse sweep applies the same stage adjoints at the same recomputed stage states whatever the
schedule, so the gradient of a taped run with the revolve schedule equals the segment schedule's
bit for bit, for every number of snapshot buffers (1, 2, 3, and one per step).  The forward steps
the sweep recomputes are Griewank's minimum T(n, s) (computed here by an independent dynamic
programme), plus one recomputation of every step's stages, less the snapshots the taped run took on
its way (the first descent of the schedule).
"""
import functools

import numpy as np
import pytest

from prandtl_amd import capi, cases
from prandtl_amd import torch_adjoint as TA

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@functools.lru_cache(maxsize=None)
def _T(n, s):
    """Fewest advance steps to reverse n steps from their start with s snapshot buffers, and the split."""
    if n <= 1:
        return 0, 1
    if s == 0:
        return n * (n - 1) // 2, 1
    best = None
    for k in range(1, n):
        c = k + _T(n - k, s - 1)[0] + _T(k, s)[0]
        if best is None or c < best[0]:
            best = (c, k)
    return best


def _expected_recompute(n, s):
    a, sl, chain = 0, s, 0
    while n - a > 1 and sl > 0:
        a += _T(n - a, sl)[1]
        chain = a
        sl -= 1
    return n + _T(n, s)[0] - chain


@pytest.mark.parametrize("scheme", [capi.LSRK45, capi.SSPRK3])
def test_revolve_matches_segments(device, scheme):
    import torch

    nsteps = 12
    _say(f"revolve vs segments, TGV 3^3 p=4 Euler + 8 modes, {nsteps} steps, scheme {scheme}")
    m = cases.tgv_mesh(n=3, order=4)
    modes = cases.fourier_modes(count=8, kmax=2, seed=11)
    with torch.no_grad():
        u0 = TA.tgv_modes_state(m, torch.tensor(0.3, dtype=torch.float64, device=device),
                                torch.full((8,), 1e-2, dtype=torch.float64, device=device), modes)
    rng = np.random.default_rng(2)
    ubT = torch.from_numpy(rng.normal(size=tuple(u0.shape))).to(device)
    op = capi.Operator(m)
    dt = 2e-3
    x = u0.clone()
    op.rk_run_tape_schedule(x, 0.0, dt, nsteps, capi.TAPE_SEGMENTS, 5, scheme=scheme)
    ref_state = x.clone()
    g_seg = ubT.clone()
    op.rk_adjoint(g_seg)
    states, rec = op.tape_info()
    per = 5 if scheme == capi.LSRK45 else 3
    print(f"[revolve] segments of 5: {states} states, {rec} recomputed steps", flush=True)
    assert rec == nsteps and states == 3 + 5 * per - 1
    for s in (1, 2, 3, nsteps):
        x = u0.clone()
        op.rk_run_tape_schedule(x, 0.0, dt, nsteps, capi.TAPE_REVOLVE, s, scheme=scheme)
        assert torch.equal(x, ref_state), "the taped run must not depend on the schedule"
        g = ubT.clone()
        op.rk_adjoint(g)
        states, rec = op.tape_info()
        exp = _expected_recompute(nsteps, s)
        same = torch.equal(g, g_seg)
        print(f"[revolve] {s} snapshots: {states} states, {rec} recomputed steps (expected {exp}), "
              f"identical gradient: {same}", flush=True)
        assert same
        assert rec == exp and states == s + 1 + per
    op.close()
