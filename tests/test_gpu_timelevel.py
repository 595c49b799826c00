"""Time-level checks on the GPU (the configurations' own runs, not only single RHS calls).

* Shu vortex (2D Euler, periodic [0,10]^2, p=3, Chandrashekar volume and surface flux,
  SSP-RK3): after one period (t = 10, the vortex advected by (1, 1) back to its start) the
  state equals the initial one up to the discretisation error, which falls at the rate p+1 = 4
  with the mesh (8, 16, 32 elements per side; dt from the reference's CFL formula at CFL 0.1,
  so the O(dt^3) time error stays below the spatial one).  The 16 x 16 run with CFL 0.5 and
  100 steps (dt = 5.84e-3) matches the oracle's 100 steps to 1e-10 relative.
* Kelvin-Helmholtz instability (2D Euler, [-1,1]^2, width 0.1, seeded noise 1e-2, SSP-RK3,
  entropy filter and positivity limiter): 32 x 32 p=3 for 40 steps at CFL 0.5 against the
  oracle's hooked steps, with the same numbers of filtered and limited elements; mass, momentum
  and energy conserved to rounding.
"""
import math

import numpy as np
import pytest

import oracle
from prandtl_amd import capi, cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _vortex_error(n, cfl, device):
    import torch

    m = cases.vortex_mesh(n=n, order=3)
    u0 = cases.vortex_state(m)
    lam = 3.342   # max(|V| + c) of the vortex state (the reference's lambda_max, C1)
    dt0 = cases.cfl_dt(m, lam, cfl)
    nsteps = int(math.ceil(10.0 / dt0))
    dt = 10.0 / nsteps
    op = capi.Operator(m)
    d = torch.from_numpy(u0).to(device)
    op.rk_run(capi.SSPRK3, d, 0.0, dt, nsteps, 100)
    u = d.cpu().numpy()
    W = m.quadrature_weights()
    err = math.sqrt((W * (u[0] - u0[0]) ** 2).sum() / W.sum())
    op.close()
    return err, nsteps


def test_vortex_one_period_convergence(device):
    _say("vortex one period, p=3, n = 8, 16, 32, CFL 0.1")
    errs = []
    for n in (8, 16, 32):
        e, ns = _vortex_error(n, 0.1, device)
        errs.append(e)
        print(f"[timelevel] vortex n={n}: {ns} steps, density L2 error after one period {e:.3e}", flush=True)
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"[timelevel] convergence rates {rates[0]:.2f}, {rates[1]:.2f}", flush=True)
    assert errs[2] < 1e-4
    assert rates[1] >= 3.5


def test_vortex_100_steps_vs_oracle(device):
    import torch

    _say("vortex 16x16 p=3, SSP-RK3 CFL 0.5, 100 steps vs oracle")
    m = cases.vortex_mesh(n=16, order=3)
    u0 = cases.vortex_state(m)
    dt = cases.cfl_dt(m, 3.342, 0.5)
    op = capi.Operator(m)
    d = torch.from_numpy(u0).to(device)
    op.rk_run(capi.SSPRK3, d, 0.0, dt, 100)
    ref = oracle.rk_run(m, oracle.phys(), u0, 0, 100, dt)
    err = np.abs(d.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"[timelevel] dt = {dt:.4e}; 100 steps: max diff vs oracle {err:.3e}", flush=True)
    assert abs(dt - 5.84e-3) < 1e-5
    assert err <= 1e-10
    op.close()


def test_khi_hooked_vs_oracle(device):
    import torch

    _say("KHI 32x32 p=3, SSP-RK3 + filter + limiter, 40 steps vs oracle")
    m = cases.khi_mesh(n=32, order=3)
    u0 = cases.khi_state(m)
    lam = cases.max_wave_speed(u0)
    dt = cases.cfl_dt(m, lam, 0.5)
    hooks = capi.HOOK_FILTER | capi.HOOK_LIMITER
    op = capi.Operator(m, stage_hooks=hooks)
    op.hook_counts()
    d = torch.from_numpy(u0).to(device)
    op.rk_run(capi.SSPRK3, d, 0.0, dt, 40)
    nf, nl = op.hook_counts()
    ref, counts = oracle.rk_run_hooked(m, oracle.phys(), u0, 0, 40, dt, hooks)
    out = d.cpu().numpy()
    err = np.abs(out - ref).max() / np.abs(ref).max()
    W = m.quadrature_weights()
    cons = [abs((W * (out[v] - u0[v])).sum()) / (W * np.abs(u0[v])).sum() for v in range(4)]
    print(f"[timelevel] KHI dt {dt:.3e}: filtered {nf} (oracle {counts[0]}), limited {nl} (oracle {counts[1]}), "
          f"max diff {err:.3e}, conservation {max(cons):.2e}", flush=True)
    assert (nf, nl) == (counts[0], counts[1])
    assert err <= 1e-10
    assert max(cons) <= 1e-13
    op.close()
