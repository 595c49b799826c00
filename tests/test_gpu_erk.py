"""MFEM's explicit Runge-Kutta steppers (ForwardEuler, RK2 with a = 2/3, classic RK4; pinned on
the CPU in tests/test_oracle_operator.py) on the GPU against the oracle's stepping: Euler and
Navier-Stokes on the periodic vortex / TGV, with the stage hooks, and on a configured case."""
import os

import numpy as np
import pytest

import oracle
from prandtl_amd import capi, cases, config

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.mark.parametrize("scheme", [capi.FORWARD_EULER, capi.RK2, capi.RK4])
@pytest.mark.parametrize("viscous", [False, True])
def test_explicit_rk_steps(device, scheme, viscous):
    import torch

    _say(f"explicit RK scheme={scheme} viscous={viscous}")
    if viscous:
        m = cases.tgv_mesh(n=4, order=3)
        u0 = cases.tgv_state(m)
        kw = cases.tgv_physics()
        ph = oracle.phys(gamma=kw["gamma"], Pr=kw["Pr"], mu=kw["mu"], viscous=True)
    else:
        m = cases.vortex_mesh(n=8, order=3)
        u0 = cases.vortex_state(m)
        kw, ph = {}, oracle.phys()
    dt = cases.cfl_dt(m, cases.max_wave_speed(u0), 0.2)
    nsteps = 3
    ref, _ = oracle.rk_run_ex(m, ph, u0, scheme, nsteps, dt)
    op = capi.Operator(m, **kw)
    d = torch.from_numpy(u0).to(device)
    t = op.rk_run(scheme, d, 0.0, dt, nsteps, nan_check=1)
    out = d.cpu().numpy()
    op.close()
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[erk] {nsteps} steps: rel.err {rel:.3e}", flush=True)
    assert rel <= 1e-11 and abs(t - nsteps * dt) < 1e-15


def test_explicit_rk_with_hooks_and_refusals(device):
    import torch

    _say("RK4 with the filter and limiter hooks; the taped run refuses it")
    m = cases.vortex_mesh(n=8, order=3)
    u0 = cases.vortex_state(m)
    dt = cases.cfl_dt(m, cases.max_wave_speed(u0), 0.2)
    ref, _ = oracle.rk_run_ex(m, oracle.phys(), u0, 4, 2, dt, hooks=3)
    op = capi.Operator(m, stage_hooks=capi.HOOK_FILTER | capi.HOOK_LIMITER)
    d = torch.from_numpy(u0).to(device)
    op.rk_run(capi.RK4, d, 0.0, dt, 2)
    out = d.cpu().numpy()
    assert np.abs(out - ref).max() / np.abs(ref).max() <= 1e-11
    with pytest.raises(capi.PrandtlError):
        op.rk_run_tape(d, 0.0, dt, 2, 1, scheme=capi.RK4)
    with pytest.raises(capi.PrandtlError, match="ODE solver"):
        op.rk_run(7, d, 0.0, dt, 1)
    op.close()
