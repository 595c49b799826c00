"""Function-valued boundary states on the GPU against the oracle (the reference's ConditionFactory
functions -- DoubleMachReflectionTopBCFunction's moving shock front and NagashimaRampingInletBC's
velocity ramp -- restated in oracle/prandtl_oracle.cpp and pinned in tests/test_oracle_bc.py),
through the C ABI.

* the RHS of the double Mach reflection (Euler, 48 x 12 lattice of [0,4] x [0,1], p = 2 and 3,
  blending off and on) at several operator times (prandtl_op_set_time): every entry within 64 eps
  of its rounding scale;
* SSP-RK3 and LSRK4(5) runs from t0 > 0 (each stage at its stage time) against the oracle's
  stepping: 1e-11 relative, the same end time;
* the velocity ramp on a subsonic-free channel inflow, Euler and Navier-Stokes (the lifting and
  viscous boundary terms read the same state), at three times;
* the reference's double-Mach-reflection configuration (its 384 x 96 lattice refined once:
  768 x 192 quadrilaterals, p = 3, 2,359,296 nodes, blending, SSP-RK3) stepped to its final time
  t = 0.04: finite, positive, and the incident shock where the top boundary's front puts it
  (density behind it near the post-shock 8, ahead of it the resting 1.4).  The configuration's
  fixed dt = 1e-4 is four times the reference's own CFL bound on this mesh (dt = cfl h /
  (lambda (p+1)^2), Simulation.cpp:906, with h = 1/192 and lambda = 12.8: 2.5e-5 at cfl 1), so
  the run takes that formula's dt at CFL 0.5 (its variable_dt setting).
"""
import os

import numpy as np
import pytest

import oracle
from prandtl_amd import capi, cases, config

pytestmark = pytest.mark.gpu
G = 1.4
HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = os.path.join(HERE, "golden", "configs")


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _dmr_tables():
    post, pre = config._dmr_states(1.0 / (G - 1.0))
    _, _, a = config._dmr_top({"x1": 1.0 / (G - 1.0)})
    spec = [(oracle.BC_SPECIFIED_STATE, post, 0, []), (oracle.BC_SPECIFIED_STATE, post, 0, []),
            (oracle.BC_SLIP_WALL, [], 0, []), (oracle.BC_SUPERSONIC_OUTFLOW, [], 0, []),
            (oracle.BC_SPECIFIED_STATE, post, oracle.PROFILE_FRONT, a)]
    obcs = [oracle.bc(t, *q, profile=pr, a=aa) for t, q, pr, aa in spec]
    cbcs = [capi.bc(t, *q, profile=pr, a=aa) for t, q, pr, aa in spec]
    return obcs, cbcs


def _rhs_compare(m, u, obcs, cbcs, device, t, blending=False, kw=None):
    import torch

    kw = kw or {}
    op = capi.Operator(m, bcs=cbcs, blending=blending, **kw)
    op.set_time(t)
    assert op.get_time() == t
    d = torch.from_numpy(u).to(device)
    du = torch.empty_like(d)
    op.mult(d, du)
    a = op.blend_coefficients(d)[0].cpu().numpy() if blending else None
    out = du.cpu().numpy()
    op.close()
    ph = oracle.phys(**{k: v for k, v in kw.items() if k in ("viscous", "mu", "Pr")}, R_gas=287.0)
    ref, _, mag = oracle.rhs_ex(m, ph, u, obcs, alpha=a, time=t)
    s = oracle.scaled_error(out, ref, mag)
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[bc-profile] t={t} rel.err {rel:.3e}, entrywise {s:.2f} eps*mag", flush=True)
    assert s <= 64
    return out


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("blending", [False, True])
def test_dmr_front_rhs(device, order, blending):
    _say(f"double Mach reflection front RHS p={order} blending={blending}")
    m = cases.dmr_mesh(48, 12, order)
    u = cases.dmr_state(m)
    obcs, cbcs = _dmr_tables()
    outs = [_rhs_compare(m, u, obcs, cbcs, device, t, blending) for t in (0.0, 0.0137, 0.04)]
    # the time reaches the kernels: the RHS moves with the front
    assert not np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("scheme,oscheme,nsteps", [(capi.SSPRK3, 0, 6), (capi.LSRK45, 1, 3)])
def test_dmr_front_steps(device, scheme, oscheme, nsteps):
    import torch

    _say(f"double Mach reflection front: {nsteps} steps of scheme {scheme} from t0 = 0.021")
    m = cases.dmr_mesh(48, 12, 3)
    u0 = cases.dmr_state(m)
    obcs, cbcs = _dmr_tables()
    t0, dt = 0.021, 2e-4
    op = capi.Operator(m, bcs=cbcs, blending=True)
    d = torch.from_numpy(u0).to(device)
    t = op.rk_run(scheme, d, t0, dt, nsteps)
    out = d.cpu().numpy()
    op.close()
    ref, _, t_ref = oracle.rk_run_ex(m, oracle.phys(), u0, oscheme, nsteps, dt, obcs, blend=True, t0=t0,
                                     return_time=True)
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[bc-profile] steps: rel.err {rel:.3e}, t {t!r} vs {t_ref!r}", flush=True)
    assert t == t_ref
    assert rel <= 1e-11


@pytest.mark.parametrize("viscous", [False, True])
def test_ramp_inflow(device, viscous):
    _say(f"ramped inflow viscous={viscous}")
    from prandtl_amd import mesh as M

    # a channel with the ramped inflow on the left, a free outflow on the right (a supersonic stream
    # at full speed; nondimensional: rho 1, p 1, full velocity 2 sqrt(gamma))
    m = M.periodic_lattice(2, 3, (6, 4), lengths=[1.2, 0.8], periodic=(False, True),
                           boundary={(0, 0): 0, (0, 1): 1})
    V = 2.0 * np.sqrt(G)
    tr = 0.05
    spec = [(oracle.BC_SUPERSONIC_INFLOW, [], oracle.PROFILE_RAMP, [tr, 1.0, V, 0.0, 1.0]),
            (oracle.BC_SUPERSONIC_OUTFLOW, [], 0, [])]
    obcs = [oracle.bc(t, *q, profile=pr, a=a) for t, q, pr, a in spec]
    cbcs = [capi.bc(t, *q, profile=pr, a=a) for t, q, pr, a in spec]
    X = m.node_coords()
    rho = 1.0 + 0.05 * np.sin(2 * np.pi * X[1] / 0.8) * np.cos(X[0])
    u = cases.prim_to_conserv(rho, [0.3 * V + 0.05 * X[1], 0.02 * np.cos(X[0])], 1.0 + 0.02 * X[0])
    kw = dict(viscous=True, mu=2e-3, Pr=0.71) if viscous else {}
    outs = [_rhs_compare(m, u, obcs, cbcs, device, t, kw=kw) for t in (0.0, 0.5 * tr, 2.0 * tr)]
    assert not np.array_equal(outs[0], outs[1])


def test_dmr_reference_configuration(device):
    import torch

    _say("double Mach reflection from its configuration: 768 x 192 p=3 to t = 0.04")
    m = cases.dmr_mesh(768, 192, 3)
    case = config.load_case(os.path.join(CONFIGS, "DoubleMachReflection.json"), mesh=m)
    u0 = config.reference_initial_state(case)
    op = case.operator()
    d = torch.from_numpy(u0).to(device)
    lam = op.max_char_speed(d)
    nsteps = int(np.ceil(case.final_time / cases.cfl_dt(m, lam, 0.5)))
    dt = case.final_time / nsteps
    torch.cuda.synchronize()
    import time

    t0 = time.perf_counter()
    t = op.rk_run(case.scheme, d, 0.0, dt, nsteps, nan_check=500)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = d.cpu().numpy()
    op.close()
    rho, vel, p = cases.conserv_to_prim(out)
    print(f"[bc-profile] DMR {m.num_nodes} nodes, {nsteps} steps to t={t:.4f} in {wall:.2f} s "
          f"({m.num_nodes * 3 * nsteps / wall / 1e9:.2f} G DOF-RHS/s); rho in [{rho.min():.3f}, {rho.max():.3f}], "
          f"p_min {p.min():.4f}", flush=True)
    assert abs(t - 0.04) < 1e-11
    assert np.isfinite(out).all() and rho.min() > 0 and p.min() > 0
    # along the top boundary the incident shock sits where the front is at t
    X = m.node_coords()
    xs = 1.0 / 6.0 + (1.0 + 20.0 * t) / np.sqrt(3)
    top = np.isclose(X[1], 1.0)
    behind = top & (X[0] < xs - 0.1)
    ahead = top & (X[0] > xs + 0.1)
    print(f"[bc-profile] top boundary: shock at x = {xs:.4f}; rho behind {rho[behind].mean():.3f}, "
          f"ahead {rho[ahead].mean():.3f}", flush=True)
    assert abs(rho[ahead].mean() - 1.4) < 0.02
    assert rho[behind].min() > 5.0
