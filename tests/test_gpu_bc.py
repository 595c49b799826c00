"""Boundary conditions on the GPU against the oracle (the reference's BdrFaceIntegrator
hierarchy restated in oracle/prandtl_oracle.cpp and pinned in tests/test_oracle_bc.py),
through the C ABI.

* the RHS with every condition: channels (x open) with supersonic inflow/outflow, specified
  states, Riemann invariants, subsonic (rho, V) and (pt, Tt, angle) inflow against subsonic
  pressure outflow; walls (y open) with slip and symmetry (Euler) and moving no-slip adiabatic
  and isothermal walls (Navier-Stokes); a box open in both x and y (corner elements with two
  boundary faces); 2D and 3D, p = 2 and 3, on states with flow through and into the
  boundaries: every entry within 64 eps of its rounding scale, the wave speed to 1e-12;
* blended RHS with boundaries (the oracle fed the GPU's coefficients): the same bound;
* SSP-RK3 and LSRK4(5) runs (a subsonic channel with slip walls, Euler; a lid-driven box with
  isothermal walls, Navier-Stokes) against the oracle's stepping: 1e-11 relative;
* the entropy filter's bounds on a mesh with boundary faces (no neighbour across them);
* the adjoint refuses a mesh with boundary faces.
"""
import numpy as np
import pytest

import oracle
from prandtl_amd import capi, cases, mesh as M

pytestmark = pytest.mark.gpu
G = 1.4


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _st(dim, rho, vel, p):
    return [rho] + [rho * v for v in vel] + [p / (G - 1) + 0.5 * rho * sum(v * v for v in vel)]


def _flow(mesh, seed, base_vel, amp=0.05):
    rng = np.random.default_rng(seed)
    X = mesh.node_coords()
    rho = 1.0 + amp * np.sin(2 * np.pi * X[0] + 0.3) * np.cos(2 * np.pi * X[1])
    vel = [base_vel[d] + amp * np.cos(2 * np.pi * (X[0] + X[-1]) + d) for d in range(mesh.dim)]
    p = 1.0 + amp * np.cos(2 * np.pi * X[1] + 0.7) + 0.01 * amp * rng.uniform(-1, 1, X[0].shape)
    return cases.prim_to_conserv(rho, vel, p)


def _box(dim, order, open_dirs, bmap):
    n = (5, 4) if dim == 2 else (4, 3, 3)
    per = tuple(d not in open_dirs for d in range(dim))
    return M.periodic_lattice(dim, order, n, lengths=[1.0, 0.8, 0.6][:dim], periodic=per, boundary=bmap)


def _pairs(dim):
    a = np.sqrt(G)
    sup = [2.0 * a] + [0.1] * (dim - 1)
    sub = [0.4] + [0.05] * (dim - 1)
    cp = G / (G - 1) * 287.0
    T = 1.0 / 287.0
    Tt = T + 0.5 * sub[0] ** 2 / cp
    pt = (Tt / T) ** (G / (G - 1))
    return [
        ("supersonic", sup, [oracle.BC_SUPERSONIC_INFLOW] + _st(dim, 1.0, sup, 1.0), [oracle.BC_SUPERSONIC_OUTFLOW]),
        ("specified", sub, [oracle.BC_SPECIFIED_STATE] + _st(dim, 1.05, sub, 0.98),
         [oracle.BC_SPECIFIED_STATE] + _st(dim, 0.97, sub, 1.02)),
        ("riemann", sub, [oracle.BC_RIEMANN_INVARIANT, 1.02] + sub + [1.01],
         [oracle.BC_RIEMANN_INVARIANT, 0.99] + sub + [0.99]),
        ("subsonic-rv", sub, [oracle.BC_SUBSONIC_INFLOW_RV, 1.03] + sub, [oracle.BC_SUBSONIC_OUTFLOW_P, 0.97]),
        ("subsonic-pttt", sub, [oracle.BC_SUBSONIC_INFLOW_PTTT, pt, Tt, 5.0, 80.0, cp],
         [oracle.BC_SUBSONIC_OUTFLOW_P, 0.99]),
    ]


def _compare(m, u, kw, obcs, device, blending=False):
    import torch

    cbcs = [capi.bc(b.type, *list(b.q)) for b in obcs]
    op = capi.Operator(m, bcs=cbcs, blending=blending, **kw)
    d = torch.from_numpy(u).to(device)
    du = torch.empty_like(d)
    op.mult(d, du)
    lam = op.get_max_char_speed()
    a = op.blend_coefficients(d)[0].cpu().numpy() if blending else None
    out = du.cpu().numpy()
    op.close()
    ph = oracle.phys(**{k: v for k, v in kw.items() if k in ("viscous", "mu", "Pr", "surface_flux")}, R_gas=287.0)
    ref, lref, mag = oracle.rhs_ex(m, ph, u, obcs, alpha=a)
    s = oracle.scaled_error(out, ref, mag)
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[bc] rel.err {rel:.3e}, entrywise {s:.2f} eps*mag, lambda {lam:.12e} vs {lref:.12e}", flush=True)
    assert s <= 64
    assert abs(lam - lref) <= 1e-12 * lref
    return a


@pytest.mark.parametrize("dim,order", [(2, 3), (2, 2), (3, 3), (3, 2)])
@pytest.mark.parametrize("viscous", [False, True])
def test_bc_channels(device, dim, order, viscous):
    kw = dict(viscous=viscous, mu=5e-3 if viscous else 0.0, Pr=0.71, R_gas=287.0)
    for name, vel, lo, hi in _pairs(dim):
        _say(f"channel {name} dim={dim} p={order} viscous={viscous}")
        m = _box(dim, order, (0,), {(0, 0): 0, (0, 1): 1})
        u = _flow(m, order + 10 * dim, vel)
        _compare(m, u, kw, [oracle.bc(*lo), oracle.bc(*hi)], device)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("surf", [0, 1, 2])
def test_bc_walls_euler(device, dim, surf):
    for kind in (oracle.BC_SLIP_WALL, oracle.BC_SYMMETRY):
        _say(f"walls type={kind} dim={dim} surf={surf}")
        m = _box(dim, 3, (1,), {(1, 0): 0, (1, 1): 0})
        # the wall-normal velocity flows into one wall and away from the other
        u = _flow(m, 5 + dim, [0.3, 0.25, 0.1][:dim], amp=0.08)
        _compare(m, u, dict(surface_flux=surf), [oracle.bc(kind)], device)


@pytest.mark.parametrize("dim", [2, 3])
def test_bc_walls_ns(device, dim):
    kw = dict(viscous=True, mu=1e-2, Pr=0.71, R_gas=287.0)
    T = 1.0 / 287.0
    for lo, hi in (([oracle.BC_NOSLIP_ADIABATIC] + [0.0] * dim + [0.0], [oracle.BC_NOSLIP_ADIABATIC, 0.3] + [0.0] * (dim - 1) + [0.02]),
                   ([oracle.BC_NOSLIP_ISOTHERMAL] + [0.0] * dim + [T], [oracle.BC_NOSLIP_ISOTHERMAL, 0.3] + [0.0] * (dim - 1) + [1.1 * T])):
        _say(f"no-slip walls type={lo[0]} dim={dim}")
        m = _box(dim, 3, (1,), {(1, 0): 0, (1, 1): 1})
        u = _flow(m, 7 + dim, [0.2, 0.1, 0.05][:dim])
        _compare(m, u, kw, [oracle.bc(*lo), oracle.bc(*hi)], device)


@pytest.mark.parametrize("viscous", [False, True])
def test_bc_box_corners_and_blending(device, viscous):
    dim = 2
    kw = dict(viscous=viscous, mu=5e-3 if viscous else 0.0, Pr=0.71, R_gas=287.0)
    vel = [0.4, 0.05]
    wall = [oracle.BC_NOSLIP_ISOTHERMAL, 0.0, 0.0, 1.0 / 287.0] if viscous else [oracle.BC_SLIP_WALL]
    bcs = [oracle.bc(oracle.BC_SUBSONIC_INFLOW_RV, 1.02, *vel), oracle.bc(oracle.BC_SUBSONIC_OUTFLOW_P, 0.98),
           oracle.bc(*wall)]
    m = _box(dim, 3, (0, 1), {(0, 0): 0, (0, 1): 1, (1, 0): 2, (1, 1): 2})
    X = m.node_coords()
    u = _flow(m, 3, vel)
    # a steep front so that blending acts in some elements
    u[0] *= 1.0 + 0.3 * np.tanh((X[0] - 0.5) / 0.02)
    _say(f"box with corners viscous={viscous}")
    _compare(m, u, kw, bcs, device)
    _say(f"box with corners, blending, viscous={viscous}")
    a = _compare(m, u, kw, bcs, device, blending=True)
    assert (a > 0).any()


@pytest.mark.parametrize("scheme,nsteps", [(capi.SSPRK3, 5), (capi.LSRK45, 3)])
def test_bc_steps(device, scheme, nsteps):
    import torch

    # Euler: subsonic channel with slip walls
    m = _box(2, 3, (0, 1), {(0, 0): 0, (0, 1): 1, (1, 0): 2, (1, 1): 2})
    vel = [0.4, 0.0]
    obcs = [oracle.bc(oracle.BC_SUBSONIC_INFLOW_RV, 1.0, *vel), oracle.bc(oracle.BC_SUBSONIC_OUTFLOW_P, 1.0),
            oracle.bc(oracle.BC_SLIP_WALL)]
    u0 = _flow(m, 11, vel)
    cases_ = [(m, u0, {}, obcs)]
    # Navier-Stokes: lid-driven box with isothermal walls
    T = 1.0 / 287.0
    mb = _box(2, 3, (0, 1), {(0, 0): 0, (0, 1): 0, (1, 0): 0, (1, 1): 1})
    ub = _flow(mb, 12, [0.0, 0.0], amp=0.02)
    ob = [oracle.bc(oracle.BC_NOSLIP_ISOTHERMAL, 0.0, 0.0, T), oracle.bc(oracle.BC_NOSLIP_ISOTHERMAL, 0.5, 0.0, T)]
    cases_.append((mb, ub, dict(viscous=True, mu=1e-2, Pr=0.71, R_gas=287.0), ob))
    for mm, uu, kw, bcs in cases_:
        _say(f"steps scheme={scheme} viscous={bool(kw)}")
        lam = cases.max_wave_speed(uu)
        dt = cases.cfl_dt(mm, lam, 0.3)
        ph = oracle.phys(**{k: v for k, v in kw.items() if k in ("viscous", "mu", "Pr")}, R_gas=287.0)
        ref, _ = oracle.rk_run_ex(mm, ph, uu, scheme, nsteps, dt, bcs)
        op = capi.Operator(mm, bcs=[capi.bc(b.type, *list(b.q)) for b in bcs], **kw)
        d = torch.from_numpy(uu).to(device)
        op.rk_run(scheme, d, 0.0, dt, nsteps)
        torch.cuda.synchronize()
        out = d.cpu().numpy()
        op.close()
        rel = np.abs(out - ref).max() / np.abs(ref).max()
        print(f"[bc] {nsteps} steps: rel.err {rel:.3e}", flush=True)
        assert rel <= 1e-11


def test_bc_filter_bounds_and_refusals(device):
    import torch

    _say("filter bounds with boundary faces; adjoint refusal")
    m = _box(2, 3, (1,), {(1, 0): 0, (1, 1): 0})
    u0 = _flow(m, 13, [0.3, 0.0])
    u = u0.copy()
    rng = np.random.default_rng(2)
    elems = [0, 7, 18]                     # 0 and 18 touch the walls
    for e in elems:
        i = int(rng.integers(0, u.shape[2]))
        u[-1, e, i] = 0.5 * (u[1, e, i] ** 2 + u[2, e, i] ** 2) / u[0, e, i] - 0.05
    bound = oracle.filter_constraints(m, u0)
    ref, _, cnt = oracle.filter_apply(m, u, bound)
    op = capi.Operator(m, bcs=[capi.bc(capi.BC_SLIP_WALL)])
    op.hook_counts()
    op.filter_constraints(torch.from_numpy(u0).to(device))
    d = torch.from_numpy(u).to(device)
    op.filter_apply(d)
    nf, _ = op.hook_counts()
    out = d.cpu().numpy()
    assert nf == cnt == len(elems)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
    with pytest.raises(capi.PrandtlError, match="UNSUPPORTED"):
        op.vjp(d, torch.ones_like(d), torch.empty_like(d))
    op.close()
