"""Subcell finite-volume blending on the GPU against the oracle (pinned in
tests/test_oracle_blend.py), through the C ABI.

* blending coefficients: k_blend_coef's eta and alpha per element against the oracle's
  Persson-Peraire indicator on states with a steep front (a mix of alpha = 0, interior values
  and alpha = alpha_max), 2D/3D, p = 2..4.  The kernel's modal transform is the sum-factorised
  tensor inverse, the oracle's the dense n x n inverse: eta within 1e-13 absolute +
  1e-11 relative; alpha, without the clips, within 1e-9 (the logistic slope is 9.2/T, about
  4,300 at p = 3); with the clips, identical;
* the blended RHS (prandtl_op_mult with blending on) against the oracle's blended RHS fed the
  GPU's own coefficients (so the comparison isolates the FV and blend arithmetic): every entry
  within 64 eps of its rounding scale (the oracle's assembly with every term in absolute
  value), Euler and Navier-Stokes, all three surface fluxes, 2D/3D, p = 1..4;
* blended SSP-RK3 and LSRK4(5) steps against the oracle's blended stepping (coefficients
  recomputed from every stage state): within 1e-11 relative;
* mixed precision with blending and the adjoint of a blended operator are refused
  (PRANDTL_ERR_UNSUPPORTED).
"""
import numpy as np
import pytest

import oracle
from prandtl_amd import capi, cases, mesh as M

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _front_state(mesh, seed=0, amp=0.3, width=0.03, noise=0.02):
    """A density/pressure front across x = L/2 plus smooth variation along the other axes."""
    rng = np.random.default_rng(seed)
    X = mesh.node_coords()
    L = X[0].max()
    step = np.tanh((X[0] - 0.5 * L) / width)
    rho = 1.0 + amp * step + 0.05 * np.sin(2 * np.pi * X[-1])
    vel = [0.2 * np.cos(2 * np.pi * X[-1] + d) for d in range(mesh.dim)]
    p = 1.0 + 0.8 * amp * step + noise * rng.uniform(-1, 1, X[0].shape)
    return cases.prim_to_conserv(rho, vel, p)


def _mesh(dim, order):
    if dim == 2:
        return M.periodic_lattice(2, order, (12, 4), lengths=[1.0, 0.5])
    return M.periodic_lattice(3, order, (6, 3, 2), lengths=[1.0, 0.5, 0.4])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [2, 3, 4])
def test_blend_coefficients(device, dim, order):
    import torch

    _say(f"blend coefficients dim={dim} p={order}")
    m = _mesh(dim, order)
    u = _front_state(m, seed=order, noise=1e-4)
    d = torch.from_numpy(u).to(device)
    # without the clips: the continuous coefficient
    op = capi.Operator(m, blending=True, alpha_max=1.0, alpha_min=0.0)
    a, eta = (t.cpu().numpy() for t in op.blend_coefficients(d))
    op.close()
    a_ref, eta_ref = oracle.blend_coef(m, u, alpha_max=1.0, alpha_min=0.0)
    deta = np.abs(eta - eta_ref)
    print(f"[blend] eta range [{eta_ref.min():.2e}, {eta_ref.max():.2e}], max |d eta| {deta.max():.2e}, "
          f"max |d alpha| {np.abs(a - a_ref).max():.2e}", flush=True)
    assert (deta <= 1e-13 + 1e-11 * np.abs(eta_ref)).all()
    assert np.abs(a - a_ref).max() <= 1e-9
    # the reference's clips
    op = capi.Operator(m, blending=True)
    a, _ = (t.cpu().numpy() for t in op.blend_coefficients(d))
    op.close()
    a_ref, _ = oracle.blend_coef(m, u)
    print(f"[blend] alpha: {int((a_ref == 0).sum())} zero, {int((a_ref == capi.ALPHA_MAX).sum())} capped, "
          f"{int(((a_ref > 0) & (a_ref < capi.ALPHA_MAX)).sum())} between", flush=True)
    assert (a_ref == 0).any() and (a_ref == capi.ALPHA_MAX).any()
    assert np.abs(a - a_ref).max() <= 1e-9


RHS_CASES = [(2, 1, False, 0), (2, 2, False, 0), (2, 3, False, 0), (2, 4, False, 0), (2, 3, False, 1),
             (2, 3, False, 2), (2, 3, True, 0), (2, 4, True, 2), (3, 1, False, 0), (3, 2, True, 0),
             (3, 3, False, 0), (3, 3, True, 0), (3, 3, True, 1), (3, 3, False, 2), (3, 4, True, 0)]


@pytest.mark.parametrize("dim,order,viscous,surf", RHS_CASES)
def test_blend_rhs(device, dim, order, viscous, surf):
    import torch

    _say(f"blended RHS dim={dim} p={order} viscous={viscous} surf={surf}")
    m = _mesh(dim, order)
    u = _front_state(m, seed=10 * dim + order)
    kw = dict(viscous=viscous, mu=2e-3 if viscous else 0.0, Pr=0.71, surface_flux=surf)
    op = capi.Operator(m, blending=True, **kw)
    d = torch.from_numpy(u).to(device)
    du = torch.empty_like(d)
    op.mult(d, du)
    a, _ = op.blend_coefficients(d)
    torch.cuda.synchronize()
    a = a.cpu().numpy()
    out = du.cpu().numpy()
    op.close()
    ph = oracle.phys(**kw)
    ref, _, mag = oracle.rhs_blend(m, ph, u, a)
    s = oracle.scaled_error(out, ref, mag)
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[blend] alpha>0 in {int((a > 0).sum())}/{a.size} elements; rel.err {rel:.3e}, entrywise {s:.2f} eps*mag",
          flush=True)
    assert (a > 0).any() or order == 1
    assert s <= 64


@pytest.mark.parametrize("scheme,nsteps", [(capi.SSPRK3, 4), (capi.LSRK45, 2)])
@pytest.mark.parametrize("dim", [2, 3])
def test_blend_steps(device, scheme, nsteps, dim):
    import torch

    _say(f"blended steps scheme={scheme} dim={dim}")
    order = 3
    m = _mesh(dim, order)
    u0 = _front_state(m, seed=3, amp=0.4)
    ph = oracle.phys()
    lam = cases.max_wave_speed(u0)
    dt = cases.cfl_dt(m, lam, 0.3)
    ref, _ = oracle.rk_run_blend(m, ph, u0, scheme, nsteps, dt)
    op = capi.Operator(m, blending=True)
    d = torch.from_numpy(u0).to(device)
    op.rk_run(scheme, d, 0.0, dt, nsteps)
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    op.close()
    rel = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"[blend] {nsteps} steps dt={dt:.3e}: rel.err {rel:.3e}", flush=True)
    assert rel <= 1e-11


def test_blend_refusals(device):
    import torch

    _say("blending refusals")
    m = _mesh(2, 3)
    with pytest.raises(capi.PrandtlError, match="UNSUPPORTED"):
        capi.Operator(m, blending=True, precision=capi.MIXED)
    with pytest.raises(capi.PrandtlError, match="ARG"):
        capi.Operator(m, blending=True, alpha_min=0.7)
    op = capi.Operator(m, blending=True)
    u = torch.from_numpy(_front_state(m)).to(device)
    with pytest.raises(capi.PrandtlError, match="UNSUPPORTED"):
        op.vjp(u, torch.ones_like(u), torch.empty_like(u))
    op.close()
