"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar, on the same seeded inputs, for every dimension, order, surface flux and the Euler /
Navier-Stokes paths, at every entry of the RHS (and of the primitive gradients):

    |gpu - oracle| <= C eps mag,   C = 64,

where mag is the oracle's rounding scale of that entry: the same assembly with every
term in absolute value (oracle.rhs_mag).  The relative max-norm difference
||gpu - oracle||_inf / ||oracle||_inf is printed beside it, with the same quantity for two
builds of the oracle itself (without and with FMA contraction): on pressure-dominated
states (Taylor-Green, p ~ 71, RHS O(1)) the momentum and energy sums cancel by three to
five digits and that spread alone reaches 7e-13 at 6^3 and 4e-11 at 32^3, so a fixed
1e-12 relative bar cannot separate rounding from error there; the entrywise bound does.
On well-conditioned states the relative difference is printed and also held to 1e-12.
RK steps (a few steps from a smooth state) are held to 1e-12 relative.  At the full
benchmark size (TGV 32^3, p=3) the RHS is compared entrywise as well, plus
size-independent properties: discrete conservation, free-stream preservation, bitwise
determinism, entropy conservation with the EC flux.
"""
import numpy as np
import pytest

import oracle
from prandtl_amd import capi, cases, mesh as M

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def relerr(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


C_ROUND = 64.0


def assert_parity(gpu, ref, ref_fma, what="", mag=None):
    """Entrywise |gpu - ref| <= C eps mag when mag is given; otherwise relative <= 1e-12."""
    floor = oracle.rounding_floor(ref, ref_fma)
    err = relerr(gpu, ref)
    if mag is not None:
        sc = oracle.scaled_error(gpu, ref, mag)
        scf = oracle.scaled_error(ref_fma, ref, mag)
        print(f"[parity] {what} rel.err={err:.3e} (oracle fma spread {floor:.3e}); "
              f"entrywise {sc:.2f} eps*mag (oracle fma spread {scf:.2f}, bound {C_ROUND:.0f})", flush=True)
        assert sc <= C_ROUND, (sc, err, floor)
        if floor <= 1e-13:
            assert err <= TOL, (err, floor)
    else:
        print(f"[parity] {what} rel.err={err:.3e} (oracle fma spread {floor:.3e}, bound {TOL:.0e})", flush=True)
        assert err <= TOL, (err, floor)


def sheared_mesh(dim, order, n):
    if dim == 2:
        A = np.array([[0.5, 0.1], [0.05, 0.45]])
    else:
        A = np.array([[0.5, 0.1, 0.0], [0.05, 0.45, 0.08], [0.02, -0.03, 0.55]])
    return M.periodic_lattice(dim, order, n, A=A)


def perturbed_state(mesh, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    X = mesh.node_coords()
    rho = 1.0 + 0.2 * np.sin(X[0]) + amp * rng.uniform(-1, 1, X[0].shape)
    vel = [0.3 * np.cos(X[(d + 1) % mesh.dim]) + amp * rng.uniform(-1, 1, X[0].shape) for d in range(mesh.dim)]
    p = 1.0 + 0.1 * np.cos(X[0] + X[1]) + amp * rng.uniform(-1, 1, X[0].shape)
    return cases.prim_to_conserv(rho, vel, p)


def gpu_rhs(mesh, u, device, grad=False, **kw):
    import torch

    op = capi.Operator(mesh, **kw)
    ud = torch.from_numpy(np.ascontiguousarray(u)).to(device)
    du = torch.empty_like(ud)
    g = None
    if grad:
        g = torch.empty((mesh.dim,) + tuple(ud.shape), dtype=torch.float64, device=device)
        op.mult_grad(ud, du, g)
    else:
        op.mult(ud, du)
    torch.cuda.synchronize()
    lam = op.get_max_char_speed()
    out = du.cpu().numpy(), lam, (g.cpu().numpy() if grad else None)
    op.close()
    return out


SURF = [oracle.SURF_CHANDRASHEKAR, oracle.SURF_EC, oracle.SURF_RUSANOV]


@pytest.mark.parametrize("surf", SURF)
def test_vortex_2d_euler(device, surf):
    _say(f"vortex 16x16 p=3 Euler surf={surf}")
    m = cases.vortex_mesh(n=16, order=3)
    u = cases.vortex_state(m)
    # element-wise jitter so surface fluxes see real jumps
    u = u * (1.0 + 1e-3 * np.random.default_rng(1).uniform(-1, 1, u.shape))
    d, lam, _ = gpu_rhs(m, u, device, surface_flux=surf)
    ref, lref, mag = oracle.rhs_mag(m, oracle.phys(surface_flux=surf), u)
    reff, _ = oracle.rhs(m, oracle.phys(surface_flux=surf), u, fma=True)
    assert_parity(d, ref, reff, "vortex", mag)
    assert abs(lam - lref) <= 1e-14 * lref


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
@pytest.mark.parametrize("viscous", [False, True])
def test_sheared_all_orders(device, dim, order, viscous):
    _say(f"sheared dim={dim} p={order} viscous={viscous}")
    n = (5, 4) if dim == 2 else (3, 2, 3)
    m = sheared_mesh(dim, order, n)
    u = perturbed_state(m, 10 * dim + order)
    kw = dict(viscous=viscous, mu=0.02, Pr=0.71)
    d, lam, _ = gpu_rhs(m, u, device, **kw)
    ref, lref, mag = oracle.rhs_mag(m, oracle.phys(**kw), u)
    reff, _ = oracle.rhs(m, oracle.phys(**kw), u, fma=True)
    assert_parity(d, ref, reff, "sheared", mag)
    assert abs(lam - lref) <= 1e-14 * lref


@pytest.mark.parametrize("surf", SURF)
def test_tgv_ns_3d(device, surf):
    _say(f"TGV 6^3 p=3 NS surf={surf}")
    m = cases.tgv_mesh(n=6, order=3)
    u = cases.tgv_state(m)
    u = u * (1.0 + 1e-4 * np.random.default_rng(2).uniform(-1, 1, u.shape))
    ph = cases.tgv_physics()
    d, lam, g = gpu_rhs(m, u, device, grad=True, surface_flux=surf, **ph)
    oph = oracle.phys(surface_flux=surf, gamma=ph["gamma"], Pr=ph["Pr"], mu=ph["mu"], viscous=True)
    ref, lref, mag, gref, gmag = oracle.rhs_mag(m, oph, u, want_grad=True)
    reff, _, greff = oracle.rhs(m, oph, u, want_grad=True, fma=True)
    assert_parity(d, ref, reff, "tgv ns", mag)
    # primitive gradients (the reference's dudx/dudy/dudz side effect)
    assert_parity(g, gref, greff, "tgv ns gradients", gmag)
    assert abs(lam - lref) <= 1e-14 * lref


def test_bug_a3_3d(device):
    _say("bug_compat A3, 3D Euler and NS")
    m = sheared_mesh(3, 3, (2, 3, 2))
    u = perturbed_state(m, 7)
    for visc in (False, True):
        kw = dict(viscous=visc, mu=0.01, Pr=0.7)
        d, _, _ = gpu_rhs(m, u, device, bug_compat=capi.BUG_A3, **kw)
        ref, _, mag = oracle.rhs_mag(m, oracle.phys(bug_compat=oracle.BUG_A3, **kw), u)
        reff, _ = oracle.rhs(m, oracle.phys(bug_compat=oracle.BUG_A3, **kw), u, fma=True)
        assert_parity(d, ref, reff, "bug A3", mag)


BUGS_A12 = [capi.BUG_A1, capi.BUG_A2, capi.BUG_A1 | capi.BUG_A2 | capi.BUG_A3]


def _tgv_ns(n, order=3, seed=2):
    m = cases.tgv_mesh(n=n, order=order)
    u = cases.tgv_state(m)
    u = u * (1.0 + 1e-4 * np.random.default_rng(seed).uniform(-1, 1, u.shape))
    return m, u, cases.tgv_physics()


@pytest.mark.parametrize("bugs", BUGS_A12)
def test_bug_a1_a2_tgv(device, bugs):
    # the reference's NS volume term as written (A1: the zeta flux store at DGSEMIntegrator.cpp:702;
    # A2: the viscous flux at the stale node, :620) on the headline Taylor-Green state
    _say(f"bug_compat {bugs}, TGV 6^3 p=3 NS")
    m, u, ph = _tgv_ns(6)
    kw = dict(gamma=ph["gamma"], Pr=ph["Pr"], mu=ph["mu"], viscous=True)
    d, lam, g = gpu_rhs(m, u, device, grad=True, bug_compat=bugs, **ph)
    oph = oracle.phys(bug_compat=bugs, **kw)
    ref, lref, mag, gref, gmag = oracle.rhs_mag(m, oph, u, want_grad=True)
    reff, _, greff = oracle.rhs(m, oph, u, want_grad=True, fma=True)
    assert_parity(d, ref, reff, f"tgv ns bug {bugs}", mag)
    # the gradients are the intended ones (the slips are in the volume assembly)
    _, _, gint = oracle.rhs(m, oracle.phys(**kw), u, want_grad=True)
    assert np.array_equal(gref, gint)
    assert_parity(g, gref, greff, "tgv ns gradients", gmag)
    # the switch changes the operator (its own path ran)
    intended, _ = oracle.rhs(m, oracle.phys(**kw), u)
    print(f"[bug] |as written - intended| / |intended| = {relerr(ref, intended):.3e}", flush=True)
    assert relerr(ref, intended) > 1e-6
    assert abs(lam - lref) <= 1e-14 * lref


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_bug_a1_a2_sheared(device, order):
    _say(f"bug_compat A1|A2|A3 sheared 3D NS p={order}, and A2 in 2D")
    bugs = capi.BUG_A1 | capi.BUG_A2 | capi.BUG_A3
    m = sheared_mesh(3, order, (2, 3, 2))
    u = perturbed_state(m, 11 + order)
    kw = dict(viscous=True, mu=0.01, Pr=0.7)
    for b in (bugs, capi.BUG_A1, capi.BUG_A2):
        d, _, _ = gpu_rhs(m, u, device, bug_compat=b, **kw)
        ref, _, mag = oracle.rhs_mag(m, oracle.phys(bug_compat=b, **kw), u)
        reff, _ = oracle.rhs(m, oracle.phys(bug_compat=b, **kw), u, fma=True)
        assert_parity(d, ref, reff, f"sheared 3D bug {b}", mag)
    m2 = sheared_mesh(2, order, (3, 4))
    u2 = perturbed_state(m2, 5 + order)
    d, _, _ = gpu_rhs(m2, u2, device, bug_compat=capi.BUG_A2, **kw)
    ref, _, mag = oracle.rhs_mag(m2, oracle.phys(bug_compat=capi.BUG_A2, **kw), u2)
    reff, _ = oracle.rhs(m2, oracle.phys(bug_compat=capi.BUG_A2, **kw), u2, fma=True)
    assert_parity(d, ref, reff, "sheared 2D bug A2", mag)


def test_bug_a1_a2_blended(device):
    # every reference configuration builds with SUBCELL_FV_BLENDING, the setting under which A2's
    # stale index is node n-1 of the same element
    import torch

    _say("bug_compat A1|A2 with blending, 3D NS")
    m = M.periodic_lattice(3, 3, (6, 3, 2), lengths=[1.0, 0.5, 0.4])
    X = m.node_coords()
    step = np.tanh((X[0] - 0.5) / 0.03)
    rng = np.random.default_rng(4)
    u = cases.prim_to_conserv(1.0 + 0.3 * step + 0.05 * np.sin(2 * np.pi * X[2]),
                              [0.2 * np.cos(2 * np.pi * X[2] + d) for d in range(3)],
                              1.0 + 0.24 * step + 0.02 * rng.uniform(-1, 1, X[0].shape))
    kw = dict(viscous=True, mu=2e-3, Pr=0.71)
    bugs = capi.BUG_A1 | capi.BUG_A2
    op = capi.Operator(m, blending=True, bug_compat=bugs, **kw)
    d = torch.from_numpy(u).to(device)
    du = torch.empty_like(d)
    op.mult(d, du)
    a, _ = op.blend_coefficients(d)
    torch.cuda.synchronize()
    a = a.cpu().numpy()
    out = du.cpu().numpy()
    op.close()
    ref, _, mag = oracle.rhs_blend(m, oracle.phys(bug_compat=bugs, **kw), u, a)
    s = oracle.scaled_error(out, ref, mag)
    print(f"[bug] blended: alpha>0 in {int((a > 0).sum())}/{a.size}; rel.err {relerr(out, ref):.3e}, "
          f"entrywise {s:.2f} eps*mag", flush=True)
    assert (a > 0).any() and s <= C_ROUND


def test_bug_a1_a2_steps(device):
    import torch

    _say("bug_compat A1|A2 LSRK4(5) and SSP-RK3 steps, TGV 4^3 NS")
    m, u, ph = _tgv_ns(4)
    bugs = capi.BUG_A1 | capi.BUG_A2
    oph = oracle.phys(bug_compat=bugs, gamma=ph["gamma"], Pr=ph["Pr"], mu=ph["mu"], viscous=True)
    for scheme, nsteps in ((capi.LSRK45, 3), (capi.SSPRK3, 3)):
        op = capi.Operator(m, bug_compat=bugs, **ph)
        d = torch.from_numpy(u.copy()).to(device)
        op.rk_run(scheme, d, 0.0, 1e-3, nsteps)
        torch.cuda.synchronize()
        out = d.cpu().numpy()
        op.close()
        ref = oracle.rk_run(m, oph, u, scheme, nsteps, 1e-3)
        print(f"[bug] scheme {scheme}, {nsteps} steps: rel.err {relerr(out, ref):.3e}", flush=True)
        assert relerr(out, ref) <= TOL
    # the adjoint differentiates the intended operator only
    op = capi.Operator(m, bug_compat=bugs, **ph)
    d = torch.from_numpy(u).to(device)
    with pytest.raises(capi.PrandtlError, match="UNSUPPORTED"):
        op.vjp(d, torch.ones_like(d), torch.empty_like(d))
    op.close()


@pytest.mark.parametrize("scheme", [capi.SSPRK3, capi.LSRK45])
@pytest.mark.parametrize("viscous", [False, True])
def test_rk_steps(device, scheme, viscous):
    import torch

    _say(f"RK scheme={scheme} viscous={viscous}")
    if viscous:
        m = cases.tgv_mesh(n=4, order=3)
        u = cases.tgv_state(m)
        ph = cases.tgv_physics()
        kw = dict(gamma=ph["gamma"], Pr=ph["Pr"], mu=ph["mu"], viscous=True)
        dt = 2e-3
    else:
        m = cases.vortex_mesh(n=8, order=3)
        u = cases.vortex_state(m)
        kw = {}
        dt = 5e-3
    op = capi.Operator(m, **kw)
    ud = torch.from_numpy(u).to(device)
    t = op.rk_run(scheme, ud, 0.0, dt, 3)
    torch.cuda.synchronize()
    assert abs(t - 3 * dt) < 1e-15
    ref = oracle.rk_run(m, oracle.phys(**kw), u, scheme, 3, dt)
    reff = oracle.rk_run(m, oracle.phys(**kw), u, scheme, 3, dt, fma=True)
    assert_parity(ud.cpu().numpy(), ref, reff, "rk")
    # single steps agree with a run
    ud2 = torch.from_numpy(u).to(device)
    for _ in range(3):
        op.rk_step(scheme, ud2, 0.0, dt)
    torch.cuda.synchronize()
    assert torch.equal(ud, ud2)
    op.close()


def test_full_size_tgv_parity(device):
    """TGV 32^3 p=3 NS (the benchmark workload): RHS against the oracle at full size."""
    import torch

    _say("TGV 32^3 p=3 NS full-size parity vs oracle")
    oracle.set_threads(16)
    m = cases.tgv_mesh(n=32, order=3)
    u = cases.tgv_state(m)
    ph = cases.tgv_physics()
    op = capi.Operator(m, **ph)
    ud = torch.from_numpy(u).to(device)
    du = torch.empty_like(ud)
    op.mult(ud, du)
    torch.cuda.synchronize()
    oph = oracle.phys(gamma=ph["gamma"], Pr=ph["Pr"], mu=ph["mu"], viscous=True)
    ref, _, mag = oracle.rhs_mag(m, oph, u)
    reff, _ = oracle.rhs(m, oph, u, fma=True)
    assert_parity(du.cpu().numpy(), ref, reff, "tgv32 ns", mag)
    op.close()


def test_full_size_tgv_invariants(device):
    """TGV 32^3 p=3 (2.1M nodes): conservation, determinism, free stream."""
    import torch

    _say("TGV 32^3 p=3 NS full-size invariants")
    m = cases.tgv_mesh(n=32, order=3)
    u = cases.tgv_state(m)
    op = capi.Operator(m, **cases.tgv_physics())
    ud = torch.from_numpy(u).to(device)
    d1 = torch.empty_like(ud)
    d2 = torch.empty_like(ud)
    op.mult(ud, d1)
    op.mult(ud, d2)
    torch.cuda.synchronize()
    assert torch.equal(d1, d2), "RHS is not bitwise reproducible"
    W = torch.from_numpy(m.quadrature_weights()).to(device)
    for v in range(5):
        s = (W * d1[v]).sum().item()
        a = (W * d1[v].abs()).sum().item()
        assert abs(s) <= 1e-12 * a, (v, s, a)
    assert torch.isfinite(d1).all()
    lam = op.get_max_char_speed()
    assert abs(lam - cases.max_wave_speed(u)) < 1e-13 * lam
    # free stream on the same mesh
    u0 = np.zeros(m.state_shape)
    u0[:] = cases.prim_to_conserv(np.ones(1), [np.full(1, 0.3), np.full(1, -0.1), np.full(1, 0.2)],
                                  np.full(1, 71.4))[:, :, None]
    ud0 = torch.from_numpy(u0).to(device)
    op.mult(ud0, d1)
    torch.cuda.synchronize()
    assert d1.abs().max().item() < 1e-10
    op.close()


def test_full_size_entropy_conservation(device):
    """Euler + EC surface flux on a jittered 32^3 TGV state: sum J w_i w(u).du = 0."""
    import torch

    _say("TGV 32^3 p=3 Euler EC entropy conservation")
    m = cases.tgv_mesh(n=32, order=3)
    u = cases.tgv_state(m) * (1.0 + 1e-3 * np.random.default_rng(9).uniform(-1, 1, m.state_shape))
    op = capi.Operator(m, surface_flux=capi.SURF_EC)
    ud = torch.from_numpy(u).to(device)
    du = torch.empty_like(ud)
    op.mult(ud, du)
    torch.cuda.synchronize()
    d = du.cpu().numpy()
    rho, vel, p = cases.conserv_to_prim(u)
    s = np.log(p) - 1.4 * np.log(rho)
    b = rho / p
    w = np.stack([(1.4 - s) / 0.4 - 0.5 * b * sum(v * v for v in vel)] + [b * v for v in vel] + [-b])
    W = m.quadrature_weights()
    rate = (W * (w * d).sum(0)).sum()
    scale = (W * np.abs(w * d).sum(0)).sum()
    assert abs(rate) <= 1e-12 * scale
    op.close()


def test_nan_detection(device):
    import torch

    _say("NaN detection in rk_run")
    m = cases.vortex_mesh(n=4, order=3)
    u = cases.vortex_state(m)
    u[0, 3, 5] = np.nan
    op = capi.Operator(m)
    ud = torch.from_numpy(u).to(device)
    with pytest.raises(capi.PrandtlError, match="NAN"):
        op.rk_run(capi.LSRK45, ud, 0.0, 1e-3, 4, nan_check=2)
    op.close()
