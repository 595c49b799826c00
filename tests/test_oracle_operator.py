"""Operator-level pins for the oracle (CPU): invariants the semi-discretisation must keep.

* free-stream preservation on Cartesian and sheared affine meshes (Euler and NS);
* discrete conservation sum_e sum_i J w_i du_i = 0 per variable (periodic);
* entropy conservation with the entropy-conservative surface flux and entropy
  dissipation with the Chandrashekar and Rusanov fluxes, on discontinuous states;
* BR1 primitive gradients of the Taylor-Green state converge to the analytic gradients;
* the Taylor-Green kinetic-energy dissipation rate at t = 0 is -3/(4 Re) to within
  the compressibility/discretisation error;
* RK schemes: observed temporal order 3 (SSP-RK3) and 4 (LSRK4(5));
* bug_compat switches change exactly the terms they name.
"""
import math

import numpy as np
import pytest

import oracle
from prandtl_amd import cases, mesh as M


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


def entropy_vars(u, gamma=1.4):
    rho, vel, p = cases.conserv_to_prim(u, gamma)
    s = np.log(p) - gamma * np.log(rho)
    b = rho / p
    v2 = sum(v * v for v in vel)
    return np.stack([(gamma - s) / (gamma - 1) - 0.5 * b * v2] + [b * v for v in vel] + [-b])


@pytest.mark.parametrize("dim,order", [(2, 3), (3, 2), (3, 3)])
@pytest.mark.parametrize("viscous", [False, True])
def test_free_stream(dim, order, viscous):
    m = sheared_mesh(dim, order, (3,) * dim)
    u = np.zeros(m.state_shape)
    vel = [0.3, -0.2, 0.1][:dim]
    u[:] = cases.prim_to_conserv(np.ones(1), [np.full(1, v) for v in vel], np.full(1, 1.3))[:, :, None]
    ph = oracle.phys(viscous=viscous, mu=0.01, Pr=0.7)
    d, _ = oracle.rhs(m, ph, u)
    assert np.abs(d).max() < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [oracle.SURF_CHANDRASHEKAR, oracle.SURF_EC, oracle.SURF_RUSANOV])
@pytest.mark.parametrize("viscous", [False, True])
def test_conservation(dim, kind, viscous):
    m = sheared_mesh(dim, 3, (3,) * dim)
    u = perturbed_state(m, 1)
    d, _ = oracle.rhs(m, oracle.phys(surface_flux=kind, viscous=viscous, mu=0.02, Pr=0.7), u)
    W = m.quadrature_weights()
    for v in range(m.neq):
        assert abs((W * d[v]).sum()) <= 1e-13 * (W * np.abs(d[v])).sum()


@pytest.mark.parametrize("dim", [2, 3])
def test_entropy_conservation_and_stability(dim):
    m = sheared_mesh(dim, 3, (3,) * dim)
    u = perturbed_state(m, 2)
    W = m.quadrature_weights()
    w = entropy_vars(u)
    rates = {}
    for kind in (oracle.SURF_EC, oracle.SURF_CHANDRASHEKAR, oracle.SURF_RUSANOV):
        d, _ = oracle.rhs(m, oracle.phys(surface_flux=kind), u)
        rates[kind] = (W * (w * d).sum(0)).sum()
        scale = (W * np.abs(w * d).sum(0)).sum()
        if kind == oracle.SURF_EC:
            assert abs(rates[kind]) <= 1e-13 * scale
        else:
            assert rates[kind] < -1e-6 * scale
    # viscous terms dissipate entropy as well
    d, _ = oracle.rhs(m, oracle.phys(surface_flux=oracle.SURF_EC, viscous=True, mu=0.05, Pr=0.7), u)
    assert (W * (w * d).sum(0)).sum() < 0.0


def test_tgv_gradients_converge():
    errs = []
    for n in (4, 8):
        m = cases.tgv_mesh(n=n, order=3)
        u = cases.tgv_state(m)
        _, _, g = oracle.rhs(m, oracle.phys(viscous=True, mu=1 / 1600, Pr=0.71), u, want_grad=True)
        x, y, z = m.node_coords()
        dudx = np.cos(x) * np.cos(y) * np.cos(z)
        dvdy = -np.cos(x) * np.cos(y) * np.cos(z)
        dudy = -np.sin(x) * np.sin(y) * np.cos(z)
        dpdx = -(1 / 16) * 2 * np.sin(2 * x) * (np.cos(2 * z) + 2)
        ev = max(np.abs(g[0][1] - dudx).max(), np.abs(g[1][2] - dvdy).max(), np.abs(g[1][1] - dudy).max())
        ep = np.abs(g[0][4] - dpdx).max()
        errs.append((ev, ep))
        assert np.abs(g[2][3]).max() < 1e-12 * max(1.0, np.abs(g).max())  # dw/dz = 0
    # velocity gradients: order ~p; the pressure gradient (amplified by 1/(gamma Ma^2) through
    # the entropy variables) is still pre-asymptotic at these sizes
    assert errs[1][0] < errs[0][0] / 5.0, errs
    assert errs[1][1] < errs[0][1] / 2.5, errs


def test_tgv_dissipation_rate():
    m = cases.tgv_mesh(n=6, order=3)
    u = cases.tgv_state(m)
    Re = 1600.0
    d, lam = oracle.rhs(m, oracle.phys(viscous=True, mu=1 / Re, Pr=0.71), u)
    W = m.quadrature_weights()
    vel = [u[1 + k] / u[0] for k in range(3)]
    ke_rate = sum(vel[k] * d[1 + k] for k in range(3)) - 0.5 * sum(v * v for v in vel) * d[0]
    rate = (W * ke_rate).sum() / W.sum()
    assert abs(cases.kinetic_energy(m, u) - 0.125) < 2e-3
    # inviscid part conserves kinetic energy only up to pressure work and discretisation
    d0, _ = oracle.rhs(m, oracle.phys(viscous=False), u)
    rate0 = (W * (sum(vel[k] * d0[1 + k] for k in range(3)) - 0.5 * sum(v * v for v in vel) * d0[0])).sum() / W.sum()
    assert abs((rate - rate0) - (-3.0 / (4.0 * Re))) < 0.03 * 3.0 / (4.0 * Re)
    assert abs(lam - 11.0) < 0.2  # |V| + c with c = 1/Ma = 10


@pytest.mark.slow
@pytest.mark.parametrize("scheme,order", [(0, 3), (1, 4)])
def test_rk_temporal_order(scheme, order):
    m = cases.vortex_mesh(n=4, order=3)
    u0 = cases.vortex_state(m)
    ph = oracle.phys()
    T = 0.4
    ref = oracle.rk_run(m, ph, u0, scheme, 64, T / 64)
    errs = []
    for nst in (4, 8):
        uu = oracle.rk_run(m, ph, u0, scheme, nst, T / nst)
        errs.append(np.abs(uu - ref).max())
    rate = math.log2(errs[0] / errs[1])
    assert rate > order - 0.35, (errs, rate)


def test_bug_compat_switches():
    m = cases.tgv_mesh(n=2, order=3)
    u = perturbed_state(m, 5)
    W = m.quadrature_weights()
    base = oracle.phys(viscous=True, mu=0.01, Pr=0.71)
    d0, _ = oracle.rhs(m, base, u)
    for bug in (oracle.BUG_A1, oracle.BUG_A2, oracle.BUG_A3):
        ph = oracle.phys(viscous=True, mu=0.01, Pr=0.71, bug_compat=bug)
        d, _ = oracle.rhs(m, ph, u)
        diff = np.abs(d - d0)
        assert diff.max() > 1e-8
        if bug == oracle.BUG_A3:
            # only the energy equation changes
            assert diff[:4].max() == 0.0
        if bug == oracle.BUG_A1:
            # the wrong zeta-direction partner fluxes break discrete conservation
            assert max(abs((W * d[v]).sum()) for v in range(5)) > 1e-6
    # A1 and A2 are Navier-Stokes-only paths: the Euler operator ignores them
    de, _ = oracle.rhs(m, oracle.phys(), u)
    de2, _ = oracle.rhs(m, oracle.phys(bug_compat=oracle.BUG_A1 | oracle.BUG_A2), u)
    assert np.array_equal(de, de2)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("viscous", [False, True])
def test_rounding_scale_bounds_build_spread(dim, viscous):
    """mag bounds the difference between two compilations of the same oracle (with and
    without FMA contraction): the entrywise yardstick the GPU parity tests use."""
    m = sheared_mesh(dim, 3, (3,) * dim)
    u = perturbed_state(m, 8)
    ph = oracle.phys(viscous=viscous, mu=0.02, Pr=0.7)
    r = oracle.rhs_mag(m, ph, u, want_grad=viscous)
    rf = oracle.rhs(m, ph, u, want_grad=viscous, fma=True)
    assert np.array_equal(r[0], oracle.rhs(m, ph, u)[0])
    assert np.all(r[2] >= np.abs(r[0]) * (1 - 1e-12))
    assert oracle.scaled_error(rf[0], r[0], r[2]) < 16.0
    if viscous:
        assert oracle.scaled_error(rf[2], r[3], r[4]) < 16.0
    # a planted mistake in a single term is far outside the bound
    bad = r[0].copy()
    bad[1, 0, 0] += 1e-9 * max(1.0, abs(bad[1, 0, 0]))
    assert oracle.scaled_error(bad, r[0], r[2]) > 100.0


@pytest.mark.parametrize("scheme,order", [(2, 1), (3, 2), (4, 4)])
def test_explicit_rk_temporal_order(scheme, order):
    """MFEM's ForwardEuler (1), RK2 (2) and RK4 (4): observed temporal order on the vortex."""
    m = cases.vortex_mesh(n=4, order=3)
    u0 = cases.vortex_state(m)
    ph = oracle.phys()
    T = 0.4
    ref, _ = oracle.rk_run_ex(m, ph, u0, 4, 128, T / 128)
    errs = []
    base = {1: 32, 2: 16, 4: 8}[order]
    for nst in (base, 2 * base):
        uu, _ = oracle.rk_run_ex(m, ph, u0, scheme, nst, T / nst)
        errs.append(np.abs(uu - ref).max())
    rate = math.log2(errs[0] / errs[1])
    assert rate > order - 0.35, (errs, rate)


def test_explicit_rk_stage_structure():
    """The RK2 and RK4 steps equal the Butcher-tableau steps written out with the oracle's RHS
    (numpy): RK2 with a = 2/3 (b = 3/4) and the classic RK4 tableau."""
    m = cases.vortex_mesh(n=3, order=2)
    u0 = cases.vortex_state(m)
    ph = oracle.phys()
    dt = 0.02
    f = lambda x: oracle.rhs_ex(m, ph, x)[0]
    k1 = f(u0)
    k2 = f(u0 + (2.0 / 3.0) * dt * k1)
    rk2 = u0 + dt * (0.25 * k1 + 0.75 * k2)
    out, _ = oracle.rk_run_ex(m, ph, u0, 3, 1, dt)
    assert np.abs(out - rk2).max() < 1e-13 * np.abs(u0).max()
    k1 = f(u0)
    k2 = f(u0 + 0.5 * dt * k1)
    k3 = f(u0 + 0.5 * dt * k2)
    k4 = f(u0 + dt * k3)
    rk4 = u0 + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6.0
    out, _ = oracle.rk_run_ex(m, ph, u0, 4, 1, dt)
    assert np.abs(out - rk4).max() < 1e-13 * np.abs(u0).max()
    out, _ = oracle.rk_run_ex(m, ph, u0, 2, 1, dt)
    assert np.abs(out - (u0 + dt * k1)).max() < 1e-14 * np.abs(u0).max()


def test_lattice_line_values_interpolates():
    """cases.lattice_line_values (the Ghia comparison's sampler) reproduces a cubic field exactly
    on a node line of a p=3 lattice."""
    m = M.periodic_lattice(2, 3, (4, 5), lengths=[2.0, 2.5], periodic=(False, False),
                           boundary={(0, 0): 0, (0, 1): 0, (1, 0): 0, (1, 1): 0})
    X = m.node_coords()
    f = lambda x, y: 1.0 + 0.3 * y - 0.2 * y ** 2 + 0.05 * y ** 3 + 0.1 * x
    ys = np.linspace(0.0, 2.5, 23)
    got = cases.lattice_line_values(m, f(X[0], X[1]), 1.0, ys)
    assert np.abs(got - f(1.0, ys)).max() < 1e-13
