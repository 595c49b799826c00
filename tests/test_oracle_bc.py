"""Pins of the oracle's boundary conditions (oracle/prandtl_oracle.cpp, the restatement of the
reference's BdrFaceIntegrator hierarchy).  None restate the oracle's formulas:

* slip wall: p* against the exact solution of the reflected Riemann problem, found here by
  root-finding Toro's shock and rarefaction relations (scipy brentq), for flow into and away
  from the wall, 2D and 3D, arbitrary scaled normals; the flux is (0, p* n, 0), the speed |V|+c;
* symmetry: zero mass and energy flux and no tangential momentum flux;
* outer states against their defining properties: supersonic outflow = the trace; Riemann
  invariants: the far-field state at a matching far field, the trace at supersonic outflow,
  the far field at supersonic inflow, and in the subsonic cases the boundary sound speed
  a_b^2 = gamma p_b / rho_b, the outgoing/incoming invariants and the entropy of the upwind side;
  subsonic outflow keeps rho and V and sets p; subsonic inflow (rho, V) keeps p; total pressure
  and temperature: cp T + V^2/2 = cp Tt and p/pt = (T/Tt)^(gamma/(gamma-1)) with R = cp (gamma-1)/gamma,
  along the inflow angles;
* no-slip walls: the lifting flux's wall state has the wall velocity and beta = rho/p (the
  trace's for adiabatic walls, 1/(R T_wall) for isothermal ones); the adiabatic wall's viscous
  energy flux is V_wall . F_m + q_n |n|;
* the operator: free streams stay steady with every inflow/outflow/far-field BC and with slip,
  symmetry and moving no-slip walls (Euler and Navier-Stokes, 2D and 3D); slip walls conserve
  mass and energy, adiabatic resting walls also under NS.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

import oracle
from prandtl_amd import cases, mesh as M

G = 1.4
EPS = np.finfo(float).eps


def _state(dim, rho, vel, p):
    ke = 0.5 * rho * sum(v * v for v in vel)
    return np.array([rho] + [rho * v for v in vel] + [p / (G - 1) + ke])


def _prim(s):
    rho = s[0]
    vel = s[1:-1] / rho
    p = (G - 1) * (s[-1] - 0.5 * rho * vel @ vel)
    return rho, vel, p


def _normals(dim, rng, k=4):
    out = []
    for _ in range(k):
        n = rng.normal(size=dim)
        out.append(n * rng.uniform(0.3, 3.0) / np.linalg.norm(n))
    return out


def _toro_f(p, rhoK, pK):
    """Toro's pressure function f_K(p) (shock or rarefaction branch)."""
    aK = np.sqrt(G * pK / rhoK)
    if p > pK:
        A = 2 / ((G + 1) * rhoK)
        B = (G - 1) / (G + 1) * pK
        return (p - pK) * np.sqrt(A / (p + B))
    return 2 * aK / (G - 1) * ((p / pK) ** ((G - 1) / (2 * G)) - 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_slip_wall_exact_riemann(dim):
    rng = np.random.default_rng(dim)
    ph = oracle.phys()
    b = oracle.bc(oracle.BC_SLIP_WALL)
    for nor in _normals(dim, rng):
        un = nor / np.linalg.norm(nor)
        for vn in (-0.8, -0.1, 0.0, 0.2, 1.5):
            t = rng.normal(size=dim)
            t -= (t @ un) * un
            vel = vn * un + 0.7 * t
            rho, p = 1.3, 0.9
            s = _state(dim, rho, vel, p)
            f, lam = oracle.bc_inviscid_flux(dim, ph, b, s, nor)
            # mirror Riemann problem: u* = 0, so f_K(p*) = vn
            if vn == 0.0:
                pstar = p
            else:
                pstar = brentq(lambda q: _toro_f(q, rho, p) - vn, 1e-12, 1e3, xtol=1e-15, rtol=1e-15)
            assert abs(f[0]) == 0.0 and abs(f[-1]) == 0.0
            assert np.allclose(f[1:-1], pstar * nor, rtol=1e-12, atol=0)
            assert np.isclose(lam, np.linalg.norm(vel) + np.sqrt(G * p / rho), rtol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("surf", [oracle.SURF_CHANDRASHEKAR, oracle.SURF_EC, oracle.SURF_RUSANOV])
def test_symmetry_flux(dim, surf):
    rng = np.random.default_rng(10 + dim)
    ph = oracle.phys(surface_flux=surf)
    b = oracle.bc(oracle.BC_SYMMETRY)
    for nor in _normals(dim, rng):
        s = _state(dim, 1.1, rng.normal(size=dim) * 0.5, 1.2)
        f, _ = oracle.bc_inviscid_flux(dim, ph, b, s, nor)
        scale = np.abs(f).max()
        un = nor / np.linalg.norm(nor)
        assert abs(f[0]) <= 1e-14 * scale and abs(f[-1]) <= 1e-14 * scale
        fm = f[1:-1]
        assert np.linalg.norm(fm - (fm @ un) * un) <= 1e-13 * scale


def test_outer_states():
    ph = oracle.phys(R_gas=287.0)
    rng = np.random.default_rng(3)
    for dim in (2, 3):
        for nor in _normals(dim, rng, 3):
            un = nor / np.linalg.norm(nor)
            s1 = _state(dim, 1.2, 0.3 * rng.normal(size=dim), 1.1)
            r1, v1, p1 = _prim(s1)
            # supersonic outflow: the trace
            assert np.array_equal(oracle.bc_outer_state(dim, ph, oracle.bc(oracle.BC_SUPERSONIC_OUTFLOW), s1, nor), s1)
            # subsonic outflow: rho and V kept, p set
            g = oracle.bc_outer_state(dim, ph, oracle.bc(oracle.BC_SUBSONIC_OUTFLOW_P, 0.8), s1, nor)
            rg, vg, pg = _prim(g)
            assert np.isclose(rg, r1) and np.allclose(vg, v1, rtol=1e-14) and np.isclose(pg, 0.8, rtol=1e-13)
            # subsonic inflow (rho, V): p kept
            V = [0.4, -0.1, 0.2][:dim]
            g = oracle.bc_outer_state(dim, ph, oracle.bc(oracle.BC_SUBSONIC_INFLOW_RV, 0.9, *V), s1, nor)
            rg, vg, pg = _prim(g)
            assert np.isclose(rg, 0.9) and np.allclose(vg, V, rtol=1e-14) and np.isclose(pg, p1, rtol=1e-13)
            # total pressure / temperature along the inflow angles
            cp, pt, Tt, th, phi = 1004.5, 1.5, 1.3e-3, 25.0, 70.0
            g = oracle.bc_outer_state(dim, ph, oracle.bc(oracle.BC_SUBSONIC_INFLOW_PTTT, pt, Tt, th, phi, cp), s1, nor)
            rg, vg, pg = _prim(g)
            R = cp * (G - 1) / G
            T = pg / (rg * R)
            assert np.isclose(pg, p1, rtol=1e-13)
            assert np.isclose(cp * T + 0.5 * vg @ vg, cp * Tt, rtol=1e-12)
            assert np.isclose(pg / pt, (T / Tt) ** (G / (G - 1)), rtol=1e-11)
            d = np.array([np.cos(np.radians(th)), np.sin(np.radians(th))])
            if dim == 3:
                d = np.array([d[0] * np.sin(np.radians(phi)), d[1] * np.sin(np.radians(phi)), np.cos(np.radians(phi))])
            assert np.allclose(vg / np.linalg.norm(vg), d, rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_riemann_invariant_states(dim):
    ph = oracle.phys()
    rng = np.random.default_rng(20 + dim)
    for nor in _normals(dim, rng, 3):
        un = nor / np.linalg.norm(nor)
        t = rng.normal(size=dim)
        t -= (t @ un) * un
        t /= np.linalg.norm(t)
        far = (1.0, 0.5 * un + 0.2 * t, 1.0)   # subsonic
        b = oracle.bc(oracle.BC_RIEMANN_INVARIANT, far[0], *far[1], far[2])
        # a matching far field: the outer state is the far field
        s = _state(dim, *far)
        assert np.allclose(oracle.bc_outer_state(dim, ph, b, s, nor), s, rtol=1e-13)
        a_far = np.sqrt(G * far[2] / far[0])
        for vn_i, kind in ((0.3, "out"), (-0.3, "in")):
            s1 = _state(dim, 1.1, vn_i * un + 0.1 * t, 0.95)
            r1, v1, p1 = _prim(s1)
            a1 = np.sqrt(G * p1 / r1)
            g = oracle.bc_outer_state(dim, ph, b, s1, nor)
            rb, vb, pb = _prim(g)
            ab = np.sqrt(G * pb / rb)
            Rp = v1 @ un + 2 * a1 / (G - 1)              # outgoing invariant from inside
            Rm = far[1] @ un - 2 * a_far / (G - 1)       # incoming invariant from the far field
            assert np.isclose(vb @ un, 0.5 * (Rp + Rm), rtol=1e-12)
            assert np.isclose(ab, 0.25 * (G - 1) * (Rp - Rm), rtol=1e-12)
            up = (r1, p1, v1) if kind == "out" else (far[0], far[2], far[1])
            assert np.isclose(pb / rb ** G, up[1] / up[0] ** G, rtol=1e-12)              # upwind entropy
            tang = vb - (vb @ un) * un
            assert np.allclose(tang, up[2] - (up[2] @ un) * un, atol=1e-13)              # upwind tangential V
        # supersonic outflow and inflow
        s1 = _state(dim, 1.0, 2.5 * un, 1.0)
        bs = oracle.bc(oracle.BC_RIEMANN_INVARIANT, 1.0, *(2.5 * un), 1.0)
        assert np.allclose(oracle.bc_outer_state(dim, ph, bs, s1, nor), s1, rtol=1e-12)
        s_in = _state(dim, 0.9, -2.2 * un + 0.1 * t, 0.8)
        bi = oracle.bc(oracle.BC_RIEMANN_INVARIANT, 0.9, *(-2.2 * un + 0.1 * t), 0.8)
        s1 = _state(dim, 1.3, -2.0 * un, 1.1)
        assert np.allclose(oracle.bc_outer_state(dim, ph, bi, s1, nor), s_in, rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_noslip_wall_terms(dim):
    R = 287.0
    ph = oracle.phys(viscous=True, mu=1e-2, Pr=0.71, R_gas=R)
    rng = np.random.default_rng(30 + dim)
    Vw = np.array([0.3, -0.2, 0.1][:dim])
    for nor in _normals(dim, rng, 3):
        s1 = _state(dim, 1.2, 0.4 * rng.normal(size=dim), 1.1)
        r1, v1, p1 = _prim(s1)
        w1 = oracle.conserv2entropy(dim, ph, s1)
        for kind, Tw in ((oracle.BC_NOSLIP_ADIABATIC, None), (oracle.BC_NOSLIP_ISOTHERMAL, 3.1e-3)):
            b = oracle.bc(kind, *Vw, 0.05 if Tw is None else Tw)
            ww = w1 + oracle.bc_lifting_flux(dim, ph, b, w1, nor)
            beta = -ww[-1]
            assert np.isclose(beta, r1 / p1 if Tw is None else 1 / (R * Tw), rtol=1e-13)
            assert np.allclose(ww[1:-1] / beta, Vw, rtol=1e-13)
            assert ww[0] == w1[0]
            dq = rng.normal(size=(dim, dim + 2))
            fv = oracle.bc_viscous_flux(dim, ph, b, s1, dq, nor)
            flux = oracle.viscous_flux(dim, ph, s1, *dq)           # (neq, dim)
            ref = flux @ nor
            assert np.allclose(fv[:-1], ref[:-1], rtol=1e-13, atol=1e-15)
            if Tw is None:
                assert np.isclose(fv[-1], Vw @ ref[1:-1] + 0.05 * np.linalg.norm(nor), rtol=1e-13)
            else:
                assert np.isclose(fv[-1], ref[-1], rtol=1e-13)


def _channel(dim, order, bcs_lo, bcs_hi, d=0, n=None):
    n = n or ((5, 3) if dim == 2 else (4, 3, 2))
    per = [True] * dim
    per[d] = False
    return M.periodic_lattice(dim, order, n, lengths=[1.0, 0.7, 0.6][:dim], periodic=per,
                              boundary={(d, 0): bcs_lo, (d, 1): bcs_hi})


def _uniform(m, rho, vel, p):
    one = np.ones((m.num_elements, m.nodes_per_element))
    return cases.prim_to_conserv(rho * one, [v * one for v in vel], p * one)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("viscous", [False, True])
def test_free_stream_with_boundaries(dim, viscous):
    ph = oracle.phys(viscous=viscous, mu=1e-2 if viscous else 0.0, Pr=0.71, R_gas=287.0)
    rho, p = 1.1, 0.9
    a = np.sqrt(G * p / rho)
    order = 3
    for mach in (0.4, 1.8):
        vel = [mach * a, 0.2 * a, 0.1 * a][:dim] if False else [mach * a] + [0.0] * (dim - 1)
        s = _state(dim, rho, vel, p)
        cp = G / (G - 1) * 287.0
        T = p / (rho * 287.0)
        Tt = T + 0.5 * (vel[0] ** 2) / cp
        pt = p * (Tt / T) ** (G / (G - 1))
        pairs = [(oracle.bc(oracle.BC_SPECIFIED_STATE, *s), oracle.bc(oracle.BC_SPECIFIED_STATE, *s)),
                 (oracle.bc(oracle.BC_RIEMANN_INVARIANT, rho, *vel, p), oracle.bc(oracle.BC_RIEMANN_INVARIANT, rho, *vel, p))]
        if mach > 1:
            pairs.append((oracle.bc(oracle.BC_SUPERSONIC_INFLOW, *s), oracle.bc(oracle.BC_SUPERSONIC_OUTFLOW)))
        else:
            pairs.append((oracle.bc(oracle.BC_SUBSONIC_INFLOW_RV, rho, *vel), oracle.bc(oracle.BC_SUBSONIC_OUTFLOW_P, p)))
            pairs.append((oracle.bc(oracle.BC_SUBSONIC_INFLOW_PTTT, pt, Tt, 0.0, 90.0, cp),
                          oracle.bc(oracle.BC_SUBSONIC_OUTFLOW_P, p)))
        for lo, hi in pairs:
            m = _channel(dim, order, 0, 1)
            u = _uniform(m, rho, vel, p)
            du, _, mag = oracle.rhs_ex(m, ph, u, [lo, hi])
            assert np.abs(du).max() <= 64 * EPS * mag.max(), (lo.type, hi.type, np.abs(du).max())
    # walls along the flow: slip and symmetry (Euler), moving no-slip walls (NS)
    vel = [0.5 * a] + [0.0] * (dim - 1)
    walls = [oracle.bc(oracle.BC_SLIP_WALL), oracle.bc(oracle.BC_SYMMETRY)] if not viscous else \
        [oracle.bc(oracle.BC_NOSLIP_ADIABATIC, *vel, 0.0),
         oracle.bc(oracle.BC_NOSLIP_ISOTHERMAL, *vel, p / (rho * 287.0))]
    for w in walls:
        m = _channel(dim, order, 0, 0, d=1)
        u = _uniform(m, rho, vel, p)
        du, _, mag = oracle.rhs_ex(m, ph, u, [w])
        assert np.abs(du).max() <= 64 * EPS * mag.max(), (w.type, np.abs(du).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_walls_conserve(dim):
    rng = np.random.default_rng(40 + dim)
    for viscous, w in ((False, oracle.bc(oracle.BC_SLIP_WALL)),
                       (True, oracle.bc(oracle.BC_NOSLIP_ADIABATIC, *([0.0] * dim), 0.0))):
        ph = oracle.phys(viscous=viscous, mu=5e-3 if viscous else 0.0, Pr=0.71)
        m = _channel(dim, 3, 0, 0, d=1)
        X = m.node_coords()
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * X[0]) * np.cos(3 * X[1])
        vel = [0.3 * np.cos(2 * np.pi * X[0] + d) * np.sin(np.pi * X[1] / 0.7) for d in range(dim)]
        pr = 1.0 + 0.1 * np.cos(2 * np.pi * X[0]) + 0.01 * rng.uniform(-1, 1, X[0].shape)
        u = cases.prim_to_conserv(rho, vel, pr)
        du, _, mag = oracle.rhs_ex(m, ph, u, [w])
        W = m.quadrature_weights()
        for v in (0, dim + 1):
            assert abs((W * du[v]).sum()) <= 1e-12 * (W * mag[v]).sum(), v


@pytest.mark.parametrize("dim", [2, 3])
def test_cut_periodic_face(dim):
    """A periodic face cut open, each side given the other side's trace as a specified state:
    the Euler RHS and the NS lifted gradients are those of the periodic mesh (the traces are
    uniform across the face, the surface flux is antisymmetric in its states and normal, and the
    default lifting flux is the interior face's 1/2 (w2 - w1))."""
    n = (2,) + (1,) * (dim - 1)
    Ls = [1.0, 0.6, 0.5][:dim]
    per = M.periodic_lattice(dim, 3, n, lengths=Ls)
    cut = M.periodic_lattice(dim, 3, n, lengths=Ls, periodic=(False,) + (True,) * (dim - 1),
                             boundary={(0, 0): 0, (0, 1): 1})
    X = per.node_coords()
    x = X[0]
    rho = 1.0 + 0.3 * np.sin(2.5 * x) + 0.2 * (x > 0.5)
    vel = [0.4 * np.cos(3 * x + d) for d in range(dim)]
    p = 1.0 + 0.2 * np.cos(2 * x)
    u = cases.prim_to_conserv(rho, vel, p)
    np1 = 4
    left = [i for i in range(per.nodes_per_element) if i % np1 == 0]
    right = [i for i in range(per.nodes_per_element) if i % np1 == np1 - 1]
    qA = u[:, 0, left[0]]          # element 0's trace at x = 0
    qB = u[:, 1, right[0]]         # element 1's trace at x = L
    bcs = [oracle.bc(oracle.BC_SPECIFIED_STATE, *qB), oracle.bc(oracle.BC_SPECIFIED_STATE, *qA)]
    for surf in (oracle.SURF_CHANDRASHEKAR, oracle.SURF_RUSANOV):
        ph = oracle.phys(surface_flux=surf)
        ref, lam_ref, mag = oracle.rhs_ex(per, ph, u)
        got, lam, _ = oracle.rhs_ex(cut, ph, u, bcs)
        assert np.abs(got - ref).max() <= 64 * EPS * mag.max()
    phv = oracle.phys(viscous=True, mu=1e-2, Pr=0.71)
    _, _, _, gref, gmag = oracle.rhs_ex(per, phv, u, want_grad=True)
    _, _, _, gcut, _ = oracle.rhs_ex(cut, phv, u, bcs, want_grad=True)
    assert np.abs(gcut - gref).max() <= 64 * EPS * gmag.max()


def _wall_case(ny, case):
    """Exact steady NS states between walls y = 0 and y = H (x periodic), R = 1, constant mu:
    heat conduction (rest, T linear from T0 to T1, uniform p) or Couette flow (u = U y/H,
    T = Tw + mu U^2/(2k) (y/H)(1 - y/H), k = mu cp / Pr, uniform p), and their wall BCs."""
    H, mu, Pr = 1.0, 0.05, 0.71
    cp = G / (G - 1)
    k = mu * cp / Pr
    m = M.periodic_lattice(2, 3, (1, ny), lengths=[0.5, H], periodic=(True, False),
                           boundary={(1, 0): 0, (1, 1): 1})
    y = m.node_coords()[1] / H
    p = 1.0
    if case == "conduction":
        T0, T1 = 1.0, 1.6
        T = T0 + (T1 - T0) * y
        vel = [0 * y, 0 * y]
        bcs = [oracle.bc(oracle.BC_NOSLIP_ISOTHERMAL, 0.0, 0.0, T0), oracle.bc(oracle.BC_NOSLIP_ISOTHERMAL, 0.0, 0.0, T1)]
    else:
        U, Tw = 0.8, 1.0
        T = Tw + mu * U * U / (2 * k) * y * (1 - y)
        vel = [U * y, 0 * y]
        bcs = [oracle.bc(oracle.BC_NOSLIP_ISOTHERMAL, 0.0, 0.0, Tw), oracle.bc(oracle.BC_NOSLIP_ISOTHERMAL, U, 0.0, Tw)]
    u = cases.prim_to_conserv(p / T, vel, p + 0 * y)
    ph = oracle.phys(viscous=True, mu=mu, Pr=Pr, R_gas=1.0)
    return m, ph, u, bcs


@pytest.mark.parametrize("case", ["conduction", "couette"])
def test_wall_steady_states_converge(case):
    """The residual of an exact steady state with isothermal no-slip walls falls under
    refinement; a wrong sign or normal in the wall's viscous or lifting terms leaves an O(1/h)
    residual at the wall."""
    res = []
    for ny in (2, 4, 8):
        m, ph, u, bcs = _wall_case(ny, case)
        du, _, _ = oracle.rhs_ex(m, ph, u, bcs)
        res.append(np.abs(du).max())
    # measured: second order (a factor 3.2-3.9 per halving of h), the rate of BR1 at these walls
    assert res[0] < 2e-2 and res[1] < res[0] / 3 and res[2] < res[1] / 3, res
