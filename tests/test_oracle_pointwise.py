"""Pins for the oracle's pointwise functions (CPU).

Each function is checked against something the mathematics fixes, chosen so that a
dropped term, wrong sign or swapped index fails:
* GLL rule: numpy's Legendre roots and exact quadrature of x^k, k <= 2N-1.
* D, D-hat, D-hat-2: polynomial exactness, the SBP identity W D + D^T W = B.
* log mean: closed form in extended precision, known answers, branch agreement.
* two-point flux: consistency with the Euler flux (independent numpy), symmetry, and
  Tadmor's entropy-conservation identity with independent entropy variables.
* surface fluxes: consistency, entropy stability, the Rusanov definition, and the
  exact size of the reference's 3D dissipation typo under bug_compat.
* entropy variables: dU/du by finite differences of U = -rho s/(gamma-1); round trip.
* gradient chain rule: finite-difference Jacobians of the forward maps.
* viscous flux: rigid rotation and dilatation carry no stress (Stokes), pure shear,
  heat flux against finite differences of p/rho, Sutherland at T0.
"""
import math

import numpy as np
import pytest

import oracle
from tests import numpy_ref as R

G = 1.4


@pytest.mark.parametrize("np_", [2, 3, 4, 5, 6, 7])
def test_gll_rule(np_):
    from numpy.polynomial import legendre as L

    x, w = oracle.gll(np_)
    N = np_ - 1
    c = np.zeros(N + 1)
    c[N] = 1
    inner = np.sort(L.legroots(L.legder(c))) if N > 1 else np.array([])
    expect = 0.5 * (np.concatenate([[-1.0], inner, [1.0]]) + 1.0)
    assert np.allclose(x, expect, rtol=0, atol=1e-15)
    for k in range(2 * N):
        assert abs(w @ x**k - 1.0 / (k + 1)) < 1e-14
    assert abs(w[0] - 1.0 / (N * (N + 1))) < 1e-16


def test_gll_known_p3_p4():
    x, w = oracle.gll(4)
    assert np.allclose(x, [0, (1 - 1 / math.sqrt(5)) / 2, (1 + 1 / math.sqrt(5)) / 2, 1], atol=1e-16)
    assert np.allclose(w, [1 / 12, 5 / 12, 5 / 12, 1 / 12], atol=1e-16)
    x, w = oracle.gll(5)
    assert np.allclose(x, [0, (1 - math.sqrt(3 / 7)) / 2, 0.5, (1 + math.sqrt(3 / 7)) / 2, 1], atol=1e-16)
    assert np.allclose(w, [1 / 20, 49 / 180, 16 / 45, 49 / 180, 1 / 20], atol=1e-16)


@pytest.mark.parametrize("np_", [2, 3, 4, 5, 6])
def test_dmatrices(np_):
    x, w = oracle.gll(np_)
    D, Dh, Dh2 = oracle.dmatrices(np_)
    for k in range(np_):
        d = D @ x**k
        expect = k * x ** (k - 1) if k > 0 else np.zeros(np_)
        assert np.allclose(d, expect, atol=1e-12)
    W = np.diag(w)
    B = np.zeros((np_, np_))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.allclose(W @ D + D.T @ W, B, atol=1e-13)
    assert abs(D[0, 0] + 1.0 / (2 * w[0])) < 1e-12
    E = np.zeros((np_, np_))
    E[0, 0], E[-1, -1] = 1.0 / w[0], -1.0 / w[-1]
    assert np.allclose(Dh, D + E, atol=1e-13)
    assert np.allclose(Dh2, 2 * D + E, atol=1e-13)
    # D-hat-2 has zero diagonal (the inert self-flux entries of the reference)
    assert np.allclose(np.diag(Dh2), 0.0, atol=1e-12)


def test_logmean():
    assert oracle.logmean(2.5, 2.5) == 2.5
    assert abs(oracle.logmean(1.0, math.e) - (math.e - 1.0)) < 1e-15
    rng = np.random.default_rng(3)
    for _ in range(2000):
        x = rng.uniform(0.01, 10)
        y = x * math.exp(rng.uniform(-3, 3) * 10 ** rng.uniform(-6, 0))
        if x == y:
            continue
        X, Y = np.longdouble(x), np.longdouble(y)
        ref = (Y - X) / np.log1p((Y - X) / X)  # y - x is exact in extended precision
        # the reference's log branch rounds xi = y/x before the log: relative error
        # ~ eps / |ln xi| is inherent to the formula being restated
        tol = 4.5e-16 * (2.0 + 1.0 / abs(math.log(y / x)))
        assert abs(oracle.logmean(x, y) - float(ref)) <= tol * float(ref)
    # symmetry and both branches near u = 1e-4: (xi-1)^2/(xi+1)^2 = 1e-4 at xi = 1.0202...
    for xi in np.linspace(1.019, 1.0215, 50):
        ref = (np.longdouble(xi) - 1) / np.log1p(np.longdouble(xi) - 1)
        assert abs(oracle.logmean(1.0, xi) - float(ref)) < 1.2e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_volume_flux_consistency_symmetry_ec(dim):
    rng = np.random.default_rng(10 + dim)
    ph = oracle.phys(gamma=G)
    for _ in range(200):
        a = R.random_state(rng, dim)
        b = R.random_state(rng, dim)
        n = rng.normal(size=dim)
        F, lam = oracle.volume_flux(dim, ph, a, a, n, n)
        assert np.allclose(F, R.euler_flux_n(a, n, G), rtol=1e-13, atol=1e-13)
        Fab, lab = oracle.volume_flux(dim, ph, a, b, n, n)
        Fba, lba = oracle.volume_flux(dim, ph, b, a, n, n)
        assert np.allclose(Fab, Fba, rtol=1e-13, atol=1e-13)
        assert lab == lba
        # Tadmor: (w_b - w_a) . F# = psi_b - psi_a, psi = rho v.n
        wa, wb = R.entropy_vars(a, G), R.entropy_vars(b, G)
        psi = lambda u: u[1:1 + dim] @ n
        lhs = (wb - wa) @ Fab
        assert abs(lhs - (psi(b) - psi(a))) < 1e-12 * (1 + np.abs(Fab).max())
        # wave speed
        ra, va, pa = R.prim(a, G)
        rb, vb, pb = R.prim(b, G)
        lam_ref = max(np.linalg.norm(va) + math.sqrt(G * pa / ra), np.linalg.norm(vb) + math.sqrt(G * pb / rb))
        assert abs(lab - lam_ref) < 1e-14 * lam_ref
    # the metric enters as the mean of the two metrics
    a, b = R.random_state(rng, dim), R.random_state(rng, dim)
    m1, m2 = rng.normal(size=dim), rng.normal(size=dim)
    F12, _ = oracle.volume_flux(dim, ph, a, b, m1, m2)
    Fm, _ = oracle.volume_flux(dim, ph, a, b, 0.5 * (m1 + m2), 0.5 * (m1 + m2))
    assert np.allclose(F12, Fm, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [oracle.SURF_CHANDRASHEKAR, oracle.SURF_EC, oracle.SURF_RUSANOV])
def test_surface_flux_consistency_and_stability(dim, kind):
    rng = np.random.default_rng(100 + dim + 7 * kind)
    ph = oracle.phys(gamma=G, surface_flux=kind)
    for _ in range(200):
        a = R.random_state(rng, dim)
        b = R.random_state(rng, dim)
        n = rng.normal(size=dim)
        f, lam = oracle.face_flux(dim, ph, a, a, n)
        assert np.allclose(f, R.euler_flux_n(a, n, G), rtol=1e-13, atol=1e-13)
        fab, _ = oracle.face_flux(dim, ph, a, b, n)
        Fec, _ = oracle.volume_flux(dim, ph, a, b, n, n)
        dw = R.entropy_vars(b, G) - R.entropy_vars(a, G)
        prod = dw @ (fab - Fec)
        if kind == oracle.SURF_EC:
            assert np.allclose(fab, Fec, rtol=1e-14, atol=1e-14)
        else:
            assert prod <= 1e-12 * (1 + np.abs(fab).max())
        # conservation across the face: f(a, b, n) = -f(b, a, -n)
        fba, _ = oracle.face_flux(dim, ph, b, a, -n)
        assert np.allclose(fab, -fba, rtol=1e-12, atol=1e-12)


def test_rusanov_definition():
    rng = np.random.default_rng(7)
    ph = oracle.phys(gamma=G, surface_flux=oracle.SURF_RUSANOV)
    for dim in (2, 3):
        a, b = R.random_state(rng, dim), R.random_state(rng, dim)
        n = rng.normal(size=dim)
        f, lam = oracle.face_flux(dim, ph, a, b, n)
        ra, va, pa = R.prim(a, G)
        rb, vb, pb = R.prim(b, G)
        l = max(np.linalg.norm(va) + math.sqrt(G * pa / ra), np.linalg.norm(vb) + math.sqrt(G * pb / rb))
        ref = 0.5 * (R.euler_flux_n(a, n, G) + R.euler_flux_n(b, n, G)) - 0.5 * l * np.linalg.norm(n) * (b - a)
        assert np.allclose(f, ref, rtol=1e-13, atol=1e-13)


def test_bug_a3_typo_size():
    """Under bug_compat the 3D dissipation uses rho_mean*dw + w_mean: for equal states the
    energy flux then differs from F(u).n by exactly -1/2 lambda |n| w."""
    rng = np.random.default_rng(11)
    ph = oracle.phys(gamma=G, bug_compat=oracle.BUG_A3)
    for _ in range(20):
        a = R.random_state(rng, 3)
        n = rng.normal(size=3)
        f, lam = oracle.face_flux(3, ph, a, a, n)
        ref = R.euler_flux_n(a, n, G)
        w = a[3] / a[0]
        assert np.allclose(f[:4], ref[:4], rtol=1e-13, atol=1e-13)
        assert abs((f[4] - ref[4]) - (-0.5 * lam * np.linalg.norm(n) * w)) < 1e-12 * (1 + abs(ref[4]))
    # 2D has no such term
    ph2 = oracle.phys(gamma=G, bug_compat=oracle.BUG_A3)
    a = R.random_state(rng, 2)
    f, _ = oracle.face_flux(2, ph2, a, a, np.array([0.3, -0.7]))
    assert np.allclose(f, R.euler_flux_n(a, np.array([0.3, -0.7]), G), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_entropy_variables(dim):
    rng = np.random.default_rng(20 + dim)
    ph = oracle.phys(gamma=G)
    for _ in range(50):
        u = R.random_state(rng, dim)
        w = oracle.conserv2entropy(dim, ph, u)
        assert np.allclose(w, R.entropy_vars_fd(u, G), rtol=1e-7, atol=1e-7)
        u2 = oracle.entropy2conserv(dim, ph, w)
        assert np.allclose(u2, u, rtol=1e-13, atol=1e-13)


def _jac(f, x, h=1e-6):
    x = np.asarray(x, dtype=np.float64)
    J = np.zeros((len(f(x)), len(x)))
    for i in range(len(x)):
        e = np.zeros_like(x)
        e[i] = h * max(1.0, abs(x[i]))
        J[:, i] = (f(x + e) - f(x - e)) / (2 * e[i])
    return J


@pytest.mark.parametrize("dim", [2, 3])
def test_entropy_grad_to_prim_grad(dim):
    rng = np.random.default_rng(30 + dim)
    ph = oracle.phys(gamma=G)

    def prim_vec(u):
        r, v, p = R.prim(u, G)
        return np.concatenate([[r], v, [p]])

    for _ in range(30):
        u = R.random_state(rng, dim)
        Jpu = _jac(prim_vec, u)
        Jwu = _jac(lambda x: R.entropy_vars(x, G), u)
        gw = rng.normal(size=dim + 2)
        expect = Jpu @ np.linalg.solve(Jwu, gw)
        got = oracle.entgrad2primgrad(dim, ph, u, gw)
        assert np.allclose(got, expect, rtol=1e-6, atol=1e-6 * np.abs(expect).max())


def test_viscous_flux_properties():
    mu, Pr = 0.37, 0.71
    ph = oracle.phys(gamma=G, Pr=Pr, mu=mu, viscous=True)
    rng = np.random.default_rng(5)
    for dim in (2, 3):
        u = R.random_state(rng, dim)
        rho, vel, p = R.prim(u, G)
        # rigid rotation + uniform dilatation: antisymmetric + isotropic velocity gradient
        A = rng.normal(size=(dim, dim))
        Gv = (A - A.T) + 0.7 * np.eye(dim)
        grads = [np.zeros(dim + 2) for _ in range(dim)]
        for c in range(dim):
            grads[c][1:1 + dim] = Gv[:, c]
        fl = oracle.viscous_flux(dim, ph, u, *grads)
        if dim == 2:
            # Stokes' hypothesis is exact in 3D only; 2D keeps the 2/3 bulk term
            assert abs(fl[1, 0] - mu * (2 * 0.7 - 2 / 3 * 1.4)) < 1e-14
        else:
            assert np.allclose(fl, 0.0, atol=1e-14)
        # pure shear du/dy = s
        grads = [np.zeros(dim + 2) for _ in range(dim)]
        grads[1][1] = 2.5
        fl = oracle.viscous_flux(dim, ph, u, *grads)
        assert abs(fl[1, 1] - mu * 2.5) < 1e-14 and abs(fl[2, 0] - mu * 2.5) < 1e-14
        assert abs(fl[-1, 1] - vel[0] * mu * 2.5) < 1e-13 and abs(fl[-1, 0] - vel[1] * mu * 2.5) < 1e-13
        assert np.all(fl[0] == 0.0)
        # heat flux: energy row = mu gamma/(Pr (gamma-1)) grad(p/rho)
        g = rng.normal(size=(dim, 2))  # d rho / dx_c, d p / dx_c
        grads = [np.zeros(dim + 2) for _ in range(dim)]
        for c in range(dim):
            grads[c][0], grads[c][-1] = g[c]
        fl = oracle.viscous_flux(dim, ph, u, *grads)
        for c in range(dim):
            h = 1e-6
            dT = ((p + h * g[c, 1]) / (rho + h * g[c, 0]) - (p - h * g[c, 1]) / (rho - h * g[c, 0])) / (2 * h)
            assert abs(fl[-1, c] - mu * G / (Pr * (G - 1)) * dT) < 1e-7 * (1 + abs(fl[-1, c]))


def test_sutherland_at_reference_temperature():
    R_gas, T0 = 287.0, 273.15
    ph = oracle.phys(gamma=G, Pr=0.72, mu=123.0, R_gas=R_gas, viscous=True, sutherland=True)
    rho = 1.2
    p = rho * R_gas * T0
    u = np.array([rho, 0.0, 0.0, 0.0, p / (G - 1)])
    grads = [np.zeros(5) for _ in range(3)]
    grads[1][1] = 1.0
    fl = oracle.viscous_flux(3, ph, u, *grads)
    assert abs(fl[1, 1] - 1.716e-5) < 1e-18
