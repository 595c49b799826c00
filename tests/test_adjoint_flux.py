"""The flux vector-Jacobian products of the adjoint (prandtl_amd/csrc/dg_adjoint.hpp), built
for the host, against central finite differences of the oracle's forward fluxes (the
reference's ChandrashekarFlux restated): for random states, normals and cotangents F_b,
s_b[k] = <F_b, (F(s + h e_k) - F(s - h e_k)) / 2h>.  The flux tests take near and far state
pairs (the log mean's short series and its closed form); test_log_mean_partials covers all three
of its branches against 40-digit values."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests import numpy_ref as R

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def vjp_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("adj") / "libadjvjp.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out,
                    os.path.join(HERE, "native", "adjoint_vjp_host.cpp")], check=True)
    L = ctypes.CDLL(out)
    dp = ctypes.POINTER(ctypes.c_double)
    L.vjp_volume.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp, dp, dp, dp, dp]
    L.vjp_face.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp, dp, dp, dp, dp]
    L.vjp_entropy_vars.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp, dp]
    L.log_mean_partials.argtypes = [ctypes.c_double, ctypes.c_double, dp]
    L.vjp_entgrad.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp, dp, dp, dp]
    L.vjp_viscous_flux.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp, dp, dp,
                                   dp]
    return L


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fd(f, s, k, h):
    e = np.zeros_like(s)
    e[k] = h * max(1.0, abs(s[k]))
    return (f(s + e) - f(s - e)) / (2 * e[k])


def _pair(rng, dim, near):
    a = R.random_state(rng, dim)
    if near:
        b = a * (1.0 + 1e-3 * rng.uniform(-1, 1, dim + 2))
    else:
        b = R.random_state(rng, dim)
    return a, b


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("near", [False, True])
def test_volume_flux_vjp(vjp_lib, dim, near):
    rng = np.random.default_rng(40 + dim + 10 * near)
    ph = oracle.phys(gamma=1.4)
    for _ in range(20):
        a, b = _pair(rng, dim, near)
        m = rng.normal(size=dim)
        Fb = rng.normal(size=dim + 2)
        ab, bb = np.zeros(dim + 2), np.zeros(dim + 2)
        vjp_lib.vjp_volume(dim, 1.4, _p(a), _p(b), _p(m), _p(Fb), ab.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           bb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        for k in range(dim + 2):
            ga = Fb @ _fd(lambda s: oracle.volume_flux(dim, ph, s, b, m, m)[0], a, k, 1e-6)
            gb = Fb @ _fd(lambda s: oracle.volume_flux(dim, ph, a, s, m, m)[0], b, k, 1e-6)
            assert abs(ab[k] - ga) <= 1e-6 * (1 + abs(ga)), (k, ab[k], ga)
            assert abs(bb[k] - gb) <= 1e-6 * (1 + abs(gb)), (k, bb[k], gb)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("near", [False, True])
@pytest.mark.parametrize("bug3", [0, 1])
def test_face_flux_vjp(vjp_lib, dim, kind, near, bug3):
    if bug3 and (dim == 2 or kind == 1):
        pytest.skip("the typo exists only in the 3D Chandrashekar dissipation")
    rng = np.random.default_rng(60 + dim + 10 * kind + 100 * near)
    ph = oracle.phys(gamma=1.4, surface_flux=kind, bug_compat=4 if bug3 else 0)
    for _ in range(20):
        a, b = _pair(rng, dim, near)
        n = rng.normal(size=dim)
        fb = rng.normal(size=dim + 2)
        ab, bb = np.zeros(dim + 2), np.zeros(dim + 2)
        vjp_lib.vjp_face(dim, kind, bug3, 1.4, _p(a), _p(b), _p(n), _p(fb),
                         ab.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         bb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        for k in range(dim + 2):
            ga = fb @ _fd(lambda s: oracle.face_flux(dim, ph, s, b, n)[0], a, k, 1e-6)
            gb = fb @ _fd(lambda s: oracle.face_flux(dim, ph, a, s, n)[0], b, k, 1e-6)
            assert abs(ab[k] - ga) <= 2e-6 * (1 + abs(ga)), (k, ab[k], ga)
            assert abs(bb[k] - gb) <= 2e-6 * (1 + abs(gb)), (k, bb[k], gb)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_face_flux_vjp_equal_traces(vjp_lib, dim, kind):
    """Equal traces (a continuous state) sit on the kink of lambda = max(lambda_1, lambda_2):
    central differences measure the mean of the one-sided derivatives, the symmetric
    subgradient the adjoint takes there."""
    rng = np.random.default_rng(90 + dim + 10 * kind)
    ph = oracle.phys(gamma=1.4, surface_flux=kind)
    for _ in range(20):
        a = R.random_state(rng, dim)
        n = rng.normal(size=dim)
        fb = rng.normal(size=dim + 2)
        ab, bb = np.zeros(dim + 2), np.zeros(dim + 2)
        vjp_lib.vjp_face(dim, kind, 0, 1.4, _p(a), _p(a), _p(n), _p(fb),
                         ab.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         bb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        for k in range(dim + 2):
            ga = fb @ _fd(lambda s: oracle.face_flux(dim, ph, s, a, n)[0], a, k, 1e-6)
            gb = fb @ _fd(lambda s: oracle.face_flux(dim, ph, a, s, n)[0], a, k, 1e-6)
            assert abs(ab[k] - ga) <= 2e-6 * (1 + abs(ga)), (k, ab[k], ga)
            assert abs(bb[k] - gb) <= 2e-6 * (1 + abs(gb)), (k, bb[k], gb)


@pytest.fixture(scope="module")
def op_vjp_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("adjop") / "libadjop.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out,
                    os.path.join(HERE, "native", "adjoint_op_host.cpp")], check=True)
    L = ctypes.CDLL(out)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.op_vjp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, dp, ctypes.c_double, ctypes.c_double,
                         ctypes.c_int, ctypes.c_int, dp, dp, dp]
    return L


def _op_vjp(L, mesh, u, lam, kind, bug3=0, gamma=1.4):
    _, w = oracle.gll(mesh.order + 1)
    _, _, Dh2 = oracle.dmatrices(mesh.order + 1)
    out = np.zeros_like(u)
    L.op_vjp(mesh.dim, mesh.order + 1, mesh.num_elements,
             np.ascontiguousarray(mesh.neighbors, dtype=np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
             _p(mesh.geom), _p(Dh2), 1.0 / w[0], gamma, kind, bug3, _p(u), _p(lam),
             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_operator_vjp_host(op_vjp_lib, dim, kind):
    """The assembled operator VJP (host build of the GPU kernel's assembly order) against
    central differences of the oracle's RHS: <lam, J v> = <J^T lam, v> on a state with jumps
    across faces (relative 1e-7)."""
    from prandtl_amd import cases, mesh as M

    rng = np.random.default_rng(7 + dim + 10 * kind)
    if dim == 2:
        m = M.periodic_lattice(2, 3, (5, 4), A=np.array([[0.5, 0.1], [0.05, 0.45]]))
        X = m.node_coords()
        u = cases.prim_to_conserv(1.0 + 0.2 * np.sin(X[0]), [0.3 * np.cos(X[1]), 0.2 * np.sin(X[0])],
                                  1.0 + 0.1 * np.cos(X[0] + X[1]))
    else:
        m = cases.tgv_mesh(n=3, order=4)
        u = cases.tgv_modes_state(m, 1e-2 * np.ones(8), cases.fourier_modes(count=8, kmax=2, seed=7), mach=0.3)
    # 5% jumps: a face point whose two wave speeds lie within eps |v| of each other (the kink of
    # max(|V|+c)) would put an O(1) error into its difference quotient; the expected number of
    # such points is ~1e-2 here
    u = u * (1.0 + 5e-2 * rng.uniform(-1, 1, u.shape))
    lam = rng.normal(size=u.shape)
    v = rng.normal(size=u.shape) * np.abs(u)
    ph = oracle.phys(gamma=1.4, surface_flux=kind)
    g = _op_vjp(op_vjp_lib, m, u, lam, kind)
    eps = 1e-6
    dp, _ = oracle.rhs(m, ph, u + eps * v)
    dm, _ = oracle.rhs(m, ph, u - eps * v)
    lhs = float((lam * (dp - dm)).sum() / (2 * eps))
    rhs = float((g * v).sum())
    assert abs(lhs - rhs) <= 1e-7 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def _out(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("dim", [2, 3])
def test_entropy_vars_vjp(vjp_lib, dim):
    """Conserv2Entropy (the reference's BasicOperations.cpp:81-104, oracle) reversed."""
    rng = np.random.default_rng(200 + dim)
    ph = oracle.phys(gamma=1.4)
    for _ in range(20):
        a = R.random_state(rng, dim)
        wb = rng.normal(size=dim + 2)
        sb = np.zeros(dim + 2)
        vjp_lib.vjp_entropy_vars(dim, 1.4, _p(a), _p(wb), _out(sb))
        for k in range(dim + 2):
            ref = wb @ _fd(lambda x: oracle.conserv2entropy(dim, ph, x), a, k, 1e-6)
            assert abs(sb[k] - ref) <= 1e-6 * (1 + abs(ref)), (k, sb[k], ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_entgrad_vjp(vjp_lib, dim):
    """EntropyGrad2PrimGrad (BasicOperations.cpp:126-170, oracle) reversed in the state and
    the entropy-variable gradient."""
    rng = np.random.default_rng(210 + dim)
    ph = oracle.phys(gamma=1.4)
    for _ in range(20):
        a = R.random_state(rng, dim)
        g = rng.normal(size=dim + 2)
        gpb = rng.normal(size=dim + 2)
        sb, gb = np.zeros(dim + 2), np.zeros(dim + 2)
        vjp_lib.vjp_entgrad(dim, 1.4, _p(a), _p(g), _p(gpb), _out(sb), _out(gb))
        for k in range(dim + 2):
            rs = gpb @ _fd(lambda x: oracle.entgrad2primgrad(dim, ph, x, g), a, k, 1e-6)
            rg = gpb @ _fd(lambda x: oracle.entgrad2primgrad(dim, ph, a, x), g, k, 1e-6)
            assert abs(sb[k] - rs) <= 1e-6 * (1 + abs(rs)), (k, sb[k], rs)
            assert abs(gb[k] - rg) <= 1e-6 * (1 + abs(rg)), (k, gb[k], rg)


@pytest.mark.parametrize("dim", [2, 3])
def test_viscous_flux_vjp(vjp_lib, dim):
    """NavierStokesFlux::ComputeViscousFlux (NavierStokesFlux.cpp:12-90, oracle; constant mu)
    reversed in the state and the primitive gradients."""
    rng = np.random.default_rng(220 + dim)
    mu, Pr = 0.013, 0.71
    ph = oracle.phys(gamma=1.4, mu=mu, Pr=Pr, R_gas=1.0, viscous=True)
    for _ in range(20):
        a = R.random_state(rng, dim)
        g = rng.normal(size=(dim, dim + 2))
        fb = rng.normal(size=(dim + 1, dim))
        sb, gb = np.zeros(dim + 2), np.zeros((dim, dim + 2))
        vjp_lib.vjp_viscous_flux(dim, 1.4, mu, Pr, _p(a), _p(g), _p(fb), _out(sb), _out(gb))

        def F(x, gg):
            return oracle.viscous_flux(dim, ph, x, *[gg[c] for c in range(dim)])[1:, :]

        for k in range(dim + 2):
            rs = np.sum(fb * _fd(lambda x: F(x, g), a, k, 1e-6))
            assert abs(sb[k] - rs) <= 1e-6 * (1 + abs(rs)), (k, sb[k], rs)
        for c in range(dim):
            for k in range(dim + 2):
                def Fg(x, c=c):
                    gg = g.copy()
                    gg[c] = x
                    return F(a, gg)
                rg = np.sum(fb * _fd(Fg, g[c], k, 1e-6))
                assert abs(gb[c, k] - rg) <= 1e-6 * (1 + abs(rg)), (c, k, gb[c, k], rg)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_operator_vjp_ns_host(op_vjp_lib, dim, kind):
    """The Navier-Stokes operator VJP (host build of the GPU kernels' assembly: the inviscid
    part plus BR1 lifting, EntropyGrad2PrimGrad, viscous flux, divergence and face average)
    against central differences of the oracle's NS RHS: <lam, J v> = <J^T lam, v>, rel. 1e-7."""
    from prandtl_amd import cases, mesh as M

    L = op_vjp_lib
    dp_ = ctypes.POINTER(ctypes.c_double)
    L.op_vjp_visc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp_, dp_, dp_,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp_, dp_, dp_]
    rng = np.random.default_rng(300 + dim + 10 * kind)
    if dim == 2:
        m = M.periodic_lattice(2, 3, (5, 4), A=np.array([[0.5, 0.1], [0.05, 0.45]]))
        X = m.node_coords()
        u = cases.prim_to_conserv(1.0 + 0.2 * np.sin(X[0]), [0.3 * np.cos(X[1]), 0.2 * np.sin(X[0])],
                                  1.0 + 0.1 * np.cos(X[0] + X[1]))
    else:
        m = M.periodic_lattice(3, 3, (3, 3, 3), A=np.array([[0.5, 0.1, 0.0], [0.05, 0.45, 0.08],
                                                            [0.02, -0.03, 0.55]]))
        X = m.node_coords()
        u = cases.prim_to_conserv(1.0 + 0.2 * np.sin(X[0]) * np.cos(X[2]),
                                  [0.3 * np.cos(X[1]), 0.2 * np.sin(X[0]), 0.1 * np.cos(X[2])],
                                  1.0 + 0.1 * np.cos(X[0] + X[1]))
    u = u * (1.0 + 5e-2 * rng.uniform(-1, 1, u.shape))
    lam = rng.normal(size=u.shape)
    v = rng.normal(size=u.shape) * np.abs(u)
    mu, Pr = 0.05, 0.71
    ph = oracle.phys(gamma=1.4, mu=mu, Pr=Pr, R_gas=1.0, viscous=True, surface_flux=kind)
    g = _op_vjp(L, m, u, lam, kind)
    D, Dh, _ = oracle.dmatrices(m.order + 1)
    _, w = oracle.gll(m.order + 1)
    L.op_vjp_visc(m.dim, m.order + 1, m.num_elements,
                  np.ascontiguousarray(m.neighbors, dtype=np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                  _p(m.geom), _p(D), _p(Dh), 1.0 / w[0], 1.4, mu, Pr, _p(u), _p(lam), g.ctypes.data_as(dp_))
    eps = 1e-6
    dp, _ = oracle.rhs(m, ph, u + eps * v)
    dm, _ = oracle.rhs(m, ph, u - eps * v)
    lhs = float((lam * (dp - dm)).sum() / (2 * eps))
    rhs = float((g * v).sum())
    assert abs(lhs - rhs) <= 1e-7 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_log_mean_partials(vjp_lib):
    """The logarithmic mean L(x, y) = (y - x)/ln(y/x), its reciprocal and its partials
    dL/dx = (L/x - 1) L/(y - x), dL/dy = (1 - L/y) L/(y - x) against 40-digit values, on all three
    of the forward's branches in u = ((y-x)/(y+x))^2: the short series (u < 1e-4), the series to
    u^11 (u < 0.05) and the closed form.  Relative errors within 64 ulp; the closed form alone
    loses digits to cancellation in the partials where u is small."""
    import mpmath

    mpmath.mp.dps = 40
    rng = np.random.default_rng(8)
    out = np.zeros(4)
    worst = {}
    for target_u in (1e-7, 5e-5, 2e-4, 3e-3, 2e-2, 4.9e-2, 0.1, 0.5):
        for _ in range(40):
            x = float(10 ** rng.uniform(-3, 3))
            f = float(np.sqrt(target_u) * rng.uniform(0.9, 1.0) * rng.choice([-1.0, 1.0]))
            y = x * (1 + f) / (1 - f)
            vjp_lib.log_mean_partials(x, y, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            X, Y = mpmath.mpf(x), mpmath.mpf(y)
            Lm = (Y - X) / mpmath.log(Y / X)
            ref = [Lm, 1 / Lm, (Lm / X - 1) * Lm / (Y - X), (1 - Lm / Y) * Lm / (Y - X)]
            for k in range(4):
                e = float(abs((mpmath.mpf(out[k]) - ref[k]) / ref[k])) / np.finfo(float).eps
                worst[(target_u, k)] = max(worst.get((target_u, k), 0.0), e)
    bad = {key: v for key, v in worst.items() if v > 64}
    assert not bad, bad
