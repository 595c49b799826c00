"""Discrete adjoint on the GPU (Euler; no counterpart in the reference).

* operator VJP: <lambda, J v> = <J^T lambda, v>, with J v from central finite differences of
  the forward RHS through the C ABI (relative 1e-7), for the Chandrashekar and EC surface
  fluxes, 2D p=3 and 3D p=4, at states with 5% jumps across faces and a step eps = 1e-6.  The
  dissipation's wave speed max(|V_1|+c_1, |V_2|+c_2) has a kink; a face point whose two
  speeds lie within eps |v| of each other puts an O(1) error into its difference quotient, so
  the jumps are large enough that such a point is unlikely (expected count ~1e-2 here);
* the GPU VJP against the host build of the same flux VJPs in the kernel's assembly order;
* RK adjoint with checkpoint segments: the dot-product identity <ubar_T, du_T> = <ubar_0, du_0>
  for a smooth (Fourier-mode) initial perturbation, which moves the face jumps of the stages
  only by O(eps) of their own size and so stays off the kink; gradients of J = integrated kinetic energy after LSRK4(5)
  steps with respect to the Mach number and Fourier-mode amplitudes of the initial velocity,
  against sixth-order central differences (Richardson over fourth-order ones at h and 2h)
  with steps where truncation and rounding are both below the 1e-6 relative tolerance: J's
  dependence on the Mach number is weak (dJ/dMa Ma / J ~ 1e-6), so a step of 1e-6 Ma leaves
  a second-order difference quotient at the rounding noise of J.  The fourth-order values at
  h and 2h are printed beside the sixth-order one.
"""
import numpy as np
import pytest

from prandtl_amd import capi, cases, mesh as M

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _state_2d(rng):
    m = M.periodic_lattice(2, 3, (5, 4), A=np.array([[0.5, 0.1], [0.05, 0.45]]))
    X = m.node_coords()
    rho = 1.0 + 0.2 * np.sin(X[0]) + 0.05 * rng.uniform(-1, 1, X[0].shape)
    vel = [0.3 * np.cos(X[1]) + 0.05 * rng.uniform(-1, 1, X[0].shape), 0.2 * np.sin(X[0])]
    p = 1.0 + 0.1 * np.cos(X[0] + X[1]) + 0.05 * rng.uniform(-1, 1, X[0].shape)
    return m, cases.prim_to_conserv(rho, vel, p)


def _state_3d(rng, jump=5e-2):
    m = cases.tgv_mesh(n=3, order=4)
    modes = cases.fourier_modes(count=8, kmax=2, seed=7)
    u = cases.tgv_modes_state(m, 1e-2 * np.ones(8), modes, mach=0.3)
    return m, u * (1.0 + jump * rng.uniform(-1, 1, u.shape))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("surf", [capi.SURF_CHANDRASHEKAR, capi.SURF_EC])
def test_operator_vjp_dot_product(device, dim, surf):
    import torch

    _say(f"operator VJP dot product dim={dim} surf={surf}")
    rng = np.random.default_rng(5 + dim)
    m, u = _state_2d(rng) if dim == 2 else _state_3d(rng)
    op = capi.Operator(m, surface_flux=surf)
    ud = torch.from_numpy(u).to(device)
    lam = torch.from_numpy(rng.normal(size=u.shape)).to(device)
    v = torch.from_numpy(rng.normal(size=u.shape) * np.abs(u)).to(device)
    g = torch.empty_like(ud)
    op.vjp(ud, lam, g)
    eps = 1e-6
    dp, dm = torch.empty_like(ud), torch.empty_like(ud)
    op.mult(ud + eps * v, dp)
    op.mult(ud - eps * v, dm)
    torch.cuda.synchronize()
    lhs = (lam * (dp - dm) / (2 * eps)).sum().item()
    rhs = (g * v).sum().item()
    print(f"[adjoint] <lam, Jv> = {lhs:.12e}, <J^T lam, v> = {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-7 * max(abs(lhs), abs(rhs))
    op.close()


def _rk_setup(device):
    import torch

    m = cases.tgv_mesh(n=3, order=4)
    modes = cases.fourier_modes(count=8, kmax=2, seed=11)
    W = torch.from_numpy(m.quadrature_weights()).to(device)
    return m, modes, W, capi.Operator(m), 2e-3, 12, 5


def test_rk_adjoint_dot_product(device):
    import torch

    from prandtl_amd import torch_adjoint as TA

    _say("RK adjoint dot product: TGV 3^3 p=4 Euler + 8 modes, 12 LSRK4(5) steps, checkpoints every 5")
    m, modes, W, op, dt, nsteps, ckpt = _rk_setup(device)
    rng = np.random.default_rng(3)
    with torch.no_grad():
        mach = torch.tensor(0.3, dtype=torch.float64, device=device)
        amps = torch.full((8,), 1e-2, dtype=torch.float64, device=device)
        u0 = TA.tgv_modes_state(m, mach, amps, modes)
        # smooth direction: per variable, 6 seeded Fourier modes (|k| <= 2) scaled by |u0|
        X = m.node_coords()
        dmodes = np.zeros(u0.shape)
        for v in range(u0.shape[0]):
            for _ in range(6):
                k = rng.integers(-2, 3, size=3)
                dmodes[v] += rng.normal() * np.sin(k[0] * X[0] + k[1] * X[1] + k[2] * X[2] + rng.uniform(0, 6.3))
        du0 = torch.from_numpy(1e-2 * dmodes).to(device) * u0.abs()
        eps = 1e-6
        up, um = u0 + eps * du0, u0 - eps * du0
        op.rk_run(capi.LSRK45, up, 0.0, dt, nsteps)
        op.rk_run(capi.LSRK45, um, 0.0, dt, nsteps)
        duT = (up - um) / (2 * eps)
        uT = u0.clone()
        op.rk_run_tape(uT, 0.0, dt, nsteps, ckpt)
        ref = u0.clone()
        op.rk_run(capi.LSRK45, ref, 0.0, dt, nsteps)
        assert torch.equal(uT, ref), "the taped run must reproduce the plain run bit for bit"
        ubT = torch.from_numpy(rng.normal(size=tuple(u0.shape))).to(device)
        ub0 = ubT.clone()
        op.rk_adjoint(ub0)
        lhs, rhs = (ubT * duT).sum().item(), (ub0 * du0).sum().item()
        print(f"[adjoint] <ubar_T, du_T> = {lhs:.12e}, <ubar_0, du_0> = {rhs:.12e}", flush=True)
        assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs))
    op.close()


def test_rk_adjoint_gradients(device):
    import torch

    from prandtl_amd import torch_adjoint as TA

    _say("RK adjoint gradients: TGV 3^3 p=4 Euler + 8 modes, 12 LSRK4(5) steps, checkpoints every 5")
    m, modes, W, op, dt, nsteps, ckpt = _rk_setup(device)

    def J_of(mach, amps):
        u0 = TA.tgv_modes_state(m, mach, amps, modes)
        return TA.kinetic_energy(TA.RKSolve.apply(u0, op, dt, nsteps, ckpt), W)

    mach = torch.tensor(0.3, dtype=torch.float64, device=device, requires_grad=True)
    amps = torch.full((8,), 1e-2, dtype=torch.float64, device=device, requires_grad=True)
    J = J_of(mach, amps)
    gM, gA = torch.autograd.grad(J, [mach, amps])
    with torch.no_grad():
        fM = lambda x: J_of(x, amps).item()
        d4, d42, fd = TA.central_fd(fM, mach, 1e-2 * 0.3)
        print(f"[adjoint] dJ/dMa adjoint {gM.item():.12e} FD6 {fd:.12e} (FD4 h {d4:.12e}, 2h {d42:.12e})",
              flush=True)
        assert abs(gM.item() - fd) <= 1e-6 * abs(fd)
        for i in (0, 3, 7):
            e = torch.zeros_like(amps)
            e[i] = 1.0
            fA = lambda x: J_of(mach, amps + (x - amps[i]) * e).item()
            d4, d42, fd = TA.central_fd(fA, amps[i], 1e-3)
            print(f"[adjoint] dJ/da_{i} adjoint {gA[i].item():.12e} FD6 {fd:.12e} (FD4 h {d4:.12e}, "
                  f"2h {d42:.12e})", flush=True)
            assert abs(gA[i].item() - fd) <= 1e-6 * abs(fd)
    op.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("surf", [capi.SURF_CHANDRASHEKAR, capi.SURF_EC])
def test_operator_vjp_vs_host_assembly(device, dim, surf):
    """The GPU VJP against the host build of the same flux VJPs assembled in the kernel's order
    (tests/native/adjoint_op_host.cpp), entrywise to 1e-11 of the largest entry, and the GPU
    forward against the oracle at the same state."""
    import ctypes
    import os
    import subprocess
    import tempfile

    import torch

    import oracle

    _say(f"operator VJP vs host assembly dim={dim} surf={surf}")
    rng = np.random.default_rng(5 + dim)
    m, u = _state_2d(rng) if dim == 2 else _state_3d(rng)
    lam = rng.normal(size=u.shape)
    here = os.path.dirname(os.path.abspath(__file__))
    so = os.path.join(tempfile.mkdtemp(), "libadjop.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", so,
                    os.path.join(here, "native", "adjoint_op_host.cpp")], check=True)
    L = ctypes.CDLL(so)
    dp = ctypes.POINTER(ctypes.c_double)
    L.op_vjp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp, dp,
                         ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp, dp]
    _, w = oracle.gll(m.order + 1)
    _, _, Dh2 = oracle.dmatrices(m.order + 1)
    href = np.zeros_like(u)
    nbr = np.ascontiguousarray(m.neighbors, dtype=np.int32)
    L.op_vjp(m.dim, m.order + 1, m.num_elements, nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
             m.geom.ctypes.data_as(dp), Dh2.ctypes.data_as(dp), 1.0 / w[0], 1.4, int(surf), 0,
             u.ctypes.data_as(dp), lam.ctypes.data_as(dp), href.ctypes.data_as(dp))
    op = capi.Operator(m, surface_flux=surf)
    ud = torch.from_numpy(u).to(device)
    g = torch.empty_like(ud)
    op.vjp(ud, torch.from_numpy(lam).to(device), g)
    du = torch.empty_like(ud)
    op.mult(ud, du)
    torch.cuda.synchronize()
    gg = g.cpu().numpy()
    diff = np.abs(gg - href)
    idx = np.unravel_index(np.argmax(diff), diff.shape)
    rel = diff.max() / np.abs(href).max()
    print(f"[adjoint] GPU VJP vs host assembly: max rel {rel:.3e} at (var, elem, node) {idx}", flush=True)
    ref, _ = oracle.rhs(m, oracle.phys(gamma=1.4, surface_flux=int(surf)), u)
    frel = np.abs(du.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"[adjoint] GPU forward vs oracle at the same state: rel {frel:.3e}", flush=True)
    assert frel <= 1e-12
    assert rel <= 1e-11
    op.close()


# ---------------------------------------------------------------------------
# Navier-Stokes (constant viscosity): the viscous part of the adjoint (BR1 lifting of the
# entropy variables, EntropyGrad2PrimGrad, viscous flux, its divergence and face average)
# ---------------------------------------------------------------------------
NS = dict(gamma=1.4, mu=0.05, Pr=0.71, R_gas=1.0, viscous=True)


def _ns_state(dim, rng):
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
    return m, u * (1.0 + 5e-2 * rng.uniform(-1, 1, u.shape))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("surf", [capi.SURF_CHANDRASHEKAR, capi.SURF_EC])
def test_operator_vjp_ns(device, dim, surf):
    """NS operator VJP: the dot-product identity against central differences of the GPU NS RHS
    (relative 1e-7), and entrywise agreement (1e-11 of the largest entry) with the host
    assembly of the same adjoint (tests/native/adjoint_op_host.cpp), itself checked against
    the oracle's NS RHS in tests/test_adjoint_flux.py."""
    import ctypes
    import os
    import subprocess
    import tempfile

    import torch

    import oracle

    _say(f"NS operator VJP dim={dim} surf={surf}")
    rng = np.random.default_rng(40 + dim + 10 * surf)
    m, u = _ns_state(dim, rng)
    op = capi.Operator(m, surface_flux=surf, **NS)
    ud = torch.from_numpy(u).to(device)
    lam_np = rng.normal(size=u.shape)
    lam = torch.from_numpy(lam_np).to(device)
    v = torch.from_numpy(rng.normal(size=u.shape) * np.abs(u)).to(device)
    g = torch.empty_like(ud)
    op.vjp(ud, lam, g)
    eps = 1e-6
    dp, dm = torch.empty_like(ud), torch.empty_like(ud)
    op.mult(ud + eps * v, dp)
    op.mult(ud - eps * v, dm)
    torch.cuda.synchronize()
    lhs = (lam * (dp - dm) / (2 * eps)).sum().item()
    rhs = (g * v).sum().item()
    print(f"[adjoint] NS <lam, Jv> = {lhs:.12e}, <J^T lam, v> = {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-7 * max(abs(lhs), abs(rhs))
    # host assembly
    here = os.path.dirname(os.path.abspath(__file__))
    so = os.path.join(tempfile.mkdtemp(), "libadjop.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", so,
                    os.path.join(here, "native", "adjoint_op_host.cpp")], check=True)
    L = ctypes.CDLL(so)
    dpp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.op_vjp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dpp, dpp, ctypes.c_double, ctypes.c_double,
                         ctypes.c_int, ctypes.c_int, dpp, dpp, dpp]
    L.op_vjp_visc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dpp, dpp, dpp, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double, dpp, dpp, dpp]
    D, Dh, Dh2 = oracle.dmatrices(m.order + 1)
    _, w = oracle.gll(m.order + 1)
    nbr = np.ascontiguousarray(m.neighbors, dtype=np.int32)
    href = np.zeros_like(u)
    args = (m.dim, m.order + 1, m.num_elements, nbr.ctypes.data_as(ip), m.geom.ctypes.data_as(dpp))
    L.op_vjp(*args, Dh2.ctypes.data_as(dpp), 1.0 / w[0], 1.4, int(surf), 0, u.ctypes.data_as(dpp),
             lam_np.ctypes.data_as(dpp), href.ctypes.data_as(dpp))
    L.op_vjp_visc(*args, D.ctypes.data_as(dpp), Dh.ctypes.data_as(dpp), 1.0 / w[0], 1.4, NS["mu"], NS["Pr"],
                  u.ctypes.data_as(dpp), lam_np.ctypes.data_as(dpp), href.ctypes.data_as(dpp))
    rel = np.abs(g.cpu().numpy() - href).max() / np.abs(href).max()
    print(f"[adjoint] NS GPU VJP vs host assembly: {rel:.3e}", flush=True)
    assert rel <= 1e-11
    op.close()


def test_rk_adjoint_ns(device):
    """RK adjoint through the NS operator (TGV 3^3 p=3 + 8 modes, mu = 0.02, 10 LSRK4(5) steps,
    checkpoints every 4): dJ/da_i and dJ/dMa against sixth-order central differences
    (1e-6 relative), and the dot-product identity for a smooth initial perturbation."""
    import torch

    from prandtl_amd import torch_adjoint as TA

    _say("NS RK adjoint: TGV 3^3 p=3 + 8 modes, mu=0.02, 10 LSRK4(5) steps")
    m = cases.tgv_mesh(n=3, order=3)
    modes = cases.fourier_modes(count=8, kmax=2, seed=13)
    W = torch.from_numpy(m.quadrature_weights()).to(device)
    op = capi.Operator(m, gamma=1.4, mu=0.02, Pr=0.71, R_gas=1.0, viscous=True)
    dt, nsteps, ckpt = 2e-3, 10, 4

    def J_of(mach, amps):
        u0 = TA.tgv_modes_state(m, mach, amps, modes)
        return TA.kinetic_energy(TA.RKSolve.apply(u0, op, dt, nsteps, ckpt), W)

    mach = torch.tensor(0.3, dtype=torch.float64, device=device, requires_grad=True)
    amps = torch.full((8,), 1e-2, dtype=torch.float64, device=device, requires_grad=True)
    J = J_of(mach, amps)
    gM, gA = torch.autograd.grad(J, [mach, amps])
    with torch.no_grad():
        d4, d42, fd = TA.central_fd(lambda x: J_of(x, amps).item(), mach, 1e-2 * 0.3)
        print(f"[adjoint] NS dJ/dMa adjoint {gM.item():.12e} FD6 {fd:.12e} (FD4 h {d4:.12e}, 2h {d42:.12e})",
              flush=True)
        assert abs(gM.item() - fd) <= 1e-6 * abs(fd)
        for i in (0, 5):
            e = torch.zeros_like(amps)
            e[i] = 1.0
            d4, d42, fd = TA.central_fd(lambda x: J_of(mach, amps + (x - amps[i]) * e).item(), amps[i], 1e-3)
            print(f"[adjoint] NS dJ/da_{i} adjoint {gA[i].item():.12e} FD6 {fd:.12e}", flush=True)
            assert abs(gA[i].item() - fd) <= 1e-6 * abs(fd)
    op.close()


@pytest.mark.parametrize("viscous", [False, True])
def test_rk_adjoint_ssprk3(device, viscous):
    """SSP-RK3 adjoint (Shu-Osher stages reversed; checkpoints every 3 of 8 steps) on the TGV
    3^3 p=3 + 8 modes, Euler and NS (mu = 0.02): dJ/da_i and dJ/dMa against sixth-order central
    differences, 1e-6 relative."""
    import torch

    from prandtl_amd import torch_adjoint as TA

    _say(f"SSP-RK3 adjoint viscous={viscous}")
    m = cases.tgv_mesh(n=3, order=3)
    modes = cases.fourier_modes(count=8, kmax=2, seed=17)
    W = torch.from_numpy(m.quadrature_weights()).to(device)
    kw = dict(gamma=1.4, mu=0.02, Pr=0.71, R_gas=1.0, viscous=True) if viscous else {}
    op = capi.Operator(m, **kw)
    dt, nsteps, ckpt = 2e-3, 8, 3

    def J_of(mach, amps):
        u0 = TA.tgv_modes_state(m, mach, amps, modes)
        return TA.kinetic_energy(TA.RKSolve.apply(u0, op, dt, nsteps, ckpt, capi.SSPRK3), W)

    mach = torch.tensor(0.3, dtype=torch.float64, device=device, requires_grad=True)
    amps = torch.full((8,), 1e-2, dtype=torch.float64, device=device, requires_grad=True)
    J = J_of(mach, amps)
    gM, gA = torch.autograd.grad(J, [mach, amps])
    with torch.no_grad():
        _, _, fd = TA.central_fd(lambda x: J_of(x, amps).item(), mach, 1e-2 * 0.3)
        print(f"[adjoint] SSP dJ/dMa adjoint {gM.item():.12e} FD6 {fd:.12e}", flush=True)
        assert abs(gM.item() - fd) <= 1e-6 * abs(fd)
        for i in (1, 6):
            e = torch.zeros_like(amps)
            e[i] = 1.0
            _, _, fd = TA.central_fd(lambda x: J_of(mach, amps + (x - amps[i]) * e).item(), amps[i], 1e-3)
            print(f"[adjoint] SSP dJ/da_{i} adjoint {gA[i].item():.12e} FD6 {fd:.12e}", flush=True)
            assert abs(gA[i].item() - fd) <= 1e-6 * abs(fd)
        # the taped forward reproduces the plain run bit for bit
        u0 = TA.tgv_modes_state(m, mach, amps, modes)
        a, b = u0.clone(), u0.clone()
        op.rk_run(capi.SSPRK3, a, 0.0, dt, nsteps)
        op.rk_run_tape(b, 0.0, dt, nsteps, ckpt, scheme=capi.SSPRK3)
        assert torch.equal(a, b)
    op.close()
