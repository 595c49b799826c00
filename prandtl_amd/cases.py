"""Synthetic workloads: meshes, initial states and physics of the benchmark cases.

All states are conserved variables stored byNODES, shape (neq, NE, n), float64.
Random numbers come from splitmix64 with the (x >> 11) * 2^-53 mapping, so every
platform draws the same values.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import Mesh, periodic_lattice

GAMMA = 1.4


def splitmix64(seed: int, count: int) -> np.ndarray:
    """Uniform [0,1) doubles from splitmix64 (Steele, Lea & Flood 2014)."""
    out = np.empty(count, dtype=np.float64)
    mask = (1 << 64) - 1
    x = seed & mask
    for i in range(count):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out


def splitmix64_np(seed: int, count: int) -> np.ndarray:
    """Vectorised splitmix64, identical to :func:`splitmix64`."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        x = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def prim_to_conserv(rho, vel, p, gamma=GAMMA):
    vel = list(vel)
    ke = 0.5 * rho * sum(v * v for v in vel)
    return np.stack([rho] + [rho * v for v in vel] + [p / (gamma - 1.0) + ke])


def conserv_to_prim(u, gamma=GAMMA):
    dim = u.shape[0] - 2
    rho = u[0]
    vel = [u[1 + d] / rho for d in range(dim)]
    p = (gamma - 1.0) * (u[-1] - 0.5 * rho * sum(v * v for v in vel))
    return rho, vel, p


# ---------------------------------------------------------------------------
# 2D isentropic (Shu) vortex on periodic [0,10]^2
# ---------------------------------------------------------------------------
def vortex_mesh(n=16, order=3) -> Mesh:
    return periodic_lattice(2, order, (n, n), lengths=(10.0, 10.0))


def vortex_state(mesh: Mesh, beta=5.0, center=(5.0, 5.0), u_inf=(1.0, 1.0), gamma=GAMMA, t=0.0,
                 period=10.0):
    """Shu's isentropic vortex advected by (u_inf) for time t (exact solution on
    the periodic domain, nearest image)."""
    X = mesh.node_coords()
    x0 = (center[0] + u_inf[0] * t) % period
    y0 = (center[1] + u_inf[1] * t) % period
    dx = (X[0] - x0 + period / 2) % period - period / 2
    dy = (X[1] - y0 + period / 2) % period - period / 2
    r2 = dx * dx + dy * dy
    f = np.exp(0.5 * (1.0 - r2))
    du = -beta / (2 * math.pi) * f * dy
    dv = beta / (2 * math.pi) * f * dx
    T = 1.0 - (gamma - 1.0) * beta**2 / (8.0 * gamma * math.pi**2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho * T
    return prim_to_conserv(rho, [u_inf[0] + du, u_inf[1] + dv], p, gamma)


# ---------------------------------------------------------------------------
# 3D Taylor-Green vortex on periodic [0, 2 pi]^3
# ---------------------------------------------------------------------------
def tgv_mesh(n=32, order=3) -> Mesh:
    L = 2 * math.pi
    return periodic_lattice(3, order, (n, n, n), lengths=(L, L, L))


def tgv_state(mesh: Mesh, mach=0.1, gamma=GAMMA):
    X = mesh.node_coords()
    x, y, z = X
    rho = np.ones_like(x)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    p = 1.0 / (gamma * mach**2) + (1.0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    return prim_to_conserv(rho, [u, v, w], p, gamma)


def tgv_physics(mach=0.1, reynolds=1600.0, prandtl=0.71, gamma=GAMMA):
    """Nondimensional TGV: rho0 = V0 = L0 = 1, so mu = 1/Re; constant viscosity."""
    return dict(gamma=gamma, Pr=prandtl, mu=1.0 / reynolds, R_gas=1.0, viscous=True)


def fourier_modes(count=64, kmax=4, seed=4242):
    """Seeded divergence-free velocity modes: integer wavevectors 1 <= |k| <= kmax,
    unit polarisations perpendicular to k, and phases in [0, 2 pi)."""
    r = splitmix64_np(seed, 16 * count * 4)
    ks, dirs, phases = [], [], []
    i = 0
    while len(ks) < count:
        k = np.floor(r[i:i + 3] * (2 * kmax + 1)).astype(int) - kmax
        i += 3
        kn = float(np.linalg.norm(k))
        if kn < 1.0 or kn > kmax:
            continue
        a = r[i:i + 3] - 0.5
        i += 3
        e = a - (a @ k) / (kn * kn) * k
        en = np.linalg.norm(e)
        if en < 1e-3:
            continue
        ks.append(k)
        dirs.append(e / en)
        phases.append(2 * math.pi * r[i])
        i += 1
    return np.array(ks, dtype=np.float64), np.array(dirs), np.array(phases)


def tgv_modes_state(mesh: Mesh, amplitudes, modes, mach=0.1, gamma=GAMMA):
    """TGV plus sum_m a_m e_m sin(k_m . x + phi_m) in the velocity."""
    ks, dirs, phases = modes
    X = mesh.node_coords()
    x, y, z = X
    rho = np.ones_like(x)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    for a, k, e, ph in zip(amplitudes, ks, dirs, phases):
        s = a * np.sin(k[0] * x + k[1] * y + k[2] * z + ph)
        u = u + e[0] * s
        v = v + e[1] * s
        w = w + e[2] * s
    p = 1.0 / (gamma * mach**2) + (1.0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    return prim_to_conserv(rho, [u, v, w], p, gamma)


# ---------------------------------------------------------------------------
# 2D Kelvin-Helmholtz instability on periodic [-1,1]^2
# ---------------------------------------------------------------------------
def khi_mesh(n=256, order=3) -> Mesh:
    return periodic_lattice(2, order, (n, n), lengths=(2.0, 2.0), origin=(-1.0, -1.0))


def khi_state(mesh: Mesh, width=0.1, noise=1e-2, seed=1234, gamma=GAMMA):
    """rho = 0.5 + 0.75 B, u = 0.5 (B - 1), v = 0.1 sin(2 pi x) + U(-noise, noise),
    p = 1, with B = tanh((y + 0.5)/width) - tanh((y - 0.5)/width)."""
    X = mesh.node_coords()
    x, y = X
    B = np.tanh((y + 0.5) / width) - np.tanh((y - 0.5) / width)
    rho = 0.5 + 0.75 * B
    u = 0.5 * (B - 1.0)
    r = splitmix64_np(seed, x.size).reshape(x.shape)
    v = 0.1 * np.sin(2 * math.pi * x) + noise * (2.0 * r - 1.0)
    p = np.ones_like(x)
    return prim_to_conserv(rho, [u, v], p, gamma)


# ---------------------------------------------------------------------------
# 2D double Mach reflection on [0,4] x [0,1] (the reference's TestCases/Euler/2D/
# DoubleMachReflection: its mesh is the 384 x 96 lattice of this domain, refined once)
# ---------------------------------------------------------------------------
DMR_BOUNDARIES = ("left", "bottom1", "bottom2", "right", "top")   # the configuration's order


def dmr_mesh(nx=384, ny=96, order=3) -> Mesh:
    """nx x ny square-element lattice of [0,4] x [0,1] (nx = 4 ny, a multiple of 24, so that x = 1/6
    falls on an element face).  Boundary faces carry the conditions in DMR_BOUNDARIES order: left
    (x = 0), bottom1 (y = 0, x < 1/6), bottom2 (y = 0, x > 1/6), right (x = 4), top (y = 1)."""
    assert nx % 24 == 0
    m = periodic_lattice(2, order, (nx, ny), lengths=(4.0, 1.0), periodic=(False, False),
                         boundary={(0, 0): 0, (0, 1): 3, (1, 0): 2, (1, 1): 4})
    e = np.arange(m.num_elements)
    first = (e // nx == 0) & (e % nx < nx // 24)   # bottom faces left of x = 1/6
    m.neighbors[first, 2] = -(1 + 1)
    return m


def dmr_state(mesh: Mesh, gamma=GAMMA):
    """DoubleMachReflectionIC: the Mach 10 post-shock state (rho 8, p 116.5, velocity 8.25 along
    (cos 30, -sin 30)) where x < 1/6 + y/sqrt 3, the resting gas (rho 1.4, p 1) elsewhere."""
    from .config import _dmr_states

    post, pre = _dmr_states(1.0 / (gamma - 1.0))
    X = mesh.node_coords()
    behind = X[0] < 1.0 / 6.0 + X[1] / math.sqrt(3)
    return np.stack([np.where(behind, post[v], pre[v]) for v in range(4)])


def cfl_dt(mesh: Mesh, lam_max: float, cfl: float) -> float:
    """dt = cfl * h_min / (lambda_max (p+1)^2) -- ref: src/Simulation/Simulation.cpp:906."""
    return cfl * mesh.h_min() / (lam_max * (mesh.order + 1) ** 2)


def max_wave_speed(u, gamma=GAMMA) -> float:
    rho, vel, p = conserv_to_prim(u, gamma)
    vm = np.sqrt(sum(v * v for v in vel))
    return float(np.max(vm + np.sqrt(gamma * p / rho)))


def kinetic_energy(mesh: Mesh, u) -> float:
    """Volume-averaged kinetic energy with the GLL quadrature."""
    W = mesh.quadrature_weights()
    rho = u[0]
    ke = 0.5 * sum(u[1 + d] ** 2 for d in range(mesh.dim)) / rho
    return float((W * ke).sum() / W.sum())


# Ghia, Ghia & Shin, J. Comput. Phys. 48 (1982) 387-411, Table I: the horizontal velocity u/U on
# the vertical centreline of the lid-driven cavity at Re = 100 (y/L from the bottom wall).
GHIA_RE100_Y = np.array([1.0000, 0.9766, 0.9688, 0.9609, 0.9531, 0.8516, 0.7344, 0.6172, 0.5000,
                         0.4531, 0.2813, 0.1719, 0.1016, 0.0703, 0.0625, 0.0547, 0.0000])
GHIA_RE100_U = np.array([1.00000, 0.84123, 0.78871, 0.73722, 0.68717, 0.23151, 0.00332, -0.13641, -0.20581,
                         -0.21090, -0.15662, -0.10150, -0.06434, -0.04775, -0.04192, -0.03717, 0.00000])


def lattice_line_values(mesh: Mesh, field: np.ndarray, x0: float, ys: np.ndarray) -> np.ndarray:
    """A nodal field (NE, n) of a 2D Cartesian lattice evaluated on the vertical line x = x0 at
    heights ys: x0 must lie on element faces (or nodes) of the lattice; along y the element's GLL
    interpolant is evaluated (Lagrange basis of its Np nodes)."""
    from .mesh import gll_nodes

    assert mesh.dim == 2 and mesh.n is not None
    np_ = mesh.np1
    xg, _ = gll_nodes(np_)
    hx = mesh.A[0, 0]
    hy = mesh.A[1, 1]
    ex = int(np.clip(np.floor((x0 - mesh.origin[0]) / hx + 1e-9), 0, mesh.n[0] - 1))
    xi = (x0 - mesh.origin[0]) / hx - ex
    i = int(np.argmin(np.abs(xg - xi)))
    assert abs(xg[i] - xi) < 1e-9, "x0 must be a node line of the lattice"
    out = np.zeros(len(ys))
    for k, y in enumerate(ys):
        ey = int(np.clip(np.floor((y - mesh.origin[1]) / hy), 0, mesh.n[1] - 1))
        eta = (y - mesh.origin[1]) / hy - ey
        e = ex + mesh.n[0] * ey
        L = np.array([np.prod([(eta - xg[m]) / (xg[j] - xg[m]) for m in range(np_) if m != j]) for j in range(np_)])
        vals = field[e].reshape(np_, np_)[:, i]          # node (i, j): index i + Np j
        out[k] = float(L @ vals)
    return out
