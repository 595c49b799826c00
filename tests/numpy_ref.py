"""Independent numpy statements of textbook gas-dynamics quantities, used to pin the
oracle (they are written from the definitions, not from the reference code)."""
import numpy as np


def prim(u, gamma):
    dim = len(u) - 2
    rho = u[0]
    vel = np.array(u[1:1 + dim]) / rho
    p = (gamma - 1.0) * (u[-1] - 0.5 * rho * vel @ vel)
    return rho, vel, p


def euler_flux_n(u, n, gamma):
    rho, vel, p = prim(u, gamma)
    vn = vel @ n
    return np.concatenate([[rho * vn], rho * vel * vn + p * np.asarray(n), [(u[-1] + p) * vn]])


def entropy_density(u, gamma):
    """Mathematical entropy U = -rho s / (gamma - 1), s = ln(p rho^-gamma)."""
    rho, vel, p = prim(u, gamma)
    s = np.log(p) - gamma * np.log(rho)
    return -rho * s / (gamma - 1.0)


def entropy_vars_fd(u, gamma, h=1e-6):
    """w = dU/du by central differences."""
    u = np.asarray(u, dtype=np.float64)
    w = np.zeros_like(u)
    for i in range(len(u)):
        e = np.zeros_like(u)
        e[i] = h * max(1.0, abs(u[i]))
        w[i] = (entropy_density(u + e, gamma) - entropy_density(u - e, gamma)) / (2 * e[i])
    return w


def entropy_vars(u, gamma):
    """Closed form of dU/du for U = -rho s/(gamma-1)."""
    rho, vel, p = prim(u, gamma)
    s = np.log(p) - gamma * np.log(rho)
    b = rho / p
    return np.concatenate([[(gamma - s) / (gamma - 1.0) - 0.5 * b * vel @ vel], b * vel, [-b]])


def random_state(rng, dim, gamma=1.4, rho=(0.5, 2.0), vel=1.0, p=(0.5, 2.0)):
    r = rng.uniform(*rho)
    v = rng.uniform(-vel, vel, dim)
    pp = rng.uniform(*p)
    return np.concatenate([[r], r * v, [pp / (gamma - 1.0) + 0.5 * r * v @ v]])
