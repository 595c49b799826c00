"""The maximum characteristic speed max(|V| + c) over the nodes (the reference's max_char_speed,
ChandrashekarFlux.cpp:61-65; DGSEMOperator.cpp:585-589): per-wave maxima reduced by one
workgroup, from prandtl_op_max_char_speed and after each Mult, against the state's own nodes."""
import numpy as np
import pytest

from prandtl_amd import capi, cases
from prandtl_amd import mesh as M

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _numpy_speed(u, gamma=1.4):
    dim = u.shape[0] - 2
    rho = u[0]
    vel = [u[1 + d] / rho for d in range(dim)]
    vsq = sum(v * v for v in vel)
    p = (gamma - 1.0) * (u[dim + 1] - 0.5 * rho * vsq)
    return float(np.max(np.sqrt(vsq) + np.sqrt(gamma * p / rho)))


# wave counts on both sides of the reduction's unrolled loop (8,192 entries and more)
@pytest.mark.parametrize("dim,n,order,viscous", [(2, 5, 3, False), (2, 64, 3, False), (3, 7, 2, False),
                                                 (3, 16, 3, False), (3, 22, 3, False), (3, 22, 3, True)])
def test_max_char_speed(device, dim, n, order, viscous):
    import torch

    _say(f"max_char_speed dim={dim} n={n} p={order} viscous={viscous}")
    m = M.periodic_lattice(dim, order, (n,) * dim)
    rng = np.random.default_rng(dim * 100 + n)
    if dim == 2:
        u = cases.vortex_state(M.periodic_lattice(2, order, (n, n), lengths=(10.0, 10.0)))
    else:
        u = cases.tgv_state(M.periodic_lattice(3, order, (n, n, n), lengths=(2 * np.pi,) * 3))
    u = u * (1 + 1e-3 * rng.uniform(-1, 1, u.shape))
    # one node above the rest, in a late element (the maximum comes from its wave): its x-velocity
    # raised by 2, the energy by the kinetic-energy change (the pressure kept)
    e, k = m.num_elements - 2, m.nodes_per_element // 2
    m_old = u[1, e, k]
    m_new = m_old + 2.0 * u[0, e, k]
    u[dim + 1, e, k] += 0.5 * (m_new * m_new - m_old * m_old) / u[0, e, k]
    u[1, e, k] = m_new
    ref = _numpy_speed(u)
    assert np.isfinite(ref) and ref == _numpy_speed(u[:, e:e + 1, k:k + 1])
    op = capi.Operator(m, viscous=viscous, mu=1e-3 if viscous else 0.0, Pr=0.71)
    d = torch.from_numpy(u).to(device)
    lam = op.max_char_speed(d)
    du = torch.empty_like(d)
    op.mult(d, du)
    lam_mult = op.get_max_char_speed()
    op.close()
    print(f"[wavespeed] {lam:.17e} vs numpy {ref:.17e}; after Mult {lam_mult:.17e}", flush=True)
    assert abs(lam - ref) <= 1e-14 * ref
    assert lam_mult == lam
