"""The extended-precision build of the oracle (oracle.rhs_truth, long double throughout) as the
exact discrete operator (CPU).

* it is the same operator: on well-conditioned states it agrees with the double oracle within a
  few eps of each entry's rounding scale (the double assembly with every term in absolute value);
* free stream: the exact operator vanishes (the scheme is free-stream preserving on affine
  meshes); the extended evaluation's residue is the double one's scaled by the ratio of the unit
  roundoffs, 2^-11;
* it resolves errors far below what double evaluations leave: a 1e-10 relative mistake planted in
  one entry of the double oracle's result is measured within 10 %, while the unplanted double
  oracle is within its own build-to-build spread (with and without FMA) of it.
"""
import numpy as np
import pytest

import oracle
from prandtl_amd import cases, mesh as M

EPS = np.finfo(np.float64).eps


def _perturbed(mesh, seed):
    rng = np.random.default_rng(seed)
    X = mesh.node_coords()
    rho = 1.0 + 0.2 * np.sin(X[0]) + 0.05 * rng.uniform(-1, 1, X[0].shape)
    vel = [0.3 * np.cos(X[(d + 1) % mesh.dim]) + 0.05 * rng.uniform(-1, 1, X[0].shape) for d in range(mesh.dim)]
    p = 1.0 + 0.1 * np.cos(X[0] + X[1]) + 0.05 * rng.uniform(-1, 1, X[0].shape)
    return cases.prim_to_conserv(rho, vel, p)


@pytest.mark.parametrize("dim,order,viscous", [(2, 3, False), (2, 4, True), (3, 2, True), (3, 3, True), (3, 3, False)])
def test_extended_operator_is_the_operator(dim, order, viscous):
    A = np.array([[0.5, 0.1], [0.05, 0.45]]) if dim == 2 else \
        np.array([[0.5, 0.1, 0.0], [0.05, 0.45, 0.08], [0.02, -0.03, 0.55]])
    m = M.periodic_lattice(dim, order, (3, 4) if dim == 2 else (2, 3, 2), A=A)
    u = _perturbed(m, 3 + order)
    ph = oracle.phys(viscous=viscous, mu=0.01, Pr=0.7)
    ref, lam, mag = oracle.rhs_mag(m, ph, u)
    hi, lo, lam_ld = oracle.rhs_truth(m, ph, u)
    s = oracle.scaled_error(hi, ref, mag)
    assert s <= 32.0, s
    assert np.abs(hi - ref).max() <= 1e-13 * np.abs(ref).max()
    assert abs(lam_ld - lam) <= 4 * EPS * lam
    # the residual is below half an ulp of the rounded value
    assert (np.abs(lo) <= 0.5 * np.spacing(np.abs(hi)) + 1e-300).all()


def test_free_stream_is_exactly_zero():
    m = cases.tgv_mesh(n=4, order=3)
    u = np.empty((5, m.num_elements, m.nodes_per_element))
    u[0], u[1], u[2], u[3], u[4] = 1.0, 0.3, -0.2, 0.1, 71.4
    ph = oracle.phys(viscous=True, mu=1e-3, Pr=0.71)
    d, _ = oracle.rhs(m, ph, u)
    hi, lo, _ = oracle.rhs_truth(m, ph, u)
    assert np.abs(d).max() > 0.0          # the double evaluation leaves rounding residue
    # the residue scales with the unit roundoff: 2^-64 / 2^-53 = 2^-11 ~ 4.9e-4
    assert np.abs(hi).max() <= 2e-3 * np.abs(d).max() + 1e-300


def test_truth_resolves_a_planted_error():
    m = cases.tgv_mesh(n=6, order=3)
    u = cases.tgv_state(m)
    ph = oracle.phys(viscous=True, mu=1.0 / 1600.0, Pr=0.71)
    hi, lo, _ = oracle.rhs_truth(m, ph, u)
    d, _ = oracle.rhs(m, ph, u)
    dfma, _ = oracle.rhs(m, ph, u, fma=True)
    base = oracle.truth_error(d, hi, lo)
    spread = np.abs(d - dfma).max() / np.abs(d).max()
    assert base <= 4.0 * spread and base < 5e-12, (base, spread)
    planted = d.copy()
    k = np.unravel_index(np.argmax(np.abs(d[4])), d[4].shape)
    planted[(4,) + k] += 1e-10 * np.abs(d).max()
    err = oracle.truth_error(planted, hi, lo)
    assert abs(err - 1e-10) <= 0.1e-10 + base, err
