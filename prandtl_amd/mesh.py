"""Conforming affine element lattices with periodic connectivity.

The operator library takes a mesh as three arrays (see ``include/prandtl_amd.h``):

* ``neighbors[e, 2*d + s]`` -- element across local face (d, s) of element e,
  s = 0 on the xi_d = 0 side, s = 1 on the xi_d = 1 side;
* ``geom[e] = [det A, adj(A) row-major]`` -- the affine map x = x0 + A xi of the
  reference element [0,1]^d (MFEM's ``Weight()`` and ``AdjugateJacobian()``);
* the aligned-face convention: node (.., Np-1 in d, ..) of e coincides with node
  (.., 0 in d, ..) of ``neighbors[e, 2*d+1]``.

Element numbering is x-fastest (e = ex + nx*(ey + ny*ez)), node numbering is the
reference's lexicographic GLL order lex = i + Np*(j + Np*k)
(src/Integrators/DGSEMIntegrator/DGSEMIntegrator.cpp:250), and a state vector is
stored byNODES: u[v, e, node] (src/Simulation/Simulation.hpp:64).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def gll_nodes(np_: int):
    """GLL nodes and weights on [0,1]: 0, 1 and the roots of P'_N(2x-1);
    weights 1/(N(N+1) P_N(2x-1)^2)."""
    from numpy.polynomial import legendre as L

    N = np_ - 1
    cN = np.zeros(N + 1)
    cN[N] = 1.0
    interior = np.sort(np.real(L.legroots(L.legder(cN)))) if N > 1 else np.zeros(0)
    z = np.concatenate([[-1.0], interior, [1.0]])
    PN = L.legval(z, cN)
    w = 1.0 / (N * (N + 1) * PN**2)
    x = 0.5 * (z + 1.0)
    x[0], x[-1] = 0.0, 1.0
    return x, w


def adjugate(A: np.ndarray) -> np.ndarray:
    d = A.shape[0]
    if d == 2:
        return np.array([[A[1, 1], -A[0, 1]], [-A[1, 0], A[0, 0]]])
    C = np.zeros((3, 3))
    for r in range(3):
        for c in range(3):
            m = np.delete(np.delete(A, r, axis=0), c, axis=1)
            C[r, c] = (-1) ** (r + c) * (m[0, 0] * m[1, 1] - m[0, 1] * m[1, 0])
    return C.T  # adj = cofactor^T


@dataclass
class Mesh:
    dim: int
    order: int
    n: tuple
    A: np.ndarray
    origin: np.ndarray
    neighbors: np.ndarray  # int32 (NE, 2*dim)
    geom: np.ndarray       # float64 (NE, 1+dim*dim)

    @property
    def np1(self) -> int:
        return self.order + 1

    @property
    def nodes_per_element(self) -> int:
        return self.np1 ** self.dim

    @property
    def num_elements(self) -> int:
        return int(np.prod(self.n))

    @property
    def num_nodes(self) -> int:
        return self.num_elements * self.nodes_per_element

    @property
    def neq(self) -> int:
        return self.dim + 2

    @property
    def state_shape(self):
        return (self.neq, self.num_elements, self.nodes_per_element)

    def node_coords(self) -> np.ndarray:
        """Physical node coordinates, shape (dim, NE, n)."""
        xg, _ = gll_nodes(self.np1)
        Np, d = self.np1, self.dim
        # reference-coordinate index of each node
        idx = np.indices((Np,) * d).reshape(d, -1)[::-1]  # lex: i fastest
        xi = xg[idx]  # (d, n)
        e = np.arange(self.num_elements)
        eidx = []
        rem = e.copy()
        for k in range(d):
            eidx.append(rem % self.n[k])
            rem //= self.n[k]
        eidx = np.array(eidx, dtype=np.float64)  # (d, NE)
        ref = eidx[:, :, None] + xi[:, None, :]  # (d, NE, n)
        return self.origin[:, None, None] + np.einsum("ab,ben->aen", self.A, ref)

    def quadrature_weights(self) -> np.ndarray:
        """J * w_i w_j (w_k) per node, shape (NE, n): the DGSEM mass matrix diagonal."""
        _, wg = gll_nodes(self.np1)
        W = wg
        for _ in range(self.dim - 1):
            W = np.multiply.outer(wg, W).reshape(-1)  # slower index outer
        W = W.reshape(-1)
        return self.geom[:, 0][:, None] * W[None, :]

    def h_min(self) -> float:
        """Smallest singular value of the element map (MFEM GetElementSize(e, 1))."""
        return float(np.linalg.svd(self.A, compute_uv=False).min())


def periodic_lattice(dim: int, order: int, n, lengths=None, A=None, origin=None, periodic=None,
                     boundary=None) -> Mesh:
    """Lattice of identical affine elements.  With ``lengths`` the elements are
    axis-aligned boxes (Cartesian mesh of [origin, origin+L]); ``A`` gives a general
    (sheared) element map instead.  Every direction is periodic unless ``periodic``
    (one flag per direction) says otherwise; the faces on the lattice's side s of a
    non-periodic direction d get the neighbour entry -(k+1), k = boundary[(d, s)] the
    index of their boundary condition."""
    n = tuple(int(v) for v in n)
    assert len(n) == dim
    if A is None:
        lengths = np.asarray(lengths if lengths is not None else [1.0] * dim, dtype=np.float64)
        A = np.diag(lengths / np.asarray(n, dtype=np.float64))
    A = np.asarray(A, dtype=np.float64)
    origin = np.zeros(dim) if origin is None else np.asarray(origin, dtype=np.float64)
    NE = int(np.prod(n))
    e = np.arange(NE)
    coords = []
    rem = e.copy()
    for k in range(dim):
        coords.append(rem % n[k])
        rem //= n[k]
    coords = np.array(coords)
    strides = np.cumprod((1,) + n[:-1])
    periodic = tuple(True for _ in range(dim)) if periodic is None else tuple(bool(p) for p in periodic)
    assert len(periodic) == dim
    nbr = np.zeros((NE, 2 * dim), dtype=np.int32)
    for d in range(dim):
        for s, off in ((0, -1), (1, +1)):
            c = coords.copy()
            c[d] = (c[d] + off) % n[d]
            nbr[:, 2 * d + s] = (c * strides[:, None]).sum(axis=0)
            if not periodic[d]:
                k = boundary[(d, s)]
                edge = coords[d] == (0 if s == 0 else n[d] - 1)
                nbr[edge, 2 * d + s] = -(k + 1)
    det = float(np.linalg.det(A))
    adj = adjugate(A)
    geom = np.zeros((NE, 1 + dim * dim))
    geom[:, 0] = det
    geom[:, 1:] = adj.reshape(-1)[None, :]
    return Mesh(dim, order, n, A, origin, np.ascontiguousarray(nbr), np.ascontiguousarray(geom))
