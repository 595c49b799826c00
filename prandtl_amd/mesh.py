"""Element meshes: conforming affine lattices, and general (curved, unstructured) meshes.

The operator library takes a mesh as arrays (see ``include/prandtl_amd.h``):

* ``neighbors[e, 2*d + s]`` -- element across local face (d, s) of element e,
  s = 0 on the xi_d = 0 side, s = 1 on the xi_d = 1 side; -(k+1) on a boundary face
  with boundary condition k;
* ``geom[e] = [det A, adj(A) row-major]`` -- the affine map x = x0 + A xi of the
  reference element [0,1]^d (MFEM's ``Weight()`` and ``AdjugateJacobian()``);
* on a lattice, the aligned-face convention: node (.., Np-1 in d, ..) of e coincides
  with node (.., 0 in d, ..) of ``neighbors[e, 2*d+1]``.

A general mesh adds ``node_geom[e, node] = [det J, adj(J) row-major]`` at every node
(the reference evaluates Tr.Weight() and Tr.AdjugateJacobian() per node,
src/Integrators/DGSEMIntegrator/DGSEMIntegrator.cpp:254-276) and
``face_orient[e, f]``: the neighbour's local face 2 d' + s' (bits 0-2), bit 3 flips
the first face coordinate, bit 4 the second, bit 5 swaps them -- the point (t1, t2)
of face f of e is the neighbour's point (u1, u2) with (a, b) = swap ? (t2, t1) :
(t1, t2), u1 = flip1 ? Np-1-a : a, u2 = flip2 ? Np-1-b : b (face coordinates are the
two other reference directions in increasing order).

Lattice element numbering is x-fastest (e = ex + nx*(ey + ny*ez)), node numbering is
the reference's lexicographic GLL order lex = i + Np*(j + Np*k)
(src/Integrators/DGSEMIntegrator/DGSEMIntegrator.cpp:250), and a state vector is
stored byNODES: u[v, e, node] (src/Simulation/Simulation.hpp:64).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

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
    n: Optional[tuple]              # lattice shape (None: general mesh)
    A: Optional[np.ndarray]         # lattice element map (None: general mesh)
    origin: Optional[np.ndarray]
    neighbors: np.ndarray           # int32 (NE, 2*dim)
    geom: np.ndarray                # float64 (NE, 1+dim*dim)
    node_geom: Optional[np.ndarray] = None    # float64 (NE, n, 1+dim*dim): per-node (det J, adj J)
    face_orient: Optional[np.ndarray] = None  # int32 (NE, 2*dim): face matching codes
    coords: Optional[np.ndarray] = None       # float64 (dim, NE, n): node coordinates
    hmin: Optional[float] = None

    @property
    def general(self) -> bool:
        return self.node_geom is not None or self.face_orient is not None

    @property
    def np1(self) -> int:
        return self.order + 1

    @property
    def nodes_per_element(self) -> int:
        return self.np1 ** self.dim

    @property
    def num_elements(self) -> int:
        return int(self.neighbors.shape[0])

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
        if self.coords is not None:
            return self.coords
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
        if self.node_geom is not None:
            return self.node_geom[:, :, 0] * W[None, :]
        return self.geom[:, 0][:, None] * W[None, :]

    def h_min(self) -> float:
        """Smallest singular value of the element map (MFEM GetElementSize(e, 1)); on a
        general mesh, of the Jacobians at the element centres."""
        if self.hmin is not None:
            return self.hmin
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


# ---------------------------------------------------------------------------
# General meshes: per-node metrics from node coordinates, faces matched by their corners.

def gll_derivative(np_: int) -> np.ndarray:
    """D[i, j] = l_j'(x_i) of the Lagrange basis on the GLL nodes of [0,1]
    (barycentric form; each row sums to zero)."""
    x, _ = gll_nodes(np_)
    c = np.array([np.prod([x[i] - x[j] for j in range(np_) if j != i]) for i in range(np_)])
    D = np.zeros((np_, np_))
    for i in range(np_):
        for j in range(np_):
            if i != j:
                D[i, j] = c[i] / (c[j] * (x[i] - x[j]))
        D[i, i] = -D[i].sum()
    return D


def reference_nodes(dim: int, np_: int) -> np.ndarray:
    """Reference coordinates of the element's nodes in lexicographic order, shape (dim, n)."""
    xg, _ = gll_nodes(np_)
    idx = np.indices((np_,) * dim).reshape(dim, -1)[::-1]   # lex: i fastest
    return xg[idx]


def _ref_derivatives(X: np.ndarray, np_: int) -> np.ndarray:
    """dX_r/dxi_c at the nodes by the GLL derivative matrix: (dim, dim, NE, n) [r][c]."""
    dim, NE, n = X.shape
    D = gll_derivative(np_)
    Xt = X.reshape((dim, NE) + (np_,) * dim)        # axes (.., k, j, i): reversed reference order
    J = np.zeros((dim, dim, NE, n))
    for c in range(dim):
        ax = 2 + (dim - 1 - c)
        Jc = np.moveaxis(np.tensordot(D, np.moveaxis(Xt, ax, 0), axes=(1, 0)), 0, ax)
        J[:, c] = Jc.reshape(dim, NE, n)
    return J


def node_geometry(X: np.ndarray, np_: int) -> np.ndarray:
    """Per-node (det J, adj J row-major), shape (NE, n, 1 + dim^2), of the element maps whose
    node coordinates are X (dim, NE, n): J[r][c] = dx_r/dxi_c of the degree-(Np-1) polynomial
    through the nodes (MFEM's Tr.Weight() and Tr.AdjugateJacobian() on a nodal mesh)."""
    dim, NE, n = X.shape
    J = np.moveaxis(_ref_derivatives(X, np_), (0, 1), (2, 3))   # (NE, n, dim, dim)
    if dim == 2:
        det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
        adj = np.empty_like(J)
        adj[..., 0, 0] = J[..., 1, 1]
        adj[..., 0, 1] = -J[..., 0, 1]
        adj[..., 1, 0] = -J[..., 1, 0]
        adj[..., 1, 1] = J[..., 0, 0]
    else:
        adj = np.empty_like(J)
        for r in range(3):
            for c in range(3):
                r1, r2 = [i for i in range(3) if i != c]
                c1, c2 = [i for i in range(3) if i != r]
                adj[..., r, c] = (-1) ** (r + c) * (J[..., r1, c1] * J[..., r2, c2] - J[..., r1, c2] * J[..., r2, c1])
        det = (J[..., 0, 0] * adj[..., 0, 0] + J[..., 0, 1] * adj[..., 1, 0] + J[..., 0, 2] * adj[..., 2, 0])
    G = np.empty((NE, n, 1 + dim * dim))
    G[..., 0] = det
    G[..., 1:] = adj.reshape(NE, n, dim * dim)
    return G


def face_nodes(dim: int, np_: int, f: int) -> np.ndarray:
    """Node indices of local face f = 2 d + s, shape (Nt2, Nt1): entry [t2, t1] is the face
    point (t1, t2), the face coordinates being the other reference directions in order."""
    d, s = divmod(f, 2)
    a = np_ - 1 if s else 0
    t = np.arange(np_)
    if dim == 2:
        i, j = (np.full(np_, a), t) if d == 0 else (t, np.full(np_, a))
        return (i + np_ * j)[None, :]
    t1, t2 = np.meshgrid(t, t, indexing="xy")       # [t2, t1]
    full = np.full_like(t1, a)
    ijk = {0: (full, t1, t2), 1: (t1, full, t2), 2: (t1, t2, full)}[d]
    return ijk[0] + np_ * (ijk[1] + np_ * ijk[2])


def _orient_map(dim: int, np_: int, code: int):
    """The neighbour's face point (u1, u2) of each point (t1, t2): arrays [t2, t1]."""
    Nt2 = np_ if dim == 3 else 1
    t2, t1 = np.meshgrid(np.arange(Nt2), np.arange(np_), indexing="ij")
    a, b = (t2, t1) if code & 32 else (t1, t2)
    u1 = np_ - 1 - a if code & 8 else a
    u2 = np_ - 1 - b if code & 16 else b
    if dim == 2:
        u2 = np.zeros_like(u1)
    return u1, u2


def _point_keys(P: np.ndarray, periods, tol: float):
    """Integer keys of points P (..., dim) equal for points equal up to tol, modulo periods."""
    keys = []
    for c in range(P.shape[-1]):
        x = P[..., c]
        L = None if periods is None else periods[c]
        if L:
            M = int(round(L / tol))
            keys.append(np.mod(np.round(np.mod(x, L) / tol).astype(np.int64), M))
        else:
            keys.append(np.round(x / tol).astype(np.int64))
    return np.stack(keys, axis=-1)


def _corner_index(dim: int, bits) -> int:
    """MFEM's (and Gmsh's) vertex index of the reference corner with coordinates ``bits``."""
    q = (0, 3, 1, 2)[2 * bits[0] + bits[1]]           # (0,0) 0, (0,1) 3, (1,0) 1, (1,1) 2
    return q + (4 * bits[2] if dim == 3 else 0)


def _face_corner_vertices(dim: int, f: int) -> np.ndarray:
    """Element vertex index at each corner (c1, c2) of local face f = 2 d + s: array [c2, c1]."""
    d, s = divmod(f, 2)
    others = [k for k in range(dim) if k != d]
    out = np.zeros((2 if dim == 3 else 1, 2), dtype=np.int64)
    for c2 in range(out.shape[0]):
        for c1 in range(2):
            bits = [0] * dim
            bits[d] = s
            bits[others[0]] = c1
            if dim == 3:
                bits[others[1]] = c2
            out[c2, c1] = _corner_index(dim, bits)
    return out


def general_mesh(dim: int, order: int, X: np.ndarray, periods=None, boundary=None, tol=None,
                 vertex_ids=None) -> Mesh:
    """A conforming mesh from the node coordinates X (dim, NE, n) of its elements (each a
    degree-``order`` polynomial map of [0,1]^dim with det J > 0).  Faces are matched by their
    corner points -- modulo ``periods`` (one length or None per direction) for periodic
    meshes -- and the orientation code of each matched face is the one that maps every face
    point onto its neighbour's.  An unmatched face is a boundary face; ``boundary(corners)``
    (corners: (2^(dim-1), dim) points) gives the index k of its condition (neighbour entry
    -(k+1)).  With ``vertex_ids`` (NE, 2^dim: each element's vertices in MFEM's order) faces are
    matched by their vertices instead, as MFEM does (periodic meshes then need no periods), and
    ``boundary`` receives the sorted tuple of the face's vertex ids.  Per-node metrics come from
    node_geometry."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    np_ = order + 1
    dimX, NE, n = X.shape
    assert dimX == dim and n == np_ ** dim
    G = node_geometry(X, np_)
    if (G[..., 0] <= 0).any():
        raise ValueError("element map with det J <= 0 at a node")
    if tol is None:
        span = float(np.max(X.max(axis=(1, 2)) - X.min(axis=(1, 2))))
        tol = 1e-9 * max(span, 1e-300)
    nf = 2 * dim
    FN = [face_nodes(dim, np_, f) for f in range(nf)]
    cidx = [0, np_ - 1]
    corners = [(FN[f][[0] if dim == 2 else cidx][:, cidx]) for f in range(nf)]   # [c2, c1] node ids
    P = np.moveaxis(X, 0, -1)                                                    # (NE, n, dim)
    table = {}
    topo = vertex_ids is not None
    if topo:
        VI = np.asarray(vertex_ids, dtype=np.int64)
        assert VI.shape == (NE, 2 ** dim)
        FCV = [_face_corner_vertices(dim, f) for f in range(nf)]                   # [c2, c1] element vertex index
    for f in range(nf):
        if topo:
            K = VI[:, FCV[f].reshape(-1)]                                           # (NE, ncorner) vertex ids
            for e in range(NE):
                table.setdefault(tuple(sorted(K[e].tolist())), []).append((e, f))
        else:
            K = _point_keys(P[:, corners[f].reshape(-1)], periods, tol)             # (NE, ncorner, dim)
            for e in range(NE):
                table.setdefault(tuple(sorted(map(tuple, K[e]))), []).append((e, f))
    nbr = np.zeros((NE, nf), dtype=np.int32)
    code = np.zeros((NE, nf), dtype=np.int32)
    cands = (0, 8, 16, 24, 32, 40, 48, 56) if dim == 3 else (0, 8)
    for key, lst in table.items():
        if len(lst) > 2:
            raise ValueError(f"{len(lst)} faces share the corners {key} (a periodic direction needs at least "
                             "3 elements for faces to be told apart by their corners)")
        if len(lst) == 1:
            e, f = lst[0]
            if boundary is None:
                raise ValueError(f"face {f} of element {e} has no neighbour and no boundary condition")
            k = int(boundary(key if topo else P[e, corners[f].reshape(-1)]))
            nbr[e, f] = -(k + 1)
            code[e, f] = f ^ 1
            continue
        for (e, f), (e2, f2) in (lst, lst[::-1]):
            found = None
            if topo:
                # corners map to corners: the code that sends each corner to the one with its vertex
                own = VI[e][FCV[f]]                                                  # [c2, c1]
                oth = VI[e2][FCV[f2]]
                for c in cands:
                    ok = True
                    for c2 in range(own.shape[0]):
                        for c1 in range(2):
                            a, b = (c2, c1) if c & 32 else (c1, c2)
                            u1 = 1 - a if c & 8 else a
                            u2 = (1 - b if c & 16 else b) if dim == 3 else 0
                            ok &= oth[u2, u1] == own[c2, c1]
                    if ok:
                        found = c
                        break
            else:
                pts = _point_keys(P[e, FN[f]], periods, tol)                      # (Nt2, Nt1, dim)
                for c in cands:
                    u1, u2 = _orient_map(dim, np_, c)
                    nb = FN[f2][u2, u1]
                    if np.array_equal(pts, _point_keys(P[e2, nb], periods, tol)):
                        found = c
                        break
            if found is None:
                raise ValueError(f"faces ({e}, {f}) and ({e2}, {f2}) share corners but not their points")
            nbr[e, f] = e2
            code[e, f] = f2 | found
    # element sizes: smallest singular value of J at the element centre
    D = gll_derivative(np_)
    xg, _ = gll_nodes(np_)
    lc = np.array([np.prod([(0.5 - xg[j]) / (xg[i] - xg[j]) for j in range(np_) if j != i]) for i in range(np_)])
    Jn = np.moveaxis(_ref_derivatives(X, np_), (0, 1), (2, 3))                   # (NE, n, dim, dim)
    W = lc
    for _ in range(dim - 1):
        W = np.multiply.outer(lc, W).reshape(-1)
    Jc = np.einsum("n,enrc->erc", W, Jn)
    hmin = float(np.linalg.svd(Jc, compute_uv=False).min())
    geom = np.ascontiguousarray(G[:, 0, :])          # node 0's values: the per-element summary
    return Mesh(dim, order, None, None, None, np.ascontiguousarray(nbr), geom, np.ascontiguousarray(G),
                np.ascontiguousarray(code), X, hmin)


def lattice_coords(dim: int, order: int, n, lengths=None, origin=None) -> np.ndarray:
    """Node coordinates (dim, NE, n) of the Cartesian lattice of periodic_lattice."""
    return periodic_lattice(dim, order, n, lengths=lengths, origin=origin).node_coords()


def multilinear_coords(dim: int, order: int, vertices: np.ndarray, cells: np.ndarray) -> np.ndarray:
    """Node coordinates (dim, NE, n) of bilinear quadrilaterals / trilinear hexahedra: cells
    (NE, 2^dim) lists each element's vertices in MFEM's (and Gmsh's) order -- counter-clockwise
    from (0,0) in 2D; the bottom face then the top face in 3D."""
    V = np.asarray(vertices, dtype=np.float64)[:, :dim]
    return cell_coords(dim, order, V[np.asarray(cells)])


# ---------------------------------------------------------------------------
# Gmsh 2.2 ASCII meshes (the reference reads them through MFEM, Simulation.cpp:272)

def read_gmsh(path: str) -> dict:
    """Parse a Gmsh 2.2 ASCII file: {'vertices': (nv, 3), 'cells': {type: (tags, vertex ids)},
    'names': {physical tag: name}}; vertex ids are 0-based."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    i = 0
    verts, names, ids = None, {}, {}
    cells = {}
    while i < len(lines):
        ln = lines[i]
        if ln == "$MeshFormat":
            ver = lines[i + 1].split()
            if not ver[0].startswith("2.") or ver[1] != "0":
                raise ValueError(f"{path}: only the Gmsh 2.x ASCII format is read (found {ver})")
            i += 3
        elif ln == "$PhysicalNames":
            cnt = int(lines[i + 1])
            for ln2 in lines[i + 2:i + 2 + cnt]:
                dim_, tag, name = ln2.split(maxsplit=2)
                names[int(tag)] = name.strip('"')
            i += cnt + 3
        elif ln == "$Nodes":
            cnt = int(lines[i + 1])
            arr = np.array([ln2.split() for ln2 in lines[i + 2:i + 2 + cnt]], dtype=np.float64)
            ids = {int(t): k for k, t in enumerate(arr[:, 0])}
            verts = arr[:, 1:4]
            i += cnt + 3
        elif ln == "$Elements":
            cnt = int(lines[i + 1])
            for ln2 in lines[i + 2:i + 2 + cnt]:
                f = [int(v) for v in ln2.split()]
                etype, ntags = f[1], f[2]
                tags = f[3:3 + ntags]
                vs = [ids[v] for v in f[3 + ntags:]]
                cells.setdefault(etype, ([], []))
                cells[etype][0].append(tags[0] if tags else 0)
                cells[etype][1].append(vs)
            i += cnt + 3
        else:
            i += 1
    if verts is None:
        raise ValueError(f"{path}: no $Nodes section")
    return {"vertices": verts, "names": names,
            "cells": {t: (np.array(a, dtype=np.int64), np.array(b, dtype=np.int64)) for t, (a, b) in cells.items()}}


def cell_coords(dim: int, order: int, corners: np.ndarray) -> np.ndarray:
    """Node coordinates (dim, NE, n) of bilinear / trilinear elements given by the coordinates
    of their corners, (NE, 2^dim, dim) in MFEM's (and Gmsh's) vertex order."""
    xi = reference_nodes(dim, order + 1)                    # (dim, n)
    if dim == 2:
        x, y = xi
        N = np.stack([(1 - x) * (1 - y), x * (1 - y), x * y, (1 - x) * y])                     # (4, n)
    else:
        x, y, z = xi
        N = np.stack([(1 - x) * (1 - y) * (1 - z), x * (1 - y) * (1 - z), x * y * (1 - z), (1 - x) * y * (1 - z),
                      (1 - x) * (1 - y) * z, x * (1 - y) * z, x * y * z, (1 - x) * y * z])
    return np.einsum("vn,evd->den", N, np.asarray(corners, dtype=np.float64)[..., :dim])


def refine_quads(cells: np.ndarray, corners: np.ndarray, bfaces=None, levels: int = 1):
    """Uniform refinement of a quadrilateral mesh (MFEM's Mesh::UniformRefinement, which the
    reference applies ser_ref_levels times, Simulation.cpp:293-321): each quadrilateral becomes
    four, with new vertices at the edge midpoints and the centre.  cells (NE, 4) are vertex ids,
    corners (NE, 4, 2) the cells' own corner coordinates (so periodic meshes, whose vertex ids
    wrap, refine correctly; the sub-cells are the bilinear map's images of the four quarter
    squares); bfaces (optional) is a list of (k, v0, v1) boundary edges, split in two.
    Returns (cells, corners, bfaces)."""
    cells = np.asarray(cells, dtype=np.int64)
    corners = np.asarray(corners, dtype=np.float64)
    nv = int(cells.max()) + 1 if cells.size else 0
    for _ in range(levels):
        mids = {}

        def mid(a, b):
            nonlocal nv
            k = (a, b) if a < b else (b, a)
            if k not in mids:
                mids[k] = nv
                nv += 1
            return mids[k]

        NE = cells.shape[0]
        C = corners
        # bilinear images of the 3 x 3 points (0, 1/2, 1)^2 of every cell
        P = {}
        for i, x in enumerate((0.0, 0.5, 1.0)):
            for j, y in enumerate((0.0, 0.5, 1.0)):
                P[i, j] = ((1 - x) * (1 - y)) * C[:, 0] + (x * (1 - y)) * C[:, 1] + (x * y) * C[:, 2] + \
                          ((1 - x) * y) * C[:, 3]
        newc = np.zeros((4 * NE, 4), dtype=np.int64)
        newx = np.zeros((4 * NE, 4, C.shape[2]))
        for e in range(NE):
            v0, v1, v2, v3 = (int(v) for v in cells[e])
            m01, m12, m23, m30 = mid(v0, v1), mid(v1, v2), mid(v2, v3), mid(v3, v0)
            c = nv
            nv += 1
            ids = {(0, 0): v0, (1, 0): m01, (2, 0): v1, (0, 1): m30, (1, 1): c, (2, 1): m12,
                   (0, 2): v3, (1, 2): m23, (2, 2): v2}
            for q, (i0, j0) in enumerate(((0, 0), (1, 0), (1, 1), (0, 1))):
                quad = [(i0, j0), (i0 + 1, j0), (i0 + 1, j0 + 1), (i0, j0 + 1)]
                newc[4 * e + q] = [ids[t] for t in quad]
                newx[4 * e + q] = [P[t][e] for t in quad]
        cells, corners = newc, newx
        if bfaces is not None:
            bfaces = [piece for k, a, b in bfaces for piece in ((k, a, mid(a, b)), (k, mid(a, b), b))]
    return cells, corners, bfaces


def _bc_lookup(bfaces, bc_index, where):
    """boundary callable of general_mesh (vertex ids): boundary face -> condition index"""
    table = {tuple(sorted(int(v) for v in vs)): name for name, *vs in bfaces}
    bc_index = bc_index or {}

    def boundary(key):
        if key not in table:
            raise ValueError(f"{where}: a boundary face with vertices {key} has no boundary element")
        name = table[key]
        if name not in bc_index:
            raise ValueError(f"{where}: no boundary condition for the boundary group {name!r}")
        return bc_index[name]

    return boundary


def _orient_cells(dim, cells, corners):
    """Reflect 2D cells whose vertex order gives det J < 0 (MFEM's
    Mesh::CheckElementOrientation(true) exchanges vertices 1 and 3)."""
    if dim != 2:
        return cells, corners
    a = corners[:, 1] - corners[:, 0]
    b = corners[:, 3] - corners[:, 0]
    neg = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0] < 0
    cells, corners = cells.copy(), corners.copy()
    cells[neg, 1], cells[neg, 3] = cells[neg, 3].copy(), cells[neg, 1].copy()
    corners[neg, 1], corners[neg, 3] = corners[neg, 3].copy(), corners[neg, 1].copy()
    return cells, corners


def _assemble(dim, order, cells, corners, bfaces, bc_index, refine, curve, where):
    cells, corners = _orient_cells(dim, np.asarray(cells, dtype=np.int64), np.asarray(corners, dtype=np.float64))
    if refine:
        if dim != 2:
            raise ValueError("uniform refinement is implemented for quadrilateral meshes")
        cells, corners, bfaces = refine_quads(cells, corners, bfaces, refine)
    X = cell_coords(dim, order, corners)
    if curve is not None:
        X = np.asarray(curve(X.reshape(dim, -1)), dtype=np.float64).reshape(X.shape)
    return general_mesh(dim, order, X, boundary=_bc_lookup(bfaces or [], bc_index, where), vertex_ids=cells)


def gmsh_mesh(path: str, order: int, bc_index=None, curve=None, refine: int = 0) -> Mesh:
    """A general mesh of the linear quadrilaterals (Gmsh type 3; 2D) or hexahedra (type 5; 3D)
    in a Gmsh 2.2 file.  Boundary faces take the physical group of the line (type 1) or
    quadrilateral (type 3, 3D) boundary element on them: ``bc_index`` maps the group's name
    (or tag) to the condition index k.  Elements whose vertex order gives det J < 0 are
    reflected (vertices 1 and 3 exchanged; MFEM's Mesh::CheckElementOrientation(true) does the
    same).  ``refine``: levels of uniform refinement (the reference's ser_ref_levels).
    ``curve`` (optional) maps node coordinates (dim, m) to curved ones."""
    g = read_gmsh(path)
    cells = g["cells"]
    if 5 in cells:
        dim, vol, bnd = 3, 5, 3
    elif 3 in cells:
        dim, vol, bnd = 2, 3, 1
    else:
        raise ValueError(f"{path}: no quadrilateral or hexahedral elements")
    V = g["vertices"][:, :dim]
    elems = cells[vol][1]
    bfaces = []
    if bnd in cells:
        for tag, vv in zip(*cells[bnd]):
            bfaces.append((g["names"].get(int(tag), int(tag)), *[int(v) for v in vv]))
    return _assemble(dim, order, elems, V[elems], bfaces, bc_index, refine, curve, path)


# ---------------------------------------------------------------------------
# MFEM mesh v1.0 files (the reference's 1D/2D generated cases: IsentropicVortex.mesh,
# KelvinHelmholtzInstability.mesh, read at Simulation.cpp:272)

def read_mfem_mesh(path: str) -> dict:
    """Parse an MFEM mesh v1.0 ASCII file: {'dim', 'cells' (NE, 2^dim vertex ids), 'attributes',
    'boundary' [(attribute, vertex ids...)], 'corners' (NE, 2^dim, dim)}.  Corner coordinates
    come from the vertices, or, for a periodic mesh, from its 'nodes' section (a first-order L2
    Gauss-Lobatto grid function: each element's own corner coordinates in lexicographic order)."""
    with open(path) as fh:
        toks = [ln.split("#")[0].strip() for ln in fh]
    lines = [t for t in toks if t]
    if not lines or not lines[0].startswith("MFEM mesh v1"):
        raise ValueError(f"{path}: not an MFEM mesh v1.x file")
    i, dim, cells, attrs, bnd, verts, nodes = 1, None, None, None, [], None, None
    geom_nv = {1: 2, 3: 4, 5: 8}
    while i < len(lines):
        key = lines[i]
        if key == "dimension":
            dim = int(lines[i + 1])
            i += 2
        elif key == "elements":
            n = int(lines[i + 1])
            rows = [[int(v) for v in lines[i + 2 + k].split()] for k in range(n)]
            for r in rows:
                if r[1] not in (3, 5) or len(r) != 2 + geom_nv[r[1]]:
                    raise ValueError(f"{path}: only quadrilaterals and hexahedra are read (geometry {r[1]})")
            attrs = np.array([r[0] for r in rows], dtype=np.int64)
            cells = np.array([r[2:] for r in rows], dtype=np.int64)
            i += 2 + n
        elif key == "boundary":
            n = int(lines[i + 1])
            for k in range(n):
                r = [int(v) for v in lines[i + 2 + k].split()]
                bnd.append((r[0], *r[2:]))
            i += 2 + n
        elif key == "vertices":
            nv = int(lines[i + 1])
            if i + 2 < len(lines) and lines[i + 2] == "nodes":
                i += 2
                continue
            vdim = int(lines[i + 2])
            verts = np.array([[float(v) for v in lines[i + 3 + k].split()] for k in range(nv)])[:, :vdim]
            i += 3 + nv
        elif key == "nodes":
            hdr = {}
            j = i + 1
            while ":" in lines[j] or lines[j] == "FiniteElementSpace":
                if ":" in lines[j]:
                    a, b = lines[j].split(":", 1)
                    hdr[a.strip()] = b.strip()
                j += 1
            fec = hdr.get("FiniteElementCollection", "")
            if not (fec.startswith("L2_T1_") and fec.endswith("_P1")) and not fec.startswith("L2_T"):
                raise ValueError(f"{path}: nodes in {fec!r} are not read (first-order L2 only)")
            if not fec.endswith("_P1"):
                raise ValueError(f"{path}: curved (order > 1) MFEM node grids are not read")
            vdim = int(hdr.get("VDim", dim))
            ordering = int(hdr.get("Ordering", 0))
            vals = np.array([float(v) for ln in lines[j:] for v in ln.split()])
            nodes = (vals, vdim, ordering)
            i = len(lines)
        else:
            i += 1
    if dim is None or cells is None:
        raise ValueError(f"{path}: missing dimension or elements")
    NE, nc = cells.shape
    lex_to_corner = [0, 1, 3, 2] if dim == 2 else [0, 1, 3, 2, 4, 5, 7, 6]
    if nodes is not None:
        vals, vdim, ordering = nodes
        total = NE * nc
        if vals.size < total * vdim:
            raise ValueError(f"{path}: nodes section too short")
        vals = vals[:total * vdim]
        arr = vals.reshape(total, vdim) if ordering == 1 else vals.reshape(vdim, total).T
        lex = arr.reshape(NE, nc, vdim)
        corners = np.zeros_like(lex)
        for L_, c in enumerate(lex_to_corner):
            corners[:, c] = lex[:, L_]
    else:
        corners = verts[cells]
    return {"dim": dim, "cells": cells, "attributes": attrs, "boundary": bnd, "corners": corners[..., :dim]}


def mfem_mesh(path: str, order: int, bc_index=None, refine: int = 0, curve=None) -> Mesh:
    """A general mesh from an MFEM mesh v1.0 file (quadrilaterals or hexahedra; periodic meshes
    through their first-order L2 node grid).  ``bc_index`` maps boundary attributes to condition
    indices; ``refine``: levels of uniform refinement (quadrilaterals)."""
    g = read_mfem_mesh(path)
    return _assemble(g["dim"], order, g["cells"], g["corners"], list(g["boundary"]), bc_index, refine, curve, path)
