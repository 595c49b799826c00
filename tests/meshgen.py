"""General test meshes: curved periodic lattices, perturbed multilinear lattices, and element
frames rotated by the symmetries of the reference square / cube (shared by the oracle and GPU
geometry tests)."""
import itertools

import numpy as np

from prandtl_amd import mesh as M


def curved_periodic(dim, order, n, L=1.0, eps=0.04):
    """The Cartesian lattice of [0, L]^dim with every node moved by a smooth periodic map
    x_c += eps L prod_d sin(2 pi x_d / L + phase_{c,d}); elements are degree-``order`` curved."""
    X = M.lattice_coords(dim, order, (n,) * dim, lengths=[L] * dim)
    Y = X.copy()
    for c in range(dim):
        s = np.ones_like(X[0])
        for d in range(dim):
            s = s * np.sin(2 * np.pi * X[d] / L + 0.7 * c + 0.3 * d)
        Y[c] = X[c] + eps * L * s
    return M.general_mesh(dim, order, Y, periods=[L] * dim)


def perturbed_multilinear(dim, order, n, L=1.0, eps=0.12, seed=3, periodic=True, boundary=None):
    """Bi/trilinear elements of a lattice of [0, L]^dim whose vertices are moved at random by up
    to eps h (periodic: the images of a vertex move together; else the boundary vertices keep
    their place along the boundary normal)."""
    rng = np.random.default_rng(seed)
    h = L / n
    grid = np.stack(np.meshgrid(*([np.arange(n + 1)] * dim), indexing="ij"), axis=-1)   # [i0, i1, (i2)] -> idx
    base = grid * h
    if periodic:
        delta = rng.uniform(-eps * h, eps * h, (n,) * dim + (dim,))
        idx = tuple(np.mod(grid[..., d], n) for d in range(dim))
        V = base + delta[idx]
    else:
        delta = rng.uniform(-eps * h, eps * h, base.shape)
        for d in range(dim):
            edge = (grid[..., d] == 0) | (grid[..., d] == n)
            delta[..., d][edge] = 0.0
        V = base + delta
    vid = np.arange((n + 1) ** dim).reshape((n + 1,) * dim)
    cells = []
    for e in itertools.product(range(n), repeat=dim):
        e = e[::-1]   # x fastest
        if dim == 2:
            i, j = e
            cells.append([vid[i, j], vid[i + 1, j], vid[i + 1, j + 1], vid[i, j + 1]])
        else:
            i, j, k = e
            cells.append([vid[i, j, k], vid[i + 1, j, k], vid[i + 1, j + 1, k], vid[i, j + 1, k],
                          vid[i, j, k + 1], vid[i + 1, j, k + 1], vid[i + 1, j + 1, k + 1], vid[i, j + 1, k + 1]])
    X = M.multilinear_coords(dim, order, V.reshape(-1, dim), np.array(cells))
    return M.general_mesh(dim, order, X, periods=[L] * dim if periodic else None, boundary=boundary)


def rotations(dim):
    """The proper symmetries of the reference square / cube as (perm, flip): reference
    coordinate c of the new frame is xi_perm[c] (1 - that when flip[c])."""
    out = []
    for perm in itertools.permutations(range(dim)):
        for flip in itertools.product((0, 1), repeat=dim):
            P = np.zeros((dim, dim))
            for c in range(dim):
                P[c, perm[c]] = -1.0 if flip[c] else 1.0
            if np.linalg.det(P) > 0:
                out.append((perm, flip))
    return out


def frame_permutations(dim, np_, rots):
    """For each element's rotation, the node permutation new_node -> old_node (NE, n)."""
    idx = np.indices((np_,) * dim).reshape(dim, -1)[::-1]     # idx[c] = the node's index along xi_c
    perms = []
    for perm, flip in rots:
        # the node at new indices a' sits where the old frame has indices a, a'_c = f(a_perm[c])
        old = np.zeros_like(idx)
        for c in range(dim):
            a = np_ - 1 - idx[c] if flip[c] else idx[c]
            old[perm[c]] = a
        lex = np.zeros(idx.shape[1], dtype=np.int64)
        for c in reversed(range(dim)):
            lex = lex * np_ + old[c]
        perms.append(lex)
    return np.array(perms)


def rotate_frames(mesh, seed=0, periods=None, boundary=None):
    """The same mesh with every element's reference frame rotated at random: returns the new
    mesh and the node permutation P (NE, n) with new node i of element e = old node P[e, i]."""
    rots = rotations(mesh.dim)
    rng = np.random.default_rng(seed)
    choice = rng.integers(0, len(rots), mesh.num_elements)
    table = frame_permutations(mesh.dim, mesh.np1, rots)
    P = table[choice]
    X = mesh.node_coords()
    Xn = np.take_along_axis(X, P[None], axis=2)
    return M.general_mesh(mesh.dim, mesh.order, Xn, periods=periods, boundary=boundary), P


def permute_state(u, P):
    return np.take_along_axis(u, np.broadcast_to(P[None], u.shape), axis=2)


def unpermute_state(u, P):
    out = np.empty_like(u)
    np.put_along_axis(out, np.broadcast_to(P[None], u.shape), u, axis=2)
    return out
