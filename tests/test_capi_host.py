"""C-ABI checks that need no GPU: the library loads, exports every symbol the public
header declares, and rejects invalid meshes and arguments before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

from prandtl_amd import capi, mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "prandtl_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prandtl_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert header_symbols() == sorted(capi.EXPORTED)


def test_library_exports_header_symbols():
    lib = capi.load()
    for name in header_symbols():
        assert hasattr(lib, name), name


def _create(m, **kw):
    lib = capi.load()
    nbr = np.ascontiguousarray(m.neighbors, dtype=np.int32)
    geom = np.ascontiguousarray(m.geom, dtype=np.float64)
    md = capi.MeshDesc(m.dim, m.order, m.num_elements, nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                       geom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    ph = capi.Physics(kw.get("gamma", 1.4), kw.get("Pr", 0.72), kw.get("mu", 0.0), 287.0, kw.get("viscous", 0), 0)
    op = capi.Opts(kw.get("surface_flux", 0), kw.get("precision", 0), kw.get("bug_compat", 0),
                   kw.get("stage_hooks", 0))
    h = ctypes.c_void_p()
    rc = lib.prandtl_op_create(ctypes.byref(md), ctypes.byref(ph), ctypes.byref(op), ctypes.byref(h))
    return rc, lib.prandtl_last_error().decode(), h


def test_rejects_non_reciprocal_neighbours():
    m = M.periodic_lattice(2, 3, (4, 4), lengths=(1, 1))
    m.neighbors[5, 1] = 7
    rc, msg, h = _create(m)
    assert rc == -4 and "reciprocal" in msg and not h.value


def test_rejects_missing_neighbour():
    m = M.periodic_lattice(3, 2, (2, 2, 2), lengths=(1, 1, 1))
    m.neighbors[0, 0] = -1
    rc, msg, _ = _create(m)
    assert rc == -4 and "neighbour" in msg


def test_rejects_nonconforming_faces_and_bad_jacobian():
    m = M.periodic_lattice(2, 3, (4, 4), lengths=(1, 1))
    m.geom[3, 1] *= 1.5
    rc, msg, _ = _create(m)
    assert rc == -4 and "non-conforming" in msg
    m = M.periodic_lattice(2, 3, (4, 4), lengths=(1, 1))
    m.geom[2, 0] = -1.0
    rc, msg, _ = _create(m)
    assert rc == -4 and "det J" in msg


@pytest.mark.parametrize("kw,code", [({"gamma": 1.0}, -1), ({"surface_flux": 9}, -1), ({"precision": 7}, -1),
                                     ({"viscous": 1, "bug_compat": 1, "precision": 1}, -2),
                                     ({"viscous": 1, "bug_compat": 2, "precision": 1}, -2), ({"bug_compat": 8}, -1),
                                     ({"viscous": 1, "Pr": 0.0}, -1), ({"stage_hooks": 8}, -1)])
def test_rejects_bad_options(kw, code):
    m = M.periodic_lattice(2, 3, (2, 2), lengths=(1, 1))
    rc, msg, _ = _create(m, **kw)
    assert rc == code, msg


def test_rejects_unsupported_order():
    m = M.periodic_lattice(2, 7, (2, 2), lengths=(1, 1))
    rc, msg, _ = _create(m)
    assert rc == -2 and "order" in msg


def test_null_arguments():
    lib = capi.load()
    assert lib.prandtl_op_mult(None, None, None, None) == -1
    assert lib.prandtl_rk_run(None, 0, None, None, 0.1, 1, 0, None) == -1
    assert lib.prandtl_op_size(None) == 0
    lib.prandtl_op_destroy(None)


def test_cpp_mirror_compiles_and_links(tmp_path):
    """The C++ mirror header compiles with g++ alone and the example links against the library."""
    import subprocess

    from tests.test_gpu_host_mirror import build_example

    exe = build_example(str(tmp_path))
    assert os.path.exists(exe)


def _create_bc(m, bcs, **kw):
    lib = capi.load()
    nbr = np.ascontiguousarray(m.neighbors, dtype=np.int32)
    geom = np.ascontiguousarray(m.geom, dtype=np.float64)
    arr = (capi.Bc * len(bcs))(*bcs)
    md = capi.MeshDesc(m.dim, m.order, m.num_elements, nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                       geom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), kw.pop("num_bc", len(bcs)), arr)
    ph = capi.Physics(1.4, 0.72, kw.get("mu", 0.0), 287.0, kw.get("viscous", 0), 0)
    op = capi.Opts(0, kw.get("precision", 0), 0, 0)
    h = ctypes.c_void_p()
    rc = lib.prandtl_op_create(ctypes.byref(md), ctypes.byref(ph), ctypes.byref(op), ctypes.byref(h))
    return rc, lib.prandtl_last_error().decode(), h


def test_boundary_condition_validation():
    m = M.periodic_lattice(2, 3, (4, 3), lengths=(1, 1), periodic=(True, False), boundary={(1, 0): 0, (1, 1): 1})
    wall = capi.bc(capi.BC_SLIP_WALL)
    # a face that names a condition the table does not hold
    rc, msg, h = _create_bc(m, [wall], num_bc=1)
    assert rc == -4 and "boundary condition" in msg and not h.value
    # an unknown type
    rc, msg, h = _create_bc(m, [wall, capi.bc(99)])
    assert rc == -1 and "unknown boundary-condition type" in msg
    # Navier-Stokes on slip walls (the reference's base class aborts there)
    rc, msg, h = _create_bc(m, [wall, wall], viscous=1, mu=1e-3)
    assert rc == -2 and "slip and symmetry" in msg
    # mixed precision
    rc, msg, h = _create_bc(m, [wall, wall], precision=capi.MIXED)
    assert rc == -2 and "FP64" in msg


def _create_general(m, node_geom, face_orient, bcs=(), **kw):
    lib = capi.load()
    nbr = np.ascontiguousarray(m.neighbors, dtype=np.int32)
    arr = (capi.Bc * max(len(bcs), 1))(*bcs)
    ng = None if node_geom is None else np.ascontiguousarray(node_geom, dtype=np.float64)
    fo = None if face_orient is None else np.ascontiguousarray(face_orient, dtype=np.int32)
    md = capi.MeshDesc(m.dim, m.order, m.num_elements, nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                       None, len(bcs), arr if bcs else None,
                       None if ng is None else ng.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                       None if fo is None else fo.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    ph = capi.Physics(1.4, 0.72, 0.0, 287.0, 0, 0)
    op = capi.Opts(0, kw.get("precision", 0), 0, 0)
    h = ctypes.c_void_p()
    rc = lib.prandtl_op_create(ctypes.byref(md), ctypes.byref(ph), ctypes.byref(op), ctypes.byref(h))
    if h.value:
        lib.prandtl_op_destroy(h)
    return rc, lib.prandtl_last_error().decode()


def test_general_mesh_validation():
    """Mesh checks of prandtl_op_create that run before any device call: a valid general mesh
    passes them (creation then succeeds, or fails at device allocation without a GPU)."""
    from . import meshgen as G

    m, _ = G.rotate_frames(G.curved_periodic(2, 3, 4), seed=5, periods=[1.0, 1.0])
    rc, msg = _create_general(m, m.node_geom, m.face_orient)
    assert rc == 0 or (rc == -3 and "device" in msg), msg
    # flipped codes on every face: the two sides' point maps no longer meet
    rc, msg = _create_general(m, m.node_geom, m.face_orient ^ 8)
    assert rc == -4 and ("non-conforming" in msg or "invert" in msg), msg
    # a face that names the wrong face of its neighbour
    w = m.face_orient.copy()
    e = int(np.argmax(m.neighbors[:, 0] >= 0))
    w[e, 0] = (w[e, 0] & ~7) | ((w[e, 0] & 7) ^ 1)
    rc, msg = _create_general(m, m.node_geom, w)
    assert rc == -4 and "reciprocal" in msg, msg
    # swap / second-coordinate bits in 2D
    w = m.face_orient.copy()
    w[e, 0] |= 32
    rc, msg = _create_general(m, m.node_geom, w)
    assert rc == -4 and "orientation code" in msg, msg
    # a node with det J <= 0
    ng = m.node_geom.copy()
    ng[0, 5, 0] = -1.0
    rc, msg = _create_general(m, ng, m.face_orient)
    assert rc == -4 and "det J" in msg, msg
    # mixed precision
    rc, msg = _create_general(m, m.node_geom, m.face_orient, precision=capi.MIXED)
    assert rc == -2 and "FP64" in msg, msg
    # 3D: swapped and flipped faces pass
    m3, _ = G.rotate_frames(G.perturbed_multilinear(3, 3, 3), seed=6, periods=[1.0, 1.0, 1.0])
    assert ((m3.face_orient >> 5) & 1).any()
    rc, msg = _create_general(m3, m3.node_geom, m3.face_orient)
    assert rc == 0 or (rc == -3 and "device" in msg), msg


class _SizedOperator(capi.Operator):
    """An Operator without a device handle whose state holds `n` values: the binding's argument
    checks run before any call into the library."""

    def __init__(self, mesh, n):
        self.mesh = mesh
        self._h = None
        self._n = n

    @property
    def size(self):
        return self._n


def test_binding_rejects_mis_sized_buffers():
    import torch

    m = M.periodic_lattice(3, 2, (2, 2, 2), lengths=(1, 1, 1))
    n = 5 * m.num_elements * 27
    op = _SizedOperator(m, n)
    good, short = torch.zeros(n, dtype=torch.float64), torch.zeros(n - 1, dtype=torch.float64)
    calls = [lambda: op.mult(short, good), lambda: op.mult(good, short), lambda: op.rk_step(1, short, 0.0, 1e-3),
             lambda: op.rk_run(1, short, 0.0, 1e-3, 2), lambda: op.vjp(good, short, good),
             lambda: op.filter_apply(short), lambda: op.limiter_apply(short), lambda: op.filter_constraints(short),
             lambda: op.max_char_speed(short), lambda: op.rk_adjoint(short),
             lambda: op.mult_grad(good, good, torch.zeros(n, dtype=torch.float64))]   # grad holds dim x size
    for c in calls:
        with pytest.raises(capi.PrandtlError, match="the operator's state holds"):
            c()
    # right size, wrong place or type: refused too (no assert that python -O strips)
    with pytest.raises(capi.PrandtlError, match="float64 device tensor"):
        op.mult(good, good.clone())
    with pytest.raises(capi.PrandtlError, match="expected a torch.Tensor"):
        op.mult(np.zeros(n), good)
