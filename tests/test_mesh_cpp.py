"""The C++ mirror's mesh builder (include/prandtl_amd_mesh.hpp) against prandtl_amd/mesh.py
(CPU): the same neighbours and face codes, and the same per-node geometry to rounding, for the
reference's Gmsh meshes (cavity, refined once; forward-facing step, refined once: flipped faces,
faces meeting in other directions), the periodic MFEM vortex mesh refined three times, and a
Gmsh file of hexahedra."""
import os
import subprocess

import numpy as np
import pytest

from prandtl_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MESHES = os.path.join(os.path.dirname(HERE), "data", "meshes")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("meshcpp")
    out = str(d / "mesh_builder_example")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "native", "mesh_builder_example.cpp"), "-o", out], check=True)
    return out


def _run(exe, tmp_path, kind, path, order, refine, groups):
    prefix = str(tmp_path / "m")
    args = [exe, kind, path, str(order), str(refine), prefix] + [f"{k}={v}" for k, v in groups.items()]
    r = subprocess.run(args, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    NE, dim = (int(x) for x in r.stdout.split())
    nbr = np.fromfile(prefix + ".nbr", dtype=np.int32).reshape(NE, 2 * dim)
    fo = np.fromfile(prefix + ".fo", dtype=np.int32).reshape(NE, 2 * dim)
    ng = np.fromfile(prefix + ".ng", dtype=np.float64).reshape(NE, (order + 1) ** dim, 1 + dim * dim)
    return nbr, fo, ng


def _compare(py, cpp):
    nbr, fo, ng = cpp
    assert np.array_equal(nbr, py.neighbors)
    inter = nbr >= 0
    assert np.array_equal(fo[inter], py.face_orient[inter])
    assert np.array_equal(fo[~inter] & 7, py.face_orient[~inter] & 7)
    scale = np.abs(py.node_geom).max()
    assert np.abs(ng - py.node_geom).max() <= 1e-13 * scale


@pytest.mark.parametrize("name,kind,order,refine,groups", [
    ("LidDrivenCavity.msh", "gmsh", 3, 1, {"left": 0, "right": 0, "bottom": 0, "lid": 1}),
    ("ForwardFacingStep.msh", "gmsh", 2, 1, {"left": 0, "right": 1, "top": 2, "bottom": 2}),
    ("IsentropicVortex.mesh", "mfem", 3, 3, {}),
    ("KelvinHelmholtzInstability.mesh", "mfem", 1, 2, {}),
])
def test_cpp_builder_matches_python(exe, tmp_path, name, kind, order, refine, groups):
    path = os.path.join(MESHES, name)
    py = (M.gmsh_mesh(path, order, groups, refine=refine) if kind == "gmsh"
          else M.mfem_mesh(path, order, None, refine=refine))
    _compare(py, _run(exe, tmp_path, kind, path, order, refine, groups))


def test_cpp_builder_hexahedra(exe, tmp_path):
    n = 3
    vid = lambda i, j, k: 1 + i + (n + 1) * (j + (n + 1) * k)
    rng = np.random.default_rng(4)
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "2", '2 1 "walls"', '3 2 "fluid"',
             "$EndPhysicalNames", "$Nodes", str((n + 1) ** 3)]
    for k in range(n + 1):
        for j in range(n + 1):
            for i in range(n + 1):
                p = np.array([i, j, k], float) / n
                inner = [0 < c < n for c in (i, j, k)]
                d = rng.uniform(-0.08, 0.08, 3) / n * np.array(inner)
                x = p + d
                lines.append(f"{vid(i, j, k)} {float(x[0])!r} {float(x[1])!r} {float(x[2])!r}")
    lines.append("$EndNodes")
    elems = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                elems.append((5, 2, [vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k),
                                     vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1),
                                     vid(i, j + 1, k + 1)]))
    for a in range(n):
        for b in range(n):
            for q in ([vid(0, a, b), vid(0, a + 1, b), vid(0, a + 1, b + 1), vid(0, a, b + 1)],
                      [vid(n, a, b), vid(n, a + 1, b), vid(n, a + 1, b + 1), vid(n, a, b + 1)],
                      [vid(a, 0, b), vid(a + 1, 0, b), vid(a + 1, 0, b + 1), vid(a, 0, b + 1)],
                      [vid(a, n, b), vid(a + 1, n, b), vid(a + 1, n, b + 1), vid(a, n, b + 1)],
                      [vid(a, b, 0), vid(a + 1, b, 0), vid(a + 1, b + 1, 0), vid(a, b + 1, 0)],
                      [vid(a, b, n), vid(a + 1, b, n), vid(a + 1, b + 1, n), vid(a, b + 1, n)]):
                elems.append((3, 1, q))
    lines += ["$Elements", str(len(elems))]
    for k, (t, tag, v) in enumerate(elems):
        lines.append(f"{k + 1} {t} 2 {tag} {tag} " + " ".join(map(str, v)))
    lines.append("$EndElements")
    path = str(tmp_path / "box.msh")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    py = M.gmsh_mesh(path, 3, {"walls": 0})
    _compare(py, _run(exe, tmp_path, "gmsh", path, 3, 0, {"walls": 0}))
