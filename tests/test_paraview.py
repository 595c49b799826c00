"""ParaView output (the reference's ParaViewDataCollection, Simulation.cpp:829-853, 980-985,
1038-1079): the VTK Lagrange point ordering against VTK's published linear-cell data, the
resampling of nodal data, round trips of the written files, the collection layout and the save
schedule."""
import os

import numpy as np
import pytest

from prandtl_amd import cases, config
from prandtl_amd import mesh as M
from prandtl_amd import paraview as P

HERE = os.path.dirname(os.path.abspath(__file__))
CFG = os.path.join(HERE, "golden", "configs")
MESHES = os.path.join(os.path.dirname(HERE), "data", "meshes")

# VTK's linear cells (vtkQuad, vtkHexahedron): corners as lattice coordinates, the edges as corner
# pairs (each edge runs from its first to its second corner) and the faces as corner quadruples
QUAD = [(0, 0), (1, 0), (1, 1), (0, 1)]
QUAD_EDGES = [(0, 1), (1, 2), (3, 2), (0, 3)]
HEX = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
HEX_EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5), (3, 7), (2, 6)]
HEX_FACES = [(0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2), (0, 3, 2, 1), (4, 5, 6, 7)]


def vtk_points(dim, n):
    """Lattice coordinates of the VTK points of a Lagrange cell of order n, in VTK order."""
    perm = P.lagrange_permutation(dim, n)
    R = n + 1
    return np.array([[(lex // R ** a) % R for a in range(dim)] for lex in perm])


def test_lagrange_low_orders_match_published_cells():
    # order 1: the linear cells' corners
    assert (vtk_points(2, 1) == np.array(QUAD)).all()
    assert (vtk_points(3, 1) == np.array(HEX)).all()
    # order 2 quadrilateral = VTK_BIQUADRATIC_QUAD: corners, the midpoints of (0,1), (1,2), (2,3),
    # (3,0), then the centre
    q = vtk_points(2, 2)
    corners = 2 * np.array(QUAD)
    mids = [(corners[a] + corners[b]) // 2 for a, b in [(0, 1), (1, 2), (2, 3), (3, 0)]]
    assert (q[:4] == corners).all() and (q[4:8] == np.array(mids)).all() and (q[8] == [1, 1]).all()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_lagrange_ordering_geometry(dim, n):
    """Corners, then each edge's interior points along the edge from its first corner, then each
    face's interior points (the lower free axis fastest), then the interior points x fastest."""
    pts = vtk_points(dim, n)
    corners = n * np.array(QUAD if dim == 2 else HEX)
    edges = QUAD_EDGES if dim == 2 else HEX_EDGES
    m = n - 1
    assert (pts[:len(corners)] == corners).all()
    k = len(corners)
    for a, b in edges:
        step = (corners[b] - corners[a]) // n
        want = np.array([corners[a] + s * step for s in range(1, n)])
        assert (pts[k:k + m] == want).all(), (a, b)
        k += m
    if dim == 3:
        for face in HEX_FACES:
            fc = corners[list(face)]
            fixed = [ax for ax in range(3) if (fc[:, ax] == fc[0, ax]).all()]
            assert len(fixed) == 1
            free = [ax for ax in range(3) if ax != fixed[0]]
            want = []
            for s2 in range(1, n):
                for s1 in range(1, n):
                    p = np.zeros(3, dtype=int)
                    p[fixed[0]] = fc[0, fixed[0]]
                    p[free[0]], p[free[1]] = s1, s2
                    want.append(p)
            assert (pts[k:k + m * m] == np.array(want)).all(), face
            k += m * m
    inner = np.array([[c[a] + 1 for a in range(dim)] for c in
                      (np.unravel_index(i, (m,) * dim)[::-1] for i in range(m ** dim))]).reshape(-1, dim)
    assert (pts[k:] == inner).all()
    assert k + m ** dim == (n + 1) ** dim


def test_resample_is_exact_for_the_element_polynomials():
    m = M.periodic_lattice(3, 3, (3, 3, 3), A=np.array([[0.5, 0.1, 0.0], [0.05, 0.45, 0.08], [0.02, -0.03, 0.55]]))
    X = m.node_coords()
    f = X[0] ** 3 + 2 * X[1] * X[2] - X[2] ** 2
    for lv in (1, 3, 5):
        Xe = P.resample(X, 3, 4, lv)
        fe = P.resample(f, 3, 4, lv)
        if lv >= 3:
            assert np.abs(fe - (Xe[0] ** 3 + 2 * Xe[1] * Xe[2] - Xe[2] ** 2)).max() < 1e-13
        # equispaced points of the affine element: origin + A (e + a / lv)
        assert np.abs(Xe[:, 0, 0] - m.origin).max() < 1e-15
        assert np.abs(Xe[:, 0, -1] - (m.origin + m.A @ np.ones(3))).max() < 1e-14


@pytest.mark.parametrize("fmt", ["binary", "ascii"])
@pytest.mark.parametrize("high_order", [True, False])
def test_vtu_round_trip(tmp_path, fmt, high_order):
    m = cases.vortex_mesh(n=4, order=3)
    u = cases.vortex_state(m)
    alpha = np.linspace(0.0, 0.5, m.num_elements)
    col = P.ParaViewDataCollection("ParaView", m, str(tmp_path), levels_of_detail=3, high_order=high_order, fmt=fmt)
    col.save_state(u, 1.4, alpha=alpha, cycle=7, t=0.25)
    d = P.read_vtu(os.path.join(str(tmp_path), "ParaView", "Cycle000007", "proc000000.vtu"))
    NE, R = m.num_elements, 4
    pts = P.resample(m.node_coords(), 2, 4, 3)
    assert d["points"].shape == (NE * R * R, 3)
    assert np.abs(d["points"][:, 0] - pts[0].ravel()).max() < 1e-14 and (d["points"][:, 2] == 0).all()
    assert np.abs(d["points"][:, 1] - pts[1].ravel()).max() < 1e-14
    prim = P.primitive_fields(u, 1.4)
    assert set(d) >= {"Density", "Horizontal V", "Vertical V", "Pressure", "Blending Coeff"}
    assert "Normal V" not in d
    for name, f in prim.items():
        assert np.abs(d[name] - P.resample(f, 2, 4, 3).ravel()).max() < 1e-12 * max(1.0, np.abs(f).max())
    assert (d["Blending Coeff"] == np.repeat(alpha, R * R)).all()
    if high_order:
        assert (d["types"] == P.VTK_LAGRANGE_QUADRILATERAL).all() and d["types"].size == NE
        assert (d["offsets"] == np.arange(1, NE + 1) * R * R).all()
        assert (d["connectivity"][:R * R] == P.lagrange_permutation(2, 3)).all()
    else:
        assert (d["types"] == P.VTK_QUAD).all() and d["types"].size == NE * 9
        assert (d["offsets"] == np.arange(1, NE * 9 + 1) * 4).all()


def test_linear_subcells_tile_the_domain(tmp_path):
    A = np.array([[0.5, 0.1, 0.0], [0.05, 0.45, 0.08], [0.02, -0.03, 0.55]])
    m = M.periodic_lattice(3, 3, (3, 3, 3), A=A)
    pts, conn, offsets, types = P.cells_of(m, 3, high_order=False)
    c = pts[conn.reshape(-1, 8)]
    # each hexahedron of this affine mesh is a parallelepiped spanned from corner 0 by corners 1, 3, 4
    vol = np.einsum("ij,ij->i", c[:, 1] - c[:, 0], np.cross(c[:, 3] - c[:, 0], c[:, 4] - c[:, 0]))
    assert (vol > 0).all()
    assert abs(vol.sum() - 27 * np.linalg.det(A)) < 1e-12
    assert (types == P.VTK_HEXAHEDRON).all() and types.size == 27 * 27


def test_collection_layout(tmp_path):
    m = cases.tgv_mesh(n=3, order=2)
    u = cases.tgv_state(m)
    col = P.ParaViewDataCollection("ParaView", m, str(tmp_path))
    col.save_state(u, 1.4, cycle=0, t=0.0)
    col.save_state(u, 1.4, cycle=100, t=0.5)
    root = os.path.join(str(tmp_path), "ParaView")
    for f in ("ParaView.pvd", "Cycle000000/data.pvtu", "Cycle000000/proc000000.vtu",
              "Cycle000100/data.pvtu", "Cycle000100/proc000000.vtu"):
        assert os.path.exists(os.path.join(root, f)), f
    import xml.etree.ElementTree as ET

    ds = ET.parse(os.path.join(root, "ParaView.pvd")).getroot().findall(".//DataSet")
    assert [(float(d.get("timestep")), d.get("file")) for d in ds] == [(0.0, "Cycle000000/data.pvtu"),
                                                                        (0.5, "Cycle000100/data.pvtu")]
    pv = ET.parse(os.path.join(root, "Cycle000100", "data.pvtu")).getroot()
    assert [p.get("Source") for p in pv.iter("Piece")] == ["proc000000.vtu"]
    names = [a.get("Name") for a in pv.find(".//PPointData")]
    assert names == ["Density", "Horizontal V", "Vertical V", "Normal V", "Pressure", "Blending Coeff"]
    d = P.read_vtu(os.path.join(root, "Cycle000100", "proc000000.vtu"))
    assert (d["types"] == P.VTK_LAGRANGE_HEXAHEDRON).all() and d["points"].shape == (27 * 27, 3)


def test_save_schedule_reference_times():
    """dt = 1/64 to t = 1 with initial_save_dt 1/8, refined_save_dt 1/4 after t = 1/2: saves at
    1/8, 1/4, 3/8, 1/2 (the save at t = 1/2 already advances by 1/4), 3/4 and at the end."""
    s = P.SaveSchedule(0.0, 0.125, 0.25, 0.5)
    t, saved = 0.0, []
    for step in range(64):
        t += 1.0 / 64
        if s.due(t, done=(step == 63)):
            saved.append(t)
            s.advance(t)
    assert saved == [0.125, 0.25, 0.375, 0.5, 0.75, 1.0]


def test_config_visualization_settings(tmp_path):
    case = config.load_case(os.path.join(CFG, "IsentropicVortex.json"), mesh_dir=MESHES)
    assert case.visualization["visualize"] is True and case.visualization["paraview"] is True
    col = case.paraview_collection(prefix_path=str(tmp_path))
    assert col.name == "ParaView" and col.levels == case.order and col.high_order
    sched = case.save_schedule()
    assert sched.next_t == pytest.approx(0.01)
