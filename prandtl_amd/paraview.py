"""ParaView output of a solution, as the reference's driver writes it.

ref: Simulation::Initialize, the ParaView data collection and its fields
(src/Simulation/Simulation.cpp:829-853), the primitive variables it saves
(:1045-1060: u = m_x/rho, v, w, p = (gamma-1)(E - rho |v|^2 / 2)), the saves of the initial state
and of the run (:980-985, :1062-1069) and their schedule (:124-128, :466, :1038, :1078-1079).

The reference hands its grid functions to MFEM's ParaViewDataCollection, whose layout this module
writes (MFEM is not vendored with the reference; the layout is restated from MFEM's documented
output):

    <prefix>/<name>/<name>.pvd                   one DataSet per saved cycle (time, file)
    <prefix>/<name>/Cycle%06d/data.pvtu          the cycle's parallel file, naming its pieces
    <prefix>/<name>/Cycle%06d/proc%06d.vtu       one piece per rank (one here: proc000000.vtu)

Point fields, one value per point: "Density", "Horizontal V", "Vertical V" (dim > 1), "Normal V"
(dim > 2), "Pressure", "Blending Coeff" (the element's alpha at each of its points).

High-order output (the reference's SetHighOrderOutput(true) with SetLevelsOfDetail(order)): one
VTK Lagrange cell per element (VTK_LAGRANGE_QUADRILATERAL = 70, VTK_LAGRANGE_HEXAHEDRON = 72) of
order ``levels_of_detail``, its points equispaced in the reference element.  The nodal (GLL)
coordinates and fields are carried there by the element's Lagrange basis, which is exact when
``levels_of_detail`` >= the order.  Points follow VTK's ordering of Lagrange cells: corners in
the linear cell's order, then the interior points of each edge in the linear cell's edge order
(each edge from its first to its second vertex), then the interior points of each face in the
linear cell's face order (x-min, x-max, y-min, y-max, z-min, z-max; the lower remaining axis
fastest), then the interior points (x fastest).  The files declare version 2.2, the version
whose readers take this ordering of the hexahedron's edges 10 = (3, 7) and 11 = (2, 6).
Without high-order output each element is cut into levels_of_detail^dim linear cells
(VTK_QUAD = 9, VTK_HEXAHEDRON = 12) on the same points.

Binary data arrays (the reference's VTKFormat::BINARY) are base64 of a UInt64 byte count
followed by the little-endian values, one base64 stream per array; ``fmt="ascii"`` writes text.
"""
from __future__ import annotations

import base64
import os
import struct
import xml.etree.ElementTree as ET
from typing import Optional

import numpy as np

from .mesh import Mesh, gll_nodes

VTK_QUAD, VTK_HEXAHEDRON = 9, 12
VTK_LAGRANGE_QUADRILATERAL, VTK_LAGRANGE_HEXAHEDRON = 70, 72
VTK_FILE_VERSION = "2.2"

# corners of the linear cells in VTK's order, as lattice coordinates (0 or 1 per axis)
QUAD_CORNERS = ((0, 0), (1, 0), (1, 1), (0, 1))
HEX_CORNERS = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))

FIELD_NAMES = ("Density", "Horizontal V", "Vertical V", "Normal V", "Pressure")


def lagrange_point_index(ijk, order: int) -> int:
    """VTK's index of the point with lattice coordinates ``ijk`` (each 0..order) in a Lagrange
    quadrilateral (two coordinates) or hexahedron (three) of ``order``."""
    n = order
    m = n - 1                                   # interior points per edge
    bd = [c == 0 or c == n for c in ijk]
    nb = sum(bd)
    if len(ijk) == 2:
        i, j = ijk
        if nb == 2:
            return (1 if i else 0) if j == 0 else (2 if i else 3)
        if nb == 1:
            # edges (0,1): j = 0; (1,2): i = n; (3,2): j = n; (0,3): i = 0, each in increasing parameter
            if not bd[0]:
                return 4 + (i - 1) + (0 if j == 0 else 2 * m)
            return 4 + (j - 1) + (m if i == n else 3 * m)
        return 4 + 4 * m + (i - 1) + m * (j - 1)
    i, j, k = ijk
    if nb == 3:
        return (2 if j else 1) + (4 if k else 0) if i else (3 if j else 0) + (4 if k else 0)
    off = 8
    if nb == 2:
        # the hexahedron's edges (0,1) (1,2) (3,2) (0,3) (4,5) (5,6) (7,6) (4,7) (0,4) (1,5) (3,7) (2,6)
        if not bd[0]:                           # along x: edges 0, 2, 4, 6
            edge = (2 if j else 0) + (4 if k else 0)
            return off + edge * m + (i - 1)
        if not bd[1]:                           # along y: edges 1 (x = n), 3 (x = 0), 5, 7
            edge = (1 if i else 3) + (4 if k else 0)
            return off + edge * m + (j - 1)
        edge = 8 + ((3 if j else 1) if i else (2 if j else 0))   # along z: edges 8..11
        return off + edge * m + (k - 1)
    off += 12 * m
    if nb == 1:
        # faces x-min, x-max (y fastest, then z), y-min, y-max (x, then z), z-min, z-max (x, then y)
        if bd[0]:
            return off + (1 if i else 0) * m * m + (j - 1) + m * (k - 1)
        if bd[1]:
            return off + (3 if j else 2) * m * m + (i - 1) + m * (k - 1)
        return off + (5 if k else 4) * m * m + (i - 1) + m * (j - 1)
    off += 6 * m * m
    return off + (i - 1) + m * ((j - 1) + m * (k - 1))


def lagrange_permutation(dim: int, order: int) -> np.ndarray:
    """perm[v] = the lexicographic lattice index (x fastest) of VTK point v of a Lagrange cell."""
    R = order + 1
    perm = np.empty(R ** dim, dtype=np.int64)
    for lex in range(R ** dim):
        ijk = tuple((lex // R ** a) % R for a in range(dim))
        perm[lagrange_point_index(ijk, order)] = lex
    return perm


def resample_matrix(np1: int, levels: int) -> np.ndarray:
    """M[a, b] = l_b(a / levels): the GLL Lagrange basis of ``np1`` nodes on [0, 1] at the
    ``levels + 1`` equispaced points."""
    xg, _ = gll_nodes(np1)
    xe = np.linspace(0.0, 1.0, levels + 1)
    M = np.ones((levels + 1, np1))
    for b in range(np1):
        for c in range(np1):
            if c != b:
                M[:, b] *= (xe - xg[c]) / (xg[b] - xg[c])
    return M


def resample(field: np.ndarray, dim: int, np1: int, levels: int) -> np.ndarray:
    """A nodal field (..., NE, np1^dim), node = i + np1 (j + np1 k), at the equispaced lattice
    of ``levels + 1`` points per axis: (..., NE, (levels + 1)^dim), x fastest."""
    M = resample_matrix(np1, levels)
    lead = field.shape[:-1]
    f = field.reshape(lead + (np1,) * dim)                 # (..., [k,] j, i)
    for ax in range(dim):
        # contract the axis of coordinate ax (the last axis is x)
        f = np.moveaxis(np.tensordot(f, M, axes=([f.ndim - 1 - ax], [1])), -1, f.ndim - 1 - ax)
    return f.reshape(lead + ((levels + 1) ** dim,))


def primitive_fields(u: np.ndarray, gamma: float) -> dict:
    """The reference's saved fields from a conserved state (dim + 2, NE, n): Density, the
    velocity components, Pressure (Simulation.cpp:1045-1060)."""
    dim = u.shape[0] - 2
    rho = u[0]
    vel = [u[1 + d] / rho for d in range(dim)]
    vsq = sum(v * v for v in vel)
    out = {"Density": rho}
    for d in range(dim):
        out[FIELD_NAMES[1 + d]] = vel[d]
    out["Pressure"] = (gamma - 1.0) * (u[dim + 1] - 0.5 * rho * vsq)
    return out


def _encode(arr: np.ndarray, fmt: str) -> str:
    if fmt == "ascii":
        if arr.dtype.kind == "f":
            return " ".join(repr(float(x)) for x in arr.ravel())
        return " ".join(str(int(x)) for x in arr.ravel())
    raw = np.ascontiguousarray(arr).astype(arr.dtype.newbyteorder("<"), copy=False).tobytes()
    return base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")


_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64",
              np.dtype(np.uint8): "UInt8"}


def _array_xml(arr: np.ndarray, fmt: str, name: Optional[str] = None, ncomp: int = 1) -> str:
    nm = f' Name="{name}"' if name is not None else ""
    return (f'<DataArray type="{_VTK_TYPES[arr.dtype]}"{nm} NumberOfComponents="{ncomp}" format="{fmt}">'
            f"{_encode(arr, fmt)}</DataArray>\n")


def cells_of(mesh: Mesh, levels: int, high_order: bool = True):
    """Points (NE (levels+1)^dim, 3), connectivity, offsets and types of the mesh's output cells."""
    dim, NE = mesh.dim, mesh.num_elements
    R = levels + 1
    X = resample(mesh.node_coords(), dim, mesh.np1, levels)              # (dim, NE, R^dim), lex order
    pts = np.zeros((NE, R ** dim, 3))
    for d in range(dim):
        pts[:, :, d] = X[d]
    base = (np.arange(NE, dtype=np.int64) * R ** dim)[:, None]
    if high_order:
        perm = lagrange_permutation(dim, levels)
        conn = (base + perm[None, :]).ravel()
        npc = R ** dim
        types = np.full(NE, VTK_LAGRANGE_QUADRILATERAL if dim == 2 else VTK_LAGRANGE_HEXAHEDRON, dtype=np.uint8)
    else:
        corners = QUAD_CORNERS if dim == 2 else HEX_CORNERS
        sub = []
        for lex in range(levels ** dim):
            c0 = [(lex // levels ** a) % levels for a in range(dim)]
            sub.append([sum((c0[a] + cc[a]) * R ** a for a in range(dim)) for cc in corners])
        sub = np.array(sub, dtype=np.int64)                               # (levels^dim, 2^dim)
        conn = (base[:, :, None] + sub[None, :, :]).ravel()
        npc = len(corners)
        types = np.full(NE * levels ** dim, VTK_QUAD if dim == 2 else VTK_HEXAHEDRON, dtype=np.uint8)
    ncell = types.size
    idx = np.int32 if NE * R ** dim < 2 ** 31 else np.int64
    offsets = (np.arange(1, ncell + 1, dtype=np.int64) * npc).astype(idx)
    return pts.reshape(-1, 3), conn.astype(idx), offsets, types


def write_vtu(path: str, mesh: Mesh, fields: dict, levels: int, high_order: bool = True, fmt: str = "binary"):
    """One piece: the mesh's cells and the nodal ``fields`` (name -> (NE, n) array, or (NE,) for
    an element constant) at the output points."""
    pts, conn, offsets, types = cells_of(mesh, levels, high_order)
    out = ['<?xml version="1.0"?>\n',
           f'<VTKFile type="UnstructuredGrid" version="{VTK_FILE_VERSION}" byte_order="LittleEndian" '
           'header_type="UInt64">\n<UnstructuredGrid>\n',
           f'<Piece NumberOfPoints="{pts.shape[0]}" NumberOfCells="{types.size}">\n<Points>\n',
           _array_xml(pts.ravel(), fmt, ncomp=3), "</Points>\n<Cells>\n",
           _array_xml(conn, fmt, "connectivity"), _array_xml(offsets, fmt, "offsets"),
           _array_xml(types, fmt, "types"), "</Cells>\n<PointData>\n"]
    R = levels + 1
    for name, f in fields.items():
        f = np.asarray(f, dtype=np.float64)
        if f.ndim == 1:                                                    # one value per element
            vals = np.repeat(f, R ** mesh.dim)
        else:
            vals = resample(f, mesh.dim, mesh.np1, levels).ravel()
        out.append(_array_xml(vals, fmt, name))
    out.append("</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as fh:
        fh.write("".join(out))


def write_pvtu(path: str, pieces, field_names):
    out = ['<?xml version="1.0"?>\n',
           f'<VTKFile type="PUnstructuredGrid" version="{VTK_FILE_VERSION}" byte_order="LittleEndian" '
           'header_type="UInt64">\n<PUnstructuredGrid GhostLevel="0">\n',
           '<PPoints>\n<PDataArray type="Float64" NumberOfComponents="3"/>\n</PPoints>\n<PPointData>\n']
    for name in field_names:
        out.append(f'<PDataArray type="Float64" Name="{name}" NumberOfComponents="1"/>\n')
    out.append("</PPointData>\n")
    for p in pieces:
        out.append(f'<Piece Source="{p}"/>\n')
    out.append("</PUnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as fh:
        fh.write("".join(out))


def write_pvd(path: str, datasets):
    """``datasets``: (time, file relative to the .pvd, name) per saved cycle."""
    out = ['<?xml version="1.0"?>\n',
           f'<VTKFile type="Collection" version="{VTK_FILE_VERSION}" byte_order="LittleEndian">\n<Collection>\n']
    for t, f, name in datasets:
        out.append(f'<DataSet timestep="{t!r}" group="" part="0" file="{f}" name="{name}"/>\n')
    out.append("</Collection>\n</VTKFile>\n")
    with open(path, "w") as fh:
        fh.write("".join(out))


class ParaViewDataCollection:
    """The reference's output collection (Simulation.cpp:829-853): ``name`` (its runtime
    "ParaView_folder", default "ParaView") under ``prefix_path`` (its "output_file_path").
    register_field / set_cycle / set_time / save as on the reference's collection; save_state
    sets the reference's fields from a conserved state and saves."""

    def __init__(self, name: str, mesh: Mesh, prefix_path: str = "", levels_of_detail: Optional[int] = None,
                 high_order: bool = True, fmt: str = "binary", rank: int = 0, nranks: int = 1):
        self.name, self.mesh, self.prefix = name, mesh, prefix_path
        self.levels = int(levels_of_detail if levels_of_detail is not None else mesh.order)
        if self.levels < 1:
            raise ValueError("levels_of_detail must be >= 1")
        self.high_order, self.fmt = high_order, fmt
        self.rank, self.nranks = rank, nranks
        self.fields = {}
        self.cycle, self.time = 0, 0.0
        self.datasets = []

    @property
    def path(self) -> str:
        return os.path.join(self.prefix, self.name)

    def register_field(self, name: str, values) -> None:
        self.fields[name] = values

    def set_cycle(self, cycle: int) -> None:
        self.cycle = int(cycle)

    def set_time(self, t: float) -> None:
        self.time = float(t)

    def save(self) -> str:
        """Write this cycle's piece, its .pvtu and the updated .pvd; returns the .pvtu path."""
        cyc = f"Cycle{self.cycle:06d}"
        cdir = os.path.join(self.path, cyc)
        os.makedirs(cdir, exist_ok=True)
        fields = {k: (v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)) for k, v in self.fields.items()}
        write_vtu(os.path.join(cdir, f"proc{self.rank:06d}.vtu"), self.mesh, fields, self.levels,
                  self.high_order, self.fmt)
        pvtu = os.path.join(cdir, "data.pvtu")
        if self.rank == 0:
            write_pvtu(pvtu, [f"proc{r:06d}.vtu" for r in range(self.nranks)], list(fields))
            self.datasets = [d for d in self.datasets if d[2] != cyc] + [(self.time, f"{cyc}/data.pvtu", cyc)]
            write_pvd(os.path.join(self.path, f"{self.name}.pvd"), self.datasets)
        return pvtu

    def save_state(self, u, gamma: float, alpha=None, cycle: Optional[int] = None, t: Optional[float] = None) -> str:
        """Register the reference's fields of the conserved state ``u`` (dim + 2, NE, n; numpy or a
        torch tensor) and the per-element blending coefficient ``alpha`` (zeros when None, as the
        reference's unset alpha), then save."""
        u = u.cpu().numpy() if hasattr(u, "cpu") else np.asarray(u)
        for k, v in primitive_fields(u, gamma).items():
            self.register_field(k, v)
        a = np.zeros(self.mesh.num_elements) if alpha is None else alpha
        self.register_field("Blending Coeff", a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a))
        if cycle is not None:
            self.set_cycle(cycle)
        if t is not None:
            self.set_time(t)
        return self.save()


class SaveSchedule:
    """The reference's output times (Simulation.cpp:124-128, 466, 1038, 1078-1079): the first save
    after the initial one is due at t0 + initial_save_dt; a step ending at t saves when the run is
    done or t >= the next save time, which then advances by initial_save_dt while t < trigger and
    by refined_save_dt after."""

    def __init__(self, t0: float, initial_save_dt: float = 0.01, refined_save_dt: Optional[float] = None,
                 refined_trigger: float = 2.0):
        self.dt1 = float(initial_save_dt)
        self.dt2 = float(refined_save_dt if refined_save_dt is not None else initial_save_dt)
        self.trigger = float(refined_trigger)
        self.next_t = t0 + self.dt1

    def due(self, t: float, done: bool = False) -> bool:
        return done or t >= self.next_t

    def advance(self, t: float) -> None:
        self.next_t += self.dt1 if t < self.trigger else self.dt2


def read_vtu(path: str) -> dict:
    """The arrays of a piece this module wrote (binary or ASCII): "points" (N, 3), "connectivity",
    "offsets", "types" and each point field by name."""
    root = ET.parse(path).getroot()
    htype = np.uint64 if root.get("header_type", "UInt32") == "UInt64" else np.uint32
    hsize = np.dtype(htype).itemsize
    types = {"Float64": np.float64, "Int32": np.int32, "Int64": np.int64, "UInt8": np.uint8}
    out = {}
    for da in root.iter("DataArray"):
        dt = np.dtype(types[da.get("type")]).newbyteorder("<")
        text = (da.text or "").strip()
        if da.get("format") == "ascii":
            vals = np.array(text.split(), dtype=np.float64).astype(dt)
        else:
            raw = base64.b64decode(text)
            n = int(np.frombuffer(raw[:hsize], dtype=np.dtype(htype).newbyteorder("<"))[0])
            vals = np.frombuffer(raw[hsize:hsize + n], dtype=dt).copy()
        ncomp = int(da.get("NumberOfComponents", "1"))
        name = da.get("Name") or "points"
        out[name] = vals.reshape(-1, ncomp) if ncomp > 1 else vals
    return out
