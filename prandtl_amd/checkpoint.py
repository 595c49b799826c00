"""Checkpoints in the reference's format.

ref: Simulation::Run (src/Simulation/Simulation.cpp:1084-1116) writes, every checkpoint_dt,
``<checkpoints_folder>/Cycle<n>/checkpoint_cycle_<n>.<rank:08d>.chk`` with ``sol->Save`` (MFEM's
GridFunction text format: the finite-element space header, then the values) and, on rank 0,
``checkpoint_cycle_<n>.json`` with the time and the cycle; the restart (:350-390) reads them back
(``checkpoint_load``, ``checkpoint_cycle``).

The solution space is the reference's ``DG_FECollection(order, dim, GaussLobatto)`` with
``num_equations`` components ordered byNODES (Simulation.cpp:338-340, Simulation.hpp:64):
header ``FiniteElementSpace`` / ``FiniteElementCollection: L2_T1_<dim>D_P<order>`` /
``VDim: <dim+2>`` / ``Ordering: 0``, a blank line, then one value per line in the state
layout u[v][e][node] -- this library's layout, so a state moves in and out without a permute.
Values are written with 17 significant digits (the reference's stream keeps the default 6; both
read back here).
"""
from __future__ import annotations

import json
import os

import numpy as np


def fec_name(dim: int, order: int) -> str:
    return f"L2_T1_{dim}D_P{order}"


def chk_path(folder: str, cycle: int, rank: int = 0) -> str:
    return os.path.join(folder, f"Cycle{cycle}", f"checkpoint_cycle_{cycle}.{rank:08d}.chk")


def meta_path(folder: str, cycle: int) -> str:
    return os.path.join(folder, f"Cycle{cycle}", f"checkpoint_cycle_{cycle}.json")


def write_gridfunction(fh, u: np.ndarray, dim: int, order: int) -> None:
    """MFEM GridFunction::Save of a byNODES DG state (u: (dim+2, NE, n) or flat)."""
    u = np.asarray(u, dtype=np.float64).reshape(-1)
    fh.write("FiniteElementSpace\n")
    fh.write(f"FiniteElementCollection: {fec_name(dim, order)}\n")
    fh.write(f"VDim: {dim + 2}\n")
    fh.write("Ordering: 0\n\n")
    fh.write("\n".join(repr(float(v)) if np.isfinite(v) else str(v) for v in u))
    fh.write("\n")


def read_gridfunction(fh, dim: int = None, order: int = None) -> tuple:
    """Parse an MFEM GridFunction text file of a DG (L2) space: (values, dim, order, vdim)."""
    head = {}
    line = fh.readline()
    if line.strip() not in ("FiniteElementSpace", "MFEM FiniteElementSpace v1.0"):
        raise ValueError(f"not an MFEM grid function (first line {line.strip()!r})")
    while True:
        line = fh.readline()
        if not line or not line.strip():
            break
        k, _, v = line.partition(":")
        head[k.strip()] = v.strip()
    fec = head.get("FiniteElementCollection", "")
    parts = fec.split("_")
    if len(parts) != 4 or parts[0] != "L2" or parts[1] != "T1" or not parts[2].endswith("D") or \
            not parts[3].startswith("P"):
        raise ValueError(f"finite-element collection {fec!r} is not the Gauss-Lobatto L2 space of the solver")
    fdim, forder = int(parts[2][:-1]), int(parts[3][1:])
    vdim = int(head.get("VDim", 1))
    if int(head.get("Ordering", 0)) != 0:
        raise ValueError("the solver's states are ordered byNODES (Ordering: 0)")
    if dim is not None and fdim != dim or order is not None and forder != order:
        raise ValueError(f"checkpoint of a dim {fdim}, order {forder} space, expected {dim}, {order}")
    vals = np.array(fh.read().split(), dtype=np.float64)
    return vals, fdim, forder, vdim


def save_checkpoint(folder: str, cycle: int, t: float, u: np.ndarray, dim: int, order: int, rank: int = 0) -> str:
    """Write the checkpoint of cycle ``cycle`` at time ``t`` (and, on rank 0, its meta file)."""
    os.makedirs(os.path.join(folder, f"Cycle{cycle}"), exist_ok=True)
    path = chk_path(folder, cycle, rank)
    with open(path, "w") as fh:
        write_gridfunction(fh, u, dim, order)
    if rank == 0:
        with open(meta_path(folder, cycle), "w") as fh:
            fh.write("{\n \"time\": %.17g,\n \"cycle\": %d\n}\n" % (t, cycle))
    return path


def load_checkpoint(folder: str, cycle: int, dim: int, order: int, num_elements: int, rank: int = 0):
    """Read the checkpoint of cycle ``cycle``: (u (dim+2, NE, n), t, cycle).  The meta file's
    cycle must match (Simulation.cpp:359-367)."""
    with open(meta_path(folder, cycle)) as fh:
        meta = json.load(fh)
    if int(meta["cycle"]) != cycle:
        raise ValueError("mismatch between the cycle number and the meta file")
    with open(chk_path(folder, cycle, rank)) as fh:
        vals, fdim, forder, vdim = read_gridfunction(fh, dim, order)
    n = (order + 1) ** dim
    if vdim != dim + 2 or vals.size != vdim * num_elements * n:
        raise ValueError(f"checkpoint holds {vals.size} values, expected {(dim + 2) * num_elements * n}")
    return vals.reshape(dim + 2, num_elements, n), float(meta["time"]), int(meta["cycle"])
