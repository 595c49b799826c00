"""The reference's JSON case configuration (``TestCases/*/config.json``) mapped onto this library.

ref: Simulation::LoadConfig and the set-up that follows it (src/Simulation/Simulation.cpp:84-880):
the compile-time switches (PARABOLIC, SUBCELL_FV_BLENDING, SUTHERLAND, USE_FILTER), the physics
constants, the mesh file with ``ser_ref_levels`` uniform refinements, the boundary conditions by
boundary name (``conditions.boundary_conditions``, :490-800), the surface flux (:212-220), the
ODE solver (:177-206), the time-step settings, the checkpoint settings (:111-116, 350-390) and the
visualization settings (:110-135; ``Case.paraview_collection`` and ``Case.save_schedule``).

``load_case`` returns a ``Case``: the mesh, the operator's arguments, the boundary-condition table
(index k <-> boundary name), the stepping settings.  Boundary values named by the configuration
(the reference's ConditionFactory keys, C++ constants of each case header) come from
``REFERENCE_BC_CONSTANTS`` for the reference's constant-valued keys, or from ``bc_values``
(key -> values).  Function-valued state conditions (``"function"``) map to the library's state
profiles (``REFERENCE_BC_FUNCTIONS``: the double-Mach-reflection shock front, the Nagashima inlet
ramp); other function-valued conditions are refused.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from typing import Optional

from . import mesh as M

SURF_CHANDRASHEKAR = 0
SSPRK3 = 0
HOOK_FILTER = 1

# The constant boundary values of the reference's cases (conserved states for the state
# conditions; wall velocity and heat flux for the walls), restated from the case headers.
REFERENCE_BC_CONSTANTS = {
    # TestCases/Euler/2D/ForwardFacingStep/ForwardFacingStep.hpp:23
    "ForwardFacingStepLeftBCVector": [1.4, 1.4 * 3.0, 0.0, 1.0 / (1.4 - 1.0) + 0.5 * 1.4 * 3.0 * 3.0],
    # TestCases/Euler/2D/Ramp/Ramp.hpp:28
    "RampInletBCVector": [1.0, 2.0, 0.0, ((1.0 / 1.4) / 0.4) + 0.5 * 1.0 * 2.0 * 2.0],
    # TestCases/Euler/2D/DoubleMachReflection/DoubleMachReflection.hpp:57
    "DoubleMachReflectionLeftBottom1BCVector": [
        8.0, 8.0 * 7.144709581221619, -8.0 * 4.125,
        116.5 * 1.0 / (1.4 - 1.0) + 0.5 * 8.0 * (7.144709581221619 * 7.144709581221619 + 4.125 * 4.125)],
    # TestCases/Euler/2D/BackwardFacingStep/BackwardFacingStep.hpp:38
    "BackwardFacingStepLeftBCVector": [5.9970, 5.9970 * 98.5914, 0.0,
                                       11666.5 / (1.4 - 1.0) + 0.5 * 5.9970 * 98.5914 * 98.5914],
    # TestCases/NavierStokes/2D/LidDrivenCavity/LidDrivenCavity.hpp:31, 43, 56
    "LidDrivenCavityAdiaBCScalar": [0.0],
    "LidDrivenCavityWallVelBCVector": [0.0, 0.0],
    "LidDrivenCavityLidVelBCVector": [1.0, 0.0],
}

# boundary-state profiles of the library (prandtl_bc.profile)
PROFILE_CONSTANT, PROFILE_FRONT, PROFILE_RAMP = 0, 1, 2


def _dmr_states(gm1i):
    """The double Mach reflection's post- and pre-shock states for 1/(gamma-1) = gm1i
    (TestCases/Euler/2D/DoubleMachReflection/DoubleMachReflection.hpp:17-27, 40-50)."""
    u, v = 7.144709581221619, 4.125
    post = [8.0, 8.0 * u, -8.0 * v, 116.5 * gm1i + 0.5 * 8.0 * (u * u + v * v)]
    pre = [1.4, 0.0, 0.0, 1.0 * gm1i]
    return post, pre


def _dmr_top(par):
    """DoubleMachReflectionTopBCFunction(gammaM1Inverse = x1) (DoubleMachReflection.hpp:32-52): the
    post-shock state where x < 1/6 + (y + 20 t)/sqrt 3, else the pre-shock state."""
    post, pre = _dmr_states(float(par["x1"]))
    return post, PROFILE_FRONT, [0, 1, 1.0 / 6.0, 20.0, math.sqrt(3), *pre]


def _nagashima_ramp(par):
    """NagashimaRampingInletBC(t_ramp = x1) (TestCases/Euler/2D/Nagashima_Ramjet/Nagashima_Ramjet.hpp:
    8-19, 40-53): the free stream of Mach 2 from p0 = 303975 Pa, a0 = 360.63 m/s (gamma 1.4,
    R 287.05), its velocity ramped over t_ramp."""
    g, R, Ma, p0, a0 = 1.4, 287.05, 2.0, 303975.0, 360.63
    C = 1 + 0.5 * (g - 1) * Ma * Ma
    a = a0 / math.sqrt(C)
    pa = p0 / math.pow(C, g / (g - 1))
    Ta = (a * a) / (g * R)
    rhoa = pa / (R * Ta)
    return [], PROFILE_RAMP, [float(par["x1"]), rhoa, Ma * a, 0.0, pa]


# The function-valued state conditions of the reference's cases (ConditionFactory keys registered by
# the case headers) as the library's state profiles: key -> (parameters -> (q, profile, a))
REFERENCE_BC_FUNCTIONS = {
    "DoubleMachReflectionTopBCFunction": _dmr_top,
    "NagashimaRampingInletBC": _nagashima_ramp,
}

# boundary-condition types by the configuration's "type" (Simulation.cpp:520-800)
BC_TYPES = {"slip": 1, "symmetry": 2, "supersonic-inflow": 3, "supersonic-outflow": 4, "specified-state": 5,
            "no-slip-adiabatic": 10, "no-slip-isothermal": 11}

# the reference's ODE solvers (Simulation.cpp:177-206) this library steps (RK6 and RK8 are not
# available: MFEM's tableaus for them are not restated here)
ODE_SOLVERS = {"RK3SSP": SSPRK3, "ForwardEuler": 2, "RK2": 3, "RK4": 4}


class ConfigError(ValueError):
    pass


@dataclass
class Case:
    path: str
    dim: int
    order: int
    mesh: M.Mesh
    physics: dict                    # gamma, Pr, mu, R_gas, viscous, sutherland
    options: dict                    # surface_flux, blending, alpha_max, stage_hooks
    bc_names: list                   # bc_names[k]: the boundary of condition k
    bc_specs: list                   # (type, q) per condition
    ode_solver: str
    scheme: Optional[int]            # the library's stepper, None when not supported
    cfl: float
    dt: Optional[float]
    variable_dt: bool
    final_time: float
    initial_condition: dict
    checkpoint: dict = field(default_factory=dict)
    visualization: dict = field(default_factory=dict)
    ignored: dict = field(default_factory=dict)
    bc_profiles: list = field(default_factory=list)   # (profile, a) per condition

    def bcs(self):
        """The condition table for prandtl_amd.capi.Operator (capi.Bc structs)."""
        from . import capi

        prof = self.bc_profiles or [(PROFILE_CONSTANT, [])] * len(self.bc_specs)
        return [capi.bc(t, *q, profile=pr, a=a) for (t, q), (pr, a) in zip(self.bc_specs, prof)]

    def operator(self, **overrides):
        """prandtl_amd.capi.Operator of the case (FP64)."""
        from . import capi

        kw = dict(self.physics)
        kw.update(self.options)
        kw.update(overrides)
        return capi.Operator(self.mesh, bcs=self.bcs() or None, **kw)

    def paraview_collection(self, prefix_path: Optional[str] = None, **kw):
        """The ParaView collection the reference's driver writes for this case
        (Simulation.cpp:829-853): "ParaView_folder" (default "ParaView") under "output_file_path"
        (or ``prefix_path``), high-order cells with levels of detail = the order."""
        from .paraview import ParaViewDataCollection

        vis = self.visualization
        prefix = prefix_path if prefix_path is not None else vis.get("output_file_path", "")
        return ParaViewDataCollection(vis.get("ParaView_folder", "ParaView"), self.mesh, prefix,
                                      levels_of_detail=self.order, **kw)

    def save_schedule(self, t0: float = 0.0):
        """The reference's output times (initial_save_dt, refined_save_dt, refined_trigger)."""
        from .paraview import SaveSchedule

        vis = self.visualization
        dt1 = float(vis.get("initial_save_dt", 0.01))
        return SaveSchedule(t0, dt1, float(vis.get("refined_save_dt", dt1)), float(vis.get("refined_trigger", 2.0)))


def _resolve_mesh_file(runtime, config_dir, mesh_dir):
    name = runtime.get("mesh_file")
    if not name:
        raise ConfigError("the configuration names no mesh_file")
    base = os.path.basename(name)
    for cand in ([os.path.join(mesh_dir, base)] if mesh_dir else []) + [os.path.join(config_dir, base),
                                                                          os.path.join(config_dir, name)]:
        if os.path.exists(cand):
            return cand
    raise ConfigError(f"mesh file {name!r} not found (looked in {mesh_dir or config_dir})")


def _bc_values(key, bc_values):
    if bc_values and key in bc_values:
        return [float(v) for v in bc_values[key]]
    if key in REFERENCE_BC_CONSTANTS:
        return list(REFERENCE_BC_CONSTANTS[key])
    raise ConfigError(f"no values for the boundary key {key!r}: pass bc_values={{{key!r}: [...]}}")


def _bc_spec(name, props, dim, bc_values):
    typ = props.get("type")
    if typ == "axis":
        raise ConfigError(f"boundary {name!r}: 'axis' belongs to the axisymmetric build, which is not supported")
    if typ not in BC_TYPES:
        raise ConfigError(f"boundary {name!r}: unknown condition type {typ!r}")
    kind = BC_TYPES[typ]
    if typ in ("slip", "symmetry", "supersonic-outflow"):
        return kind, [], PROFILE_CONSTANT, []
    if typ in ("supersonic-inflow", "specified-state"):
        if "function" in props:
            # Simulation.cpp:664-790: a VectorFunctionCoefficient (time_dependent: a
            # VectorTDFunctionCoefficient) from the ConditionFactory key and its parameters
            key = props["function"]
            if key not in REFERENCE_BC_FUNCTIONS:
                raise ConfigError(f"boundary {name!r}: the state function {key!r} has no profile in this library")
            if dim != 2:
                raise ConfigError(f"boundary {name!r}: {key!r} is a 2D condition")
            q, profile, a = REFERENCE_BC_FUNCTIONS[key](props.get("params", {}))
            return kind, q, profile, a
        if "vector" not in props:
            raise ConfigError(f"boundary {name!r}: the condition names neither a vector nor a function")
        q = _bc_values(props["vector"], bc_values)
        if len(q) != dim + 2:
            raise ConfigError(f"boundary {name!r}: {len(q)} state values for {dim + 2} equations")
        return kind, q, PROFILE_CONSTANT, []
    # no-slip walls: velocity vector, then the heat-flux (adiabatic) or temperature (isothermal) scalar
    vel = props.get("velocity", {})
    sec = props.get("heat" if typ == "no-slip-adiabatic" else "temperature", {})
    if "vector" not in vel or "scalar" not in sec:
        raise ConfigError(f"boundary {name!r}: function-valued wall conditions are not supported")
    v = _bc_values(vel["vector"], bc_values)
    s = _bc_values(sec["scalar"], bc_values)
    if len(v) != dim or len(s) != 1:
        raise ConfigError(f"boundary {name!r}: wall values of the wrong size")
    return kind, v + s, PROFILE_CONSTANT, []


def load_case(path: str, mesh_dir: Optional[str] = None, bc_values=None, order: Optional[int] = None,
              mesh: Optional[M.Mesh] = None) -> Case:
    """Read a reference configuration file.  ``mesh_dir``: where to look for the mesh file (by its
    base name; default the configuration's directory); ``bc_values``: values of boundary keys
    that REFERENCE_BC_CONSTANTS does not hold; ``order`` overrides the configuration's; ``mesh``: a
    mesh to use instead of the configuration's mesh file (its boundary faces numbered in the order
    of the configuration's boundary_conditions)."""
    with open(path) as fh:
        cfg = json.load(fh)
    ct = cfg.get("compileTime", {})
    rt = cfg.get("runTime", {})
    ignored = {}
    if ct.get("AXISYMMETRIC"):
        raise ConfigError("the axisymmetric build (AXISYMMETRIC) is not supported")
    dim = int(rt.get("dim", 2))
    if dim not in (2, 3):
        raise ConfigError(f"dim {dim} is not supported (2 or 3)")
    p = int(order if order is not None else rt.get("order", 3))
    neq = int(rt.get("num_equations", 4))   # the reference's default (Simulation.cpp:102)
    if neq != dim + 2:
        raise ConfigError(f"num_equations {neq} != dim + 2")
    if rt.get("lifting_scheme", "BR1") != "BR1":
        raise ConfigError(f"lifting scheme {rt.get('lifting_scheme')!r} is not supported (BR1)")
    if rt.get("numerical_flux", "Chandrashekar") != "Chandrashekar":
        raise ConfigError(f"numerical flux {rt.get('numerical_flux')!r} (the reference accepts only Chandrashekar)")
    viscous = bool(ct.get("PARABOLIC", False))
    physics = dict(gamma=float(rt.get("gamma", 1.4)), Pr=float(rt.get("Pr", 0.72)),
                   mu=float(rt.get("mu", 0.02)), R_gas=float(rt.get("R_gas", 287.05)), viscous=viscous,
                   sutherland=bool(ct.get("SUTHERLAND", False)))
    options = dict(surface_flux=SURF_CHANDRASHEKAR, blending=bool(ct.get("SUBCELL_FV_BLENDING", False)),
                   alpha_max=float(rt.get("alpha_max", 0.5)),
                   stage_hooks=HOOK_FILTER if ct.get("USE_FILTER", False) else 0)
    # boundary conditions in the configuration's order; k indexes the condition table
    bcs_cfg = rt.get("conditions", {}).get("boundary_conditions", {}) or {}
    bc_names, bc_specs, bc_profiles = [], [], []
    for name, props in bcs_cfg.items():
        bc_names.append(name)
        kind, q, profile, a = _bc_spec(name, props, dim, bc_values)
        bc_specs.append((kind, q))
        bc_profiles.append((profile, a))
    bc_index = {n: k for k, n in enumerate(bc_names)}
    # serial refinement, then the parallel mesh's uniform refinement (Simulation.cpp:329-336): both
    # split every element into 2^dim, so they fold into one count
    refine = int(rt.get("ser_ref_levels", 0) or 0) + int(rt.get("par_ref_levels", 0) or 0)
    mpath = None if mesh is not None else _resolve_mesh_file(rt, os.path.dirname(os.path.abspath(path)), mesh_dir)
    if mesh is not None:
        if mesh.dim != dim or mesh.order != p:
            raise ConfigError(f"the given mesh is {mesh.dim}D of order {mesh.order}; the case is {dim}D of order {p}")
        ignored["mesh_file"] = "replaced by the given mesh"
    elif mpath.endswith(".msh"):
        mesh = M.gmsh_mesh(mpath, p, bc_index, refine=refine)
    else:
        # the reference's 2D MFEM meshes are periodic (no boundary); boundary names of MFEM meshes are
        # attribute sets the reference defines for its generated 1D meshes only
        if bc_names:
            raise ConfigError("boundary conditions on an MFEM .mesh file are not supported (Gmsh meshes name them)")
        mesh = M.mfem_mesh(mpath, p, None, refine=refine)
    if "mesh_ordering" in rt:
        ignored["mesh_ordering"] = "element reordering permutes elements without changing the RHS"
    solver = rt.get("ode_solver", "RK3SSP")
    ic = rt.get("conditions", {}).get("initial_conditions", {}) or {}
    ckpt = {k: rt[k] for k in ("checkpoint_save", "checkpoint_load", "checkpoint_dt", "checkpoints_folder",
                               "checkpoint_cycle", "output_file_path") if k in rt}
    # Simulation.cpp:110-135: the visualization switches and the output schedule
    vis = {k: rt[k] for k in ("visualize", "paraview", "visit", "vis_steps", "initial_save_dt", "refined_save_dt",
                              "refined_trigger", "ParaView_folder", "output_file_path") if k in rt}
    if vis.get("visualize") and vis.get("visit") and not vis.get("paraview"):
        ignored["visit"] = "VisIt output is not written (ParaView output is)"
    return Case(path=path, dim=dim, order=p, mesh=mesh, physics=physics, options=options, bc_names=bc_names,
                bc_specs=bc_specs, bc_profiles=bc_profiles, ode_solver=solver, scheme=ODE_SOLVERS.get(solver),
                cfl=float(rt.get("cfl", 1.0)), dt=float(rt["dt"]) if "dt" in rt else None,
                variable_dt=bool(rt.get("variable_dt", False)), final_time=float(rt.get("final_time", 0.0)),
                initial_condition=dict(ic), checkpoint=ckpt, visualization=vis, ignored=ignored)


def reference_initial_state(case: Case):
    """The initial state (dim+2, NE, n) of the reference's constant-state cases, from their
    initial-condition functions and the configuration's parameters:
    LidDrivenCavityIC(Ma = x1, gamma = x2): rho = 1 at rest, p = 1/(Ma^2 gamma)
    (TestCases/NavierStokes/2D/LidDrivenCavity/LidDrivenCavity.hpp:9-19);
    ForwardFacingStepIC(1/(gamma-1) = x1): rho = 1.4, u = 3, v = 0, p = 1
    (TestCases/Euler/2D/ForwardFacingStep/ForwardFacingStep.hpp:9-20);
    DoubleMachReflectionIC(1/(gamma-1) = x1): the post-shock state where x < 1/6 + y/sqrt 3, the
    pre-shock state elsewhere, at the mesh's node coordinates
    (TestCases/Euler/2D/DoubleMachReflection/DoubleMachReflection.hpp:9-29).  Other cases' initial
    conditions are functions of x the caller evaluates."""
    import numpy as np

    ic = case.initial_condition
    name, par = ic.get("function"), ic.get("params", {})
    shape = (case.mesh.num_elements, case.mesh.nodes_per_element)
    u = np.zeros((case.dim + 2,) + shape)
    if name == "LidDrivenCavityIC":
        Ma, gamma = float(par.get("x1", 0.0)), float(par.get("x2", 0.0))
        u[0] = 1.0
        u[case.dim + 1] = 1.0 / (Ma * Ma * gamma) / (gamma - 1.0)
        return u
    if name == "ForwardFacingStepIC":
        gm1i = float(par.get("x1", 0.0))
        u[0] = 1.4
        u[1] = 1.4 * 3.0
        u[case.dim + 1] = 1.0 * gm1i + 0.5 * 1.4 * 3.0 * 3.0
        return u
    if name == "DoubleMachReflectionIC":
        post, pre = _dmr_states(float(par.get("x1", 0.0)))
        X = case.mesh.node_coords()
        behind = X[0] < 1.0 / 6.0 + X[1] / math.sqrt(3)
        for v in range(case.dim + 2):
            u[v] = np.where(behind, post[v], pre[v])
        return u
    raise ConfigError(f"initial condition {name!r} is a function of x: evaluate it at mesh.node_coords()")
