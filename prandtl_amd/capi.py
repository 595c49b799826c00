"""ctypes binding of the C ABI in include/prandtl_amd.h (thin test/bench binding).

Device memory comes from torch (ROCm) tensors; the library only sees raw device
pointers and a HIP stream handle, exactly as a C/C++ caller would pass them.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .mesh import Mesh

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libprandtl_amd.so")

PRANDTL_OK = 0
ERRORS = {-1: "PRANDTL_ERR_ARG", -2: "PRANDTL_ERR_UNSUPPORTED", -3: "PRANDTL_ERR_HIP", -4: "PRANDTL_ERR_MESH",
          -5: "PRANDTL_ERR_NAN"}
SURF_CHANDRASHEKAR, SURF_EC, SURF_RUSANOV = 0, 1, 2
FP64, MIXED = 0, 1
BUG_A1, BUG_A2, BUG_A3 = 1, 2, 4
SSPRK3, LSRK45, FORWARD_EULER, RK2, RK4 = 0, 1, 2, 3, 4
HOOK_FILTER, HOOK_LIMITER, HOOK_UNFUSED = 1, 2, 4
TAPE_SEGMENTS, TAPE_REVOLVE = 0, 1

# every symbol the public header declares
EXPORTED = [
    "prandtl_op_create", "prandtl_op_destroy", "prandtl_op_size", "prandtl_op_mult", "prandtl_op_mult_grad",
    "prandtl_op_get_max_char_speed", "prandtl_op_max_char_speed", "prandtl_rk_step", "prandtl_rk_run",
    "prandtl_op_profile", "prandtl_op_vjp", "prandtl_rk_run_tape", "prandtl_rk_adjoint", "prandtl_last_error",
    "prandtl_filter_constraints", "prandtl_filter_apply", "prandtl_limiter_apply", "prandtl_hook_counts",
    "prandtl_device_alloc", "prandtl_device_free", "prandtl_memcpy", "prandtl_stream_synchronize",
    "prandtl_blend_coefficients", "prandtl_rk_tape_id", "prandtl_op_vjp_noalloc", "prandtl_op_timing",
    "prandtl_op_timing_read", "prandtl_rk_run_tape_schedule", "prandtl_rk_tape_info", "prandtl_op_set_time",
    "prandtl_op_get_time",
]
ALPHA_MAX, ALPHA_MIN = 0.5, 0.001   # the reference's blending defaults


# boundary conditions (prandtl_bc); q as listed in include/prandtl_amd.h
BC_SLIP_WALL, BC_SYMMETRY, BC_SUPERSONIC_INFLOW, BC_SUPERSONIC_OUTFLOW, BC_SPECIFIED_STATE = 1, 2, 3, 4, 5
BC_RIEMANN_INVARIANT, BC_SUBSONIC_INFLOW_RV, BC_SUBSONIC_INFLOW_PTTT, BC_SUBSONIC_OUTFLOW_P = 6, 7, 8, 9
BC_NOSLIP_ADIABATIC, BC_NOSLIP_ISOTHERMAL = 10, 11


# boundary-state profiles (prandtl_bc.profile): the function-valued states of the reference's test cases
PROFILE_CONSTANT, PROFILE_FRONT, PROFILE_RAMP = 0, 1, 2


class Bc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("q", ctypes.c_double * 8), ("profile", ctypes.c_int),
                ("a", ctypes.c_double * 16)]


def bc(type_: int, *q, profile: int = PROFILE_CONSTANT, a=()) -> Bc:
    """A boundary condition: type, its constant values q, and (inflow / specified state) a profile
    with its parameters a, as include/prandtl_amd.h lists them."""
    b = Bc()
    b.type = int(type_)
    for i, v in enumerate(q):
        b.q[i] = float(v)
    b.profile = int(profile)
    for i, v in enumerate(a):
        b.a[i] = float(v)
    return b


def bc_front(type_: int, behind, ahead, i0: int, i1: int, c0: float, c1: float, c2: float) -> Bc:
    """A state with a planar front: `behind` where x[i0] < c0 + (x[i1] + c1 t) / c2, else `ahead`."""
    return bc(type_, *behind, profile=PROFILE_FRONT, a=[i0, i1, c0, c1, c2, *ahead])


def bc_ramp(type_: int, t_ramp: float, rho: float, vel, p: float) -> Bc:
    """An inflow whose velocity rises as clamp((1 + tanh((t - t_r/2)/(t_r/10)))/2, 0, 1) times vel."""
    return bc(type_, profile=PROFILE_RAMP, a=[t_ramp, rho, *vel, p])


class MeshDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("order", ctypes.c_int), ("num_elements", ctypes.c_int),
                ("neighbors", ctypes.POINTER(ctypes.c_int)), ("geometry", ctypes.POINTER(ctypes.c_double)),
                ("num_bc", ctypes.c_int), ("bc", ctypes.POINTER(Bc)),
                ("node_geometry", ctypes.POINTER(ctypes.c_double)), ("face_orientation", ctypes.POINTER(ctypes.c_int)),
                ("node_coords", ctypes.POINTER(ctypes.c_double))]


class Physics(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("Pr", ctypes.c_double), ("mu", ctypes.c_double),
                ("R_gas", ctypes.c_double), ("viscous", ctypes.c_int), ("sutherland", ctypes.c_int)]


class Opts(ctypes.Structure):
    _fields_ = [("surface_flux", ctypes.c_int), ("precision", ctypes.c_int), ("bug_compat", ctypes.c_int),
                ("stage_hooks", ctypes.c_int), ("blending", ctypes.c_int), ("alpha_max", ctypes.c_double),
                ("alpha_min", ctypes.c_double)]


class PrandtlError(RuntimeError):
    pass


_lib = None


def load():
    """Load libprandtl_amd.so (or the build PRANDTL_AMD_LIB names).  Raises if it has not been built:
    there is no fallback."""
    global _lib
    if _lib is None:
        path = os.environ.get("PRANDTL_AMD_LIB") or LIB_PATH
        if not os.path.exists(path):
            raise PrandtlError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; "
                               f"g.build()'` (hipcc, gfx950)")
        L = ctypes.CDLL(path)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        L.prandtl_op_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(Physics), ctypes.POINTER(Opts),
                                        ctypes.POINTER(vp)]
        L.prandtl_op_create.restype = ctypes.c_int
        L.prandtl_op_destroy.argtypes = [vp]
        L.prandtl_op_destroy.restype = None
        L.prandtl_op_size.argtypes = [vp]
        L.prandtl_op_size.restype = ctypes.c_size_t
        L.prandtl_op_mult.argtypes = [vp, vp, vp, vp]
        L.prandtl_op_mult_grad.argtypes = [vp, vp, vp, vp, vp]
        L.prandtl_op_get_max_char_speed.argtypes = [vp, dp]
        L.prandtl_op_max_char_speed.argtypes = [vp, vp, dp, vp]
        L.prandtl_rk_step.argtypes = [vp, ctypes.c_int, vp, dp, ctypes.c_double, vp]
        L.prandtl_rk_run.argtypes = [vp, ctypes.c_int, vp, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, vp]
        L.prandtl_op_profile.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, dp, vp]
        L.prandtl_op_vjp.argtypes = [vp, vp, vp, vp, vp]
        L.prandtl_op_vjp_noalloc.argtypes = [vp, vp, vp, vp, vp]
        L.prandtl_rk_run_tape_schedule.argtypes = [vp, ctypes.c_int, vp, dp, ctypes.c_double, ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_int, vp]
        L.prandtl_rk_tape_info.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
        L.prandtl_op_timing.argtypes = [vp, ctypes.c_int]
        L.prandtl_op_timing_read.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_longlong)]
        L.prandtl_rk_run_tape.argtypes = [vp, ctypes.c_int, vp, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, vp]
        L.prandtl_rk_adjoint.argtypes = [vp, vp, vp]
        L.prandtl_rk_tape_id.argtypes = [vp, ctypes.POINTER(ctypes.c_ulonglong)]
        L.prandtl_filter_constraints.argtypes = [vp, vp, vp]
        L.prandtl_filter_apply.argtypes = [vp, vp, vp]
        L.prandtl_limiter_apply.argtypes = [vp, vp, vp]
        L.prandtl_hook_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
        L.prandtl_blend_coefficients.argtypes = [vp, vp, vp, vp, vp]
        L.prandtl_op_set_time.argtypes = [vp, ctypes.c_double]
        L.prandtl_op_get_time.argtypes = [vp, dp]
        L.prandtl_last_error.argtypes = []
        L.prandtl_last_error.restype = ctypes.c_char_p
        for n in ("prandtl_op_mult", "prandtl_op_mult_grad", "prandtl_op_get_max_char_speed",
                  "prandtl_op_max_char_speed", "prandtl_rk_step", "prandtl_rk_run", "prandtl_op_profile",
                  "prandtl_op_vjp", "prandtl_rk_run_tape", "prandtl_rk_adjoint", "prandtl_filter_constraints",
                  "prandtl_filter_apply", "prandtl_limiter_apply", "prandtl_hook_counts",
                  "prandtl_blend_coefficients", "prandtl_rk_tape_id", "prandtl_op_vjp_noalloc",
                  "prandtl_op_timing", "prandtl_op_timing_read", "prandtl_rk_run_tape_schedule",
                  "prandtl_rk_tape_info", "prandtl_op_set_time", "prandtl_op_get_time"):
            getattr(L, n).restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc: int):
    if rc != PRANDTL_OK:
        msg = load().prandtl_last_error().decode()
        raise PrandtlError(f"{ERRORS.get(rc, rc)}: {msg}")


def _ptr(t, numel=None, name="tensor") -> ctypes.c_void_p:
    """Device pointer of a contiguous float64 CUDA tensor of `numel` values (checked: the kernels
    index the whole state, so a short buffer would be written past its end)."""
    import torch

    if not isinstance(t, torch.Tensor):
        raise PrandtlError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if numel is not None and t.numel() != numel:
        raise PrandtlError(f"{name}: {t.numel()} values; the operator's state holds {numel}")
    if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
        raise PrandtlError(f"{name}: expected a contiguous float64 device tensor "
                           f"(got {t.dtype} on {t.device}, contiguous={t.is_contiguous()})")
    return ctypes.c_void_p(t.data_ptr())


def _ptrs(*specs):
    """_ptr over (tensor, numel, name) triples, every size checked before any placement."""
    import torch

    for t, n, name in specs:
        if isinstance(t, torch.Tensor) and n is not None and t.numel() != n:
            raise PrandtlError(f"{name}: {t.numel()} values; the operator's state holds {n}")
    return [_ptr(t, n, name) for t, n, name in specs]


def _stream(stream=None) -> ctypes.c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Operator:
    """Owns a prandtl_op handle: the DGSEMOperator of one mesh and physics."""

    def __init__(self, mesh: Mesh, gamma=1.4, viscous=False, Pr=0.72, mu=0.0, R_gas=287.0, sutherland=False,
                 surface_flux=SURF_CHANDRASHEKAR, precision=FP64, bug_compat=0, stage_hooks=0, blending=False,
                 alpha_max=ALPHA_MAX, alpha_min=ALPHA_MIN, bcs=None):
        L = load()
        self.mesh = mesh
        self._nbr = np.ascontiguousarray(mesh.neighbors, dtype=np.int32)
        self._geom = np.ascontiguousarray(mesh.geom, dtype=np.float64)
        self._bcs = (Bc * len(bcs))(*bcs) if bcs else None
        ng = getattr(mesh, "node_geom", None)
        fo = getattr(mesh, "face_orient", None)
        self._ngeom = None if ng is None else np.ascontiguousarray(ng, dtype=np.float64)
        self._forient = None if fo is None else np.ascontiguousarray(fo, dtype=np.int32)
        # node coordinates [e][node][dim], for boundary states that vary in space
        self._xyz = None
        if bcs and any(b.profile == PROFILE_FRONT for b in bcs):
            self._xyz = np.ascontiguousarray(np.transpose(mesh.node_coords(), (1, 2, 0)), dtype=np.float64)
        md = MeshDesc(mesh.dim, mesh.order, mesh.num_elements,
                      self._nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                      self._geom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      len(bcs) if bcs else 0, self._bcs,
                      None if self._ngeom is None else self._ngeom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      None if self._forient is None else self._forient.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                      None if self._xyz is None else self._xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        ph = Physics(gamma, Pr, mu, R_gas, int(viscous), int(sutherland))
        op = Opts(int(surface_flux), int(precision), int(bug_compat), int(stage_hooks), int(blending),
                  float(alpha_max), float(alpha_min))
        h = ctypes.c_void_p()
        _check(L.prandtl_op_create(ctypes.byref(md), ctypes.byref(ph), ctypes.byref(op), ctypes.byref(h)))
        self._h = h
        self.viscous = bool(viscous)

    def close(self):
        if getattr(self, "_h", None):
            load().prandtl_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self) -> int:
        return int(load().prandtl_op_size(self._h))

    def mult(self, u, dudt, stream=None):
        pu, pd = _ptrs((u, self.size, "u"), (dudt, self.size, "dudt"))
        _check(load().prandtl_op_mult(self._h, pu, pd, _stream(stream)))

    def mult_grad(self, u, dudt, grad, stream=None):
        pu, pd, pg = _ptrs((u, self.size, "u"), (dudt, self.size, "dudt"), (grad, self.mesh.dim * self.size, "grad"))
        _check(load().prandtl_op_mult_grad(self._h, pu, pd, pg, _stream(stream)))

    def set_time(self, t: float):
        """The time of Mult's boundary states (TimeDependentOperator::SetTime)."""
        _check(load().prandtl_op_set_time(self._h, float(t)))

    def get_time(self) -> float:
        v = ctypes.c_double(0.0)
        _check(load().prandtl_op_get_time(self._h, ctypes.byref(v)))
        return v.value

    def get_max_char_speed(self) -> float:
        v = ctypes.c_double(0.0)
        _check(load().prandtl_op_get_max_char_speed(self._h, ctypes.byref(v)))
        return v.value

    def max_char_speed(self, u, stream=None) -> float:
        v = ctypes.c_double(0.0)
        _check(load().prandtl_op_max_char_speed(self._h, _ptr(u, self.size, "u"), ctypes.byref(v), _stream(stream)))
        return v.value

    def profile(self, u, epilogue=0, iters=20, stream=None):
        """Average ms per launch of (lift pass, flux-differencing pass)."""
        ms = (ctypes.c_double * 2)()
        _check(load().prandtl_op_profile(self._h, _ptr(u, self.size, "u"), int(epilogue), int(iters), ms, _stream(stream)))
        return float(ms[0]), float(ms[1])

    def timing(self, enable: bool):
        """Bracket every lifting / flux-differencing launch of this handle with HIP events."""
        _check(load().prandtl_op_timing(self._h, int(bool(enable))))

    def timing_read(self):
        """((lift ms, flux ms) average durations, (lift launches, flux launches)) since the last read."""
        ms = (ctypes.c_double * 2)()
        n = (ctypes.c_longlong * 2)()
        _check(load().prandtl_op_timing_read(self._h, ms, n))
        return (float(ms[0]), float(ms[1])), (int(n[0]), int(n[1]))

    def rk_step(self, scheme, u, t, dt, stream=None) -> float:
        tt = ctypes.c_double(t)
        _check(load().prandtl_rk_step(self._h, int(scheme), _ptr(u, self.size, "u"), ctypes.byref(tt), float(dt), _stream(stream)))
        return tt.value

    def rk_run(self, scheme, u, t, dt, nsteps, nan_check=0, stream=None) -> float:
        tt = ctypes.c_double(t)
        _check(load().prandtl_rk_run(self._h, int(scheme), _ptr(u, self.size, "u"), ctypes.byref(tt), float(dt), int(nsteps),
                                     int(nan_check), _stream(stream)))
        return tt.value

    def vjp(self, u, lam, out, stream=None):
        """out = (dL/du)^T lam at u."""
        pu, pl, po = _ptrs((u, self.size, "u"), (lam, self.size, "lam"), (out, self.size, "out"))
        _check(load().prandtl_op_vjp(self._h, pu, pl, po, _stream(stream)))

    def rk_run_tape(self, u, t, dt, nsteps, ckpt_every, stream=None, scheme=LSRK45) -> float:
        tt = ctypes.c_double(t)
        _check(load().prandtl_rk_run_tape(self._h, int(scheme), _ptr(u, self.size, "u"), ctypes.byref(tt), float(dt), int(nsteps),
                                          int(ckpt_every), _stream(stream)))
        return tt.value

    def rk_run_tape_schedule(self, u, t, dt, nsteps, schedule, param, stream=None, scheme=LSRK45) -> float:
        """A taped run with checkpoint schedule TAPE_SEGMENTS (param = steps per segment) or TAPE_REVOLVE
        (param = snapshot buffers)."""
        tt = ctypes.c_double(t)
        _check(load().prandtl_rk_run_tape_schedule(self._h, int(scheme), _ptr(u, self.size, "u"), ctypes.byref(tt),
                                                   float(dt), int(nsteps), int(schedule), int(param),
                                                   _stream(stream)))
        return tt.value

    def tape_info(self):
        """(device states held by the reverse sweep, forward steps the last reverse sweep ran)."""
        a, b = ctypes.c_longlong(0), ctypes.c_longlong(0)
        _check(load().prandtl_rk_tape_info(self._h, ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def tape_id(self) -> int:
        """Identifier of the handle's last taped run (0: none, or it failed)."""
        v = ctypes.c_ulonglong(0)
        _check(load().prandtl_rk_tape_id(self._h, ctypes.byref(v)))
        return int(v.value)

    def rk_adjoint(self, ubar, stream=None):
        """ubar: dJ/du_final on entry, dJ/du_initial on return (for the last taped run)."""
        _check(load().prandtl_rk_adjoint(self._h, _ptr(ubar, self.size, "ubar"), _stream(stream)))

    # subcell blending -- DGSEMOperator::ComputeBlendingCoefficient (the alpha and eta grid functions)
    def blend_coefficients(self, u, stream=None):
        """(alpha, eta) per element of u, device tensors."""
        import torch

        alpha = torch.empty(self.mesh.num_elements, dtype=torch.float64, device=u.device)
        eta = torch.empty_like(alpha)
        _check(load().prandtl_blend_coefficients(self._h, _ptr(u, self.size, "u"), _ptr(alpha), _ptr(eta), _stream(stream)))
        return alpha, eta

    # stage hooks -- Filter::GetFilterConstraints / FilterSolution, Limiter::LimitSolution
    def filter_constraints(self, u, stream=None):
        _check(load().prandtl_filter_constraints(self._h, _ptr(u, self.size, "u"), _stream(stream)))

    def filter_apply(self, u, stream=None):
        _check(load().prandtl_filter_apply(self._h, _ptr(u, self.size, "u"), _stream(stream)))

    def limiter_apply(self, u, stream=None):
        _check(load().prandtl_limiter_apply(self._h, _ptr(u, self.size, "u"), _stream(stream)))

    def hook_counts(self):
        """(elements filtered, elements limited) since the previous call."""
        f, l_ = ctypes.c_longlong(0), ctypes.c_longlong(0)
        _check(load().prandtl_hook_counts(self._h, ctypes.byref(f), ctypes.byref(l_)))
        return f.value, l_.value
