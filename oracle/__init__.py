"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is a line-traceable C++ restatement of the reference solver's
arithmetic, built from ``oracle/prandtl_oracle.cpp`` by ``oracle/Makefile``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module; the product path never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_FMA_PATH = os.path.join(_HERE, "liboracle_fma.so")
_LD_PATH = os.path.join(_HERE, "liboracle_ld.so")
_lib = None
_lib_fma = None
_lib_ld = None

SURF_CHANDRASHEKAR, SURF_EC, SURF_RUSANOV = 0, 1, 2
BUG_A1, BUG_A2, BUG_A3 = 1, 2, 4


# boundary-condition types (BdrFaceIntegrator subclasses); q as listed in include/prandtl_amd.h
BC_SLIP_WALL, BC_SYMMETRY, BC_SUPERSONIC_INFLOW, BC_SUPERSONIC_OUTFLOW, BC_SPECIFIED_STATE = 1, 2, 3, 4, 5
BC_RIEMANN_INVARIANT, BC_SUBSONIC_INFLOW_RV, BC_SUBSONIC_INFLOW_PTTT, BC_SUBSONIC_OUTFLOW_P = 6, 7, 8, 9
BC_NOSLIP_ADIABATIC, BC_NOSLIP_ISOTHERMAL = 10, 11


# boundary-state profiles (constant, planar front, velocity ramp), parameters as in include/prandtl_amd.h
PROFILE_CONSTANT, PROFILE_FRONT, PROFILE_RAMP = 0, 1, 2


class OrcBc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("q", ctypes.c_double * 8), ("profile", ctypes.c_int),
                ("a", ctypes.c_double * 16)]


def bc(type_: int, *q, profile: int = PROFILE_CONSTANT, a=()) -> OrcBc:
    b = OrcBc()
    b.type = int(type_)
    for i, v in enumerate(q):
        b.q[i] = float(v)
    b.profile = int(profile)
    for i, v in enumerate(a):
        b.a[i] = float(v)
    return b


def _bcs(bcs):
    if not bcs:
        return 0, None
    arr = (OrcBc * len(bcs))(*bcs)
    return len(bcs), arr


class OrcPhys(ctypes.Structure):
    _fields_ = [
        ("gamma", ctypes.c_double),
        ("Pr", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("R_gas", ctypes.c_double),
        ("viscous", ctypes.c_int),
        ("sutherland", ctypes.c_int),
        ("surface_flux", ctypes.c_int),
        ("bug_compat", ctypes.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so, liboracle_fma.so and liboracle_ld.so with make (g++)."""
    if force or not all(os.path.exists(p) for p in (_LIB_PATH, _FMA_PATH, _LD_PATH)):
        args = ["make", "-s"] + (["-B"] if force else []) + ["all"]
        subprocess.run(args, cwd=_HERE, check=True)
    return _LIB_PATH


def _configure(L):
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    pp = ctypes.POINTER(OrcPhys)
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_set_threads.restype = ctypes.c_int
    L.orc_gll.argtypes = [ctypes.c_int, dp, dp]
    L.orc_dmatrices.argtypes = [ctypes.c_int, dp, dp, dp]
    L.orc_logmean.argtypes = [ctypes.c_double, ctypes.c_double]
    L.orc_logmean.restype = ctypes.c_double
    for name in ("orc_volume_flux",):
        getattr(L, name).argtypes = [ctypes.c_int, pp, dp, dp, dp, dp, dp]
        getattr(L, name).restype = ctypes.c_double
    L.orc_face_flux.argtypes = [ctypes.c_int, pp, dp, dp, dp, dp]
    L.orc_face_flux.restype = ctypes.c_double
    L.orc_euler_flux_dot_n.argtypes = [ctypes.c_int, pp, dp, dp, dp]
    L.orc_euler_flux_dot_n.restype = ctypes.c_double
    for name in ("orc_conserv2entropy", "orc_entropy2conserv", "orc_entgrad2primgrad"):
        getattr(L, name).argtypes = [ctypes.c_int, pp, dp, dp]
        getattr(L, name).restype = None
    L.orc_viscous_flux.argtypes = [ctypes.c_int, pp, dp, dp, dp, dp, dp]
    L.orc_viscous_flux.restype = None
    L.orc_rhs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, pp, dp, dp, dp, dp]
    L.orc_rhs.restype = ctypes.c_int
    L.orc_rhs_mag.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, pp, dp, dp, dp, dp, dp, dp]
    L.orc_rhs_mag.restype = ctypes.c_int
    L.orc_rk_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, pp, ctypes.c_int,
                             ctypes.c_int, ctypes.c_double, dp]
    L.orc_rk_run.restype = ctypes.c_int
    # stage hooks (prandtl_oracle_hooks.cpp)
    L.orc_modal_basis.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ip]
    L.orc_modal_basis.restype = ctypes.c_int
    L.orc_legendre01.argtypes = [ctypes.c_int, ctypes.c_double, dp]
    L.orc_legendre01.restype = None
    L.orc_filter_constraints.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, ctypes.c_double, dp]
    L.orc_filter_constraints.restype = ctypes.c_int
    L.orc_filter_apply.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, dp, dp]
    L.orc_filter_apply.restype = ctypes.c_int
    L.orc_limiter_apply.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double]
    L.orc_limiter_apply.restype = ctypes.c_int
    L.orc_limiter_apply_gen.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double,
                                        dp]
    L.orc_limiter_apply_gen.restype = ctypes.c_int
    L.orc_rk_run_hooked.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, pp, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, ip]
    L.orc_rk_run_hooked.restype = ctypes.c_int
    bp = ctypes.POINTER(OrcBc)
    L.orc_rhs_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, pp, ctypes.c_int, bp, dp, dp, dp, dp,
                             dp, dp, dp]
    L.orc_rhs_ex.restype = ctypes.c_int
    L.orc_rk_run_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, pp, ctypes.c_int, bp, ctypes.c_int,
                                ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, dp, ip]
    L.orc_rk_run_ex.restype = ctypes.c_int
    L.orc_rhs_gen.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, dp, ip, pp, ctypes.c_int, bp, dp, dp,
                              dp, dp, dp, dp, dp]
    L.orc_rhs_gen.restype = ctypes.c_int
    L.orc_rk_run_gen.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, dp, ip, pp, ctypes.c_int, bp,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, dp, ip]
    L.orc_rk_run_gen.restype = ctypes.c_int
    L.orc_bc_outer_state.argtypes = [ctypes.c_int, pp, bp, dp, dp, dp]
    L.orc_bc_outer_state.restype = ctypes.c_int
    L.orc_bc_inviscid_flux.argtypes = [ctypes.c_int, pp, bp, dp, dp, dp]
    L.orc_bc_inviscid_flux.restype = ctypes.c_double
    for name in ("orc_bc_viscous_flux",):
        getattr(L, name).argtypes = [ctypes.c_int, pp, bp, dp, dp, dp, dp]
        getattr(L, name).restype = None
    L.orc_bc_lifting_flux.argtypes = [ctypes.c_int, pp, bp, dp, dp, dp]
    L.orc_bc_lifting_flux.restype = None
    L.orc_set_boundary_context.argtypes = [dp, ctypes.c_double]
    L.orc_set_boundary_context.restype = None
    L.orc_get_time.argtypes = []
    L.orc_get_time.restype = ctypes.c_double
    L.orc_set_time.argtypes = [ctypes.c_double]
    L.orc_set_time.restype = None
    L.orc_bc_state.argtypes = [ctypes.c_int, pp, bp, dp, ctypes.c_double, dp]
    L.orc_bc_state.restype = ctypes.c_int
    L.orc_blend_coef.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, dp, dp]
    L.orc_blend_coef.restype = ctypes.c_int
    L.orc_rhs_blend_mag.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, pp, dp, dp, dp, dp, dp, dp, dp]
    L.orc_rhs_blend_mag.restype = ctypes.c_int
    L.orc_rk_run_blend.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, pp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   dp, ip]
    L.orc_rk_run_blend.restype = ctypes.c_int
    return L


def lib(fma: bool = False):
    """The oracle library; fma=True gives the FMA-contracted build of the same source."""
    global _lib, _lib_fma
    if _lib is None or _lib_fma is None:
        if not (os.path.exists(_LIB_PATH) and os.path.exists(_FMA_PATH)):
            build()
        _lib = _configure(ctypes.CDLL(_LIB_PATH))
        _lib_fma = _configure(ctypes.CDLL(_FMA_PATH))
    return _lib_fma if fma else _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if a is not None else None


def _lattice(mesh):
    # entry points without the general-mesh arguments take aligned affine lattices only
    assert not getattr(mesh, "general", False), "a general mesh needs rhs_ex / rk_run_ex"
    return mesh


def _gen(mesh):
    ng = getattr(mesh, "node_geom", None)
    fo = getattr(mesh, "face_orient", None)
    ng = None if ng is None else np.ascontiguousarray(ng, dtype=np.float64)
    fo = None if fo is None else np.ascontiguousarray(fo, dtype=np.int32)
    return ng, fo


def phys(gamma=1.4, Pr=0.72, mu=0.0, R_gas=287.0, viscous=False, sutherland=False,
         surface_flux=SURF_CHANDRASHEKAR, bug_compat=0) -> OrcPhys:
    return OrcPhys(gamma, Pr, mu, R_gas, int(viscous), int(sutherland), int(surface_flux), int(bug_compat))


def set_threads(n: int) -> int:
    return lib().orc_set_threads(int(n))


def gll(np_: int):
    x = np.zeros(np_)
    w = np.zeros(np_)
    assert lib().orc_gll(np_, _dp(x), _dp(w)) == 0
    return x, w


def dmatrices(np_: int):
    D, Dh, Dh2 = (np.zeros((np_, np_)) for _ in range(3))
    assert lib().orc_dmatrices(np_, _dp(D), _dp(Dh), _dp(Dh2)) == 0
    return D, Dh, Dh2


def logmean(x: float, y: float) -> float:
    return lib().orc_logmean(float(x), float(y))


def volume_flux(dim, ph, s1, s2, m1, m2):
    s1, s2, m1, m2 = (np.ascontiguousarray(a, dtype=np.float64) for a in (s1, s2, m1, m2))
    F = np.zeros(dim + 2)
    lam = lib().orc_volume_flux(dim, ctypes.byref(ph), _dp(s1), _dp(s2), _dp(m1), _dp(m2), _dp(F))
    return F, lam


def face_flux(dim, ph, s1, s2, nor):
    s1, s2, nor = (np.ascontiguousarray(a, dtype=np.float64) for a in (s1, s2, nor))
    F = np.zeros(dim + 2)
    lam = lib().orc_face_flux(dim, ctypes.byref(ph), _dp(s1), _dp(s2), _dp(nor), _dp(F))
    return F, lam


def euler_flux_dot_n(dim, ph, s, nor):
    s, nor = (np.ascontiguousarray(a, dtype=np.float64) for a in (s, nor))
    F = np.zeros(dim + 2)
    lam = lib().orc_euler_flux_dot_n(dim, ctypes.byref(ph), _dp(s), _dp(nor), _dp(F))
    return F, lam


def conserv2entropy(dim, ph, s):
    s = np.ascontiguousarray(s, dtype=np.float64)
    w = np.zeros(dim + 2)
    lib().orc_conserv2entropy(dim, ctypes.byref(ph), _dp(s), _dp(w))
    return w


def entropy2conserv(dim, ph, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    s = np.zeros(dim + 2)
    lib().orc_entropy2conserv(dim, ctypes.byref(ph), _dp(w), _dp(s))
    return s


def entgrad2primgrad(dim, ph, s, g):
    s = np.ascontiguousarray(s, dtype=np.float64)
    g = np.array(g, dtype=np.float64)
    lib().orc_entgrad2primgrad(dim, ctypes.byref(ph), _dp(s), _dp(g))
    return g


def viscous_flux(dim, ph, s, gx, gy, gz=None):
    s, gx, gy = (np.ascontiguousarray(a, dtype=np.float64) for a in (s, gx, gy))
    gz = np.ascontiguousarray(gz, dtype=np.float64) if gz is not None else None
    flux = np.zeros((dim + 2, dim))
    lib().orc_viscous_flux(dim, ctypes.byref(ph), _dp(s), _dp(gx), _dp(gy), _dp(gz), _dp(flux))
    return flux


def rhs(mesh, ph, u, want_grad=False, fma=False):
    """dudt = L(u) on ``mesh`` (a prandtl_amd.mesh.Mesh); returns (dudt, lambda_max[, grad])."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    dudt = np.zeros_like(u)
    grad = np.zeros((mesh.dim,) + u.shape) if want_grad else None
    lam = ctypes.c_double(0.0)
    rc = lib(fma).orc_rhs(mesh.dim, mesh.order, mesh.num_elements, _ip(_lattice(mesh).neighbors), _dp(mesh.geom),
                       ctypes.byref(ph), _dp(u), _dp(dudt), _dp(grad), ctypes.byref(lam))
    assert rc == 0, rc
    if want_grad:
        return dudt, lam.value, grad
    return dudt, lam.value


def rhs_mag(mesh, ph, u, want_grad=False):
    """RHS with the rounding scale of each entry: (dudt, lambda, mag[, grad, gmag]).
    mag is the same assembly with every term in absolute value (see orc_rhs_mag)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    dudt = np.zeros_like(u)
    mag = np.zeros_like(u)
    grad = np.zeros((mesh.dim,) + u.shape) if want_grad else None
    gmag = np.zeros((mesh.dim,) + u.shape) if want_grad else None
    lam = ctypes.c_double(0.0)
    rc = lib().orc_rhs_mag(mesh.dim, mesh.order, mesh.num_elements, _ip(_lattice(mesh).neighbors), _dp(mesh.geom),
                           ctypes.byref(ph), _dp(u), _dp(dudt), _dp(grad), ctypes.byref(lam), _dp(mag), _dp(gmag))
    assert rc == 0, rc
    if want_grad:
        return dudt, lam.value, mag, grad, gmag
    return dudt, lam.value, mag


def rhs_truth(mesh, ph, u):
    """The operator in extended precision (liboracle_ld.so, long double throughout): (hi, lo, lambda)
    with hi the value rounded to double and lo the rounding residual, so that the error of a double
    result x is (x - hi) - lo.  Affine periodic meshes."""
    global _lib_ld
    if _lib_ld is None:
        if not os.path.exists(_LD_PATH):
            build()
        L = ctypes.CDLL(_LD_PATH)
        L.orc_ld_rhs.restype = ctypes.c_int
        _lib_ld = L
    u = np.ascontiguousarray(u, dtype=np.float64)
    hi = np.zeros_like(u)
    lo = np.zeros_like(u)
    lam = ctypes.c_double(0.0)
    rc = _lib_ld.orc_ld_rhs(mesh.dim, mesh.order, mesh.num_elements, _ip(_lattice(mesh).neighbors), _dp(mesh.geom),
                            ctypes.byref(ph), _dp(u), _dp(hi), _dp(lo), ctypes.byref(lam))
    assert rc == 0, rc
    return hi, lo, lam.value


def truth_error(x, hi, lo) -> float:
    """||x - truth||_inf / ||truth||_inf for a double result x and rhs_truth's (hi, lo)."""
    return float(np.abs((x - hi) - lo).max() / np.abs(hi).max())


def scaled_error(a, ref, mag) -> float:
    """max |a - ref| / (eps * mag): the difference in units of the entry's rounding scale."""
    eps = np.finfo(np.float64).eps
    return float(np.max(np.abs(a - ref) / (eps * mag + np.finfo(np.float64).tiny)))


def rk_run(mesh, ph, u, scheme: int, nsteps: int, dt: float, fma=False):
    """Advance u (copy) by nsteps fixed steps; scheme 0 = SSP-RK3, 1 = LSRK4(5)."""
    u = np.array(u, dtype=np.float64, copy=True)
    rc = lib(fma).orc_rk_run(mesh.dim, mesh.order, mesh.num_elements, _ip(_lattice(mesh).neighbors), _dp(mesh.geom),
                          ctypes.byref(ph), int(scheme), int(nsteps), float(dt), _dp(u))
    assert rc == 0, rc
    return u


def rounding_floor(a, b) -> float:
    """Relative max-norm difference of two oracle builds' results (their rounding spread)."""
    return float(np.abs(a - b).max() / np.abs(a).max())


# ---------------------------------------------------------------------------
# stage hooks: entropy filter and positivity limiter (prandtl_oracle_hooks.cpp)
# ---------------------------------------------------------------------------
HOOK_FILTER, HOOK_LIMITER = 1, 2
LIMITER_EPS = 1e-8


def legendre01(p: int, x: float):
    """P_n(2x - 1), n = 0..p (MFEM Poly_1D::CalcLegendre)."""
    L = np.zeros(p + 1)
    lib().orc_legendre01(int(p), float(x), _dp(L))
    return L


def modal_basis(dim: int, order: int):
    """(V, V^-1, maxdeg) of the ModalBasis on the tensor GLL nodes."""
    n = (order + 1) ** dim
    V, Vi = np.zeros((n, n)), np.zeros((n, n))
    md = np.zeros(n, dtype=np.int32)
    assert lib().orc_modal_basis(dim, order, _dp(V), _dp(Vi), _ip(md)) == 0
    return V, Vi, md


def filter_constraints(mesh, u, gamma=1.4):
    """Per-element entropy bounds: min nodal rho ln(p rho^-gamma) over the element and its face
    neighbours (EntropyFilter::GetEntropyConstraints with the intended element minima)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    b = np.zeros(mesh.num_elements)
    assert lib().orc_filter_constraints(mesh.dim, mesh.order, mesh.num_elements, _ip(mesh.neighbors), _dp(u),
                                        float(gamma), _dp(b)) == 0
    return b


def filter_apply(mesh, u, bound, gamma=1.4):
    """EntropyFilter::FilterSolution on a copy of u: (filtered u, zeta per element, count)."""
    u = np.array(u, dtype=np.float64, copy=True)
    z = np.zeros(mesh.num_elements)
    bound = np.ascontiguousarray(bound, dtype=np.float64)
    c = lib().orc_filter_apply(mesh.dim, mesh.order, mesh.num_elements, _dp(u), float(gamma), _dp(bound), _dp(z))
    assert c >= 0, c
    return u, z, c


def limiter_apply(mesh, u, gamma=1.4, eps=LIMITER_EPS):
    """Zhang-Shu positivity limiter on a copy of u: (limited u, count)."""
    u = np.array(u, dtype=np.float64, copy=True)
    ng, _ = _gen(mesh)
    if ng is not None:
        c = lib().orc_limiter_apply_gen(mesh.dim, mesh.order, mesh.num_elements, _dp(u), float(gamma), float(eps),
                                        _dp(ng))
    else:
        c = lib().orc_limiter_apply(mesh.dim, mesh.order, mesh.num_elements, _dp(u), float(gamma), float(eps))
    assert c >= 0, c
    return u, c


def rk_run_hooked(mesh, ph, u, scheme: int, nsteps: int, dt: float, hooks: int):
    """rk_run with the stage hooks (filter constraints from the step's initial state; after every
    stage the filter, then the limiter).  Returns (u, [elements filtered, elements limited])."""
    u = np.array(u, dtype=np.float64, copy=True)
    counts = np.zeros(2, dtype=np.int32)
    rc = lib().orc_rk_run_hooked(mesh.dim, mesh.order, mesh.num_elements, _ip(_lattice(mesh).neighbors), _dp(mesh.geom),
                                 ctypes.byref(ph), int(scheme), int(nsteps), float(dt), int(hooks), _dp(u),
                                 _ip(counts))
    assert rc == 0, rc
    return u, counts


# ---------------------------------------------------------------------------
# subcell finite-volume blending (prandtl_oracle.cpp, prandtl_oracle_hooks.cpp)
ALPHA_MAX = 0.5     # Simulation.cpp:477-484 (runtime "alpha_max", default 0.5)
ALPHA_MIN = 0.001   # DGSEMOperator.hpp:99


def blend_coef(mesh, u, gamma=1.4, alpha_max=ALPHA_MAX, alpha_min=ALPHA_MIN):
    """Persson-Peraire indicator and blending coefficient per element: (alpha, eta)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    a = np.zeros(mesh.num_elements)
    eta = np.zeros(mesh.num_elements)
    assert lib().orc_blend_coef(mesh.dim, mesh.order, mesh.num_elements, _dp(u), float(gamma), float(alpha_max),
                                float(alpha_min), _dp(a), _dp(eta)) == 0
    return a, eta


def rhs_blend(mesh, ph, u, alpha, want_grad=False, fma=False):
    """Blended RHS with the given per-element coefficients: (dudt, lambda, mag[, grad, gmag])."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    assert alpha.shape == (mesh.num_elements,)
    dudt = np.zeros_like(u)
    mag = np.zeros_like(u)
    grad = np.zeros((mesh.dim,) + u.shape) if want_grad else None
    gmag = np.zeros((mesh.dim,) + u.shape) if want_grad else None
    lam = ctypes.c_double(0.0)
    rc = lib(fma).orc_rhs_blend_mag(mesh.dim, mesh.order, mesh.num_elements, _ip(_lattice(mesh).neighbors), _dp(mesh.geom),
                                    ctypes.byref(ph), _dp(u), _dp(alpha), _dp(dudt), _dp(grad), ctypes.byref(lam),
                                    _dp(mag), _dp(gmag))
    assert rc == 0, rc
    if want_grad:
        return dudt, lam.value, mag, grad, gmag
    return dudt, lam.value, mag


def rk_run_blend(mesh, ph, u, scheme: int, nsteps: int, dt: float, hooks: int = 0,
                 alpha_max=ALPHA_MAX, alpha_min=ALPHA_MIN):
    """rk_run with blending in every RHS (coefficients from the stage state), and the stage hooks.
    Returns (u, [elements filtered, elements limited])."""
    u = np.array(u, dtype=np.float64, copy=True)
    counts = np.zeros(2, dtype=np.int32)
    rc = lib().orc_rk_run_blend(mesh.dim, mesh.order, mesh.num_elements, _ip(_lattice(mesh).neighbors), _dp(mesh.geom),
                                ctypes.byref(ph), int(scheme), int(nsteps), float(dt), int(hooks), 1,
                                float(alpha_max), float(alpha_min), _dp(u), _ip(counts))
    assert rc == 0, rc
    return u, counts


# ---------------------------------------------------------------------------
# boundary conditions (prandtl_oracle.cpp)
def _xyz(mesh, bcs):
    """Node coordinates [e][node][dim] when a condition's state has a front (else None)."""
    if not bcs or not any(b.profile == PROFILE_FRONT for b in bcs):
        return None
    return np.ascontiguousarray(np.transpose(mesh.node_coords(), (1, 2, 0)), dtype=np.float64)


def rhs_ex(mesh, ph, u, bcs=None, alpha=None, want_grad=False, fma=False, time=0.0):
    """The general RHS with boundary conditions (and blending) at `time` (the function-valued
    boundary states): (dudt, lambda, mag[, grad, gmag])."""
    xyz = _xyz(mesh, bcs)
    lib(fma).orc_set_boundary_context(_dp(xyz), float(time))
    try:
        return _rhs_ex(mesh, ph, u, bcs, alpha, want_grad, fma)
    finally:
        lib(fma).orc_set_boundary_context(None, 0.0)


def _rhs_ex(mesh, ph, u, bcs, alpha, want_grad, fma):
    u = np.ascontiguousarray(u, dtype=np.float64)
    dudt = np.zeros_like(u)
    mag = np.zeros_like(u)
    grad = np.zeros((mesh.dim,) + u.shape) if want_grad else None
    gmag = np.zeros((mesh.dim,) + u.shape) if want_grad else None
    if alpha is not None:
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    lam = ctypes.c_double(0.0)
    nbc, arr = _bcs(bcs)
    ng, fo = _gen(mesh)
    rc = lib(fma).orc_rhs_gen(mesh.dim, mesh.order, mesh.num_elements, _ip(mesh.neighbors), _dp(mesh.geom), _dp(ng),
                              _ip(fo), ctypes.byref(ph), nbc, arr, _dp(u), _dp(alpha), _dp(dudt), _dp(grad),
                              ctypes.byref(lam), _dp(mag), _dp(gmag))
    assert rc == 0, rc
    if want_grad:
        return dudt, lam.value, mag, grad, gmag
    return dudt, lam.value, mag


def rk_run_ex(mesh, ph, u, scheme: int, nsteps: int, dt: float, bcs=None, hooks: int = 0, blend=False,
              alpha_max=ALPHA_MAX, alpha_min=ALPHA_MIN, t0=0.0, return_time=False):
    """The general stepper (boundary conditions, blending, stage hooks) from time t0, each stage at
    its stage time: (u, counts[, t_end])."""
    u = np.array(u, dtype=np.float64, copy=True)
    counts = np.zeros(2, dtype=np.int32)
    nbc, arr = _bcs(bcs)
    ng, fo = _gen(mesh)
    xyz = _xyz(mesh, bcs)
    L = lib()
    L.orc_set_boundary_context(_dp(xyz), float(t0))
    try:
        rc = L.orc_rk_run_gen(mesh.dim, mesh.order, mesh.num_elements, _ip(mesh.neighbors), _dp(mesh.geom), _dp(ng),
                              _ip(fo), ctypes.byref(ph), nbc, arr, int(scheme), int(nsteps), float(dt), int(hooks),
                              int(blend), float(alpha_max), float(alpha_min), _dp(u), _ip(counts))
        t_end = L.orc_get_time()
    finally:
        L.orc_set_boundary_context(None, 0.0)
    assert rc == 0, rc
    return (u, counts, t_end) if return_time else (u, counts)


def bc_state(dim, ph, b, x, t):
    """The state of condition b at the point x and time t (constant or profile)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    s2 = np.zeros(dim + 2)
    assert lib().orc_bc_state(dim, ctypes.byref(ph), ctypes.byref(b), _dp(x), float(t), _dp(s2)) == 0
    return s2


def bc_outer_state(dim, ph, b, s1, nor):
    """The BC's outer (ghost) state for the trace s1 and outward normal nor."""
    s1, nor = (np.ascontiguousarray(a, dtype=np.float64) for a in (s1, nor))
    s2 = np.zeros(dim + 2)
    assert lib().orc_bc_outer_state(dim, ctypes.byref(ph), ctypes.byref(b), _dp(s1), _dp(nor), _dp(s2)) == 0
    return s2


def bc_inviscid_flux(dim, ph, b, s1, nor):
    """(f, wave speed) of the BC's inviscid flux at one face point (outward normal nor)."""
    s1, nor = (np.ascontiguousarray(a, dtype=np.float64) for a in (s1, nor))
    f = np.zeros(dim + 2)
    lam = lib().orc_bc_inviscid_flux(dim, ctypes.byref(ph), ctypes.byref(b), _dp(s1), _dp(nor), _dp(f))
    return f, lam


def bc_viscous_flux(dim, ph, b, s1, dq, nor):
    """F_v . nor of the BC at one face point; dq: (dim, neq) primitive gradients."""
    s1, dq, nor = (np.ascontiguousarray(a, dtype=np.float64) for a in (s1, dq, nor))
    fv = np.zeros(dim + 2)
    lib().orc_bc_viscous_flux(dim, ctypes.byref(ph), ctypes.byref(b), _dp(s1), _dp(dq), _dp(nor), _dp(fv))
    return fv


def bc_lifting_flux(dim, ph, b, w1, nor):
    """The BC's BR1 lifting flux at one face point (w1: the element's entropy variables)."""
    w1, nor = (np.ascontiguousarray(a, dtype=np.float64) for a in (w1, nor))
    fl = np.zeros(dim + 2)
    lib().orc_bc_lifting_flux(dim, ctypes.byref(ph), ctypes.byref(b), _dp(w1), _dp(nor), _dp(fl))
    return fl
