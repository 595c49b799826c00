/* prandtl_amd.h -- C ABI of the MI355X (gfx950) DGSEM right-hand-side and RK-stage library.
 *
 * This is the drop-in boundary beneath the reference's operator API: each entry
 * point names the reference interface it replaces (paths relative to the Prandtl
 * repository root).  Plain pointers and sizes only; no exceptions or aborts cross
 * it.  Return value 0 is success, negative values are errors, and
 * prandtl_last_error() describes the last error of the calling thread.
 *
 * Device buffers are caller-owned and use the reference's byNODES layout
 * (src/Simulation/Simulation.hpp:64):
 *     u[v * NE * n + e * n + node],  node = i + Np * (j + Np * k),  n = Np^dim,
 * with v over the dim + 2 conserved variables (rho, rho u, [rho v, [rho w,]] E).
 * The handle owns its scratch (gradient/viscous-flux buffers, RK registers).
 * One handle is not thread-safe (like the reference's DGSEMOperator, whose Mult
 * mutates internal state); distinct handles are independent.  Every call runs on
 * the given HIP stream (NULL = the default stream).
 */
#ifndef PRANDTL_AMD_H
#define PRANDTL_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct prandtl_op_s *prandtl_op;
typedef void *prandtl_stream; /* a hipStream_t */

enum {
    PRANDTL_OK = 0,
    PRANDTL_ERR_ARG = -1,         /* invalid argument */
    PRANDTL_ERR_UNSUPPORTED = -2, /* valid request outside what this library implements */
    PRANDTL_ERR_HIP = -3,         /* HIP runtime error */
    PRANDTL_ERR_MESH = -4,        /* inconsistent mesh description */
    PRANDTL_ERR_NAN = -5          /* non-finite state detected */
};

/* Surface (Riemann) flux plug-ins -- ref: NumericalFlux::ComputeFaceFlux,
 * src/RiemannSolvers/NumericalFlux.hpp:25.  CHANDRASHEKAR is the reference's
 * ChandrashekarFlux::ComputeFaceFlux (ChandrashekarFlux.cpp:85-161). */
enum { PRANDTL_SURF_CHANDRASHEKAR = 0, PRANDTL_SURF_EC = 1, PRANDTL_SURF_RUSANOV = 2 };
enum { PRANDTL_FP64 = 0, PRANDTL_MIXED = 1 };
/* Reproduce the reference as written instead of the intended formula:
 *   A1: the zeta-direction flux store H_inviscid(id2).SetCol(k, g) at DGSEMIntegrator.cpp:702;
 *   A2: the viscous volume flux read at the stale node id1 (DGSEMIntegrator.cpp:620);
 *   A3: the 3D dissipation term "rho_mean * dw + w_mean" at ChandrashekarFlux.cpp:131. */
enum { PRANDTL_BUG_A1 = 1, PRANDTL_BUG_A2 = 2, PRANDTL_BUG_A3 = 4 };
/* ODE solvers -- ref: the MFEM ODESolver chosen at src/Simulation/Simulation.cpp:177-206
 * and stepped at :1010.  SSPRK3 is "RK3SSP" (stage update fused into the RHS kernel);
 * LSRK45 is the Carpenter-Kennedy 2N-storage RK4(3)5 (fused); FORWARD_EULER, RK2 (MFEM's
 * RK2Solver with its default a = 2/3) and RK4 (MFEM's classic RK4Solver tableau) are
 * "ForwardEuler", "RK2" and "RK4": RHS evaluations plus one fused stage-combination pass. */
enum { PRANDTL_SSPRK3 = 0, PRANDTL_LSRK45 = 1, PRANDTL_FORWARD_EULER = 2, PRANDTL_RK2 = 3, PRANDTL_RK4 = 4 };

/* Boundary conditions -- the reference's BdrFaceIntegrator subclasses
 * (src/Integrators/BoundaryConditions/, chosen per boundary attribute at
 * src/Simulation/Simulation.cpp:505-800), with constant values q:
 *   SLIP_WALL          "slip"               (none; Euler only, as in the reference)
 *   SYMMETRY           "symmetry" / "axis"  (none; Euler only)
 *   SUPERSONIC_INFLOW  q[0..neq-1] conserved state
 *   SUPERSONIC_OUTFLOW (none)
 *   SPECIFIED_STATE    q[0..neq-1] conserved state
 *   RIEMANN_INVARIANT  q[0..neq-1] primitive far-field state (rho, velocity, p)
 *   SUBSONIC_INFLOW_RV q[0] rho, q[1..dim] velocity
 *   SUBSONIC_INFLOW_PTTT q[0] total pressure, q[1] total temperature, q[2] theta, q[3] phi
 *                      (degrees), q[4] cp
 *   SUBSONIC_OUTFLOW_P q[0] pressure
 *   NOSLIP_ADIABATIC   q[0..dim-1] wall velocity, q[dim] normal heat flux
 *   NOSLIP_ISOTHERMAL  q[0..dim-1] wall velocity, q[dim] wall temperature (with R_gas) */
enum {
    PRANDTL_BC_SLIP_WALL = 1, PRANDTL_BC_SYMMETRY = 2, PRANDTL_BC_SUPERSONIC_INFLOW = 3,
    PRANDTL_BC_SUPERSONIC_OUTFLOW = 4, PRANDTL_BC_SPECIFIED_STATE = 5, PRANDTL_BC_RIEMANN_INVARIANT = 6,
    PRANDTL_BC_SUBSONIC_INFLOW_RV = 7, PRANDTL_BC_SUBSONIC_INFLOW_PTTT = 8, PRANDTL_BC_SUBSONIC_OUTFLOW_P = 9,
    PRANDTL_BC_NOSLIP_ADIABATIC = 10, PRANDTL_BC_NOSLIP_ISOTHERMAL = 11
};
/* Boundary states that vary over the boundary and in time -- the reference's function-valued
 * conditions (ConditionFactory, src/Simulation/ConditionFactory.hpp:16-40, registered by a test
 * case and evaluated at each face point's physical coordinates x and the operator's time t,
 * Simulation.cpp:664-790).  For SUPERSONIC_INFLOW and SPECIFIED_STATE, whose value is a state:
 *   PROFILE_CONSTANT  the state q (a zero-initialised profile);
 *   PROFILE_FRONT     a planar front: q where x[i0] < c0 + (x[i1] + c1 t) / c2, else the state
 *                     a[5 .. 4+neq] ahead of it; a[0] = i0, a[1] = i1, a[2] = c0, a[3] = c1,
 *                     a[4] = c2 (DoubleMachReflectionTopBCFunction, TestCases/Euler/2D/
 *                     DoubleMachReflection/DoubleMachReflection.hpp:32-52: i0 = 0, i1 = 1,
 *                     c0 = 1/6, c1 = 20, c2 = sqrt 3);
 *   PROFILE_RAMP      an inflow whose velocity rises over a time t_r:
 *                     r = clamp((1 + tanh((t - t_r/2) / (t_r/10))) / 2, 0, 1) and the state
 *                     (rho, rho v r, p / (gamma - 1) + rho |v r|^2 / 2) with a[0] = t_r,
 *                     a[1] = rho, a[2 .. 1+dim] = v, a[2+dim] = p (NagashimaRampingInletBC,
 *                     TestCases/Euler/2D/Nagashima_Ramjet/Nagashima_Ramjet.hpp:40-53).
 * A non-constant profile needs the mesh's node coordinates (prandtl_mesh_desc.node_coords). */
enum { PRANDTL_PROFILE_CONSTANT = 0, PRANDTL_PROFILE_FRONT = 1, PRANDTL_PROFILE_RAMP = 2 };
typedef struct {
    int type;     /* PRANDTL_BC_* */
    double q[8];
    int profile;  /* PRANDTL_PROFILE_* (SUPERSONIC_INFLOW, SPECIFIED_STATE; 0 otherwise) */
    double a[16]; /* the profile's parameters */
} prandtl_bc;

/* Conforming mesh -- the reference's ParMesh and its finite-element space as the
 * integrators see them (src/Simulation/Simulation.cpp:272-330, DGSEMIntegrator.cpp:8-149).
 *   neighbors[e * 2*dim + 2*d + s]: element across local face (d, s) of e,
 *       s = 0 on the xi_d = 0 side, s = 1 on the xi_d = 1 side; -(k+1) marks a
 *       boundary face with condition bc[k] (the reference's boundary attributes);
 *   geometry[e * (1 + dim*dim)]: det J, then adj J row-major (MFEM Weight() and
 *       AdjugateJacobian(), DGSEMIntegrator.cpp:254-256) of an affine element;
 *   bc[0..num_bc-1]: the boundary conditions (num_bc = 0 for a periodic mesh).
 * Affine lattice (node_geometry and face_orientation NULL): node (.., Np-1 in d, ..)
 * of e coincides with node (.., 0 in d, ..) of its +d neighbour.
 * General mesh (curved or unstructured; either array non-NULL):
 *   node_geometry[(e * n + node) * (1 + dim*dim)]: det J and adj J at every node
 *       (the reference evaluates them per node, DGSEMIntegrator.cpp:254-276); NULL:
 *       geometry at every node (geometry may be NULL when this is given);
 *   face_orientation[e * 2*dim + f]: the neighbour's local face f' = 2 d' + s' in
 *       bits 0-2; bit 3 flips the first face coordinate, bit 4 the second, bit 5 swaps
 *       them: point (t1, t2) of face f (t1, t2 the indices along the other reference
 *       directions, in increasing order) is the neighbour's point (u1, u2) of face f'
 *       with (a, b) = swap ? (t2, t1) : (t1, t2), u1 = flip1 ? Np-1-a : a,
 *       u2 = flip2 ? Np-1-b : b; NULL: the aligned-lattice matching (f' = f ^ 1).
 * node_coords[(e * n + node) * dim]: the physical coordinates of every node (MFEM's
 *     Tr.Transform of the face points, BdrFaceIntegrator.cpp:83); needed only when a
 *     boundary condition has a non-constant profile (NULL otherwise).
 * All arrays are host pointers, copied at create time.  Boundary faces and general
 * meshes run in FP64 (mixed precision and the adjoint refuse them). */
typedef struct {
    int dim;
    int order;
    int num_elements;
    const int *neighbors;
    const double *geometry;
    int num_bc;
    const prandtl_bc *bc;
    const double *node_geometry;
    const int *face_orientation;
    const double *node_coords;
} prandtl_mesh_desc;

/* ref: src/Physics/Physics.hpp:10-38 and NavierStokesFlux(dim, gamma, Pr, mu, ...),
 * src/Fluxes/NavierStokesFlux.hpp:20-25 */
typedef struct {
    double gamma;
    double Pr;
    double mu;
    double R_gas;
    int viscous;    /* 0: Euler; 1: Navier-Stokes with BR1 lifting (PARABOLIC) */
    int sutherland; /* 1: Sutherland viscosity (SUTHERLAND) */
} prandtl_physics;

/* Stage hooks the steppers (prandtl_rk_step, prandtl_rk_run) apply after every RK stage:
 * the entropy filter (bounds from the state at the start of each step), then the
 * positivity limiter -- the Filter and Limiter hooks of src/Filters/Filter.hpp:11-30 and
 * src/Limiters/Limiter.hpp:10-16.  LSRK4(5) and SSP-RK3 detect the elements to filter or
 * limit in the stage update itself (all shapes but 3D p = 4); PRANDTL_HOOK_UNFUSED asks for
 * separate detection scans after every stage instead (the same elements and results, more
 * launches: for checking and measurement). */
enum { PRANDTL_HOOK_FILTER = 1, PRANDTL_HOOK_LIMITER = 2, PRANDTL_HOOK_UNFUSED = 4 };

/* Subcell finite-volume blending (the reference built with SUBCELL_FV_BLENDING): every RHS
 * evaluation computes the per-element coefficient alpha from its input state (Persson-Peraire
 * indicator on rho p) and blends the inviscid volume term as (1 - alpha) DG + alpha FV --
 * ref: DGSEMOperator::Mult, DGSEMOperator.cpp:547-549; DGSEMIntegrator.cpp:240-241, 394-396,
 * 972-1177.  FP64 only.  The reference's defaults: alpha_max 0.5 (runtime "alpha_max",
 * Simulation.cpp:477-484), alpha_min 0.001 (DGSEMOperator.hpp:99); they are taken as given
 * here (a zero-initialised struct has alpha_max = 0, which blends nothing). */
typedef struct {
    int surface_flux; /* PRANDTL_SURF_* */
    int precision;    /* PRANDTL_FP64 | PRANDTL_MIXED */
    int bug_compat;   /* PRANDTL_BUG_* bitmask */
    int stage_hooks;  /* PRANDTL_HOOK_* bitmask */
    int blending;     /* 1: subcell finite-volume blending */
    double alpha_max; /* in [0, 1] */
    double alpha_min; /* in [0, 0.5) */
} prandtl_opts;

/* Build an operator -- replaces the DGSEMOperator constructor,
 * src/Operators/DGSEMOperator.hpp:87-99 (with its DGSEMIntegrator,
 * DGSEMIntegrator.cpp:8-149, and the LiftingBR1 set-up, LiftingScheme.cpp:6-42). */
int prandtl_op_create(const prandtl_mesh_desc *mesh, const prandtl_physics *phys, const prandtl_opts *opts,
                      prandtl_op *out);

/* ref: DGSEMOperator::~DGSEMOperator, DGSEMOperator.cpp:52-58 */
void prandtl_op_destroy(prandtl_op op);

/* Number of doubles in a state vector -- ref: vfes->GetTrueVSize(), DGSEMOperator.cpp:18 */
size_t prandtl_op_size(prandtl_op op);

/* dudt = L(u), dudt fully overwritten -- ref: DGSEMOperator::Mult(const Vector &u,
 * Vector &dudt) const, DGSEMOperator.cpp:537-591 (mfem::TimeDependentOperator::Mult). */
int prandtl_op_mult(prandtl_op op, const double *u, double *dudt, prandtl_stream stream);

/* Mult plus the primitive gradients the reference leaves in its dudx/dudy/dudz grid
 * functions (DGSEMOperator.cpp:568-570).  grad holds dim state-sized blocks:
 * grad[c * size + v * NE * n + e * n + node] = d(rho, vel, p)_v / dx_c.
 * Navier-Stokes operators only. */
int prandtl_op_mult_grad(prandtl_op op, const double *u, double *dudt, double *grad, prandtl_stream stream);

/* The time at which prandtl_op_mult and prandtl_op_mult_grad evaluate time-dependent boundary
 * states (default 0) -- ref: TimeDependentOperator::SetTime / DGSEMOperator::GetTimeRef,
 * DGSEMOperator.hpp:111-114 (the boundary integrators hold a reference to that time,
 * Simulation.cpp:522).  The steppers take their stage times from their own t argument instead:
 * t, t + dt, t + dt/2 for SSPRK3 (MFEM RK3SSPSolver), t + c_i dt for LSRK45 (the
 * Carpenter-Kennedy c_i), t, t + a dt for RK2, t, t + dt/2, t + dt/2, t + dt for RK4, with t
 * advanced by dt after each step. */
int prandtl_op_set_time(prandtl_op op, double t);
int prandtl_op_get_time(prandtl_op op, double *t);

/* Max characteristic speed max(|V| + c) of the last Mult (Euler and NS) --
 * ref: DGSEMOperator::GetMaxCharSpeed, DGSEMOperator.hpp:106-109.  Waits for the stream the
 * last Mult ran on; prandtl_op_max_char_speed does not change it. */
int prandtl_op_get_max_char_speed(prandtl_op op, double *lambda_host);

/* max(|V| + c) over the nodes of u (the same quantity, computed from u alone). */
int prandtl_op_max_char_speed(prandtl_op op, const double *u, double *lambda_host, prandtl_stream stream);

/* One step of an ODE solver, u advanced in place, *t += dt -- ref: ode_solver->Step(*sol,
 * t, dt_real), src/Simulation/Simulation.cpp:1010.  The stage update is fused into the
 * RHS kernel's epilogue. */
int prandtl_rk_step(prandtl_op op, int scheme, double *u, double *t, double dt, prandtl_stream stream);

/* nsteps fixed-dt steps (the reference's time loop without output,
 * Simulation.cpp:995-1125).  Checks for a non-finite state every nan_check steps
 * (0 = never), like the density scan at Simulation.cpp:1025-1035; returns
 * PRANDTL_ERR_NAN if one is found. */
int prandtl_rk_run(prandtl_op op, int scheme, double *u, double *t, double dt, int nsteps, int nan_check,
                   prandtl_stream stream);

/* Persson-Peraire indicator eta and blending coefficient alpha of every element of u (device
 * arrays of num_elements doubles; either may be NULL) -- ref: DGSEMOperator::
 * ComputeBlendingCoefficient, DGSEMOperator.cpp:60-80, and PerssonPeraireIndicator::
 * CheckSmoothness, src/Indicators/PerssonPeraireIndicator.cpp:21-58 (the reference's alpha and
 * eta grid functions).  Uses the operator's alpha_max and alpha_min. */
int prandtl_blend_coefficients(prandtl_op op, const double *u, double *alpha, double *eta, prandtl_stream stream);

/* Entropy-filter bounds from u: per element, the minimum nodal entropy rho ln(p rho^-gamma)
 * over the element and its face neighbours -- ref: Filter::GetFilterConstraints(const
 * Vector&), src/Filters/Filter.hpp:27; EntropyFilter::GetEntropyConstraints,
 * src/Filters/EntropyFilter.cpp:81-158 (whose empty ComputeElementMinima, :275-277, is
 * implemented as intended). */
int prandtl_filter_constraints(prandtl_op op, const double *u, prandtl_stream stream);

/* Entropy filter of u in place, against the last constraints -- ref: Filter::FilterSolution
 * (Vector&), Filter.hpp:21; EntropyFilter::FilterSolution, EntropyFilter.cpp:171-242 (the
 * root-find against the element's own bound, :338).  Elements with rho, p >= 1e-8 and
 * entropy >= bound - 1e-4 at every node are untouched. */
int prandtl_filter_apply(prandtl_op op, double *u, prandtl_stream stream);

/* Positivity limiter of u in place -- ref: Limiter::LimitSolution(Vector&),
 * src/Limiters/Limiter.hpp:14 (the reference has the interface only): the Zhang-Shu scaling
 * toward the element's GLL mean so that rho, p >= 1e-8 at every node. */
int prandtl_limiter_apply(prandtl_op op, double *u, prandtl_stream stream);

/* Elements changed by the filter and by the limiter since the previous call (resets the
 * counters; synchronises the device).  The filter and limiter of one operator are stream-ordered:
 * calls on different streams must be synchronised in between; an application that finds its
 * element list out of date changes nothing and this call then returns PRANDTL_ERR_ARG. */
int prandtl_hook_counts(prandtl_op op, long long *filtered, long long *limited);

/* Discrete adjoint (no counterpart in the reference, which has no derivatives).
 * out = (dL/du)^T lam at state u: the vector-Jacobian product of the semi-discrete operator
 * (volume flux differencing, surface flux, and for Navier-Stokes with constant viscosity the
 * BR1 lifting, viscous flux and viscous face terms; FP64, Chandrashekar or EC surface flux).
 * out must not alias u or lam. */
int prandtl_op_vjp(prandtl_op op, const double *u, const double *lam, double *out, prandtl_stream stream);

/* prandtl_op_vjp without allocation (no counterpart in the reference): for callers that cannot
 * call hipMalloc -- code under HIP graph capture, tight serving loops, harnesses that forbid it --
 * while prandtl_op_vjp allocates the handle's work arrays on its first call.  The work arrays live
 * in a fixed 64 MiB device array the library holds on every GPU it runs on until the process ends
 * (per node: 56 B for a 3D Euler operator; 3D Navier-Stokes at p = 3 adds 140 B); an operator that
 * needs more is refused with PRANDTL_ERR_ARG.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race.  Same result as prandtl_op_vjp. */
int prandtl_op_vjp_noalloc(prandtl_op op, const double *u, const double *lam, double *out, prandtl_stream stream);

/* prandtl_rk_run (PRANDTL_LSRK45 or PRANDTL_SSPRK3; no stage hooks) that also keeps the state
 * at the start of every ckpt_every-step segment in the handle, for prandtl_rk_adjoint. */
int prandtl_rk_run_tape(prandtl_op op, int scheme, double *u, double *t, double dt, int nsteps, int ckpt_every,
                        prandtl_stream stream);

/* Checkpoint schedules of a taped run (no counterpart in the reference):
 *   SEGMENTS: the state at the start of every param-step segment; the reverse sweep recomputes
 *             each segment once keeping its stage inputs (memory: nsteps/param + 5 param states
 *             for LSRK4(5), 3 param for SSP-RK3; one forward run recomputed);
 *   REVOLVE:  Griewank's binomial checkpointing with param snapshot buffers (memory: param + 1
 *             checkpoints + one step's stage inputs + a work state; the fewest recomputed steps for
 *             that memory), at most 4096 steps. */
enum { PRANDTL_TAPE_SEGMENTS = 0, PRANDTL_TAPE_REVOLVE = 1 };

/* prandtl_rk_run_tape with a chosen checkpoint schedule (prandtl_rk_run_tape is SEGMENTS with
 * param = ckpt_every). */
int prandtl_rk_run_tape_schedule(prandtl_op op, int scheme, double *u, double *t, double dt, int nsteps, int schedule,
                                 int param, prandtl_stream stream);

/* The last taped run's memory and recomputation: the states the reverse sweep holds on the device
 * (checkpoints, stage inputs, work; the three adjoint registers not counted) and the forward steps
 * the last prandtl_rk_adjoint ran (-1 before it ran). */
int prandtl_rk_tape_info(prandtl_op op, long long *states, long long *recomputed_steps);

/* Identifier of the last taped run of the handle (no counterpart in the reference): unique over
 * the process's handles; 0 when there is none, when a taped run failed, or while one is being
 * written.  A caller that tapes more than one run on a handle compares it with the identifier read
 * after its own run before calling prandtl_rk_adjoint, which sweeps the handle's last tape. */
int prandtl_rk_tape_id(prandtl_op op, unsigned long long *id);

/* Reverse sweep through the last taped run: ubar holds dJ/du at the final state on entry
 * and dJ/du at the initial state on return.  Each segment is recomputed forward from its
 * checkpoint keeping its stage inputs (5 x ckpt_every states of scratch), then the stage
 * adjoints (LSRK4(5) or SSP-RK3) run backwards, fused into the operator-adjoint kernel.
 * Euler and constant-viscosity Navier-Stokes operators, FP64.  Synchronises. */
int prandtl_rk_adjoint(prandtl_op op, double *ubar, prandtl_stream stream);

/* Per-kernel timing of one RHS evaluation, with HIP events on `stream`: ms[0] = average
 * duration of the lifting/viscous-flux pass (0 for Euler), ms[1] = the flux-differencing
 * pass, each launched `iters` times back to back on input u.  epilogue 0 times the Mult
 * kernels, 1 the LSRK4(5) stage kernels (writing only handle scratch).  Instrumentation
 * in place of the reference's run timer (src/Simulation/Simulation.cpp:910-914,1162-1171). */
int prandtl_op_profile(prandtl_op op, const double *u, int epilogue, int iters, double *ms, prandtl_stream stream);

/* Launch timing inside a caller's own loop (instrumentation, like prandtl_op_profile; the reference
 * times whole runs, Simulation.cpp:910-914).  While enabled, every lifting pass (kind 0) and
 * flux-differencing pass (kind 1) this handle launches is bracketed by HIP events on its stream.
 * prandtl_op_timing_read waits for the recorded launches and returns each kind's average duration
 * in ms (ms[2]) and, when count is not NULL, the number of launches (count[2]); it then forgets
 * them.  Not for use under HIP graph capture. */
int prandtl_op_timing(prandtl_op op, int enable);
int prandtl_op_timing_read(prandtl_op op, double *ms, long long *count);

/* Device memory and copies for host programs that do not use HIP themselves (the C++ mirror
 * in prandtl_amd.hpp).  No counterpart in the reference, whose vectors live in host memory.
 * prandtl_memcpy is synchronous; kind is one of PRANDTL_COPY_*. */
enum { PRANDTL_COPY_H2D = 0, PRANDTL_COPY_D2H = 1, PRANDTL_COPY_D2D = 2 };
int prandtl_device_alloc(size_t bytes, void **ptr);
void prandtl_device_free(void *ptr);
int prandtl_memcpy(void *dst, const void *src, size_t bytes, int kind);
int prandtl_stream_synchronize(prandtl_stream stream);

/* Description of the last error on this thread. */
const char *prandtl_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PRANDTL_AMD_H */
