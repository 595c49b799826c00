// prandtl_oracle.h -- shared C declarations of the CPU oracle (test infrastructure only).
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double gamma, Pr, mu, R_gas;
    int viscous;       // 0 Euler, 1 Navier-Stokes (BR1)
    int sutherland;    // 1: Sutherland viscosity
    int surface_flux;  // 0 Chandrashekar (reference), 1 EC, 2 Rusanov
    int bug_compat;    // bit 0: A1, bit 1: A2, bit 2: A3
} orc_phys;

// A boundary condition; the meaning of q by type, and of the state profile and its parameters a,
// is listed in include/prandtl_amd.h (prandtl_bc).  A mesh marks a boundary face with neighbour
// -(k+1), k the index of its condition.
typedef struct {
    int type;
    double q[8];
    int profile;   // 0 constant, 1 front, 2 ramp
    double a[16];
} orc_bc;

// The context of function-valued boundary states for the following RHS evaluations and runs:
// node coordinates xyz[(e * n + node) * dim] (may be null when no condition has a front) and the
// time t (the runs start there and leave it at their end time).  Not thread-safe.
void orc_set_boundary_context(const double *xyz, double t);
double orc_get_time(void);
void orc_set_time(double t);
// the state of a condition at the point x and time t (constant or profile)
int orc_bc_state(int dim, const orc_phys *p, const orc_bc *bc, const double *x, double t, double *s2);

int orc_gll(int np, double *x, double *w);
int orc_rhs(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, const double *u,
            double *dudt, double *grad, double *lambda_max);
// The general RHS: boundary conditions (nbc, bcs), blending (alpha, may be null), rounding
// scales (mag, gmag, may be null)
int orc_rhs_ex(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, int nbc,
               const orc_bc *bcs, const double *u, const double *alpha, double *dudt, double *grad,
               double *lambda_max, double *mag, double *gmag);
// orc_rhs_ex on a general mesh: per-node geometry ngeom [e][node][1 + dim^2] (replaces geom
// when non-null) and face orientation codes forient [e][2 dim] (null: an aligned lattice)
int orc_rhs_gen(int dim, int order, int NE, const int *nbr, const double *geom, const double *ngeom,
                const int *forient, const orc_phys *p, int nbc, const orc_bc *bcs, const double *u,
                const double *alpha, double *dudt, double *grad, double *lambda_max, double *mag, double *gmag);
// RHS with subcell finite-volume blending; alpha: one blending coefficient per element
int orc_rhs_blend_mag(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p,
                      const double *u, const double *alpha, double *dudt, double *grad, double *lambda_max,
                      double *mag, double *gmag);

#ifdef __cplusplus
}
#endif
