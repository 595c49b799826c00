// prandtl_oracle_ld.cpp -- the oracle's operator evaluated in extended precision.
//
// TEST INFRASTRUCTURE ONLY.  The same restatement as prandtl_oracle.cpp (which it includes), with
// every quantity of the arithmetic in long double (x87 extended precision, 64-bit significand,
// eps = 2^-64): the GLL rule, the derivative matrices, the fluxes and the assembly.  Its result is
// the discrete operator of the given double-precision inputs (state, geometry, physics constants)
// to about 1e-16 relative even where the double-precision sums cancel by several digits, so both
// the double oracle and the GPU path are measured against it.
#define ORACLE_REAL long double
#define ORACLE_NO_C_API
#include "prandtl_oracle.cpp"

extern "C" {

// dudt = L(u) in extended precision, returned as a double pair: dudt_hi = the value rounded to
// double, dudt_lo (optional) = the rounding residual (value - hi, rounded to double), so that
// errors of a double result x are x - hi - lo.  lambda_max (optional): the max wave speed.
int orc_ld_rhs(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, const double *u,
               double *dudt_hi, double *dudt_lo, double *lambda_max) {
    if (dim < 2 || dim > 3 || order < 1 || order + 1 > MAXNP) return -1;
    int nn = 1;
    for (int d = 0; d < dim; d++) nn *= order + 1;
    const size_t size = (size_t)(dim + 2) * NE * nn, ng = (size_t)NE * (1 + dim * dim);
    std::vector<real> g(geom, geom + ng), x(u, u + size), y(size, 0.0L);
    Mesh M;
    MakeMesh(M, dim, order, NE, nbr, g.data());
    Phys P = MakePhys(dim, p->gamma, p->Pr, p->mu, p->R_gas, p->viscous, p->sutherland, p->surface_flux,
                      p->bug_compat);
    const real lam = RhsEval(M, P, x.data(), y.data(), nullptr);
    for (size_t i = 0; i < size; i++) {
        const double hi = (double)y[i];
        dudt_hi[i] = hi;
        if (dudt_lo) dudt_lo[i] = (double)(y[i] - (real)hi);
    }
    if (lambda_max) *lambda_max = (double)lam;
    return 0;
}

}  // extern "C"
