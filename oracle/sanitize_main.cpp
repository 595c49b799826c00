// sanitize_main.cpp -- test infrastructure: drives every entry point of the CPU oracle on small
// meshes so that a build with AddressSanitizer and UndefinedBehaviorSanitizer (`make sanitize`)
// checks the oracle's memory accesses and arithmetic.  Inputs are built here (affine lattices,
// periodic or with boundary faces in x; smooth seeded states; inadmissible nodes planted for the
// hooks); every call must succeed and return finite values.  Built with -DSAN_LD it drives the
// extended-precision operator instead.  Exit status 0 when every check passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "prandtl_oracle.h"

extern "C" {
int orc_dmatrices(int np, double *D, double *Dhat, double *Dhat2);
int orc_rk_run(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, int scheme,
               int nsteps, double dt, double *u);
int orc_blend_coef(int dim, int order, int NE, const double *u, double gamma, double alpha_max, double alpha_min,
                   double *alpha, double *eta);
int orc_modal_basis(int dim, int order, double *V, double *Vinv, int *maxdeg);
int orc_filter_constraints(int dim, int order, int NE, const int *nbr, const double *u, double gamma, double *bound);
int orc_filter_apply(int dim, int order, int NE, double *u, double gamma, const double *bound, double *zeta);
int orc_limiter_apply(int dim, int order, int NE, double *u, double gamma, double eps);
int orc_limiter_apply_gen(int dim, int order, int NE, double *u, double gamma, double eps, const double *ngeom);
int orc_rk_run_gen(int dim, int order, int NE, const int *nbr, const double *geom, const double *ngeom,
                   const int *forient, const orc_phys *p, int nbc, const orc_bc *bcs, int scheme, int nsteps,
                   double dt, int hooks, int blend, double alpha_max, double alpha_min, double *u, int *counts);
int orc_ld_rhs(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, const double *u,
               double *dudt_hi, double *dudt_lo, double *lambda_max);
}

static int g_fail = 0;

static void check(bool ok, const char *what, int dim, int order, int extra) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s (dim %d, order %d, case %d)\n", what, dim, order, extra);
        g_fail++;
    }
}

static bool finite(const std::vector<double> &v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

struct Lattice {
    int dim, order, NE, nn;
    std::vector<int> nbr;
    std::vector<double> geom, ngeom;
};

// n[d] elements per direction, element map A (sheared: A[0][1] = 0.1 h); bnd: the faces on the
// lattice's two x ends get conditions 0 and 1
static Lattice make_lattice(int dim, int order, int n, bool bnd) {
    Lattice L;
    L.dim = dim;
    L.order = order;
    L.NE = dim == 2 ? n * n : n * n * n;
    L.nn = dim == 2 ? (order + 1) * (order + 1) : (order + 1) * (order + 1) * (order + 1);
    const double h = 1.0 / n;
    double A[3][3] = {{h, 0.1 * h, 0.0}, {0.0, h, 0.0}, {0.0, 0.0, h}};
    double det, adj[3][3];
    if (dim == 2) {
        det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        adj[0][0] = A[1][1];
        adj[0][1] = -A[0][1];
        adj[1][0] = -A[1][0];
        adj[1][1] = A[0][0];
    } else {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
                adj[i][j] = A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1];
            }
        det = A[0][0] * adj[0][0] + A[0][1] * adj[1][0] + A[0][2] * adj[2][0];
    }
    const int g1 = 1 + dim * dim;
    L.nbr.resize((size_t)L.NE * 2 * dim);
    L.geom.resize((size_t)L.NE * g1);
    L.ngeom.resize((size_t)L.NE * L.nn * g1);
    for (int e = 0; e < L.NE; e++) {
        int c[3] = {e % n, (e / n) % n, e / (n * n)};
        for (int d = 0; d < dim; d++)
            for (int s = 0; s < 2; s++) {
                int cc[3] = {c[0], c[1], c[2]};
                cc[d] = (cc[d] + (s ? 1 : n - 1)) % n;
                int nb = cc[0] + n * cc[1] + (dim == 3 ? n * n * cc[2] : 0);
                if (bnd && d == 0 && c[0] == (s ? n - 1 : 0)) nb = -(s + 1);
                L.nbr[(size_t)e * 2 * dim + 2 * d + s] = nb;
            }
        double *g = &L.geom[(size_t)e * g1];
        g[0] = det;
        for (int r = 0; r < dim; r++)
            for (int q = 0; q < dim; q++) g[1 + r * dim + q] = adj[r][q];
        for (int k = 0; k < L.nn; k++)
            for (int i = 0; i < g1; i++) L.ngeom[((size_t)e * L.nn + k) * g1 + i] = g[i];
    }
    return L;
}

// a smooth state with a seeded perturbation, byNODES layout [var][element][node]
static std::vector<double> make_state(const Lattice &L, unsigned seed, double gamma) {
    const int neq = L.dim + 2;
    const size_t vs = (size_t)L.NE * L.nn;
    std::vector<double> u(neq * vs);
    unsigned x = seed * 2654435761u + 1u;
    auto rnd = [&]() {
        x = x * 1664525u + 1013904223u;
        return (double)(x >> 8) / 16777216.0 - 0.5;
    };
    for (size_t i = 0; i < vs; i++) {
        const double rho = 1.0 + 0.2 * std::sin(0.37 * (double)i) + 0.02 * rnd();
        double v[3], ke = 0.0;
        for (int d = 0; d < L.dim; d++) {
            v[d] = 0.3 * std::cos(0.21 * (double)i + d) + 0.02 * rnd();
            ke += 0.5 * rho * v[d] * v[d];
        }
        const double p = 1.0 + 0.1 * std::cos(0.13 * (double)i) + 0.02 * rnd();
        u[i] = rho;
        for (int d = 0; d < L.dim; d++) u[(1 + d) * vs + i] = rho * v[d];
        u[(neq - 1) * vs + i] = p / (gamma - 1.0) + ke;
    }
    return u;
}

static orc_phys make_phys(int viscous, int surf, int bug) {
    orc_phys p;
    p.gamma = 1.4;
    p.Pr = 0.72;
    p.mu = 0.01;
    p.R_gas = 287.05;
    p.viscous = viscous;
    p.sutherland = 0;
    p.surface_flux = surf;
    p.bug_compat = bug;
    return p;
}

#ifndef SAN_LD
static void run_operator_cases(int dim, int order) {
    const double gamma = 1.4;
    Lattice L = make_lattice(dim, order, dim == 2 ? 4 : 3, false);
    const int neq = dim + 2;
    const size_t size = (size_t)neq * L.NE * L.nn, gsize = (size_t)dim * neq * L.NE * L.nn;
    std::vector<double> u = make_state(L, 7u + dim * 10 + order, gamma);
    int cs = 0;
    for (int viscous = 0; viscous < 2; viscous++)
        for (int surf = 0; surf < 3; surf++)
            for (int bug : {0, 4, viscous && dim == 3 ? 7 : 4}) {
                cs++;
                orc_phys p = make_phys(viscous, surf, bug);
                std::vector<double> du(size), grad(gsize), mag(size), gmag(gsize), alpha(L.NE, 0.25);
                double lam = 0.0;
                check(orc_rhs(dim, order, L.NE, L.nbr.data(), L.geom.data(), &p, u.data(), du.data(), grad.data(), &lam) == 0 &&
                          finite(du) && finite(grad) && lam > 0.0,
                      "orc_rhs", dim, order, cs);
                check(orc_rhs_ex(dim, order, L.NE, L.nbr.data(), L.geom.data(), &p, 0, nullptr, u.data(), nullptr,
                                 du.data(), grad.data(), &lam, mag.data(), gmag.data()) == 0 &&
                          finite(du) && finite(mag),
                      "orc_rhs_ex", dim, order, cs);
                check(orc_rhs_gen(dim, order, L.NE, L.nbr.data(), nullptr, L.ngeom.data(), nullptr, &p, 0, nullptr,
                                  u.data(), nullptr, du.data(), grad.data(), &lam, mag.data(), gmag.data()) == 0 &&
                          finite(du),
                      "orc_rhs_gen", dim, order, cs);
                if (!(bug & 3)) {
                    check(orc_rhs_blend_mag(dim, order, L.NE, L.nbr.data(), L.geom.data(), &p, u.data(), alpha.data(),
                                            du.data(), grad.data(), &lam, mag.data(), gmag.data()) == 0 &&
                              finite(du),
                          "orc_rhs_blend_mag", dim, order, cs);
                }
            }
    // steppers: the five schemes, Euler and NS
    for (int viscous = 0; viscous < 2; viscous++) {
        orc_phys p = make_phys(viscous, 0, 0);
        for (int scheme = 0; scheme < 2; scheme++) {
            std::vector<double> x = u;
            check(orc_rk_run(dim, order, L.NE, L.nbr.data(), L.geom.data(), &p, scheme, 2, 1e-4, x.data()) == 0 && finite(x),
                  "orc_rk_run", dim, order, scheme);
        }
        for (int scheme = 0; scheme < 5; scheme++) {
            std::vector<double> x = u;
            int counts[2] = {0, 0};
            check(orc_rk_run_gen(dim, order, L.NE, L.nbr.data(), L.geom.data(), nullptr, nullptr, &p, 0, nullptr, scheme,
                                 2, 1e-4, 0, 0, 0.0, 0.0, x.data(), counts) == 0 &&
                      finite(x),
                  "orc_rk_run_gen", dim, order, 10 + scheme);
        }
    }
    // boundary faces in x: every condition type on the lower end, an outflow on the upper one
    Lattice B = make_lattice(dim, order, dim == 2 ? 4 : 3, true);
    std::vector<double> ub = make_state(B, 99u, gamma);
    for (int type = 1; type <= 11; type++) {
        const int viscous = type >= 10 ? 1 : 0;
        if (type == 2 && viscous) continue;
        orc_phys p = make_phys(viscous, 0, 0);
        orc_bc bc[2];
        for (int k = 0; k < 2; k++)
            for (int i = 0; i < 8; i++) bc[k].q[i] = 0.0;
        bc[0].type = type;
        bc[1].type = 4;
        if (type == 3 || type == 5) {
            bc[0].q[0] = 1.0;
            bc[0].q[1] = 2.0;
            bc[0].q[dim + 1] = 1.0 / (gamma - 1.0) + 2.0;
        } else if (type == 6) {
            bc[0].q[0] = 1.0;
            bc[0].q[1] = 0.3;
            bc[0].q[dim + 1] = 1.0;
        } else if (type == 7) {
            bc[0].q[0] = 1.0;
            bc[0].q[1] = 0.2;
        } else if (type == 8) {
            bc[0].q[0] = 1.1;
            bc[0].q[1] = 0.0035;
            bc[0].q[4] = 1004.5;
        } else if (type == 9) {
            bc[0].q[0] = 0.9;
        } else if (type == 10 || type == 11) {
            bc[0].q[0] = 0.1;
            bc[0].q[dim] = type == 11 ? 0.0035 : 0.0;
        }
        const size_t size = (size_t)(dim + 2) * B.NE * B.nn, gsize = (size_t)dim * size;
        std::vector<double> du(size), grad(gsize), mag(size), gmag(gsize);
        double lam = 0.0;
        check(orc_rhs_ex(dim, order, B.NE, B.nbr.data(), B.geom.data(), &p, 2, bc, ub.data(), nullptr, du.data(),
                         grad.data(), &lam, mag.data(), gmag.data()) == 0 &&
                  finite(du),
              "orc_rhs_ex boundary", dim, order, type);
        std::vector<double> x = ub;
        int counts[2] = {0, 0};
        check(orc_rk_run_gen(dim, order, B.NE, B.nbr.data(), B.geom.data(), nullptr, nullptr, &p, 2, bc, 0, 1, 1e-4, 0,
                             1, 0.5, 0.001, x.data(), counts) == 0 &&
                  finite(x),
              "orc_rk_run_gen boundary + blending", dim, order, type);
    }
    // hooks: inadmissible nodes planted in a few elements
    {
        std::vector<double> x = u;
        const size_t vs = (size_t)L.NE * L.nn;
        for (int e = 0; e < L.NE; e += 3) x[(size_t)e * L.nn + (e % L.nn)] = -0.05;
        std::vector<double> bound(L.NE), zeta(L.NE), alpha(L.NE), eta(L.NE);
        check(orc_filter_constraints(dim, order, L.NE, L.nbr.data(), u.data(), gamma, bound.data()) == 0 && finite(bound),
              "orc_filter_constraints", dim, order, 0);
        std::vector<double> y = x;
        check(orc_filter_apply(dim, order, L.NE, y.data(), gamma, bound.data(), zeta.data()) >= 0 && finite(y),
              "orc_filter_apply", dim, order, 0);
        y = x;
        check(orc_limiter_apply(dim, order, L.NE, y.data(), gamma, 1e-8) >= 0 && finite(y), "orc_limiter_apply", dim, order, 0);
        y = x;
        check(orc_limiter_apply_gen(dim, order, L.NE, y.data(), gamma, 1e-8, L.ngeom.data()) >= 0 && finite(y),
              "orc_limiter_apply_gen", dim, order, 0);
        check(orc_blend_coef(dim, order, L.NE, u.data(), gamma, 0.5, 0.001, alpha.data(), eta.data()) == 0 &&
                  finite(alpha) && finite(eta),
              "orc_blend_coef", dim, order, 0);
        orc_phys p = make_phys(0, 0, 0);
        for (int scheme = 0; scheme < 2; scheme++) {
            std::vector<double> z = x;
            for (size_t i = 0; i < vs; i++) z[i] = std::fabs(z[i]) + 1e-3;   // positive density, then a low pressure
            for (int e = 0; e < L.NE; e += 4) z[(size_t)(dim + 1) * vs + (size_t)e * L.nn] = 1e-9;
            int counts[2] = {0, 0};
            check(orc_rk_run_gen(dim, order, L.NE, L.nbr.data(), L.geom.data(), nullptr, nullptr, &p, 0, nullptr, scheme,
                                 2, 1e-5, 3, 1, 0.5, 0.001, z.data(), counts) == 0,
                  "orc_rk_run_gen hooks + blending", dim, order, scheme);
        }
    }
}
#endif

int main() {
    for (int np = 2; np <= 5; np++) {
        std::vector<double> x(np), w(np), D(np * np), Dh(np * np), D2(np * np);
#ifndef SAN_LD
        check(orc_gll(np, x.data(), w.data()) == 0 && finite(x) && finite(w), "orc_gll", 0, np - 1, 0);
        check(orc_dmatrices(np, D.data(), Dh.data(), D2.data()) == 0 && finite(D) && finite(D2), "orc_dmatrices", 0,
              np - 1, 0);
#endif
    }
    for (int dim = 2; dim <= 3; dim++)
        for (int order = 1; order <= 4; order++) {
#ifdef SAN_LD
            Lattice L = make_lattice(dim, order, dim == 2 ? 4 : 3, false);
            std::vector<double> u = make_state(L, 5u + order, 1.4);
            for (int viscous = 0; viscous < 2; viscous++) {
                orc_phys p = make_phys(viscous, 0, 0);
                std::vector<double> hi(u.size()), lo(u.size());
                double lam = 0.0;
                check(orc_ld_rhs(dim, order, L.NE, L.nbr.data(), L.geom.data(), &p, u.data(), hi.data(), lo.data(), &lam) == 0 &&
                          finite(hi) && finite(lo),
                      "orc_ld_rhs", dim, order, viscous);
            }
#else
            run_operator_cases(dim, order);
            std::vector<double> V(125 * 125), Vi(125 * 125);
            std::vector<int> md(125);
            check(orc_modal_basis(dim, order, V.data(), Vi.data(), md.data()) == 0, "orc_modal_basis", dim, order, 0);
#endif
        }
    std::printf("%s: %d failed checks\n", g_fail ? "FAILED" : "ok", g_fail);
    return g_fail ? 1 : 0;
}
