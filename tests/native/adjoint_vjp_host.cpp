// Host build of the flux vector-Jacobian products (prandtl_amd/csrc/dg_adjoint.hpp) for the
// finite-difference check in tests/test_adjoint_flux.py.  Built with g++ by that test.
#include "../../prandtl_amd/csrc/dg_adjoint.hpp"

extern "C" {
void vjp_volume(int dim, double gamma, const double *s1, const double *s2, const double *m, const double *Fb,
                double *s1b, double *s2b) {
    for (int i = 0; i < dim + 2; i++) s1b[i] = s2b[i] = 0.0;
    if (dim == 2)
        prandtl::adj::volume_flux_vjp<2>(s1, s2, m, gamma - 1.0, Fb, s1b, s2b);
    else
        prandtl::adj::volume_flux_vjp<3>(s1, s2, m, gamma - 1.0, Fb, s1b, s2b);
}
void vjp_face(int dim, int kind, int bug3, double gamma, const double *s1, const double *s2, const double *n,
              const double *fb, double *s1b, double *s2b) {
    for (int i = 0; i < dim + 2; i++) s1b[i] = s2b[i] = 0.0;
    if (dim == 2)
        prandtl::adj::face_flux_vjp<2>(kind, s1, s2, n, gamma, bug3 != 0, fb, s1b, s2b);
    else
        prandtl::adj::face_flux_vjp<3>(kind, s1, s2, n, gamma, bug3 != 0, fb, s1b, s2b);
}
}

extern "C" {
// the logarithmic mean's value, reciprocal and partials: out = (L, 1/L, dL/dx, dL/dy)
void log_mean_partials(double x, double y, double *out) {
    prandtl::adj::log_mean_dq(x, y, 1.0 / x, 1.0 / y, out[0], out[1], out[2], out[3]);
}
void vjp_entropy_vars(int dim, double gamma, const double *s, const double *w_b, double *s_b) {
    for (int i = 0; i < dim + 2; i++) s_b[i] = 0.0;
    if (dim == 2)
        prandtl::adj::entropy_vars_vjp<2>(s, gamma, w_b, s_b);
    else
        prandtl::adj::entropy_vars_vjp<3>(s, gamma, w_b, s_b);
}
void vjp_entgrad(int dim, double gamma, const double *s, const double *g, const double *gp_b, double *s_b,
                 double *g_b) {
    for (int i = 0; i < dim + 2; i++) s_b[i] = g_b[i] = 0.0;
    if (dim == 2)
        prandtl::adj::entgrad_vjp<2>(s, g, gamma, gp_b, s_b, g_b);
    else
        prandtl::adj::entgrad_vjp<3>(s, g, gamma, gp_b, s_b, g_b);
}
// g and g_b: [c][v] with c < dim, v < dim + 2; flux_b: [r][c], r <= dim
void vjp_viscous_flux(int dim, double gamma, double mu, double Pr, const double *s, const double *g,
                      const double *flux_b, double *s_b, double *g_b) {
    for (int i = 0; i < dim + 2; i++) s_b[i] = 0.0;
    for (int i = 0; i < dim * (dim + 2); i++) g_b[i] = 0.0;
    if (dim == 2)
        prandtl::adj::viscous_flux_vjp<2>(s, reinterpret_cast<const double(*)[4]>(g), gamma, mu, 1.0 / Pr, 2.0 / 3.0,
                                          reinterpret_cast<const double(*)[2]>(flux_b), s_b,
                                          reinterpret_cast<double(*)[4]>(g_b));
    else
        prandtl::adj::viscous_flux_vjp<3>(s, reinterpret_cast<const double(*)[5]>(g), gamma, mu, 1.0 / Pr, 2.0 / 3.0,
                                          reinterpret_cast<const double(*)[3]>(flux_b), s_b,
                                          reinterpret_cast<double(*)[5]>(g_b));
}
}
