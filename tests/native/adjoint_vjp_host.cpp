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
