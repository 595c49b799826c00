// Host assembly of the Euler operator's vector-Jacobian product out = (dL/du)^T lam from the
// flux VJPs of prandtl_amd/csrc/dg_adjoint.hpp, in the order of the GPU kernel (per element,
// direction and GLL line: each symmetric pair once, then both line ends' face points).  Built
// with g++ by tests/test_adjoint_flux.py and checked there against central differences of the
// oracle's RHS.  Layout: u[v][e][node], lexicographic nodes, nbr[e][2*dir+side].
#include <vector>

#include "../../prandtl_amd/csrc/dg_adjoint.hpp"

namespace {

template <int DIM>
void op_vjp_t(int np, int NE, const int *nbr, const double *geom, const double *Dh2, double inv_w0, double gamma,
              int kind, int bug3, const double *u, const double *lam, double *out) {
    using namespace prandtl::adj;
    constexpr int NEQ = DIM + 2;
    int nn = 1, L = 1;
    for (int d = 0; d < DIM; d++) nn *= np;
    L = nn / np;
    const size_t vs = (size_t)NE * nn;
    const double gm1 = gamma - 1.0;
    for (size_t i = 0; i < vs * NEQ; i++) out[i] = 0.0;
    auto state = [&](int e, int node, double *s, const double *src) {
        for (int v = 0; v < NEQ; v++) s[v] = src[v * vs + (size_t)e * nn + node];
    };
    for (int e = 0; e < NE; e++) {
        const double *g = geom + (size_t)e * (1 + DIM * DIM);
        const double invJ = 1.0 / g[0];
        for (int dir = 0; dir < DIM; dir++) {
            const double *m = g + 1 + dir * DIM;
            const int st = dir == 0 ? 1 : (dir == 1 ? np : np * np);
            for (int l = 0; l < L; l++) {
                const int base = (l % st) + (l / st) * st * np;
                for (int a = 0; a < np; a++)
                    for (int b = a + 1; b < np; b++) {
                        double sa[NEQ], sb[NEQ], la[NEQ], lb[NEQ], Fb[NEQ], ab[NEQ] = {}, bb[NEQ] = {};
                        state(e, base + a * st, sa, u);
                        state(e, base + b * st, sb, u);
                        state(e, base + a * st, la, lam);
                        state(e, base + b * st, lb, lam);
                        for (int v = 0; v < NEQ; v++)
                            Fb[v] = -invJ * (Dh2[a * np + b] * la[v] + Dh2[b * np + a] * lb[v]);
                        volume_flux_vjp<DIM>(sa, sb, m, gm1, Fb, ab, bb);
                        for (int v = 0; v < NEQ; v++) {
                            out[v * vs + (size_t)e * nn + base + a * st] += ab[v];
                            out[v * vs + (size_t)e * nn + base + b * st] += bb[v];
                        }
                    }
                for (int side = 1; side >= 0; side--) {
                    const int en = nbr[(size_t)e * 2 * DIM + 2 * dir + side];
                    const int nown = base + (side ? (np - 1) * st : 0);
                    const int nnb = base + (side ? 0 : (np - 1) * st);
                    const double swo = inv_w0 * invJ, swn = inv_w0 / geom[(size_t)en * (1 + DIM * DIM)];
                    double so[NEQ], sn[NEQ], lo[NEQ], ln[NEQ], fb[NEQ], b1[NEQ] = {}, b2[NEQ] = {};
                    state(e, nown, so, u);
                    state(en, nnb, sn, u);
                    state(e, nown, lo, lam);
                    state(en, nnb, ln, lam);
                    for (int v = 0; v < NEQ; v++) fb[v] = side ? (-swo * lo[v] + swn * ln[v]) : (-swn * ln[v] + swo * lo[v]);
                    if (side)
                        face_flux_vjp<DIM>(kind, so, sn, m, gamma, bug3 != 0, fb, b1, b2);
                    else
                        face_flux_vjp<DIM>(kind, sn, so, m, gamma, bug3 != 0, fb, b1, b2);
                    for (int v = 0; v < NEQ; v++) out[v * vs + (size_t)e * nn + nown] += side ? b1[v] : b2[v];
                }
            }
        }
    }
}

}  // namespace

extern "C" void op_vjp(int dim, int np, int NE, const int *nbr, const double *geom, const double *Dh2, double inv_w0,
                       double gamma, int kind, int bug3, const double *u, const double *lam, double *out) {
    if (dim == 2)
        op_vjp_t<2>(np, NE, nbr, geom, Dh2, inv_w0, gamma, kind, bug3, u, lam, out);
    else
        op_vjp_t<3>(np, NE, nbr, geom, Dh2, inv_w0, gamma, kind, bug3, u, lam, out);
}
