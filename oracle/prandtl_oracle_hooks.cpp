// prandtl_oracle_hooks.cpp -- CPU oracle for the stage hooks: the entropy filter and the
// positivity limiter.
//
// TEST INFRASTRUCTURE ONLY (see prandtl_oracle.cpp): the checker the HIP filter and limiter
// kernels are compared against; the product library never links or calls it.
//
// Entropy filter: a restatement of the reference's EntropyFilter (src/Filters/
// EntropyFilter.cpp) and ModalBasis (src/ModalBasis/ModalBasis.cpp), with the three defects
// of the reference replaced by the evidently intended behaviour (the Dzanic-Witherden filter):
//   * ComputeElementMinima (EntropyFilter.cpp:275-277) is empty, so every interior element's
//     entropy bound stays DBL_MAX; here the bound is the minimum nodal entropy over the
//     element and its face neighbours (the loop at :113-157 visits exactly those elements);
//   * ComputeRoot compares with (*entropy)(0), the first element's bound (:338), where the
//     detection at :191 and :208 uses the element's own bound ent_dof(0); here both use the
//     element's bound;
//   * ComputeEntropy returns 0 under PARABOLIC (BasicOperations.cpp:39-46); here the entropy
//     rho ln(p rho^-gamma) is used for Euler and Navier-Stokes states alike.
// MFEM's Poly_1D::CalcLegendre (unvendored) is restated from its closed form: the Legendre
// polynomials on [0,1], P_n(2x-1), by the three-term recurrence, unnormalised.
//
// Positivity limiter: the reference declares only the Limiter interface
// (src/Limiters/Limiter.hpp:10-16; its MinmodLimiter is an unfinished no-op).  The limiter
// here is the published scaling limiter of X. Zhang and C.-W. Shu, "On positivity-preserving
// high order discontinuous Galerkin schemes for compressible Euler equations on rectangular
// meshes", J. Comput. Phys. 229 (2010) 8918-8934, Section 3: per element, with the GLL-mean
// state ubar, rho is scaled toward rhobar by theta_1 = min(1, (rhobar - eps)/(rhobar -
// rho_min)), then the whole state toward ubar by theta_2 = min_i t_i, with t_i the root in
// [0, 1] of p(ubar + t (u_i - ubar)) = eps (found by bisection in FP64 to the bracket width
// 2^-52; the lower, admissible end is taken).  eps = 1e-8, the filter's density and pressure
// bounds (EntropyFilter.hpp:59-60).

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "prandtl_oracle.h"

namespace {

// Constants -- ref: EntropyFilter.hpp:59-64
constexpr double kRhoMin = 1e-8;
constexpr double kPMin = 1e-8;
constexpr double kEntTol = 1e-4;
constexpr double kEpsilon = 1e-10;
constexpr int kNIter = 20;

// ComputePressure -- ref: BasicOperations.cpp:31-37
double Pressure(const double *s, int neq, double gammaM1) {
    double v2 = 0.0;
    for (int d = 1; d < neq - 1; d++) v2 += (s[d] / s[0]) * (s[d] / s[0]);
    return gammaM1 * (s[neq - 1] - 0.5 * s[0] * v2);
}

// ComputeEntropy -- ref: BasicOperations.cpp:39-46 (the non-PARABOLIC branch)
double Entropy(double rho, double p, double gamma) {
    return rho > 0.0 && p > 0.0 ? rho * std::log(p * std::pow(rho, -gamma)) : std::numeric_limits<double>::infinity();
}

// Poly_1D::CalcLegendre (MFEM): P_n(2x - 1), n = 0..p
void Legendre01(int p, double x, double *L) {
    const double z = 2.0 * x - 1.0;
    L[0] = 1.0;
    if (p >= 1) L[1] = z;
    for (int n = 1; n < p; n++) L[n + 1] = ((2.0 * n + 1.0) * z * L[n] - n * L[n - 1]) / (n + 1.0);
}

// ModalBasis -- ref: ModalBasis.cpp:5-22 (constructor), :42-108 (ComputeUBDegs: quad i outer,
// j inner; hex i, j, k), :110-135 (ComputeVDM: V(i,k) = prod_d L_{deg_d(k)}(x_d(i)), then
// V_inv = V^-1), :146-150 (ComputeModes: modes = V_inv nodes).  The element's nodal points
// are the tensor GLL nodes in lexicographic order (x fastest), as DGSEMIntegrator reads them.
struct ModalBasis {
    int dim, np, n;
    std::vector<int> maxdeg;      // max over dimensions of mode k's degrees (EntropyFilter.cpp:59-73)
    std::vector<double> V, Vinv;  // n x n, row-major
};

bool Invert(std::vector<double> A, int n, std::vector<double> &inv) {
    // Gauss-Jordan with partial pivoting (DenseMatrix::Invert's result, to rounding)
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        if (A[(size_t)piv * n + c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < n; k++) {
                std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
                std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
            }
        const double d = 1.0 / A[(size_t)c * n + c];
        for (int k = 0; k < n; k++) {
            A[(size_t)c * n + k] *= d;
            inv[(size_t)c * n + k] *= d;
        }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; k++) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
            }
        }
    }
    return true;
}

ModalBasis MakeModalBasis(int dim, int order) {
    ModalBasis B;
    B.dim = dim;
    B.np = order + 1;
    B.n = dim == 2 ? B.np * B.np : B.np * B.np * B.np;
    std::vector<double> xg(B.np), wg(B.np);
    orc_gll(B.np, xg.data(), wg.data());
    // ubdegs (ModalBasis.cpp:71-104)
    std::vector<int> deg((size_t)B.n * dim);
    int m = 0;
    if (dim == 2) {
        for (int i = 0; i < B.np; i++)
            for (int j = 0; j < B.np; j++, m++) deg[(size_t)m * dim] = i, deg[(size_t)m * dim + 1] = j;
    } else {
        for (int i = 0; i < B.np; i++)
            for (int j = 0; j < B.np; j++)
                for (int k = 0; k < B.np; k++, m++)
                    deg[(size_t)m * dim] = i, deg[(size_t)m * dim + 1] = j, deg[(size_t)m * dim + 2] = k;
    }
    B.maxdeg.resize(B.n);
    for (int k = 0; k < B.n; k++) {
        int md = 0;
        for (int d = 0; d < dim; d++) md = std::max(md, deg[(size_t)k * dim + d]);
        B.maxdeg[k] = md;
    }
    // V(i, k) (ModalBasis.cpp:114-130); node i = i0 + np (i1 + np i2)
    B.V.assign((size_t)B.n * B.n, 1.0);
    std::vector<double> L(B.np);
    for (int i = 0; i < B.n; i++) {
        int idx[3] = {i % B.np, (i / B.np) % B.np, i / (B.np * B.np)};
        for (int d = 0; d < dim; d++) {
            Legendre01(order, xg[idx[d]], L.data());
            for (int k = 0; k < B.n; k++) B.V[(size_t)i * B.n + k] *= L[deg[(size_t)k * dim + d]];
        }
    }
    Invert(B.V, B.n, B.Vinv);
    return B;
}

// GetEntropyConstraints -- ref: EntropyFilter.cpp:81-158, with the element minima of the
// element and its face neighbours (the intended ComputeElementMinima, :275-277).
void FilterConstraints(int dim, int order, int NE, const int *nbr, const double *u, double gamma, double *bound) {
    const int np = order + 1, n = dim == 2 ? np * np : np * np * np, neq = dim + 2;
    const size_t vs = (size_t)NE * n;
    std::vector<double> emin(NE);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < NE; e++) {
        double m = std::numeric_limits<double>::max();   // el_min_e = __DBL_MAX__ (:120)
        for (int i = 0; i < n; i++) {
            double s[5];
            for (int v = 0; v < neq; v++) s[v] = u[v * vs + (size_t)e * n + i];
            m = std::min(m, Entropy(s[0], Pressure(s, neq, gamma - 1.0), gamma));   // ComputeMinima (:262-273)
        }
        emin[e] = m;
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < NE; e++) {
        double m = emin[e];
        for (int f = 0; f < 2 * dim; f++) {
            const int nb = nbr[(size_t)e * 2 * dim + f];
            if (nb >= 0) m = std::min(m, emin[nb]);   // boundary faces have no neighbour
        }
        bound[e] = m;   // entropy->SetSubVector(ent_indices, el_min_e) (:153)
    }
}

// FilterModes -- ref: EntropyFilter.cpp:160-169: state = H_0 + sum_{j>=1} exp(-zeta j^2) H_j
void FilterModes(const std::vector<double> &H, int np, int neq, double zeta, double *state) {
    for (int k = 0; k < neq; k++) state[k] = H[k];
    for (int j = 1; j < np; j++) {
        const double f = std::exp(-zeta * std::pow(j, 2));
        for (int k = 0; k < neq; k++) state[k] += H[(size_t)j * neq + k] * f;
    }
}

bool Violates(const double *s, int neq, double gamma, double ebound) {
    const double rho = s[0], p = Pressure(s, neq, gamma - 1.0), e = Entropy(rho, p, gamma);
    return rho < kRhoMin || p < kPMin || e < ebound - kEntTol;
}

// FilterSolution -- ref: EntropyFilter.cpp:171-242 (with ComputeHierarchStates :244-260 and
// ComputeRoot :325-349, the latter against the element's own bound).  Returns the number of
// elements filtered; zeta_out[e] receives the element's filter strength (0 if untouched).
int FilterApply(int dim, int order, int NE, double *u, double gamma, const double *bound, double *zeta_out) {
    const ModalBasis B = MakeModalBasis(dim, order);
    const int np = B.np, n = B.n, neq = dim + 2;
    const size_t vs = (size_t)NE * n;
    int count = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : count)
    for (int e = 0; e < NE; e++) {
        std::vector<double> el((size_t)n * neq), modes((size_t)n * neq), H((size_t)np * neq);
        for (int i = 0; i < n; i++)
            for (int v = 0; v < neq; v++) el[(size_t)i * neq + v] = u[v * vs + (size_t)e * n + i];
        const double eb = bound[e];
        // element minima (:183-189) and the detection (:191)
        double mr = std::numeric_limits<double>::infinity(), mp = mr, me = mr;
        for (int i = 0; i < n; i++) {
            const double *s = &el[(size_t)i * neq];
            const double p = Pressure(s, neq, gamma - 1.0);
            mr = std::min(mr, s[0]);
            mp = std::min(mp, p);
            me = std::min(me, Entropy(s[0], p, gamma));
        }
        if (zeta_out) zeta_out[e] = 0.0;
        if (!(mr < kRhoMin || mp < kPMin || me < eb - kEntTol)) continue;
        count++;
        // modes = V^-1 u per variable (:193-198)
        for (int l = 0; l < n; l++)
            for (int k = 0; k < neq; k++) {
                double t = 0.0;
                for (int i = 0; i < n; i++) t += B.Vinv[(size_t)l * n + i] * el[(size_t)i * neq + k];
                modes[(size_t)l * neq + k] = t;
            }
        double zeta = 0.0;   // :200
        double st[5];
        for (int i = 0; i < n; i++) {
            // ComputeHierarchStates (:244-260)
            std::fill(H.begin(), H.end(), 0.0);
            for (int j = 0; j < np; j++)
                for (int k = 0; k < neq; k++)
                    for (int l = 0; l < n; l++)
                        if (B.maxdeg[l] == j) H[(size_t)j * neq + k] += B.V[(size_t)i * n + l] * modes[(size_t)l * neq + k];
            FilterModes(H, np, neq, zeta, st);   // :204
            if (Violates(st, neq, gamma, eb)) {
                // ComputeRoot (:325-349)
                double z1 = zeta, z2 = -std::log(kEpsilon);
                for (int it = 0; it < kNIter; it++) {
                    const double z3 = 0.5 * (z1 + z2);
                    FilterModes(H, np, neq, z3, st);
                    if (Violates(st, neq, gamma, eb))
                        z1 = z3;
                    else
                        z2 = z3;
                }
                zeta = z2;
            }
        }
        // reconstruction with ffac(j) = exp(-zeta j^2), ffac(0) = 1 (:214-237)
        std::vector<double> ffac(np, 1.0);
        for (int d = 1; d < np; d++) ffac[d] = std::exp(-zeta * std::pow(d, 2));
        for (int i = 0; i < n; i++)
            for (int k = 0; k < neq; k++) {
                double t = 0.0;
                for (int j = 0; j < np; j++)
                    for (int l = 0; l < n; l++)
                        if (B.maxdeg[l] == j) t += ffac[j] * B.V[(size_t)i * n + l] * modes[(size_t)l * neq + k];
                u[k * vs + (size_t)e * n + i] = t;
            }
        if (zeta_out) zeta_out[e] = zeta;
    }
    return count;
}

// Zhang-Shu positivity limiter (see the header).  Returns the number of elements changed.
// ngeom (per-node det J, [e][node][1 + dim^2]) non-null: the element mean is the conserved one,
// sum J w u / sum J w (on affine elements J cancels).
int LimiterApply(int dim, int order, int NE, double *u, double gamma, double eps, const double *ngeom = nullptr) {
    const int np = order + 1, n = dim == 2 ? np * np : np * np * np, neq = dim + 2;
    const size_t vs = (size_t)NE * n;
    std::vector<double> xg(np), wg(np);
    orc_gll(np, xg.data(), wg.data());
    int count = 0;
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int e = 0; e < NE; e++) {
        double s[125][5], mean[5] = {0, 0, 0, 0, 0}, wsum = 0.0;
        for (int i = 0; i < n; i++) {
            double w = wg[i % np] * wg[(i / np) % np] * (dim == 3 ? wg[i / (np * np)] : 1.0);
            if (ngeom) w *= ngeom[((size_t)e * n + i) * (1 + dim * dim)];
            wsum += w;
            for (int v = 0; v < neq; v++) {
                s[i][v] = u[v * vs + (size_t)e * n + i];
                mean[v] += w * s[i][v];
            }
        }
        if (ngeom)
            for (int v = 0; v < neq; v++) mean[v] /= wsum;
        bool changed = false;
        // density: theta_1
        double rmin = std::numeric_limits<double>::infinity();
        for (int i = 0; i < n; i++) rmin = std::min(rmin, s[i][0]);
        if (rmin < eps) {
            // an inadmissible mean (rhobar <= eps) leaves only the mean itself
            const double th = std::max(0.0, std::min(1.0, (mean[0] - eps) / (mean[0] - rmin)));
            for (int i = 0; i < n; i++) s[i][0] = mean[0] + th * (s[i][0] - mean[0]);
            changed = true;
        }
        // pressure: theta_2 = min_i t_i
        double th2 = 1.0;
        for (int i = 0; i < n; i++) {
            if (Pressure(s[i], neq, gamma - 1.0) >= eps) continue;
            double lo = 0.0, hi = 1.0, st[5];
            while (hi - lo > 0x1p-52) {
                const double mid = 0.5 * (lo + hi);
                for (int v = 0; v < neq; v++) st[v] = mean[v] + mid * (s[i][v] - mean[v]);
                if (Pressure(st, neq, gamma - 1.0) >= eps)
                    lo = mid;
                else
                    hi = mid;
            }
            th2 = std::min(th2, lo);
        }
        if (th2 < 1.0) {
            for (int i = 0; i < n; i++)
                for (int v = 0; v < neq; v++) s[i][v] = mean[v] + th2 * (s[i][v] - mean[v]);
            changed = true;
        }
        if (changed) {
            count++;
            for (int i = 0; i < n; i++)
                for (int v = 0; v < neq; v++) u[v * vs + (size_t)e * n + i] = s[i][v];
        }
    }
    return count;
}

// ---------------------------------------------------------------------------
// Blending coefficient of the subcell finite-volume blending -- ref:
// PerssonPeraireIndicator::CheckSmoothness (src/Indicators/PerssonPeraireIndicator.cpp:21-58)
// and DGSEMOperator::ComputeBlendingCoefficient (src/Operators/DGSEMOperator.cpp:60-80), with
// modalThreshold T = 0.5 10^(-1.8 p^(1/4)) (:25) and sharpness_fac 9.21024
// (DGSEMOperator.hpp:50).  Per element: q = rho p at the nodes (ComputePressure,
// BasicOperations.cpp:31-37), its modes V^-1 q (ModalBasis::ComputeModes, :146-150); modesM1
// drops the modes with a degree above p - 1, modesM2 those above p - 2;
// eta = max(1 - |M1|^2/|M|^2, 1 - |M2|^2/|M1|^2);  alpha = 1/(1 + exp(-9.21024 (eta - T)/T)),
// set to 0 below alpha_min, to 1 above 1 - alpha_min, then min(alpha, alpha_max).
// At p = 1 every mode has a degree above p - 2, so |M2| = 0 and eta = 1 (as in the reference).
// ---------------------------------------------------------------------------
void BlendingCoefficient(int dim, int order, int NE, const double *u, double gamma, double alpha_max,
                         double alpha_min, double *alpha, double *eta) {
    const ModalBasis B = MakeModalBasis(dim, order);
    const int n = B.n, neq = dim + 2;
    const size_t vs = (size_t)NE * n;
    const double gammaM1 = gamma - 1.0;
    const double T = 0.5 * std::pow(10.0, -1.8 * std::pow((double)order, 0.25));
    const double sharpness = 9.21024;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < NE; e++) {
        std::vector<double> q(n), modes(n, 0.0);
        for (int i = 0; i < n; i++) {
            double s[5];
            for (int v = 0; v < neq; v++) s[v] = u[v * vs + (size_t)e * n + i];
            const double inv = 1.0 / s[0];                 // V /= state(0)
            double vv = 0.0;
            for (int d = 1; d < neq - 1; d++) vv += (s[d] * inv) * (s[d] * inv);
            q[i] = s[0] * (gammaM1 * (s[neq - 1] - 0.5 * s[0] * vv));
        }
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) modes[i] += B.Vinv[(size_t)i * n + j] * q[j];
        double mm = 0.0, m1 = 0.0, m2 = 0.0;              // modes*modes, modesM1*modesM1, modesM2*modesM2
        for (int i = 0; i < n; i++) {
            const double a = modes[i];
            mm += a * a;
            if (B.maxdeg[i] <= order - 1) m1 += a * a;
            if (B.maxdeg[i] <= order - 2) m2 += a * a;
        }
        double ind = 1.0 - m1 / mm;
        ind = std::max(ind, 1.0 - m2 / m1);
        double a = 1.0 / (1.0 + std::exp(-sharpness * (ind - T) / T));
        if (a < alpha_min)
            a = 0.0;
        else if (a > (1.0 - alpha_min))
            a = 1.0;
        a = std::min(a, alpha_max);
        if (alpha) alpha[e] = a;
        if (eta) eta[e] = ind;
    }
}

}  // namespace

extern "C" {

// Persson-Peraire indicator eta and blending coefficient alpha of every element
int orc_blend_coef(int dim, int order, int NE, const double *u, double gamma, double alpha_max, double alpha_min,
                   double *alpha, double *eta) {
    if ((dim != 2 && dim != 3) || order < 1 || order > 4) return -1;
    BlendingCoefficient(dim, order, NE, u, gamma, alpha_max, alpha_min, alpha, eta);
    return 0;
}

// V (n x n, row-major) and the max degree of every mode (ModalBasis + EntropyFilter.cpp:59-73)
int orc_modal_basis(int dim, int order, double *V, double *Vinv, int *maxdeg) {
    if ((dim != 2 && dim != 3) || order < 1 || order > 4) return -1;
    const ModalBasis B = MakeModalBasis(dim, order);
    std::copy(B.V.begin(), B.V.end(), V);
    std::copy(B.Vinv.begin(), B.Vinv.end(), Vinv);
    std::copy(B.maxdeg.begin(), B.maxdeg.end(), maxdeg);
    return 0;
}

void orc_legendre01(int p, double x, double *L) { Legendre01(p, x, L); }

int orc_filter_constraints(int dim, int order, int NE, const int *nbr, const double *u, double gamma, double *bound) {
    if ((dim != 2 && dim != 3) || order < 1 || order > 4) return -1;
    FilterConstraints(dim, order, NE, nbr, u, gamma, bound);
    return 0;
}

int orc_filter_apply(int dim, int order, int NE, double *u, double gamma, const double *bound, double *zeta) {
    if ((dim != 2 && dim != 3) || order < 1 || order > 4) return -1;
    return FilterApply(dim, order, NE, u, gamma, bound, zeta);
}

int orc_limiter_apply(int dim, int order, int NE, double *u, double gamma, double eps) {
    if ((dim != 2 && dim != 3) || order < 1 || order > 4) return -1;
    return LimiterApply(dim, order, NE, u, gamma, eps);
}

// the limiter on a general mesh (per-node det J in ngeom: the conserved element mean)
int orc_limiter_apply_gen(int dim, int order, int NE, double *u, double gamma, double eps, const double *ngeom) {
    if ((dim != 2 && dim != 3) || order < 1 || order > 4) return -1;
    return LimiterApply(dim, order, NE, u, gamma, eps, ngeom);
}

// orc_rk_run (prandtl_oracle.cpp) with the stage hooks, in the order of the stepper: the
// filter's entropy bounds from the state at the start of each step (GetFilterConstraints),
// then after every stage the filter (FilterSolution) and the limiter (LimitSolution).
// hooks: bit 0 filter, bit 1 limiter.  counts[0], counts[1] accumulate the elements changed.
// blend != 0: every RHS evaluation with subcell finite-volume blending, its coefficients from
// the stage state (DGSEMOperator::Mult, DGSEMOperator.cpp:547-549).
int orc_rk_run_blend(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p,
                     int scheme, int nsteps, double dt, int hooks, int blend, double alpha_max, double alpha_min,
                     double *u, int *counts);

int orc_rk_run_hooked(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p,
                      int scheme, int nsteps, double dt, int hooks, double *u, int *counts) {
    return orc_rk_run_blend(dim, order, NE, nbr, geom, p, scheme, nsteps, dt, hooks, 0, 0.0, 0.0, u, counts);
}

int orc_rk_run_ex(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, int nbc,
                  const orc_bc *bcs, int scheme, int nsteps, double dt, int hooks, int blend, double alpha_max,
                  double alpha_min, double *u, int *counts);

int orc_rk_run_blend(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p,
                     int scheme, int nsteps, double dt, int hooks, int blend, double alpha_max, double alpha_min,
                     double *u, int *counts) {
    return orc_rk_run_ex(dim, order, NE, nbr, geom, p, 0, nullptr, scheme, nsteps, dt, hooks, blend, alpha_max,
                         alpha_min, u, counts);
}

int orc_rk_run_gen(int dim, int order, int NE, const int *nbr, const double *geom, const double *ngeom,
                   const int *forient, const orc_phys *p, int nbc, const orc_bc *bcs, int scheme, int nsteps,
                   double dt, int hooks, int blend, double alpha_max, double alpha_min, double *u, int *counts);

int orc_rk_run_ex(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, int nbc,
                  const orc_bc *bcs, int scheme, int nsteps, double dt, int hooks, int blend, double alpha_max,
                  double alpha_min, double *u, int *counts) {
    return orc_rk_run_gen(dim, order, NE, nbr, geom, nullptr, nullptr, p, nbc, bcs, scheme, nsteps, dt, hooks, blend,
                          alpha_max, alpha_min, u, counts);
}

// The general stepper: boundary conditions, blending, stage hooks, general meshes (ngeom,
// forient as in orc_rhs_gen).
int orc_rk_run_gen(int dim, int order, int NE, const int *nbr, const double *geom, const double *ngeom,
                   const int *forient, const orc_phys *p, int nbc, const orc_bc *bcs, int scheme, int nsteps,
                   double dt, int hooks, int blend, double alpha_max, double alpha_min, double *u, int *counts) {
    if ((dim != 2 && dim != 3) || order < 1 || order > 4) return -1;
    const double gamma = p->gamma;
    const int np = order + 1, n = dim == 2 ? np * np : np * np * np;
    const size_t size = (size_t)(dim + 2) * NE * n;
    std::vector<double> k(size), y(size), kk(size), bound(NE), alpha(NE);
    auto hook = [&](double *x) {
        if (hooks & 1) counts[0] += FilterApply(dim, order, NE, x, gamma, bound.data(), nullptr);
        if (hooks & 2) counts[1] += LimiterApply(dim, order, NE, x, gamma, 1e-8, ngeom);
    };
    // the stage times of MFEM's steppers (ODESolver::Step sets f's time before each Mult), from
    // the time of the boundary context; t += dt after each step
    double t = orc_get_time();
    auto rhs_at = [&](const double *x, double c) {
        orc_set_time(t + c * dt);
        if (blend) BlendingCoefficient(dim, order, NE, x, gamma, alpha_max, alpha_min, alpha.data(), nullptr);
        return orc_rhs_gen(dim, order, NE, nbr, geom, ngeom, forient, p, nbc, bcs, x, blend ? alpha.data() : nullptr,
                           k.data(), nullptr, nullptr, nullptr, nullptr);
    };
    int stage = 0;   // the stage of the current step (its time: the scheme's table below)
    auto rhs = [&](const double *x) {
        // SSP-RK3 t, t + dt, t + dt/2 (RK3SSPSolver); LSRK4(5) the Carpenter-Kennedy c_i; forward
        // Euler t; RK2 t, t + a dt (a = 2/3); RK4 t, t + dt/2, t + dt/2, t + dt
        static const double C0[3] = {0.0, 1.0, 0.5};
        static const double C1[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                                     2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};
        static const double C3[2] = {0.0, 2.0 / 3.0};
        static const double C4[4] = {0.0, 0.5, 0.5, 1.0};
        const double c = scheme == 0 ? C0[stage] : scheme == 1 ? C1[stage] : scheme == 3 ? C3[stage]
                                                                   : scheme == 4 ? C4[stage] : 0.0;
        stage++;
        return rhs_at(x, c);
    };
    for (int s = 0; s < nsteps; s++, t += dt) {
        stage = 0;
        if (hooks & 1) FilterConstraints(dim, order, NE, nbr, u, gamma, bound.data());
        if (scheme == 0) {
            // SSP-RK3, Shu-Osher form (as orc_rk_run)
            if (rhs(u)) return -3;
            for (size_t i = 0; i < size; i++) y[i] = u[i] + dt * k[i];
            hook(y.data());
            if (rhs(y.data())) return -3;
            for (size_t i = 0; i < size; i++) {
                y[i] = y[i] + dt * k[i];
                y[i] = 3. / 4 * u[i] + 1. / 4 * y[i];
            }
            hook(y.data());
            if (rhs(y.data())) return -3;
            for (size_t i = 0; i < size; i++) {
                y[i] = y[i] + dt * k[i];
                u[i] = 1. / 3 * u[i] + 2. / 3 * y[i];
            }
            hook(u);
        } else if (scheme == 1) {
            static const double A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
            static const double B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                        2277821191437.0 / 14882151754819.0};
            std::fill(kk.begin(), kk.end(), 0.0);
            for (int st = 0; st < 5; st++) {
                if (rhs(u)) return -3;
                for (size_t i = 0; i < size; i++) {
                    kk[i] = A[st] * kk[i] + dt * k[i];
                    u[i] = u[i] + B[st] * kk[i];
                }
                hook(u);
            }
        } else if (scheme == 2) {
            // MFEM ForwardEulerSolver::Step: x.Add(dt, dxdt)
            if (rhs(u)) return -3;
            for (size_t i = 0; i < size; i++) u[i] = u[i] + dt * k[i];
            hook(u);
        } else if (scheme == 3) {
            // MFEM RK2Solver::Step with its default a = 2/3, b = 1/(2a): x1 = x + (1-b) dt k;
            // x += a dt k; x = x1 + b dt f(x)
            const double a = 2.0 / 3.0, b = 0.5 / a;
            if (rhs(u)) return -3;
            for (size_t i = 0; i < size; i++) {
                kk[i] = u[i] + ((1.0 - b) * dt) * k[i];
                y[i] = u[i] + (a * dt) * k[i];
            }
            hook(y.data());
            if (rhs(y.data())) return -3;
            for (size_t i = 0; i < size; i++) u[i] = kk[i] + (b * dt) * k[i];
            hook(u);
        } else if (scheme == 4) {
            // MFEM RK4Solver (ExplicitRKSolver::Step with a = {1/2; 0, 1/2; 0, 0, 1},
            // b = {1/6, 1/3, 1/3, 1/6}): y_l = x + a dt k (the zero entries add nothing), then
            // x += b_0 dt k_0, += b_1 dt k_1, ... in order; accumulated here stage by stage
            static const double A4[3] = {0.5, 0.5, 1.0};
            static const double B4[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
            const double *in = u;
            for (int st = 0; st < 4; st++) {
                if (rhs(in)) return -3;
                for (size_t i = 0; i < size; i++) {
                    const double acc = (st == 0 ? u[i] : kk[i]) + (B4[st] * dt) * k[i];
                    if (st == 3)
                        u[i] = acc;
                    else {
                        kk[i] = acc;
                        y[i] = u[i] + (A4[st] * dt) * k[i];
                    }
                }
                hook(st == 3 ? u : y.data());
                in = y.data();
            }
        } else {
            return -2;
        }
    }
    orc_set_time(t);
    return 0;
}

}  // extern "C"
