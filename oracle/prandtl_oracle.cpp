// prandtl_oracle.cpp -- CPU oracle for the DGSEM right-hand side and RK stage.
//
// TEST INFRASTRUCTURE ONLY.  This file is a line-traceable C++17 restatement of
// the reference solver's arithmetic (Prandtl, MFEM-based).  It is the checker
// against which the HIP path is compared, and the CPU baseline that bench.py
// times beside the GPU.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product library never links or calls it.
//
// Every function cites the reference file:line it restates.  Paths are
// relative to the reference repository root.
//
// MFEM-side arithmetic (the GLL rule, CalcOrtho, element Jacobians, the ODE
// solver) is not vendored in the reference, so it is restated here from its
// closed forms:
//   * GLL nodes on [0,1]: 0, 1 and the roots of P'_N(2x-1); weights
//     1/(N(N+1) P_N(2x-1)^2).
//   * affine element x = x0 + A xi on [0,1]^d: Weight() = det A,
//     AdjugateJacobian() = adj A, and CalcOrtho of the face xi_d = 1 is row d
//     of adj A (the face area vector, outward).
//   * RK3SSPSolver in Shu-Osher form; the 5-stage 2N-storage RK4(3)5 of
//     Carpenter & Kennedy (NASA TM-109112, 1994).
//
// Parity pins for every function live in tests/test_oracle_*.py (closed forms,
// invariants, brute force); see the header of each test file.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp, portable x86-64-v3 code).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "prandtl_oracle.h"

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// The arithmetic type: double (the oracle), or long double for the extended-precision build
// that serves as the exact discrete operator (prandtl_oracle_ld.cpp).
using real = ORACLE_REAL;

constexpr int MAXNP = 10;
constexpr int MAXEQ = 5;

// ---------------------------------------------------------------------------
// Physics constants -- ref: src/Physics/Physics.hpp:10-38
// ---------------------------------------------------------------------------
struct Phys {
    int dim, neq;
    real gamma, gammaM1, gammaM1Inverse, Pr, PrInverse, mu, R_gas;
    real mu_bulk = 2.0 / 3.0;   // Physics.hpp:34
    real mu0 = 1.716e-5;        // Physics.hpp:35
    real T0 = 273.15;           // Physics.hpp:36
    real Ts = 110.4;            // Physics.hpp:37
    int viscous, sutherland, surface_flux, bug;
};

// ---------------------------------------------------------------------------
// ComputeLogMean -- ref: src/BasicOperations/BasicOperations.cpp:6-12
// ---------------------------------------------------------------------------
inline real LogMean(real x, real y, real eps = 1e-4) {
    real xi = y / x;
    real u = (xi * (xi - 2.0) + 1.0) / (xi * (xi + 2.0) + 1);
    real ln_mean = (u < eps) ? (x + y) * 52.5 / (105.0 + u * (35.0 + u * (21.0 + 15.0 * u)))
                               : (y - x) / std::log(xi);
    return ln_mean;
}

// ComputeSoundSpeed -- ref: BasicOperations.cpp:53-56 (b = 0)
inline real SoundSpeed(real p, real rho, real gamma) {
    return std::sqrt(gamma * p / (1.0 - 0.0 * rho) / rho);
}

// ---------------------------------------------------------------------------
// ChandrashekarFlux::ComputeVolumeFlux -- ref: src/RiemannSolvers/ChandrashekarFlux.cpp:11-83
// ---------------------------------------------------------------------------
real VolumeFlux(const Phys &P, const real *state1, const real *state2,
                  const real *metric1, const real *metric2, real *F_tilde) {
    const int dim = P.dim, neq = P.neq;
    real metric[3], mom[3];
    for (int d = 0; d < dim; d++) metric[d] = (metric1[d] + metric2[d]) * 0.5;  // ComputeMean :15, BasicOperations.cpp:24-29

    real rho1 = state1[0], rho2 = state2[0];
    real rho_ln = LogMean(rho1, rho2);
    real u1 = state1[1] / rho1, u2 = state2[1] / rho2;
    real u_mean = 0.5 * (u1 + u2);
    real V_sq1 = u1 * u1, V_sq2 = u2 * u2;
    real Vn = u_mean * metric[0];
    mom[0] = rho_ln * u_mean;
    real h_hat = -0.25 * (u1 * u1 + u2 * u2) + u_mean * u_mean;
    if (dim > 1) {
        real v1 = state1[2] / rho1, v2 = state2[2] / rho2;
        real v_mean = 0.5 * (v1 + v2);
        V_sq1 += v1 * v1;
        V_sq2 += v2 * v2;
        Vn += v_mean * metric[1];
        mom[1] = rho_ln * v_mean;
        h_hat += -0.25 * (v1 * v1 + v2 * v2) + v_mean * v_mean;
        if (dim > 2) {
            real w1 = state1[3] / rho1, w2 = state2[3] / rho2;
            real w_mean = 0.5 * (w1 + w2);
            V_sq1 += w1 * w1;
            V_sq2 += w2 * w2;
            Vn += w_mean * metric[2];
            mom[2] = rho_ln * w_mean;
            h_hat += -0.25 * (w1 * w1 + w2 * w2) + w_mean * w_mean;
        }
    }
    real p1 = P.gammaM1 * (state1[neq - 1] - 0.5 * rho1 * V_sq1);   // :58
    real p2 = P.gammaM1 * (state2[neq - 1] - 0.5 * rho2 * V_sq2);   // :59
    real V_mag1 = std::sqrt(V_sq1), V_mag2 = std::sqrt(V_sq2);
    real lambda_max = V_mag1 + SoundSpeed(p1, rho1, P.gamma);
    lambda_max = std::max(lambda_max, V_mag2 + SoundSpeed(p2, rho2, P.gamma));
    real beta1 = 0.5 * rho1 / p1, beta2 = 0.5 * rho2 / p2;           // :67-68
    real beta_ln = LogMean(beta1, beta2);
    real p_hat = 0.5 * (rho1 + rho2) / (beta1 + beta2);              // :71
    h_hat += 0.5 / beta_ln * P.gammaM1Inverse + p_hat / rho_ln;        // :73
    F_tilde[0] = rho_ln * Vn;
    for (int d = 0; d < dim; d++) F_tilde[1 + d] = Vn * mom[d] + p_hat * metric[d];
    F_tilde[1 + dim] = rho_ln * Vn * h_hat;
    return lambda_max;
}

// ---------------------------------------------------------------------------
// ChandrashekarFlux::ComputeFaceFlux -- ref: ChandrashekarFlux.cpp:85-161
// bug & 4 reproduces the 3D dissipation typo at :131 ("rho_mean*dw + w_mean").
// ---------------------------------------------------------------------------
real FaceFluxChandrashekar(const Phys &P, const real *state1, const real *state2,
                             const real *nor, real *flux) {
    const int dim = P.dim, neq = P.neq;
    real mom[3];
    real nor_mag = 0.0;
    for (int d = 0; d < dim; d++) nor_mag += nor[d] * nor[d];
    nor_mag = std::sqrt(nor_mag);                                      // :88 Norml2

    real rho1 = state1[0], rho2 = state2[0];
    real rho_mean = 0.5 * (rho1 + rho2);
    real rho_ln = LogMean(rho1, rho2);
    real drho = rho2 - rho1;
    real u1 = state1[1] / rho1, u2 = state2[1] / rho2;
    real u_mean = 0.5 * (u1 + u2);
    real du = u2 - u1;
    real V_sq1 = u1 * u1, V_sq2 = u2 * u2;
    real Vn = u_mean * nor[0];
    mom[0] = rho_ln * u_mean;
    real h_hat = -0.25 * (u1 * u1 + u2 * u2) + u_mean * u_mean;
    real diss = 0.5 * drho * u1 * u2 + rho_mean * du * u_mean;       // :105
    if (dim > 1) {
        real v1 = state1[2] / rho1, v2 = state2[2] / rho2;
        real v_mean = 0.5 * (v1 + v2);
        real dv = v2 - v1;
        V_sq1 += v1 * v1;
        V_sq2 += v2 * v2;
        Vn += v_mean * nor[1];
        mom[1] = rho_ln * v_mean;
        h_hat += -0.25 * (v1 * v1 + v2 * v2) + v_mean * v_mean;
        diss += 0.5 * drho * v1 * v2 + rho_mean * dv * v_mean;        // :118
        if (dim > 2) {
            real w1 = state1[3] / rho1, w2 = state2[3] / rho2;
            real w_mean = 0.5 * (w1 + w2);
            real dw = w2 - w1;
            V_sq1 += w1 * w1;
            V_sq2 += w2 * w2;
            Vn += w_mean * nor[2];
            mom[2] = rho_ln * w_mean;
            h_hat += -0.25 * (w1 * w1 + w2 * w2) + w_mean * w_mean;
            if (P.bug & 4)
                diss += 0.5 * drho * w1 * w2 + rho_mean * dw + w_mean;     // :131 as written
            else
                diss += 0.5 * drho * w1 * w2 + rho_mean * dw * w_mean;     // :131 intended
        }
    }
    real p1 = P.gammaM1 * (state1[neq - 1] - 0.5 * rho1 * V_sq1);
    real p2 = P.gammaM1 * (state2[neq - 1] - 0.5 * rho2 * V_sq2);
    real V_mag1 = std::sqrt(V_sq1), V_mag2 = std::sqrt(V_sq2);
    real lambda_max = V_mag1 + SoundSpeed(p1, rho1, P.gamma);
    lambda_max = std::max(lambda_max, V_mag2 + SoundSpeed(p2, rho2, P.gamma));
    real beta1 = 0.5 * rho1 / p1, beta2 = 0.5 * rho2 / p2;
    real beta_ln = LogMean(beta1, beta2);
    real p_hat = 0.5 * (rho1 + rho2) / (beta1 + beta2);
    h_hat += 0.5 / beta_ln * P.gammaM1Inverse + p_hat / rho_ln;
    diss += 0.5 * drho * P.gammaM1Inverse / beta_ln +
            0.5 * rho_mean * P.gammaM1Inverse * (1.0 / beta2 - 1.0 / beta1);   // :151
    flux[0] = rho_ln * Vn - 0.5 * lambda_max * (rho2 - rho1) * nor_mag;       // :153
    for (int d = 0; d < dim; d++)
        flux[1 + d] = Vn * mom[d] + p_hat * nor[d] - 0.5 * lambda_max * (state2[1 + d] - state1[1 + d]) * nor_mag;
    flux[1 + dim] = rho_ln * Vn * h_hat - 0.5 * lambda_max * diss * nor_mag;   // :158
    return lambda_max;
}

// Euler physical flux dotted with n (used by the Rusanov plug-in and by the
// consistency tests).  The reference gets it from mfem::EulerFlux
// (NavierStokesFlux : EulerFlux, src/Fluxes/NavierStokesFlux.hpp:10); the formula is
// the textbook one.
real EulerFluxDotN(const Phys &P, const real *s, const real *nor, real *fn) {
    const int dim = P.dim, neq = P.neq;
    real rho = s[0], vel[3], vsq = 0.0, vn = 0.0;
    for (int d = 0; d < dim; d++) {
        vel[d] = s[1 + d] / rho;
        vsq += vel[d] * vel[d];
        vn += vel[d] * nor[d];
    }
    real p = P.gammaM1 * (s[neq - 1] - 0.5 * rho * vsq);
    fn[0] = rho * vn;
    for (int d = 0; d < dim; d++) fn[1 + d] = s[1 + d] * vn + p * nor[d];
    fn[neq - 1] = (s[neq - 1] + p) * vn;
    return std::sqrt(vsq) + SoundSpeed(p, rho, P.gamma);
}

// Surface-flux plug-ins.  kind 0 = the reference's ChandrashekarFlux::ComputeFaceFlux;
// kind 1 = the entropy-conservative flux alone (ComputeVolumeFlux with metric = nor);
// kind 2 = Rusanov / local Lax-Friedrichs on the central flux:
//   f* = 1/2 (F(u1) + F(u2)).n - 1/2 lambda_max |n| (u2 - u1).
real FaceFlux(const Phys &P, const real *s1, const real *s2, const real *nor, real *flux) {
    if (P.surface_flux == 0) return FaceFluxChandrashekar(P, s1, s2, nor, flux);
    if (P.surface_flux == 1) return VolumeFlux(P, s1, s2, nor, nor, flux);
    real f1[MAXEQ], f2[MAXEQ];
    real l1 = EulerFluxDotN(P, s1, nor, f1);
    real l2 = EulerFluxDotN(P, s2, nor, f2);
    real lam = std::max(l1, l2);
    real nm = 0.0;
    for (int d = 0; d < P.dim; d++) nm += nor[d] * nor[d];
    nm = std::sqrt(nm);
    for (int v = 0; v < P.neq; v++) flux[v] = 0.5 * (f1[v] + f2[v]) - 0.5 * lam * nm * (s2[v] - s1[v]);
    return lam;
}

// ---------------------------------------------------------------------------
// Conserv2Entropy -- ref: BasicOperations.cpp:81-104
// ---------------------------------------------------------------------------
void Conserv2Entropy(const Phys &P, const real *state, real *ent) {
    const int neq = P.neq, dim = P.dim;
    real vel[3], v_sq = 0.0;
    for (int d = 0; d < dim; d++) vel[d] = state[1 + d] / state[0];
    for (int d = 0; d < dim; d++) v_sq += vel[d] * vel[d];
    v_sq = 0.5 * v_sq;
    real p = P.gammaM1 * (state[neq - 1] - state[0] * v_sq);
    real beta = state[0] / p;
    real s = std::log(p * std::pow(state[0], -P.gamma));
    ent[0] = (P.gamma - s) * P.gammaM1Inverse - beta * v_sq;
    for (int d = 0; d < dim; d++) ent[1 + d] = beta * vel[d];
    ent[neq - 1] = -beta;
}

// Entropy2Conserv -- ref: BasicOperations.cpp:106-124
void Entropy2Conserv(const Phys &P, const real *ent, real *state) {
    const int dim = P.dim;
    const real beta = -ent[dim + 1];
    real vel[3], vv = 0.0;
    for (int d = 0; d < dim; d++) vel[d] = ent[1 + d] / beta;
    for (int d = 0; d < dim; d++) vv += vel[d] * vel[d];
    const real s = P.gamma - (ent[0] + 0.5 * beta * vv) * P.gammaM1;
    state[0] = std::pow(std::exp(-s) / beta, P.gammaM1Inverse);
    for (int d = 0; d < dim; d++) state[1 + d] = state[0] * vel[d];
    state[dim + 1] = state[0] * (1.0 / beta * P.gammaM1Inverse + 0.5 * vv);
}

// EntropyGrad2PrimGrad (one node, one direction, in place) -- ref: BasicOperations.cpp:126-170
void EntropyGrad2PrimGrad(const Phys &P, const real *state, real *g) {
    const int dim = P.dim, neq = P.neq;
    real vel[3], v_sq = 0.0;
    for (int d = 0; d < dim; d++) vel[d] = state[1 + d] / state[0];
    for (int d = 0; d < dim; d++) v_sq += vel[d] * vel[d];
    v_sq = 0.5 * v_sq;
    real p = P.gammaM1 * (state[neq - 1] - state[0] * v_sq);
    real KE = v_sq * state[0];
    real gs[MAXEQ] = {};
    for (int v = 0; v < neq; v++) gs[v] = g[v];
    const real ge = gs[neq - 1];
    real mdot = 0.0;
    for (int d = 0; d < dim; d++) {
        g[1 + d] = p / state[0] * (gs[1 + d] + vel[d] * ge);
    }
    for (int d = 0; d < dim; d++) mdot += state[1 + d] * g[1 + d];
    g[0] = state[0] * gs[0] - ge * (KE - p * P.gammaM1Inverse) + state[0] / p * mdot;
    g[neq - 1] = p / state[0] * (g[0] + p * ge);
}

// ---------------------------------------------------------------------------
// NavierStokesFlux::ComputeViscousFlux -- ref: src/Fluxes/NavierStokesFlux.cpp:12-58 (3D),
// :60-90 (2D).  Sutherland: NavierStokesFlux.hpp:32-36.  flux is neq x dim,
// row-major (flux[v*dim + d]); row 0 stays zero (DGSEMIntegrator.cpp:106).
// ---------------------------------------------------------------------------
void ViscousFlux(const Phys &P, const real *state, const real *dqdx, const real *dqdy,
                 const real *dqdz, real *flux) {
    const int dim = P.dim, neq = P.neq;
    real prim[MAXEQ];                                         // Conserv2Prim BasicOperations.cpp:172-186
    prim[0] = state[0];
    for (int d = 0; d < dim; d++) prim[1 + d] = state[1 + d] / state[0];
    {
        real vv = 0.0;
        for (int d = 0; d < dim; d++) vv += prim[1 + d] * prim[1 + d];
        prim[neq - 1] = P.gammaM1 * (state[neq - 1] - 0.5 * state[0] * vv);  // ComputePressure :31-37
    }
    real mu = P.mu;
    if (P.sutherland) {
        real T = prim[neq - 1] / (prim[0] * P.R_gas);
        mu = P.mu0 * (P.T0 + P.Ts) / (T + P.Ts) * (T / P.T0) * std::sqrt(T / P.T0);
    }
    const real lambda = mu * P.gamma * P.PrInverse;
    for (int i = 0; i < neq * dim; i++) flux[i] = 0.0;
    auto F = [&](int r, int c) -> real & { return flux[r * dim + c]; };
    if (dim == 3) {
        const real drdx = dqdx[0], dudx = dqdx[1], dvdx = dqdx[2], dwdx = dqdx[3], dpdx = dqdx[4];
        const real drdy = dqdy[0], dudy = dqdy[1], dvdy = dqdy[2], dwdy = dqdy[3], dpdy = dqdy[4];
        const real drdz = dqdz[0], dudz = dqdz[1], dvdz = dqdz[2], dwdz = dqdz[3], dpdz = dqdz[4];
        real div = dudx + dvdy + dwdz;
        real cv_dTdx = P.gammaM1Inverse / prim[0] * (dpdx - prim[4] / prim[0] * drdx);
        real cv_dTdy = P.gammaM1Inverse / prim[0] * (dpdy - prim[4] / prim[0] * drdy);
        real cv_dTdz = P.gammaM1Inverse / prim[0] * (dpdz - prim[4] / prim[0] * drdz);
        F(1, 0) = mu * (2.0 * dudx - P.mu_bulk * div);
        F(2, 0) = mu * (dudy + dvdx);
        F(3, 0) = mu * (dudz + dwdx);
        F(4, 0) = prim[1] * F(1, 0) + prim[2] * F(2, 0) + prim[3] * F(3, 0) + lambda * cv_dTdx;
        F(1, 1) = mu * (dvdx + dudy);
        F(2, 1) = mu * (2.0 * dvdy - P.mu_bulk * div);
        F(3, 1) = mu * (dvdz + dwdy);
        F(4, 1) = prim[1] * F(1, 1) + prim[2] * F(2, 1) + prim[3] * F(3, 1) + lambda * cv_dTdy;
        F(1, 2) = mu * (dwdx + dudz);
        F(2, 2) = mu * (dwdy + dvdz);
        F(3, 2) = mu * (2.0 * dwdz - P.mu_bulk * div);
        F(4, 2) = prim[1] * F(1, 2) + prim[2] * F(2, 2) + prim[3] * F(3, 2) + lambda * cv_dTdz;
    } else {
        const real drdx = dqdx[0], dudx = dqdx[1], dvdx = dqdx[2], dpdx = dqdx[3];
        const real drdy = dqdy[0], dudy = dqdy[1], dvdy = dqdy[2], dpdy = dqdy[3];
        real div = dudx + dvdy;
        real cv_dTdx = P.gammaM1Inverse / prim[0] * (dpdx - prim[3] / prim[0] * drdx);
        real cv_dTdy = P.gammaM1Inverse / prim[0] * (dpdy - prim[3] / prim[0] * drdy);
        F(1, 0) = mu * (2.0 * dudx - P.mu_bulk * div);
        F(2, 0) = mu * (dudy + dvdx);
        F(3, 0) = prim[1] * F(1, 0) + prim[2] * F(2, 0) + lambda * cv_dTdx;
        F(1, 1) = mu * (dvdx + dudy);
        F(2, 1) = mu * (2.0 * dvdy - P.mu_bulk * div);
        F(3, 1) = prim[1] * F(1, 1) + prim[2] * F(2, 1) + lambda * cv_dTdy;
    }
}

// Rounding scale of ViscousFlux: the same formula with every term in absolute value and
// the gradients replaced by their rounding scales (gmag >= |grad| entrywise).
void ViscousFluxMag(const Phys &P, const real *state, const real *ax, const real *ay, const real *az,
                    real *flux) {
    const int dim = P.dim, neq = P.neq;
    const real *g[3] = {ax, ay, az};
    const real rho = state[0];
    real vel[3], vsq = 0.0;
    for (int d = 0; d < dim; d++) {
        vel[d] = state[1 + d] / rho;
        vsq += vel[d] * vel[d];
    }
    const real p = P.gammaM1 * (state[neq - 1] - 0.5 * rho * vsq);
    real mu = P.mu;
    if (P.sutherland) {
        real T = p / (rho * P.R_gas);
        mu = P.mu0 * (P.T0 + P.Ts) / (T + P.Ts) * (T / P.T0) * std::sqrt(T / P.T0);
    }
    mu = std::fabs(mu);
    const real lambda = mu * P.gamma * P.PrInverse;
    real div = 0.0;
    for (int d = 0; d < dim; d++) div += g[d][1 + d];
    for (int i = 0; i < neq * dim; i++) flux[i] = 0.0;
    for (int c = 0; c < dim; c++) {
        real e = 0.0;
        for (int r = 0; r < dim; r++) {
            const real t = (r == c) ? mu * (2.0 * g[c][1 + r] + P.mu_bulk * div) : mu * (g[c][1 + r] + g[r][1 + c]);
            flux[(1 + r) * dim + c] = t;
            e += std::fabs(vel[r]) * t;
        }
        const real cv = P.gammaM1Inverse / rho * (g[c][neq - 1] + std::fabs(p / rho) * g[c][0]);
        flux[(neq - 1) * dim + c] = e + lambda * cv;
    }
}

// ---------------------------------------------------------------------------
// GLL rule on [0,1] (MFEM IntegrationRules(0, GaussLobatto), DGSEMIntegrator.cpp:18-37),
// restated from its closed form: interior nodes are the roots of P'_N, found by
// Newton's method in long double on [-1,1] then mapped to [0,1].
// ---------------------------------------------------------------------------
void LegendreAndDeriv(int N, long double z, long double &P, long double &dP) {
    long double p0 = 1.0L, p1 = z;
    if (N == 0) { P = 1.0L; dP = 0.0L; return; }
    for (int n = 1; n < N; n++) {
        long double p2 = ((2 * n + 1) * z * p1 - n * p0) / (n + 1);
        p0 = p1;
        p1 = p2;
    }
    P = p1;
    dP = N * (z * p1 - p0) / (z * z - 1.0L);  // valid for |z| < 1
}

void GLL(int np, real *x, real *w) {
    const int N = np - 1;
    std::vector<long double> z(np);
    z[0] = -1.0L;
    z[N] = 1.0L;
    for (int j = 1; j < N; j++) {
        // Chebyshev-Gauss-Lobatto initial guess, Newton on q(z) = P_{N-1}(z) - P_{N+1}(z)
        long double zz = -std::cos(3.14159265358979323846264338327950288L * j / N);
        for (int it = 0; it < 100; it++) {
            // P'_N roots: use q = (1-z^2) P'_N = N (P_{N-1} - z P_N)
            long double P, dP;
            LegendreAndDeriv(N, zz, P, dP);
            // second derivative from the Legendre ODE: (1-z^2)P'' = 2zP' - N(N+1)P
            long double d2P = (2.0L * zz * dP - N * (N + 1.0L) * P) / (1.0L - zz * zz);
            long double dz = dP / d2P;
            zz -= dz;
            if (std::fabs((real)dz) < 1e-30) break;
        }
        z[j] = zz;
    }
    for (int j = 0; j < np; j++) {
        long double P, dP;
        if (j == 0 || j == N) {
            P = (j == 0 && (N % 2 == 1)) ? -1.0L : 1.0L;
        } else {
            LegendreAndDeriv(N, z[j], P, dP);
        }
        long double wj = 1.0L / (N * (N + 1.0L) * P * P);  // on [0,1]: half of 2/(N(N+1)P^2)
        x[j] = (real)(0.5L * (z[j] + 1.0L));
        w[j] = (real)wj;
    }
    x[0] = 0.0;
    x[N] = 1.0;
}

// D, Dhat, Dhat2 -- ref: DGSEMIntegrator.cpp:41-82.  Stored as the mathematical
// matrices (D[i*np+j] = D_ij), i.e. before the reference's transposes.
void DMatrices(int np, const real *x, const real *w, real *D, real *Dhat, real *Dhat2) {
    real wBary[MAXNP];
    for (int i = 0; i < np; i++) wBary[i] = 1.0;
    for (int i = 1; i < np; i++)
        for (int j = 0; j < i; j++) {
            wBary[j] *= (x[j] - x[i]);
            wBary[i] *= (x[i] - x[j]);
        }
    for (int i = 0; i < np; i++) wBary[i] = 1.0 / wBary[i];       // Reciprocal :57
    for (int i = 0; i < np * np; i++) D[i] = 0.0;
    for (int iL = 0; iL < np; iL++)
        for (int i = 0; i < np; i++)
            if (iL != i) {
                D[i * np + iL] = wBary[iL] / wBary[i] / (x[i] - x[iL]);   // :65
                D[i * np + i] -= D[i * np + iL];                          // :66
            }
    for (int i = 0; i < np * np; i++) {
        Dhat[i] = D[i];
        Dhat2[i] = 2.0 * D[i];
    }
    Dhat[0] += 1.0 / w[0];                                    // :72
    Dhat[(np - 1) * np + np - 1] -= 1.0 / w[np - 1];          // :73
    Dhat2[0] += 1.0 / w[0];                                   // :78
    Dhat2[(np - 1) * np + np - 1] -= 1.0 / w[np - 1];         // :79
}

// ---------------------------------------------------------------------------
// Boundary conditions -- ref: src/Integrators/BoundaryConditions/ (the BdrFaceIntegrator
// hierarchy, one integrator per boundary attribute, Simulation.cpp:505-800), with constant
// boundary values.  nor is the element's outward face normal (CalcOrtho, scaled by the face
// area), state1 the element's trace.  Each BC supplies
//   ComputeBdrFaceInviscidFlux  -- the surface flux f(state1, state2; nor) and a wave speed;
//   ComputeBdrFaceViscousFlux   -- F_v . nor at the face (Navier-Stokes);
//   ComputeBdrFaceLiftingFlux   -- the BR1 face term (entropy variables), default
//                                  1/2 (w(state2) - w1) (BdrFaceIntegrator.cpp:275-282).
// The slip wall and symmetry BCs override only the inviscid flux; their viscous and lifting
// fluxes reach the base class's mfem_error (Euler only in the reference).
// ---------------------------------------------------------------------------
enum BcType {
    BC_SLIP_WALL = 1, BC_SYMMETRY = 2, BC_SUPERSONIC_INFLOW = 3, BC_SUPERSONIC_OUTFLOW = 4,
    BC_SPECIFIED_STATE = 5, BC_RIEMANN_INVARIANT = 6, BC_SUBSONIC_INFLOW_RV = 7,
    BC_SUBSONIC_INFLOW_PTTT = 8, BC_SUBSONIC_OUTFLOW_P = 9, BC_NOSLIP_ADIABATIC = 10,
    BC_NOSLIP_ISOTHERMAL = 11
};

// ComputePressure -- BasicOperations.cpp:31-37 (V = m/rho as Vector /= rho, i.e. times 1/rho)
real PressureOf(const Phys &P, const real *s) {
    const real inv = 1.0 / s[0];
    real vv = 0.0;
    for (int d = 0; d < P.dim; d++) vv += (s[1 + d] * inv) * (s[1 + d] * inv);
    return P.gammaM1 * (s[P.neq - 1] - 0.5 * s[0] * vv);
}

void NormalizeVec(int dim, real *v) {   // Normalize, BasicOperations.hpp:39-42
    real n = 0.0;
    for (int d = 0; d < dim; d++) n += v[d] * v[d];
    n = std::sqrt(n);
    for (int d = 0; d < dim; d++) v[d] /= n;
}

// Normal (BasicOperations.cpp:221-235), Cross (:211-219), RotateState (:238-259): the momentum
// in the frame (n, t1[, n x t1]).
void RotateStateRef(int dim, real *state, const real *n) {
    real t1[3];
    if (dim == 2) {
        t1[0] = -n[1];
        t1[1] = n[0];
    } else {
        const real r = n[1] / n[0];
        t1[0] = std::abs(n[0]) > 1e-10 ? -r / std::sqrt(1.0 + std::pow(r, 2)) : -1.0;
        t1[1] = std::abs(n[0]) > 1e-10 ? 1.0 / std::sqrt(1.0 + std::pow(r, 2)) : 0.0;
        t1[2] = 0.0;
    }
    real *m = state + 1;
    real mu = 0.0, mv = 0.0;
    for (int d = 0; d < dim; d++) {
        mu += m[d] * n[d];
        mv += m[d] * t1[d];
    }
    if (dim == 3) {
        const real t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
        m[2] = m[0] * t2[0] + m[1] * t2[1] + m[2] * t2[2];
    }
    m[0] = mu;
    m[1] = mv;
}

// The value of a state-valued condition at the face point x and time t: the constant vector, or
// the function a test case registered with the ConditionFactory (src/Simulation/
// ConditionFactory.hpp:16-40), which the reference evaluates per face point through a
// VectorFunctionCoefficient / VectorTDFunctionCoefficient at the operator's time
// (Simulation.cpp:664-790; SpecifiedStateBdrFaceIntegrator.cpp:26-40).  Double arithmetic, as
// the registered functions compute.
enum BcProfile { PROFILE_CONSTANT = 0, PROFILE_FRONT = 1, PROFILE_RAMP = 2 };
void BcState(int dim, double gamma, const orc_bc &bc, const double *x, double t, double *y) {
    const int neq = dim + 2;
    if (bc.profile == PROFILE_FRONT) {
        // DoubleMachReflectionTopBCFunction (TestCases/Euler/2D/DoubleMachReflection/
        // DoubleMachReflection.hpp:35-51): if (x(0) < 1.0 / 6.0 + (x(1) + 20.0 * t) / sqrt(3)) the
        // post-shock state, else the pre-shock state; here with the coordinates, offset, speed and
        // divisor as parameters
        const int i0 = (int)bc.a[0], i1 = (int)bc.a[1];
        if (x[i0] < bc.a[2] + (x[i1] + bc.a[3] * t) / bc.a[4])
            for (int v = 0; v < neq; v++) y[v] = bc.q[v];
        else
            for (int v = 0; v < neq; v++) y[v] = bc.a[5 + v];
    } else if (bc.profile == PROFILE_RAMP) {
        // NagashimaRampingInletBC (TestCases/Euler/2D/Nagashima_Ramjet/Nagashima_Ramjet.hpp:43-53):
        // s = tanh((t - 0.5 t_ramp)/(0.1 t_ramp)), ramp = clamp(0.5 (1 + s), 0, 1), u = V ramp,
        // y = (rho, rho u, pi/(gamma - 1) + 0.5 rho u u)
        const double t_ramp = bc.a[0];
        const double sv = std::tanh((t - 0.5 * t_ramp) / (0.1 * t_ramp));
        const double ramp = std::clamp(0.5 * (1.0 + sv), 0.0, 1.0);
        const double rho = bc.a[1];
        y[0] = rho;
        double e = bc.a[2 + dim] / (gamma - 1.0);
        for (int d = 0; d < dim; d++) {
            const double u = bc.a[2 + d] * ramp;
            y[1 + d] = rho * u;
            e += 0.5 * rho * u * u;
        }
        y[neq - 1] = e;
    } else {
        for (int v = 0; v < neq; v++) y[v] = bc.q[v];
    }
}

// ComputeOuterInviscidState of the BCs that define an outer state; x, t: the face point's
// coordinates and the operator's time (read by the function-valued states only).
void OuterState(const Phys &P, const orc_bc &bc, const real *s1, const real *nor, real *s2, const double *x,
                double t) {
    const int dim = P.dim, neq = P.neq;
    switch (bc.type) {
        case BC_SUPERSONIC_INFLOW:    // SupersonicInflowBdrFaceIntegrator.cpp:34-48
        case BC_SPECIFIED_STATE: {    // SpecifiedStateBdrFaceIntegrator.cpp:26-40
            double y[MAXEQ];
            BcState(dim, (double)P.gamma, bc, x, t, y);
            for (int v = 0; v < neq; v++) s2[v] = y[v];
            break;
        }
        case BC_SUPERSONIC_OUTFLOW:   // SupersonicOutflowBdrFaceIntegrator.cpp:12-15
            for (int v = 0; v < neq; v++) s2[v] = s1[v];
            break;
        case BC_RIEMANN_INVARIANT: {  // RiemannInvariantBdrFaceIntegrator.cpp:44-117
            real un[3], Vo[3], Vi[3];
            for (int d = 0; d < dim; d++) un[d] = nor[d];
            NormalizeVec(dim, un);
            const real rho_o = bc.q[0], p_o = bc.q[neq - 1];
            for (int d = 0; d < dim; d++) Vo[d] = bc.q[1 + d];
            const real inv = 1.0 / s1[0];
            real Vn_i = 0.0, Vn_o = 0.0;
            for (int d = 0; d < dim; d++) {
                Vi[d] = s1[1 + d] * inv;
                Vn_i += Vi[d] * un[d];
                Vn_o += Vo[d] * un[d];
            }
            const real p_i = PressureOf(P, s1);
            const real a_o = SoundSpeed(p_o, rho_o, P.gamma), a_i = SoundSpeed(p_i, s1[0], P.gamma);
            const real gI = P.gammaM1Inverse;
            const real R_minus = (std::abs(Vn_o) >= a_o && Vn_i >= 0.0) ? Vn_i - 2.0 * a_i * gI : Vn_o - 2.0 * a_o * gI;
            const real R_plus = (std::abs(Vn_o) >= a_o && Vn_i < 0.0) ? Vn_o + 2.0 * a_o * gI : Vn_i + 2.0 * a_i * gI;
            const real Vn_b = 0.5 * (R_plus + R_minus);
            const real a_b = 0.25 * P.gammaM1 * (R_plus - R_minus);
            const real dVn_i = Vn_b - Vn_i, dVn_o = Vn_b - Vn_o;
            const real gInv = 1.0 / P.gamma;
            s2[0] = (Vn_i < 0.0) ? std::pow(a_b * a_b * gInv * rho_o / p_o, gI) * rho_o
                                 : std::pow(a_b * a_b * gInv * s1[0] / p_i, gI) * s1[0];
            real e = 0.0;
            for (int d = 0; d < dim; d++) {
                s2[1 + d] = (Vn_i >= 0.0) ? s2[0] * (Vi[d] + dVn_i * un[d]) : s2[0] * (Vo[d] + dVn_o * un[d]);
                e += s2[1 + d] * s2[1 + d];
            }
            const real p_b = a_b * a_b * s2[0] * gInv;
            s2[neq - 1] = e * (0.5 / s2[0]) + p_b * gI;
            break;
        }
        case BC_SUBSONIC_INFLOW_RV: { // SubsonicInflowRVBdrFaceIntegrator.cpp:22-52
            const real r = bc.q[0];
            real dke = 0.0, uu = 0.0;
            s2[0] = r;
            s2[neq - 1] = s1[neq - 1];
            for (int d = 0; d < dim; d++) {
                s2[1 + d] = r * bc.q[1 + d];
                dke -= s1[1 + d] * s1[1 + d];
                uu += bc.q[1 + d] * bc.q[1 + d];
            }
            dke *= 0.5 / s1[0];
            dke += 0.5 * uu * r;
            s2[neq - 1] += dke;
            break;
        }
        case BC_SUBSONIC_INFLOW_PTTT: {  // SubsonicInflowPtTtAngBdrFaceIntegrator.cpp:48-73
            const real pt = bc.q[0], Tt = bc.q[1], theta = bc.q[2], phi = bc.q[3], cp = bc.q[4];
            real Vc[3] = {std::cos(theta * M_PI / 180.0), std::sin(theta * M_PI / 180.0), 0.0};
            if (dim > 2) {
                // the reference writes cos(phi) to V_comps(3), outside the 3-vector (:19, :39); the
                // intended z component is V_comps(2)
                Vc[0] *= std::sin(phi * M_PI / 180.0);
                Vc[1] *= std::sin(phi * M_PI / 180.0);
                Vc[2] = std::cos(phi * M_PI / 180.0);
            }
            const real p = PressureOf(P, s1);
            real V_sq = 2.0 * cp * Tt * (1.0 - std::pow(p / pt, P.gammaM1 / P.gamma));
            V_sq = std::max<real>(0.0, V_sq);
            s2[0] = (P.gamma * P.gammaM1Inverse) * p / (cp * Tt - 0.5 * V_sq);
            for (int d = 0; d < dim; d++) s2[1 + d] = s2[0] * std::sqrt(V_sq) * Vc[d];
            s2[neq - 1] = p * P.gammaM1Inverse + 0.5 * s2[0] * V_sq;
            break;
        }
        case BC_SUBSONIC_OUTFLOW_P: {  // SubsonicOutflowPBdrFaceIntegrator.cpp:17-28
            const real inv = 1.0 / s1[0];
            real vv = 0.0;
            for (int d = 0; d < dim; d++) vv += (s1[1 + d] * inv) * (s1[1 + d] * inv);
            for (int v = 0; v < neq; v++) s2[v] = s1[v];
            s2[neq - 1] = bc.q[0] * P.gammaM1Inverse + 0.5 * s1[0] * vv;
            break;
        }
        default:
            for (int v = 0; v < neq; v++) s2[v] = s1[v];
    }
}

// ComputeBdrFaceInviscidFlux: the slip wall's (SlipWallBdrFaceIntegrator.cpp:15-57, inherited by
// the no-slip walls), the symmetry BC's (SymmetryBdrFaceIntegrator.cpp:10-32), otherwise the
// surface flux with the outer state (BdrFaceIntegrator.cpp:245-253).
real BdrInviscidFlux(const Phys &P, const orc_bc &bc, const real *s1, const real *nor, real *flux, const double *x,
                     double t) {
    const int dim = P.dim, neq = P.neq;
    real s2[MAXEQ];
    if (bc.type == BC_SLIP_WALL || bc.type == BC_NOSLIP_ADIABATIC || bc.type == BC_NOSLIP_ISOTHERMAL) {
        real un[3];
        for (int d = 0; d < dim; d++) un[d] = nor[d];
        NormalizeVec(dim, un);
        for (int v = 0; v < neq; v++) s2[v] = s1[v];
        RotateStateRef(dim, s2, un);
        // Conserv2Prim (BasicOperations.cpp:172-186)
        real prim[MAXEQ];
        prim[0] = s2[0];
        for (int d = 0; d < dim; d++) prim[1 + d] = s2[1 + d] / s2[0];
        prim[neq - 1] = PressureOf(P, s2);
        real vv = 0.0;
        for (int d = 0; d < dim; d++) vv += prim[1 + d] * prim[1 + d];
        const real an = SoundSpeed(prim[neq - 1], prim[0], P.gamma);
        const real gP1 = 1.0 + P.gamma, gP1I = 1.0 / gP1;
        real p_star;
        if (prim[1] > 0.0)
            p_star = prim[neq - 1] + 0.25 * prim[1] * gP1 * prim[0] *
                                         (prim[1] + std::sqrt(std::pow(prim[1], 2) +
                                                              8.0 * gP1I / prim[0] * prim[neq - 1] * (P.gammaM1 * gP1I + 1.0)));
        else
            p_star = prim[neq - 1] *
                     std::pow(std::max<real>(1.0 + 0.5 * P.gammaM1 * prim[1] / an, 0.0001), 2.0 * P.gamma * P.gammaM1Inverse);
        for (int v = 0; v < neq; v++) flux[v] = 0.0;
        for (int d = 0; d < dim; d++) flux[1 + d] = p_star * nor[d];
        return std::sqrt(vv) + an;
    }
    if (bc.type == BC_SYMMETRY) {
        real un[3], mn = 0.0;
        for (int d = 0; d < dim; d++) un[d] = nor[d];
        NormalizeVec(dim, un);
        for (int v = 0; v < neq; v++) s2[v] = s1[v];
        for (int d = 0; d < dim; d++) mn += s2[1 + d] * un[d];
        for (int d = 0; d < dim; d++) s2[1 + d] += -2.0 * mn * un[d];
        return FaceFlux(P, s1, s2, nor, flux);
    }
    OuterState(P, bc, s1, nor, s2, x, t);
    return FaceFlux(P, s1, s2, nor, flux);
}

// ComputeBdrFaceViscousFlux: F_v . nor (row 0 zero); dq[c] the element's primitive gradients.
void BdrViscousFlux(const Phys &P, const orc_bc &bc, const real *s1, const real (*dq)[MAXEQ], const real *nor,
                    real *fn, const double *x, double t) {
    const int dim = P.dim, neq = P.neq;
    real s2[MAXEQ], zero[3][MAXEQ] = {}, flux[MAXEQ * 3];
    const real(*g)[MAXEQ] = dq;
    const real *sv = s1;
    switch (bc.type) {
        case BC_SUPERSONIC_OUTFLOW:    // state2 = state1, the element's gradients
        case BC_NOSLIP_ADIABATIC:
        case BC_NOSLIP_ISOTHERMAL:
            break;
        case BC_SUBSONIC_INFLOW_PTTT:  // the outer state with the element's gradients (:75-99)
            OuterState(P, bc, s1, nor, s2, x, t);
            sv = s2;
            break;
        default:                       // the outer state with zero gradients
            OuterState(P, bc, s1, nor, s2, x, t);
            sv = s2;
            g = zero;
    }
    ViscousFlux(P, sv, g[0], g[1], dim > 2 ? g[2] : nullptr, flux);
    for (int v = 0; v < neq; v++) {
        real a = 0.0;
        for (int c = 0; c < dim; c++) a += flux[v * dim + c] * nor[c];   // flux_mat.Mult(nor)
        fn[v] = a;
    }
    if (bc.type == BC_NOSLIP_ADIABATIC) {   // NoSlipAdiabWallBdrFaceIntegrator.cpp:30-44
        real nn = 0.0;
        for (int c = 0; c < dim; c++) nn += nor[c] * nor[c];
        const real qn = bc.q[dim] * std::sqrt(nn);
        real e = 0.0;
        for (int d = 0; d < dim; d++) e += bc.q[d] * fn[1 + d];
        fn[neq - 1] = e + qn;
    }
}

// ComputeBdrFaceLiftingFlux: w1 the element's entropy variables at the face point.
void BdrLiftingFlux(const Phys &P, const orc_bc &bc, const real *w1, const real *nor, real *fn, const double *x,
                    double t) {
    const int dim = P.dim, neq = P.neq;
    if (bc.type == BC_NOSLIP_ADIABATIC || bc.type == BC_NOSLIP_ISOTHERMAL) {
        // NoSlipAdiabWallBdrFaceIntegrator.cpp:74-95, NoSlipIsothWallBdrFaceIntegrator.cpp:41-66
        const real v = bc.type == BC_NOSLIP_ADIABATIC ? -w1[neq - 1] : 1.0 / (P.R_gas * bc.q[dim]);
        fn[0] = w1[0];
        for (int d = 0; d < dim; d++) fn[1 + d] = bc.q[d] * v;
        fn[neq - 1] = -v;
        for (int k = 0; k < neq; k++) fn[k] -= w1[k];
        return;
    }
    // BdrFaceIntegrator.cpp:275-282
    real s1[MAXEQ], s2[MAXEQ], w2[MAXEQ];
    Entropy2Conserv(P, w1, s1);
    OuterState(P, bc, s1, nor, s2, x, t);
    Conserv2Entropy(P, s2, w2);
    for (int k = 0; k < neq; k++) fn[k] = (w2[k] - w1[k]) * 0.5;
}

// ---------------------------------------------------------------------------
// Mesh: conforming affine elements with aligned faces.  neighbors[e*2*dim + 2*d + s]
// is the element across local face (d, s) (s = 0: xi_d = 0 side, s = 1: xi_d = 1 side).
// geom[e*(1+dim*dim)] = det A, followed by adj A row-major.
// ---------------------------------------------------------------------------
struct Mesh {
    int dim, np, nn, NE;
    const int *nbr;                 // neighbour, or -(k+1) for a boundary face with condition bc[k]
    const real *geom;
    const orc_bc *bc = nullptr;
    int nbc = 0;
    real x[MAXNP], w[MAXNP], D[MAXNP * MAXNP], Dhat[MAXNP * MAXNP], Dhat2[MAXNP * MAXNP];
    real J(int e) const { return geom[(size_t)e * (1 + dim * dim)]; }
    real adj(int e, int r, int c) const { return geom[(size_t)e * (1 + dim * dim) + 1 + r * dim + c]; }
    int lex(int i, int j, int k) const { return i + np * (j + np * k); }
    // general geometry: per-node (det J, adj J) -- the reference evaluates Tr.Weight() and
    // Tr.AdjugateJacobian() at every node (DGSEMIntegrator.cpp:254-276) -- and per-face
    // orientation codes (neighbour's local face 2 d' + s', bit 3 flips the first face
    // coordinate, bit 4 the second, bit 5 swaps them); null: affine elements, aligned faces.
    const real *ngeom = nullptr;
    const int *forient = nullptr;
    // function-valued boundary states: node coordinates [e][node][dim] and the operator's time
    const double *xyz = nullptr;
    double time = 0.0;
    const double *xat(int e, int node) const { return xyz ? xyz + ((size_t)e * nn + node) * dim : nullptr; }
    real Jn(int e, int node) const {
        return ngeom ? ngeom[((size_t)e * nn + node) * (1 + dim * dim)] : J(e);
    }
    real adjn(int e, int node, int r, int c) const {
        return ngeom ? ngeom[((size_t)e * nn + node) * (1 + dim * dim) + 1 + r * dim + c] : adj(e, r, c);
    }
    int nbface(int e, int f) const { return forient ? (forient[(size_t)e * 2 * dim + f] & 7) : (f ^ 1); }
    // the neighbour's face coordinates of the point (t1, t2) of face f of e
    void nbpoint(int e, int f, int t1, int t2, int &u1, int &u2) const {
        const int code = forient ? forient[(size_t)e * 2 * dim + f] : 0;
        int a = t1, b = t2;
        if (code & 32) std::swap(a, b);
        u1 = (code & 8) ? np - 1 - a : a;
        u2 = (code & 16) ? np - 1 - b : b;
        if (dim == 2) u2 = 0;
    }
};

inline size_t gidx(const Mesh &M, int v, int e, int node) {
    return ((size_t)v * M.NE + e) * M.nn + node;
}

// node index of the point with index a along direction d and tangential
// indices (t1, t2) (in increasing direction order)
inline int line_node(const Mesh &M, int d, int a, int t1, int t2) {
    if (M.dim == 2) return d == 0 ? M.lex(a, t1, 0) : M.lex(t1, a, 0);
    if (d == 0) return M.lex(a, t1, t2);
    if (d == 1) return M.lex(t1, a, t2);
    return M.lex(t1, t2, a);
}

// ---------------------------------------------------------------------------
// Euler element term -- ref: DGSEMIntegrator::AssembleElementVector, DGSEMIntegrator.cpp:224-433
// (blending off; AXISYMMETRIC off).  Adds into el_dudt (n x neq, column-major as
// the reference's DenseMatrix el_dudt_mat).  Returns the running lambda max.
// NS 3D element -- ref: :601-750 (with `visc` = the per-node contravariant viscous
// fluxes F_viscous/G_viscous/H_viscous, and the A1 bug under P.bug & 1).
// ---------------------------------------------------------------------------
real ElementInviscid(const Mesh &M, const Phys &P, int e, const real *el_u, const real *Fvis,
                       real *el_dudt, real *el_mag, const real *FvisMag, const real *alpha) {
    const int np = M.np, n = M.nn, dim = P.dim, neq = P.neq;
    const int Npy = dim > 1 ? np : 1, Npz = dim > 2 ? np : 1;
    real lam = 0.0;
    // F_inviscid(id)(v, m): DenseTensor(num_equations, Np, n) -- :120-122
    std::vector<real> Fi((size_t)n * np * neq), Gi((size_t)n * np * neq), Hi((size_t)n * np * neq);
    auto FI = [&](std::vector<real> &T, int id, int v, int m) -> real & {
        return T[((size_t)id * np + m) * neq + v];
    };
    real s1[MAXEQ], s2[MAXEQ], f[MAXEQ], g[MAXEQ], h[MAXEQ], m1[3], m2[3];
    for (int k = 0; k < Npz; k++)
        for (int j = 0; j < Npy; j++)
            for (int i = 0; i < np; i++) {
                const int id1 = M.lex(i, j, k);
                const real J = M.Jn(e, id1);                             // Tr.Weight() at id1 (:254)
                for (int v = 0; v < neq; v++) s1[v] = el_u[v * n + id1];
                for (int c = 0; c < dim; c++) m1[c] = M.adjn(e, id1, 0, c);
                for (int v = 0; v < neq; v++) FI(Fi, id1, v, i) = 0.0;      // f = 0; SetCol(i, f)
                for (int m = i + 1; m < np; m++) {
                    const int id2 = M.lex(m, j, k);
                    for (int v = 0; v < neq; v++) s2[v] = el_u[v * n + id2];
                    for (int c = 0; c < dim; c++) m2[c] = M.adjn(e, id2, 0, c);
                    lam = std::max(lam, VolumeFlux(P, s1, s2, m1, m2, f));
                    for (int v = 0; v < neq; v++) {
                        FI(Fi, id1, v, m) = f[v];
                        FI(Fi, id2, v, i) = f[v];
                    }
                }
                if (dim > 1) {
                    for (int c = 0; c < dim; c++) m1[c] = M.adjn(e, id1, 1, c);
                    for (int v = 0; v < neq; v++) g[v] = 0.0;
                    for (int v = 0; v < neq; v++) FI(Gi, id1, v, j) = 0.0;
                    for (int m = j + 1; m < Npy; m++) {
                        const int id2 = M.lex(i, m, k);
                        for (int v = 0; v < neq; v++) s2[v] = el_u[v * n + id2];
                        for (int c = 0; c < dim; c++) m2[c] = M.adjn(e, id2, 1, c);
                        lam = std::max(lam, VolumeFlux(P, s1, s2, m1, m2, g));
                        for (int v = 0; v < neq; v++) {
                            FI(Gi, id1, v, m) = g[v];
                            FI(Gi, id2, v, j) = g[v];
                        }
                    }
                    if (dim > 2) {
                        for (int c = 0; c < dim; c++) m1[c] = M.adjn(e, id1, 2, c);
                        for (int v = 0; v < neq; v++) FI(Hi, id1, v, k) = 0.0;
                        for (int m = k + 1; m < Npz; m++) {
                            const int id2 = M.lex(i, j, m);
                            for (int v = 0; v < neq; v++) s2[v] = el_u[v * n + id2];
                            for (int c = 0; c < dim; c++) m2[c] = M.adjn(e, id2, 2, c);
                            lam = std::max(lam, VolumeFlux(P, s1, s2, m1, m2, h));
                            for (int v = 0; v < neq; v++) {
                                FI(Hi, id1, v, m) = h[v];
                                // :702 as written stores g (NS 3D only); :363 (Euler) stores h
                                FI(Hi, id2, v, k) = (Fvis && (P.bug & 1)) ? g[v] : h[v];
                            }
                        }
                    }
                }
                // D-hat-2 accumulation -- :368-390 / :707-738
                real dUi[MAXEQ], dUv[MAXEQ];
                for (int v = 0; v < neq; v++) {
                    real acc = 0.0;
                    for (int m = 0; m < np; m++) acc += FI(Fi, id1, v, m) * M.Dhat2[i * np + m];
                    dUi[v] = acc;
                }
                for (int v = 0; v < neq; v++) dUv[v] = 0.0;
                if (Fvis) {
                    for (int l = 0; l < np; l++) {
                        const int idl = M.lex(l, j, k);
                        for (int v = 0; v < neq; v++) dUv[v] += Fvis[(0 * n + idl) * neq + v] * M.Dhat[i * np + l];
                    }
                }
                if (dim > 1) {
                    for (int m = 0; m < np; m++)                          // AddMult: y += A x
                        for (int v = 0; v < neq; v++) dUi[v] += FI(Gi, id1, v, m) * M.Dhat2[j * np + m];
                    if (Fvis)
                        for (int l = 0; l < np; l++) {
                            const int idl = M.lex(i, l, k);
                            for (int v = 0; v < neq; v++) dUv[v] += Fvis[(1 * n + idl) * neq + v] * M.Dhat[j * np + l];
                        }
                    if (dim > 2) {
                        for (int m = 0; m < np; m++)
                            for (int v = 0; v < neq; v++) dUi[v] += FI(Hi, id1, v, m) * M.Dhat2[k * np + m];
                        if (Fvis)
                            for (int l = 0; l < np; l++) {
                                const int idl = M.lex(i, j, l);
                                for (int v = 0; v < neq; v++) dUv[v] += Fvis[(2 * n + idl) * neq + v] * M.Dhat[k * np + l];
                            }
                    }
                }
                for (int v = 0; v < neq; v++) {
                    real dU = -dUi[v];                                  // Neg :392 / :740
                    if (alpha) dU *= (1.0 - *alpha);                      // :395 / :742
                    if (Fvis) dU = dU + dUv[v];                           // add :744
                    el_dudt[v * n + id1] += dU * (1.0 / J);               // :403 / :745 (Vector /= c is *= 1/c), AddRow
                }
                if (el_mag) {
                    // rounding scale: the same sums with every term in absolute value
                    const real sa = alpha ? std::fabs(1.0 - *alpha) : 1.0;
                    for (int v = 0; v < neq; v++) {
                        real a = 0.0;
                        for (int m = 0; m < np; m++) {
                            a += sa * std::fabs(FI(Fi, id1, v, m) * M.Dhat2[i * np + m]);
                            if (dim > 1) a += sa * std::fabs(FI(Gi, id1, v, m) * M.Dhat2[j * np + m]);
                            if (dim > 2) a += sa * std::fabs(FI(Hi, id1, v, m) * M.Dhat2[k * np + m]);
                            if (FvisMag) {
                                a += FvisMag[(0 * n + M.lex(m, j, k)) * neq + v] * std::fabs(M.Dhat[i * np + m]);
                                if (dim > 1) a += FvisMag[(1 * n + M.lex(i, m, k)) * neq + v] * std::fabs(M.Dhat[j * np + m]);
                                if (dim > 2) a += FvisMag[(2 * n + M.lex(i, j, m)) * neq + v] * std::fabs(M.Dhat[k * np + m]);
                            }
                        }
                        el_mag[v * n + id1] += a / std::fabs(J);
                    }
                }
            }
    return lam;
}

// ---------------------------------------------------------------------------
// Subcell finite-volume term of the blended scheme -- ref: DGSEMIntegrator::ComputeFVFluxes
// (DGSEMIntegrator.cpp:972-1177) with the subcell metrics of ComputeSubcellMetrics
// (:1180-1304); AXISYMMETRIC off.  Along every line in direction d the interface between
// nodes i and i+1 carries the surface flux f*(u_i, u_{i+1}; n_{i+1}) (the surface-flux
// plug-in, :1011); node i receives (f_{i-1/2} - f_{i+1/2}) / (J w_i) (:1042-1057), the two
// outer interfaces of the line excluded (the DG surface term carries them).  The element's
// sum is scaled by alpha (:1177).  Subcell normal n_i = adj_d(x_first) + sum_{l<i} w_l
// sum_m D_lm adj_d(x_m) (:1194-1211; D_T holds D^T, so its column l is row l of D), evaluated
// as written: on the affine elements here adj is constant and the sums vanish to rounding.
// ---------------------------------------------------------------------------
real ElementSubcellFV(const Mesh &M, const Phys &P, int e, const real *el_u, real alpha, real *el_dudt,
                        real *el_mag) {
    const int np = M.np, n = M.nn, dim = P.dim, neq = P.neq;
    const int Nt1 = dim > 1 ? np : 1, Nt2 = dim > 2 ? np : 1;
    real lam = 0.0;
    std::vector<real> fv((size_t)n * neq, 0.0), fm((size_t)n * neq, 0.0);
    for (int d = 0; d < dim; d++) {
        for (int t2 = 0; t2 < Nt2; t2++)
            for (int t1 = 0; t1 < Nt1; t1++) {
                // subcell normals of this line (ComputeSubcellMetrics :1194-1211): metric1 at the
                // line's first node plus the weighted D-sums of the line's adj rows
                real nor[MAXNP][3];
                for (int i = 0; i < np; i++) {
                    for (int c = 0; c < dim; c++) nor[i][c] = M.adjn(e, line_node(M, d, 0, t1, t2), d, c);
                    for (int l = 0; l < i; l++) {
                        real tmp[3] = {0.0, 0.0, 0.0};
                        for (int m = 0; m < np; m++)
                            for (int c = 0; c < dim; c++) tmp[c] += M.adjn(e, line_node(M, d, m, t1, t2), d, c) * M.D[l * np + m];
                        for (int c = 0; c < dim; c++) nor[i][c] += tmp[c] * M.w[l];
                    }
                }
                real dU[MAXEQ], aU[MAXEQ], s1[MAXEQ], s2[MAXEQ], f[MAXEQ];
                for (int v = 0; v < neq; v++) dU[v] = aU[v] = 0.0;
                int id1 = line_node(M, d, 0, t1, t2);
                for (int v = 0; v < neq; v++) s1[v] = el_u[v * n + id1];
                for (int i = 0; i < np - 1; i++) {
                    const int id2 = line_node(M, d, i + 1, t1, t2);
                    for (int v = 0; v < neq; v++) s2[v] = el_u[v * n + id2];
                    lam = std::max(lam, FaceFlux(P, s1, s2, nor[i + 1], f));
                    const real J = M.Jn(e, id1);                           // Tr.Weight() at id1
                    const real sc = 1.0 / (J * M.w[i]);                  // Vector /= c multiplies by 1/c
                    for (int v = 0; v < neq; v++) {
                        dU[v] -= f[v];
                        fv[(size_t)v * n + id1] += dU[v] * sc;             // SetRow (xi) / AddRow (eta, zeta)
                        fm[(size_t)v * n + id1] += (aU[v] + std::fabs(f[v])) / std::fabs(J * M.w[i]);
                        dU[v] = f[v];
                        aU[v] = std::fabs(f[v]);
                    }
                    for (int v = 0; v < neq; v++) s1[v] = s2[v];
                    id1 = id2;
                }
                const real J = M.Jn(e, id1);
                const real sc = 1.0 / (J * M.w[np - 1]);
                for (int v = 0; v < neq; v++) {
                    fv[(size_t)v * n + id1] += dU[v] * sc;
                    fm[(size_t)v * n + id1] += aU[v] / std::fabs(J * M.w[np - 1]);
                }
            }
    }
    for (size_t i = 0; i < (size_t)n * neq; i++) {
        el_dudt[i] += fv[i] * alpha;                                        // el_dudt_mat *= alpha
        if (el_mag) el_mag[i] += fm[i] * std::fabs(alpha);
    }
    return lam;
}

struct Rhs {
    const Mesh &M;
    const Phys &P;
    int nthreads;
};

// el_u gather (GetElementVDofs + GetSubVector, byNODES): el_u[v*n + node]
inline void gather(const Mesh &M, int neq, const real *u, int e, real *el) {
    for (int v = 0; v < neq; v++)
        std::memcpy(el + (size_t)v * M.nn, u + gidx(M, v, e, 0), sizeof(real) * M.nn);
}

// ---------------------------------------------------------------------------
// BR1 lifting of the entropy variables -- ref: DGSEMOperator::Mult :551-571,
// ComputeGlobalEntropyVector :82-94, DGSEMNonlinearForm::MultLifting :424-645,
// LiftingBR1::AssembleLiftingElementVector (3D :156-219, 2D :222-272) and
// AssembleLiftingFaceVector (3D :10-66, 2D :68-115), then
// ComputeGlobalPrimitiveGradVector :96-159.  grad holds dim arrays in the
// layout of u.
// ---------------------------------------------------------------------------
void LiftingAndPrimGrad(const Mesh &M, const Phys &P, const real *u, real *grad, real *gmag) {
    const int np = M.np, n = M.nn, dim = P.dim, neq = P.neq, NE = M.NE;
    const size_t size = (size_t)neq * NE * n;
    std::vector<real> went(size);
    // ComputeGlobalEntropyVector
#pragma omp parallel for schedule(static)
    for (int e = 0; e < NE; e++) {
        real s[MAXEQ], w[MAXEQ];
        for (int node = 0; node < n; node++) {
            for (int v = 0; v < neq; v++) s[v] = u[gidx(M, v, e, node)];
            Conserv2Entropy(P, s, w);
            for (int v = 0; v < neq; v++) went[gidx(M, v, e, node)] = w[v];
        }
    }
    // element lifting
#pragma omp parallel for schedule(static)
    for (int e = 0; e < NE; e++) {
        const int Npy = dim > 1 ? np : 1, Npz = dim > 2 ? np : 1;
        for (int k = 0; k < Npz; k++)
            for (int j = 0; j < Npy; j++)
                for (int i = 0; i < np; i++) {
                    const int id = M.lex(i, j, k);
                    const real J = M.Jn(e, id);
                    real gm1[MAXEQ][3];                          // grad_mat1 (neq x dim)
                    for (int v = 0; v < neq; v++)
                        for (int c = 0; c < 3; c++) gm1[v][c] = 0.0;
                    for (int l = 0; l < np; l++) {
                        for (int v = 0; v < neq; v++) {
                            gm1[v][0] += went[gidx(M, v, e, M.lex(l, j, k))] * M.D[i * np + l];   // :192-194
                            if (dim > 1) gm1[v][1] += went[gidx(M, v, e, M.lex(i, l, k))] * M.D[j * np + l];
                            if (dim > 2) gm1[v][2] += went[gidx(M, v, e, M.lex(i, j, l))] * M.D[k * np + l];
                        }
                    }
                    // grad_mat2 = grad_mat1 * adj; *= 1/J  -- :205-206
                    for (int v = 0; v < neq; v++)
                        for (int c = 0; c < dim; c++) {
                            real acc = 0.0;
                            for (int r = 0; r < dim; r++) acc += gm1[v][r] * M.adjn(e, id, r, c);
                            grad[(size_t)c * size + gidx(M, v, e, id)] = acc * (1.0 / J);
                        }
                    if (gmag) {
                        real am[MAXEQ][3];
                        for (int v = 0; v < neq; v++)
                            for (int r = 0; r < 3; r++) am[v][r] = 0.0;
                        for (int l = 0; l < np; l++)
                            for (int v = 0; v < neq; v++) {
                                am[v][0] += std::fabs(went[gidx(M, v, e, M.lex(l, j, k))] * M.D[i * np + l]);
                                if (dim > 1) am[v][1] += std::fabs(went[gidx(M, v, e, M.lex(i, l, k))] * M.D[j * np + l]);
                                if (dim > 2) am[v][2] += std::fabs(went[gidx(M, v, e, M.lex(i, j, l))] * M.D[k * np + l]);
                            }
                        for (int v = 0; v < neq; v++)
                            for (int c = 0; c < dim; c++) {
                                real acc = 0.0;
                                for (int r = 0; r < dim; r++) acc += am[v][r] * std::fabs(M.adjn(e, id, r, c));
                                gmag[(size_t)c * size + gidx(M, v, e, id)] = acc / std::fabs(J);
                            }
                    }
                }
    }
    // face lifting: 1/2 (w2 - w1) n_c / (w0 J) added to both sides
    const int Nt1 = dim > 1 ? np : 1, Nt2 = dim > 2 ? np : 1;
    if (M.forient) {
        // general connectivity: every interior face once, from its lower-numbered side (serial:
        // a face may meet the neighbour in another direction)
        for (int e1 = 0; e1 < NE; e1++)
            for (int f1 = 0; f1 < 2 * dim; f1++) {
                const int e2 = M.nbr[(size_t)e1 * 2 * dim + f1];
                const int f2 = M.nbface(e1, f1);
                if (e2 < 0 || e2 < e1 || (e2 == e1 && f2 < f1)) continue;
                const int d1 = f1 >> 1, s1 = f1 & 1, d2 = f2 >> 1, s2 = f2 & 1;
                for (int t2 = 0; t2 < Nt2; t2++)
                    for (int t1 = 0; t1 < Nt1; t1++) {
                        int u1, u2;
                        M.nbpoint(e1, f1, t1, t2, u1, u2);
                        const int n1 = line_node(M, d1, s1 ? np - 1 : 0, t1, t2), n2 = line_node(M, d2, s2 ? np - 1 : 0, u1, u2);
                        const real J1 = M.Jn(e1, n1), J2 = M.Jn(e2, n2);
                        real nor[3];
                        for (int c = 0; c < dim; c++) nor[c] = (s1 ? 1.0 : -1.0) * M.adjn(e1, n1, d1, c);   // outward from e1
                        for (int v = 0; v < neq; v++) {
                            const real f = 0.5 * (went[gidx(M, v, e2, n2)] - went[gidx(M, v, e1, n1)]);
                            for (int c = 0; c < dim; c++) {
                                const real fc = f * nor[c];
                                grad[(size_t)c * size + gidx(M, v, e1, n1)] += 1.0 / (M.w[0] * J1) * fc;
                                grad[(size_t)c * size + gidx(M, v, e2, n2)] += 1.0 / (M.w[0] * J2) * fc;
                                if (gmag) {
                                    const real a = 0.5 * (std::fabs(went[gidx(M, v, e2, n2)]) +
                                                            std::fabs(went[gidx(M, v, e1, n1)])) * std::fabs(nor[c]);
                                    gmag[(size_t)c * size + gidx(M, v, e1, n1)] += a / (M.w[0] * std::fabs(J1));
                                    gmag[(size_t)c * size + gidx(M, v, e2, n2)] += a / (M.w[0] * std::fabs(J2));
                                }
                            }
                        }
                    }
            }
    }
    for (int d = 0; d < dim && !M.forient; d++) {
#pragma omp parallel for schedule(static)
        for (int e1 = 0; e1 < NE; e1++) {
            const int e2 = M.nbr[(size_t)e1 * 2 * dim + 2 * d + 1];
            if (e2 < 0) continue;   // a boundary face (below)
            real nor[3];
            for (int t2 = 0; t2 < Nt2; t2++)
                for (int t1 = 0; t1 < Nt1; t1++) {
                    const int n1 = line_node(M, d, np - 1, t1, t2), n2 = line_node(M, d, 0, t1, t2);
                    const real J1 = M.Jn(e1, n1), J2 = M.Jn(e2, n2);
                    for (int c = 0; c < dim; c++) nor[c] = M.adjn(e1, n1, d, c);   // CalcOrtho
                    for (int v = 0; v < neq; v++) {
                        const real f = 0.5 * (went[gidx(M, v, e2, n2)] - went[gidx(M, v, e1, n1)]);  // subtract(0.5,...) :51
                        for (int c = 0; c < dim; c++) {
                            const real fc = f * nor[c];
                            grad[(size_t)c * size + gidx(M, v, e1, n1)] += 1.0 / (M.w[0] * J1) * fc;
                            grad[(size_t)c * size + gidx(M, v, e2, n2)] += 1.0 / (M.w[0] * J2) * fc;
                            if (gmag) {
                                const real a = 0.5 * (std::fabs(went[gidx(M, v, e2, n2)]) +
                                                        std::fabs(went[gidx(M, v, e1, n1)])) * std::fabs(nor[c]);
                                gmag[(size_t)c * size + gidx(M, v, e1, n1)] += a / (M.w[0] * std::fabs(J1));
                                gmag[(size_t)c * size + gidx(M, v, e2, n2)] += a / (M.w[0] * std::fabs(J2));
                            }
                        }
                    }
                }
        }
    }
    // boundary faces -- LiftingBR1::AssembleLiftingBdrFaceVector (LiftingBR1.cpp:311-407): the
    // BC's lifting flux f times the outward normal, f n_c / (w0 J), to the element's face point.
    // Each face belongs to one element, so the elements run in parallel.
    if (M.bc) {
#pragma omp parallel for schedule(static)
        for (int e = 0; e < NE; e++) {
            for (int d = 0; d < dim; d++)
                for (int side = 0; side < 2; side++) {
                    const int nb = M.nbr[(size_t)e * 2 * dim + 2 * d + side];
                    if (nb >= 0) continue;
                    const orc_bc &bc = M.bc[-nb - 1];
                    for (int t2 = 0; t2 < Nt2; t2++)
                        for (int t1 = 0; t1 < Nt1; t1++) {
                            const int n1 = line_node(M, d, side ? np - 1 : 0, t1, t2);
                            const real J = M.Jn(e, n1);
                            real nor[3];
                            for (int c = 0; c < dim; c++) nor[c] = (side ? 1.0 : -1.0) * M.adjn(e, n1, d, c);   // outward
                            real w1[MAXEQ], f[MAXEQ];
                            for (int v = 0; v < neq; v++) w1[v] = went[gidx(M, v, e, n1)];
                            BdrLiftingFlux(P, bc, w1, nor, f, M.xat(e, n1), M.time);
                            for (int v = 0; v < neq; v++)
                                for (int c = 0; c < dim; c++) {
                                    grad[(size_t)c * size + gidx(M, v, e, n1)] += 1.0 / (M.w[0] * J) * (f[v] * nor[c]);
                                    if (gmag)
                                        gmag[(size_t)c * size + gidx(M, v, e, n1)] +=
                                            (std::fabs(f[v]) + std::fabs(w1[v])) * std::fabs(nor[c]) / (M.w[0] * std::fabs(J));
                                }
                        }
                }
        }
    }
    // ComputeGlobalPrimitiveGradVector: EntropyGrad2PrimGrad per direction
#pragma omp parallel for schedule(static)
    for (int e = 0; e < NE; e++) {
        real s[MAXEQ], g[MAXEQ];
        for (int node = 0; node < n; node++) {
            for (int v = 0; v < neq; v++) s[v] = u[gidx(M, v, e, node)];
            for (int c = 0; c < dim; c++) {
                for (int v = 0; v < neq; v++) g[v] = grad[(size_t)c * size + gidx(M, v, e, node)];
                EntropyGrad2PrimGrad(P, s, g);
                for (int v = 0; v < neq; v++) grad[(size_t)c * size + gidx(M, v, e, node)] = g[v];
                if (gmag) {
                    // the chain rule applied to the magnitudes, term by term in absolute value
                    real a[MAXEQ], o[MAXEQ];
                    for (int v = 0; v < neq; v++) a[v] = gmag[(size_t)c * size + gidx(M, v, e, node)];
                    const real rho = s[0];
                    real vel[3], vsq = 0.0;
                    for (int d = 0; d < dim; d++) {
                        vel[d] = s[1 + d] / rho;
                        vsq += vel[d] * vel[d];
                    }
                    const real p = P.gammaM1 * (s[neq - 1] - 0.5 * rho * vsq);
                    const real por = std::fabs(p / rho), ae = a[neq - 1];
                    real mdot = 0.0;
                    for (int d = 0; d < dim; d++) {
                        o[1 + d] = por * (a[1 + d] + std::fabs(vel[d]) * ae);
                        mdot += std::fabs(s[1 + d]) * o[1 + d];
                    }
                    o[0] = rho * a[0] + ae * (0.5 * rho * vsq + std::fabs(p) * P.gammaM1Inverse) +
                           std::fabs(rho / p) * mdot;
                    o[neq - 1] = por * (o[0] + std::fabs(p) * ae);
                    for (int v = 0; v < neq; v++) gmag[(size_t)c * size + gidx(M, v, e, node)] = o[v];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Full RHS -- ref: DGSEMOperator::Mult (src/Operators/DGSEMOperator.cpp:537-591) driving
// DGSEMNonlinearForm::Mult (Euler :647-849, NS 3D :1302-1545, NS 2D :1070-1300):
// zero dudt, element loop, interior-face loop.  Face term -- ref:
// DGSEMIntegrator::AssembleFaceVector (Euler :151-222, NS 3D :435-494, NS 2D :496-549).
// Returns lambda_max over all pair and face flux evaluations (:270,:188).
// ---------------------------------------------------------------------------
real RhsEval(const Mesh &M, const Phys &P, const real *u, real *dudt, real *grad_out, real *mag = nullptr,
               real *gmag = nullptr, const real *alpha = nullptr) {
    const int np = M.np, n = M.nn, dim = P.dim, neq = P.neq, NE = M.NE;
    const size_t size = (size_t)neq * NE * n;
    std::vector<real> grad_local, gmag_local;
    real *grad = grad_out;
    if (P.viscous) {
        if (!grad) {
            grad_local.resize(size * dim);
            grad = grad_local.data();
        }
        if (mag && !gmag) {
            gmag_local.resize(size * dim);
            gmag = gmag_local.data();
        }
        LiftingAndPrimGrad(M, P, u, grad, gmag);
    }
    std::fill(dudt, dudt + size, 0.0);
    if (mag) std::fill(mag, mag + size, 0.0);
    real lam = 0.0;
    // element loop
#pragma omp parallel for schedule(static) reduction(max : lam)
    for (int e = 0; e < NE; e++) {
        std::vector<real> el_u((size_t)n * neq), el_dudt((size_t)n * neq, 0.0), el_mag, Fvis;
        if (mag) el_mag.assign((size_t)n * neq, 0.0);
        gather(M, neq, u, e, el_u.data());
        std::vector<real> FvisMag;
        if (P.viscous) {
            // per-node viscous flux, mapped with adj -- :617-641
            Fvis.assign((size_t)3 * n * neq, 0.0);
            if (mag) FvisMag.assign((size_t)3 * n * neq, 0.0);
            real s[MAXEQ], dq[3][MAXEQ], flux[MAXEQ * 3];
            for (int i = 0; i < n; i++) {
                // A2: the reference reads state1 at the stale id1 (:620); under bug_compat
                // this is taken as node n-1 of the same element (blending on leaves id1 = n-1).
                const int sid = (P.bug & 2) ? n - 1 : i;
                for (int v = 0; v < neq; v++) s[v] = el_u[v * n + sid];
                for (int c = 0; c < dim; c++)
                    for (int v = 0; v < neq; v++) dq[c][v] = grad[(size_t)c * size + gidx(M, v, e, i)];
                ViscousFlux(P, s, dq[0], dq[1], dim > 2 ? dq[2] : nullptr, flux);
                // MultABt(adj1, flux_mat1, flux_mat): flux_mat(r, v) = sum_c adj(r,c) flux(v,c)
                for (int r = 0; r < dim; r++)
                    for (int v = 0; v < neq; v++) {
                        real acc = 0.0;
                        for (int c = 0; c < dim; c++) acc += M.adjn(e, i, r, c) * flux[v * dim + c];
                        Fvis[((size_t)r * n + i) * neq + v] = acc;
                    }
                if (mag) {
                    real gq[3][MAXEQ], fm[MAXEQ * 3];
                    for (int c = 0; c < dim; c++)
                        for (int v = 0; v < neq; v++) gq[c][v] = gmag[(size_t)c * size + gidx(M, v, e, i)];
                    ViscousFluxMag(P, s, gq[0], gq[1], dim > 2 ? gq[2] : nullptr, fm);
                    for (int r = 0; r < dim; r++)
                        for (int v = 0; v < neq; v++) {
                            real acc = 0.0;
                            for (int c = 0; c < dim; c++) acc += std::fabs(M.adjn(e, i, r, c)) * fm[v * dim + c];
                            FvisMag[((size_t)r * n + i) * neq + v] = acc;
                        }
                }
            }
        }
        if (alpha) {
            // SUBCELL_FV_BLENDING: the FV term first, into the zeroed element vector (:240-241)
            const real l = ElementSubcellFV(M, P, e, el_u.data(), alpha[e], el_dudt.data(),
                                              mag ? el_mag.data() : nullptr);
            lam = std::max(lam, l);
        }
        real l = ElementInviscid(M, P, e, el_u.data(), P.viscous ? Fvis.data() : nullptr, el_dudt.data(),
                                   mag ? el_mag.data() : nullptr, FvisMag.empty() ? nullptr : FvisMag.data(),
                                   alpha ? alpha + e : nullptr);
        lam = std::max(lam, l);
        for (int v = 0; v < neq; v++)
            for (int node = 0; node < n; node++) dudt[gidx(M, v, e, node)] += el_dudt[(size_t)v * n + node];
        if (mag)
            for (int v = 0; v < neq; v++)
                for (int node = 0; node < n; node++) mag[gidx(M, v, e, node)] = el_mag[(size_t)v * n + node];
    }
    // interior-face loop.  Aligned lattice: faces of one direction touch disjoint node sets
    // of each element, so a direction's faces run in parallel; directions run in turn.  General
    // connectivity: every face once, from its lower-numbered side (e1 < e2, or e1 == e2 and the
    // lower local face), in one serial pass (a face may meet its neighbour in any direction).
    const int Nt1 = dim > 1 ? np : 1, Nt2 = dim > 2 ? np : 1;
    const int npass = M.forient ? 1 : dim;
    for (int d = 0; d < npass; d++) {
        const int nf = M.forient ? 2 * dim : 1;
#pragma omp parallel for schedule(static) reduction(max : lam) if (!M.forient)
        for (int e1 = 0; e1 < NE; e1++)
            for (int fi = 0; fi < nf; fi++) {
            const int f1 = M.forient ? fi : 2 * d + 1;
            const int e2 = M.nbr[(size_t)e1 * 2 * dim + f1];
            if (e2 < 0) continue;   // a boundary face (below)
            const int f2 = M.nbface(e1, f1);
            if (M.forient && (e2 < e1 || (e2 == e1 && f2 < f1))) continue;
            const int d1 = f1 >> 1, sd1 = f1 & 1, d2 = f2 >> 1, sd2 = f2 & 1;
            real nor[3], s1[MAXEQ], s2[MAXEQ], fnum[MAXEQ], fl1[MAXEQ * 3], fl2[MAXEQ * 3];
            real dq[3][MAXEQ];
            for (int t2 = 0; t2 < Nt2; t2++)
                for (int t1 = 0; t1 < Nt1; t1++) {
                    int u1, u2;
                    M.nbpoint(e1, f1, t1, t2, u1, u2);
                    const int n1 = line_node(M, d1, sd1 ? np - 1 : 0, t1, t2);
                    const int n2 = line_node(M, d2, sd2 ? np - 1 : 0, u1, u2);
                    const real J1 = M.Jn(e1, n1), J2 = M.Jn(e2, n2);
                    // CalcOrtho of the face Jacobian: e1's adj row d1 at the face point, outward
                    for (int c = 0; c < dim; c++) nor[c] = (sd1 ? 1.0 : -1.0) * M.adjn(e1, n1, d1, c);
                    for (int v = 0; v < neq; v++) {
                        s1[v] = u[gidx(M, v, e1, n1)];
                        s2[v] = u[gidx(M, v, e2, n2)];
                    }
                    lam = std::max(lam, FaceFlux(P, s1, s2, nor, fnum));
                    real dU1[MAXEQ], dU2[MAXEQ], aU[MAXEQ];
                    for (int v = 0; v < neq; v++) {
                        dU1[v] = -fnum[v];
                        dU2[v] = fnum[v];
                        aU[v] = std::fabs(fnum[v]);
                    }
                    if (P.viscous) {
                        for (int c = 0; c < dim; c++)
                            for (int v = 0; v < neq; v++) dq[c][v] = grad[(size_t)c * size + gidx(M, v, e1, n1)];
                        ViscousFlux(P, s1, dq[0], dq[1], dim > 2 ? dq[2] : nullptr, fl1);
                        for (int c = 0; c < dim; c++)
                            for (int v = 0; v < neq; v++) dq[c][v] = grad[(size_t)c * size + gidx(M, v, e2, n2)];
                        ViscousFlux(P, s2, dq[0], dq[1], dim > 2 ? dq[2] : nullptr, fl2);
                        real fm1[MAXEQ * 3], fm2[MAXEQ * 3];
                        if (mag) {
                            for (int c = 0; c < dim; c++)
                                for (int v = 0; v < neq; v++) dq[c][v] = gmag[(size_t)c * size + gidx(M, v, e1, n1)];
                            ViscousFluxMag(P, s1, dq[0], dq[1], dim > 2 ? dq[2] : nullptr, fm1);
                            for (int c = 0; c < dim; c++)
                                for (int v = 0; v < neq; v++) dq[c][v] = gmag[(size_t)c * size + gidx(M, v, e2, n2)];
                            ViscousFluxMag(P, s2, dq[0], dq[1], dim > 2 ? dq[2] : nullptr, fm2);
                        }
                        for (int v = 0; v < neq; v++) {
                            real fv = 0.0;                   // (flux_mat1 + flux_mat2) * 0.5, Mult(nor)
                            real af = 0.0;
                            for (int c = 0; c < dim; c++) {
                                fv += ((fl1[v * dim + c] + fl2[v * dim + c]) * 0.5) * nor[c];
                                if (mag) af += 0.5 * (fm1[v * dim + c] + fm2[v * dim + c]) * std::fabs(nor[c]);
                            }
                            dU1[v] += fv;
                            dU2[v] -= fv;
                            aU[v] += af;
                        }
                    }
                    for (int v = 0; v < neq; v++) {
                        dudt[gidx(M, v, e1, n1)] += 1.0 / (M.w[0] * J1) * dU1[v];   // AddMult_a_VWt :219
                        dudt[gidx(M, v, e2, n2)] += 1.0 / (M.w[0] * J2) * dU2[v];   // :220
                        if (mag) {
                            mag[gidx(M, v, e1, n1)] += aU[v] / (M.w[0] * std::fabs(J1));
                            mag[gidx(M, v, e2, n2)] += aU[v] / (M.w[0] * std::fabs(J2));
                        }
                    }
                }
            }
    }
    // boundary faces -- BdrFaceIntegrator::AssembleFaceVector (Euler BdrFaceIntegrator.cpp:68-139,
    // NS :148-231): dU = -f(state1, nor) [+ F_v . nor], added as dU / (w0 J) at the face point.
    if (M.bc) {
#pragma omp parallel for schedule(static) reduction(max : lam)
        for (int e = 0; e < NE; e++) {
            for (int d = 0; d < dim; d++)
                for (int side = 0; side < 2; side++) {
                    const int nb = M.nbr[(size_t)e * 2 * dim + 2 * d + side];
                    if (nb >= 0) continue;
                    const orc_bc &bc = M.bc[-nb - 1];
                    for (int t2 = 0; t2 < Nt2; t2++)
                        for (int t1 = 0; t1 < Nt1; t1++) {
                            const int n1 = line_node(M, d, side ? np - 1 : 0, t1, t2);
                            const real J = M.Jn(e, n1);
                            real nor[3];
                            for (int c = 0; c < dim; c++) nor[c] = (side ? 1.0 : -1.0) * M.adjn(e, n1, d, c);
                            real s1[MAXEQ], f[MAXEQ], dU[MAXEQ], aU[MAXEQ];
                            for (int v = 0; v < neq; v++) s1[v] = u[gidx(M, v, e, n1)];
                            lam = std::max(lam, BdrInviscidFlux(P, bc, s1, nor, f, M.xat(e, n1), M.time));
                            for (int v = 0; v < neq; v++) {
                                dU[v] = -f[v];
                                aU[v] = std::fabs(f[v]);
                            }
                            if (P.viscous) {
                                real dq[3][MAXEQ], fv[MAXEQ];
                                for (int c = 0; c < dim; c++)
                                    for (int v = 0; v < neq; v++) dq[c][v] = grad[(size_t)c * size + gidx(M, v, e, n1)];
                                BdrViscousFlux(P, bc, s1, dq, nor, fv, M.xat(e, n1), M.time);
                                for (int v = 0; v < neq; v++) {
                                    dU[v] += fv[v];
                                    if (mag) {
                                        // rounding scale: the element-side viscous flux's scale (a bound for
                                        // each BC's flux of the trace and its gradients)
                                        real gq[3][MAXEQ], fm[MAXEQ * 3];
                                        for (int c = 0; c < dim; c++)
                                            for (int k = 0; k < neq; k++) gq[c][k] = gmag[(size_t)c * size + gidx(M, k, e, n1)];
                                        ViscousFluxMag(P, s1, gq[0], gq[1], dim > 2 ? gq[2] : nullptr, fm);
                                        real a = std::fabs(fv[v]);
                                        for (int c = 0; c < dim; c++) a += fm[v * dim + c] * std::fabs(nor[c]);
                                        aU[v] += a;
                                    }
                                }
                            }
                            for (int v = 0; v < neq; v++) {
                                dudt[gidx(M, v, e, n1)] += 1.0 / (M.w[0] * J) * dU[v];
                                if (mag) mag[gidx(M, v, e, n1)] += aU[v] / (M.w[0] * std::fabs(J));
                            }
                        }
                }
        }
    }
    return lam;
}

Phys MakePhys(int dim, real gamma, real Pr, real mu, real R_gas, int viscous, int sutherland,
              int surface_flux, int bug) {
    Phys P;
    P.dim = dim;
    P.neq = dim + 2;
    P.gamma = gamma;
    P.gammaM1 = gamma - 1.0;
    P.gammaM1Inverse = 1.0 / P.gammaM1;
    P.Pr = Pr;
    P.PrInverse = 1.0 / Pr;
    P.mu = mu;
    P.R_gas = R_gas;
    P.viscous = viscous;
    P.sutherland = sutherland;
    P.surface_flux = surface_flux;
    P.bug = bug;
    return P;
}

void MakeMesh(Mesh &M, int dim, int order, int NE, const int *nbr, const real *geom) {
    M.dim = dim;
    M.np = order + 1;
    M.nn = 1;
    for (int d = 0; d < dim; d++) M.nn *= M.np;
    M.NE = NE;
    M.nbr = nbr;
    M.geom = geom;
    GLL(M.np, M.x, M.w);
    DMatrices(M.np, M.x, M.w, M.D, M.Dhat, M.Dhat2);
}

}  // namespace

#ifndef ORACLE_NO_C_API
// ===========================================================================
// C interface (ctypes)
// ===========================================================================
extern "C" {


static Phys phys_of(int dim, const orc_phys *p) {
    return MakePhys(dim, p->gamma, p->Pr, p->mu, p->R_gas, p->viscous, p->sutherland, p->surface_flux,
                    p->bug_compat);
}

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int orc_gll(int np, double *x, double *w) {
    if (np < 2 || np > MAXNP) return -1;
    GLL(np, x, w);
    return 0;
}

int orc_dmatrices(int np, double *D, double *Dhat, double *Dhat2) {
    if (np < 2 || np > MAXNP) return -1;
    double x[MAXNP], w[MAXNP];
    GLL(np, x, w);
    DMatrices(np, x, w, D, Dhat, Dhat2);
    return 0;
}

double orc_logmean(double x, double y) { return LogMean(x, y); }

double orc_volume_flux(int dim, const orc_phys *p, const double *s1, const double *s2, const double *m1,
                       const double *m2, double *F) {
    Phys P = phys_of(dim, p);
    return VolumeFlux(P, s1, s2, m1, m2, F);
}

double orc_face_flux(int dim, const orc_phys *p, const double *s1, const double *s2, const double *nor,
                     double *F) {
    Phys P = phys_of(dim, p);
    return FaceFlux(P, s1, s2, nor, F);
}

double orc_euler_flux_dot_n(int dim, const orc_phys *p, const double *s, const double *nor, double *F) {
    Phys P = phys_of(dim, p);
    return EulerFluxDotN(P, s, nor, F);
}

void orc_conserv2entropy(int dim, const orc_phys *p, const double *s, double *w) {
    Phys P = phys_of(dim, p);
    Conserv2Entropy(P, s, w);
}

void orc_entropy2conserv(int dim, const orc_phys *p, const double *w, double *s) {
    Phys P = phys_of(dim, p);
    Entropy2Conserv(P, w, s);
}

void orc_entgrad2primgrad(int dim, const orc_phys *p, const double *s, double *g) {
    Phys P = phys_of(dim, p);
    EntropyGrad2PrimGrad(P, s, g);
}

void orc_viscous_flux(int dim, const orc_phys *p, const double *s, const double *gx, const double *gy,
                      const double *gz, double *flux) {
    Phys P = phys_of(dim, p);
    ViscousFlux(P, s, gx, gy, gz, flux);
}

// RHS of the semi-discretisation: dudt = L(u).  grad (optional, dim*size) receives
// the primitive gradients (NS only).  lambda_max (optional) receives the max wave speed.
int orc_rhs_mag(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p,
                const double *u, double *dudt, double *grad, double *lambda_max, double *mag, double *gmag);

int orc_rhs(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p,
            const double *u, double *dudt, double *grad, double *lambda_max) {
    return orc_rhs_mag(dim, order, NE, nbr, geom, p, u, dudt, grad, lambda_max, nullptr, nullptr);
}

// As orc_rhs, plus the rounding scale of every entry: mag (size of dudt) holds the same
// assembly with every term in absolute value, gmag (size of grad) the same for the lifted
// gradients and their chain rule.  |a - b| <= C eps mag bounds the difference of two
// evaluations of the same sums in different order or contraction.
int orc_rhs_mag(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p,
                const double *u, double *dudt, double *grad, double *lambda_max, double *mag, double *gmag) {
    if (dim < 2 || dim > 3 || order < 1 || order + 1 > MAXNP) return -1;
    Mesh M;
    MakeMesh(M, dim, order, NE, nbr, geom);
    Phys P = phys_of(dim, p);
    std::vector<double> gm_local;
    if (gmag && !grad) return -3;
    double lam = RhsEval(M, P, u, dudt, grad, mag, gmag);
    if (lambda_max) *lambda_max = lam;
    return 0;
}

// One boundary face point: the outer state, the inviscid flux (returns the wave speed), the
// viscous flux (dq: dim x neq primitive gradients) and the lifting flux (w1: entropy variables).
int orc_bc_outer_state(int dim, const orc_phys *p, const orc_bc *bc, const double *s1, const double *nor,
                       double *s2) {
    Phys P = phys_of(dim, p);
    OuterState(P, *bc, s1, nor, s2, nullptr, 0.0);
    return 0;
}

double orc_bc_inviscid_flux(int dim, const orc_phys *p, const orc_bc *bc, const double *s1, const double *nor,
                            double *f) {
    Phys P = phys_of(dim, p);
    return BdrInviscidFlux(P, *bc, s1, nor, f, nullptr, 0.0);
}

void orc_bc_viscous_flux(int dim, const orc_phys *p, const orc_bc *bc, const double *s1, const double *dq,
                         const double *nor, double *fv) {
    Phys P = phys_of(dim, p);
    double g[3][MAXEQ] = {};
    for (int c = 0; c < dim; c++)
        for (int v = 0; v < P.neq; v++) g[c][v] = dq[c * P.neq + v];
    BdrViscousFlux(P, *bc, s1, g, nor, fv, nullptr, 0.0);
}

void orc_bc_lifting_flux(int dim, const orc_phys *p, const orc_bc *bc, const double *w1, const double *nor,
                         double *fl) {
    Phys P = phys_of(dim, p);
    BdrLiftingFlux(P, *bc, w1, nor, fl, nullptr, 0.0);
}

// the boundary context of the following evaluations (orc_set_boundary_context)
static const double *g_bxyz = nullptr;
static double g_btime = 0.0;

void orc_set_boundary_context(const double *xyz, double t) {
    g_bxyz = xyz;
    g_btime = t;
}
double orc_get_time(void) { return g_btime; }
void orc_set_time(double t) { g_btime = t; }

int orc_bc_state(int dim, const orc_phys *p, const orc_bc *bc, const double *x, double t, double *s2) {
    if (dim < 2 || dim > 3 || !bc || !s2) return -1;
    if (bc->profile == PROFILE_FRONT && !x) return -1;
    BcState(dim, p->gamma, *bc, x, t, s2);
    return 0;
}

// The general RHS: boundary conditions bcs[0..nbc-1] for the faces whose neighbour entry is
// -(k+1), blending when alpha is non-null, rounding scales when mag (gmag) are non-null.
int orc_rhs_gen(int dim, int order, int NE, const int *nbr, const double *geom, const double *ngeom,
                const int *forient, const orc_phys *p, int nbc, const orc_bc *bcs, const double *u,
                const double *alpha, double *dudt, double *grad, double *lambda_max, double *mag, double *gmag);

int orc_rhs_ex(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, int nbc,
               const orc_bc *bcs, const double *u, const double *alpha, double *dudt, double *grad,
               double *lambda_max, double *mag, double *gmag) {
    return orc_rhs_gen(dim, order, NE, nbr, geom, nullptr, nullptr, p, nbc, bcs, u, alpha, dudt, grad, lambda_max,
                       mag, gmag);
}

// orc_rhs_ex on a general mesh: ngeom (per node (det J, adj J), [e][node][1 + dim^2]) replaces
// the per-element geom when non-null; forient (per face [e][2 dim]: the neighbour's local face
// 2 d + s in bits 0-2, bit 3 flips the first face coordinate, bit 4 the second, bit 5 swaps
// them) gives the face matching when non-null, else faces are those of an aligned lattice.
int orc_rhs_gen(int dim, int order, int NE, const int *nbr, const double *geom, const double *ngeom,
                const int *forient, const orc_phys *p, int nbc, const orc_bc *bcs, const double *u,
                const double *alpha, double *dudt, double *grad, double *lambda_max, double *mag, double *gmag) {
    if (dim < 2 || dim > 3 || order < 1 || order + 1 > MAXNP) return -1;
    if (!geom && !ngeom) return -1;
    Mesh M;
    MakeMesh(M, dim, order, NE, nbr, geom);
    M.ngeom = ngeom;
    M.forient = forient;
    Phys P = phys_of(dim, p);
    for (size_t i = 0; i < (size_t)NE * 2 * dim; i++)
        if (nbr[i] < 0 && (-nbr[i] - 1 >= nbc || !bcs)) return -4;   // a boundary face without its condition
    M.bc = nbc > 0 ? bcs : nullptr;
    M.nbc = nbc;
    M.xyz = g_bxyz;
    M.time = g_btime;
    for (int k = 0; k < nbc; k++)
        if (bcs[k].profile == PROFILE_FRONT && !M.xyz) return -5;   // a front without node coordinates
    if (gmag && !grad) return -3;
    double lam = RhsEval(M, P, u, dudt, grad, mag, gmag, alpha);
    if (lambda_max) *lambda_max = lam;
    return 0;
}

// As orc_rhs_mag with subcell finite-volume blending (the reference built with
// SUBCELL_FV_BLENDING): alpha holds one blending coefficient per element, as
// DGSEMOperator::ComputeBlendingCoefficient leaves it in the alpha grid function
// (orc_blend_coef computes it).
int orc_rhs_blend_mag(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p,
                      const double *u, const double *alpha, double *dudt, double *grad, double *lambda_max,
                      double *mag, double *gmag) {
    if (dim < 2 || dim > 3 || order < 1 || order + 1 > MAXNP || !alpha) return -1;
    Mesh M;
    MakeMesh(M, dim, order, NE, nbr, geom);
    Phys P = phys_of(dim, p);
    if (gmag && !grad) return -3;
    double lam = RhsEval(M, P, u, dudt, grad, mag, gmag, alpha);
    if (lambda_max) *lambda_max = lam;
    return 0;
}

// Time integration with a fixed dt.  scheme 0: SSP-RK3 (MFEM RK3SSPSolver,
// Shu-Osher form; the reference calls it at Simulation.cpp:1010); scheme 1:
// LSRK4(5) 2N-storage (Carpenter & Kennedy 1994).  u is advanced in place.
static const double CK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double CK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};

int orc_rk_run(int dim, int order, int NE, const int *nbr, const double *geom, const orc_phys *p, int scheme,
               int nsteps, double dt, double *u) {
    if (dim < 2 || dim > 3 || order < 1 || order + 1 > MAXNP) return -1;
    Mesh M;
    MakeMesh(M, dim, order, NE, nbr, geom);
    Phys P = phys_of(dim, p);
    const size_t size = (size_t)P.neq * NE * M.nn;
    std::vector<double> k(size), y(size), x0(size);
    for (int s = 0; s < nsteps; s++) {
        if (scheme == 0) {
            RhsEval(M, P, u, k.data(), nullptr);
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < size; i++) y[i] = u[i] + dt * k[i];
            RhsEval(M, P, y.data(), k.data(), nullptr);
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < size; i++) {
                y[i] = y[i] + dt * k[i];
                y[i] = 3. / 4 * u[i] + 1. / 4 * y[i];
            }
            RhsEval(M, P, y.data(), k.data(), nullptr);
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < size; i++) {
                y[i] = y[i] + dt * k[i];
                u[i] = 1. / 3 * u[i] + 2. / 3 * y[i];
            }
        } else if (scheme == 1) {
            std::vector<double> &kk = x0;  // 2N register
            std::fill(kk.begin(), kk.end(), 0.0);
            for (int st = 0; st < 5; st++) {
                RhsEval(M, P, u, k.data(), nullptr);
#pragma omp parallel for schedule(static)
                for (size_t i = 0; i < size; i++) {
                    kk[i] = CK_A[st] * kk[i] + dt * k[i];
                    u[i] = u[i] + CK_B[st] * kk[i];
                }
            }
        } else {
            return -2;
        }
    }
    return 0;
}

}  // extern "C"
#endif  // ORACLE_NO_C_API
