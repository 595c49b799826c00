// dg_adjoint.hpp -- reverse-mode derivatives (vector-Jacobian products) of the pointwise
// Euler fluxes, for the discrete adjoint of the DGSEM operator.
//
// Each *_vjp function recomputes the forward quantities of its flux from the conserved
// states and accumulates  s_b += (dF/ds)^T F_b  for the given cotangent F_b.  The volume
// kernel splits the same steps: node_q once per node, ec_core_back once per pair (adjoints of
// density, velocity and beta), prim_back once per node.  The forward
// formulas are those of dg_physics.hpp (the reference's ChandrashekarFlux.cpp:11-161 and
// ComputeLogMean, BasicOperations.cpp:6-12); the derivatives are of the same mathematics,
// with the logarithmic mean differentiated through the same two branches (series for
// u = ((y-x)/(y+x))^2 < 1e-4, closed form otherwise).  The wave speed max(|V|+c) of the
// surface dissipation is differentiated through the branch the forward takes (both halves at
// a tie).
//
// The functions are host+device so the same code is checked against finite differences on
// the CPU (tests/native/adjoint_vjp_host.cpp).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else  // host-only build of the same functions (the finite-difference check in tests/native)
#define __host__
#define __device__
#endif

#include <cmath>

namespace prandtl {
namespace adj {

__host__ __device__ inline double sq(double x) { return x * x; }

// reciprocal: on the device the v_rcp_f64 seed with two Newton steps (within an ulp or two of
// the division); on the host the division
__host__ __device__ inline double inv(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

// per-node quantities the volume adjoint reads for every pair: density, velocity, beta =
// rho/(2p) and their reciprocals (computed once per node, not once per pair)
template <int DIM>
struct NodeQ {
    double rho, v[DIM], beta, irho, ibeta;
};

template <int DIM>
__host__ __device__ inline NodeQ<DIM> node_q(const double *s, double gm1) {
    NodeQ<DIM> q;
    q.rho = s[0];
    q.irho = inv(q.rho);
    double vsq = 0.0;
    for (int d = 0; d < DIM; d++) {
        q.v[d] = s[1 + d] * q.irho;
        vsq += q.v[d] * q.v[d];
    }
    const double p = gm1 * (s[DIM + 1] - 0.5 * q.rho * vsq);
    q.beta = 0.5 * q.rho * inv(p);
    q.ibeta = inv(q.beta);
    return q;
}

// logarithmic mean L of (x, y), its reciprocal and its partials, given 1/x and 1/y; the three
// branches of the forward, in u = ((y-x)/(y+x))^2: the series P = 1 + u/3 + u^2/5 + u^3/7 below
// 1e-4, the series to u^11 below kLogMeanSeriesU = 0.05, the closed form with ln(y/x) beyond.  On
// the series branches L = (x+y)/(2P), so 1/L = 2P/(x+y) needs no division; below 1e-4, 1/P is
// itself the series 1 - u/3 - 4u^2/45 - 44u^3/945 (the omitted u^4 term is below 1e-17 relative).
// With u = f^2, f = (y-x)/(x+y): dL/dx = 1/(2P) + c y, dL/dy = 1/(2P) - c x,
// c = 2 P'(u) f / (P^2 (x+y)).  The closed form's partials (y-x)/x - ln(y/x) over ln(y/x)^2 cancel
// to O(u) in the numerator, so in the series range they would lose digits that the series keeps.
constexpr double kLogMeanSeriesU = 0.05;   // the forward's switch to the closed form
__host__ __device__ inline void log_mean_dq(double x, double y, double ix, double iy, double &L, double &iL,
                                            double &Lx, double &Ly) {
    const double s = x + y, r = inv(s), f = (y - x) * r, u = f * f;
    if (u < kLogMeanSeriesU) {
        double P, dP, iP;
        if (u < 1e-4) {
            P = 1.0 + u * (1.0 / 3.0 + u * (1.0 / 5.0 + u * (1.0 / 7.0)));
            dP = 1.0 / 3.0 + u * (2.0 / 5.0 + u * (3.0 / 7.0));
            iP = 1.0 - u * (1.0 / 3.0 + u * (4.0 / 45.0 + u * (44.0 / 945.0)));
        } else {
            // P = 1 + sum_{k=1..11} u^k/(2k+1), P' = sum_{k=1..11} k u^(k-1)/(2k+1)
            double p = 1.0 / 23.0, dp = 11.0 / 23.0;
            for (int k = 10; k >= 1; k--) {
                p = p * u + 1.0 / (2 * k + 1);
                dp = dp * u + (double)k / (2 * k + 1);
            }
            P = 1.0 + u * p;
            dP = dp;
            iP = inv(P);
        }
        L = 0.5 * s * iP;
        iL = 2.0 * r * P;
        const double c = 2.0 * dP * iP * iP * f * r;
        Lx = 0.5 * iP + c * y;
        Ly = 0.5 * iP - c * x;
    } else {
        const double lg = log(y * ix), ilg = inv(lg), d = y - x;
        L = d * ilg;
        iL = lg * inv(d);
        Lx = (d * ix - lg) * ilg * ilg;
        Ly = (lg - d * iy) * ilg * ilg;
    }
}

// adjoints of a node's (rho, v, beta)
template <int DIM>
struct PrimB {
    double rho, v[DIM], beta;
};

template <int DIM>
__host__ __device__ inline void primb_zero(PrimB<DIM> &g) {
    g.rho = g.beta = 0.0;
    for (int d = 0; d < DIM; d++) g.v[d] = 0.0;
}

// core of the entropy-conservative two-point flux
//   F = (rho_ln Vn, rho_ln Vn vbar + p_hat m, rho_ln Vn h_hat),
//   Vn = vbar . m, p_hat = (rho_1 + rho_2) / (2 (beta_1 + beta_2)),
//   h_hat = 1/(2 beta_ln (gamma-1)) - |v_1|^2/4 - |v_2|^2/4 + |vbar|^2 + p_hat / rho_ln
//         = 1/(2 beta_ln (gamma-1)) + (v_1 . v_2)/2 + p_hat / rho_ln,
// reversed: accumulates the adjoints of (rho, v, beta) of both nodes for the cotangent Fb.
template <int DIM>
__host__ __device__ inline void ec_core_back(const NodeQ<DIM> &a, const NodeQ<DIM> &b, const double *m,
                                             double gm1inv, const double *Fb, PrimB<DIM> &ga, PrimB<DIM> &gb) {
    double Lr, iLr, Lr1, Lr2, Lb, iLb, Lb1, Lb2;
    log_mean_dq(a.rho, b.rho, a.irho, b.irho, Lr, iLr, Lr1, Lr2);
    log_mean_dq(a.beta, b.beta, a.ibeta, b.ibeta, Lb, iLb, Lb1, Lb2);
    double vs[DIM], Vn2 = 0.0, vdot = 0.0;   // vs = v_1 + v_2 = 2 vbar
    for (int d = 0; d < DIM; d++) {
        vs[d] = a.v[d] + b.v[d];
        Vn2 += vs[d] * m[d];
        vdot += a.v[d] * b.v[d];
    }
    const double Vn = 0.5 * Vn2;
    const double ibs = inv(a.beta + b.beta);
    const double p_hat = 0.5 * (a.rho + b.rho) * ibs;
    const double h = 0.5 * (vdot + iLb * gm1inv) + p_hat * iLr;
    const double mflux = Lr * Vn;
    // reverse
    double mflux_b = Fb[0] + Fb[DIM + 1] * h;
    double p_hat_b = 0.0;
    for (int d = 0; d < DIM; d++) {
        mflux_b += 0.5 * Fb[1 + d] * vs[d];
        p_hat_b += Fb[1 + d] * m[d];
    }
    const double h_b = Fb[DIM + 1] * mflux;
    const double Lb_b = h_b * (-0.5 * gm1inv * iLb * iLb);
    p_hat_b += h_b * iLr;
    const double Lr_b = h_b * (-p_hat * iLr * iLr) + mflux_b * Vn;
    const double Vn_b = mflux_b * Lr;
    const double hh = 0.5 * h_b;
    for (int d = 0; d < DIM; d++) {
        // through vbar (Vn and the momentum flux), and h_hat's (v_1 . v_2)/2
        const double vb = 0.5 * (Vn_b * m[d] + Fb[1 + d] * mflux);
        ga.v[d] += hh * b.v[d] + vb;
        gb.v[d] += hh * a.v[d] + vb;
    }
    const double rb = p_hat_b * 0.5 * ibs, bb = -p_hat_b * p_hat * ibs;
    ga.rho += rb + Lr_b * Lr1;
    gb.rho += rb + Lr_b * Lr2;
    ga.beta += bb + Lb_b * Lb1;
    gb.beta += bb + Lb_b * Lb2;
}

// adjoints of a node's (rho, v, beta) -> adjoint of its conserved state (accumulated):
// v = m/rho, p = (gamma-1)(E - rho|v|^2/2), beta = rho/(2p)
template <int DIM>
__host__ __device__ inline void prim_back(const NodeQ<DIM> &q, double gm1, const PrimB<DIM> &g, double *s_b) {
    // beta = rho/(2p): d/drho = 1/(2p) = beta/rho, d/dp = -beta/p = -2 beta^2/rho
    double rho_b = g.rho + g.beta * q.beta * q.irho;
    const double pg = -g.beta * 2.0 * q.beta * q.beta * q.irho * gm1;   // adjoint of p, times gm1
    s_b[DIM + 1] += pg;
    double vsq = 0.0;
    for (int d = 0; d < DIM; d++) {
        vsq += q.v[d] * q.v[d];
        const double vb = g.v[d] - pg * q.rho * q.v[d];
        s_b[1 + d] += vb * q.irho;
        rho_b -= vb * q.v[d] * q.irho;
    }
    rho_b -= 0.5 * pg * vsq;
    s_b[0] += rho_b;
}

// (dF#/ds1)^T Fb and (dF#/ds2)^T Fb of the two-point volume flux with metric m
template <int DIM>
__host__ __device__ inline void volume_flux_vjp(const double *s1, const double *s2, const double *m, double gm1,
                                                const double *Fb, double *s1b, double *s2b) {
    const NodeQ<DIM> a = node_q<DIM>(s1, gm1), b = node_q<DIM>(s2, gm1);
    PrimB<DIM> ga, gb;
    primb_zero<DIM>(ga);
    primb_zero<DIM>(gb);
    ec_core_back<DIM>(a, b, m, 1.0 / gm1, Fb, ga, gb);
    prim_back<DIM>(a, gm1, ga, s1b);
    prim_back<DIM>(b, gm1, gb, s2b);
}

// (df*/ds1)^T fb and (df*/ds2)^T fb of the reference's Chandrashekar surface flux
// (ChandrashekarFlux.cpp:85-161); bug3 follows the :131 typo.  kind 1 = the EC flux.
template <int DIM>
__host__ __device__ inline void face_flux_vjp(int kind, const double *s1, const double *s2, const double *n,
                                              double gamma, bool bug3, const double *fb, double *s1b, double *s2b) {
    const double gm1 = gamma - 1.0, gm1inv = 1.0 / gm1;
    const NodeQ<DIM> a = node_q<DIM>(s1, gm1), b = node_q<DIM>(s2, gm1);
    PrimB<DIM> ga, gb;
    primb_zero<DIM>(ga);
    primb_zero<DIM>(gb);
    ec_core_back<DIM>(a, b, n, gm1inv, fb, ga, gb);
    if (kind == 0) {
        // dissipation  -1/2 lam |n| [drho, dm, diss]
        double nmag = 0.0;
        for (int d = 0; d < DIM; d++) nmag += n[d] * n[d];
        nmag = sqrt(nmag);
        const double hl = 0.5 * nmag;
        double Lb, iLb, Lb1, Lb2;
        log_mean_dq(a.beta, b.beta, a.ibeta, b.ibeta, Lb, iLb, Lb1, Lb2);
        const double rho_mean = 0.5 * (a.rho + b.rho), drho = b.rho - a.rho;
        double diss = 0.0, vsq1 = 0.0, vsq2 = 0.0;
        for (int d = 0; d < DIM; d++) {
            const double vm = 0.5 * (a.v[d] + b.v[d]), dv = b.v[d] - a.v[d];
            if (DIM == 3 && d == 2 && bug3)
                diss += 0.5 * drho * a.v[d] * b.v[d] + rho_mean * dv + vm;
            else
                diss += 0.5 * drho * a.v[d] * b.v[d] + rho_mean * dv * vm;
            vsq1 += a.v[d] * a.v[d];
            vsq2 += b.v[d] * b.v[d];
        }
        diss += 0.5 * drho * gm1inv * iLb + 0.5 * rho_mean * gm1inv * (b.ibeta - a.ibeta);
        // p = rho / (2 beta), c = sqrt(gamma p / rho) = sqrt(gamma / (2 beta))
        const double c1 = sqrt(0.5 * gamma * a.ibeta), c2 = sqrt(0.5 * gamma * b.ibeta);
        const double vn1 = sqrt(vsq1), vn2 = sqrt(vsq2);
        const double l1 = vn1 + c1, l2 = vn2 + c2;
        const double lam = l1 >= l2 ? l1 : l2;
        // reverse of the dissipation terms
        double lam_b = -hl * fb[0] * drho;
        double drho_b = -hl * lam * fb[0];
        for (int d = 0; d < DIM; d++) {
            const double dm = s2[1 + d] - s1[1 + d];
            lam_b += -hl * fb[1 + d] * dm;
            s2b[1 + d] += -hl * lam * fb[1 + d];
            s1b[1 + d] += hl * lam * fb[1 + d];
        }
        lam_b += -hl * fb[DIM + 1] * diss;
        const double diss_b = -hl * lam * fb[DIM + 1];
        double rho_mean_b = 0.0;
        for (int d = 0; d < DIM; d++) {
            const double vm = 0.5 * (a.v[d] + b.v[d]), dv = b.v[d] - a.v[d];
            drho_b += diss_b * 0.5 * a.v[d] * b.v[d];
            ga.v[d] += diss_b * 0.5 * drho * b.v[d];
            gb.v[d] += diss_b * 0.5 * drho * a.v[d];
            double dv_b, vm_b;
            if (DIM == 3 && d == 2 && bug3) {
                rho_mean_b += diss_b * dv;
                dv_b = diss_b * rho_mean;
                vm_b = diss_b;
            } else {
                rho_mean_b += diss_b * dv * vm;
                dv_b = diss_b * rho_mean * vm;
                vm_b = diss_b * rho_mean * dv;
            }
            gb.v[d] += dv_b + 0.5 * vm_b;
            ga.v[d] += -dv_b + 0.5 * vm_b;
        }
        drho_b += diss_b * 0.5 * gm1inv * iLb;
        const double Lb_b = diss_b * (-0.5 * drho * gm1inv * iLb * iLb);
        ga.beta += Lb_b * Lb1;
        gb.beta += Lb_b * Lb2;
        rho_mean_b += diss_b * 0.5 * gm1inv * (b.ibeta - a.ibeta);
        gb.beta += diss_b * 0.5 * rho_mean * gm1inv * (-b.ibeta * b.ibeta);
        ga.beta += diss_b * 0.5 * rho_mean * gm1inv * (a.ibeta * a.ibeta);
        ga.rho += 0.5 * rho_mean_b - drho_b;
        gb.rho += 0.5 * rho_mean_b + drho_b;
        // lam = max(|V_1| + c_1, |V_2| + c_2) through the branch taken; at a tie the mean of the
        // two one-sided derivatives (what central differences of the forward measure)
        const double l1b = l1 > l2 ? lam_b : (l1 < l2 ? 0.0 : 0.5 * lam_b);
        const double l2b = lam_b - l1b;
        if (vn1 > 0.0) {
            const double g1 = l1b * inv(vn1);
            for (int d = 0; d < DIM; d++) ga.v[d] += g1 * a.v[d];
        }
        if (vn2 > 0.0) {
            const double g2 = l2b * inv(vn2);
            for (int d = 0; d < DIM; d++) gb.v[d] += g2 * b.v[d];
        }
        // dc/dbeta = -c / (2 beta)
        ga.beta += -l1b * c1 * 0.5 * a.ibeta;
        gb.beta += -l2b * c2 * 0.5 * b.ibeta;
    }
    prim_back<DIM>(a, gm1, ga, s1b);
    prim_back<DIM>(b, gm1, gb, s2b);
}

// ---------------------------------------------------------------------------
// Navier-Stokes pointwise adjoints (constant viscosity).  Forward formulas: conserv_to_entropy
// (BasicOperations.cpp:81-104), entgrad_to_primgrad (:126-170), viscous_flux
// (NavierStokesFlux.cpp:12-90), as in dg_physics.hpp.  Each accumulates s_b (and g_b).
// ---------------------------------------------------------------------------

// w = (dU/du)^-1: w0 = (gamma - s)/(gamma-1) - beta q, w_d = beta v_d, w_E = -beta,
// s = ln p - gamma ln rho, beta = rho/p, q = |v|^2/2.  Given w_b, accumulate s_b.
template <int DIM>
__host__ __device__ inline void entropy_vars_vjp(const double *s, double gamma, const double *w_b, double *s_b) {
    const double gm1 = gamma - 1.0, rho = s[0], irho = inv(rho);
    double v[DIM], q = 0.0;
    for (int d = 0; d < DIM; d++) {
        v[d] = s[1 + d] * irho;
        q += v[d] * v[d];
    }
    q *= 0.5;
    const double p = gm1 * (s[DIM + 1] - rho * q), ip = inv(p);
    const double beta = rho * ip;
    double beta_b = -w_b[DIM + 1] - w_b[0] * q, v_b[DIM];
    for (int d = 0; d < DIM; d++) {
        beta_b += w_b[1 + d] * v[d];
        v_b[d] = w_b[1 + d] * beta;
    }
    double q_b = -w_b[0] * beta;
    const double ent_b = -w_b[0] / gm1;
    double p_b = ent_b * ip - beta_b * beta * ip;
    double rho_b = -gamma * ent_b * irho + beta_b * ip;
    // p = (gamma-1)(E - rho q)
    s_b[DIM + 1] += gm1 * p_b;
    rho_b += -gm1 * p_b * q;
    q_b += -gm1 * p_b * rho;
    for (int d = 0; d < DIM; d++) {
        v_b[d] += q_b * v[d];
        s_b[1 + d] += v_b[d] * irho;
        rho_b += -v_b[d] * v[d] * irho;
    }
    s_b[0] += rho_b;
}

// EntropyGrad2PrimGrad of one direction: gp = T(s, g) (g: gradient of w, gp: gradient of
// (rho, v, p)); given gp_b, accumulate s_b and g_b.
template <int DIM>
__host__ __device__ inline void entgrad_vjp(const double *s, const double *g, double gamma, const double *gp_b,
                                            double *s_b, double *g_b) {
    const double gm1 = gamma - 1.0, gm1inv = 1.0 / gm1;
    const double rho = s[0], irho = inv(rho);
    double vel[DIM], vsq = 0.0;
    for (int d = 0; d < DIM; d++) {
        vel[d] = s[1 + d] * irho;
        vsq += vel[d] * vel[d];
    }
    vsq *= 0.5;   // v_sq of the forward
    const double p = gm1 * (s[DIM + 1] - rho * vsq);
    const double KE = vsq * rho, ge = g[DIM + 1], por = p * irho, ipor = inv(por);
    double outd[DIM], mdot = 0.0;
    for (int d = 0; d < DIM; d++) {
        outd[d] = por * (g[1 + d] + vel[d] * ge);
        mdot += s[1 + d] * outd[d];
    }
    const double out0 = rho * g[0] - ge * (KE - p * gm1inv) + ipor * mdot;
    // reverse
    const double outE_b = gp_b[DIM + 1];
    double por_b = outE_b * (out0 + p * ge);
    const double out0_b = gp_b[0] + outE_b * por;
    double p_b = outE_b * por * ge;
    double ge_b = outE_b * por * p;
    double rho_b = out0_b * g[0];
    g_b[0] += out0_b * rho;
    ge_b += -out0_b * (KE - p * gm1inv);
    const double KE_b = -out0_b * ge;
    p_b += out0_b * ge * gm1inv;
    const double ipor_b = out0_b * mdot;
    const double mdot_b = out0_b * ipor;
    double vel_b[DIM], m_b[DIM];
    for (int d = 0; d < DIM; d++) {
        m_b[d] = mdot_b * outd[d];
        const double od_b = gp_b[1 + d] + mdot_b * s[1 + d];
        por_b += od_b * (g[1 + d] + vel[d] * ge);
        g_b[1 + d] += od_b * por;
        vel_b[d] = od_b * por * ge;
        ge_b += od_b * por * vel[d];
    }
    por_b += -ipor_b * ipor * ipor;
    p_b += por_b * irho;
    rho_b += -por_b * p * irho * irho;
    double vsq_b = KE_b * rho;
    rho_b += KE_b * vsq;
    s_b[DIM + 1] += gm1 * p_b;
    rho_b += -gm1 * p_b * vsq;
    vsq_b += -gm1 * p_b * rho;
    g_b[DIM + 1] += ge_b;
    for (int d = 0; d < DIM; d++) {
        vel_b[d] += vsq_b * vel[d];
        m_b[d] += vel_b[d] * irho;
        rho_b += -vel_b[d] * vel[d] * irho;
        s_b[1 + d] += m_b[d];
    }
    s_b[0] += rho_b;
}

// NavierStokesFlux::ComputeViscousFlux with constant mu: flux[r][c] (r < DIM momentum, r = DIM
// energy) from the state and the primitive gradients g[c][(rho, v, p)].  Given flux_b,
// accumulate s_b and g_b.
template <int DIM>
__host__ __device__ inline void viscous_flux_vjp(const double *s, const double (*g)[DIM + 2], double gamma, double mu,
                                                 double Pr_inv, double mu_bulk, const double (*flux_b)[DIM],
                                                 double *s_b, double (*g_b)[DIM + 2]) {
    const double gm1 = gamma - 1.0, gm1inv = 1.0 / gm1;
    const double rho = s[0], irho = inv(rho);
    double vel[DIM], vsq = 0.0;
    for (int d = 0; d < DIM; d++) {
        vel[d] = s[1 + d] * irho;
        vsq += vel[d] * vel[d];
    }
    const double p = gm1 * (s[DIM + 1] - 0.5 * rho * vsq);
    const double lambda = mu * gamma * Pr_inv, por = p * irho;
    double div = g[0][1];
    for (int d = 1; d < DIM; d++) div += g[d][1 + d];
    double vel_b[DIM], div_b = 0.0, por_b = 0.0, irho_b = 0.0;
    for (int d = 0; d < DIM; d++) vel_b[d] = 0.0;
    for (int c = 0; c < DIM; c++) {
        const double e_b = flux_b[DIM][c];
        const double cv_b = lambda * e_b;
        for (int r = 0; r < DIM; r++) {
            const double t = (r == c) ? mu * (2.0 * g[c][1 + r] - mu_bulk * div) : mu * (g[c][1 + r] + g[r][1 + c]);
            const double t_b = flux_b[r][c] + e_b * vel[r];
            vel_b[r] += e_b * t;
            if (r == c) {
                g_b[c][1 + r] += 2.0 * mu * t_b;
                div_b += -mu * mu_bulk * t_b;
            } else {
                g_b[c][1 + r] += mu * t_b;
                g_b[r][1 + c] += mu * t_b;
            }
        }
        // cv_dT = (gamma-1)^-1 irho (g[c][p] - por g[c][rho])
        g_b[c][DIM + 1] += cv_b * gm1inv * irho;
        g_b[c][0] += -cv_b * gm1inv * irho * por;
        por_b += -cv_b * gm1inv * irho * g[c][0];
        irho_b += cv_b * gm1inv * (g[c][DIM + 1] - por * g[c][0]);
    }
    for (int d = 0; d < DIM; d++) g_b[d][1 + d] += div_b;
    const double p_b = por_b * irho;
    irho_b += por_b * p;
    s_b[DIM + 1] += gm1 * p_b;
    double rho_b = -gm1 * 0.5 * vsq * p_b;
    for (int d = 0; d < DIM; d++) {
        vel_b[d] += -gm1 * rho * vel[d] * p_b;
        s_b[1 + d] += vel_b[d] * irho;
        irho_b += vel_b[d] * s[1 + d];
    }
    rho_b += -irho_b * irho * irho;
    s_b[0] += rho_b;
}

}  // namespace adj
}  // namespace prandtl
