// dg_physics.hpp -- pointwise compressible-flow physics for the gfx950 kernels.
//
// Device functions restating the reference's pointwise arithmetic (file:line
// cited per function, paths relative to the reference repository root).  All
// are templated on the spatial dimension; arithmetic is FP64 unless a caller
// instantiates the float variants used by the mixed-precision path.
#pragma once

#include <hip/hip_runtime.h>

namespace prandtl {

struct PhysConst {
    double gamma, gammaM1, gammaM1Inverse;
    double mu, PrInverse, R_gas, mu_bulk;
    double mu0, T0, Ts;
    int sutherland;
    // the logarithmic mean's series constants (lm_delta): kernel arguments, so that they enter the
    // flux arithmetic as scalar operands instead of being rebuilt in vector registers at every use
    double lm_k = 7.0 / 5.0, lm_c3 = 1.0 / 3.0, lm_c7 = 1.0 / 7.0;
    // the longer series' coefficients 1/(2k+1), k = 1..11 (lm_delta_long)
    double lm_cl[11] = {1.0 / 3.0,  1.0 / 5.0,  1.0 / 7.0,  1.0 / 9.0,  1.0 / 11.0, 1.0 / 13.0,
                        1.0 / 15.0, 1.0 / 17.0, 1.0 / 19.0, 1.0 / 21.0, 1.0 / 23.0};
};

// ---------------------------------------------------------------------------
// Arithmetic helpers.  Divisions in the flux kernels are reciprocals: v_rcp_f64 seed
// refined by two Newton-Raphson steps (within ~1 ulp; every argument is a positive,
// normal density, pressure or sum of them).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Primitive state of one node as the two-point flux consumes it.
template <int DIM>
struct NodePrim {
    double rho;
    double vel[DIM];
    double beta;  // rho / (2 p)  (ChandrashekarFlux.cpp:67)
};

// Primitive state plus what a surface flux needs of one side.
template <int DIM>
struct SideState {
    double rho, irho, vel[DIM], vsq, p, beta, lam;
    double por;   // p / rho
};

template <int DIM>
__device__ __forceinline__ SideState<DIM> make_side(const double *s, const PhysConst &P) {
    SideState<DIM> q;
    q.rho = s[0];
    q.irho = rcp(q.rho);
    q.vsq = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        q.vel[d] = s[1 + d] * q.irho;
        q.vsq += q.vel[d] * q.vel[d];
    }
    q.p = P.gammaM1 * (s[DIM + 1] - 0.5 * q.rho * q.vsq);       // ChandrashekarFlux.cpp:58
    const double por = q.p * q.irho;
    q.por = por;
    q.beta = 0.5 * rcp(por);                                   // :67  rho / (2 p)
    q.lam = sqrt(q.vsq) + sqrt(P.gamma * por);                 // :61-65
    return q;
}

// The surface fluxes' view of a node whose primitives a kernel already holds (density, velocity,
// beta = rho / (2 p), max wave speed): p / rho = 1 / (2 beta), p = rho p / rho.  irho and vsq are
// not set (the surface fluxes do not read them).
template <int DIM>
__device__ __forceinline__ SideState<DIM> side_from_prim(double rho, double beta, const double *vel, double lam) {
    SideState<DIM> q;
    q.rho = rho;
    q.beta = beta;
#pragma unroll
    for (int d = 0; d < DIM; d++) q.vel[d] = vel[d];
    q.lam = lam;
    q.por = 0.5 * rcp(beta);
    q.p = rho * q.por;
    q.irho = 0.0;
    q.vsq = 0.0;
    return q;
}

template <int DIM>
__device__ __forceinline__ NodePrim<DIM> to_prim(const SideState<DIM> &q) {
    NodePrim<DIM> r;
    r.rho = q.rho;
#pragma unroll
    for (int d = 0; d < DIM; d++) r.vel[d] = q.vel[d];
    r.beta = q.beta;
    return r;
}

// delta = u/3 + u^2/5 + u^3/7 = u (1/3 + u (u + 7/5) / 7): one scalar constant per instruction
__device__ __forceinline__ double lm_delta(double u, const PhysConst &P) {
    return u * fma(u, (u + P.lm_k) * P.lm_c7, P.lm_c3);
}

// The same series to u^N for 1e-4 <= u < kLmLongU: delta = sum_{k=1..N} u^k/(2k+1).  Its
// tail sum_{k>N} u^k/(2k+1) < u^(N+1)/((2N+3)(1-u)) is below 1.1e-17 (a tenth of the
// double's half-ulp at 1) for u < 0.05, N = 11.  In this range it is also the more accurate of
// the two forms: ln(y/x) of a ratio near 1 turns the ratio's rounding error e into a relative
// error e/ln(y/x), about 50 ulp at u = 1e-4.  A wave whose pairs all fall below kLmLongU skips
// the logarithm, which a single lane with a ratio a little off 1 otherwise makes every lane wait
// for (the density and temperature jumps of a shear layer put such pairs in most waves).
constexpr double kLmLongU = 0.05;
constexpr int kLmLongN = 11;
__device__ __forceinline__ double lm_delta_long(double u, const PhysConst &P) {
    static_assert(sizeof(P.lm_cl) / sizeof(double) == kLmLongN, "one coefficient per term");
    double s = P.lm_cl[kLmLongN - 1];
#pragma unroll
    for (int k = kLmLongN - 2; k >= 0; k--) s = fma(s, u, P.lm_cl[k]);
    return u * s;
}

// 1 / L(x, y), L the logarithmic mean, given r = 1/(x+y).
// ref: ComputeLogMean, src/BasicOperations/BasicOperations.cpp:6-12.  With f = (y-x)/(y+x)
// and u = f^2 (= the reference's (xi-1)^2/(xi+1)^2, xi = y/x), ln(y/x) = 2 artanh f, so
// L = (x+y)/2 / (1 + u/3 + u^2/5 + u^3/7 + ...); the reference's series branch
// (x+y) 52.5/(105 + u(35 + u(21 + 15u))) is that expansion to u^3.  Same branch point u < 1e-4,
// then the longer series to u < kLmLongU, the closed form beyond; the reciprocal form lets one
// 1/(x+y) serve the flux's other terms.
__device__ __forceinline__ double inv_log_mean(double x, double y, double r, const PhysConst &P) {
    const double f = (y - x) * r;
    const double u = f * f;
    double il;
    if (u < 1e-4) {
        const double r2 = 2.0 * r;
        il = fma(r2, lm_delta(u, P), r2);
    } else if (u < kLmLongU) {
        const double r2 = 2.0 * r;
        il = fma(r2, lm_delta_long(u, P), r2);
    } else {
        il = log(y * rcp(x)) * rcp(y - x);
    }
    return il;
}

// 1 / L(x, y) and L(x, y) together, given r = 1/(x+y).  On the short series branch
// L = (x+y)/2 / (1 + delta), delta = u/3 + u^2/5 + u^3/7 < 3.4e-5, and 1/(1 + delta) =
// 1 - delta (1 - delta (1 - delta)) + O(delta^4) (the omitted term is below 2e-18 relative),
// so L needs no reciprocal there.
__device__ __forceinline__ void log_mean_both(double x, double y, double r, const PhysConst &P, double &il,
                                              double &l) {
    const double f = (y - x) * r;
    const double u = f * f;
    if (u < 1e-4) {
        const double d = lm_delta(u, P), r2 = 2.0 * r;
        il = fma(r2, d, r2);
        l = 0.5 * (x + y) * fma(-d, fma(-d, 1.0 - d, 1.0), 1.0);
    } else if (u < kLmLongU) {
        const double r2 = 2.0 * r;
        il = fma(r2, lm_delta_long(u, P), r2);
        l = rcp(il);
    } else {
        il = log(y * rcp(x)) * rcp(y - x);
        l = rcp(il);
    }
}

// ChandrashekarFlux::ComputeVolumeFlux -- ref: src/RiemannSolvers/ChandrashekarFlux.cpp:11-83.
// m is the (already averaged) metric row Ja^dir; F receives DIM+2 values.  LN_SERIES: the
// logarithmic mean of the density from log_mean_both (one reciprocal fewer; the line kernel,
// whose threads hold whole lines, keeps the reciprocal to stay within its registers).
template <int DIM, bool LN_SERIES = false>
__device__ __forceinline__ void volume_flux(const NodePrim<DIM> &a, const NodePrim<DIM> &b, const double *m,
                                            const PhysConst &P, double *F) {
    const double gammaM1Inverse = P.gammaM1Inverse;
    const double rsum = a.rho + b.rho;
    const double r_rho = rcp(rsum);
    const double r_b = rcp(a.beta + b.beta);
    double il_rho, rho_ln;
    if constexpr (LN_SERIES) {
        log_mean_both(a.rho, b.rho, r_rho, P, il_rho, rho_ln);
    } else {
        il_rho = inv_log_mean(a.rho, b.rho, r_rho, P);
        rho_ln = rcp(il_rho);
    }
    const double il_b = inv_log_mean(a.beta, b.beta, r_b, P);
    // with vs = v_a + v_b the mean velocity is vs/2, and the kinetic part of h-hat,
    // |mean|^2 - (|v_a|^2 + |v_b|^2)/4, is (v_a . v_b)/2: one dot product, no cancellation
    double vs[DIM];
    double vdot = 0.0, Vn2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vs[d] = a.vel[d] + b.vel[d];
        vdot += a.vel[d] * b.vel[d];
        Vn2 += vs[d] * m[d];
    }
    const double p_hat = 0.5 * rsum * r_b;                                // :71
    const double h_hat = 0.5 * (vdot + il_b * gammaM1Inverse) + p_hat * il_rho;   // :73
    const double mflux = 0.5 * rho_ln * Vn2;                              // rho_ln (mean . m)
    const double hf = 0.5 * mflux;
    F[0] = mflux;
#pragma unroll
    for (int d = 0; d < DIM; d++) F[1 + d] = hf * vs[d] + p_hat * m[d];
    F[DIM + 1] = mflux * h_hat;
}

// Euler flux . n (textbook; the reference takes it from mfem::EulerFlux)
template <int DIM>
__device__ __forceinline__ void euler_flux_n(const double *s, const SideState<DIM> &q, const double *n, double *fn) {
    double vn = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) vn += q.vel[d] * n[d];
    fn[0] = q.rho * vn;
#pragma unroll
    for (int d = 0; d < DIM; d++) fn[1 + d] = s[1 + d] * vn + q.p * n[d];
    fn[DIM + 1] = (s[DIM + 1] + q.p) * vn;
}

enum SurfaceKind { SURF_CHANDRASHEKAR = 0, SURF_EC = 1, SURF_RUSANOV = 2 };

// Surface flux plug-ins.  SURF_CHANDRASHEKAR -- ref: ChandrashekarFlux::ComputeFaceFlux,
// ChandrashekarFlux.cpp:85-161 (bug3 reproduces the :131 typo "rho_mean*dw + w_mean");
// SURF_EC -- the volume flux with metric = n; SURF_RUSANOV -- local Lax-Friedrichs on the
// central flux.  s1/s2 are conserved states with their SideStates, n the face area vector
// outward from side 1, nmag = |n|.
template <int DIM, int SURF>
__device__ __forceinline__ void face_flux(const double *s1, const SideState<DIM> &q1, const double *s2,
                                          const SideState<DIM> &q2, const double *n, double nmag,
                                          const PhysConst &P, bool bug3, double *flux) {
    if constexpr (SURF == SURF_CHANDRASHEKAR) {
        const double rsum = q1.rho + q2.rho;
        const double r_rho = rcp(rsum);
        const double r_b = rcp(q1.beta + q2.beta);
        double il_rho, rho_ln;
        log_mean_both(q1.rho, q2.rho, r_rho, P, il_rho, rho_ln);
        const double il_b = inv_log_mean(q1.beta, q2.beta, r_b, P);
        const double rho_mean = 0.5 * rsum;
        const double drho = q2.rho - q1.rho;
        // the kinetic parts through v1 . v2: h-hat's |mean|^2 - (|v1|^2 + |v2|^2)/4 = (v1 . v2)/2,
        // and the energy jump's sum of drho v1_d v2_d / 2 is drho (v1 . v2)/2
        double vm[DIM];
        double Vn = 0.0, vdot = 0.0, dvm = 0.0, diss = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            vm[d] = 0.5 * (q1.vel[d] + q2.vel[d]);
            const double dv = q2.vel[d] - q1.vel[d];
            Vn += vm[d] * n[d];
            vdot += q1.vel[d] * q2.vel[d];
            if (DIM == 3 && d == 2 && bug3)
                diss += rho_mean * dv + vm[d];                                          // :131 as written
            else
                dvm += dv * vm[d];                                                      // :105,:118,:131
        }
        const double lam = fmax(q1.lam, q2.lam);
        const double p_hat = rho_mean * r_b;
        const double h_hat = 0.5 * (vdot + il_b * P.gammaM1Inverse) + p_hat * il_rho;
        // 1/beta = 2 p / rho
        diss += 0.5 * drho * (vdot + P.gammaM1Inverse * il_b) + rho_mean * dvm +
                rho_mean * P.gammaM1Inverse * (q2.por - q1.por);                        // :151
        const double hl = 0.5 * lam * nmag;
        const double mflux = rho_ln * Vn;
        flux[0] = mflux - hl * drho;
#pragma unroll
        for (int d = 0; d < DIM; d++) flux[1 + d] = mflux * vm[d] + p_hat * n[d] - hl * (s2[1 + d] - s1[1 + d]);
        flux[DIM + 1] = mflux * h_hat - hl * diss;
    } else if constexpr (SURF == SURF_EC) {
        volume_flux<DIM>(to_prim(q1), to_prim(q2), n, P, flux);
    } else {
        double f1[DIM + 2], f2[DIM + 2];
        euler_flux_n<DIM>(s1, q1, n, f1);
        euler_flux_n<DIM>(s2, q2, n, f2);
        const double hl = 0.5 * fmax(q1.lam, q2.lam) * nmag;
#pragma unroll
        for (int v = 0; v < DIM + 2; v++) flux[v] = 0.5 * (f1[v] + f2[v]) - hl * (s2[v] - s1[v]);
    }
}

// ---------------------------------------------------------------------------
// Stage-hook pointwise tests (the entropy filter's detection and the positivity limiter's
// trigger), shared by the hook kernels and the stage epilogue that detects for them.
// ---------------------------------------------------------------------------
constexpr double kFilterRhoMin = 1e-8;   // EntropyFilter.hpp:59
constexpr double kFilterPMin = 1e-8;     // :60
constexpr double kFilterEntTol = 1e-4;   // :61
constexpr double kFilterEpsilon = 1e-10; // :62 (zeta_max = -ln epsilon)
constexpr int kFilterNIter = 20;         // :63
constexpr double kLimiterEps = 1e-8;

template <int DIM>
__device__ __forceinline__ double hook_pressure(const double *s, double gammaM1) {
    double v2 = 0.0;
#pragma unroll
    for (int d = 1; d <= DIM; d++) v2 += (s[d] / s[0]) * (s[d] / s[0]);
    return gammaM1 * (s[DIM + 1] - 0.5 * s[0] * v2);
}

// rho ln(p rho^-gamma) for rho, p > 0, else +inf (BasicOperations.cpp:39-46)
__device__ __forceinline__ double hook_entropy(double rho, double p, double gamma) {
    return (rho > 0.0 && p > 0.0) ? rho * (log(p) - gamma * log(rho)) : __builtin_inf();
}

// ---------------------------------------------------------------------------
// Mixed precision (PRANDTL_MIXED).  The logarithmic means, the reciprocals of the sums
// they share and p-hat stay FP64; the rest of the two-point and surface fluxes is FP32.
// The momentum fluxes carry p-hat - p_ref with p_ref a constant of the element whose
// equation is being assembled: a constant pressure contributes nothing to the assembled
// RHS on an affine element (the volume sum of D-hat-2 rows and the surface terms cancel,
// i.e. free-stream preservation), and subtracting it in FP64 keeps the FP32 momentum
// sums from cancelling three digits of a ~1/(gamma Ma^2) pressure.  Likewise the energy
// fluxes carry F_E - h_ref F_rho (h_ref a constant enthalpy of the element); the caller adds
// h_ref times the assembled mass RHS back in FP64, which restores the energy RHS exactly in
// exact arithmetic while the FP32 sums see only the enthalpy's variation.
// ---------------------------------------------------------------------------
template <int DIM>
struct NodePrimMx {
    double rho, beta;
    float vel[DIM];
};

template <int DIM>
__device__ __forceinline__ void volume_flux_mx(const NodePrimMx<DIM> &a, const NodePrimMx<DIM> &b, const float *m,
                                               const PhysConst &P, double p_ref, double h_ref, float *F) {
    const double gammaM1Inverse = P.gammaM1Inverse;
    const double rsum = a.rho + b.rho;
    const double r_rho = rcp(rsum);
    const double r_b = rcp(a.beta + b.beta);
    const double il_rho = inv_log_mean(a.rho, b.rho, r_rho, P);
    const double il_b = inv_log_mean(a.beta, b.beta, r_b, P);
    const double p_hat = 0.5 * rsum * r_b;
    const float ilr = (float)il_rho, ilb = (float)il_b;
    const float rho_ln = rcp(ilr);
    // as volume_flux: vs = v_a + v_b, the kinetic part of h-hat is (v_a . v_b)/2
    float vs[DIM];
    float vdot = 0.f, Vn2 = 0.f;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vs[d] = a.vel[d] + b.vel[d];
        vdot += a.vel[d] * b.vel[d];
        Vn2 += vs[d] * m[d];
    }
    const float h_hat = 0.5f * vdot + (float)(0.5 * il_b * gammaM1Inverse + p_hat * il_rho - h_ref);
    const float pd = (float)(p_hat - p_ref);
    const float mflux = 0.5f * rho_ln * Vn2;
    const float hf = 0.5f * mflux;
    F[0] = mflux;
#pragma unroll
    for (int d = 0; d < DIM; d++) F[1 + d] = hf * vs[d] + pd * m[d];
    F[DIM + 1] = mflux * h_hat;
}

template <int DIM, int SURF>
__device__ __forceinline__ void face_flux_mx(const double *s1, const SideState<DIM> &q1, const double *s2,
                                             const SideState<DIM> &q2, const float *n, float nmag, const PhysConst &P,
                                             bool bug3, double p_ref, double h_ref, float *flux) {
    if constexpr (SURF == SURF_CHANDRASHEKAR) {
        const double rsum = q1.rho + q2.rho;
        const double r_rho = rcp(rsum);
        const double r_b = rcp(q1.beta + q2.beta);
        const double il_rho = inv_log_mean(q1.rho, q2.rho, r_rho, P);
        const double il_b = inv_log_mean(q1.beta, q2.beta, r_b, P);
        const double p_hat = 0.5 * rsum * r_b;
        const float ilr = (float)il_rho, ilb = (float)il_b;
        const float rho_ln = rcp(ilr);
        const float rho_mean = (float)(0.5 * rsum);
        const float drho = (float)(q2.rho - q1.rho);
        const float g1 = (float)P.gammaM1Inverse;
        float vm[DIM];
        float Vn = 0.f, vdot = 0.f, dvm = 0.f, diss = 0.f;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const float v1 = (float)q1.vel[d], v2 = (float)q2.vel[d];
            vm[d] = 0.5f * (v1 + v2);
            const float dv = (float)(q2.vel[d] - q1.vel[d]);
            Vn += vm[d] * n[d];
            vdot += v1 * v2;
            if (DIM == 3 && d == 2 && bug3)
                diss += rho_mean * dv + vm[d];
            else
                dvm += dv * vm[d];
        }
        diss += 0.5f * drho * vdot + rho_mean * dvm;
        const float lam = (float)fmax(q1.lam, q2.lam);
        const float h_hat = 0.5f * vdot + (float)(0.5 * il_b * P.gammaM1Inverse + p_hat * il_rho - h_ref);
        // diss - h_ref drho: the energy jump measured from h_ref times the density jump
        const double drho64 = q2.rho - q1.rho;
        diss += (float)(drho64 * (0.5 * P.gammaM1Inverse * il_b - h_ref)) +
                rho_mean * g1 * (float)(q2.por - q1.por);
        const float hl = 0.5f * lam * nmag;
        const float mflux = rho_ln * Vn;
        const float pd = (float)(p_hat - p_ref);
        flux[0] = mflux - hl * drho;
#pragma unroll
        for (int d = 0; d < DIM; d++) flux[1 + d] = mflux * vm[d] + pd * n[d] - hl * (float)(s2[1 + d] - s1[1 + d]);
        flux[DIM + 1] = mflux * h_hat - hl * diss;
    } else if constexpr (SURF == SURF_EC) {
        NodePrimMx<DIM> a, b;
        a.rho = q1.rho;
        a.beta = q1.beta;
        b.rho = q2.rho;
        b.beta = q2.beta;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            a.vel[d] = (float)q1.vel[d];
            b.vel[d] = (float)q2.vel[d];
        }
        volume_flux_mx<DIM>(a, b, n, P, p_ref, h_ref, flux);
    } else {
        double nd[DIM], f1[DIM + 2], f2[DIM + 2];
#pragma unroll
        for (int d = 0; d < DIM; d++) nd[d] = n[d];
        euler_flux_n<DIM>(s1, q1, nd, f1);
        euler_flux_n<DIM>(s2, q2, nd, f2);
        const double hl = 0.5 * fmax(q1.lam, q2.lam) * nmag;
#pragma unroll
        for (int v = 0; v < DIM + 2; v++) {
            double fv = 0.5 * (f1[v] + f2[v]) - hl * (s2[v] - s1[v]);
            if (v >= 1 && v <= DIM) fv -= p_ref * nd[v - 1];
            flux[v] = (float)fv;
        }
        flux[DIM + 1] = (float)(0.5 * (f1[DIM + 1] + f2[DIM + 1]) - hl * (s2[DIM + 1] - s1[DIM + 1]) -
                                h_ref * (0.5 * (f1[0] + f2[0]) - hl * (s2[0] - s1[0])));
    }
}

// ln x near a reference x0 whose logarithm lx0 and reciprocal ix0 are known:
// ln x = ln x0 + 2 artanh(t / (2 + t)), t = x/x0 - 1, for |t| <= 1/32 by the artanh series to
// z^11 (|z| < 1/63, so the first omitted term z^13/13 is below 2^-80 relative to ln x); the
// library log otherwise.  x - x0 is exact there (Sterbenz), so the result is within about
// half an ulp of ln x as the library log is.
__device__ __forceinline__ double log_near(double x, double x0, double lx0, double ix0) {
    const double t = (x - x0) * ix0;
    if (fabs(t) <= 0.03125) {
        const double z = t * rcp(2.0 + t);
        const double z2 = z * z;
        const double q = 1.0 / 3.0 + z2 * (1.0 / 5.0 + z2 * (1.0 / 7.0 + z2 * (1.0 / 9.0 + z2 * (1.0 / 11.0))));
        return lx0 + 2.0 * fma(z * z2, q, z);
    }
    return log(x);
}

// Reference state of an element for log_near: p0, rho0 with their logarithms and reciprocals.
struct LogRef {
    double p0, lp0, ip0, r0, lr0, ir0;
};

template <int DIM>
__device__ __forceinline__ LogRef make_log_ref(const double *s, const PhysConst &P) {
    LogRef R;
    R.r0 = s[0];
    R.ir0 = rcp(R.r0);
    double v_sq = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) v_sq += s[1 + d] * s[1 + d];
    R.p0 = P.gammaM1 * (s[DIM + 1] - 0.5 * v_sq * R.ir0);
    R.ip0 = rcp(R.p0);
    R.lp0 = log(R.p0);
    R.lr0 = log(R.r0);
    return R;
}

// Conserv2Entropy -- ref: BasicOperations.cpp:81-104.  s = ln(p rho^-gamma) is evaluated as
// ln p - gamma ln rho (the same value to rounding), each logarithm near the element's
// reference state (log_near), the divisions as reciprocals.
template <int DIM>
__device__ __forceinline__ void conserv_to_entropy_ref(const double *s, const PhysConst &P, const LogRef &R,
                                                       double *w) {
    const double rho = s[0];
    const double irho = rcp(rho);
    double vel[DIM], v_sq = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vel[d] = s[1 + d] * irho;
        v_sq += vel[d] * vel[d];
    }
    v_sq = 0.5 * v_sq;
    const double p = P.gammaM1 * (s[DIM + 1] - rho * v_sq);
    const double beta = rho * rcp(p);
    const double ent = log_near(p, R.p0, R.lp0, R.ip0) - P.gamma * log_near(rho, R.r0, R.lr0, R.ir0);
    w[0] = (P.gamma - ent) * P.gammaM1Inverse - beta * v_sq;
#pragma unroll
    for (int d = 0; d < DIM; d++) w[1 + d] = beta * vel[d];
    w[DIM + 1] = -beta;
}

// EntropyGrad2PrimGrad for one direction, in place -- ref: BasicOperations.cpp:126-170
template <int DIM, typename T = double>
__device__ __forceinline__ void entgrad_to_primgrad(const T *s, const PhysConst &P, T *g) {
    const T rho = s[0];
    const T irho = rcp(rho);
    T vel[DIM], v_sq = T(0);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vel[d] = s[1 + d] * irho;
        v_sq += vel[d] * vel[d];
    }
    v_sq = T(0.5) * v_sq;
    const T p = T(P.gammaM1) * (s[DIM + 1] - rho * v_sq);
    const T KE = v_sq * rho;
    const T ge = g[DIM + 1];
    const T por = p * irho;
    T mdot = T(0);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        g[1 + d] = por * (g[1 + d] + vel[d] * ge);
        mdot += s[1 + d] * g[1 + d];
    }
    g[0] = rho * g[0] - ge * (KE - p * T(P.gammaM1Inverse)) + rcp(por) * mdot;
    g[DIM + 1] = por * (g[0] + p * ge);
}

// NavierStokesFlux::ComputeViscousFlux -- ref: src/Fluxes/NavierStokesFlux.cpp:12-58 (3D),
// :60-90 (2D); Sutherland NavierStokesFlux.hpp:32-36.  g[c][v] are the primitive
// gradients (rho, vel, p) along x_c.  flux[v-1][c] for v = 1..DIM+1 (row 0 is zero).
template <int DIM, typename T = double>
__device__ __forceinline__ void viscous_flux(const T *s, const T (*g)[DIM + 2], const PhysConst &P, T (*flux)[DIM]) {
    const T rho = s[0];
    const T irho = rcp(rho);
    T vel[DIM], vsq = T(0);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vel[d] = s[1 + d] * irho;
        vsq += vel[d] * vel[d];
    }
    const T p = T(P.gammaM1) * (s[DIM + 1] - T(0.5) * rho * vsq);
    T mu = T(P.mu);
    if (P.sutherland) {
        const double Tk = double(p) / (double(rho) * P.R_gas);
        mu = T(P.mu0 * (P.T0 + P.Ts) / (Tk + P.Ts) * (Tk / P.T0) * sqrt(Tk / P.T0));
    }
    const T lambda = mu * T(P.gamma * P.PrInverse);
    const T por = p * irho;
    T div = g[0][1];
#pragma unroll
    for (int d = 1; d < DIM; d++) div += g[d][1 + d];
#pragma unroll
    for (int c = 0; c < DIM; c++) {
        const T cv_dT = T(P.gammaM1Inverse) * irho * (g[c][DIM + 1] - por * g[c][0]);
        T e = T(0);
#pragma unroll
        for (int r = 0; r < DIM; r++) {
            // tau_rc = mu (d vel_r / dx_c + d vel_c / dx_r) - mu mu_bulk div delta_rc
            T t = (r == c) ? mu * (T(2) * g[c][1 + r] - T(P.mu_bulk) * div) : mu * (g[c][1 + r] + g[r][1 + c]);
            flux[r][c] = t;
            e += vel[r] * t;
        }
        flux[DIM][c] = e + lambda * cv_dT;
    }
}

}  // namespace prandtl
