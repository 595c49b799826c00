// dg_bc.hpp -- boundary conditions for the node RHS and lift kernels (gfx950, FP64).
//
// The reference's BdrFaceIntegrator hierarchy (src/Integrators/BoundaryConditions/), one
// condition per boundary attribute (Simulation.cpp:505-800), with constant boundary values.
// A boundary face of element e is the neighbour entry -(k+1), k the index into the operator's
// table of conditions.  The state of an inflow or specified-state condition may vary over the
// boundary and in time (bc_state).  Each condition supplies, at a face point with the element's trace s1
// and the element's outward normal n (the face area vector, CalcOrtho):
//   the inviscid flux f(s1, n) and a wave speed   ComputeBdrFaceInviscidFlux
//                                                   (BdrFaceIntegrator.cpp:245-253: the
//                                                   surface flux with the outer state; the
//                                                   slip and symmetry walls override it);
//   the viscous flux F_v . n                       ComputeBdrFaceViscousFlux;
//   the BR1 lifting term (entropy variables)       ComputeBdrFaceLiftingFlux, default
//                                                   1/2 (w(outer) - w1) (:275-282).
// The slip wall and symmetry conditions define only the inviscid flux (Euler only, as in the
// reference, whose base class aborts for their viscous and lifting terms).
#pragma once

#include "dg_physics.hpp"

namespace prandtl {

struct BcDev {          // the layout of prandtl_bc
    int type;
    double q[8];
    int profile;
    double a[16];
};

enum BcProfile { BC_PROFILE_CONSTANT = 0, BC_PROFILE_FRONT = 1, BC_PROFILE_RAMP = 2 };

enum BcKind {
    BC_SLIP_WALL = 1, BC_SYMMETRY = 2, BC_SUPERSONIC_INFLOW = 3, BC_SUPERSONIC_OUTFLOW = 4,
    BC_SPECIFIED_STATE = 5, BC_RIEMANN_INVARIANT = 6, BC_SUBSONIC_INFLOW_RV = 7,
    BC_SUBSONIC_INFLOW_PTTT = 8, BC_SUBSONIC_OUTFLOW_P = 9, BC_NOSLIP_ADIABATIC = 10,
    BC_NOSLIP_ISOTHERMAL = 11
};

// ComputePressure (BasicOperations.cpp:31-37)
template <int DIM>
__device__ __forceinline__ double bc_pressure(const double *s, const PhysConst &P) {
    const double inv = 1.0 / s[0];
    double vv = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) vv += (s[1 + d] * inv) * (s[1 + d] * inv);
    return P.gammaM1 * (s[DIM + 1] - 0.5 * s[0] * vv);
}

template <int DIM>
__device__ __forceinline__ void bc_unit(const double *n, double *un) {
    double a = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) a += n[d] * n[d];
    a = sqrt(a);
#pragma unroll
    for (int d = 0; d < DIM; d++) un[d] = n[d] / a;
}

// The state of a SUPERSONIC_INFLOW / SPECIFIED_STATE condition at the face point x and time t: the
// constant q, or the test cases' function-valued states (the VectorFunctionCoefficient /
// VectorTDFunctionCoefficient the reference evaluates per face point, Simulation.cpp:664-790):
//   FRONT (DoubleMachReflectionTopBCFunction, DoubleMachReflection.hpp:32-52): q behind the
//     planar front x[i0] < c0 + (x[i1] + c1 t) / c2, the state a[5..] ahead of it;
//   RAMP (NagashimaRampingInletBC, Nagashima_Ramjet.hpp:40-53): the velocity a[2..1+dim] scaled by
//     r = clamp((1 + tanh((t - t_r/2) / (t_r/10))) / 2, 0, 1), density a[1], pressure a[2+dim].
// The operations are the reference functions' own, in their order.
template <int DIM>
__device__ __forceinline__ void bc_state(const BcDev &bc, const PhysConst &P, const double *x, double t, double *s2) {
    constexpr int NEQ = DIM + 2;
    if (bc.profile == BC_PROFILE_FRONT) {
        const int i0 = (int)bc.a[0], i1 = (int)bc.a[1];
        const bool behind = x[i0] < bc.a[2] + (x[i1] + bc.a[3] * t) / bc.a[4];
#pragma unroll
        for (int v = 0; v < NEQ; v++) s2[v] = behind ? bc.q[v] : bc.a[5 + v];
    } else if (bc.profile == BC_PROFILE_RAMP) {
        const double tr = bc.a[0];
        const double sh = tanh((t - 0.5 * tr) / (0.1 * tr));
        const double r = fmin(fmax(0.5 * (1.0 + sh), 0.0), 1.0);
        const double rho = bc.a[1];
        double e = bc.a[2 + DIM] / P.gammaM1;
        s2[0] = rho;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double v = bc.a[2 + d] * r;
            s2[1 + d] = rho * v;
            e += 0.5 * rho * v * v;
        }
        s2[NEQ - 1] = e;
    } else {
#pragma unroll
        for (int v = 0; v < NEQ; v++) s2[v] = bc.q[v];
    }
}

// ComputeOuterInviscidState (the conditions that define an outer state); x, t: the face point's
// coordinates and the time (read by the function-valued states only)
template <int DIM>
__device__ void bc_outer_state(const BcDev &bc, const PhysConst &P, const double *s1, const double *n, double *s2,
                               const double *x, double t) {
    constexpr int NEQ = DIM + 2;
    switch (bc.type) {
        case BC_SUPERSONIC_INFLOW:
        case BC_SPECIFIED_STATE:
            bc_state<DIM>(bc, P, x, t, s2);
            return;
        case BC_RIEMANN_INVARIANT: {   // RiemannInvariantBdrFaceIntegrator.cpp:44-117
            double un[DIM], Vi[DIM];
            bc_unit<DIM>(n, un);
            const double rho_o = bc.q[0], p_o = bc.q[NEQ - 1];
            const double inv = 1.0 / s1[0];
            double Vn_i = 0.0, Vn_o = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                Vi[d] = s1[1 + d] * inv;
                Vn_i += Vi[d] * un[d];
                Vn_o += bc.q[1 + d] * un[d];
            }
            const double p_i = bc_pressure<DIM>(s1, P);
            const double a_o = sqrt(P.gamma * p_o / rho_o), a_i = sqrt(P.gamma * p_i / s1[0]);
            const double gI = P.gammaM1Inverse;
            const double R_minus = (fabs(Vn_o) >= a_o && Vn_i >= 0.0) ? Vn_i - 2.0 * a_i * gI : Vn_o - 2.0 * a_o * gI;
            const double R_plus = (fabs(Vn_o) >= a_o && Vn_i < 0.0) ? Vn_o + 2.0 * a_o * gI : Vn_i + 2.0 * a_i * gI;
            const double Vn_b = 0.5 * (R_plus + R_minus);
            const double a_b = 0.25 * P.gammaM1 * (R_plus - R_minus);
            const double gInv = 1.0 / P.gamma;
            const bool in = Vn_i < 0.0;
            s2[0] = in ? pow(a_b * a_b * gInv * rho_o / p_o, gI) * rho_o : pow(a_b * a_b * gInv * s1[0] / p_i, gI) * s1[0];
            double e = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                s2[1 + d] = in ? s2[0] * (bc.q[1 + d] + (Vn_b - Vn_o) * un[d]) : s2[0] * (Vi[d] + (Vn_b - Vn_i) * un[d]);
                e += s2[1 + d] * s2[1 + d];
            }
            const double p_b = a_b * a_b * s2[0] * gInv;
            s2[NEQ - 1] = e * (0.5 / s2[0]) + p_b * gI;
            return;
        }
        case BC_SUBSONIC_INFLOW_RV: {  // SubsonicInflowRVBdrFaceIntegrator.cpp:22-52
            const double r = bc.q[0];
            double dke = 0.0, uu = 0.0;
            s2[0] = r;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                s2[1 + d] = r * bc.q[1 + d];
                dke -= s1[1 + d] * s1[1 + d];
                uu += bc.q[1 + d] * bc.q[1 + d];
            }
            s2[NEQ - 1] = s1[NEQ - 1] + (dke * (0.5 / s1[0]) + 0.5 * uu * r);
            return;
        }
        case BC_SUBSONIC_INFLOW_PTTT: {  // SubsonicInflowPtTtAngBdrFaceIntegrator.cpp:48-73
            const double pt = bc.q[0], Tt = bc.q[1], th = bc.q[2] * (M_PI / 180.0), ph = bc.q[3] * (M_PI / 180.0);
            const double cp = bc.q[4];
            double Vc[3] = {cos(th), sin(th), 0.0};
            if (DIM > 2) {
                // the reference writes cos(phi) to V_comps(3), past the 3-vector; the z component
                // it intends is V_comps(2)
                Vc[0] *= sin(ph);
                Vc[1] *= sin(ph);
                Vc[2] = cos(ph);
            }
            const double p = bc_pressure<DIM>(s1, P);
            const double V_sq = fmax(0.0, 2.0 * cp * Tt * (1.0 - pow(p / pt, P.gammaM1 / P.gamma)));
            s2[0] = (P.gamma * P.gammaM1Inverse) * p / (cp * Tt - 0.5 * V_sq);
            const double Vm = sqrt(V_sq);
#pragma unroll
            for (int d = 0; d < DIM; d++) s2[1 + d] = s2[0] * Vm * Vc[d];
            s2[NEQ - 1] = p * P.gammaM1Inverse + 0.5 * s2[0] * V_sq;
            return;
        }
        case BC_SUBSONIC_OUTFLOW_P: {  // SubsonicOutflowPBdrFaceIntegrator.cpp:17-28
            const double inv = 1.0 / s1[0];
            double vv = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) vv += (s1[1 + d] * inv) * (s1[1 + d] * inv);
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) s2[v] = s1[v];
            s2[NEQ - 1] = bc.q[0] * P.gammaM1Inverse + 0.5 * s1[0] * vv;
            return;
        }
        default:   // supersonic outflow (SupersonicOutflowBdrFaceIntegrator.cpp:12-15)
#pragma unroll
            for (int v = 0; v < NEQ; v++) s2[v] = s1[v];
    }
}

// ComputeBdrFaceInviscidFlux; returns the wave speed.  q1 is make_side(s1); x, t as bc_outer_state.
template <int DIM, int SURF>
__device__ double bc_inviscid_flux(const BcDev &bc, const PhysConst &P, bool bug3, const double *s1,
                                   const SideState<DIM> &q1, const double *n, double *f, const double *x, double t) {
    constexpr int NEQ = DIM + 2;
    double un[DIM];
    bc_unit<DIM>(n, un);
    if (bc.type == BC_SLIP_WALL || bc.type == BC_NOSLIP_ADIABATIC || bc.type == BC_NOSLIP_ISOTHERMAL) {
        // SlipWallBdrFaceIntegrator.cpp:15-57 (inherited by the no-slip walls): the pressure of
        // the reflected Riemann problem, from the normal velocity of the rotated trace
        double m[DIM], mn = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) mn += s1[1 + d] * un[d];
        double t1[3];
        if (DIM == 2) {
            t1[0] = -un[1];
            t1[1] = un[0];
        } else {
            const double r = un[1] / un[0];
            const bool ok = fabs(un[0]) > 1e-10;
            t1[0] = ok ? -r / sqrt(1.0 + r * r) : -1.0;
            t1[1] = ok ? 1.0 / sqrt(1.0 + r * r) : 0.0;
            t1[2] = 0.0;
        }
        double mt = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) mt += s1[1 + d] * t1[d];
        m[0] = mn;
        m[1] = mt;
        if (DIM == 3) {
            const double t2[3] = {un[1] * t1[2] - un[DIM - 1] * t1[1], un[DIM - 1] * t1[0] - un[0] * t1[2],
                                  un[0] * t1[1] - un[1] * t1[0]};
            m[DIM - 1] = s1[1] * t2[0] + s1[2] * t2[1] + s1[DIM] * t2[2];
        }
        const double rho = s1[0];
        double vv = 0.0;
        double prim[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            prim[d] = m[d] / rho;
            vv += prim[d] * prim[d];
        }
        double srot[NEQ];
        srot[0] = rho;
#pragma unroll
        for (int d = 0; d < DIM; d++) srot[1 + d] = m[d];
        srot[NEQ - 1] = s1[NEQ - 1];
        const double p = bc_pressure<DIM>(srot, P);
        const double an = sqrt(P.gamma * p / rho);
        const double gP1 = 1.0 + P.gamma, gP1I = 1.0 / gP1;
        const double u = prim[0];
        double p_star;
        if (u > 0.0)
            p_star = p + 0.25 * u * gP1 * rho * (u + sqrt(u * u + 8.0 * gP1I / rho * p * (P.gammaM1 * gP1I + 1.0)));
        else
            p_star = p * pow(fmax(1.0 + 0.5 * P.gammaM1 * u / an, 0.0001), 2.0 * P.gamma * P.gammaM1Inverse);
        f[0] = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) f[1 + d] = p_star * n[d];
        f[NEQ - 1] = 0.0;
        return sqrt(vv) + an;
    }
    double s2[NEQ];
    if (bc.type == BC_SYMMETRY) {   // SymmetryBdrFaceIntegrator.cpp:10-32: the mirrored trace
        double mn = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) mn += s1[1 + d] * un[d];
#pragma unroll
        for (int v = 0; v < NEQ; v++) s2[v] = s1[v];
#pragma unroll
        for (int d = 0; d < DIM; d++) s2[1 + d] += -2.0 * mn * un[d];
    } else {
        bc_outer_state<DIM>(bc, P, s1, n, s2, x, t);
    }
    const SideState<DIM> q2 = make_side<DIM>(s2, P);
    double nmag = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) nmag += n[d] * n[d];
    nmag = sqrt(nmag);
    face_flux<DIM, SURF>(s1, q1, s2, q2, n, nmag, P, bug3, f);
    return fmax(q1.lam, q2.lam);
}

// Entropy variables of a state (BasicOperations.cpp:81-104), logarithms taken directly.
template <int DIM>
__device__ __forceinline__ void bc_entropy(const double *s, const PhysConst &P, double *w) {
    const double rho = s[0];
    double vel[DIM], v_sq = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vel[d] = s[1 + d] / rho;
        v_sq += vel[d] * vel[d];
    }
    const double p = P.gammaM1 * (s[DIM + 1] - 0.5 * rho * v_sq);
    const double beta = rho / p;
    const double ent = log(p) - P.gamma * log(rho);
    w[0] = (P.gamma - ent) * P.gammaM1Inverse - 0.5 * beta * v_sq;
#pragma unroll
    for (int d = 0; d < DIM; d++) w[1 + d] = beta * vel[d];
    w[DIM + 1] = -beta;
}

// ComputeBdrFaceLiftingFlux: w1 the element's entropy variables, s1 its conserved trace; x, t as
// bc_outer_state.
template <int DIM>
__device__ void bc_lifting_flux(const BcDev &bc, const PhysConst &P, const double *w1, const double *s1,
                                const double *n, double *fl, const double *x, double t) {
    constexpr int NEQ = DIM + 2;
    if (bc.type == BC_NOSLIP_ADIABATIC || bc.type == BC_NOSLIP_ISOTHERMAL) {
        // NoSlipAdiabWallBdrFaceIntegrator.cpp:74-95, NoSlipIsothWallBdrFaceIntegrator.cpp:41-66:
        // the wall state's entropy variables (w1_0, V beta_w, -beta_w) minus w1
        const double v = bc.type == BC_NOSLIP_ADIABATIC ? -w1[NEQ - 1] : 1.0 / (P.R_gas * bc.q[DIM]);
        fl[0] = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) fl[1 + d] = bc.q[d] * v - w1[1 + d];
        fl[NEQ - 1] = -v - w1[NEQ - 1];
        return;
    }
    double s2[NEQ], w2[NEQ];
    bc_outer_state<DIM>(bc, P, s1, n, s2, x, t);
    bc_entropy<DIM>(s2, P, w2);
#pragma unroll
    for (int k = 0; k < NEQ; k++) fl[k] = (w2[k] - w1[k]) * 0.5;
}

// The face-trace value the lift kernel stores for a boundary face point (the RHS kernel uses it
// as the viscous flux along +adj_d): side ? F : -F of ComputeBdrFaceViscousFlux with the outward
// normal n = (side ? 1 : -1) adj_d.  s the trace, g its primitive gradients, Fc the trace's own
// contravariant viscous flux adj_d . F_v (rows 1..dim+1).
template <int DIM>
__device__ void bc_viscous_trace(const BcDev &bc, const PhysConst &P, const double *s, const double (*g)[DIM + 2],
                                 const double *adj_d, int side, const double *Fc, double *out) {
    constexpr int NEQ = DIM + 2;
    switch (bc.type) {
        case BC_SUPERSONIC_OUTFLOW:   // the trace and its gradients
        case BC_NOSLIP_ISOTHERMAL:
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) out[v] = Fc[v];
            return;
        case BC_NOSLIP_ADIABATIC: {   // NoSlipAdiabWallBdrFaceIntegrator.cpp:30-44
            double nm = 0.0, e = 0.0;
#pragma unroll
            for (int c = 0; c < DIM; c++) nm += adj_d[c] * adj_d[c];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                out[d] = Fc[d];
                e += bc.q[d] * Fc[d];
            }
            out[NEQ - 2] = e + (side ? 1.0 : -1.0) * bc.q[DIM] * sqrt(nm);
            return;
        }
        case BC_SUBSONIC_INFLOW_PTTT: {   // the outer state with the trace's gradients (:75-99)
            double n[DIM], s2[NEQ], flux[DIM + 1][DIM];
#pragma unroll
            for (int c = 0; c < DIM; c++) n[c] = (side ? 1.0 : -1.0) * adj_d[c];
            bc_outer_state<DIM>(bc, P, s, n, s2, nullptr, 0.0);   // a constant state
            viscous_flux<DIM, double>(s2, g, P, flux);
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) {
                double a = 0.0;
#pragma unroll
                for (int c = 0; c < DIM; c++) a += adj_d[c] * flux[v][c];
                out[v] = a;
            }
            return;
        }
        default:   // the outer state with zero gradients: no viscous flux
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) out[v] = 0.0;
    }
}

}  // namespace prandtl
