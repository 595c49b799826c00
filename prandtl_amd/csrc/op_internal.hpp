// op_internal.hpp -- internals of the C ABI shared by its translation units: the operator
// handle, the error channel, the derivative tables, and the stage launcher (kernel dispatch
// of one RHS evaluation / fused RK stage).  The stage launcher is instantiated per (dim, Np)
// in stage_kernels.hip, compiled once per pair so the kernel variants build in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/prandtl_amd.h"
#include "dg_hooks.hpp"
#include "dg_kernels.hpp"

namespace prandtl {

extern thread_local std::string g_last_error;   // prandtl_last_error()

inline int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

constexpr int MAX_NP = 5;

// GLL rule on [0,1] (prandtl_capi.hip)
void gll_rule(int np, double *x, double *w);

// binomial checkpointing plan of a taped run (prandtl_capi.hip)
struct RevolvePlan;

}  // namespace prandtl

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return prandtl::fail(PRANDTL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

using namespace prandtl;

struct prandtl_op_s {
    int dim = 0, order = 0, np = 0, nn = 0, NE = 0, neq = 0;
    size_t size = 0;       // doubles per state vector
    prandtl_physics phys{};
    prandtl_opts opts{};
    PhysConst P{};
    double D[MAX_NP * MAX_NP], Dh[MAX_NP * MAX_NP], Dh2[MAX_NP * MAX_NP], inv_w0 = 0.0;
    double w[MAX_NP] = {};        // GLL weights on [0,1]
    BcDev *d_bc = nullptr;        // boundary conditions (neighbour -(k+1) -> d_bc[k])
    int nbc = 0;
    double *d_bxyz = nullptr;     // face-point coordinates [e][2 dim][Np^(dim-1)][dim], only when a
                                  // boundary state varies in space
    double time = 0.0;            // prandtl_op_set_time: the time of Mult's boundary states
    double *d_alpha = nullptr;    // blending: per-element coefficient of the current RHS
    int *d_nbr = nullptr;
    double *d_geom = nullptr;
    double *d_ngeom = nullptr;     // general mesh: per-node (det J, adj J), NE x n x (1 + dim^2)
    int *d_forient = nullptr;     // general mesh: face orientation codes, NE x 2 dim
    double *d_dv = nullptr;       // NS: viscous volume term, (neq-1) x NE x n
    double *d_Ff = nullptr;       // NS: viscous face traces, dim x 2 x (neq-1) x NE x Np^(dim-1)
    double *d_k = nullptr;        // RK register
    double *d_ub = nullptr;       // RK alternate state buffer
    double *d_y2 = nullptr;       // SSP-RK3 second stage state
    double *d_emin = nullptr;     // stage hooks: per-element minimum nodal entropy
    double *d_fbound = nullptr;   // stage hooks: per-element entropy bound of the filter
    int *d_hflag = nullptr;       // stage hooks: per-element trigger flags
    unsigned long long *d_hcount = nullptr;   // stage hooks: elements filtered, limited
    int *d_hlist = nullptr;       // stage hooks: the elements a hook applies to
    int *d_hlist2 = nullptr;      // stage-fused hooks: the limiter's list (the filter's is d_hlist)
    unsigned int *d_hlistn = nullptr;         // their counters: two per hook (filter, limiter)
    int hook_par[2] = {0, 0};     // which of its two counters a hook's next scan fills
    bool have_bounds = false;     // filter constraints computed
    double *d_vq = nullptr;       // adjoint: per-node quantities, (dim+6) x NE x n
    double *d_vgf = nullptr;      // NS adjoint: lifting adjoint at face points, dim x 2 x neq x NE x Np^(dim-1)
    double *d_vwb = nullptr;      // NS adjoint: partial adjoint of the entropy variables
    double *d_vsb = nullptr;      // NS adjoint: partial adjoint of the state
    double *d_vlam = nullptr;     // adjoint: stage cotangent lambda
    unsigned long long *d_lam = nullptr;   // [0]: max |V| + c of the last Mult; [1]: of the last max_char_speed
    hipStream_t lam_stream = nullptr;      // the stream of the last Mult (its d_lam[0] is written there)
    double *d_lam_part = nullptr; // per-wave maxima of |V| + c (2 NE + 64 entries, at least 8192)
    int lam_waves = 0;            // waves of the last flux kernel launched (their maxima in d_lam_part)
    int *d_nan = nullptr;
    bool lam_valid = false;
    // taped forward run for the discrete adjoint: the state at the start of every segment
    int tape_steps = 0, tape_every = 0, tape_scheme = PRANDTL_LSRK45;
    double tape_dt = 0.0, tape_t0 = 0.0;
    unsigned long long tape_id = 0;   // identifies the taped run (0: none, or the last one failed)
    std::vector<double *> tape_ckpt;
    // adjoint sweep storage, kept between calls: the stage inputs of one segment (5 per step;
    // slot 0 points at the segment's checkpoint) and the double-buffered adjoint registers
    std::vector<double *> adj_stage;
    double *adj_ub2 = nullptr, *adj_kb = nullptr, *adj_kb2 = nullptr, *adj_work = nullptr;
    // checkpoint schedule of the taped run: the steps whose states tape_ckpt holds after the run
    // (segments: every segment start; revolve: the initial state and the first snapshots), the
    // schedule and its parameter, the revolve plan, and the forward steps the last reverse sweep ran
    std::vector<int> tape_marks;
    int tape_schedule = 0, tape_param = 0;
    std::shared_ptr<prandtl::RevolvePlan> revolve;
    long long tape_recomputed = -1;
    // launch timing (prandtl_op_timing): HIP events around every lifting (kind 0) and
    // flux-differencing (kind 1) launch while enabled; pool reused between reads
    bool timing = false;
    std::vector<hipEvent_t> tev;
    size_t tev_used = 0;
    std::vector<int> tkind;   // kind of event pair i (events 2i, 2i+1)
};

namespace prandtl {

template <int NP>
DTables<NP> make_tables(const prandtl_op_s *op) {
    DTables<NP> T;
    for (int i = 0; i < NP; i++)
        for (int j = 0; j < NP; j++) {
            T.D[i][j] = op->D[i * NP + j];
            T.Dhat[i][j] = op->Dh[i * NP + j];
            T.Dhat2[i][j] = op->Dh2[i * NP + j];
        }
    T.inv_w0 = op->inv_w0;
    for (int i = 0; i < NP; i++) T.inv_w[i] = 1.0 / op->w[i];
    // subcell metric weights: sub[i][m] = delta_m0 + sum_{l<=i} w_l D_lm
    for (int i = 0; i < NP; i++)
        for (int m = 0; m < NP; m++) {
            double c = (m == 0) ? 1.0 : 0.0;
            for (int l = 0; l <= i && l < NP - 1; l++) c += op->w[l] * op->D[l * NP + m];
            T.sub[i][m] = c;
        }
    return T;
}

// 1D inverse Vandermonde of the Legendre basis on the GLL nodes: Vi[k][a] = w_a P_k(x_a) / gamma_k,
// gamma_k = sum_b w_b P_k(x_b)^2 (the GLL rule integrates P_k P_l exactly for k + l <= 2 Np - 3 and
// the discrete products vanish for k != l <= Np - 1, so Vi V1 = I).
template <int NP>
BlendParams<NP> make_blend_params(const prandtl_op_s *op) {
    BlendParams<NP> B{};
    double x[NP], w[NP], V[NP][NP];
    gll_rule(NP, x, w);
    for (int i = 0; i < NP; i++) {
        const double z = 2.0 * x[i] - 1.0;
        V[i][0] = 1.0;
        if (NP > 1) V[i][1] = z;
        for (int n = 1; n + 1 < NP; n++) V[i][n + 1] = ((2.0 * n + 1.0) * z * V[i][n] - n * V[i][n - 1]) / (n + 1.0);
    }
    for (int k = 0; k < NP; k++) {
        double g = 0.0;
        for (int b = 0; b < NP; b++) g += w[b] * V[b][k] * V[b][k];
        for (int a = 0; a < NP; a++) B.Vi[k][a] = w[a] * V[a][k] / g;
    }
    B.T = 0.5 * std::pow(10.0, -1.8 * std::pow((double)op->order, 0.25));   // DGSEMOperator.cpp:25
    B.sharpness = 9.21024;                                                    // DGSEMOperator.hpp:50
    B.alpha_max = op->opts.alpha_max;
    B.alpha_min = op->opts.alpha_min;
    return B;
}

struct StageSpec {
    int mode;            // EPI_*
    const double *u;     // stage input
    double *out;
    double *k;
    const double *x;
    double ca, cb, dt;
    double *grad;        // optional primitive-gradient export (NS)
    bool want_lambda;
    bool nan_check;
    int phases = 3;      // bit 0: k_lift (NS), bit 1: k_rhs
    int hook_mode = 0;   // HOOK_DETECT_* bits: stage hooks detected in the epilogue
    double time = 0.0;   // the stage's time (function-valued boundary states)
};

template <int DIM, int NP, bool VISC, int SURF, bool MIX>
int launch_stage(prandtl_op_s *op, const StageSpec &S, hipStream_t st) {
    if constexpr (VISC) {
      if (S.phases & 1) {
        using NC = LiftCfg<DIM, NP>;
        LiftArgs<NP> la;
        la.T = make_tables<NP>(op);
        la.P = op->P;
        la.NE = op->NE;
        la.nbr = op->d_nbr;
        la.geom = op->d_geom;
        la.u = S.u;
        la.dv = op->d_dv;
        la.Ff = op->d_Ff;
        la.dvf = reinterpret_cast<float *>(op->d_dv);
        la.Fff = reinterpret_cast<float *>(op->d_Ff);
        la.grad = S.grad;
        la.bc = op->d_bc;
        la.bxyz = op->d_bxyz;
        la.time = S.time;
        la.ngeom = op->d_ngeom;
        la.forient = op->d_forient;
        const int nblk = (op->NE + NC::EPB - 1) / NC::EPB;
        if constexpr (!MIX) {
            if (op->opts.bug_compat & PRANDTL_BUG_A2)   // affine, no boundary faces (checked at create)
                hipLaunchKernelGGL((k_lift<DIM, NP, false, false, false, true>), dim3(nblk), dim3(NC::BLOCK), 0, st, la);
            else if (op->d_ngeom && op->nbc)
                hipLaunchKernelGGL((k_lift<DIM, NP, false, true, true>), dim3(nblk), dim3(NC::BLOCK), 0, st, la);
            else if (op->d_ngeom)
                hipLaunchKernelGGL((k_lift<DIM, NP, false, false, true>), dim3(nblk), dim3(NC::BLOCK), 0, st, la);
            else if (op->nbc)
                hipLaunchKernelGGL((k_lift<DIM, NP, false, true>), dim3(nblk), dim3(NC::BLOCK), 0, st, la);
            else
                hipLaunchKernelGGL((k_lift<DIM, NP, false>), dim3(nblk), dim3(NC::BLOCK), 0, st, la);
        } else {
            hipLaunchKernelGGL((k_lift<DIM, NP, MIX>), dim3(nblk), dim3(NC::BLOCK), 0, st, la);
        }
        HIP_TRY(hipGetLastError());
      }
    }
    op->lam_waves = 0;
    if (!(S.phases & 2)) return PRANDTL_OK;
    RhsArgs<NP> ra;
    ra.T = make_tables<NP>(op);
    ra.P = op->P;
    ra.NE = op->NE;
    ra.bug3 = (op->opts.bug_compat & PRANDTL_BUG_A3) ? 1 : 0;
    ra.mode = S.mode;
    ra.want_lambda = S.want_lambda ? 1 : 0;
    ra.nbr = op->d_nbr;
    ra.geom = op->d_geom;
    ra.u = S.u;
    ra.dv = op->d_dv;
    ra.Ff = op->d_Ff;
    ra.dvf = reinterpret_cast<const float *>(op->d_dv);
    ra.Fff = reinterpret_cast<const float *>(op->d_Ff);
    ra.out = S.out;
    ra.k = S.k;
    ra.x = S.x;
    ra.ca = S.ca;
    ra.cb = S.cb;
    ra.dt = S.dt;
    ra.lam_part = op->d_lam_part;
    ra.nan_flag = S.nan_check ? op->d_nan : nullptr;
    ra.alpha = op->d_alpha;
    ra.bc = op->d_bc;
    ra.bxyz = op->d_bxyz;
    ra.time = S.time;
    ra.ngeom = op->d_ngeom;
    ra.forient = op->d_forient;
    ra.hook_mode = S.hook_mode;
    ra.hbound = op->d_fbound;
    ra.hemin = op->d_emin;
    ra.hlist_f = op->d_hlist;
    ra.hlist_l = op->d_hlist2;
    ra.hcnt_f = ra.hcnt_f_next = ra.hcnt_l = ra.hcnt_l_next = nullptr;
    if (S.hook_mode) {
        ra.hcnt_f = op->d_hlistn + op->hook_par[0];
        ra.hcnt_f_next = op->d_hlistn + (1 - op->hook_par[0]);
        ra.hcnt_l = op->d_hlistn + 2 + op->hook_par[1];
        ra.hcnt_l_next = op->d_hlistn + 2 + (1 - op->hook_par[1]);
    }
    if constexpr (!MIX) {
        if (op->d_ngeom) {
            // general mesh: per-node metrics and oriented faces (boundary faces included)
            if (op->opts.blending) ra.blend = make_blend_params<NP>(op);
            using LC = NodeCfg<DIM, NP>;
            const int nblk = (op->NE + LC::EPB - 1) / LC::EPB;
            op->lam_waves = nblk * LC::BLOCK / 64;
            // boundary faces only where the mesh has them (the boundary path costs registers)
            if (op->opts.blending && op->nbc)
                hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, true, true, true>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
            else if (op->opts.blending)
                hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, true, false, true>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
            else if (op->nbc)
                hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, false, true, true>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
            else
                hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, false, false, true>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
            HIP_TRY(hipGetLastError());
            return PRANDTL_OK;
        }
        if constexpr (VISC && DIM == 3) {
            if (op->opts.bug_compat & PRANDTL_BUG_A1) {
                // the reference's zeta flux store (affine, no boundary faces: checked at create)
                if (op->opts.blending) ra.blend = make_blend_params<NP>(op);
                using LC = NodeCfg<DIM, NP>;
                const int nblk = (op->NE + LC::EPB - 1) / LC::EPB;
                op->lam_waves = nblk * LC::BLOCK / 64;
                if (op->opts.blending)
                    hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, true, false, false, true>), dim3(nblk),
                                       dim3(LC::BLOCK), 0, st, ra);
                else
                    hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, false, false, false, true>), dim3(nblk),
                                       dim3(LC::BLOCK), 0, st, ra);
                HIP_TRY(hipGetLastError());
                return PRANDTL_OK;
            }
        }
        if (op->opts.blending || op->nbc) {
            // blending: the kernel computes this evaluation's coefficients from its input state
            // (DGSEMOperator.cpp:547-549) and blends; boundary faces: the conditions' fluxes.
            // Both run the node kernel (for Euler too).
            if (op->opts.blending) ra.blend = make_blend_params<NP>(op);
            using LC = NodeCfg<DIM, NP>;
            const int nblk = (op->NE + LC::EPB - 1) / LC::EPB;
            op->lam_waves = nblk * LC::BLOCK / 64;
            if (op->opts.blending && op->nbc)
                hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, true, true>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
            else if (op->opts.blending)
                hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, true, false>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
            else
                hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, false, false, true>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
            HIP_TRY(hipGetLastError());
            return PRANDTL_OK;
        }
    }
    // Navier-Stokes and mixed precision: one thread per node (the node's state stays in registers
    // through the fused epilogue, no LDS accumulator: 4 waves/SIMD at TGV 32^3 p=3).
    // Euler: the node kernel where its lines lie in one wave (each symmetric pair evaluated once
    // through cross-lane shuffles; KHI 256^2 p=3: 38.2 against 47.7 us per stage for the line
    // kernel), and on small meshes (fewer line-kernel workgroups than two per CU: the launch is
    // latency-bound there); otherwise the line kernel (each pair once, without shuffles)
    constexpr int kSmallGrid = 512;
    constexpr bool wave_lines = NodeCfg<DIM, NP>::BLOCK == 64 || NodeCfg<DIM, NP>::NN == 64;
    const bool small = (op->NE + LineCfg<DIM, NP>::EPB - 1) / LineCfg<DIM, NP>::EPB < kSmallGrid;
    if (VISC || MIX || wave_lines || small) {
        using LC = NodeCfg<DIM, NP>;
        const int nblk = (op->NE + LC::EPB - 1) / LC::EPB;
        op->lam_waves = nblk * LC::BLOCK / 64;
        hipLaunchKernelGGL((k_rhs_node<DIM, NP, VISC, SURF, MIX>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
    } else {
        using LC = LineCfg<DIM, NP>;
        const int nblk = (op->NE + LC::EPB - 1) / LC::EPB;
        op->lam_waves = nblk * LC::BLOCK / 64;
        hipLaunchKernelGGL((k_rhs<DIM, NP, VISC, SURF>), dim3(nblk), dim3(LC::BLOCK), 0, st, ra);
    }
    HIP_TRY(hipGetLastError());
    return PRANDTL_OK;
}

template <int DIM, int NP, bool VISC, bool MIX>
int dispatch_surf(prandtl_op_s *op, const StageSpec &S, hipStream_t st) {
    switch (op->opts.surface_flux) {
        case PRANDTL_SURF_CHANDRASHEKAR: return launch_stage<DIM, NP, VISC, SURF_CHANDRASHEKAR, MIX>(op, S, st);
        case PRANDTL_SURF_EC: return launch_stage<DIM, NP, VISC, SURF_EC, MIX>(op, S, st);
        case PRANDTL_SURF_RUSANOV: return launch_stage<DIM, NP, VISC, SURF_RUSANOV, MIX>(op, S, st);
    }
    return fail(PRANDTL_ERR_ARG, "unknown surface flux");
}


// one (dim, Np) pair of the stage launcher: defined and instantiated in stage_kernels.hip
template <int DIM, int NP>
int dispatch_visc(prandtl_op_s *op, const StageSpec &S, hipStream_t st);

}  // namespace prandtl
