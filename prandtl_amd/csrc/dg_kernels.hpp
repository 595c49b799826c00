// dg_kernels.hpp -- DGSEM right-hand-side kernels for gfx950 (CDNA4).
//
// HBM layout (the reference's byNODES layout, src/Simulation/Simulation.hpp:64):
//   u[v][e][node], node = i + Np*(j + Np*k) (DGSEMIntegrator.cpp:250).
// Every kernel reads and writes whole element blocks of each variable, so a wave's
// global accesses are contiguous runs of Np^d doubles per variable.
//
// Work decomposition of the flux-differencing kernel (k_rhs): a workgroup owns EPB
// elements; thread t owns GLL line (t % L) of element (t / L), L = Np^(d-1).  For each
// direction the thread loads its line's Np node primitives from LDS, evaluates every
// symmetric pair flux once (Np(Np-1)/2 per line: 4.5 pair fluxes per node at p=3 in
// 3D, the reference's count, DGSEMIntegrator.cpp:260-300) and both surface fluxes at
// the line's ends, and adds its line's contributions into an LDS accumulator that
// only it touches for that direction.  No atomics: results are bitwise reproducible.
//
// Navier-Stokes runs two grid passes (BR1 needs every gradient before any viscous
// face flux): k_lift computes entropy variables, the BR1 gradient (volume + face
// jump), the primitive-gradient chain rule and the viscous flux per node, stages the
// contravariant viscous fluxes F~^d = F_v . Ja^d in LDS, and writes (a) the viscous
// volume term (1/J) sum_d D-hat F~^d (neq-1 values per node) and (b) F~^d at the
// element's face points in a face-trace layout [d][side][v][e][t] (t = the point's
// index on the face), so k_rhs reads both sides of every face coalesced.  On a
// conforming affine mesh F~^d at a face point is F_v . n for that face.
#pragma once

#include <type_traits>

#include "dg_adjoint.hpp"
#include "dg_bc.hpp"
#include "dg_physics.hpp"

#ifndef PRANDTL_LINE_WAVES
#define PRANDTL_LINE_WAVES 2
#endif
#ifndef PRANDTL_NODE_WAVES
#define PRANDTL_NODE_WAVES 4
#endif
#ifndef PRANDTL_LIFT_MFMA
#define PRANDTL_LIFT_MFMA 1
#endif
#ifndef PRANDTL_LIFT_WAVES
#define PRANDTL_LIFT_WAVES 4   // 128 VGPRs: at 3D p=3 FP64, 4 waves/SIMD beat 3 without the cap
#endif


namespace prandtl {

// the adjoint differentiates the log mean on the branches the forward takes
static_assert(adj::kLogMeanSeriesU == kLmLongU, "the forward and the adjoint switch to the closed form at one u");

constexpr int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; i++) r *= b;
    return r;
}

// Smallest block of 64, 128 or 256 threads whose whole-element packing uses >= 90 % of it.
constexpr int choose_block(int L) {
    for (int B = 64; B <= 256; B *= 2)
        if ((B / L) * L * 10 >= B * 9) return B;
    return 256;
}

template <int DIM, int NP>
struct LineCfg {
    static constexpr int NN = ipow(NP, DIM);      // nodes per element
    static constexpr int L = ipow(NP, DIM - 1);   // lines per direction
    static constexpr int BLOCK = choose_block(L);
    static constexpr int EPB = BLOCK / L;         // elements per workgroup
};

template <int DIM, int NP>
struct NodeCfg {
    static constexpr int NN = ipow(NP, DIM);
    static constexpr int BLOCK = NN <= 32 ? 64 : (NN <= 64 ? 128 : 64 * ((NN + 63) / 64));
    static constexpr int EPB = BLOCK / NN;
};

template <int NP>
struct BlendParams {
    double Vi[NP][NP];     // V1^-1: Vi[k][a] = w_a P_k(x_a) / sum_b w_b P_k(x_b)^2
    double T;              // modalThreshold 0.5 10^(-1.8 p^(1/4))
    double sharpness;      // 9.21024
    double alpha_max, alpha_min;
};

template <int NP>
struct DTables {
    double D[NP][NP];      // D_ij (DGSEMIntegrator.cpp:45-69)
    double Dhat[NP][NP];   // D + e0 e0^T / w0 - eN eN^T / wN  (:71-74)
    double Dhat2[NP][NP];  // 2 D + the same corrections      (:76-80)
    double inv_w0;         // 1 / w_0
    double inv_w[NP];      // 1 / w_a (subcell finite-volume term)
    // subcell metric of interface i+1 of a line (between its nodes i and i+1) on a general
    // element: sum_m sub[i][m] adj_d(node m), sub[i][m] = delta_m0 + sum_{l<=i} w_l D_lm
    // (ComputeSubcellMetrics, DGSEMIntegrator.cpp:1199-1304); rows 0..Np-2
    double sub[NP][NP];
};

enum EpilogueMode { EPI_MULT = 0, EPI_LSRK = 1, EPI_SSP = 2 };

template <int NP>
struct RhsArgs {
    DTables<NP> T;
    PhysConst P;
    int NE;
    int bug3;
    int mode;
    int want_lambda;
    const int *nbr;
    const double *geom;
    const double *u;       // stage input state
    const double *dv;      // NS: viscous volume term [v-1][e][node]
    const double *Ff;      // NS: viscous face traces F~^d [d][side][v-1][e][t]
    const float *dvf;      // the same, FP32 (mixed precision)
    const float *Fff;
    double *out;           // EPI_MULT: dudt; otherwise the stage output state
    double *k;             // EPI_LSRK: 2N register
    const double *x;       // EPI_SSP: base state of the Shu-Osher combination
    double ca, cb, dt;
    double *lam_part;      // want_lambda: each wave's max |V| + c, at its global wave index
    int *nan_flag;
    const BcDev *bc;       // boundary conditions (neighbour -(k+1) -> bc[k]; k_rhs_node<..., BND = true>)
    const double *bxyz;    // coordinates of the face points [e][2 dim][Np^(dim-1)][dim] (function-valued
                           // boundary states; null otherwise)
    double time;           // the time of the boundary states
    double *alpha;         // k_rhs_node<..., BLEND = true>: the coefficients it computed (may be null)
    BlendParams<NP> blend; // the same: Persson-Peraire indicator and blending constants
    const double *ngeom;   // general mesh (k_rhs_node<..., GEN = true>): per-node (det J, adj J) [e][node][1+dim^2]
    const int *forient;    // general mesh: face orientation codes [e][2 dim]
    // stage hooks detected in the epilogue of a stage update (k_rhs_node, elements within a wave):
    int hook_mode;         // HOOK_DETECT_* bits
    const double *hbound;  // the entropy filter's bound per element
    double *hemin;         // minimum nodal entropy of each element of the stage output
    int *hlist_f;          // elements the filter must visit, and their counter (the other counter,
    unsigned *hcnt_f;      // hcnt_f_next, is cleared for the next application)
    unsigned *hcnt_f_next;
    int *hlist_l;          // elements the limiter must visit (the filter off), and counters
    unsigned *hcnt_l;
    unsigned *hcnt_l_next;
};

enum HookDetect { HOOK_DETECT_FILTER = 1, HOOK_DETECT_LIMITER = 2, HOOK_DETECT_EMIN = 4 };

template <int NP>
struct LiftArgs {
    DTables<NP> T;
    PhysConst P;
    int NE;
    const int *nbr;
    const double *geom;
    const double *u;
    double *dv;            // out: viscous volume term [v-1][e][node]
    double *Ff;            // out: viscous face traces [d][side][v-1][e][t]
    float *dvf;            // out, FP32 (mixed precision)
    float *Fff;
    double *grad;          // optional out: primitive gradients [c][v][e][node]
    const BcDev *bc;       // boundary conditions (neighbour -(k+1) -> bc[k])
    const double *bxyz;    // face-point coordinates for function-valued boundary states (as RhsArgs)
    double time;
    const double *ngeom;   // general mesh (k_lift<..., GEN = true>): per-node (det J, adj J)
    const int *forient;    // general mesh: face orientation codes
};

// Workgroups are dispatched round-robin over the 8 XCDs; remap so each XCD walks one
// contiguous range of elements (face neighbours then share that XCD's L2).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

// xcd_block with each XCD's range walked from its end.  The lifting pass runs between two flux
// passes, which walk the ranges forward; walking backwards it starts on the state the previous
// flux pass wrote last, and leaves the flux pass its freshest outputs (dv, Ff) to start on: the
// arrays of a stage outgrow the 256 MB MALL at the sizes that matter, so what was touched last
// is what is still there.
__device__ __forceinline__ int xcd_block_rev(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, i = b >> 3;
    const int qx = q + (x < r ? 1 : 0);   // element blocks of XCD x
    return x * q + (x < r ? x : r) + (qx - 1 - i);
}

// node index of point a on line l of direction dir
template <int NP>
__device__ __forceinline__ constexpr int line_base(int l, int dir) {
    const int st = dir == 0 ? 1 : (dir == 1 ? NP : NP * NP);
    return (l % st) + (l / st) * st * NP;
}

// line_base for a direction known only at run time: the compile-time-stride forms, selected
// (a division by a run-time stride would be a generic integer division)
template <int DIM, int NP>
__device__ __forceinline__ int line_base_rt(int l, int dir) {
    const int b0 = line_base<NP>(l, 0), b1 = line_base<NP>(l, 1);
    if constexpr (DIM == 2) {
        return dir == 0 ? b0 : b1;
    } else {
        const int b2 = line_base<NP>(l, 2);
        return dir == 0 ? b0 : (dir == 1 ? b1 : b2);
    }
}

// LDS slot of node n (0..63, n = i + 4 j + 16 k, 3D p = 3) of a one-wave element: k XORed into
// the xi bits and into the eta bits, a permutation of each 16-node k-plane.  With 64 lanes reading
// FP64 values, each of these patterns then hits distinct LDS banks within every 16 lanes (32 banks
// of 4 bytes) and every 32 lanes (64 banks): the own node, a line partner in any direction, a
// face's nodes (one index fixed at 0 or 3), the node with its xi and zeta (or eta and zeta)
// indices exchanged.  In node order the last two are 2- to 4-way conflicts.
__device__ __forceinline__ int node_swz(int n) { return n ^ (5 * ((n >> 4) & 3)); }

template <int NP>
__device__ __forceinline__ constexpr int dir_stride(int dir) {
    return dir == 0 ? 1 : (dir == 1 ? NP : NP * NP);
}

// index t of the line (of direction dir) through a node: the inverse of line_base
template <int NP>
__device__ __forceinline__ int line_index(int node, int dir) {
    const int st = dir_stride<NP>(dir);
    return (node % st) + (node / (st * NP)) * st;
}

// offset of face-trace value (dir, side, v, e, t)
template <int DIM, int NP>
__device__ __forceinline__ size_t ftrace(int dir, int side, int v, int e, int t, int NE) {
    constexpr int L = ipow(NP, DIM - 1);
    return ((((size_t)(dir * 2 + side) * (DIM + 1) + v) * NE + e) * L) + t;
}

// General mesh: the neighbour's face point tn (face coordinates u1 + Np u2) of the point
// t = t1 + Np t2 of a face with orientation code `code` (bits 0-2 the neighbour's local face,
// bit 3 flips the first face coordinate, bit 4 the second, bit 5 swaps them).
template <int DIM, int NP>
__device__ __forceinline__ int face_point_map(int code, int t) {
    const int t1 = t % NP, t2 = t / NP;
    const bool sw = (code & 32) != 0;
    const int a = sw ? t2 : t1, b = sw ? t1 : t2;
    const int u1 = (code & 8) ? NP - 1 - a : a;
    const int u2 = (code & 16) ? NP - 1 - b : b;
    return DIM == 2 ? u1 : u1 + NP * u2;
}

// Streamed (non-temporal) access for the arrays a kernel touches once: in the 3D kernels, whose
// stage arrays (the state, the RK register, the viscous terms) together exceed the 256 MB MALL at
// the sizes they run, streaming the outputs keeps their dirty lines from delaying the next
// kernel's reads; in 2D the arrays of the cases here fit the MALL, and cached stores let the next
// kernel find them there.  A compile-time choice: a runtime flag could let the compiler merge the
// two forms of a store and drop the hint.
template <bool STREAM>
__device__ __forceinline__ void put(double *p, double v) {
    if constexpr (STREAM)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
template <bool STREAM>
__device__ __forceinline__ double take(const double *p) {
    if constexpr (STREAM)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Stage epilogue: writes the RHS (Mult) or applies the fused RK stage update to one value;
// x is the RK register (LSRK) or the base state of the Shu-Osher combination (SSP).  The outputs
// are streamed in 3D: the next kernel (the lifting pass) reads the new state once, and the RK
// register is read again only a stage later.
template <int DIM, int NP>
__device__ __forceinline__ double stage_update_x(const RhsArgs<NP> &A, size_t gi, double u_in, double du, double x,
                                                 int &bad) {
    constexpr bool S = DIM == 3;
    if (A.mode == EPI_MULT) {
        put<S>(&A.out[gi], du);
        return du;
    } else if (A.mode == EPI_LSRK) {
        const double kk = A.ca * x + A.dt * du;
        put<S>(&A.k[gi], kk);
        const double un = u_in + A.cb * kk;
        put<S>(&A.out[gi], un);
        bad |= !(un == un);
        return un;
    } else {
        const double un = A.ca * x + A.cb * (u_in + A.dt * du);
        put<S>(&A.out[gi], un);
        bad |= !(un == un);
        return un;
    }
}

template <int DIM, int NP>
__device__ __forceinline__ double stage_update(const RhsArgs<NP> &A, size_t gi, double u_in, double du, int &bad) {
    const double x = A.mode == EPI_LSRK ? take<DIM == 3>(&A.k[gi]) : (A.mode == EPI_SSP ? A.x[gi] : 0.0);
    return stage_update_x<DIM, NP>(A, gi, u_in, du, x, bad);
}

// Stage hooks detected on the stage output s (one node per lane, an element's NN nodes in
// consecutive lanes of one wave, its first node at `node` 0): the entropy filter's detection
// against the element's bound (EntropyFilter.cpp:183-191), the limiter's trigger, the minimum nodal
// entropy (the next step's filter constraints when this is a step's last stage).  Per element:
// a segmented tree reduction over its lanes; the elements found are appended to the hooks' lists
// with one atomic per wave.  Every lane of the wave calls this.
template <int DIM, int NP, int NN>
__device__ __forceinline__ void stage_hook_detect(const RhsArgs<NP> &A, bool active, int e, int node,
                                                  const double *s) {
    const int lane = threadIdx.x & 63;
    double ent = __builtin_inf();
    int fv = 0, lv = 0;
    if (active) {
        const double p = hook_pressure<DIM>(s, A.P.gammaM1);
        if (A.hook_mode & HOOK_DETECT_EMIN) {
            ent = hook_entropy(s[0], p, A.P.gamma);
            if (A.hook_mode & HOOK_DETECT_FILTER)
                fv = (s[0] < kFilterRhoMin || p < kFilterPMin || ent < A.hbound[e] - kFilterEntTol) ? 1 : 0;
        } else if (A.hook_mode & HOOK_DETECT_FILTER) {
            fv = (s[0] < kFilterRhoMin || p < kFilterPMin) ? 1 : 0;
            if (!fv) {
                // the entropy test: an FP32 estimate decides wherever it lies clearly above the
                // bound (rho, p > 0 here); near it, the FP64 entropy decides, as the scan does
                const double thr = A.hbound[e] - kFilterEntTol;
                const float rf = (float)s[0], pf = (float)p;
                const float lr = __builtin_amdgcn_logf(rf) * 0.69314718f, lp = __builtin_amdgcn_logf(pf) * 0.69314718f;
                const double est = (double)rf * ((double)lp - A.P.gamma * (double)lr);
                const double err = 4e-6 * s[0] * (1.0 + fabs((double)lp) + A.P.gamma * fabs((double)lr));
                if (!(est - err >= thr)) fv = hook_entropy(s[0], p, A.P.gamma) < thr ? 1 : 0;
            }
        }
        if (A.hook_mode & HOOK_DETECT_LIMITER) lv = (s[0] < kLimiterEps || p < kLimiterEps) ? 1 : 0;
    }
#pragma unroll
    for (int off = 1; off < NN; off *= 2) {
        const double eo = __shfl(ent, lane + off, 64);
        const int fo = __shfl(fv, lane + off, 64);
        const int lo = __shfl(lv, lane + off, 64);
        if (node + off < NN) {
            ent = fmin(ent, eo);
            fv |= fo;
            lv |= lo;
        }
    }
    const bool head = active && node == 0;
    if (head && (A.hook_mode & HOOK_DETECT_EMIN)) A.hemin[e] = fmin(1.7976931348623157e308, ent);   // DBL_MAX start
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the counters the previous application of each hook read are free for the next one
        if (A.hook_mode & HOOK_DETECT_FILTER) *A.hcnt_f_next = 0u;
        if (A.hook_mode & HOOK_DETECT_LIMITER) *A.hcnt_l_next = 0u;
    }
    auto append = [&](bool mine, int *list, unsigned *cnt) {
        const unsigned long long m = __ballot(mine);
        if (m == 0ull) return;
        const int leader = __ffsll((long long)m) - 1;
        unsigned base = 0u;
        if (lane == leader) base = atomicAdd(cnt, (unsigned)__popcll(m));
        base = __shfl(base, leader, 64);
        if (mine) {
            const unsigned slot = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
            if (slot < (unsigned)A.NE) list[slot] = e;
        }
    };
    if (A.hook_mode & HOOK_DETECT_FILTER) append(head && fv, A.hlist_f, A.hcnt_f);
    if (A.hook_mode & HOOK_DETECT_LIMITER) append(head && lv, A.hlist_l, A.hcnt_l);
}

template <int NP>
__device__ __forceinline__ void finish_block(const RhsArgs<NP> &A, double lam, int bad) {
    if (A.want_lambda) {
        // one plain store per wave (a single reduction kernel follows): atomics on one address
        // from every wave of the grid serialise
        for (int o = 32; o > 0; o >>= 1) lam = fmax(lam, __shfl_xor(lam, o));
        if ((threadIdx.x & 63) == 0) A.lam_part[((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6] = lam;
    }
    if (A.nan_flag && __any(bad)) {
        if ((threadIdx.x & 63) == 0) atomicOr(A.nan_flag, 1);
    }
}

}  // namespace prandtl

#include "dg_blend.hpp"

namespace prandtl {

// ---------------------------------------------------------------------------
// k_rhs (line variant): inviscid flux differencing + surface flux (+ viscous divergence
// and the viscous face average for NS) + a fused stage epilogue.
// ref: DGSEMIntegrator::AssembleElementVector (Euler :224-433, NS 3D :601-750),
//      AssembleFaceVector (Euler :151-222, NS 3D :435-494, NS 2D :496-549),
//      DGSEMNonlinearForm::Mult (:647-849, :1302-1545).
// ---------------------------------------------------------------------------
template <int DIM, int NP, bool VISC, int SURF>
__global__ void __launch_bounds__((LineCfg<DIM, NP>::BLOCK), PRANDTL_LINE_WAVES) k_rhs(const RhsArgs<NP> A) {
    using C = LineCfg<DIM, NP>;
    constexpr int NEQ = DIM + 2, NN = C::NN, L = C::L, EPB = C::EPB, NQ = DIM + 2;
    __shared__ double sP[NQ][EPB * NN];   // rho, vel[DIM], beta
    __shared__ double sD[NEQ][EPB * NN];  // dudt accumulator

    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int e0 = blk * EPB;
    const int tid = threadIdx.x;
    const int NE = A.NE;
    const size_t vs = (size_t)NE * NN;  // variable stride
    const size_t off0 = (size_t)e0 * NN;
    const int nvalid = min(EPB, NE - e0) * NN;

    double lam = 0.0;
    for (int idx = tid; idx < nvalid; idx += C::BLOCK) {
        double s[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) s[v] = A.u[v * vs + off0 + idx];
        const SideState<DIM> q = make_side<DIM>(s, A.P);
        sP[0][idx] = q.rho;
#pragma unroll
        for (int d = 0; d < DIM; d++) sP[1 + d][idx] = q.vel[d];
        sP[DIM + 1][idx] = q.beta;
        lam = fmax(lam, q.lam);
    }
    __syncthreads();

    const int eb = tid / L, l = tid - (tid / L) * L;
    const int e = e0 + eb;
    const bool active = (eb < EPB) && (e < NE);
    double invJ = 0.0, adj[DIM][DIM];
    if (active) {
        const double *g = A.geom + (size_t)e * (1 + DIM * DIM);
        invJ = rcp(g[0]);
#pragma unroll
        for (int r = 0; r < DIM; r++)
#pragma unroll
            for (int c = 0; c < DIM; c++) adj[r][c] = g[1 + r * DIM + c];
    }
    const double sw = A.T.inv_w0 * invJ;

#pragma unroll
    for (int dir = 0; dir < DIM; dir++) {
        if (active) {
            const int base = line_base<NP>(l, dir);
            const int st = dir_stride<NP>(dir);
            const int lb = eb * NN + base;
            NodePrim<DIM> q[NP];
#pragma unroll
            for (int a = 0; a < NP; a++) {
                const int li = lb + a * st;
                q[a].rho = sP[0][li];
#pragma unroll
                for (int d = 0; d < DIM; d++) q[a].vel[d] = sP[1 + d][li];
                q[a].beta = sP[DIM + 1][li];
            }
            double acc[NP][NEQ];
#pragma unroll
            for (int a = 0; a < NP; a++)
#pragma unroll
                for (int v = 0; v < NEQ; v++) acc[a][v] = 0.0;
            const double *m = adj[dir];
#pragma unroll
            for (int a = 0; a < NP; a++)
#pragma unroll
                for (int b = a + 1; b < NP; b++) {
                    double F[NEQ];
                    volume_flux<DIM>(q[a], q[b], m, A.P, F);
#pragma unroll
                    for (int v = 0; v < NEQ; v++) {
                        acc[a][v] += A.T.Dhat2[a][b] * F[v];
                        acc[b][v] += A.T.Dhat2[b][a] * F[v];
                    }
                }
#pragma unroll
            for (int a = 0; a < NP; a++)
#pragma unroll
                for (int v = 0; v < NEQ; v++) acc[a][v] = -acc[a][v] * invJ;

            double nmag = m[0] * m[0];
#pragma unroll
            for (int d = 1; d < DIM; d++) nmag += m[d] * m[d];
            nmag = sqrt(nmag);
            // surface terms at both ends of the line
#pragma unroll
            for (int side = 1; side >= 0; side--) {
                // side 1: this element is el1 (outward normal Ja^dir), the neighbour el2;
                // side 0: the neighbour is el1, this element el2
                const int en = A.nbr[(size_t)e * 2 * DIM + 2 * dir + side];
                const int nown = base + (side ? (NP - 1) * st : 0);
                const int nnb = base + (side ? 0 : (NP - 1) * st);
                double so[NEQ], sn[NEQ], f[NEQ];
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    so[v] = A.u[v * vs + (size_t)e * NN + nown];
                    sn[v] = A.u[v * vs + (size_t)en * NN + nnb];
                }
                const SideState<DIM> qo = make_side<DIM>(so, A.P);
                const SideState<DIM> qn = make_side<DIM>(sn, A.P);
                if (side)
                    face_flux<DIM, SURF>(so, qo, sn, qn, m, nmag, A.P, A.bug3 != 0, f);
                else
                    face_flux<DIM, SURF>(sn, qn, so, qo, m, nmag, A.P, A.bug3 != 0, f);
                if constexpr (VISC) {
#pragma unroll
                    for (int v = 0; v < NEQ - 1; v++)
                        f[1 + v] -= 0.5 * (A.Ff[ftrace<DIM, NP>(dir, side, v, e, l, NE)] +
                                           A.Ff[ftrace<DIM, NP>(dir, 1 - side, v, en, l, NE)]);
                }
                const double sg = side ? -sw : sw;
#pragma unroll
                for (int v = 0; v < NEQ; v++) acc[side ? NP - 1 : 0][v] += f[v] * sg;
            }
#pragma unroll
            for (int a = 0; a < NP; a++) {
                const int li = lb + a * st;
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    if (dir == 0)
                        sD[v][li] = acc[a][v];
                    else
                        sD[v][li] += acc[a][v];
                }
            }
        }
        __syncthreads();
    }

    // epilogue: one thread per node, coalesced per variable
    int bad = 0;
    for (int idx = tid; idx < nvalid; idx += C::BLOCK) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            const size_t gi = v * vs + off0 + idx;
            double du = sD[v][idx];
            if (VISC && v > 0) du += A.dv[(v - 1) * vs + off0 + idx];
            stage_update<DIM, NP>(A, gi, A.mode == EPI_MULT ? 0.0 : A.u[gi], du, bad);
        }
    }
    finish_block<NP>(A, lam, bad);
}

// ---------------------------------------------------------------------------
// k_rhs_node (node variant): the same operator with one thread per node.  Each node
// evaluates the two-point flux with its Np-1 line partners in every direction (the
// symmetric pair is evaluated by both of its nodes), which needs no LDS accumulator
// and keeps the node's own state in registers through the fused epilogue.  The surface
// fluxes of the element's 2 dim Np^(dim-1) face points are work items spread over all
// threads of the workgroup (a face node's thread would otherwise run them while the
// interior nodes of its wave idle), with the element's own traces and metric from LDS and
// the neighbour table read once per workgroup; the face nodes read their fluxes from LDS.
//
// BLEND (subcell finite-volume blending, FP64): the inviscid volume term becomes
// (1 - alpha) DG + alpha FV (ref: DGSEMIntegrator.cpp:240-241, 394-396, 741-742, and
// ComputeFVFluxes :972-1177), alpha per element from k_blend_coef.  FV: along every line the
// interface between nodes i and i+1 carries the surface flux f*(u_i, u_{i+1}; adj_d) (on an
// affine element the subcell metric of ComputeSubcellMetrics, :1180-1304, is the adj row);
// node i receives (f_{i-1/2} - f_{i+1/2}) / (J w_i), the line's outer interfaces excluded.
// The interface fluxes are work items spread over the workgroup one direction at a time,
// staged in the surface-flux LDS buffer after the DG accumulation has read it; workgroups
// whose elements all have alpha = 0 skip them.
//
// GEN (general meshes, FP64): per-node metrics (det J, adj J at every node, as the reference
// evaluates Tr.Weight() and Tr.AdjugateJacobian() per node, DGSEMIntegrator.cpp:254-276) and
// faces that meet their neighbour's face in any direction, flipped or swapped (an orientation
// code per face).  A pair flux takes the averaged metric 1/2 (adj_d(i) + adj_d(j)) (ComputeMean,
// ChandrashekarFlux.cpp:15); a face flux is evaluated as the reference's face integrator does,
// once per face with the normal of its lower-numbered element (CalcOrtho: that element's adj row
// at the face point, outward), by both elements' work items; the FV subcell metrics come from
// the line's per-node adj rows (the DTables sub weights).
// ---------------------------------------------------------------------------
//
// BUG1 (bug_compat A1, 3D Navier-Stokes on affine FP64 meshes): the reference's NS element stores,
// for the zeta pair of node (i, j, k) with its partner (i, j, m > k), the last eta flux of node
// (i, j, k) in the partner's slot (H_inviscid(id2).SetCol(k, g), DGSEMIntegrator.cpp:702): the
// flux between (i, j, k) and (i, Np-1, k) with the eta metric, or 0 when j = Np-1 (g is zeroed
// before the eta loop).  A node's lower-zeta partners then contribute that flux instead.
template <int DIM, int NP, bool VISC, int SURF, bool MIX, bool BLEND = false, bool BND = false, bool GEN = false,
          bool BUG1 = false>
__global__ void __launch_bounds__((NodeCfg<DIM, NP>::BLOCK), PRANDTL_NODE_WAVES) k_rhs_node(const RhsArgs<NP> A) {
    static_assert(!(BUG1 && (DIM != 3 || !VISC || MIX || BND || GEN)), "bug_compat A1: 3D NS, affine, FP64");
    using C = NodeCfg<DIM, NP>;
    using R = typename std::conditional<MIX, float, double>::type;  // flux arithmetic
    static_assert(!(MIX && BLEND), "subcell blending is FP64 only");
    static_assert(!(MIX && BND), "boundary faces run in FP64");
    static_assert(!(MIX && GEN), "general meshes run in FP64");
    constexpr int NEQ = DIM + 2, NN = C::NN, EPB = C::EPB;
    __shared__ double sPd[2][EPB * NN];   // rho, beta
    __shared__ R sPv[DIM][EPB * NN];      // velocity
    __shared__ R sT[NP * NP];             // Dhat2
    __shared__ double sPref[EPB];         // element reference pressure (mixed precision)
    __shared__ double sHref[EPB];         // element reference enthalpy gamma p / ((gamma-1) rho)
    __shared__ int sNbr[EPB][2 * DIM];    // face neighbours of the workgroup's elements
    constexpr int L = ipow(NP, DIM - 1), FITEMS = 2 * DIM * L;
    // conserved state (GEN without blending reads the element's own traces from global memory
    // instead: the LDS it saves keeps 4 waves per SIMD in 3D)
    constexpr bool SSTATE = !(GEN && !BLEND);
    __shared__ double sS[NEQ][SSTATE ? EPB * NN : 1];
    __shared__ R sM[GEN ? 1 : EPB][DIM][DIM];              // adj J rows (affine elements)
    __shared__ double sMr[GEN ? EPB * NN : 1][DIM];        // GEN: adj J row of the current direction, per node
    __shared__ int sOr[GEN ? EPB : 1][2 * DIM];            // GEN: face orientation codes
    constexpr int G1 = 1 + DIM * DIM;                      // per-node geometry record (GEN)
    __shared__ R sF[EPB * FITEMS][NEQ];   // surface flux (with the viscous face average) per face point
    constexpr int SITEMS = (NP - 1) * L;  // subcell interfaces per element and direction
    static_assert(SITEMS <= FITEMS, "subcell fluxes are staged in the surface-flux buffer");
    static_assert(!BLEND || EPB * FITEMS * NEQ >= 3 * EPB * NN, "the coefficient scratch fits the surface-flux buffer");
    __shared__ double sVi[BLEND ? NP * NP : 1];     // V1^-1 of the blending indicator
    __shared__ double sAlpha[BLEND ? EPB : 1][2];   // alpha, eta
    // affine FP64 interior faces take the element's own side from LDS: the primitives above and
    // each node's max wave speed, and each face direction's |adj row|
    constexpr bool LSIDE = !GEN && !MIX;
    __shared__ double sLam[LSIDE ? EPB * NN : 1];
    __shared__ R sNm[GEN ? 1 : EPB][DIM];
    // volume pairs: each symmetric pair evaluated once when the lines of an element lie in one
    // wave (the partner's value arrives by a cross-lane shuffle)
    constexpr bool ROT = !GEN && !BUG1 && (C::BLOCK == 64 || NN == 64);
    constexpr bool ROTG = GEN && (C::BLOCK == 64 || NN == 64);   // the same on general meshes
    // even Np: the Np/2-apart pairs of directions 0 and 1 share one pass (half of each direction's
    // lanes would otherwise evaluate their partner's pair again)
    constexpr bool HALF2 = ROT && NP % 2 == 0;
    // the node arrays (sS, sPd, sPv, sLam) stay in node order: swizzled by node_swz, the face
    // items' own-side reads lose their bank conflicts (4.3 M -> 1.0 M conflict cycles per launch at
    // TGV 32^3 p = 3), but the address arithmetic costs a 12-byte spill at 128 VGPRs, 6 % per launch

    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x;
    const int eb = tid / NN, node = tid - (tid / NN) * NN;
    const int e = blk * EPB + eb;
    const int NE = A.NE;
    const bool active = (eb < EPB) && (e < NE);
    const size_t vs = (size_t)NE * NN;
    if (tid < NP * NP) sT[tid] = (R)A.T.Dhat2[tid / NP][tid % NP];
    if constexpr (BLEND) {
        if (tid < NP * NP) sVi[tid] = A.blend.Vi[tid / NP][tid % NP];
    }
    // the neighbour table rows and metric rows of the workgroup's elements, read once here
    // rather than by every face item (EPB (2 dim + dim^2) entries can exceed the block size)
    for (int i = tid; i < EPB * 2 * DIM; i += C::BLOCK) {
        const int ib = i / (2 * DIM), f = i - ib * 2 * DIM;
        const int ee = min(blk * EPB + ib, NE - 1);
        sNbr[ib][f] = A.nbr[(size_t)ee * 2 * DIM + f];
    }
    if constexpr (GEN) {
        for (int i = tid; i < EPB * 2 * DIM; i += C::BLOCK) {
            const int ib = i / (2 * DIM), f = i - ib * 2 * DIM;
            const int ee = min(blk * EPB + ib, NE - 1);
            sOr[ib][f] = A.forient[(size_t)ee * 2 * DIM + f];
        }
    } else {
        for (int i = tid; i < EPB * DIM * DIM; i += C::BLOCK) {
            const int ib = i / (DIM * DIM), r = i - ib * DIM * DIM;
            const int ee = min(blk * EPB + ib, NE - 1);
            sM[ib][r / DIM][r % DIM] = (R)A.geom[(size_t)ee * (1 + DIM * DIM) + 1 + r];
        }
        for (int i = tid; i < EPB * DIM; i += C::BLOCK) {
            const int ib = i / DIM, d = i - ib * DIM;
            const int ee = min(blk * EPB + ib, NE - 1);
            const double *g = A.geom + (size_t)ee * (1 + DIM * DIM) + 1 + d * DIM;
            R nm = (R)g[0] * (R)g[0];
#pragma unroll
            for (int c = 1; c < DIM; c++) nm += (R)g[c] * (R)g[c];
            sNm[ib][d] = sqrt(nm);
        }
    }
    double lam = 0.0;
    double qrp = 0.0;   // rho p at the node (BLEND)
    double detJ = 0.0;  // det J at the node (GEN)
    if constexpr (GEN) {
        if (active) detJ = A.ngeom[((size_t)e * NN + node) * G1];
    }
    // the conserved state at node nd of the workgroup's element ib
    auto own_state = [&](int v, int ib, int nd) -> double {
        if constexpr (SSTATE)
            return sS[v][ib * NN + nd];
        else
            return A.u[v * vs + (size_t)(blk * EPB + ib) * NN + nd];
    };
    // GEN: the adj J row `dir` of every node of the workgroup into sMr (callers synchronise)
    auto load_rows = [&](int dir) {
        if constexpr (GEN) {
            if (active) {
                const double *g = A.ngeom + ((size_t)e * NN + node) * G1 + 1 + dir * DIM;
#pragma unroll
                for (int c = 0; c < DIM; c++) sMr[tid][c] = g[c];
            }
        }
    };
    if (active) {
        double s[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            s[v] = A.u[v * vs + (size_t)e * NN + node];
            if constexpr (SSTATE) sS[v][tid] = s[v];
        }
        const SideState<DIM> qs = make_side<DIM>(s, A.P);
        sPd[0][tid] = qs.rho;
        sPd[1][tid] = qs.beta;
#pragma unroll
        for (int d = 0; d < DIM; d++) sPv[d][tid] = (R)qs.vel[d];
        if (MIX && node == 0) {
            sPref[eb] = qs.p;
            sHref[eb] = A.P.gamma * A.P.gammaM1Inverse * qs.p * qs.irho;
        }
        lam = qs.lam;
        if constexpr (LSIDE) sLam[tid] = qs.lam;
        if constexpr (BLEND) qrp = qs.rho * qs.p;   // the indicator's rho p from the node's primitives
    }
    __syncthreads();
    if constexpr (BLEND) {
        // the elements' blending coefficients (DGSEMOperator.cpp:547-549), the surface-flux
        // buffer as scratch before the face items fill it
        double *scr = &sF[0][0];
        blend_coef_block<DIM, NP, EPB, C::BLOCK == 64 || 64 % NN == 0>(qrp, active, tid, eb, node, A.blend, sVi, scr,
                                                                       scr + EPB * NN, scr + 2 * EPB * NN, sAlpha);
        __syncthreads();
        if (active && node == 0 && A.alpha) A.alpha[e] = sAlpha[eb][0];
    }
    // surface fluxes: work items (element ib, direction, side, line t) over the whole workgroup,
    // the element's own traces from LDS, the neighbour's from HBM
    for (int it = tid; it < EPB * FITEMS; it += C::BLOCK) {
        const int ib = it / FITEMS, r = it - ib * FITEMS;
        const int ee = blk * EPB + ib;
        if (ee >= NE) continue;
        const int dir = r / (2 * L), side = (r / L) & 1, t = r - (r / L) * L;
        const int st = dir_stride<NP>(dir);
        const int base = line_base_rt<DIM, NP>(t, dir);
        const int nown = base + (side ? (NP - 1) * st : 0);
        const int nnb = base + (side ? 0 : (NP - 1) * st);
        const int en = sNbr[ib][2 * dir + side];
        if constexpr (BND) {
            if (en < 0) {
                // boundary face (BdrFaceIntegrator::AssembleFaceVector, Euler :68-139, NS :148-231):
                // f = side ? fb : -fb along +adj_d, fb the condition's inviscid flux with the
                // outward normal, minus the viscous trace the lift kernel stored for it
                double so[NEQ], nor[DIM], fb[NEQ];
#pragma unroll
                for (int v = 0; v < NEQ; v++) so[v] = own_state(v, ib, nown);
                const SideState<DIM> qo = make_side<DIM>(so, A.P);
                const R *m = GEN ? A.ngeom + ((size_t)ee * NN + nown) * G1 + 1 + dir * DIM : sM[ib][dir];
#pragma unroll
                for (int c = 0; c < DIM; c++) nor[c] = side ? m[c] : -m[c];
                const double *xb = A.bxyz ? A.bxyz + (((size_t)ee * 2 * DIM + 2 * dir + side) * L + t) * DIM : nullptr;
                lam = fmax(lam, bc_inviscid_flux<DIM, SURF>(A.bc[-en - 1], A.P, A.bug3 != 0, so, qo, nor, fb, xb,
                                                            A.time));
                const double sg = side ? 1.0 : -1.0;
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    double fv = 0.0;
                    if (VISC && v > 0) fv = A.Ff[ftrace<DIM, NP>(dir, side, v - 1, ee, t, NE)];
                    sF[it][v] = sg * fb[v] - fv;
                }
                continue;
            }
        }
        if constexpr (GEN) {
            // the neighbour's face (d2, s2) and point tn; the flux once per face, with the normal
            // of the lower-numbered element (ties: the lower local face), stored along +adj_dir
            const int code = sOr[ib][2 * dir + side];
            const int f2 = code & 7, d2 = f2 >> 1, s2 = f2 & 1;
            const int tn = face_point_map<DIM, NP>(code, t);
            const int st2 = dir_stride<NP>(d2);
            const int nn2 = line_base<NP>(tn, d2) + (s2 ? (NP - 1) * st2 : 0);
            const bool first = (ee < en) || (ee == en && 2 * dir + side < f2);
            double so[NEQ], sn[NEQ], n1[DIM];
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                so[v] = own_state(v, ib, nown);
                sn[v] = A.u[v * vs + (size_t)en * NN + nn2];
            }
            {
                // CalcOrtho: the lower-numbered element's adj row at its face point, outward
                const double *g = first ? A.ngeom + ((size_t)ee * NN + nown) * G1 + 1 + dir * DIM
                                        : A.ngeom + ((size_t)en * NN + nn2) * G1 + 1 + d2 * DIM;
                const bool out = first ? (side == 1) : (s2 == 1);
#pragma unroll
                for (int c = 0; c < DIM; c++) n1[c] = out ? g[c] : -g[c];
            }
            double nmag = n1[0] * n1[0];
#pragma unroll
            for (int d = 1; d < DIM; d++) nmag += n1[d] * n1[d];
            nmag = sqrt(nmag);
            const SideState<DIM> qo = make_side<DIM>(so, A.P);
            const SideState<DIM> qn = make_side<DIM>(sn, A.P);
            double f[NEQ];
            if (first)
                face_flux<DIM, SURF>(so, qo, sn, qn, n1, nmag, A.P, A.bug3 != 0, f);
            else
                face_flux<DIM, SURF>(sn, qn, so, qo, n1, nmag, A.P, A.bug3 != 0, f);
            // the own element's outward flux is (first ? f : -f); along +adj_dir: times (side ? 1 : -1)
            const double sg = (first == (side == 1)) ? 1.0 : -1.0;
#pragma unroll
            for (int v = 0; v < NEQ; v++) f[v] *= sg;
            if constexpr (VISC) {
                // 1/2 (F_own + F_nb) along the own +adj_dir: the neighbour's trace is stored along its
                // +adj_d2, which points the other way across the face when s2 == side
                const double sig = (side != s2) ? 1.0 : -1.0;
#pragma unroll
                for (int v = 0; v < NEQ - 1; v++)
                    f[1 + v] -= 0.5 * (A.Ff[ftrace<DIM, NP>(dir, side, v, ee, t, NE)] +
                                       sig * A.Ff[ftrace<DIM, NP>(d2, s2, v, en, tn, NE)]);
            }
#pragma unroll
            for (int v = 0; v < NEQ; v++) sF[it][v] = f[v];
            continue;
        }
        double so[NEQ], sn[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            so[v] = own_state(v, ib, nown);
            sn[v] = A.u[v * vs + (size_t)en * NN + nnb];
        }
        R fv[NEQ - 1];
        if constexpr (VISC) {
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) {
                if constexpr (MIX)
                    fv[v] = 0.5f * (A.Fff[ftrace<DIM, NP>(dir, side, v, ee, t, NE)] +
                                    A.Fff[ftrace<DIM, NP>(dir, 1 - side, v, en, t, NE)]);
                else
                    fv[v] = 0.5 * (A.Ff[ftrace<DIM, NP>(dir, side, v, ee, t, NE)] +
                                   A.Ff[ftrace<DIM, NP>(dir, 1 - side, v, en, t, NE)]);
            }
        }
        SideState<DIM> qo;
        if constexpr (LSIDE) {
            const int lo = ib * NN + nown;
            double vo[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) vo[d] = sPv[d][lo];
            qo = side_from_prim<DIM>(sPd[0][lo], sPd[1][lo], vo, sLam[lo]);
        } else {
            qo = make_side<DIM>(so, A.P);
        }
        const SideState<DIM> qn = make_side<DIM>(sn, A.P);
        const R *m = sM[ib][dir];
        const R nmag = sNm[ib][dir];
        R f[NEQ];
        if constexpr (MIX) {
            const double p_ref = sPref[ib], h_ref = sHref[ib];
            if (side)
                face_flux_mx<DIM, SURF>(so, qo, sn, qn, m, nmag, A.P, A.bug3 != 0, p_ref, h_ref, f);
            else
                face_flux_mx<DIM, SURF>(sn, qn, so, qo, m, nmag, A.P, A.bug3 != 0, p_ref, h_ref, f);
        } else {
            if (side)
                face_flux<DIM, SURF>(so, qo, sn, qn, m, nmag, A.P, A.bug3 != 0, f);
            else
                face_flux<DIM, SURF>(sn, qn, so, qo, m, nmag, A.P, A.bug3 != 0, f);
        }
        if constexpr (VISC) {
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) f[1 + v] -= fv[v];
        }
#pragma unroll
        for (int v = 0; v < NEQ; v++) sF[it][v] = f[v];
    }
    // The volume terms and the stage epilogue, one body instantiated for the blended path (BL) and
    // for the unblended one: a blended workgroup whose elements all have alpha = 0 (the usual case
    // away from shocks) takes the unblended body, which needs neither the (1 - alpha) weights nor
    // the registers of the finite-volume work.
    // (A by value: a copy of the kernel arguments that the body only reads, so that their loads are
    // not repeated after every store the body makes.  The copy is indexed at fixed offsets only,
    // else it would be placed in scratch memory: the two tables the finite-volume terms index per
    // lane are read through pointers into the kernel arguments themselves.)
    const double *const kinvw = A.T.inv_w;
    const double(*const ksub)[NP] = A.T.sub;
    int bad = 0;
    auto body = [&, A](auto tag) __attribute__((always_inline)) {
        constexpr bool BL = decltype(tag)::value;
        // Affine elements: the epilogue's operands (the RK register or base state, the viscous
        // volume term) are loaded here, so their HBM latency overlaps the volume loop instead of
        // stalling the epilogue; the face terms are folded into acc (FOLD) to keep the register
        // budget.  General meshes keep those registers for their per-node metrics.
        constexpr bool PRE = !GEN;
        double pre_x[NEQ];
        R pre_dv[NEQ - 1];
        if constexpr (PRE) {
            if (active) {
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    const size_t gi = v * vs + (size_t)e * NN + node;
                    pre_x[v] = A.mode == EPI_LSRK ? take<DIM == 3>(&A.k[gi]) : (A.mode == EPI_SSP ? A.x[gi] : 0.0);
                }
                if constexpr (VISC) {
#pragma unroll
                    for (int v = 0; v < NEQ - 1; v++) {
                        if constexpr (MIX)
                            pre_dv[v] = A.dvf[v * vs + (size_t)e * NN + node];
                        else   // its last use (FP64; the FP32 array of the mixed mode stays cached)
                            pre_dv[v] = take<DIM == 3>(&A.dv[v * vs + (size_t)e * NN + node]);
                    }
                }
            }
        }
        // affine elements fold the surface terms into the volume accumulator; blended ones weight the
        // volume pairs by (1 - alpha) as they accumulate them, so that the fold holds for them too
        constexpr bool FOLD = !GEN;
        __syncthreads();
        R oma = R(1);   // 1 - alpha of the node's element (BLEND on affine elements)
        if constexpr (BL && FOLD) {
            if (active) oma = (R)(1.0 - sAlpha[eb][0]);
        }
        int ai[3];
        ai[0] = node % NP;
        ai[1] = (node / NP) % NP;
        ai[2] = node / (NP * NP);
        R acc[NEQ], accs[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) acc[v] = accs[v] = R(0);
        double invJ = 0.0;
        if constexpr (GEN) {
            // per direction: the nodes' adj rows into sMr, then the pair fluxes with the averaged
            // metric 1/2 (adj_d(i) + adj_d(j)) (ComputeMean, ChandrashekarFlux.cpp:15)
            if (active) invJ = rcp(detJ);
#pragma unroll
            for (int dir = 0; dir < DIM; dir++) {
                if (dir > 0) __syncthreads();   // the previous direction's readers of sMr are done
                load_rows(dir);
                __syncthreads();
                if (active) {
                    const int st = dir_stride<NP>(dir);
                    const int a = ai[dir];
                    const int lb = eb * NN + node - a * st;
                    double mo[DIM];
#pragma unroll
                    for (int c = 0; c < DIM; c++) mo[c] = sMr[tid][c];
                    // the pair (a, b) with the averaged metric: symmetric in a and b like the flux
                    auto gen_pair = [&](int li, double *F) {
                        NodePrim<DIM> qo, qb;
                        qo.rho = sPd[0][tid];
                        qo.beta = sPd[1][tid];
#pragma unroll
                        for (int d = 0; d < DIM; d++) qo.vel[d] = sPv[d][tid];
                        qb.rho = sPd[0][li];
                        qb.beta = sPd[1][li];
#pragma unroll
                        for (int d = 0; d < DIM; d++) qb.vel[d] = sPv[d][li];
                        double mp[DIM];
#pragma unroll
                        for (int c = 0; c < DIM; c++) mp[c] = (mo[c] + sMr[li][c]) * 0.5;
                        volume_flux<DIM, true>(qo, qb, mp, A.P, F);
                    };
                    if constexpr (ROTG) {
                        // each symmetric pair once where an element's lines lie in one wave: node a
                        // evaluates (a, a+k), k <= Np/2, and receives (a-k, a) from the lane of a-k
                        const int lane = tid & 63;
#pragma unroll
                        for (int k = 1; 2 * k <= NP; k++) {
                            const int b = (a + k) % NP;
                            double F[NEQ];
                            gen_pair(lb + b * st, F);
                            const double dd = sT[a * NP + b];
#pragma unroll
                            for (int v = 0; v < NEQ; v++) acc[v] += dd * F[v];
                            if (2 * k != NP) {
                                const int c = (a + NP - k) % NP;
                                const int src = lane + (c - a) * st;
                                const double dc = sT[a * NP + c];
#pragma unroll
                                for (int v = 0; v < NEQ; v++) acc[v] += dc * __shfl(F[v], src, 64);
                            }
                        }
                    } else {
#pragma unroll 1
                        for (int k = 1; k < NP; k++) {
                            const int b = (a + k) % NP;
                            double F[NEQ];
                            gen_pair(lb + b * st, F);
                            const double dd = sT[a * NP + b];
#pragma unroll
                            for (int v = 0; v < NEQ; v++) acc[v] += dd * F[v];
                        }
                    }
                    if (a == 0 || a == NP - 1) {
                        const int side = (a == NP - 1) ? 1 : 0;
                        const int it = eb * FITEMS + (dir * 2 + side) * L + line_index<NP>(node, dir);
                        const double sg = side ? -1.0 : 1.0;
#pragma unroll
                        for (int v = 0; v < NEQ; v++) accs[v] += sg * sF[it][v];
                    }
                }
            }
        } else if (active) {
            invJ = rcp(A.geom[(size_t)e * (1 + DIM * DIM)]);
            const double p_ref = MIX ? sPref[eb] : 0.0;
            const double h_ref = MIX ? sHref[eb] : 0.0;
#pragma unroll
            for (int dir = 0; dir < DIM; dir++) {
                const int st = dir_stride<NP>(dir);
                const int a = ai[dir];
                const int lb = eb * NN + node - a * st;
                const R *m = sM[GEN ? 0 : eb][dir];   // (GEN takes the branch above)
                // the two-point flux with the node at LDS index li (the line partner b)
                auto pair_flux = [&](int b, int li, R *F) {
                    if constexpr (MIX) {
                        NodePrimMx<DIM> qo, qb;
                        qo.rho = sPd[0][tid];
                        qo.beta = sPd[1][tid];
                        qb.rho = sPd[0][li];
                        qb.beta = sPd[1][li];
#pragma unroll
                        for (int d = 0; d < DIM; d++) {
                            qo.vel[d] = sPv[d][tid];
                            qb.vel[d] = sPv[d][li];
                        }
                        volume_flux_mx<DIM>(qo, qb, m, A.P, p_ref, h_ref, F);
                    } else if (BUG1 && dir == 2 && b < a) {
                        // the lower-zeta partner's slot holds its own last eta flux (:702 as written)
                        const int j = ai[1];
                        if (j == NP - 1) {
#pragma unroll
                            for (int v = 0; v < NEQ; v++) F[v] = R(0);
                        } else {
                            const int lj = li + (NP - 1 - j) * NP;   // node (i, Np-1, b)
                            NodePrim<DIM> qo, qb;
                            qo.rho = sPd[0][li];
                            qo.beta = sPd[1][li];
                            qb.rho = sPd[0][lj];
                            qb.beta = sPd[1][lj];
#pragma unroll
                            for (int d = 0; d < DIM; d++) {
                                qo.vel[d] = sPv[d][li];
                                qb.vel[d] = sPv[d][lj];
                            }
                            volume_flux<DIM, true>(qo, qb, sM[eb][1], A.P, F);
                        }
                    } else {
                        NodePrim<DIM> qo, qb;
                        qo.rho = sPd[0][tid];
                        qo.beta = sPd[1][tid];
#pragma unroll
                        for (int d = 0; d < DIM; d++) qo.vel[d] = sPv[d][tid];
                        qb.rho = sPd[0][li];
                        qb.beta = sPd[1][li];
#pragma unroll
                        for (int d = 0; d < DIM; d++) qb.vel[d] = sPv[d][li];
                        volume_flux<DIM, true>(qo, qb, m, A.P, F);
                    }
                };
                if constexpr (ROT) {
                    // each symmetric pair once: node a evaluates (a, a+k) for k <= Np/2 and receives
                    // (a-k, a) from the lane of node a-k (the same flux, F(u_i,u_j) = F(u_j,u_i));
                    // 2k = Np pairs are evaluated by both of their nodes in the last direction, and in
                    // one shared pass after this loop in directions 0 and 1 (HALF2)
                    const int lane = tid & 63;
                    const int KMAX = (HALF2 && dir < 2) ? (NP - 1) / 2 : NP / 2;
#pragma unroll
                    for (int k = 1; k <= KMAX; k++) {
                        const int b = (a + k) % NP;
                        R F[NEQ];
                        pair_flux(b, lb + b * st, F);
                        const R dd = BL ? sT[a * NP + b] * oma : sT[a * NP + b];
#pragma unroll
                        for (int v = 0; v < NEQ; v++) acc[v] += dd * F[v];
                        if (2 * k != NP) {
                            const int c = (a + NP - k) % NP;
                            const int src = lane + (c - a) * st;
                            const R dc = BL ? sT[a * NP + c] * oma : sT[a * NP + c];
#pragma unroll
                            for (int v = 0; v < NEQ; v++) acc[v] += dc * __shfl(F[v], src, 64);
                        }
                    }
                } else {
#pragma unroll 1
                    for (int k = 1; k < NP; k++) {
                        const int b = (a + k) % NP;
                        R F[NEQ];
                        pair_flux(b, lb + b * st, F);
                        const R dd = BL ? sT[a * NP + b] * oma : sT[a * NP + b];
#pragma unroll
                        for (int v = 0; v < NEQ; v++) acc[v] += dd * F[v];
                    }
                }
                if (a == 0 || a == NP - 1) {
                    const int side = (a == NP - 1) ? 1 : 0;
                    const int it = eb * FITEMS + (dir * 2 + side) * L + line_index<NP>(node, dir);
                    if constexpr (FOLD) {
                        // du = -(1/J)(acc - accs/w0): the face term enters acc as -/+ f/w0
                        const R sg = (R)(side ? A.T.inv_w0 : -A.T.inv_w0);
#pragma unroll
                        for (int v = 0; v < NEQ; v++) acc[v] += sg * sF[it][v];
                    } else {
                        const R sg = side ? R(-1) : R(1);
#pragma unroll
                        for (int v = 0; v < NEQ; v++) accs[v] += sg * sF[it][v];
                    }
                }
            }
            if constexpr (HALF2) {
                // the pairs (a, a + Np/2) of directions 0 and 1, each once: a lane with i < Np/2 takes
                // its node's direction-0 pair (i, j) - (i + Np/2, j); each other lane one
                // direction-1 pair (i', j') - (i', j' + Np/2) with j' < Np/2, numbered
                // (i - Np/2) + (Np/2) j = i' + Np j'.  Both nodes of a pair receive it by shuffles.
                constexpr int H = NP / 2;
                const int lane = tid & 63;
                const int i = ai[0], j = ai[1];
                const int n00 = node - i - j * NP;   // node (0, 0, k)
                int na, nb, dl;
                if (i < H) {
                    na = node;
                    nb = node + H;
                    dl = 0;
                } else {
                    const int idx = (i - H) + H * j;
                    na = n00 + (idx % NP) + (idx / NP) * NP;
                    nb = na + H * NP;
                    dl = 1;
                }
                R F[NEQ];
                {
                    const int la = eb * NN + na, lb2 = eb * NN + nb;
                    const R *m = sM[eb][dl];
                    if constexpr (MIX) {
                        NodePrimMx<DIM> qa, qb;
                        qa.rho = sPd[0][la];
                        qa.beta = sPd[1][la];
                        qb.rho = sPd[0][lb2];
                        qb.beta = sPd[1][lb2];
#pragma unroll
                        for (int d = 0; d < DIM; d++) {
                            qa.vel[d] = sPv[d][la];
                            qb.vel[d] = sPv[d][lb2];
                        }
                        volume_flux_mx<DIM>(qa, qb, m, A.P, p_ref, h_ref, F);
                    } else {
                        NodePrim<DIM> qa, qb;
                        qa.rho = sPd[0][la];
                        qa.beta = sPd[1][la];
                        qb.rho = sPd[0][lb2];
                        qb.beta = sPd[1][lb2];
#pragma unroll
                        for (int d = 0; d < DIM; d++) {
                            qa.vel[d] = sPv[d][la];
                            qb.vel[d] = sPv[d][lb2];
                        }
                        volume_flux<DIM, true>(qa, qb, m, A.P, F);
                    }
                }
                // the direction-0 pair of this node was evaluated by the lane of (i mod H, j), the
                // direction-1 pair by the lane that numbers (i, j mod H)
                const int xl = lane - (i >= H ? H : 0);
                const int jm = j < H ? j : j - H;
                const int idy = i + NP * jm;
                const int yl = lane - (node - n00) + (H + idy % H) + (idy / H) * NP;
                const R dx = sT[i * NP + (i < H ? i + H : i - H)] * oma;
                const R dy = sT[j * NP + (j < H ? j + H : j - H)] * oma;
#pragma unroll
                for (int v = 0; v < NEQ; v++) acc[v] += dx * __shfl(F[v], xl, 64) + dy * __shfl(F[v], yl, 64);
            }
        }
        if constexpr (BLEND) {
            if (active) {
                // the DG terms, (1 - alpha) volume + surface, as the RHS so far; acc then collects the
                // RHS and the FV term is added into it
                if constexpr (FOLD) {
#pragma unroll
                    for (int v = 0; v < NEQ; v++) acc[v] = -acc[v] * invJ;
                } else {
                    const double al = BL ? sAlpha[eb][0] : 0.0, sw = A.T.inv_w0 * invJ;
#pragma unroll
                    for (int v = 0; v < NEQ; v++) {
                        acc[v] = (-acc[v] * invJ) * (1.0 - al) + accs[v] * sw;
                        accs[v] = 0.0;
                    }
                }
            }
            if constexpr (BL) {
#pragma unroll 1
                for (int dir = 0; dir < DIM; dir++) {
                    __syncthreads();   // the previous readers of sF (and sMr) are done
                    if constexpr (GEN) {
                        load_rows(dir);
                        __syncthreads();
                    }
                    const int st = dir_stride<NP>(dir);
                    for (int it = tid; it < EPB * SITEMS; it += C::BLOCK) {
                        const int ib = it / SITEMS, r = it - ib * SITEMS;
                        if (blk * EPB + ib >= NE || sAlpha[ib][0] == 0.0) continue;
                        const int i = r / L, t = r - i * L;
                        const int nl = ib * NN + line_base<NP>(t, dir) + i * st;
                        double sl[NEQ], sr[NEQ];
#pragma unroll
                        for (int v = 0; v < NEQ; v++) {
                            sl[v] = sS[v][nl];
                            sr[v] = sS[v][nl + st];
                        }
                        const SideState<DIM> ql = make_side<DIM>(sl, A.P);
                        const SideState<DIM> qr = make_side<DIM>(sr, A.P);
                        R mg[DIM];
                        if constexpr (GEN) {
                            // the subcell metric of interface i+1 from the line's per-node adj rows
                            const int l0 = ib * NN + line_base<NP>(t, dir);
#pragma unroll
                            for (int c = 0; c < DIM; c++) mg[c] = 0.0;
#pragma unroll
                            for (int k = 0; k < NP; k++) {
                                const double wk = ksub[i][k];
#pragma unroll
                                for (int c = 0; c < DIM; c++) mg[c] += wk * sMr[l0 + k * st][c];
                            }
                        } else {
#pragma unroll
                            for (int c = 0; c < DIM; c++) mg[c] = sM[ib][dir][c];
                        }
                        const R *m = mg;
                        R nmag = m[0] * m[0];
#pragma unroll
                        for (int d = 1; d < DIM; d++) nmag += m[d] * m[d];
                        nmag = sqrt(nmag);
                        R f[NEQ];
                        face_flux<DIM, SURF>(sl, ql, sr, qr, m, nmag, A.P, A.bug3 != 0, f);
#pragma unroll
                        for (int v = 0; v < NEQ; v++) sF[it][v] = f[v];
                    }
                    __syncthreads();
                    if (active && sAlpha[eb][0] != 0.0) {
                        const int a = (node / st) % NP;   // not ai[dir]: dir is not unrolled here
                        const int it = eb * SITEMS + line_index<NP>(node, dir);
                        const double sc = sAlpha[eb][0] * invJ * kinvw[a];   // alpha / (J w_a)
#pragma unroll
                        for (int v = 0; v < NEQ; v++) {
                            double g = 0.0;
                            if (a > 0) g += sF[it + (a - 1) * L][v];
                            if (a < NP - 1) g -= sF[it + a * L][v];
                            acc[v] += g * sc;
                        }
                    }
                }
            }
        }
        double snew[NEQ];   // the node's stage output (stage hooks)
        if (active) {
            const R sw = (R)(A.T.inv_w0 * invJ);
            const R rinvJ = (R)invJ;
            const double h_ref = MIX ? sHref[eb] : 0.0;
            double du0 = 0.0;
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                R du;
                if constexpr (BLEND) {
                    du = acc[v];
                } else if constexpr (FOLD) {
                    du = -acc[v] * rinvJ;
                } else {
                    du = -acc[v] * rinvJ + accs[v] * sw;
                }
                double dud;
                if constexpr (MIX) {
                    // the energy sums carried F_E - h_ref F_rho: add h_ref times the inviscid mass RHS
                    if (v == 0) du0 = (double)du;
                    dud = (v == NEQ - 1) ? (double)du + h_ref * du0 : (double)du;
                } else {
                    dud = du;
                }
                if constexpr (VISC) {
                    if (v > 0) {
                        if constexpr (PRE)
                            dud += pre_dv[v - 1];
                        else if constexpr (MIX)
                            dud += A.dvf[(v - 1) * vs + (size_t)e * NN + node];
                        else
                            dud += take<DIM == 3>(&A.dv[(v - 1) * vs + (size_t)e * NN + node]);
                    }
                }
                // the RK accumulation is FP64 in both modes
                const size_t gi = v * vs + (size_t)e * NN + node;
                if constexpr (PRE)
                    snew[v] = stage_update_x<DIM, NP>(A, gi, own_state(v, eb, node), dud, pre_x[v], bad);
                else
                    snew[v] = stage_update<DIM, NP>(A, gi, own_state(v, eb, node), dud, bad);
            }
        }
        if constexpr (C::BLOCK == 64 || NN == 64) {
            // an element's nodes lie in one wave: the hooks' detection from the registers
            if (A.hook_mode) stage_hook_detect<DIM, NP, NN>(A, active, e, node, snew);
        }
    };
    if constexpr (BLEND) {
        bool any = false;
#pragma unroll
        for (int ib = 0; ib < EPB; ib++) any |= (blk * EPB + ib < NE) && sAlpha[ib][0] != 0.0;
        if (any)   // the same value in every thread of the workgroup
            body(std::true_type{});
        else
            body(std::false_type{});
    } else {
        body(std::false_type{});
    }
    finish_block<NP>(A, A.want_lambda ? lam : 0.0, bad);
}

// ---------------------------------------------------------------------------
// k_lift: BR1 lifting of the entropy variables + chain rule + viscous flux.
// ref: DGSEMOperator::Mult (:551-571), ComputeGlobalEntropyVector (:82-94),
//      LiftingBR1::AssembleLiftingElementVector (3D :156-219, 2D :222-272),
//      LiftingBR1::AssembleLiftingFaceVector (3D :10-66, 2D :68-115),
//      ComputeGlobalPrimitiveGradVector (:96-159), NavierStokesFlux::ComputeViscousFlux,
//      the adj(J) mapping of the viscous flux (DGSEMIntegrator.cpp:632-640) and its
//      D-hat divergence (:705-738).
// One thread per node.  The entropy variables of the face neighbours (2*dim*Np^(dim-1)
// face points per element) are evaluated as cooperative items shared by all threads of
// the workgroup (full lane occupancy) and staged in LDS.
// ---------------------------------------------------------------------------
template <int DIM, int NP>
struct LiftCfg {
    static constexpr int NN = ipow(NP, DIM);
    static constexpr int L = ipow(NP, DIM - 1);
    static constexpr int FITEMS = 2 * DIM * L;   // face points per element
    static constexpr int BLOCK = NN <= 32 ? 64 : (NN <= 64 ? 128 : 64 * ((NN + 63) / 64));
    static constexpr int EPB = BLOCK / NN;
};

// GEN (general meshes): per-node metrics and oriented face matching, as in k_rhs_node; the
// jump term of a face point uses the node's own outward normal (the reference's face normal up
// to rounding on a conforming mesh).
//
// BUG2 (bug_compat A2, affine FP64 meshes): the viscous flux of the volume term is evaluated with
// the conserved state of the element's last node (the reference reads el_u at a stale id1,
// DGSEMIntegrator.cpp:620 in 3D and :770 in 2D, which the blended build leaves at n-1) and each
// node's own gradient; the face traces keep the node's own state (AssembleFaceVector).
template <int DIM, int NP, bool MIX, bool BND = false, bool GEN = false, bool BUG2 = false>
__global__ void __launch_bounds__((LiftCfg<DIM, NP>::BLOCK), PRANDTL_LIFT_WAVES) k_lift(const LiftArgs<NP> A) {
    static_assert(!(MIX && BND), "boundary faces run in FP64");
    static_assert(!(BUG2 && (MIX || BND || GEN)), "bug_compat A2 runs on affine FP64 meshes");
    using C = LiftCfg<DIM, NP>;
    // derivative contractions, chain rule and viscous flux in R; entropy variables in FP64.
    // Mixed precision contracts w - w_ref (w_ref = the element's node-0 value; D annihilates
    // constants) so the FP32 values carry the variation, not the ~1e1 offset of w.
    using R = typename std::conditional<MIX, float, double>::type;
    constexpr int NEQ = DIM + 2, NN = C::NN, EPB = C::EPB, L = C::L;
    constexpr int NW = NEQ * EPB * C::FITEMS, NFV = DIM * (NEQ - 1) * EPB * NN;
    // FP64 3D p = 3: the tensor contractions on the matrix cores (an element is one wave)
    constexpr bool LIFT_MFMA = PRANDTL_LIFT_MFMA && DIM == 3 && NP == 4 && !MIX;
    __shared__ R sW[NEQ][EPB * NN];
    __shared__ R sU[NW > NFV ? NW : NFV];  // neighbour w at face points, then F~^d
    __shared__ double sWref[EPB][NEQ];
    __shared__ LogRef sLR[EPB];           // reference state of the logarithms (the element's node 0)
    __shared__ int sNbr[EPB][2 * DIM];    // face neighbours of the workgroup's elements
    __shared__ int sOr[GEN ? EPB : 1][2 * DIM];   // GEN: face orientation codes

    const int blk = xcd_block_rev(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x;
    const int eb = tid / NN, node = tid - (tid / NN) * NN;
    const int e0 = blk * EPB;
    const int e = e0 + eb;
    const int NE = A.NE;
    const int nel = min(EPB, NE - e0);
    const bool active = (eb < EPB) && (e < NE);
    const size_t vs = (size_t)NE * NN;

    for (int i = tid; i < EPB * 2 * DIM; i += C::BLOCK) {
        const int ib = i / (2 * DIM), f = i - ib * 2 * DIM;
        sNbr[ib][f] = A.nbr[(size_t)min(e0 + ib, NE - 1) * 2 * DIM + f];
        if constexpr (GEN) sOr[ib][f] = A.forient[(size_t)min(e0 + ib, NE - 1) * 2 * DIM + f];
    }
    double s[NEQ], w[NEQ];
    if (active) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) s[v] = A.u[v * vs + (size_t)e * NN + node];
        if (node == 0) sLR[eb] = make_log_ref<DIM>(s, A.P);
    }
    __syncthreads();
    if (active) {
        conserv_to_entropy_ref<DIM>(s, A.P, sLR[eb], w);
        if (MIX && node == 0) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) sWref[eb][v] = w[v];
        }
    }
    if (MIX) __syncthreads();
    double wref[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) wref[v] = (MIX && active) ? sWref[eb][v] : 0.0;
    // the MFMA path reads sW only at bit-swapped nodes of the wave's element; it stores node n at
    // n ^ 5 (n >> 4) (the zeta index XORed into the xi and eta bits), where the own, the
    // xi-swapped and the eta-swapped patterns each touch every LDS bank once
    const int lsw = LIFT_MFMA ? (tid & ~63) | node_swz(tid & 63) : tid;
    if (active) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) sW[v][lsw] = (R)(w[v] - wref[v]);
    }
    for (int it = tid; it < nel * C::FITEMS; it += C::BLOCK) {
        const int fe = it / C::FITEMS, i = it - fe * C::FITEMS;
        const int dir = i / (2 * L), side = (i / L) & 1, t = i % L;
        const int st = dir_stride<NP>(dir);
        const int en = sNbr[fe][2 * dir + side];
        const int nnb = line_base<NP>(t, dir) + (side ? 0 : (NP - 1) * st);
        double sn[NEQ], wn[NEQ];
        if (BND && en < 0) {
            // boundary face (LiftingBR1.cpp:311-407): the condition's lifting flux f enters the
            // node's face term 1/2 (wn - w) as wn = w + 2 f, w the element's own trace
            const int nown = line_base<NP>(t, dir) + (side ? (NP - 1) * st : 0);
            double so[NEQ], wo[NEQ], nor[DIM], fl[NEQ];
#pragma unroll
            for (int v = 0; v < NEQ; v++) so[v] = A.u[v * vs + (size_t)(e0 + fe) * NN + nown];
            conserv_to_entropy_ref<DIM>(so, A.P, sLR[fe], wo);
            const double *gm = GEN ? A.ngeom + ((size_t)(e0 + fe) * NN + nown) * (1 + DIM * DIM) + 1 + dir * DIM
                                   : A.geom + (size_t)(e0 + fe) * (1 + DIM * DIM) + 1 + dir * DIM;
#pragma unroll
            for (int c = 0; c < DIM; c++) nor[c] = side ? gm[c] : -gm[c];
            constexpr int LF = ipow(NP, DIM - 1);
            const double *xb = A.bxyz ? A.bxyz + (((size_t)(e0 + fe) * 2 * DIM + 2 * dir + side) * LF + t) * DIM
                                      : nullptr;
            bc_lifting_flux<DIM>(A.bc[-en - 1], A.P, wo, so, nor, fl, xb, A.time);
#pragma unroll
            for (int v = 0; v < NEQ; v++) wn[v] = wo[v] + 2.0 * fl[v];
        } else {
            int nb = nnb;
            if constexpr (GEN) {
                const int code = sOr[fe][2 * dir + side];
                const int d2 = (code & 7) >> 1, s2 = code & 1;
                const int st2 = dir_stride<NP>(d2);
                nb = line_base<NP>(face_point_map<DIM, NP>(code, t), d2) + (s2 ? (NP - 1) * st2 : 0);
            }
#pragma unroll
            for (int v = 0; v < NEQ; v++) sn[v] = A.u[v * vs + (size_t)en * NN + nb];
            conserv_to_entropy_ref<DIM>(sn, A.P, sLR[fe], wn);
        }
#pragma unroll
        for (int v = 0; v < NEQ; v++) sU[(v * EPB + fe) * C::FITEMS + i] = (R)(wn[v] - (MIX ? sWref[fe][v] : 0.0));
    }
    __syncthreads();

    int ai[3];
    ai[0] = node % NP;
    ai[1] = (node / NP) % NP;
    ai[2] = node / (NP * NP);
    R invJ = R(0), adj[DIM][DIM], Fc[DIM][NEQ - 1];
    if (active) {
        const double *gm = GEN ? A.ngeom + ((size_t)e * NN + node) * (1 + DIM * DIM) : A.geom + (size_t)e * (1 + DIM * DIM);
        invJ = (R)rcp(gm[0]);
#pragma unroll
        for (int r = 0; r < DIM; r++)
#pragma unroll
            for (int c = 0; c < DIM; c++) adj[r][c] = (R)gm[1 + r * DIM + c];
        R wrel[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) wrel[v] = (R)(w[v] - wref[v]);
        R gref[DIM][NEQ];
        if constexpr (LIFT_MFMA) {
            // the D-line contractions on the matrix cores: one v_mfma_f64_4x4x4f64 (four 4x4x4
            // blocks; A at lane m + 4 blk + 16 k, B at lane n + 4 blk + 16 k, D at lane
            // n + 4 blk + 16 m) per direction and variable for the element's 64 nodes (lane = node
            // i + 4 j + 16 k).  zeta: A = D, B = the node's own w; xi: A = w at the node with the
            // lane's bits 0-1 and 4-5 swapped, B = D; eta: A = D, B = w at the node with bits 2-3 and
            // 4-5 swapped, the result on the lane with the same swap
            const int lw = tid & 63, base = tid - lw;
            const int sx = ((lw >> 4) & 3) | (lw & 12) | ((lw & 3) << 4);
            const int sy = (lw & 3) | (((lw >> 4) & 3) << 2) | (((lw >> 2) & 3) << 4);
            const int ax = base + node_swz(sx), ay = base + node_swz(sy);
            const double Dc = A.T.D[lw & 3][lw >> 4];
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                gref[2][v] = __builtin_amdgcn_mfma_f64_4x4x4f64(Dc, (double)wrel[v], 0.0, 0, 0, 0);
                gref[0][v] = __builtin_amdgcn_mfma_f64_4x4x4f64((double)sW[v][ax], Dc, 0.0, 0, 0, 0);
                const double gy = __builtin_amdgcn_mfma_f64_4x4x4f64(Dc, (double)sW[v][ay], 0.0, 0, 0, 0);
                gref[1][v] = __shfl(gy, sy, 64);
            }
        }
#pragma unroll
        for (int dir = 0; dir < DIM; dir++) {
            const int st = dir_stride<NP>(dir);
            const int a = ai[dir];
            const int lb = tid - a * st;
            if constexpr (!LIFT_MFMA) {
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    R t = R(0);
#pragma unroll
                    for (int b = 0; b < NP; b++) t += (R)A.T.D[a][b] * sW[v][lb + b * st];
                    gref[dir][v] = t;
                }
            }
            if (a == NP - 1 || a == 0) {
                // face jump 1/2 (w+ - w-) / w0, the same sign on both sides (LiftingBR1.cpp:51-65)
                const bool plus = (a == NP - 1);
                const int i = dir * 2 * L + (plus ? L : 0) + line_index<NP>(node, dir);
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    const R wn = sU[(v * EPB + eb) * C::FITEMS + i];
                    gref[dir][v] += R(0.5) * (plus ? (wn - wrel[v]) : (wrel[v] - wn)) * (R)A.T.inv_w0;
                }
            }
        }
        R g[DIM][NEQ];
#pragma unroll
        for (int c = 0; c < DIM; c++)
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                R t = R(0);
#pragma unroll
                for (int r = 0; r < DIM; r++) t += gref[r][v] * adj[r][c];
                g[c][v] = t * invJ;
            }
        R sr[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) sr[v] = (R)s[v];
#pragma unroll
        for (int c = 0; c < DIM; c++) entgrad_to_primgrad<DIM, R>(sr, A.P, g[c]);
        if (A.grad) {
#pragma unroll
            for (int c = 0; c < DIM; c++)
#pragma unroll
                for (int v = 0; v < NEQ; v++)
                    A.grad[((size_t)c * NEQ + v) * vs + (size_t)e * NN + node] = (double)g[c][v];
        }
        R flux[DIM + 1][DIM];
        viscous_flux<DIM, R>(sr, g, A.P, flux);
        R fluxb[BUG2 ? DIM + 1 : 1][DIM];
        if constexpr (BUG2) {
            R sl[NEQ];
#pragma unroll
            for (int v = 0; v < NEQ; v++) sl[v] = (R)A.u[v * vs + (size_t)e * NN + NN - 1];
            viscous_flux<DIM, R>(sl, g, A.P, fluxb);
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const int a = ai[d];
            const int t = line_index<NP>(node, d);
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) {
                R f = R(0);
#pragma unroll
                for (int c = 0; c < DIM; c++) f += adj[d][c] * flux[v][c];
                if constexpr (BUG2) {
                    R fb = R(0);
#pragma unroll
                    for (int c = 0; c < DIM; c++) fb += adj[d][c] * fluxb[v][c];
                    Fc[d][v] = fb;
                } else {
                    Fc[d][v] = f;
                }
                if constexpr (MIX) {
                    if (a == 0) A.Fff[ftrace<DIM, NP>(d, 0, v, e, t, NE)] = f;
                    if (a == NP - 1) A.Fff[ftrace<DIM, NP>(d, 1, v, e, t, NE)] = f;
                } else if constexpr (!BND) {
                    if (a == 0) A.Ff[ftrace<DIM, NP>(d, 0, v, e, t, NE)] = f;
                    if (a == NP - 1) A.Ff[ftrace<DIM, NP>(d, 1, v, e, t, NE)] = f;
                }
            }
            if constexpr (BND) {
                // a face point on a boundary stores the condition's viscous flux (along +adj_d)
#pragma unroll
                for (int side = 0; side < 2; side++) {
                    if (a != (side ? NP - 1 : 0)) continue;
                    const int nb = sNbr[eb][2 * d + side];
                    double tr[NEQ - 1];
                    if (nb < 0) {
                        double adjd[DIM];
#pragma unroll
                        for (int c = 0; c < DIM; c++) adjd[c] = adj[d][c];
                        bc_viscous_trace<DIM>(A.bc[-nb - 1], A.P, sr, g, adjd, side, Fc[d], tr);
                    } else {
#pragma unroll
                        for (int v = 0; v < NEQ - 1; v++) tr[v] = Fc[d][v];
                    }
#pragma unroll
                    for (int v = 0; v < NEQ - 1; v++) A.Ff[ftrace<DIM, NP>(d, side, v, e, t, NE)] = tr[v];
                }
            }
        }
    }
    __syncthreads();  // sU (neighbour w) dead: reuse it for F~^d
    if (active) {
        // the MFMA divergence reads the xi and eta fluxes at swizzled nodes, the zeta flux from
        // the node's own register
#pragma unroll
        for (int d = 0; d < (LIFT_MFMA ? 2 : DIM); d++)
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) sU[(d * (NEQ - 1) + v) * EPB * NN + lsw] = Fc[d][v];
    }
    __syncthreads();
    if (!active) return;
    R dvv[NEQ - 1];
    if constexpr (LIFT_MFMA) {
        // the D-hat divergence on the matrix cores, as the gradient above: zeta and xi accumulate
        // in one result (both land on the node's lane), eta is moved to its lane
        const int lw = tid & 63, base = tid - lw;
        const int sx = ((lw >> 4) & 3) | (lw & 12) | ((lw & 3) << 4);
        const int sy = (lw & 3) | (((lw >> 4) & 3) << 2) | (((lw >> 2) & 3) << 4);
        const int ax = base + node_swz(sx), ay = base + node_swz(sy);
        const double Dh = A.T.Dhat[lw & 3][lw >> 4];
#pragma unroll
        for (int v = 0; v < NEQ - 1; v++) {
            double acc = __builtin_amdgcn_mfma_f64_4x4x4f64(Dh, (double)Fc[2][v], 0.0, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f64_4x4x4f64((double)sU[(0 * (NEQ - 1) + v) * EPB * NN + ax], Dh, acc, 0, 0, 0);
            const double dy = __builtin_amdgcn_mfma_f64_4x4x4f64(Dh, (double)sU[(1 * (NEQ - 1) + v) * EPB * NN + ay],
                                                                0.0, 0, 0, 0);
            dvv[v] = acc + __shfl(dy, sy, 64);
        }
    } else {
#pragma unroll
        for (int v = 0; v < NEQ - 1; v++) dvv[v] = R(0);
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const int st = dir_stride<NP>(d);
            const int a = ai[d];
            const int lb = tid - a * st;
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++)
#pragma unroll
                for (int b = 0; b < NP; b++)
                    dvv[v] += (R)A.T.Dhat[a][b] * sU[(d * (NEQ - 1) + v) * EPB * NN + lb + b * st];
        }
    }
#pragma unroll
    for (int v = 0; v < NEQ - 1; v++) {
        if constexpr (MIX)
            A.dvf[(size_t)v * vs + (size_t)e * NN + node] = dvv[v] * invJ;
        else
            A.dv[(size_t)v * vs + (size_t)e * NN + node] = dvv[v] * invJ;
    }
}

// ---------------------------------------------------------------------------
// k_vjp: the discrete adjoint of the Euler operator, ubar = (dL/du)^T lambda, with an
// optional fused LSRK4(5) stage adjoint.  Forward stage s (2N storage):
//     k_s = A_s k_{s-1} + dt L(u_{s-1}),   u_s = u_{s-1} + B_s k_s;
// reverse, given the adjoints (ub, kb) of (u_s, k_s):
//     lambda = kb + B_s ub,   ub_{s-1} = ub + dt (dL/du(u_{s-1}))^T lambda,   kb_{s-1} = A_s lambda.
// k_vjp_prep evaluates, once per node, the quantities every pair reads (density, velocity,
// beta and their reciprocals; SoA, NQ = DIM + 4 arrays) and lambda.  k_vjp then
// works as k_rhs (one thread per GLL line, each symmetric pair once): a pair's cotangent is
// Fb = -(1/J)(Dhat2[a][b] lambda_a + Dhat2[b][a] lambda_b), its reverse accumulates the
// adjoints of (rho, v, beta) of the line's nodes in registers, and those are mapped to
// conserved adjoints once per node and line.  A face point's cotangent is
// fb = -lambda_1/(w0 J1) + lambda_2/(w0 J2).  Each element accumulates only its own nodes (face
// adjoints are evaluated by both neighbours), so no atomics.
// ---------------------------------------------------------------------------
// VJP_SSP: the SSP-RK3 stage adjoint in the form out = x + B (lambda + dt (dL/du)^T lambda)
// (x optional), lambda the plain cotangent.
enum VjpMode { VJP_PLAIN = 0, VJP_LSRK = 1, VJP_SSP = 2 };

template <int NP>
struct VjpArgs {
    DTables<NP> T;
    PhysConst P;
    int NE, bug3, mode;
    const int *nbr;
    const double *geom;
    const double *u;      // state at which the Jacobian is taken (the stage input)
    const double *lam;    // VJP_PLAIN, VJP_SSP: the cotangent
    const double *x;      // VJP_SSP: the additive term (or null)
    const double *ub;     // VJP_LSRK: adjoint of u_s
    const double *kb;     // VJP_LSRK: adjoint of k_s
    double *q;            // per-node quantities, NQ x NE x Np^DIM (written by k_vjp_prep)
    double *lamw;         // VJP_LSRK: lambda = kb + B_s ub (written by k_vjp_prep)
    double *out;          // VJP_PLAIN: (dL/du)^T lambda; VJP_LSRK: adjoint of u_{s-1}
    double *kb_out;       // VJP_LSRK: adjoint of k_{s-1}
    double ca, cb, dt;    // A_s, B_s, dt
};

template <int DIM, int NP>
__global__ void __launch_bounds__(256) k_vjp_prep(const VjpArgs<NP> A) {
    constexpr int NEQ = DIM + 2, NN = ipow(NP, DIM);
    const size_t nodes = (size_t)A.NE * NN;
    for (size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gi < nodes; gi += (size_t)gridDim.x * blockDim.x) {
        double s[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) s[v] = A.u[v * nodes + gi];
        const adj::NodeQ<DIM> q = adj::node_q<DIM>(s, A.P.gammaM1);
        A.q[gi] = q.rho;
#pragma unroll
        for (int d = 0; d < DIM; d++) A.q[(1 + d) * nodes + gi] = q.v[d];
        A.q[(DIM + 1) * nodes + gi] = q.beta;
        A.q[(DIM + 2) * nodes + gi] = q.irho;
        A.q[(DIM + 3) * nodes + gi] = q.ibeta;
        if (A.mode == VJP_LSRK) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) A.lamw[v * nodes + gi] = A.kb[v * nodes + gi] + A.cb * A.ub[v * nodes + gi];
        }
    }
}

template <int DIM>
__device__ __forceinline__ adj::NodeQ<DIM> load_q(const double *Q, size_t nodes, size_t gi) {
    adj::NodeQ<DIM> q;
    q.rho = Q[gi];
#pragma unroll
    for (int d = 0; d < DIM; d++) q.v[d] = Q[(1 + d) * nodes + gi];
    q.beta = Q[(DIM + 1) * nodes + gi];
    q.irho = Q[(DIM + 2) * nodes + gi];
    q.ibeta = Q[(DIM + 3) * nodes + gi];
    return q;
}

template <int DIM, int NP, int SURF>
__global__ void __launch_bounds__((LineCfg<DIM, NP>::BLOCK), PRANDTL_LINE_WAVES) k_vjp(const VjpArgs<NP> A) {
    using C = LineCfg<DIM, NP>;
    constexpr int NEQ = DIM + 2, NN = C::NN, L = C::L, EPB = C::EPB;
    __shared__ double sB[NEQ][EPB * NN];

    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int e0 = blk * EPB;
    const int tid = threadIdx.x;
    const int NE = A.NE;
    const size_t vs = (size_t)NE * NN;
    const size_t off0 = (size_t)e0 * NN;
    const int nvalid = min(EPB, NE - e0) * NN;
    const int eb = tid / L, l = tid - (tid / L) * L;
    const int e = e0 + eb;
    const bool active = (eb < EPB) && (e < NE);
    const double *lamv = A.mode == VJP_LSRK ? A.lamw : A.lam;
    double invJ = 0.0, adjm[DIM][DIM];
    if (active) {
        const double *g = A.geom + (size_t)e * (1 + DIM * DIM);
        invJ = adj::inv(g[0]);
#pragma unroll
        for (int r = 0; r < DIM; r++)
#pragma unroll
            for (int c = 0; c < DIM; c++) adjm[r][c] = g[1 + r * DIM + c];
    }
    const double gm1 = A.P.gammaM1, gm1inv = A.P.gammaM1Inverse;
#pragma unroll
    for (int dir = 0; dir < DIM; dir++) {
        if (active) {
            const int base = line_base<NP>(l, dir);
            const int st = dir_stride<NP>(dir);
            const size_t ebase = (size_t)e * NN + base;
            const double *m = adjm[dir];
            adj::PrimB<DIM> acc[NP];
#pragma unroll
            for (int a = 0; a < NP; a++) adj::primb_zero<DIM>(acc[a]);
#pragma unroll 1
            for (int a = 0; a < NP - 1; a++) {
                const size_t ga = ebase + a * st;
                const adj::NodeQ<DIM> qa = load_q<DIM>(A.q, vs, ga);
                double la[NEQ];
#pragma unroll
                for (int v = 0; v < NEQ; v++) la[v] = lamv[v * vs + ga];
#pragma unroll 1
                for (int b = a + 1; b < NP; b++) {
                    const size_t gb = ebase + b * st;
                    const adj::NodeQ<DIM> qb = load_q<DIM>(A.q, vs, gb);
                    const double dab = A.T.Dhat2[a][b], dba = A.T.Dhat2[b][a];
                    double Fb[NEQ];
#pragma unroll
                    for (int v = 0; v < NEQ; v++) Fb[v] = -invJ * (dab * la[v] + dba * lamv[v * vs + gb]);
                    adj::PrimB<DIM> pa, pb;
                    adj::primb_zero<DIM>(pa);
                    adj::primb_zero<DIM>(pb);
                    adj::ec_core_back<DIM>(qa, qb, m, gm1inv, Fb, pa, pb);
                    // dynamic register indices avoided: add through a switch on the line position
#pragma unroll
                    for (int r = 0; r < NP; r++) {
                        if (r == a) {
                            acc[r].rho += pa.rho;
                            acc[r].beta += pa.beta;
#pragma unroll
                            for (int d = 0; d < DIM; d++) acc[r].v[d] += pa.v[d];
                        }
                        if (r == b) {
                            acc[r].rho += pb.rho;
                            acc[r].beta += pb.beta;
#pragma unroll
                            for (int d = 0; d < DIM; d++) acc[r].v[d] += pb.v[d];
                        }
                    }
                }
            }
            // conserved adjoints of the line's nodes
#pragma unroll
            for (int a = 0; a < NP; a++) {
                const size_t ga = ebase + a * st;
                const adj::NodeQ<DIM> qa = load_q<DIM>(A.q, vs, ga);
                double sb[NEQ];
#pragma unroll
                for (int v = 0; v < NEQ; v++) sb[v] = 0.0;
                adj::prim_back<DIM>(qa, gm1, acc[a], sb);
                const int li = eb * NN + base + a * st;
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    if (dir == 0)
                        sB[v][li] = sb[v];
                    else
                        sB[v][li] += sb[v];
                }
            }
            // surface terms at both ends of the line
#pragma unroll
            for (int side = 1; side >= 0; side--) {
                const int en = A.nbr[(size_t)e * 2 * DIM + 2 * dir + side];
                const int nown = base + (side ? (NP - 1) * st : 0);
                const int nnb = base + (side ? 0 : (NP - 1) * st);
                const double swo = A.T.inv_w0 * invJ;
                const double swn = A.T.inv_w0 * adj::inv(A.geom[(size_t)en * (1 + DIM * DIM)]);
                double so[NEQ], sn[NEQ], fb[NEQ], b1[NEQ], b2[NEQ];
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    const size_t go = v * vs + (size_t)e * NN + nown, gn = v * vs + (size_t)en * NN + nnb;
                    so[v] = A.u[go];
                    sn[v] = A.u[gn];
                    // side 1: own = el1 (gets -f/(w0 J)), neighbour = el2; side 0: the reverse
                    fb[v] = side ? (-swo * lamv[go] + swn * lamv[gn]) : (-swn * lamv[gn] + swo * lamv[go]);
                    b1[v] = b2[v] = 0.0;
                }
                if (side)
                    adj::face_flux_vjp<DIM>(SURF, so, sn, m, A.P.gamma, A.bug3 != 0, fb, b1, b2);
                else
                    adj::face_flux_vjp<DIM>(SURF, sn, so, m, A.P.gamma, A.bug3 != 0, fb, b1, b2);
                const int li = eb * NN + nown;
#pragma unroll
                for (int v = 0; v < NEQ; v++) sB[v][li] += side ? b1[v] : b2[v];
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < nvalid; idx += C::BLOCK) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            const size_t gi = v * vs + off0 + idx;
            const double g = sB[v][idx];
            if (A.mode == VJP_PLAIN) {
                A.out[gi] = g;
            } else if (A.mode == VJP_LSRK) {
                A.out[gi] = A.ub[gi] + A.dt * g;
                A.kb_out[gi] = A.ca * A.lamw[gi];
            } else {
                A.out[gi] = (A.x ? A.x[gi] : 0.0) + A.cb * (A.lam[gi] + A.dt * g);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The viscous part of the Navier-Stokes operator adjoint (constant viscosity), in two passes
// around k_vjp.  Forward (k_lift and the viscous terms of the RHS): w = w(u); BR1 lifting
// g_ref[r] = D w along r-lines + 1/2 (w+ - w-)/w0 at faces; g_w[c] = (1/J) sum_r adj[r][c]
// g_ref[r]; g_p[c] = EntropyGrad2PrimGrad(u, g_w[c]); Fv = viscous_flux(u, g_p); Fc[d] =
// adj[d] . Fv; du += (1/J) sum_d Dhat Fc[d] + the face average -/+ 1/2 (Fc- + Fc+)/(w0 J).
// k_vjp_visc1 (element per workgroup, thread per node) recomputes the forward lifting and
// reverses everything except the neighbour side of the lifting jump: it writes
// the adjoint of g_ref at the element's face points (Gf, for the neighbours), and per node the
// partial adjoints of w (WB) and of u (SB).  k_vjp_visc2 (thread per node) adds the
// neighbours' jump contributions to WB, reverses w(u), and adds the result to the output of
// k_vjp (times dt for the LSRK stage adjoint).
// ---------------------------------------------------------------------------
template <int NP>
struct VjpViscArgs {
    DTables<NP> T;
    PhysConst P;
    int NE, mode;
    double dt, cb;
    const int *nbr;
    const double *geom;
    const double *u;
    const double *lam;    // the cotangent lambda (VJP_PLAIN, VJP_SSP) or kb + B_s ub (VJP_LSRK)
    double *Gf;           // adjoint of g_ref at face points [r][side][v][e][t]
    double *WB, *SB;      // partial adjoints of w and u [v][e][node]
    double *out;          // k_vjp's output, completed by k_vjp_visc2
};

template <int DIM, int NP>
__device__ __forceinline__ size_t gslot(int dir, int side, int v, int e, int t, int NE) {
    constexpr int L = ipow(NP, DIM - 1);
    return ((((size_t)(dir * 2 + side) * (DIM + 2) + v) * NE + e) * L) + t;
}

template <int DIM, int NP>
__global__ void __launch_bounds__((LiftCfg<DIM, NP>::BLOCK)) k_vjp_visc1(const VjpViscArgs<NP> A) {
    using C = LiftCfg<DIM, NP>;
    constexpr int NEQ = DIM + 2, NN = C::NN, EPB = C::EPB, L = C::L;
    __shared__ double sW[NEQ][EPB * NN];               // w, then lambda, then adjoint of g_ref
    __shared__ double sU[NEQ][EPB * C::FITEMS];        // neighbour w at face points
    __shared__ double sG[DIM][NEQ][EPB * NN];          // adjoint of g_ref
    __shared__ LogRef sLR[EPB];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x;
    const int eb = tid / NN, node = tid - (tid / NN) * NN;
    const int e0 = blk * EPB, e = e0 + eb;
    const int NE = A.NE;
    const int nel = min(EPB, NE - e0);
    const bool active = (eb < EPB) && (e < NE);
    const size_t vs = (size_t)NE * NN;
    double s[NEQ];
    if (active) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) s[v] = A.u[v * vs + (size_t)e * NN + node];
        if (node == 0) sLR[eb] = make_log_ref<DIM>(s, A.P);
    }
    __syncthreads();
    double w[NEQ];
    if (active) {
        conserv_to_entropy_ref<DIM>(s, A.P, sLR[eb], w);
#pragma unroll
        for (int v = 0; v < NEQ; v++) sW[v][tid] = w[v];
    }
    for (int it = tid; it < nel * C::FITEMS; it += C::BLOCK) {
        const int fe = it / C::FITEMS, i = it - fe * C::FITEMS;
        const int dir = i / (2 * L), side = (i / L) & 1, t = i % L;
        const int st = dir_stride<NP>(dir);
        const int en = A.nbr[(size_t)(e0 + fe) * 2 * DIM + 2 * dir + side];
        const int nnb = line_base<NP>(t, dir) + (side ? 0 : (NP - 1) * st);
        double sn[NEQ], wn[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) sn[v] = A.u[v * vs + (size_t)en * NN + nnb];
        conserv_to_entropy_ref<DIM>(sn, A.P, sLR[fe], wn);
#pragma unroll
        for (int v = 0; v < NEQ; v++) sU[v][fe * C::FITEMS + i] = wn[v];
    }
    __syncthreads();
    int ai[3];
    ai[0] = node % NP;
    ai[1] = (node / NP) % NP;
    ai[2] = node / (NP * NP);
    double invJ = 0.0, adj[DIM][DIM], gw[DIM][NEQ], gp[DIM][NEQ];
    if (active) {
        const double *gm = A.geom + (size_t)e * (1 + DIM * DIM);
        invJ = rcp(gm[0]);
#pragma unroll
        for (int r = 0; r < DIM; r++)
#pragma unroll
            for (int c = 0; c < DIM; c++) adj[r][c] = gm[1 + r * DIM + c];
        double gref[DIM][NEQ];
#pragma unroll
        for (int r = 0; r < DIM; r++) {
            const int st = dir_stride<NP>(r);
            const int a = ai[r];
            const int lb = tid - a * st;
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < NP; b++) t += A.T.D[a][b] * sW[v][lb + b * st];
                gref[r][v] = t;
            }
            if (a == NP - 1 || a == 0) {
                const bool plus = (a == NP - 1);
                const int i = r * 2 * L + (plus ? L : 0) + line_index<NP>(node, r);
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    const double wn = sU[v][eb * C::FITEMS + i];
                    gref[r][v] += 0.5 * (plus ? (wn - w[v]) : (w[v] - wn)) * A.T.inv_w0;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < DIM; c++) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                double t = 0.0;
#pragma unroll
                for (int r = 0; r < DIM; r++) t += gref[r][v] * adj[r][c];
                gw[c][v] = t * invJ;
                gp[c][v] = gw[c][v];
            }
            entgrad_to_primgrad<DIM, double>(s, A.P, gp[c]);
        }
    }
    __syncthreads();   // sW (w) dead: load lambda
    if (active) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) sW[v][tid] = A.lam[v * vs + (size_t)e * NN + node];
    }
    __syncthreads();
    double sb[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) sb[v] = 0.0;
    if (active) {
        const double sw = A.T.inv_w0 * invJ;
        double Fcb[DIM][NEQ - 1];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const int st = dir_stride<NP>(d);
            const int b = ai[d];
            const int lb = tid - b * st;
#pragma unroll
            for (int v = 0; v < NEQ - 1; v++) {
                double t = 0.0;
#pragma unroll
                for (int a = 0; a < NP; a++) t += A.T.Dhat[a][b] * sW[1 + v][lb + a * st];
                Fcb[d][v] = t * invJ;
            }
#pragma unroll
            for (int side = 0; side < 2; side++) {
                if (b != (side ? NP - 1 : 0)) continue;
                const int en = A.nbr[(size_t)e * 2 * DIM + 2 * d + side];
                const double swn = A.T.inv_w0 * rcp(A.geom[(size_t)en * (1 + DIM * DIM)]);
                const int jn = node + (side ? -(NP - 1) : (NP - 1)) * st;
#pragma unroll
                for (int v = 0; v < NEQ - 1; v++) {
                    const double ln = A.lam[(1 + v) * vs + (size_t)en * NN + jn];
                    Fcb[d][v] += 0.5 * (side ? 1.0 : -1.0) * (sw * sW[1 + v][tid] - swn * ln);
                }
            }
        }
        double fb[DIM + 1][DIM];
#pragma unroll
        for (int v = 0; v < NEQ - 1; v++)
#pragma unroll
            for (int c = 0; c < DIM; c++) {
                double t = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) t += adj[d][c] * Fcb[d][v];
                fb[v][c] = t;
            }
        double gpb[DIM][NEQ], gwb[DIM][NEQ];
#pragma unroll
        for (int c = 0; c < DIM; c++)
#pragma unroll
            for (int v = 0; v < NEQ; v++) gpb[c][v] = gwb[c][v] = 0.0;
        adj::viscous_flux_vjp<DIM>(s, gp, A.P.gamma, A.P.mu, A.P.PrInverse, A.P.mu_bulk, fb, sb, gpb);
#pragma unroll
        for (int c = 0; c < DIM; c++) adj::entgrad_vjp<DIM>(s, gw[c], A.P.gamma, gpb[c], sb, gwb[c]);
#pragma unroll
        for (int r = 0; r < DIM; r++)
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                double t = 0.0;
#pragma unroll
                for (int c = 0; c < DIM; c++) t += adj[r][c] * gwb[c][v];
                t *= invJ;
                sG[r][v][tid] = t;
                const int a = ai[r];
                if (a == 0) A.Gf[gslot<DIM, NP>(r, 0, v, e, line_index<NP>(node, r), NE)] = t;
                if (a == NP - 1) A.Gf[gslot<DIM, NP>(r, 1, v, e, line_index<NP>(node, r), NE)] = t;
            }
    }
    __syncthreads();
    if (!active) return;
    double wb[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) wb[v] = 0.0;
#pragma unroll
    for (int r = 0; r < DIM; r++) {
        const int st = dir_stride<NP>(r);
        const int b = ai[r];
        const int lb = tid - b * st;
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < NP; a++) t += A.T.D[a][b] * sG[r][v][lb + a * st];
            wb[v] += t;
        }
        if (b == 0 || b == NP - 1) {
            const double c = (b == NP - 1 ? -0.5 : 0.5) * A.T.inv_w0;
#pragma unroll
            for (int v = 0; v < NEQ; v++) wb[v] += c * sG[r][v][tid];
        }
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
        A.WB[v * vs + (size_t)e * NN + node] = wb[v];
        A.SB[v * vs + (size_t)e * NN + node] = sb[v];
    }
}

template <int DIM, int NP>
__global__ void __launch_bounds__(256) k_vjp_visc2(const VjpViscArgs<NP> A) {
    constexpr int NEQ = DIM + 2, NN = ipow(NP, DIM);
    const size_t vs = (size_t)A.NE * NN;
    const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= vs) return;
    const int e = (int)(gi / NN), node = (int)(gi - (size_t)e * NN);
    double s[NEQ], wb[NEQ], sb[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
        s[v] = A.u[v * vs + gi];
        wb[v] = A.WB[v * vs + gi];
        sb[v] = A.SB[v * vs + gi];
    }
    // the neighbours' lifting jumps use this node's w: -/+ 1/2 / w0 times their g_ref adjoint
#pragma unroll
    for (int r = 0; r < DIM; r++) {
        const int b = (node / dir_stride<NP>(r)) % NP;
#pragma unroll
        for (int side = 0; side < 2; side++) {
            if (b != (side ? NP - 1 : 0)) continue;
            const int en = A.nbr[(size_t)e * 2 * DIM + 2 * r + side];
            const double c = (side ? -0.5 : 0.5) * A.T.inv_w0;
            const int t = line_index<NP>(node, r);
#pragma unroll
            for (int v = 0; v < NEQ; v++) wb[v] += c * A.Gf[gslot<DIM, NP>(r, 1 - side, v, en, t, A.NE)];
        }
    }
    adj::entropy_vars_vjp<DIM>(s, A.P.gamma, wb, sb);
    const double f = A.mode == VJP_LSRK ? A.dt : (A.mode == VJP_SSP ? A.cb * A.dt : 1.0);
#pragma unroll
    for (int v = 0; v < NEQ; v++) A.out[v * vs + gi] += f * sb[v];
}

}  // namespace prandtl
