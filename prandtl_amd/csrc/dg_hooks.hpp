// dg_hooks.hpp -- stage hooks on the GPU: the entropy filter and the positivity limiter.
//
// Entropy filter -- ref: src/Filters/EntropyFilter.cpp (GetEntropyConstraints :81-158,
// FilterSolution :171-242, ComputeHierarchStates :244-260, ComputeRoot :325-349) and
// src/ModalBasis/ModalBasis.cpp, with the reference's defects replaced by the intended
// behaviour: the entropy bound of an element is the minimum nodal entropy over the element
// and its face neighbours (ComputeElementMinima, :275-277, is empty in the reference), the
// root-find compares with the element's own bound (the reference reads the first element's,
// :338), and the entropy rho ln(p rho^-gamma) is used for Euler and Navier-Stokes states.
//
// The filtered state of node i is  u_i(zeta) = sum_j exp(-zeta j^2) H_j(i),  with H_j the part
// of the element's Legendre expansion whose modes have maximum degree j.  The reference forms
// H_j from V^-1 u and V; here H_j = S_j - S_{j-1} with S_j = (P_j x P_j [x P_j]) u, P_j the 1D
// projection onto Legendre degrees <= j on the GLL nodes (a tensor product, applied one
// direction at a time through LDS): the same polynomial, O(n Np) work per element instead of
// O(n^2).  The root-find is the reference's node-by-node scan: nodes are visited in order at
// the current zeta; the first inadmissible one runs the 20-step bisection on [zeta,
// -ln 1e-10] and raises zeta to the bracket's admissible end; the scan resumes after it.  A
// workgroup (one element, one thread per node) evaluates all remaining nodes at once and
// picks the first inadmissible one with an LDS atomicMin, which visits exactly the nodes the
// sequential scan visits, in its order, at its zeta.  The bisection itself runs in rounds of up
// to log2(block size) steps: in a round every lane evaluates one node of the round's bisection
// tree (heap order: node 1 the bracket's midpoint, 2l and 2l+1 the midpoints of node l's left and
// right halves), each computed by the halvings the sequential bisection would perform to reach
// it, and the path through the tree (right where the midpoint is inadmissible) ends on the
// sequential bisection's bracket, bit for bit.  Root-finding and the constraint tests are FP64.
//
// The detection scans append the elements that need a hook to a list (one counter per hook;
// the scan that fills one counter clears the other, which the previous application of that
// hook has finished reading), and the filter and limiter walk the list with a fixed grid, so
// elements that need nothing cost no workgroup.
//
// Positivity limiter -- the reference declares only the Limiter interface
// (src/Limiters/Limiter.hpp:10-16).  This is the scaling limiter of Zhang & Shu (J. Comput.
// Phys. 229 (2010) 8918-8934): with the GLL-mean state ubar of an element, rho is scaled
// toward rhobar by theta_1 = min(1, (rhobar - eps)/(rhobar - rho_min)), then the state toward
// ubar by theta_2 = min_i t_i, t_i the admissible end of a bisection for
// p(ubar + t (u_i - ubar)) = eps on [0, 1] down to a bracket of 2^-52.
#pragma once

#include "dg_kernels.hpp"

namespace prandtl {


template <int NP>
struct HookTables {
    double P[NP][NP][NP];   // P[j]: 1D projection onto Legendre degrees <= j at the GLL nodes
    double w[NP];           // GLL weights on [0,1]
};

template <int NP>
struct HookArgs {
    HookTables<NP> T;
    int NE;
    double gamma, gammaM1;
    const int *nbr;
    double *u;                     // state (filtered / limited in place)
    double *emin;                  // per-element minimum nodal entropy
    double *bound;                 // per-element entropy bound
    int *flag;                     // per-element: the element needs the hook
    unsigned long long *count;     // elements changed: [0] filter, [1] limiter; [2] nonzero: a
                                   // hook found its list counter above NE (hooks run out of order)
    int *list;                     // the elements that need the hook (k_hook_scan MODE 1, 2)
    unsigned int *nlist;           // their number (appended by the scan, read by the hook)
    unsigned int *nlist_next;      // the other counter of this hook: cleared by the scan
    const double *ngeom;           // general mesh: per-node (det J, adj J) -- the limiter's mean is
                                   // sum J w u / sum J w there; null on affine elements (J cancels)
    int inline_limit;              // k_filter: apply the limiter to a filtered element that is
                                   // still inadmissible (the stage-fused hooks: no limiter pass)
    int track_emin;                // k_filter, k_limiter: rewrite emin of the elements they change
};


template <int DIM>
__device__ __forceinline__ bool filter_violates(const double *s, double gamma, double gammaM1, double eb) {
    const double p = hook_pressure<DIM>(s, gammaM1);
    return s[0] < kFilterRhoMin || p < kFilterPMin || hook_entropy(s[0], p, gamma) < eb - kFilterEntTol;
}

template <int DIM, int NP>
struct HookCfg {
    static constexpr int NN = ipow(NP, DIM);
    static constexpr int EPB = 256 / NN > 0 ? 256 / NN : 1;   // elements per 256-thread block
    static constexpr int FBLOCK = NN <= 64 ? 64 : 128;         // filter / limiter: one element per workgroup
};

// Per-element pass over the nodes, one thread per node, EPB elements per block:
//   MODE 0: emin[e] = minimum nodal entropy (GetEntropyConstraints, EntropyFilter.cpp:113-157);
//   MODE 1: flag[e] = the element fails the filter's detection (:183-191);
//   MODE 2: flag[e] = some node has rho < eps or p < eps (the limiter's trigger).
template <int DIM, int NP, int MODE>
__global__ void __launch_bounds__(256) k_hook_scan(const HookArgs<NP> A) {
    using C = HookCfg<DIM, NP>;
    constexpr int NN = C::NN, EPB = C::EPB;
    __shared__ double sm[EPB * NN];
    __shared__ unsigned s_cnt, s_base;   // the block's appends to the list (MODE 1, 2)
    const int tid = threadIdx.x, eb = tid / NN, node = tid - eb * NN;
    const int e = blockIdx.x * EPB + eb;
    const size_t vs = (size_t)A.NE * NN;
    const bool active = eb < EPB && e < A.NE;
    if (active) {
        double s[DIM + 2];
#pragma unroll
        for (int v = 0; v < DIM + 2; v++) s[v] = A.u[v * vs + (size_t)e * NN + node];
        const double p = hook_pressure<DIM>(s, A.gammaM1);
        if (MODE == 0)
            sm[tid] = hook_entropy(s[0], p, A.gamma);
        else if (MODE == 1)
            sm[tid] = filter_violates<DIM>(s, A.gamma, A.gammaM1, A.bound[e]) ? 1.0 : 0.0;
        else
            sm[tid] = (s[0] < kLimiterEps || p < kLimiterEps) ? 1.0 : 0.0;
    }
    if (MODE != 0 && tid == 0) {
        s_cnt = 0u;
        if (blockIdx.x == 0) *A.nlist_next = 0u;
    }
    __syncthreads();
    if (MODE == 0) {
        if (active && node == 0) {
            double m = 1.7976931348623157e308;   // el_min_e = DBL_MAX (EntropyFilter.cpp:120)
            for (int i = 0; i < NN; i++) m = fmin(m, sm[eb * NN + i]);
            A.emin[e] = m;
        }
        return;
    }
    // the block's flagged elements take consecutive list slots: one global atomic per block
    int slot = -1;
    if (active && node == 0) {
        double f = 0.0;
        for (int i = 0; i < NN; i++) f = fmax(f, sm[eb * NN + i]);
        A.flag[e] = f > 0.0 ? 1 : 0;
        if (f > 0.0) slot = (int)atomicAdd(&s_cnt, 1u);
    }
    __syncthreads();
    if (tid == 0 && s_cnt > 0u) s_base = atomicAdd(A.nlist, s_cnt);
    __syncthreads();
    // a counter that did not start at zero (a hook application that never ran) cannot write
    // past the list
    if (slot >= 0 && s_base + (unsigned)slot < (unsigned)A.NE) A.list[s_base + (unsigned)slot] = e;
}

// entropy bound: minimum over the element and its 2 dim face neighbours
template <int DIM, int NP>
__global__ void __launch_bounds__(256) k_filter_bound(const HookArgs<NP> A) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= A.NE) return;
    double m = A.emin[e];
#pragma unroll
    for (int f = 0; f < 2 * DIM; f++) {
        const int nb = A.nbr[(size_t)e * 2 * DIM + f];
        if (nb >= 0) m = fmin(m, A.emin[nb]);   // boundary faces have no neighbour
    }
    A.bound[e] = m;
}

template <int NP>
__device__ __forceinline__ double filter_sigma(double zeta, int j) {
    return j == 0 ? 1.0 : exp(-zeta * (double)(j * j));
}

// FilterSolution: one element per workgroup at a time, one thread per node; the workgroups
// walk the list of detected elements
template <int DIM, int NP>
__global__ void __launch_bounds__((HookCfg<DIM, NP>::FBLOCK)) k_filter(const HookArgs<NP> A) {
    constexpr int NEQ = DIM + 2, NN = ipow(NP, DIM);
    constexpr int BL = HookCfg<DIM, NP>::FBLOCK, NW = BL / 64;
    static_assert(NN <= BL && (BL == 64 || BL == 128), "one element per workgroup of one or two waves");
    constexpr int tdepth = BL == 64 ? 6 : 7;   // steps per round: 2^tdepth - 1 lanes evaluate
    __shared__ double sU[NEQ][NN], sT[NEQ][NN];
    __shared__ double sH[NP][NEQ];      // the expansion of the node being root-found
    __shared__ double sP[NP][NP][NP];   // the 1D projections, read per lane
    __shared__ unsigned long long sMask[NW];   // a round's verdicts, one bit per heap node
    __shared__ int s_first;
    const int i = threadIdx.x;
    const bool node_ok = i < NN;
    for (int t = i; t < NP * NP * NP; t += BL) sP[t / (NP * NP)][(t / NP) % NP][t % NP] = A.T.P[t / (NP * NP)][(t / NP) % NP][t % NP];
    const size_t vs = (size_t)A.NE * NN;
    // the scan before this hook (on the same stream) started the counter at zero, so it holds at
    // most NE; more means the hooks of this operator ran out of order: report, change nothing
    unsigned nlist = *A.nlist;
    if (nlist > (unsigned)A.NE) {
        if (blockIdx.x == 0 && threadIdx.x == 0) A.count[2] = 1ull;
        nlist = 0u;
    }
    for (unsigned li = blockIdx.x; li < nlist; li += gridDim.x) {
    const int e = A.list[li];   // detected (EntropyFilter.cpp:183-191) by k_hook_scan<.., 1>
    const double eb = A.bound[e];
    if (node_ok) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) sU[v][i] = A.u[v * vs + (size_t)e * NN + i];
    }
    __syncthreads();
    const int i0 = i % NP, i1 = (i / NP) % NP, i2 = i / (NP * NP);
    // H_j (registers) from S_j = tensor projection of u
    double H[NP][NEQ], prev[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) prev[v] = 0.0;
#pragma unroll
    for (int j = 0; j < NP; j++) {
        double t[NEQ];
        // direction 0
        if (node_ok) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                double a = 0.0;
#pragma unroll
                for (int k = 0; k < NP; k++) a += sP[j][i0][k] * sU[v][i - i0 + k];
                t[v] = a;
            }
        }
        __syncthreads();
        if (node_ok)
#pragma unroll
            for (int v = 0; v < NEQ; v++) sT[v][i] = t[v];
        __syncthreads();
        // direction 1
        if (node_ok) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                double a = 0.0;
#pragma unroll
                for (int k = 0; k < NP; k++) a += sP[j][i1][k] * sT[v][i + (k - i1) * NP];
                t[v] = a;
            }
        }
        if constexpr (DIM == 3) {
            __syncthreads();
            if (node_ok)
#pragma unroll
                for (int v = 0; v < NEQ; v++) sT[v][i] = t[v];
            __syncthreads();
            if (node_ok) {
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    double a = 0.0;
#pragma unroll
                    for (int k = 0; k < NP; k++) a += sP[j][i2][k] * sT[v][i + (k - i2) * NP * NP];
                    t[v] = a;
                }
            }
        }
        __syncthreads();   // sT is rewritten by the next j
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            H[j][v] = t[v] - prev[v];
            prev[v] = t[v];
        }
    }
    // node-by-node scan (EntropyFilter.cpp:200-212) with ComputeRoot (:325-349)
    const double zmax = -log(kFilterEpsilon);
    double zeta = 0.0;
    int start = 0;
    for (int guard = 0; guard <= NN; guard++) {
        if (i == 0) s_first = NN;
        __syncthreads();
        if (node_ok && i >= start) {
            double st[NEQ];
#pragma unroll
            for (int v = 0; v < NEQ; v++) {
                double a = H[0][v];
#pragma unroll
                for (int j = 1; j < NP; j++) a += H[j][v] * filter_sigma<NP>(zeta, j);
                st[v] = a;
            }
            if (filter_violates<DIM>(st, A.gamma, A.gammaM1, eb)) atomicMin(&s_first, i);
        }
        __syncthreads();
        const int first = s_first;
        if (first >= NN) break;
        if (i == first) {
#pragma unroll
            for (int j = 0; j < NP; j++)
#pragma unroll
                for (int v = 0; v < NEQ; v++) sH[j][v] = H[j][v];
        }
        __syncthreads();
        // the 20-step bisection on [zeta, zmax] for node `first`, in rounds of tree evaluations
        double z1 = zeta, z2 = zmax;
        for (int done = 0; done < kFilterNIter;) {
            const int d = min(tdepth, kFilterNIter - done);
            bool viol = false;
            if (i >= 1 && i < (1 << d)) {
                // heap node i: follow its path from the root, halving as the bisection does
                double a = z1, b = z2;
                const int depth = 31 - __clz(i);
                for (int k = depth - 1; k >= 0; k--) {
                    const double m = 0.5 * (a + b);
                    if ((i >> k) & 1)
                        a = m;
                    else
                        b = m;
                }
                const double z3 = 0.5 * (a + b);
                double st[NEQ];
#pragma unroll
                for (int v = 0; v < NEQ; v++) {
                    double acc = sH[0][v];
#pragma unroll
                    for (int j = 1; j < NP; j++) acc += sH[j][v] * filter_sigma<NP>(z3, j);
                    st[v] = acc;
                }
                viol = filter_violates<DIM>(st, A.gamma, A.gammaM1, eb);
            }
            // the verdicts as bits (heap node n: bit n % 64 of wave n / 64), then the walk down
            // the tree on registers
            unsigned long long m[NW];
            m[0] = __ballot(viol);
            if constexpr (NW > 1) {
                if ((i & 63) == 0) sMask[i >> 6] = m[0];
                __syncthreads();
#pragma unroll
                for (int w = 0; w < NW; w++) m[w] = sMask[w];
            }
            int nd = 1;
            for (int k = 0; k < d; k++) {
                const double z3 = 0.5 * (z1 + z2);
                const unsigned long long mw = (NW == 1 || nd < 64) ? m[0] : m[NW - 1];
                if ((mw >> (nd & 63)) & 1ull) {
                    z1 = z3;
                    nd = 2 * nd + 1;
                } else {
                    z2 = z3;
                    nd = 2 * nd;
                }
            }
            done += d;
            if constexpr (NW > 1) __syncthreads();   // sMask is rewritten by the next round
        }
        zeta = z2;
        start = first + 1;
    }
    // reconstruction (EntropyFilter.cpp:214-238)
    double sf[NEQ];
    if (node_ok) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            double a = H[0][v];
#pragma unroll
            for (int j = 1; j < NP; j++) a += H[j][v] * filter_sigma<NP>(zeta, j);
            A.u[v * vs + (size_t)e * NN + i] = a;
            sf[v] = a;
        }
    }
    if (i == 0) atomicAdd(A.count, 1ull);
    __syncthreads();   // sU is rewritten for the next element
    if (A.inline_limit) {
        // the limiter's trigger on the filtered state (elements the filter did not visit pass its
        // detection, hence rho, p >= 1e-8: only filtered elements can need the limiter)
        const int bad = node_ok && (sf[0] < kLimiterEps || hook_pressure<DIM>(sf, A.gammaM1) < kLimiterEps);
        if (__syncthreads_or(bad)) limit_element<DIM, NP, NN>(A, e, i, sf, sU, sT[0], sT[1], sT[2]);
    }
    if (A.track_emin) element_emin<DIM, NP, NN>(A, e, i, sf, sU[0]);
    }
}

// The Zhang-Shu limiter of element e in place: node i (< NN) holds its state in s; the workgroup
// calls this together (it synchronises).  sR (NEQ x NN), sMean (NEQ), sW (NN), sScal (2) are LDS.
template <int DIM, int NP, int NN>
__device__ void limit_element(const HookArgs<NP> &A, int e, int i, double *s, double (*sR)[NN], double *sMean,
                              double *sW, double *sScal) {
    constexpr int NEQ = DIM + 2;
    const bool node_ok = i < NN;
    const size_t vs = (size_t)A.NE * NN;
    const double eps = kLimiterEps;
    if (node_ok) {
        double w = A.T.w[i % NP] * A.T.w[(i / NP) % NP] * (DIM == 3 ? A.T.w[i / (NP * NP)] : 1.0);
        if (A.ngeom) w *= A.ngeom[((size_t)e * NN + i) * (1 + DIM * DIM)];
        sW[i] = w;
#pragma unroll
        for (int v = 0; v < NEQ; v++) sR[v][i] = w * s[v];
    }
    __syncthreads();
    if (i < NEQ) {
        double m = 0.0;
        for (int k = 0; k < NN; k++) m += sR[i][k];
        if (A.ngeom) {
            double ws = 0.0;
            for (int k = 0; k < NN; k++) ws += sW[k];
            m /= ws;
        }
        sMean[i] = m;
    }
    __syncthreads();
    if (node_ok) sR[0][i] = s[0];
    __syncthreads();
    if (i == 0) {
        double r = sR[0][0];
        for (int k = 1; k < NN; k++) r = fmin(r, sR[0][k]);
        sScal[0] = r;     // rho_min
        sScal[1] = 1.0;   // theta_2
    }
    __syncthreads();
    double mean[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) mean[v] = sMean[v];
    // density (theta_1)
    const double rmin = sScal[0];
    if (rmin < eps && node_ok) {
        const double th = fmax(0.0, fmin(1.0, (mean[0] - eps) / (mean[0] - rmin)));
        s[0] = mean[0] + th * (s[0] - mean[0]);
    }
    // pressure (theta_2 = min_i t_i)
    if (node_ok && hook_pressure<DIM>(s, A.gammaM1) < eps) {
        double lo = 0.0, hi = 1.0, st[NEQ];
        while (hi - lo > 0x1p-52) {
            const double mid = 0.5 * (lo + hi);
#pragma unroll
            for (int v = 0; v < NEQ; v++) st[v] = mean[v] + mid * (s[v] - mean[v]);
            if (hook_pressure<DIM>(st, A.gammaM1) >= eps)
                lo = mid;
            else
                hi = mid;
        }
        // fmin over the workgroup through the ordered bit pattern of a non-negative double
        atomicMin(reinterpret_cast<unsigned long long *>(&sScal[1]), (unsigned long long)__double_as_longlong(lo));
    }
    __syncthreads();
    const double th2 = sScal[1];
    if (node_ok) {
        if (th2 < 1.0)
#pragma unroll
            for (int v = 0; v < NEQ; v++) s[v] = mean[v] + th2 * (s[v] - mean[v]);
#pragma unroll
        for (int v = 0; v < NEQ; v++) A.u[v * vs + (size_t)e * NN + i] = s[v];
    }
    if (i == 0) atomicAdd(A.count + 1, 1ull);
    __syncthreads();   // the LDS scratch is rewritten by the caller's next use
}

// The minimum nodal entropy of element e (node i holds its state in s) into A.emin[e], with the
// scan's starting value DBL_MAX; sR (NN) is LDS scratch; the workgroup calls this together.
template <int DIM, int NP, int NN>
__device__ void element_emin(const HookArgs<NP> &A, int e, int i, const double *s, double *sR) {
    if (i < NN) sR[i] = hook_entropy(s[0], hook_pressure<DIM>(s, A.gammaM1), A.gamma);
    __syncthreads();
    if (i == 0) {
        double m = 1.7976931348623157e308;
        for (int k = 0; k < NN; k++) m = fmin(m, sR[k]);
        A.emin[e] = m;
    }
    __syncthreads();
}

// Zhang-Shu limiter: one element per workgroup at a time, one thread per node; the workgroups
// walk the list of elements with a node below the bounds (admissible elements are untouched)
template <int DIM, int NP>
__global__ void __launch_bounds__(128) k_limiter(const HookArgs<NP> A) {
    constexpr int NEQ = DIM + 2, NN = ipow(NP, DIM);
    static_assert(NN <= 128, "one element per workgroup of at most 128 threads");
    __shared__ double sR[NEQ][NN];
    __shared__ double sMean[NEQ];
    __shared__ double sScal[2];
    __shared__ double sW[NN];
    const int i = threadIdx.x;
    const bool node_ok = i < NN;
    const size_t vs = (size_t)A.NE * NN;
    // the detection before this hook (on the same stream) started the counter at zero, so it holds
    // at most NE; more means the hooks of this operator ran out of order: report, change nothing
    unsigned nlist = *A.nlist;
    if (nlist > (unsigned)A.NE) {
        if (blockIdx.x == 0 && threadIdx.x == 0) A.count[2] = 1ull;
        nlist = 0u;
    }
    for (unsigned li = blockIdx.x; li < nlist; li += gridDim.x) {
        const int e = A.list[li];
        double s[NEQ];
        if (node_ok) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) s[v] = A.u[v * vs + (size_t)e * NN + i];
        }
        limit_element<DIM, NP, NN>(A, e, i, s, sR, sMean, sW, sScal);
        if (A.track_emin) element_emin<DIM, NP, NN>(A, e, i, s, sW);
    }
}

}  // namespace prandtl
