// dg_blend.hpp -- blending coefficient of the subcell finite-volume blending (gfx950).
//
// ref: PerssonPeraireIndicator::CheckSmoothness (src/Indicators/PerssonPeraireIndicator.cpp:21-58)
//      and DGSEMOperator::ComputeBlendingCoefficient (src/Operators/DGSEMOperator.cpp:60-80),
//      modalThreshold (:25), sharpness_fac (DGSEMOperator.hpp:50).
//
// Per element: q = rho p at the nodes; its modal coefficients in the tensor Legendre basis
// P_k0(x) P_k1(y) P_k2(z) on [0,1]^d (the reference's ModalBasis, ModalBasis.cpp:110-135, whose
// Vandermonde is the tensor product of the 1D one, V1[a][k] = P_k(2 x_a - 1)), computed by sum
// factorisation: V1^-1 applied along each direction in turn (Np multiply-adds per node and
// direction instead of the dense n x n product).  After the last pass the thread at node
// (a0, a1, a2) holds the mode of degrees (a0, a1, a2).  The three energies |M|^2, |M1|^2 (modes
// of maximum degree <= p - 1) and |M2|^2 (<= p - 2) are summed per element in a fixed order, so
// the result is reproducible; then eta = max(1 - |M1|^2/|M|^2, 1 - |M2|^2/|M1|^2) and
// alpha = 1/(1 + exp(-s (eta - T)/T)), clipped as the reference does.
//
// Included by dg_kernels.hpp after its layout helpers (NodeCfg, dir_stride, xcd_block): the
// blended RHS kernel computes the coefficients of its own elements with blend_coef_block, and
// k_blend_coef serves prandtl_blend_coefficients.
#pragma once

namespace prandtl {

template <int NP>
struct BlendArgs {
    BlendParams<NP> B;
    const double *u;
    double *alpha;         // out (may be null)
    double *eta;           // out (may be null)
    int NE;
    PhysConst P;
};

// rho p of one state (ComputePressure, BasicOperations.cpp:31-37: V = m / rho,
// p = (gamma-1)(E - rho V.V / 2))
template <int DIM>
__device__ __forceinline__ double rho_p(const double *s, const PhysConst &P) {
    const double gammaM1 = P.gammaM1;
    const double inv = 1.0 / s[0];
    double vv = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double vd = s[1 + d] * inv;
        vv += vd * vd;
    }
    return s[0] * (gammaM1 * (s[DIM + 1] - 0.5 * s[0] * vv));
}

// Orders the coefficient block's LDS accesses: a workgroup barrier, or, where every element's
// nodes lie in one wave (each step reads and writes only its own element's slots), the wave's own
// ordering of its LDS instructions.
template <bool WAVE>
__device__ __forceinline__ void coef_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Partial sums per modal energy: the largest divisor K of NN with 3 K <= NN, so that each of the
// 3 K partial sums (NN / K squares) has a lane of its own.
constexpr int energy_chunks(int nn) {
    for (int k = nn / 3; k > 1; k--)
        if (nn % k == 0) return k;
    return 1;
}

// The workgroup's elements (EPB of NN = Np^dim nodes, thread tid = eb NN + node) from q = rho p
// at the nodes: modes by three sum-factorised passes (double-buffered in sA / sB), the three
// modal energies as K partial sums over consecutive nodes then their sum in chunk order (a fixed
// order, so the result is reproducible), eta and alpha into sRes[eb] = {alpha, eta}.  Every
// thread of the workgroup calls it (it may contain barriers); sVi holds V1^-1 (Np x Np, row k =
// degree k) in LDS; sA, sB and sC hold EPB NN doubles each.  WAVE: every element lies in one wave
// (wave-local ordering in place of the barriers).  sRes is valid after the caller's next barrier.
template <int DIM, int NP, int EPB, bool WAVE>
__device__ __forceinline__ void blend_coef_block(double q, bool active, int tid, int eb, int node,
                                                 const BlendParams<NP> &B, const double *sVi, double *sA, double *sB,
                                                 double *sC, double (*sRes)[2]) {
    constexpr int NN = ipow(NP, DIM);
    constexpr int K = energy_chunks(NN), CH = NN / K;
    int ai[3];
    ai[0] = node % NP;
    ai[1] = (node / NP) % NP;
    ai[2] = node / (NP * NP);
    if (active) sA[tid] = q;
    coef_sync<WAVE>();
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double *src = (d & 1) ? sB : sA;
        double *dst = (d & 1) ? sA : sB;
        if (active) {
            const int st = dir_stride<NP>(d);
            const int a = ai[d];
            const int lb = tid - a * st;
            double t = 0.0;
#pragma unroll
            for (int b = 0; b < NP; b++) t += sVi[a * NP + b] * src[lb + b * st];
            q = t;
            if (d + 1 < DIM) dst[tid] = t;
        }
        coef_sync<WAVE>();   // the last pass's reads are done before sA / sB are reused below
    }
    // q: the mode of degrees (ai[0], ai[1], ai[2])
    if (active) {
        int md = ai[0];
#pragma unroll
        for (int d = 1; d < DIM; d++) md = max(md, ai[d]);
        const double q2 = q * q;
        sC[tid] = q2;
        sB[tid] = md <= NP - 2 ? q2 : 0.0;   // modesM1: no degree above p - 1
        sA[tid] = md <= NP - 3 ? q2 : 0.0;   // modesM2: no degree above p - 2
    }
    coef_sync<WAVE>();
    if (active && node < 3 * K) {
        // partial sum c of energy set: its chunk's squares, into the chunk's first slot (no
        // other lane reads that chunk)
        const int set = node / K, c = node - set * K;
        double *src = (set == 0 ? sC : (set == 1 ? sB : sA)) + eb * NN + c * CH;
        double acc = src[0];
#pragma unroll
        for (int i = 1; i < CH; i++) acc += src[i];
        src[0] = acc;
    }
    coef_sync<WAVE>();
    if (active && node == 0) {
        double mm = 0.0, m1 = 0.0, m2 = 0.0;
#pragma unroll
        for (int c = 0; c < K; c++) {
            mm += sC[eb * NN + c * CH];
            m1 += sB[eb * NN + c * CH];
            m2 += sA[eb * NN + c * CH];
        }
        double ind = 1.0 - m1 / mm;
        const double ind2 = 1.0 - m2 / m1;
        ind = ind2 > ind ? ind2 : ind;                       // std::max(ind, ind2)
        double a = 1.0 / (1.0 + exp(-B.sharpness * (ind - B.T) / B.T));
        if (a < B.alpha_min)
            a = 0.0;
        else if (a > (1.0 - B.alpha_min))
            a = 1.0;
        a = (B.alpha_max < a) ? B.alpha_max : a;             // std::min(a, alpha_max)
        sRes[eb][0] = a;
        sRes[eb][1] = ind;
    }
}

template <int DIM, int NP>
__global__ void __launch_bounds__((NodeCfg<DIM, NP>::BLOCK)) k_blend_coef(const BlendArgs<NP> A) {
    using C = NodeCfg<DIM, NP>;
    constexpr int NEQ = DIM + 2, NN = C::NN, EPB = C::EPB;
    __shared__ double sA[EPB * NN], sB[EPB * NN], sC[EPB * NN];
    __shared__ double sVi[NP * NP];
    __shared__ double sRes[EPB][2];

    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x;
    const int eb = tid / NN, node = tid - eb * NN;
    const int e = blk * EPB + eb;
    const bool active = (eb < EPB) && (e < A.NE);
    const size_t vs = (size_t)A.NE * NN;

    if (tid < NP * NP) sVi[tid] = A.B.Vi[tid / NP][tid % NP];
    double q = 0.0;
    if (active) {
        double s[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) s[v] = A.u[v * vs + (size_t)e * NN + node];
        q = rho_p<DIM>(s, A.P);
    }
    __syncthreads();   // sVi
    blend_coef_block<DIM, NP, EPB, C::BLOCK == 64 || 64 % NN == 0>(q, active, tid, eb, node, A.B, sVi, sA, sB, sC,
                                                                   sRes);
    if (active && node == 0) {
        if (A.alpha) A.alpha[e] = sRes[eb][0];
        if (A.eta) A.eta[e] = sRes[eb][1];
    }
}

}  // namespace prandtl
