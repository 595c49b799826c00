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
#pragma once

#include "dg_kernels.hpp"

namespace prandtl {

template <int NP>
struct BlendArgs {
    double Vi[NP][NP];     // V1^-1: Vi[k][a] = w_a P_k(x_a) / sum_b w_b P_k(x_b)^2
    const double *u;
    double *alpha;         // out (may be null)
    double *eta;           // out (may be null)
    int NE;
    double gammaM1;
    double T;              // modalThreshold 0.5 10^(-1.8 p^(1/4))
    double sharpness;      // 9.21024
    double alpha_max, alpha_min;
};

template <int DIM, int NP>
__global__ void __launch_bounds__((NodeCfg<DIM, NP>::BLOCK)) k_blend_coef(const BlendArgs<NP> A) {
    using C = NodeCfg<DIM, NP>;
    constexpr int NEQ = DIM + 2, NN = C::NN, EPB = C::EPB;
    __shared__ double sQ[EPB * NN];
    __shared__ double sE[3][EPB * NN];
    __shared__ double sSum[EPB][3];

    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x;
    const int eb = tid / NN, node = tid - eb * NN;
    const int e = blk * EPB + eb;
    const bool active = (eb < EPB) && (e < A.NE);
    const size_t vs = (size_t)A.NE * NN;

    double q = 0.0;
    if (active) {
        double s[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) s[v] = A.u[v * vs + (size_t)e * NN + node];
        // ComputePressure (BasicOperations.cpp:31-37): V = m / rho, p = (gamma-1)(E - rho V.V / 2)
        const double inv = 1.0 / s[0];
        double vv = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double vd = s[1 + d] * inv;
            vv += vd * vd;
        }
        q = s[0] * (A.gammaM1 * (s[NEQ - 1] - 0.5 * s[0] * vv));
    }
    int ai[3];
    ai[0] = node % NP;
    ai[1] = (node / NP) % NP;
    ai[2] = node / (NP * NP);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        if (active) sQ[tid] = q;
        __syncthreads();
        if (active) {
            const int st = dir_stride<NP>(d);
            const int a = ai[d];
            const int lb = tid - a * st;
            double t = 0.0;
#pragma unroll
            for (int b = 0; b < NP; b++) t += A.Vi[a][b] * sQ[lb + b * st];
            q = t;
        }
        __syncthreads();
    }
    if (active) {
        int md = ai[0];
#pragma unroll
        for (int d = 1; d < DIM; d++) md = max(md, ai[d]);
        const double q2 = q * q;
        sE[0][tid] = q2;
        sE[1][tid] = md <= NP - 2 ? q2 : 0.0;   // modesM1: no degree above p - 1
        sE[2][tid] = md <= NP - 3 ? q2 : 0.0;   // modesM2: no degree above p - 2
    }
    __syncthreads();
    if (active && node < 3) {
        double acc = 0.0;
        for (int i = 0; i < NN; i++) acc += sE[node][eb * NN + i];
        sSum[eb][node] = acc;
    }
    __syncthreads();
    if (active && node == 0) {
        const double mm = sSum[eb][0], m1 = sSum[eb][1], m2 = sSum[eb][2];
        double ind = 1.0 - m1 / mm;
        const double ind2 = 1.0 - m2 / m1;
        ind = ind2 > ind ? ind2 : ind;                       // std::max(ind, ind2)
        double a = 1.0 / (1.0 + exp(-A.sharpness * (ind - A.T) / A.T));
        if (a < A.alpha_min)
            a = 0.0;
        else if (a > (1.0 - A.alpha_min))
            a = 1.0;
        a = (A.alpha_max < a) ? A.alpha_max : a;             // std::min(a, alpha_max)
        if (A.alpha) A.alpha[e] = a;
        if (A.eta) A.eta[e] = ind;
    }
}

}  // namespace prandtl
