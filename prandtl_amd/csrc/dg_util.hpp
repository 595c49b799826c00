// dg_util.hpp -- small utility kernels of the C ABI (fills, copies, the adjoint sweep's
// combinations, the max wave speed).  Included by prandtl_capi.hip only: non-template
// kernels are defined once per library.
#pragma once

#include "dg_kernels.hpp"

namespace prandtl {

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------
// y = a x1 + b x2 (the adjoint sweep's combinations)
__global__ void k_axpby(double *y, double a, const double *x1, double b, const double *x2, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = a * x1[i] + b * x2[i];
}

// One stage of an explicit Runge-Kutta method in MFEM's ExplicitRKSolver order, from the stage's
// RHS k: acc = (first ? x : acc) + (cb dt) k, the running x + sum_j b_j dt k_j; unless last,
// y = x + (ca dt) k, the next stage's input; when last, the new state acc goes to x.
__global__ void k_erk_update(double *x, double *acc, double *y, const double *k, double ca, double cb, double dt,
                             int first, int last, size_t n) {
    const double bd = cb * dt, ad = ca * dt;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double xi = x[i];
        const double a = (first ? xi : acc[i]) + bd * k[i];
        if (last) {
            x[i] = a;
        } else {
            acc[i] = a;
            y[i] = xi + ad * k[i];
        }
    }
}

// non-finite scan of a state (the explicit RK steppers' NaN check)
__global__ void k_nan_scan(const double *x, size_t n, int *flag) {
    int bad = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad |= !(x[i] == x[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int *p, int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *p, double v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

__global__ void k_copy_f64(double *__restrict__ dst, const double *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) dst[i] = src[i];
}

// max over nodes of |V| + c (the reference's max_char_speed, ChandrashekarFlux.cpp:61-65): each
// wave stores its maximum at its global wave index; k_reduce_max combines them
template <int DIM>
__global__ void k_max_speed(const double *u, size_t nnodes, double gamma, double *lam_part) {
    constexpr int NEQ = DIM + 2;
    double lam = 0.0;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < nnodes; i += st) {
        double s[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) s[v] = u[v * nnodes + i];
        PhysConst P{};
        P.gamma = gamma;
        P.gammaM1 = gamma - 1.0;
        lam = fmax(lam, make_side<DIM>(s, P).lam);
    }
    for (int o = 32; o > 0; o >>= 1) lam = fmax(lam, __shfl_xor(lam, o));
    if ((threadIdx.x & 63) == 0) lam_part[((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6] = lam;
}

// *out = the bits of max(0, part[0..n-1]) (positive doubles order like their bit patterns); one
// workgroup
__global__ void __launch_bounds__(1024) k_reduce_max(const double *part, int n, unsigned long long *out) {
    __shared__ double s[16];
    // eight independent maxima per thread keep eight loads in flight
    const int bd = blockDim.x;
    double mk[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int i = threadIdx.x;
    for (; i + 7 * bd < n; i += 8 * bd) {
#pragma unroll
        for (int k = 0; k < 8; k++) mk[k] = fmax(mk[k], part[i + k * bd]);
    }
    for (; i < n; i += bd) mk[0] = fmax(mk[0], part[i]);
    double m = mk[0];
#pragma unroll
    for (int k = 1; k < 8; k++) m = fmax(m, mk[k]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r = fmax(r, s[w]);
        *out = (unsigned long long)__double_as_longlong(r);
    }
}

}  // namespace prandtl
