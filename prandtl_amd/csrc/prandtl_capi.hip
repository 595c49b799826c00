// prandtl_capi.hip -- C ABI of the gfx950 DGSEM library: operator handle, mesh
// validation, kernel dispatch and the fused RK steppers.
//
// Reference interfaces replaced are cited per entry point in include/prandtl_amd.h.

#include "op_internal.hpp"
#include "dg_util.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <memory>
#include <mutex>

using namespace prandtl;

namespace prandtl {

thread_local std::string g_last_error;

// roctx ranges around the library's phases (an RHS evaluation, the stage hooks, a run of RK
// steps, a VJP, an adjoint sweep) when PRANDTL_ROCTX is set to a nonzero value; rocprofv3
// --marker-trace records them.  They mark the host's enqueue intervals: the kernels a range
// launched run on the stream after it.
bool roctx_enabled() {
    static const bool on = [] {
        const char *v = std::getenv("PRANDTL_ROCTX");
        return v && std::atoi(v) != 0;
    }();
    return on;
}
struct RoctxRange {
    const bool on;
    explicit RoctxRange(const char *name) : on(roctx_enabled()) {
        if (on) roctxRangePushA(name);
    }
    ~RoctxRange() {
        if (on) roctxRangePop();
    }
    RoctxRange(const RoctxRange &) = delete;
    RoctxRange &operator=(const RoctxRange &) = delete;
};

// GLL rule on [0,1] from its closed form (0, 1 and the roots of P'_N(2x-1), weights
// 1/(N(N+1) P_N^2)), Newton's method in long double; the reference takes it from
// MFEM's IntegrationRules(0, GaussLobatto) (DGSEMIntegrator.cpp:18-37).
void gll_rule(int np, double *x, double *w) {
    const int N = np - 1;
    auto legendre = [N](long double z, long double &P, long double &dP) {
        long double p0 = 1.0L, p1 = z;
        for (int n = 1; n < N; n++) {
            long double p2 = ((2 * n + 1) * z * p1 - n * p0) / (n + 1);
            p0 = p1;
            p1 = p2;
        }
        P = p1;
        dP = N * (z * p1 - p0) / (z * z - 1.0L);
    };
    std::vector<long double> z(np);
    z[0] = -1.0L;
    z[N] = 1.0L;
    for (int j = 1; j < N; j++) {
        long double zz = -std::cos(3.14159265358979323846264338327950288L * j / N);
        for (int it = 0; it < 100; it++) {
            long double P, dP;
            legendre(zz, P, dP);
            const long double d2P = (2.0L * zz * dP - N * (N + 1.0L) * P) / (1.0L - zz * zz);
            const long double dz = dP / d2P;
            zz -= dz;
            if (std::fabs((double)dz) < 1e-30) break;
        }
        z[j] = zz;
    }
    for (int j = 0; j < np; j++) {
        long double P = 1.0L, dP;
        if (j == 0)
            P = (N % 2 == 1) ? -1.0L : 1.0L;
        else if (j < N)
            legendre(z[j], P, dP);
        w[j] = (double)(1.0L / (N * (N + 1.0L) * P * P));
        x[j] = (double)(0.5L * (z[j] + 1.0L));
    }
    x[0] = 0.0;
    x[N] = 1.0;
}


constexpr int kRevolveMaxSteps = 4096;

// Binomial checkpointing (Griewank, "Achieving logarithmic growth of temporal and spatial
// complexity in reverse automatic differentiation", 1992; the schedule of Griewank & Walther's
// revolve): T(n, s) = fewest forward steps to reverse n steps from their start state with s
// snapshot buffers, T(1, s) = 0, T(n, 0) = n (n - 1) / 2,
// T(n, s) = min_k [k + T(n - k, s - 1) + T(k, s)] (advance k steps, snapshot, reverse the n - k
// after it with s - 1 buffers, then the k before it with s); split(n, s) is the minimising k.
// Each reversed step also recomputes its own stages once (not counted in T).
struct RevolvePlan {
    int N = 0, S = 0;
    std::vector<long long> T;   // T[s][n], s = 0..S, n = 0..N
    std::vector<int> K;         // argmin k
    long long cost(int n, int s) const { return T[(size_t)s * (N + 1) + n]; }
    int split(int n, int s) const { return K[(size_t)s * (N + 1) + n]; }
    void build(int n_max, int s_max) {
        N = n_max;
        S = s_max;
        T.assign((size_t)(S + 1) * (N + 1), 0);
        K.assign((size_t)(S + 1) * (N + 1), 1);
        for (int n = 1; n <= N; n++) T[n] = (long long)n * (n - 1) / 2;
        for (int s = 1; s <= S; s++) {
            long long *Ts = &T[(size_t)s * (N + 1)];
            const long long *Tp = &T[(size_t)(s - 1) * (N + 1)];
            int *Ks = &K[(size_t)s * (N + 1)];
            for (int n = 2; n <= N; n++) {
                long long best = -1;
                int bk = 1;
                for (int k = 1; k < n; k++) {
                    const long long c = k + Tp[n - k] + Ts[k];
                    if (best < 0 || c < best) best = c, bk = k;
                }
                Ts[n] = best;
                Ks[n] = bk;
            }
        }
    }
};

}  // namespace prandtl

namespace {

// the plan of the handle's taped run (kept with the handle)
const RevolvePlan &revolve_plan(prandtl_op_s *op, int nsteps, int snaps) {
    if (!op->revolve || op->revolve->N < nsteps || op->revolve->S < snaps) {
        op->revolve = std::make_shared<RevolvePlan>();
        op->revolve->build(nsteps, std::min(snaps, nsteps));
    }
    return *op->revolve;
}

// identifiers of taped runs, unique over the handles of the process
unsigned long long next_tape_id() {
    static std::atomic<unsigned long long> counter{0};
    return ++counter;
}

// Barycentric differentiation matrix and its surface-corrected forms
// (DGSEMIntegrator.cpp:45-80), stored as mathematical matrices M[i][j].
void diff_matrices(int np, const double *x, const double *w, double *D, double *Dh, double *Dh2) {
    double lam[MAX_NP + 1];
    for (int i = 0; i < np; i++) lam[i] = 1.0;
    for (int i = 1; i < np; i++)
        for (int j = 0; j < i; j++) {
            lam[j] *= (x[j] - x[i]);
            lam[i] *= (x[i] - x[j]);
        }
    for (int i = 0; i < np; i++) lam[i] = 1.0 / lam[i];
    for (int i = 0; i < np * np; i++) D[i] = 0.0;
    for (int j = 0; j < np; j++)
        for (int i = 0; i < np; i++)
            if (i != j) {
                D[i * np + j] = lam[j] / lam[i] / (x[i] - x[j]);
                D[i * np + i] -= D[i * np + j];
            }
    for (int i = 0; i < np * np; i++) {
        Dh[i] = D[i];
        Dh2[i] = 2.0 * D[i];
    }
    Dh[0] += 1.0 / w[0];
    Dh[np * np - 1] -= 1.0 / w[np - 1];
    Dh2[0] += 1.0 / w[0];
    Dh2[np * np - 1] -= 1.0 / w[np - 1];
}

template <int DIM, int NP>
int launch_blend_coef(prandtl_op_s *op, const double *u, double *alpha, double *eta, hipStream_t st) {
    BlendArgs<NP> B{};
    B.B = make_blend_params<NP>(op);
    B.u = u;
    B.alpha = alpha;
    B.eta = eta;
    B.NE = op->NE;
    B.P = op->P;
    using C = NodeCfg<DIM, NP>;
    const int nblk = (op->NE + C::EPB - 1) / C::EPB;
    hipLaunchKernelGGL((k_blend_coef<DIM, NP>), dim3(nblk), dim3(C::BLOCK), 0, st, B);
    HIP_TRY(hipGetLastError());
    return PRANDTL_OK;
}

template <int DIM>
int dispatch_np(prandtl_op_s *op, const StageSpec &S, hipStream_t st) {
    switch (op->np) {
        case 2: return dispatch_visc<DIM, 2>(op, S, st);
        case 3: return dispatch_visc<DIM, 3>(op, S, st);
        case 4: return dispatch_visc<DIM, 4>(op, S, st);
        case 5: return dispatch_visc<DIM, 5>(op, S, st);
    }
    return fail(PRANDTL_ERR_UNSUPPORTED, "order out of range");
}

// the next pair of timing events (created on first use)
int timing_events(prandtl_op_s *op, int kind, hipEvent_t &a, hipEvent_t &b) {
    while (op->tev.size() < op->tev_used + 2) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        op->tev.push_back(e);
    }
    a = op->tev[op->tev_used];
    b = op->tev[op->tev_used + 1];
    op->tev_used += 2;
    op->tkind.push_back(kind);
    return PRANDTL_OK;
}

int dispatch_stage(prandtl_op_s *op, const StageSpec &S, hipStream_t st) {
    return op->dim == 2 ? dispatch_np<2>(op, S, st) : dispatch_np<3>(op, S, st);
}

int run_stage(prandtl_op_s *op, const StageSpec &S, hipStream_t st) {
    RoctxRange range("prandtl.rhs");
    int rc;
    if (op->timing && S.phases == 3) {
        // each pass between its own pair of events
        rc = PRANDTL_OK;
        for (int ph = op->phys.viscous ? 0 : 1; ph < 2 && rc == PRANDTL_OK; ph++) {
            hipEvent_t a, b;
            rc = timing_events(op, ph, a, b);
            if (rc) break;
            StageSpec P = S;
            P.phases = 1 << ph;
            HIP_TRY(hipEventRecord(a, st));
            rc = dispatch_stage(op, P, st);
            HIP_TRY(hipEventRecord(b, st));
        }
    } else {
        rc = dispatch_stage(op, S, st);
    }
    if (rc == PRANDTL_OK && S.want_lambda) {
        // the flux kernel left one maximum per wave
        hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, st, op->d_lam_part, op->lam_waves, op->d_lam);
        HIP_TRY(hipGetLastError());
    }
    return rc;
}

int ensure_rk_buffers(prandtl_op_s *op, int scheme) {
    const size_t bytes = op->size * sizeof(double);
    if (!op->d_ub) {
        HIP_TRY(hipMalloc(&op->d_ub, bytes));
    }
    if (scheme == PRANDTL_LSRK45 && !op->d_k) {
        HIP_TRY(hipMalloc(&op->d_k, bytes));
        hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, 0, op->d_k, 0.0, op->size);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
    }
    const bool erk = scheme == PRANDTL_FORWARD_EULER || scheme == PRANDTL_RK2 || scheme == PRANDTL_RK4;
    if (erk && !op->d_k) HIP_TRY(hipMalloc(&op->d_k, bytes));   // the stage RHS
    if ((scheme == PRANDTL_SSPRK3 || erk) && !op->d_y2) {
        HIP_TRY(hipMalloc(&op->d_y2, bytes));
    }
    return PRANDTL_OK;
}

void release_adjoint_storage(prandtl_op_s *op) {
    for (size_t i = 1; i < op->adj_stage.size(); i++) (void)hipFree(op->adj_stage[i]);
    op->adj_stage.clear();
    (void)hipFree(op->adj_ub2);
    (void)hipFree(op->adj_kb);
    (void)hipFree(op->adj_kb2);
    (void)hipFree(op->adj_work);
    op->adj_ub2 = op->adj_kb = op->adj_kb2 = op->adj_work = nullptr;
}

struct VjpSpec {
    int mode;
    const double *u, *lam, *ub, *kb;
    double *out, *kb_out;
    double ca, cb, dt;
    const double *x = nullptr;   // VJP_SSP additive term
};

// device work arrays of one operator-adjoint evaluation
struct VjpBuffers {
    double *q = nullptr;      // per-node quantities, (dim + 4) x nodes
    double *lamw = nullptr;   // VJP_LSRK: the stage cotangent (state size; unused by VJP_PLAIN, VJP_SSP)
    double *gf = nullptr;     // NS: lifting adjoint at face points, dim x 2 x neq x NE x Np^(dim-1)
    double *wb = nullptr;     // NS: partial adjoint of the entropy variables (state size)
    double *sb = nullptr;     // NS: partial adjoint of the state (state size)
};

// doubles of each array (gf, wb, sb only for Navier-Stokes)
struct VjpSizes {
    size_t q, lamw, gf, wb, sb;
};

VjpSizes vjp_sizes(const prandtl_op_s *op) {
    const size_t nodes = (size_t)op->NE * op->nn;
    VjpSizes z{(size_t)(op->dim + 4) * nodes, op->size, 0, 0, 0};
    if (op->phys.viscous) {
        z.gf = (size_t)op->dim * 2 * op->neq * op->NE * (op->nn / op->np);
        z.wb = z.sb = op->size;
    }
    return z;
}

// the handle's own work arrays, allocated on first use (the allocating entry points)
int vjp_handle_buffers(prandtl_op_s *op, VjpBuffers &B) {
    const VjpSizes z = vjp_sizes(op);
    if (!op->d_vq) HIP_TRY(hipMalloc(&op->d_vq, z.q * sizeof(double)));
    if (!op->d_vlam) HIP_TRY(hipMalloc(&op->d_vlam, z.lamw * sizeof(double)));
    if (op->phys.viscous && !op->d_vgf) {
        HIP_TRY(hipMalloc(&op->d_vgf, z.gf * sizeof(double)));
        HIP_TRY(hipMalloc(&op->d_vwb, z.wb * sizeof(double)));
        HIP_TRY(hipMalloc(&op->d_vsb, z.sb * sizeof(double)));
    }
    B.q = op->d_vq;
    B.lamw = op->d_vlam;
    B.gf = op->d_vgf;
    B.wb = op->d_vwb;
    B.sb = op->d_vsb;
    return PRANDTL_OK;
}

template <int DIM, int NP>
int launch_vjp(prandtl_op_s *op, const VjpSpec &S, hipStream_t st, const VjpBuffers &B) {
    const size_t nodes = (size_t)op->NE * op->nn;
    VjpArgs<NP> a;
    a.T = make_tables<NP>(op);
    a.P = op->P;
    a.NE = op->NE;
    a.bug3 = (op->opts.bug_compat & PRANDTL_BUG_A3) ? 1 : 0;
    a.mode = S.mode;
    a.nbr = op->d_nbr;
    a.geom = op->d_geom;
    a.u = S.u;
    a.lam = S.lam;
    a.x = S.x;
    a.ub = S.ub;
    a.kb = S.kb;
    a.q = B.q;
    a.lamw = B.lamw;
    a.out = S.out;
    a.kb_out = S.kb_out;
    a.ca = S.ca;
    a.cb = S.cb;
    a.dt = S.dt;
    const int pblk = (int)std::min<size_t>((nodes + 255) / 256, 8192);
    hipLaunchKernelGGL((k_vjp_prep<DIM, NP>), dim3(pblk), dim3(256), 0, st, a);
    VjpViscArgs<NP> va;
    if (op->phys.viscous) {
        // the viscous part: recompute the lifting and reverse it (k_vjp_visc1) before the
        // inviscid adjoint writes the output, and complete it afterwards (k_vjp_visc2)
        va.T = a.T;
        va.P = op->P;
        va.NE = op->NE;
        va.mode = S.mode;
        va.dt = S.dt;
        va.cb = S.cb;
        va.nbr = op->d_nbr;
        va.geom = op->d_geom;
        va.u = S.u;
        va.lam = S.mode == VJP_LSRK ? B.lamw : S.lam;
        va.Gf = B.gf;
        va.WB = B.wb;
        va.SB = B.sb;
        va.out = S.out;
        using VC = LiftCfg<DIM, NP>;
        hipLaunchKernelGGL((k_vjp_visc1<DIM, NP>), dim3((op->NE + VC::EPB - 1) / VC::EPB), dim3(VC::BLOCK), 0, st, va);
    }
    using LC = LineCfg<DIM, NP>;
    const int nblk = (op->NE + LC::EPB - 1) / LC::EPB;
    if (op->opts.surface_flux == PRANDTL_SURF_CHANDRASHEKAR)
        hipLaunchKernelGGL((k_vjp<DIM, NP, 0>), dim3(nblk), dim3(LC::BLOCK), 0, st, a);
    else
        hipLaunchKernelGGL((k_vjp<DIM, NP, 1>), dim3(nblk), dim3(LC::BLOCK), 0, st, a);
    if (op->phys.viscous)
        hipLaunchKernelGGL((k_vjp_visc2<DIM, NP>), dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, st, va);
    HIP_TRY(hipGetLastError());
    return PRANDTL_OK;
}

// the adjoint's requirements on the operator
int vjp_supported(const prandtl_op_s *op) {
    if (op->opts.blending) return fail(PRANDTL_ERR_UNSUPPORTED, "the adjoint does not differentiate the blending");
    if (op->nbc || op->d_ngeom) return fail(PRANDTL_ERR_UNSUPPORTED, "the adjoint covers periodic affine meshes");
    if (op->phys.viscous && op->phys.sutherland)
        return fail(PRANDTL_ERR_UNSUPPORTED, "the Navier-Stokes adjoint covers constant viscosity");
    if (op->opts.surface_flux == PRANDTL_SURF_RUSANOV)
        return fail(PRANDTL_ERR_UNSUPPORTED, "the adjoint covers the Chandrashekar and EC surface fluxes");
    if (op->opts.precision != PRANDTL_FP64) return fail(PRANDTL_ERR_UNSUPPORTED, "the adjoint is FP64");
    if (op->phys.viscous && (op->opts.bug_compat & (PRANDTL_BUG_A1 | PRANDTL_BUG_A2)))
        return fail(PRANDTL_ERR_UNSUPPORTED, "the adjoint differentiates the intended operator, not bug_compat A1/A2");
    return PRANDTL_OK;
}

int dispatch_vjp(prandtl_op_s *op, const VjpSpec &S, hipStream_t st, const VjpBuffers &B) {
    if (op->dim == 2) {
        switch (op->np) {
            case 2: return launch_vjp<2, 2>(op, S, st, B);
            case 3: return launch_vjp<2, 3>(op, S, st, B);
            case 4: return launch_vjp<2, 4>(op, S, st, B);
            case 5: return launch_vjp<2, 5>(op, S, st, B);
        }
    } else {
        switch (op->np) {
            case 2: return launch_vjp<3, 2>(op, S, st, B);
            case 3: return launch_vjp<3, 3>(op, S, st, B);
            case 4: return launch_vjp<3, 4>(op, S, st, B);
            case 5: return launch_vjp<3, 5>(op, S, st, B);
        }
    }
    return fail(PRANDTL_ERR_UNSUPPORTED, "order out of range");
}

int run_vjp(prandtl_op_s *op, const VjpSpec &S, hipStream_t st) {
    RoctxRange range("prandtl.vjp");
    int rc = vjp_supported(op);
    if (rc) return rc;
    VjpBuffers B;
    rc = vjp_handle_buffers(op, B);
    if (rc) return rc;
    return dispatch_vjp(op, S, st, B);
}

// ---------------------------------------------------------------------------
// Allocation-free operator adjoint (prandtl_op_vjp_noalloc): the work arrays live in one
// file-scope device array of fixed size instead of handle memory allocated on first use.  Each
// device holds its own copy at its own address, looked up once per device with
// hipGetSymbolAddress; host code uses only that address (the array's name is a host-side
// stand-in).  Every array is fully written before it is read, so nothing is cleared.
// ---------------------------------------------------------------------------
constexpr size_t kVjpScratchBytes = (size_t)64 << 20;   // 64 MiB per device
static __device__ __align__(256) double scratch[kVjpScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void **out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PRANDTL_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PRANDTL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PRANDTL_OK;
}

// the work arrays carved from the device's scratch array, each 256-byte aligned
int vjp_scratch_buffers(const prandtl_op_s *op, VjpBuffers &B) {
    const VjpSizes z = vjp_sizes(op);
    auto up = [](size_t n) { return (n * sizeof(double) + 255) & ~(size_t)255; };
    const size_t need = up(z.q) + up(z.gf) + up(z.wb) + up(z.sb);   // VJP_PLAIN reads no lamw
    if (need > kVjpScratchBytes)
        return fail(PRANDTL_ERR_ARG, "prandtl_op_vjp_noalloc: this operator needs " + std::to_string(need) +
                                         " bytes of work space, above its fixed " +
                                         std::to_string(kVjpScratchBytes) + " (use prandtl_op_vjp)");
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    char *p = static_cast<char *>(base);
    B.q = reinterpret_cast<double *>(p);
    p += up(z.q);
    if (op->phys.viscous) {
        B.gf = reinterpret_cast<double *>(p);
        p += up(z.gf);
        B.wb = reinterpret_cast<double *>(p);
        p += up(z.wb);
        B.sb = reinterpret_cast<double *>(p);
    }
    B.lamw = nullptr;
    return PRANDTL_OK;
}

// ---------------------------------------------------------------------------
// stage hooks: entropy filter and positivity limiter (dg_hooks.hpp)
// ---------------------------------------------------------------------------
// P[j] = V1[:, :j+1] V1^-1[:j+1, :], V1[i][k] = P_k(2 x_i - 1) on the GLL nodes x_i of [0,1]
template <int NP>
HookTables<NP> make_hook_tables() {
    HookTables<NP> T;
    double x[NP], w[NP], V[NP][NP], A[NP][NP], Vi[NP][NP];
    gll_rule(NP, x, w);
    for (int i = 0; i < NP; i++) {
        const double z = 2.0 * x[i] - 1.0;
        V[i][0] = 1.0;
        if (NP > 1) V[i][1] = z;
        for (int n = 1; n + 1 < NP; n++) V[i][n + 1] = ((2.0 * n + 1.0) * z * V[i][n] - n * V[i][n - 1]) / (n + 1.0);
        T.w[i] = w[i];
    }
    // Gauss-Jordan inverse with partial pivoting
    for (int i = 0; i < NP; i++)
        for (int k = 0; k < NP; k++) A[i][k] = V[i][k], Vi[i][k] = (i == k) ? 1.0 : 0.0;
    for (int c = 0; c < NP; c++) {
        int piv = c;
        for (int r = c + 1; r < NP; r++)
            if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        for (int k = 0; k < NP; k++) std::swap(A[c][k], A[piv][k]), std::swap(Vi[c][k], Vi[piv][k]);
        const double d = 1.0 / A[c][c];
        for (int k = 0; k < NP; k++) A[c][k] *= d, Vi[c][k] *= d;
        for (int r = 0; r < NP; r++) {
            if (r == c) continue;
            const double f = A[r][c];
            for (int k = 0; k < NP; k++) A[r][k] -= f * A[c][k], Vi[r][k] -= f * Vi[c][k];
        }
    }
    for (int j = 0; j < NP; j++)
        for (int i = 0; i < NP; i++)
            for (int a = 0; a < NP; a++) {
                double t = 0.0;
                for (int k = 0; k <= j; k++) t += V[i][k] * Vi[k][a];
                T.P[j][i][a] = t;
            }
    return T;
}

int ensure_hook_buffers(prandtl_op_s *op) {
    if (!op->d_emin) {
        HIP_TRY(hipMalloc(&op->d_emin, (size_t)op->NE * sizeof(double)));
        HIP_TRY(hipMalloc(&op->d_fbound, (size_t)op->NE * sizeof(double)));
        HIP_TRY(hipMalloc(&op->d_hflag, (size_t)op->NE * sizeof(int)));
        HIP_TRY(hipMalloc(&op->d_hcount, 3 * sizeof(unsigned long long)));
        HIP_TRY(hipMalloc(&op->d_hlist, (size_t)op->NE * sizeof(int)));
        HIP_TRY(hipMalloc(&op->d_hlist2, (size_t)op->NE * sizeof(int)));
        HIP_TRY(hipMalloc(&op->d_hlistn, 4 * sizeof(unsigned int)));
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, 0, op->d_hcount, 0ull, 3);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, 0, reinterpret_cast<int *>(op->d_hlistn), 0, 4);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
    }
    return PRANDTL_OK;
}

enum HookOp {
    HOOK_CONSTRAINTS = 0, HOOK_FILTER = 1, HOOK_LIMITER = 2,
    // the stage-fused hooks (detection in the stage epilogue): the filter's bounds from the tracked
    // minima; the filter (with the limiter inline when both hooks are on), at a step's last stage
    // also rewriting the minima of the elements it changes; the limiter alone
    HOOK_BOUND = 3, HOOK_FUSED_FILTER = 4, HOOK_FUSED_FILTER_LAST = 5, HOOK_FUSED_LIMITER = 6
};

template <int DIM, int NP>
int launch_hook(prandtl_op_s *op, int what, double *u, hipStream_t st) {
    HookArgs<NP> a;
    a.T = make_hook_tables<NP>();
    a.NE = op->NE;
    a.gamma = op->phys.gamma;
    a.gammaM1 = op->phys.gamma - 1.0;
    a.nbr = op->d_nbr;
    a.u = u;
    a.emin = op->d_emin;
    a.bound = op->d_fbound;
    a.flag = op->d_hflag;
    a.count = op->d_hcount;
    a.ngeom = op->d_ngeom;
    a.list = nullptr;
    a.nlist = a.nlist_next = nullptr;
    a.inline_limit = 0;
    a.track_emin = 0;
    using C = HookCfg<DIM, NP>;
    const int sblk = (op->NE + C::EPB - 1) / C::EPB;
    constexpr int ebl = C::FBLOCK;   // one element per workgroup at a time
    // the filter and limiter walk the scan's list of elements with a fixed grid (the limiter's
    // list is usually empty: a smaller grid costs less to launch)
    const int lblk = std::min(op->NE, 8192), lblk_lim = std::min(op->NE, 512);
    const int h = what == HOOK_FILTER ? 0 : 1, par = op->hook_par[h];
    if (what != HOOK_CONSTRAINTS) {
        // the scan appends to one of the hook's two counters and clears the other, which the
        // previous application of the hook (earlier on this stream: the hooks of one operator are
        // stream-ordered) has finished reading
        a.list = op->d_hlist;
        a.nlist = op->d_hlistn + 2 * h + par;
        a.nlist_next = op->d_hlistn + 2 * h + (1 - par);
    }
    if (what == HOOK_CONSTRAINTS || what == HOOK_BOUND) {
        // the bounds from the elements' minimum entropies: scanned from u, or (HOOK_BOUND) as the
        // fused hooks of the previous step's last stage left them
        if (what == HOOK_CONSTRAINTS) hipLaunchKernelGGL((k_hook_scan<DIM, NP, 0>), dim3(sblk), dim3(256), 0, st, a);
        hipLaunchKernelGGL((k_filter_bound<DIM, NP>), dim3((op->NE + 255) / 256), dim3(256), 0, st, a);
    } else if (what == HOOK_FUSED_FILTER || what == HOOK_FUSED_LIMITER || what == HOOK_FUSED_FILTER_LAST) {
        // the lists the stage epilogue filled (stage_hook_detect): no scan
        const bool filt = what != HOOK_FUSED_LIMITER;
        a.list = filt ? op->d_hlist : op->d_hlist2;
        a.nlist = op->d_hlistn + (filt ? 0 : 2) + op->hook_par[filt ? 0 : 1];
        a.inline_limit = filt && (op->opts.stage_hooks & PRANDTL_HOOK_LIMITER) ? 1 : 0;
        a.track_emin = what == HOOK_FUSED_FILTER_LAST ? 1 : 0;
        if (filt)
            hipLaunchKernelGGL((k_filter<DIM, NP>), dim3(lblk), dim3(ebl), 0, st, a);
        else
            hipLaunchKernelGGL((k_limiter<DIM, NP>), dim3(lblk_lim), dim3(ebl), 0, st, a);
        HIP_TRY(hipGetLastError());
        // the hook whose counters the epilogue used switches them (the other's are untouched)
        op->hook_par[filt ? 0 : 1] = 1 - op->hook_par[filt ? 0 : 1];
        return PRANDTL_OK;
    } else if (what == HOOK_FILTER) {
        hipLaunchKernelGGL((k_hook_scan<DIM, NP, 1>), dim3(sblk), dim3(256), 0, st, a);
        hipLaunchKernelGGL((k_filter<DIM, NP>), dim3(lblk), dim3(ebl), 0, st, a);
    } else {
        hipLaunchKernelGGL((k_hook_scan<DIM, NP, 2>), dim3(sblk), dim3(256), 0, st, a);
        hipLaunchKernelGGL((k_limiter<DIM, NP>), dim3(lblk_lim), dim3(ebl), 0, st, a);
    }
    HIP_TRY(hipGetLastError());
    // the counters switch roles only once this application is queued
    if (what == HOOK_FILTER || what == HOOK_LIMITER) op->hook_par[h] = 1 - par;
    return PRANDTL_OK;
}

int run_hook(prandtl_op_s *op, int what, double *u, hipStream_t st) {
    int rc = ensure_hook_buffers(op);
    if (rc) return rc;
    if (what == HOOK_FILTER && !op->have_bounds)
        return fail(PRANDTL_ERR_ARG, "no filter constraints: call prandtl_filter_constraints first");
    if (op->dim == 2) {
        switch (op->np) {
            case 2: return launch_hook<2, 2>(op, what, u, st);
            case 3: return launch_hook<2, 3>(op, what, u, st);
            case 4: return launch_hook<2, 4>(op, what, u, st);
            case 5: return launch_hook<2, 5>(op, what, u, st);
        }
    } else {
        switch (op->np) {
            case 2: return launch_hook<3, 2>(op, what, u, st);
            case 3: return launch_hook<3, 3>(op, what, u, st);
            case 4: return launch_hook<3, 4>(op, what, u, st);
            case 5: return launch_hook<3, 5>(op, what, u, st);
        }
    }
    return fail(PRANDTL_ERR_UNSUPPORTED, "order out of range");
}

// the hooks the stepper applies after a stage: filter, then limiter
int stage_hooks(prandtl_op_s *op, double *u, hipStream_t st) {
    RoctxRange range("prandtl.hooks");
    int rc = PRANDTL_OK;
    if (op->opts.stage_hooks & PRANDTL_HOOK_FILTER) rc = run_hook(op, HOOK_FILTER, u, st);
    if (!rc && (op->opts.stage_hooks & PRANDTL_HOOK_LIMITER)) rc = run_hook(op, HOOK_LIMITER, u, st);
    return rc;
}

// Carpenter & Kennedy (1994) 2N-storage RK4(3)5 coefficients.
const double CK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double CK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};
// the stage times (as fractions of dt) of the same scheme
const double CK_C[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                        2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

// Advance u by nsteps; the state ping-pongs between u and the handle's alternate buffer
// (a stage reads its neighbours' face nodes, so it cannot update in place).  *t: the time at the
// start, advanced by dt after every step (MFEM's ODESolver::Step); each stage runs at its stage time.
int rk_advance(prandtl_op_s *op, int scheme, double *u, double *t, double dt, int nsteps, int nan_check,
               hipStream_t st) {
    RoctxRange range("prandtl.rk_run");
    int rc = ensure_rk_buffers(op, scheme);
    if (rc) return rc;
    if (nan_check > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, op->d_nan, 0, 1);
        HIP_TRY(hipGetLastError());
    }
    double *cur = u;
    const bool hooks = (op->opts.stage_hooks & (PRANDTL_HOOK_FILTER | PRANDTL_HOOK_LIMITER)) != 0;
    // the fused steppers detect the hooks in the stage epilogue when the flux kernel holds each
    // element in one wave (all but 3D p = 4), unless the caller asked for separate scans
    const bool fused = hooks && (scheme == PRANDTL_LSRK45 || scheme == PRANDTL_SSPRK3) &&
                       !(op->dim == 3 && op->np == 5) && !(op->opts.stage_hooks & PRANDTL_HOOK_UNFUSED);
    const bool hf = (op->opts.stage_hooks & PRANDTL_HOOK_FILTER) != 0;
    const int hmode = !fused ? 0 : (hf ? HOOK_DETECT_FILTER : HOOK_DETECT_LIMITER);
    if (fused) {
        rc = ensure_hook_buffers(op);
        if (rc) return rc;
    }
    // after a fused stage: the filter (and the limiter inline) or the limiter on the detected lists
    auto fused_hooks = [&](double *x, bool last) -> int {
        return run_hook(op, hf ? (last ? HOOK_FUSED_FILTER_LAST : HOOK_FUSED_FILTER) : HOOK_FUSED_LIMITER, x, st);
    };
    for (int s = 0; s < nsteps; s++) {
        const bool chk = nan_check > 0 && ((s + 1) % nan_check == 0 || s + 1 == nsteps);
        if (hf) {
            // GetFilterConstraints on the state at the start of the step: scanned, or (fused, after
            // the first step) the minima the previous step's last stage left
            rc = run_hook(op, (fused && s > 0) ? HOOK_BOUND : HOOK_CONSTRAINTS, cur, st);
            if (rc) return rc;
            op->have_bounds = true;
        }
        if (scheme == PRANDTL_FORWARD_EULER || scheme == PRANDTL_RK2 || scheme == PRANDTL_RK4) {
            // MFEM's explicit RK steppers: ForwardEulerSolver (x += dt f(x)); RK2Solver with a = 2/3
            // (x1 = x + (1-b) dt k1, y = x + a dt k1, x = x1 + b dt f(y), b = 1/(2a)); RK4Solver, the
            // classic tableau a = {1/2; 0, 1/2; 0, 0, 1}, b = {1/6, 1/3, 1/3, 1/6} of ExplicitRKSolver
            // (y_l = x + a_ll dt k_l, the zero entries contributing nothing; x += sum_l b_l dt k_l).
            // The stage inputs y ping-pong between d_ub and the caller's buffer; acc in d_y2.
            static const double rk2a = 2.0 / 3.0, rk2b = 0.5 / rk2a;
            int ns;
            double ca[4], cb[4];
            if (scheme == PRANDTL_FORWARD_EULER) {
                ns = 1;
                cb[0] = 1.0;
            } else if (scheme == PRANDTL_RK2) {
                ns = 2;
                ca[0] = rk2a;
                cb[0] = 1.0 - rk2b;
                cb[1] = rk2b;
            } else {
                ns = 4;
                ca[0] = 0.5, ca[1] = 0.5, ca[2] = 1.0;
                cb[0] = 1.0 / 6.0, cb[1] = 1.0 / 3.0, cb[2] = 1.0 / 3.0, cb[3] = 1.0 / 6.0;
            }
            double *x = cur, *yin = cur;
            const unsigned gblk = (unsigned)std::min<size_t>((op->size + 255) / 256, 8192);
            // MFEM's stage times: RK2 t, t + a dt; RK4 t + c dt, c = 0, 1/2, 1/2, 1
            const double ct[4] = {0.0, scheme == PRANDTL_RK2 ? rk2a : 0.5, 0.5, 1.0};
            for (int stg = 0; stg < ns; stg++) {
                StageSpec S{EPI_MULT, yin, op->d_k, nullptr, nullptr, 0.0, 0.0, dt, nullptr, false, false};
                S.time = *t + ct[stg] * dt;
                rc = run_stage(op, S, st);
                if (rc) return rc;
                double *ynext = (x == u) ? op->d_ub : u;
                const bool last = stg == ns - 1;
                hipLaunchKernelGGL(k_erk_update, dim3(gblk), dim3(256), 0, st, x, op->d_y2, ynext, op->d_k,
                                   last ? 0.0 : ca[stg], cb[stg], dt, stg == 0 ? 1 : 0, last ? 1 : 0, op->size);
                HIP_TRY(hipGetLastError());
                if (hooks) rc = stage_hooks(op, last ? x : ynext, st);
                if (rc) return rc;
                yin = ynext;
            }
            if (chk) {
                hipLaunchKernelGGL(k_nan_scan, dim3(gblk), dim3(256), 0, st, x, op->size, op->d_nan);
                HIP_TRY(hipGetLastError());
            }
        } else if (scheme == PRANDTL_LSRK45) {
            for (int stg = 0; stg < 5; stg++) {
                double *nxt = (cur == u) ? op->d_ub : u;
                StageSpec S{EPI_LSRK, cur, nxt, op->d_k, nullptr, CK_A[stg], CK_B[stg], dt, nullptr, false,
                            chk && stg == 4};
                S.hook_mode = hmode | (hf && stg == 4 ? HOOK_DETECT_EMIN : 0);
                S.time = *t + CK_C[stg] * dt;
                rc = run_stage(op, S, st);
                if (!rc && hooks) rc = fused ? fused_hooks(nxt, stg == 4) : stage_hooks(op, nxt, st);
                if (rc) return rc;
                cur = nxt;
            }
        } else {
            // MFEM RK3SSPSolver (Shu-Osher): y = x + dt L(x); y2 = 3/4 x + 1/4 (y + dt L(y));
            // x = 1/3 x + 2/3 (y2 + dt L(y2)).  x stays in `cur`.
            double *y = (cur == u) ? op->d_ub : u;
            // stage times t, t + dt, t + dt/2 (RK3SSPSolver::Step)
            StageSpec S1{EPI_SSP, cur, y, nullptr, cur, 0.0, 1.0, dt, nullptr, false, false};
            S1.hook_mode = hmode;
            S1.time = *t;
            rc = run_stage(op, S1, st);
            if (!rc && hooks) rc = fused ? fused_hooks(y, false) : stage_hooks(op, y, st);
            if (rc) return rc;
            StageSpec S2{EPI_SSP, y, op->d_y2, nullptr, cur, 3.0 / 4.0, 1.0 / 4.0, dt, nullptr, false, false};
            S2.hook_mode = hmode;
            S2.time = *t + dt;
            rc = run_stage(op, S2, st);
            if (!rc && hooks) rc = fused ? fused_hooks(op->d_y2, false) : stage_hooks(op, op->d_y2, st);
            if (rc) return rc;
            StageSpec S3{EPI_SSP, op->d_y2, cur, nullptr, cur, 1.0 / 3.0, 2.0 / 3.0, dt, nullptr, false, chk};
            S3.hook_mode = hmode | (hf ? HOOK_DETECT_EMIN : 0);
            S3.time = *t + dt / 2;
            rc = run_stage(op, S3, st);
            if (!rc && hooks) rc = fused ? fused_hooks(cur, true) : stage_hooks(op, cur, st);
            if (rc) return rc;
        }
        *t += dt;
        if (chk) {
            int flag = 0;
            HIP_TRY(hipMemcpyAsync(&flag, op->d_nan, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (flag) {
                if (cur != u) {
                    hipLaunchKernelGGL(k_copy_f64, dim3(2048), dim3(256), 0, st, u, cur, op->size);
                    HIP_TRY(hipGetLastError());
                }
                return fail(PRANDTL_ERR_NAN, "non-finite state after step " + std::to_string(s + 1));
            }
        }
    }
    if (cur != u) {
        hipLaunchKernelGGL(k_copy_f64, dim3(2048), dim3(256), 0, st, u, cur, op->size);
        HIP_TRY(hipGetLastError());
    }
    return PRANDTL_OK;
}

// the neighbour's face point of face point t for orientation code `code` (host copy of
// face_point_map in dg_kernels.hpp)
int host_face_point(int dim, int np, int code, int t) {
    const int t1 = t % np, t2 = t / np;
    const bool sw = (code & 32) != 0;
    const int a = sw ? t2 : t1, b = sw ? t1 : t2;
    const int u1 = (code & 8) ? np - 1 - a : a;
    const int u2 = (code & 16) ? np - 1 - b : b;
    return dim == 2 ? u1 : u1 + np * u2;
}

// node of face point t of local face f = 2 d + s
int host_face_node(int np, int f, int t) {
    const int d = f >> 1, s = f & 1;
    const int st = d == 0 ? 1 : (d == 1 ? np : np * np);
    return (t % st) + (t / st) * st * np + (s ? (np - 1) * st : 0);
}

// A general mesh (per-node geometry and/or face orientation codes), with the codes of an
// aligned lattice where face_orientation is null and the element geometry at every node where
// node_geometry is null: positive det J at every node, reciprocal face matching whose point maps
// invert each other, and area vectors that agree across every interior face point.
int validate_general_mesh(const prandtl_mesh_desc *m, const std::vector<double> &ng, const std::vector<int> &fo) {
    const int dim = m->dim, NE = m->num_elements, nf = 2 * dim, np = m->order + 1;
    int nn = 1;
    for (int d = 0; d < dim; d++) nn *= np;
    const int G = 1 + dim * dim, L = nn / np;
    for (size_t i = 0; i < (size_t)NE * nn; i++)
        if (!(ng[i * G] > 0.0) || !std::isfinite(ng[i * G]))
            return fail(PRANDTL_ERR_MESH, "element " + std::to_string(i / nn) + ": det J must be positive at every node");
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < nf; f++) {
            const int en = m->neighbors[(size_t)e * nf + f];
            const std::string where = "element " + std::to_string(e) + " face " + std::to_string(f);
            if (en >= NE || (en < 0 && -en - 1 >= m->num_bc))
                return fail(PRANDTL_ERR_MESH, where + ": a neighbour entry is neither an element nor a boundary condition");
            if (en < 0) continue;
            const int code = fo[(size_t)e * nf + f], f2 = code & 7;
            if ((code & ~63) || f2 >= nf || (dim == 2 && (code & 48)))
                return fail(PRANDTL_ERR_MESH, where + ": bad face orientation code");
            if (m->neighbors[(size_t)en * nf + f2] != e || (fo[(size_t)en * nf + f2] & 7) != f)
                return fail(PRANDTL_ERR_MESH, where + ": neighbour table is not reciprocal");
            const int back = fo[(size_t)en * nf + f2];
            for (int t = 0; t < L; t++) {
                const int tn = host_face_point(dim, np, code, t);
                if (host_face_point(dim, np, back, tn) != t)
                    return fail(PRANDTL_ERR_MESH, where + ": face orientation codes of the two sides do not invert");
                // outward area vectors of the two sides at the shared point: equal and opposite
                const double *a = &ng[((size_t)e * nn + host_face_node(np, f, t)) * G + 1 + (f >> 1) * dim];
                const double *b = &ng[((size_t)en * nn + host_face_node(np, f2, tn)) * G + 1 + (f2 >> 1) * dim];
                const double sa = (f & 1) ? 1.0 : -1.0, sb = (f2 & 1) ? 1.0 : -1.0;
                double nrm = 0.0, diff = 0.0;
                for (int c = 0; c < dim; c++) {
                    nrm += a[c] * a[c];
                    diff += (sa * a[c] + sb * b[c]) * (sa * a[c] + sb * b[c]);
                }
                if (!(diff <= 1e-20 * nrm))
                    return fail(PRANDTL_ERR_MESH, where + ": face area vectors of neighbours disagree (non-conforming face)");
            }
        }
    return PRANDTL_OK;
}

int validate_mesh(const prandtl_mesh_desc *m) {
    const int dim = m->dim, NE = m->num_elements, nf = 2 * dim, ng = 1 + dim * dim;
    for (int e = 0; e < NE; e++) {
        const double J = m->geometry[(size_t)e * ng];
        if (!(J > 0.0) || !std::isfinite(J))
            return fail(PRANDTL_ERR_MESH, "element " + std::to_string(e) + ": det J must be positive");
        for (int d = 0; d < dim; d++) {
            const int ep = m->neighbors[(size_t)e * nf + 2 * d + 1];
            const int em = m->neighbors[(size_t)e * nf + 2 * d];
            for (int nb : {ep, em}) {
                if (nb >= NE || (nb < 0 && -nb - 1 >= m->num_bc))
                    return fail(PRANDTL_ERR_MESH, "element " + std::to_string(e) +
                                                      ": a neighbour entry is neither an element nor a boundary condition");
            }
            if (em >= 0 && m->neighbors[(size_t)em * nf + 2 * d + 1] != e)
                return fail(PRANDTL_ERR_MESH, "element " + std::to_string(e) + ": neighbour table is not reciprocal");
            if (ep < 0) continue;   // a boundary face
            if (m->neighbors[(size_t)ep * nf + 2 * d] != e)
                return fail(PRANDTL_ERR_MESH, "element " + std::to_string(e) + ": neighbour table is not reciprocal");
            // the shared face's area vector must agree from both sides (conforming affine faces)
            double nrm = 0.0, diff = 0.0;
            for (int c = 0; c < dim; c++) {
                const double a = m->geometry[(size_t)e * ng + 1 + d * dim + c];
                const double b = m->geometry[(size_t)ep * ng + 1 + d * dim + c];
                nrm += a * a;
                diff += (a - b) * (a - b);
            }
            if (diff > 1e-24 * nrm)
                return fail(PRANDTL_ERR_MESH, "element " + std::to_string(e) +
                                                  ": face area vectors of neighbours disagree (non-conforming face)");
        }
    }
    return PRANDTL_OK;
}

}  // namespace

extern "C" {

const char *prandtl_last_error(void) { return g_last_error.c_str(); }

int prandtl_op_create(const prandtl_mesh_desc *mesh, const prandtl_physics *phys, const prandtl_opts *opts,
                      prandtl_op *out) {
    if (!mesh || !phys || !out) return fail(PRANDTL_ERR_ARG, "null argument");
    *out = nullptr;
    if (mesh->dim < 2 || mesh->dim > 3) return fail(PRANDTL_ERR_UNSUPPORTED, "dim must be 2 or 3");
    if (mesh->order < 1 || mesh->order + 1 > MAX_NP) return fail(PRANDTL_ERR_UNSUPPORTED, "order must be 1..4");
    if (mesh->num_elements <= 0 || !mesh->neighbors || (!mesh->geometry && !mesh->node_geometry))
        return fail(PRANDTL_ERR_ARG, "empty mesh");
    const bool general = mesh->node_geometry || mesh->face_orientation;
    if (!(phys->gamma > 1.0)) return fail(PRANDTL_ERR_ARG, "gamma must exceed 1");
    prandtl_opts o{};
    if (opts) o = *opts;
    if (o.surface_flux < 0 || o.surface_flux > 2) return fail(PRANDTL_ERR_ARG, "unknown surface flux");
    if (o.precision != PRANDTL_FP64 && o.precision != PRANDTL_MIXED) return fail(PRANDTL_ERR_ARG, "unknown precision");
    if (o.stage_hooks & ~(PRANDTL_HOOK_FILTER | PRANDTL_HOOK_LIMITER | PRANDTL_HOOK_UNFUSED))
        return fail(PRANDTL_ERR_ARG, "unknown stage hook");
    if (o.blending != 0 && o.blending != 1) return fail(PRANDTL_ERR_ARG, "blending must be 0 or 1");
    if (!(o.alpha_min >= 0.0 && o.alpha_min < 0.5) || !(o.alpha_max >= 0.0 && o.alpha_max <= 1.0))
        return fail(PRANDTL_ERR_ARG, "alpha_min must lie in [0, 0.5) and alpha_max in [0, 1]");
    if (o.blending && o.precision != PRANDTL_FP64)
        return fail(PRANDTL_ERR_UNSUPPORTED, "subcell blending runs in FP64");
    if (o.bug_compat & ~(PRANDTL_BUG_A1 | PRANDTL_BUG_A2 | PRANDTL_BUG_A3))
        return fail(PRANDTL_ERR_ARG, "unknown bug_compat bit");
    if (phys->viscous && (o.bug_compat & (PRANDTL_BUG_A1 | PRANDTL_BUG_A2)) &&
        (general || mesh->num_bc > 0 || o.precision != PRANDTL_FP64))
        return fail(PRANDTL_ERR_UNSUPPORTED, "bug_compat A1/A2 run on periodic affine meshes in FP64");
    if (phys->viscous && (!(phys->Pr > 0.0) || phys->mu < 0.0)) return fail(PRANDTL_ERR_ARG, "bad Pr or mu");
    if (mesh->num_bc < 0 || (mesh->num_bc > 0 && !mesh->bc)) return fail(PRANDTL_ERR_ARG, "bad boundary-condition table");
    for (int k = 0; k < mesh->num_bc; k++) {
        const int ty = mesh->bc[k].type;
        if (ty < PRANDTL_BC_SLIP_WALL || ty > PRANDTL_BC_NOSLIP_ISOTHERMAL)
            return fail(PRANDTL_ERR_ARG, "unknown boundary-condition type " + std::to_string(ty));
        if (phys->viscous && (ty == PRANDTL_BC_SLIP_WALL || ty == PRANDTL_BC_SYMMETRY))
            return fail(PRANDTL_ERR_UNSUPPORTED, "slip and symmetry walls define no viscous or lifting flux "
                                                 "(the reference's base class aborts for them)");
        const prandtl_bc &b = mesh->bc[k];
        if (b.profile < PRANDTL_PROFILE_CONSTANT || b.profile > PRANDTL_PROFILE_RAMP)
            return fail(PRANDTL_ERR_ARG, "boundary condition " + std::to_string(k) + ": unknown profile " +
                                             std::to_string(b.profile));
        if (b.profile != PRANDTL_PROFILE_CONSTANT && ty != PRANDTL_BC_SUPERSONIC_INFLOW &&
            ty != PRANDTL_BC_SPECIFIED_STATE)
            return fail(PRANDTL_ERR_ARG, "boundary condition " + std::to_string(k) +
                                             ": profiles apply to supersonic-inflow and specified-state conditions");
        if (b.profile == PRANDTL_PROFILE_FRONT) {
            const double i0 = b.a[0], i1 = b.a[1];
            if (!(i0 == 0.0 || i0 == 1.0 || (mesh->dim == 3 && i0 == 2.0)) ||
                !(i1 == 0.0 || i1 == 1.0 || (mesh->dim == 3 && i1 == 2.0)) || !(b.a[4] != 0.0))
                return fail(PRANDTL_ERR_ARG, "boundary condition " + std::to_string(k) +
                                                 ": a front needs coordinate indices below dim and c2 != 0");
            if (!mesh->node_coords)
                return fail(PRANDTL_ERR_ARG, "boundary condition " + std::to_string(k) +
                                                 ": a front needs the node coordinates (node_coords)");
        }
        if (b.profile == PRANDTL_PROFILE_RAMP && !(b.a[0] > 0.0 && b.a[1] > 0.0))
            return fail(PRANDTL_ERR_ARG, "boundary condition " + std::to_string(k) +
                                             ": a ramp needs t_r > 0 and a positive density");
    }
    if (mesh->num_bc > 0 && o.precision != PRANDTL_FP64)
        return fail(PRANDTL_ERR_UNSUPPORTED, "boundary faces run in FP64");
    if (general && o.precision != PRANDTL_FP64)
        return fail(PRANDTL_ERR_UNSUPPORTED, "general meshes run in FP64");
    // general mesh: per-node geometry and orientation codes, filled in from the element geometry
    // and the aligned-lattice convention where the caller gives none
    std::vector<double> hng;
    std::vector<int> hfo;
    int rc;
    if (general) {
        const int dim = mesh->dim, np = mesh->order + 1, G = 1 + dim * dim, nf = 2 * dim;
        int nn = 1;
        for (int d = 0; d < dim; d++) nn *= np;
        const size_t NE = (size_t)mesh->num_elements;
        if (mesh->node_geometry) {
            hng.assign(mesh->node_geometry, mesh->node_geometry + NE * nn * G);
        } else {
            hng.resize(NE * nn * G);
            for (size_t e = 0; e < NE; e++)
                for (int i = 0; i < nn; i++)
                    std::copy(mesh->geometry + e * G, mesh->geometry + (e + 1) * G, hng.begin() + (e * nn + i) * G);
        }
        if (mesh->face_orientation) {
            hfo.assign(mesh->face_orientation, mesh->face_orientation + NE * nf);
        } else {
            hfo.resize(NE * nf);
            for (size_t e = 0; e < NE; e++)
                for (int f = 0; f < nf; f++) hfo[e * nf + f] = f ^ 1;
        }
        rc = validate_general_mesh(mesh, hng, hfo);
    } else {
        rc = validate_mesh(mesh);
    }
    if (rc) return rc;

    prandtl_op_s *op = new prandtl_op_s();
    op->dim = mesh->dim;
    op->order = mesh->order;
    op->np = mesh->order + 1;
    op->nn = 1;
    for (int d = 0; d < op->dim; d++) op->nn *= op->np;
    op->NE = mesh->num_elements;
    op->neq = op->dim + 2;
    op->size = (size_t)op->neq * op->NE * op->nn;
    op->phys = *phys;
    op->opts = o;
    op->P.gamma = phys->gamma;
    op->P.gammaM1 = phys->gamma - 1.0;
    op->P.gammaM1Inverse = 1.0 / (phys->gamma - 1.0);
    op->P.mu = phys->mu;
    op->P.PrInverse = phys->viscous ? 1.0 / phys->Pr : 0.0;
    op->P.R_gas = phys->R_gas;
    op->P.mu_bulk = 2.0 / 3.0;   // Physics.hpp:34
    op->P.mu0 = 1.716e-5;        // Physics.hpp:35
    op->P.T0 = 273.15;           // Physics.hpp:36
    op->P.Ts = 110.4;            // Physics.hpp:37
    op->P.sutherland = phys->sutherland;
    double x[MAX_NP], w[MAX_NP];
    gll_rule(op->np, x, w);
    diff_matrices(op->np, x, w, op->D, op->Dh, op->Dh2);
    op->inv_w0 = 1.0 / w[0];
    for (int i = 0; i < op->np; i++) op->w[i] = w[i];

    auto cleanup = [&](int code) {
        prandtl_op_destroy(op);
        return code;
    };
    const size_t nb = (size_t)op->NE * 2 * op->dim * sizeof(int);
    const size_t G = 1 + op->dim * op->dim;
    const size_t gb = (size_t)op->NE * G * sizeof(double);
    // per-element geometry: the caller's, else node 0's of the per-node geometry (general meshes
    // read only the per-node arrays)
    std::vector<double> hgeom;
    if (!mesh->geometry) {
        hgeom.resize((size_t)op->NE * G);
        for (size_t el = 0; el < (size_t)op->NE; el++)
            std::copy(hng.begin() + el * op->nn * G, hng.begin() + el * op->nn * G + G, hgeom.begin() + el * G);
    }
    const double *geometry = mesh->geometry ? mesh->geometry : hgeom.data();
    hipError_t e;
    if ((e = hipMalloc(&op->d_nbr, nb)) != hipSuccess ||
        (e = hipMemcpy(op->d_nbr, mesh->neighbors, nb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMalloc(&op->d_geom, gb)) != hipSuccess ||
        (e = hipMemcpy(op->d_geom, geometry, gb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMalloc(&op->d_lam, 2 * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc(&op->d_lam_part, std::max<size_t>(2 * (size_t)op->NE + 64, 8192) * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&op->d_nan, sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&op->d_alpha, (size_t)op->NE * sizeof(double))) != hipSuccess)
        return cleanup(fail(PRANDTL_ERR_HIP, std::string("device allocation: ") + hipGetErrorString(e)));
    if (general) {
        if ((e = hipMalloc(&op->d_ngeom, hng.size() * sizeof(double))) != hipSuccess ||
            (e = hipMemcpy(op->d_ngeom, hng.data(), hng.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMalloc(&op->d_forient, hfo.size() * sizeof(int))) != hipSuccess ||
            (e = hipMemcpy(op->d_forient, hfo.data(), hfo.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(fail(PRANDTL_ERR_HIP, std::string("device allocation: ") + hipGetErrorString(e)));
    }
    if (mesh->num_bc > 0) {
        static_assert(sizeof(BcDev) == sizeof(prandtl_bc), "BcDev mirrors prandtl_bc");
        op->nbc = mesh->num_bc;
        if ((e = hipMalloc(&op->d_bc, (size_t)op->nbc * sizeof(BcDev))) != hipSuccess ||
            (e = hipMemcpy(op->d_bc, mesh->bc, (size_t)op->nbc * sizeof(BcDev), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(fail(PRANDTL_ERR_HIP, std::string("device allocation: ") + hipGetErrorString(e)));
        bool spatial = false;
        for (int k = 0; k < op->nbc; k++) spatial |= mesh->bc[k].profile == PRANDTL_PROFILE_FRONT;
        if (spatial) {
            // the coordinates of every element's face points, [e][2 dim][t][dim] (t the point's index on
            // the face, as the face traces): read by the boundary faces' function-valued states
            const int dim = op->dim, np = op->np, L = op->nn / np;
            std::vector<double> hx((size_t)op->NE * 2 * dim * L * dim);
            for (size_t el = 0; el < (size_t)op->NE; el++)
                for (int d = 0; d < dim; d++) {
                    const int st = d == 0 ? 1 : (d == 1 ? np : np * np);
                    for (int side = 0; side < 2; side++)
                        for (int t = 0; t < L; t++) {
                            const int node = (t % st) + (t / st) * st * np + (side ? (np - 1) * st : 0);
                            const double *src = mesh->node_coords + (el * op->nn + node) * dim;
                            double *dst = &hx[((el * 2 * dim + 2 * d + side) * L + t) * dim];
                            for (int c = 0; c < dim; c++) dst[c] = src[c];
                        }
                }
            if ((e = hipMalloc(&op->d_bxyz, hx.size() * sizeof(double))) != hipSuccess ||
                (e = hipMemcpy(op->d_bxyz, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
                return cleanup(fail(PRANDTL_ERR_HIP, std::string("device allocation: ") + hipGetErrorString(e)));
        }
    }
    if (phys->viscous) {
        const size_t nface = op->nn / op->np;  // points per face
        if ((e = hipMalloc(&op->d_dv, (size_t)(op->neq - 1) * op->NE * op->nn * sizeof(double))) != hipSuccess ||
            (e = hipMalloc(&op->d_Ff, (size_t)op->dim * 2 * (op->neq - 1) * op->NE * nface * sizeof(double))) !=
                hipSuccess)
            return cleanup(fail(PRANDTL_ERR_HIP, std::string("device allocation: ") + hipGetErrorString(e)));
    }
    *out = op;
    return PRANDTL_OK;
}

void prandtl_op_destroy(prandtl_op op) {
    if (!op) return;
    (void)hipFree(op->d_nbr);
    (void)hipFree(op->d_geom);
    (void)hipFree(op->d_ngeom);
    (void)hipFree(op->d_forient);
    (void)hipFree(op->d_dv);
    (void)hipFree(op->d_Ff);
    (void)hipFree(op->d_k);
    (void)hipFree(op->d_ub);
    (void)hipFree(op->d_y2);
    (void)hipFree(op->d_vq);
    (void)hipFree(op->d_vlam);
    (void)hipFree(op->d_vgf);
    (void)hipFree(op->d_vwb);
    (void)hipFree(op->d_vsb);
    (void)hipFree(op->d_emin);
    (void)hipFree(op->d_fbound);
    (void)hipFree(op->d_hflag);
    (void)hipFree(op->d_hcount);
    (void)hipFree(op->d_hlist);
    (void)hipFree(op->d_hlist2);
    (void)hipFree(op->d_hlistn);
    (void)hipFree(op->d_lam);
    (void)hipFree(op->d_lam_part);
    (void)hipFree(op->d_nan);
    (void)hipFree(op->d_alpha);
    (void)hipFree(op->d_bc);
    (void)hipFree(op->d_bxyz);
    for (hipEvent_t e : op->tev) (void)hipEventDestroy(e);
    for (double *p : op->tape_ckpt) (void)hipFree(p);
    release_adjoint_storage(op);
    delete op;
}

size_t prandtl_op_size(prandtl_op op) { return op ? op->size : 0; }

int prandtl_op_mult(prandtl_op op, const double *u, double *dudt, prandtl_stream stream) {
    if (!op || !u || !dudt) return fail(PRANDTL_ERR_ARG, "null argument");
    if (u == dudt) return fail(PRANDTL_ERR_ARG, "u and dudt must not alias");
    StageSpec S{EPI_MULT, u, dudt, nullptr, nullptr, 0.0, 0.0, 0.0, nullptr, true, false};
    S.time = op->time;
    int rc = run_stage(op, S, (hipStream_t)stream);
    op->lam_valid = (rc == PRANDTL_OK);
    op->lam_stream = (hipStream_t)stream;
    return rc;
}

int prandtl_op_mult_grad(prandtl_op op, const double *u, double *dudt, double *grad, prandtl_stream stream) {
    if (!op || !u || !dudt || !grad) return fail(PRANDTL_ERR_ARG, "null argument");
    if (!op->phys.viscous) return fail(PRANDTL_ERR_ARG, "gradients exist only for Navier-Stokes operators");
    StageSpec S{EPI_MULT, u, dudt, nullptr, nullptr, 0.0, 0.0, 0.0, grad, true, false};
    S.time = op->time;
    int rc = run_stage(op, S, (hipStream_t)stream);
    op->lam_valid = (rc == PRANDTL_OK);
    op->lam_stream = (hipStream_t)stream;
    return rc;
}

int prandtl_op_get_max_char_speed(prandtl_op op, double *lambda_host) {
    if (!op || !lambda_host) return fail(PRANDTL_ERR_ARG, "null argument");
    if (!op->lam_valid) return fail(PRANDTL_ERR_ARG, "no Mult has run on this operator");
    // ordered after the Mult on its own stream (a non-blocking caller stream is not ordered with
    // a plain copy)
    unsigned long long bits = 0;
    HIP_TRY(hipMemcpyAsync(&bits, op->d_lam, sizeof(bits), hipMemcpyDeviceToHost, op->lam_stream));
    HIP_TRY(hipStreamSynchronize(op->lam_stream));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *lambda_host = v;
    return PRANDTL_OK;
}

int prandtl_op_max_char_speed(prandtl_op op, const double *u, double *lambda_host, prandtl_stream stream) {
    if (!op || !u || !lambda_host) return fail(PRANDTL_ERR_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const size_t nnodes = (size_t)op->NE * op->nn;
    const int nblk = (int)std::min<size_t>(2048, (nnodes + 255) / 256);   // <= 8192 waves
    if (op->dim == 2)
        hipLaunchKernelGGL(k_max_speed<2>, dim3(nblk), dim3(256), 0, st, u, nnodes, op->P.gamma, op->d_lam_part);
    else
        hipLaunchKernelGGL(k_max_speed<3>, dim3(nblk), dim3(256), 0, st, u, nnodes, op->P.gamma, op->d_lam_part);
    // its own result word: the last Mult's speed (d_lam[0]) stays what get_max_char_speed returns
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, st, op->d_lam_part, nblk * 4, op->d_lam + 1);
    HIP_TRY(hipGetLastError());
    unsigned long long bits = 0;
    HIP_TRY(hipMemcpyAsync(&bits, op->d_lam + 1, sizeof(bits), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *lambda_host = v;
    return PRANDTL_OK;
}

int prandtl_op_profile(prandtl_op op, const double *u, int epilogue, int iters, double *ms, prandtl_stream stream) {
    if (!op || !u || !ms || iters <= 0) return fail(PRANDTL_ERR_ARG, "bad argument");
    if (epilogue != EPI_MULT && epilogue != EPI_LSRK) return fail(PRANDTL_ERR_ARG, "epilogue must be 0 or 1");
    hipStream_t st = (hipStream_t)stream;
    int rc = ensure_rk_buffers(op, PRANDTL_LSRK45);
    if (rc) return rc;
    StageSpec S{epilogue, u, op->d_ub, op->d_k, nullptr, 0.5, 0.25, 1e-6, nullptr, false, false};
    // the lift pass must have produced the viscous fluxes once before timing the RHS pass alone
    rc = run_stage(op, S, st);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int ph = 0; ph < 2; ph++) {
        ms[ph] = 0.0;
        if (ph == 0 && !op->phys.viscous) continue;
        S.phases = 1 << ph;
        HIP_TRY(hipEventRecord(e0, st));
        for (int i = 0; i < iters; i++) {
            rc = run_stage(op, S, st);
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        ms[ph] = (double)t / iters;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return PRANDTL_OK;
}

int prandtl_op_timing(prandtl_op op, int enable) {
    if (!op) return fail(PRANDTL_ERR_ARG, "null argument");
    op->timing = enable != 0;
    return PRANDTL_OK;
}

int prandtl_op_timing_read(prandtl_op op, double *ms, long long *count) {
    if (!op || !ms) return fail(PRANDTL_ERR_ARG, "null argument");
    double sum[2] = {0.0, 0.0};
    long long n[2] = {0, 0};
    for (size_t i = 0; i < op->tkind.size(); i++) {
        HIP_TRY(hipEventSynchronize(op->tev[2 * i + 1]));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, op->tev[2 * i], op->tev[2 * i + 1]));
        sum[op->tkind[i]] += t;
        n[op->tkind[i]]++;
    }
    for (int k = 0; k < 2; k++) {
        ms[k] = n[k] ? sum[k] / n[k] : 0.0;
        if (count) count[k] = n[k];
    }
    op->tkind.clear();
    op->tev_used = 0;
    return PRANDTL_OK;
}

int prandtl_blend_coefficients(prandtl_op op, const double *u, double *alpha, double *eta, prandtl_stream stream) {
    if (!op || !u || (!alpha && !eta)) return fail(PRANDTL_ERR_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    if (op->dim == 2) {
        switch (op->np) {
            case 2: return launch_blend_coef<2, 2>(op, u, alpha, eta, st);
            case 3: return launch_blend_coef<2, 3>(op, u, alpha, eta, st);
            case 4: return launch_blend_coef<2, 4>(op, u, alpha, eta, st);
            case 5: return launch_blend_coef<2, 5>(op, u, alpha, eta, st);
        }
    } else {
        switch (op->np) {
            case 2: return launch_blend_coef<3, 2>(op, u, alpha, eta, st);
            case 3: return launch_blend_coef<3, 3>(op, u, alpha, eta, st);
            case 4: return launch_blend_coef<3, 4>(op, u, alpha, eta, st);
            case 5: return launch_blend_coef<3, 5>(op, u, alpha, eta, st);
        }
    }
    return fail(PRANDTL_ERR_UNSUPPORTED, "order out of range");
}

int prandtl_filter_constraints(prandtl_op op, const double *u, prandtl_stream stream) {
    if (!op || !u) return fail(PRANDTL_ERR_ARG, "null argument");
    int rc = run_hook(op, HOOK_CONSTRAINTS, const_cast<double *>(u), (hipStream_t)stream);   // reads u only
    if (rc) return rc;
    op->have_bounds = true;
    return PRANDTL_OK;
}

int prandtl_filter_apply(prandtl_op op, double *u, prandtl_stream stream) {
    if (!op || !u) return fail(PRANDTL_ERR_ARG, "null argument");
    return run_hook(op, HOOK_FILTER, u, (hipStream_t)stream);
}

int prandtl_limiter_apply(prandtl_op op, double *u, prandtl_stream stream) {
    if (!op || !u) return fail(PRANDTL_ERR_ARG, "null argument");
    return run_hook(op, HOOK_LIMITER, u, (hipStream_t)stream);
}

int prandtl_hook_counts(prandtl_op op, long long *filtered, long long *limited) {
    if (!op) return fail(PRANDTL_ERR_ARG, "null argument");
    int rc = ensure_hook_buffers(op);
    if (rc) return rc;
    unsigned long long c[3] = {0, 0, 0};
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(c, op->d_hcount, sizeof(c), hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, 0, op->d_hcount, 0ull, 3);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    if (c[2])
        return fail(PRANDTL_ERR_ARG, "a filter or limiter application found its element list out of date: the "
                                     "hooks of one operator must run in order on one stream (or synchronise "
                                     "between streams); the affected applications changed nothing");
    if (filtered) *filtered = (long long)c[0];
    if (limited) *limited = (long long)c[1];
    return PRANDTL_OK;
}

int prandtl_device_alloc(size_t bytes, void **ptr) {
    if (!ptr) return fail(PRANDTL_ERR_ARG, "null argument");
    *ptr = nullptr;
    if (bytes == 0) return PRANDTL_OK;
    HIP_TRY(hipMalloc(ptr, bytes));
    return PRANDTL_OK;
}

void prandtl_device_free(void *ptr) {
    if (ptr) (void)hipFree(ptr);
}

int prandtl_memcpy(void *dst, const void *src, size_t bytes, int kind) {
    if ((!dst || !src) && bytes) return fail(PRANDTL_ERR_ARG, "null argument");
    hipMemcpyKind k;
    switch (kind) {
        case PRANDTL_COPY_H2D: k = hipMemcpyHostToDevice; break;
        case PRANDTL_COPY_D2H: k = hipMemcpyDeviceToHost; break;
        case PRANDTL_COPY_D2D: k = hipMemcpyDeviceToDevice; break;
        default: return fail(PRANDTL_ERR_ARG, "unknown copy kind");
    }
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, k));
    return PRANDTL_OK;
}

int prandtl_stream_synchronize(prandtl_stream stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return PRANDTL_OK;
}

int prandtl_op_vjp(prandtl_op op, const double *u, const double *lam, double *out, prandtl_stream stream) {
    if (!op || !u || !lam || !out) return fail(PRANDTL_ERR_ARG, "null argument");
    if (out == u || out == lam) return fail(PRANDTL_ERR_ARG, "out must not alias u or lam");
    VjpSpec S{VJP_PLAIN, u, lam, nullptr, nullptr, out, nullptr, 0.0, 0.0, 0.0};
    return run_vjp(op, S, (hipStream_t)stream);
}

int prandtl_op_vjp_noalloc(prandtl_op op, const double *u, const double *lam, double *out, prandtl_stream stream) {
    if (!op || !u || !lam || !out) return fail(PRANDTL_ERR_ARG, "null argument");
    if (out == u || out == lam) return fail(PRANDTL_ERR_ARG, "out must not alias u or lam");
    int rc = vjp_supported(op);
    if (rc) return rc;
    VjpBuffers B;
    rc = vjp_scratch_buffers(op, B);
    if (rc) return rc;
    VjpSpec S{VJP_PLAIN, u, lam, nullptr, nullptr, out, nullptr, 0.0, 0.0, 0.0};
    return dispatch_vjp(op, S, (hipStream_t)stream, B);
}

int prandtl_rk_run_tape(prandtl_op op, int scheme, double *u, double *t, double dt, int nsteps, int ckpt_every,
                        prandtl_stream stream) {
    return prandtl_rk_run_tape_schedule(op, scheme, u, t, dt, nsteps, PRANDTL_TAPE_SEGMENTS, ckpt_every, stream);
}

int prandtl_rk_run_tape_schedule(prandtl_op op, int scheme, double *u, double *t, double dt, int nsteps, int schedule,
                                 int param, prandtl_stream stream) {
    if (!op || !u) return fail(PRANDTL_ERR_ARG, "null argument");
    if (scheme != PRANDTL_LSRK45 && scheme != PRANDTL_SSPRK3) return fail(PRANDTL_ERR_ARG, "unknown scheme");
    if (schedule != PRANDTL_TAPE_SEGMENTS && schedule != PRANDTL_TAPE_REVOLVE)
        return fail(PRANDTL_ERR_ARG, "unknown checkpoint schedule");
    if (op->opts.stage_hooks & (PRANDTL_HOOK_FILTER | PRANDTL_HOOK_LIMITER))
        return fail(PRANDTL_ERR_UNSUPPORTED, "the adjoint does not differentiate the filter or limiter hooks");
    if (op->opts.blending) return fail(PRANDTL_ERR_UNSUPPORTED, "the adjoint does not differentiate the blending");
    if (op->nbc || op->d_ngeom) return fail(PRANDTL_ERR_UNSUPPORTED, "the adjoint covers periodic affine meshes");
    if (nsteps <= 0 || param <= 0 || !(dt > 0.0)) return fail(PRANDTL_ERR_ARG, "bad steps, schedule parameter or dt");
    if (schedule == PRANDTL_TAPE_REVOLVE && nsteps > kRevolveMaxSteps)
        return fail(PRANDTL_ERR_ARG, "the revolve schedule takes at most " + std::to_string(kRevolveMaxSteps) + " steps");
    hipStream_t st = (hipStream_t)stream;
    double tt = t ? *t : 0.0;   // the stage times (the adjoint's operators are time-independent)
    // the previous tape is gone from here on: a run that fails leaves no tape
    op->tape_id = 0;
    op->tape_steps = 0;
    // checkpoint buffers and the steps whose states the forward run stores in them
    std::vector<int> marks;
    if (schedule == PRANDTL_TAPE_SEGMENTS) {
        for (int k = 0; k * param < nsteps; k++) marks.push_back(k * param);
    } else {
        // revolve: the initial state and the first snapshot of every level of the schedule's
        // rightmost branch (the reverse sweep's first descent), taken on the way
        marks.push_back(0);
        const RevolvePlan &plan = revolve_plan(op, nsteps, param);
        int a = 0, sl = param;
        while (nsteps - a > 1 && sl > 0) {
            a += plan.split(nsteps - a, sl);
            marks.push_back(a);
            sl--;
        }
    }
    const size_t nbuf = schedule == PRANDTL_TAPE_SEGMENTS ? marks.size() : (size_t)param + 1;
    if (op->tape_ckpt.size() != nbuf) {
        release_adjoint_storage(op);            // its slot 0 points at a checkpoint
        for (double *p : op->tape_ckpt) (void)hipFree(p);
        op->tape_ckpt.assign(nbuf, nullptr);
        for (size_t k = 0; k < nbuf; k++)
            if (hipMalloc(&op->tape_ckpt[k], op->size * sizeof(double)) != hipSuccess) {
                for (double *p : op->tape_ckpt) (void)hipFree(p);
                op->tape_ckpt.clear();
                return fail(PRANDTL_ERR_HIP, "checkpoint storage allocation failed");
            }
    }
    int cur = 0;
    for (size_t k = 0; k < marks.size(); k++) {
        int rc = rk_advance(op, scheme, u, &tt, dt, marks[k] - cur, 0, st);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(op->tape_ckpt[k], u, op->size * sizeof(double), hipMemcpyDeviceToDevice, st));
        cur = marks[k];
    }
    int rc = rk_advance(op, scheme, u, &tt, dt, nsteps - cur, 0, st);
    if (rc) return rc;
    op->tape_marks = marks;
    op->tape_schedule = schedule;
    op->tape_param = param;
    op->tape_steps = nsteps;
    op->tape_every = schedule == PRANDTL_TAPE_SEGMENTS ? param : 1;
    op->tape_dt = dt;
    op->tape_scheme = scheme;
    op->tape_t0 = t ? *t : 0.0;
    op->tape_id = next_tape_id();
    op->tape_recomputed = -1;
    if (t) *t = tt;
    return PRANDTL_OK;
}

int prandtl_rk_adjoint(prandtl_op op, double *ubar, prandtl_stream stream) {
    if (!op || !ubar) return fail(PRANDTL_ERR_ARG, "null argument");
    if (op->tape_ckpt.empty() || op->tape_id == 0) return fail(PRANDTL_ERR_ARG, "no taped run on this operator");
    RoctxRange range("prandtl.adjoint");
    hipStream_t st = (hipStream_t)stream;
    const size_t bytes = op->size * sizeof(double);
    const int C = op->tape_every, nsteps = op->tape_steps;
    const bool ssp = op->tape_scheme == PRANDTL_SSPRK3;
    const int per = ssp ? 3 : 5;   // stage inputs kept per step
    int rc = ensure_rk_buffers(op, op->tape_scheme);
    if (rc) return rc;
    if (!ssp) rc = ensure_rk_buffers(op, PRANDTL_LSRK45);
    if (rc) return rc;
    // stage inputs of one block of steps (slot 0 is the block's start state itself), the adjoint
    // registers (double-buffered: a stage reads its neighbours' adjoints) and, for revolve, a work
    // state for advancing
    const size_t nstage = (size_t)C * per;
    if (op->adj_stage.size() != nstage) {
        release_adjoint_storage(op);
        op->adj_stage.assign(nstage, nullptr);
        for (size_t i = 1; i < op->adj_stage.size(); i++)
            if (hipMalloc(&op->adj_stage[i], bytes) != hipSuccess) {
                release_adjoint_storage(op);
                return fail(PRANDTL_ERR_HIP, "adjoint stage storage allocation failed");
            }
        if (hipMalloc(&op->adj_ub2, bytes) != hipSuccess || hipMalloc(&op->adj_kb, bytes) != hipSuccess ||
            hipMalloc(&op->adj_kb2, bytes) != hipSuccess || hipMalloc(&op->adj_work, bytes) != hipSuccess) {
            release_adjoint_storage(op);
            return fail(PRANDTL_ERR_HIP, "adjoint register allocation failed");
        }
    }
    std::vector<double *> &stage = op->adj_stage;
    const double dt = op->tape_dt;
    long long recomputed = 0;
    // LSRK4(5): adjoint registers (ub, kb); SSP-RK3: (U, Y2, Y1, X)
    double *ub_cur = ubar, *ub_nxt = op->adj_ub2, *kb_cur = op->adj_kb, *kb_nxt = op->adj_kb2;
    double *U = ubar, *Y2 = op->adj_kb, *Y1 = op->adj_kb2, *X = op->adj_ub2;
    if (!ssp) hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, st, kb_cur, 0.0, op->size);
    // reverse the n steps that start from state x0 (not modified): recompute them keeping every
    // stage input, then run the stage adjoints backwards
    auto reverse_block = [&](double *x0, int n) -> int {
        int r = PRANDTL_OK;
        recomputed += n;
        stage[0] = x0;
        if (ssp) {
            // SSP-RK3 (Shu-Osher): y1 = u + dt L(u), y2 = 3/4 u + 1/4 (y1 + dt L(y1)),
            // u' = 1/3 u + 2/3 (y2 + dt L(y2)).  Reverse, given U = adjoint of u':
            //   yb2 = 2/3 (U + dt J(y2)^T U),  yb1 = 1/4 (yb2 + dt J(y1)^T yb2),
            //   U <- 1/3 U + 3/4 yb2 + yb1 + dt J(u)^T yb1.
            for (int j = 0; j < n && r == PRANDTL_OK; j++) {
                double *u0 = stage[3 * j];
                StageSpec S1{EPI_SSP, u0, stage[3 * j + 1], nullptr, u0, 0.0, 1.0, dt, nullptr, false, false};
                r = run_stage(op, S1, st);
                StageSpec S2{EPI_SSP, stage[3 * j + 1], stage[3 * j + 2], nullptr, u0, 0.75, 0.25, dt, nullptr,
                             false, false};
                if (!r) r = run_stage(op, S2, st);
                if (!r && j + 1 < n) {
                    StageSpec S3{EPI_SSP, stage[3 * j + 2], stage[3 * j + 3], nullptr, u0, 1.0 / 3.0, 2.0 / 3.0, dt,
                                 nullptr, false, false};
                    r = run_stage(op, S3, st);
                }
            }
            for (int j = n - 1; j >= 0 && r == PRANDTL_OK; j--) {
                VjpSpec Va{VJP_SSP, stage[3 * j + 2], U, nullptr, nullptr, Y2, nullptr, 0.0, 2.0 / 3.0, dt, nullptr};
                r = run_vjp(op, Va, st);
                VjpSpec Vb{VJP_SSP, stage[3 * j + 1], Y2, nullptr, nullptr, Y1, nullptr, 0.0, 0.25, dt, nullptr};
                if (!r) r = run_vjp(op, Vb, st);
                if (!r) {
                    hipLaunchKernelGGL(k_axpby, dim3(2048), dim3(256), 0, st, X, 1.0 / 3.0, (const double *)U, 0.75,
                                       (const double *)Y2, op->size);
                    VjpSpec Vc{VJP_SSP, stage[3 * j], Y1, nullptr, nullptr, X, nullptr, 0.0, 1.0, dt, X};
                    r = run_vjp(op, Vc, st);
                }
                std::swap(U, X);
            }
            return r;
        }
        // LSRK4(5): recompute the stage inputs (the last stage's output is not needed)
        for (int j = 0; j < 5 * n - 1 && r == PRANDTL_OK; j++) {
            const int stg = j % 5;
            StageSpec S{EPI_LSRK, stage[j], stage[j + 1], op->d_k, nullptr, CK_A[stg], CK_B[stg], dt, nullptr, false,
                        false};
            r = run_stage(op, S, st);
        }
        for (int j = 5 * n - 1; j >= 0 && r == PRANDTL_OK; j--) {
            const int stg = j % 5;
            VjpSpec V{VJP_LSRK, stage[j], nullptr, ub_cur, kb_cur, ub_nxt, kb_nxt, CK_A[stg], CK_B[stg], dt};
            r = run_vjp(op, V, st);
            std::swap(ub_cur, ub_nxt);
            std::swap(kb_cur, kb_nxt);
        }
        return r;
    };
    if (op->tape_schedule == PRANDTL_TAPE_SEGMENTS) {
        const int nseg = (int)op->tape_ckpt.size();
        for (int seg = nseg - 1; seg >= 0 && rc == PRANDTL_OK; seg--)
            rc = reverse_block(op->tape_ckpt[seg], std::min(C, nsteps - seg * C));
    } else {
        // Griewank's binomial schedule: reverse(a, b, s) reverses steps [a, b) from the state at a
        // with s free snapshot buffers -- a snapshot at the optimal split m, the steps after it
        // with s-1 buffers, then the steps before it with s
        const RevolvePlan &plan = revolve_plan(op, nsteps, op->tape_param);
        std::vector<double *> pool(op->tape_ckpt.begin() + 1, op->tape_ckpt.end());
        std::vector<int> taken(op->tape_marks.begin() + 1, op->tape_marks.end());   // pool[i] holds step taken[i]
        taken.resize(pool.size(), -1);
        std::function<int(int, int, int, double *)> rev = [&](int a, int b, int sl, double *xa) -> int {
            const int n = b - a;
            if (n == 1) return reverse_block(xa, 1);
            if (sl == 0) {
                // no buffer left: advance from a to every step, last first
                for (int j = b - 1; j >= a; j--) {
                    int r;
                    if (j == a) {
                        r = reverse_block(xa, 1);
                    } else {
                        HIP_TRY(hipMemcpyAsync(op->adj_work, xa, bytes, hipMemcpyDeviceToDevice, st));
                        double tj = op->tape_t0 + a * dt;   // (no time-dependent terms on the adjoint's meshes)
                        r = rk_advance(op, op->tape_scheme, op->adj_work, &tj, dt, j - a, 0, st);
                        recomputed += j - a;
                        if (!r) r = reverse_block(op->adj_work, 1);
                    }
                    if (r) return r;
                }
                return PRANDTL_OK;
            }
            const int m = a + plan.split(n, sl);
            int slot = -1;
            for (size_t i = 0; i < pool.size(); i++)
                if (taken[i] == m) slot = (int)i;
            if (slot < 0) {
                for (size_t i = 0; i < pool.size() && slot < 0; i++)
                    if (taken[i] < 0) slot = (int)i;
                if (slot < 0) return fail(PRANDTL_ERR_HIP, "revolve: no free snapshot buffer");
                HIP_TRY(hipMemcpyAsync(pool[slot], xa, bytes, hipMemcpyDeviceToDevice, st));
                double tm = op->tape_t0 + a * dt;
                const int r = rk_advance(op, op->tape_scheme, pool[slot], &tm, dt, m - a, 0, st);
                if (r) return r;
                recomputed += m - a;
                taken[slot] = m;
            }
            int r = rev(m, b, sl - 1, pool[slot]);
            taken[slot] = -1;
            if (r) return r;
            return rev(a, m, sl, xa);
        };
        rc = rev(0, nsteps, op->tape_param, op->tape_ckpt[0]);
    }
    double *res = ssp ? U : ub_cur;
    if (rc == PRANDTL_OK && res != ubar) (void)hipMemcpyAsync(ubar, res, bytes, hipMemcpyDeviceToDevice, st);
    hipError_t e = hipStreamSynchronize(st);
    if (rc) return rc;
    if (e != hipSuccess) return fail(PRANDTL_ERR_HIP, hipGetErrorString(e));
    op->tape_recomputed = recomputed;
    return PRANDTL_OK;
}

int prandtl_rk_tape_info(prandtl_op op, long long *states, long long *recomputed_steps) {
    if (!op) return fail(PRANDTL_ERR_ARG, "null argument");
    if (op->tape_id == 0) return fail(PRANDTL_ERR_ARG, "no taped run on this operator");
    const int per = op->tape_scheme == PRANDTL_SSPRK3 ? 3 : 5;
    // checkpoints + the stage inputs of one block (its first is a checkpoint) + the work state
    const long long held = (long long)op->tape_ckpt.size() + (long long)op->tape_every * per - 1 +
                           (op->tape_schedule == PRANDTL_TAPE_REVOLVE ? 1 : 0);
    if (states) *states = held;
    if (recomputed_steps) *recomputed_steps = op->tape_recomputed;
    return PRANDTL_OK;
}

int prandtl_rk_tape_id(prandtl_op op, unsigned long long *id) {
    if (!op || !id) return fail(PRANDTL_ERR_ARG, "null argument");
    *id = op->tape_id;
    return PRANDTL_OK;
}

int prandtl_rk_step(prandtl_op op, int scheme, double *u, double *t, double dt, prandtl_stream stream) {
    return prandtl_rk_run(op, scheme, u, t, dt, 1, 0, stream);
}

int prandtl_rk_run(prandtl_op op, int scheme, double *u, double *t, double dt, int nsteps, int nan_check,
                   prandtl_stream stream) {
    if (!op || !u) return fail(PRANDTL_ERR_ARG, "null argument");
    if (scheme < PRANDTL_SSPRK3 || scheme > PRANDTL_RK4) return fail(PRANDTL_ERR_ARG, "unknown ODE solver");
    if (nsteps < 0 || !(dt >= 0.0)) return fail(PRANDTL_ERR_ARG, "bad step count or dt");
    double tt = t ? *t : op->time;
    int rc = rk_advance(op, scheme, u, &tt, dt, nsteps, nan_check, (hipStream_t)stream);
    if (rc == PRANDTL_OK && t) *t = tt;
    return rc;
}

int prandtl_op_set_time(prandtl_op op, double t) {
    if (!op) return fail(PRANDTL_ERR_ARG, "null argument");
    if (!std::isfinite(t)) return fail(PRANDTL_ERR_ARG, "the time must be finite");
    op->time = t;
    return PRANDTL_OK;
}

int prandtl_op_get_time(prandtl_op op, double *t) {
    if (!op || !t) return fail(PRANDTL_ERR_ARG, "null argument");
    *t = op->time;
    return PRANDTL_OK;
}

}  // extern "C"
