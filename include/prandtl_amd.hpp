// prandtl_amd.hpp -- C++ host-side mirror of the reference's operator, solver, filter and
// limiter classes over the C ABI of prandtl_amd.h.  Header-only; a host program compiles with
// any C++17 compiler and links libprandtl_amd.so, with no HIP headers of its own.
//
// Reference interfaces mirrored (paths relative to the Prandtl repository root):
//   mfem::TimeDependentOperator / DGSEMOperator   src/Operators/DGSEMOperator.hpp:15-115
//   NumericalFlux (surface flux plug-in)          src/RiemannSolvers/NumericalFlux.hpp:21-29
//   mfem::ODESolver Init / Step                   src/Simulation/Simulation.cpp:177-206, 807, 1010
//   Filter / EntropyFilter                        src/Filters/Filter.hpp:11-30, EntropyFilter.hpp
//   Limiter                                       src/Limiters/Limiter.hpp:10-16
//   blending coefficient (Indicator / alpha)      src/Operators/DGSEMOperator.cpp:60-80,
//                                                 src/Indicators/PerssonPeraireIndicator.cpp:21-58
// The virtual per-point flux calls of the reference cannot cross to the device; the surface
// flux is chosen at construction (SurfaceFlux) and compiled into the kernels.
//
// Errors: the reference aborts through mfem_error; here a failing call throws prandtl::Error
// carrying prandtl_last_error(), which a caller that wants the reference's behaviour lets
// escape.  Device vectors live in HBM in the reference's byNODES layout.
#pragma once

#include <array>
#include <cstddef>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "prandtl_amd.h"

namespace prandtl {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc) {
    if (rc != PRANDTL_OK) throw Error(rc, prandtl_last_error());
}

// A device-resident FP64 vector (the role of mfem::Vector with its device memory).
class Vector {
public:
    Vector() = default;
    explicit Vector(size_t n) : n_(n) {
        void *p = nullptr;
        check(prandtl_device_alloc(n * sizeof(double), &p));
        d_ = static_cast<double *>(p);
    }
    Vector(const Vector &) = delete;
    Vector &operator=(const Vector &) = delete;
    Vector(Vector &&o) noexcept : d_(o.d_), n_(o.n_) { o.d_ = nullptr, o.n_ = 0; }
    Vector &operator=(Vector &&o) noexcept {
        std::swap(d_, o.d_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~Vector() { prandtl_device_free(d_); }

    size_t Size() const { return n_; }
    const double *Read() const { return d_; }
    double *Write() { return d_; }
    double *ReadWrite() { return d_; }
    void CopyFromHost(const double *h) { check(prandtl_memcpy(d_, h, n_ * sizeof(double), PRANDTL_COPY_H2D)); }
    void CopyToHost(double *h) const { check(prandtl_memcpy(h, d_, n_ * sizeof(double), PRANDTL_COPY_D2H)); }
    std::vector<double> HostCopy() const {
        std::vector<double> h(n_);
        CopyToHost(h.data());
        return h;
    }

private:
    double *d_ = nullptr;
    size_t n_ = 0;
};

// The surface flux plug-in (NumericalFlux::ComputeFaceFlux).  The two-point volume flux is
// the reference's Chandrashekar flux (ChandrashekarFlux::ComputeVolumeFlux) in every case.
enum class SurfaceFlux { Chandrashekar = PRANDTL_SURF_CHANDRASHEKAR, EC = PRANDTL_SURF_EC, Rusanov = PRANDTL_SURF_RUSANOV };
enum class Precision { FP64 = PRANDTL_FP64, Mixed = PRANDTL_MIXED };

// Conforming mesh: per element its 2 dim face neighbours and (det J, adj J); a general
// (curved or unstructured) mesh adds the per-node geometry and the face orientation codes
// (their layout: prandtl_mesh_desc).
struct Mesh {
    int dim = 0, order = 0, num_elements = 0;
    std::vector<int> neighbors;         // [e][2 d + s]; -(k+1): a boundary face with condition bc[k]
    std::vector<double> geometry;       // [e][1 + dim^2] (may be empty when node_geometry is given)
    std::vector<prandtl_bc> bc;         // boundary conditions (empty for a periodic mesh)
    std::vector<double> node_geometry;  // general mesh: [e][node][1 + dim^2] (empty: affine elements)
    std::vector<int> face_orientation;  // general mesh: [e][2 dim] (empty: aligned lattice faces)
    std::vector<double> node_coords;    // [e][node][dim]: needed when a boundary state has a front profile

    // Periodic lattice of n[0] x n[1] (x n[2]) axis-aligned boxes of the domain
    // [origin, origin + length] (the generated meshes of the benchmark cases).
    static Mesh PeriodicLattice(int dim, int order, const std::array<int, 3> &n, const std::array<double, 3> &length) {
        Mesh m;
        m.dim = dim;
        m.order = order;
        const int nx = n[0], ny = n[1], nz = dim == 3 ? n[2] : 1;
        m.num_elements = nx * ny * nz;
        m.neighbors.resize((size_t)m.num_elements * 2 * dim);
        m.geometry.assign((size_t)m.num_elements * (1 + dim * dim), 0.0);
        double h[3] = {length[0] / nx, length[1] / ny, dim == 3 ? length[2] / nz : 1.0};
        double det = 1.0;
        for (int d = 0; d < dim; d++) det *= h[d];
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const int e = i + nx * (j + ny * k);
                    const int c[3] = {i, j, k}, nn[3] = {nx, ny, nz};
                    for (int d = 0; d < dim; d++)
                        for (int s = 0; s < 2; s++) {
                            int cc[3] = {c[0], c[1], c[2]};
                            cc[d] = (cc[d] + (s ? 1 : nn[d] - 1)) % nn[d];
                            m.neighbors[(size_t)e * 2 * dim + 2 * d + s] = cc[0] + nx * (cc[1] + ny * cc[2]);
                        }
                    double *g = &m.geometry[(size_t)e * (1 + dim * dim)];
                    g[0] = det;
                    for (int d = 0; d < dim; d++) g[1 + d * dim + d] = det / h[d];   // adj of diag(h)
                }
        return m;
    }
    int NodesPerElement() const {
        int n = 1;
        for (int d = 0; d < dim; d++) n *= order + 1;
        return n;
    }
    size_t StateSize() const { return (size_t)(dim + 2) * num_elements * NodesPerElement(); }
};

struct Physics {
    double gamma = 1.4, Pr = 0.72, mu = 0.0, R_gas = 287.0;
    bool viscous = false, sutherland = false;
};

struct Options {
    SurfaceFlux surface_flux = SurfaceFlux::Chandrashekar;
    Precision precision = Precision::FP64;
    int bug_compat = 0;     // PRANDTL_BUG_* bitmask
    int stage_hooks = 0;    // PRANDTL_HOOK_*: filter / limiter inside the fused steppers
    bool blending = false;  // subcell finite-volume blending (SUBCELL_FV_BLENDING), FP64
    double alpha_max = 0.5; // Simulation.cpp:477-484
    double alpha_min = 0.001;  // DGSEMOperator.hpp:99
};

// mfem::TimeDependentOperator's surface as the reference uses it.
class TimeDependentOperator {
public:
    explicit TimeDependentOperator(int height) : height_(height) {}
    virtual ~TimeDependentOperator() = default;
    virtual void Mult(const Vector &u, Vector &dudt) const = 0;
    virtual void SetTime(double t) { t_ = t; }
    double GetTime() const { return t_; }
    double &GetTimeRef() { return t_; }
    int Height() const { return height_; }

protected:
    int height_;
    double t_ = 0.0;
};

// DGSEMOperator -- DGSEMOperator.hpp:87-115 over prandtl_op_*.
class DGSEMOperator : public TimeDependentOperator {
public:
    DGSEMOperator(const Mesh &mesh, const Physics &phys, const Options &opts = {})
        : TimeDependentOperator((int)mesh.StateSize()) {
        prandtl_mesh_desc md{mesh.dim, mesh.order, mesh.num_elements, mesh.neighbors.data(),
                             mesh.geometry.empty() ? nullptr : mesh.geometry.data(),
                             (int)mesh.bc.size(), mesh.bc.empty() ? nullptr : mesh.bc.data(),
                             mesh.node_geometry.empty() ? nullptr : mesh.node_geometry.data(),
                             mesh.face_orientation.empty() ? nullptr : mesh.face_orientation.data(),
                             mesh.node_coords.empty() ? nullptr : mesh.node_coords.data()};
        prandtl_physics ph{phys.gamma, phys.Pr, phys.mu, phys.R_gas, phys.viscous ? 1 : 0, phys.sutherland ? 1 : 0};
        prandtl_opts o{(int)opts.surface_flux, (int)opts.precision, opts.bug_compat, opts.stage_hooks,
                       opts.blending ? 1 : 0, opts.alpha_max, opts.alpha_min};
        check(prandtl_op_create(&md, &ph, &o, &op_));
        num_elements_ = mesh.num_elements;
    }
    DGSEMOperator(const DGSEMOperator &) = delete;
    DGSEMOperator &operator=(const DGSEMOperator &) = delete;
    ~DGSEMOperator() override { prandtl_op_destroy(op_); }

    // dudt = L(u), fully overwritten (DGSEMOperator.cpp:537-591), with time-dependent boundary
    // states at the operator's time (SetTime, or GetTimeRef as the reference's integrators hold it)
    void Mult(const Vector &u, Vector &dudt) const override {
        CheckState(u, "u");
        CheckState(dudt, "dudt");
        check(prandtl_op_set_time(op_, t_));
        check(prandtl_op_mult(op_, u.Read(), dudt.Write(), stream_));
    }
    // a state vector of this operator: (dim + 2) x elements x nodes doubles (GetTrueVSize)
    void CheckState(const Vector &x, const char *what) const {
        if (x.Size() != (size_t)Height())
            throw Error(PRANDTL_ERR_ARG, std::string(what) + " holds " + std::to_string(x.Size()) +
                                             " values; the operator's state holds " + std::to_string(Height()));
    }
    // max(|V| + c) of the last Mult (DGSEMOperator.hpp:106-109)
    double GetMaxCharSpeed() const {
        double l = 0.0;
        check(prandtl_op_get_max_char_speed(op_, &l));
        return l;
    }
    // the same quantity from a state alone
    double MaxCharSpeed(const Vector &u) const {
        CheckState(u, "u");
        double l = 0.0;
        check(prandtl_op_max_char_speed(op_, u.Read(), &l, stream_));
        return l;
    }
    // the blending coefficient alpha and indicator eta of every element of u (the reference's
    // alpha and eta grid functions, DGSEMOperator::ComputeBlendingCoefficient); each Vector has
    // num_elements entries
    void ComputeBlendingCoefficient(const Vector &u, Vector &alpha, Vector &eta) const {
        CheckState(u, "u");
        if (alpha.Size() != (size_t)num_elements_ || eta.Size() != (size_t)num_elements_)
            throw Error(PRANDTL_ERR_ARG, "alpha and eta hold one value per element");
        check(prandtl_blend_coefficients(op_, u.Read(), alpha.Write(), eta.Write(), stream_));
    }
    void SetStream(prandtl_stream s) { stream_ = s; }
    prandtl_stream Stream() const { return stream_; }
    prandtl_op Handle() const { return op_; }

private:
    prandtl_op op_ = nullptr;
    prandtl_stream stream_ = nullptr;
    int num_elements_ = 0;
};

// mfem::ODESolver's Init / Step.  The steppers run the fused stage kernels of the library, so
// they step a DGSEMOperator (Init throws for any other operator).
class ODESolver {
public:
    virtual ~ODESolver() = default;
    virtual void Init(TimeDependentOperator &f) {
        op_ = dynamic_cast<DGSEMOperator *>(&f);
        if (!op_) throw Error(PRANDTL_ERR_UNSUPPORTED, "the fused steppers advance a prandtl::DGSEMOperator");
    }
    // x advanced by one step in place; t += dt (dt unchanged: fixed-step solvers)
    virtual void Step(Vector &x, double &t, double &dt) {
        if (!op_) throw Error(PRANDTL_ERR_ARG, "ODESolver::Step before Init");
        op_->CheckState(x, "x");
        check(prandtl_rk_step(op_->Handle(), scheme(), x.ReadWrite(), &t, dt, op_->Stream()));
        op_->SetTime(t);
    }
    // nsteps steps at once (the time loop without output), NaN scan every nan_check steps
    void Run(Vector &x, double &t, double dt, int nsteps, int nan_check = 0) {
        if (!op_) throw Error(PRANDTL_ERR_ARG, "ODESolver::Run before Init");
        op_->CheckState(x, "x");
        check(prandtl_rk_run(op_->Handle(), scheme(), x.ReadWrite(), &t, dt, nsteps, nan_check, op_->Stream()));
        op_->SetTime(t);
    }

protected:
    virtual int scheme() const = 0;
    DGSEMOperator *op_ = nullptr;
};

// MFEM RK3SSPSolver (Shu-Osher form)
class SSPRK3Solver : public ODESolver {
protected:
    int scheme() const override { return PRANDTL_SSPRK3; }
};

// Carpenter-Kennedy 2N-storage RK4(3)5
class LSRK45Solver : public ODESolver {
protected:
    int scheme() const override { return PRANDTL_LSRK45; }
};

// MFEM ForwardEulerSolver, RK2Solver (a = 2/3) and RK4Solver (the classic tableau)
class ForwardEulerSolver : public ODESolver {
protected:
    int scheme() const override { return PRANDTL_FORWARD_EULER; }
};

class RK2Solver : public ODESolver {
protected:
    int scheme() const override { return PRANDTL_RK2; }
};

class RK4Solver : public ODESolver {
protected:
    int scheme() const override { return PRANDTL_RK4; }
};

// Filter -- Filter.hpp:11-30
class Filter {
public:
    virtual ~Filter() = default;
    virtual void FilterSolution(Vector &sol) = 0;
    virtual void GetFilterConstraints(const Vector &x) = 0;
};

// EntropyFilter -- EntropyFilter.hpp / EntropyFilter.cpp, on the operator's mesh
class EntropyFilter : public Filter {
public:
    explicit EntropyFilter(const DGSEMOperator &op) : op_(op) {}
    void GetFilterConstraints(const Vector &x) override {
        op_.CheckState(x, "x");
        check(prandtl_filter_constraints(op_.Handle(), x.Read(), op_.Stream()));
    }
    void FilterSolution(Vector &sol) override {
        op_.CheckState(sol, "sol");
        check(prandtl_filter_apply(op_.Handle(), sol.ReadWrite(), op_.Stream()));
    }

private:
    const DGSEMOperator &op_;
};

// Limiter -- Limiter.hpp:10-16
class Limiter {
public:
    virtual ~Limiter() = default;
    virtual void LimitSolution(Vector &sol) = 0;
};

// Zhang-Shu positivity limiter (the reference has only the interface)
class PositivityLimiter : public Limiter {
public:
    explicit PositivityLimiter(const DGSEMOperator &op) : op_(op) {}
    void LimitSolution(Vector &sol) override {
        op_.CheckState(sol, "sol");
        check(prandtl_limiter_apply(op_.Handle(), sol.ReadWrite(), op_.Stream()));
    }

private:
    const DGSEMOperator &op_;
};

// Completes the work queued on the operator's stream.
inline void Synchronize(const DGSEMOperator &op) { check(prandtl_stream_synchronize(op.Stream())); }

// Checkpoints in the reference's format -- Simulation::Run, src/Simulation/Simulation.cpp:1084-1116
// (write) and :350-390 (restart): <folder>/Cycle<n>/checkpoint_cycle_<n>.<rank:08d>.chk holds
// MFEM's GridFunction::Save text of the byNODES DG state (space header L2_T1_<dim>D_P<order>,
// VDim dim+2, Ordering 0, then one value per line), checkpoint_cycle_<n>.json the time and cycle.
// Host arrays in the state layout u[v][e][node]; values are written with 17 significant digits.
inline std::string CheckpointFile(const std::string &folder, int cycle, int rank = 0) {
    char r[16];
    std::snprintf(r, sizeof r, "%08d", rank);
    return folder + "/Cycle" + std::to_string(cycle) + "/checkpoint_cycle_" + std::to_string(cycle) + "." + r + ".chk";
}

inline void SaveCheckpoint(const std::string &folder, int cycle, double t, const double *u, size_t n, int dim,
                           int order, int rank = 0) {
    const std::string dir = folder + "/Cycle" + std::to_string(cycle);
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) throw Error(PRANDTL_ERR_ARG, "cannot create " + dir + ": " + ec.message());
    std::ofstream out(CheckpointFile(folder, cycle, rank));
    if (!out) throw Error(PRANDTL_ERR_ARG, "cannot write " + CheckpointFile(folder, cycle, rank));
    out << "FiniteElementSpace\nFiniteElementCollection: L2_T1_" << dim << "D_P" << order << "\nVDim: " << dim + 2
        << "\nOrdering: 0\n\n";
    out.precision(17);
    for (size_t i = 0; i < n; i++) out << u[i] << '\n';
    if (rank == 0) {
        std::ofstream meta(dir + "/checkpoint_cycle_" + std::to_string(cycle) + ".json");
        meta.precision(17);
        meta << "{\n \"time\": " << t << ",\n \"cycle\": " << cycle << "\n}\n";
    }
}

// Reads the checkpoint of `cycle` into u (resized to the file's values); returns its time.  The
// space must be the DG space of (dim, order) and the meta file's cycle must match.
inline double LoadCheckpoint(const std::string &folder, int cycle, int dim, int order, std::vector<double> &u,
                             int rank = 0) {
    const std::string dir = folder + "/Cycle" + std::to_string(cycle);
    std::ifstream meta(dir + "/checkpoint_cycle_" + std::to_string(cycle) + ".json");
    if (!meta) throw Error(PRANDTL_ERR_ARG, "no meta file in " + dir);
    std::stringstream ms;
    ms << meta.rdbuf();
    const std::string m = ms.str();
    auto field = [&](const char *key) {
        const size_t k = m.find(std::string("\"") + key + "\"");
        if (k == std::string::npos) throw Error(PRANDTL_ERR_ARG, std::string("meta file lacks ") + key);
        return std::stod(m.substr(m.find(':', k) + 1));
    };
    const double t = field("time");
    if ((int)field("cycle") != cycle) throw Error(PRANDTL_ERR_ARG, "mismatch between cycle number and meta file");
    std::ifstream in(CheckpointFile(folder, cycle, rank));
    if (!in) throw Error(PRANDTL_ERR_ARG, "cannot read " + CheckpointFile(folder, cycle, rank));
    std::string line;
    std::getline(in, line);
    if (line != "FiniteElementSpace" && line != "MFEM FiniteElementSpace v1.0")
        throw Error(PRANDTL_ERR_ARG, "not an MFEM grid function");
    const std::string fec = "L2_T1_" + std::to_string(dim) + "D_P" + std::to_string(order);
    bool fec_ok = false, ord_ok = false, vdim_ok = false;
    while (std::getline(in, line) && !line.empty()) {
        if (line == "FiniteElementCollection: " + fec) fec_ok = true;
        if (line == "Ordering: 0") ord_ok = true;
        if (line == "VDim: " + std::to_string(dim + 2)) vdim_ok = true;
    }
    if (!fec_ok || !ord_ok || !vdim_ok)
        throw Error(PRANDTL_ERR_ARG, "checkpoint space is not " + fec + " with VDim " + std::to_string(dim + 2) +
                                         ", Ordering 0");
    u.clear();
    double v;
    while (in >> v) u.push_back(v);
    return t;
}

}  // namespace prandtl
