// A host program written against the C++ mirror (include/prandtl_amd.hpp) the way the
// reference's Simulation drives its operator: build the mesh and the operator, Mult once,
// then advance with an ODE solver's Init/Step, and apply the filter and limiter hooks.
// Built with g++ (no HIP headers) by tests/test_gpu_host_mirror.py, which checks the files it
// writes against the oracle.
//
// usage: host_mirror_example <dir>
//   reads  <dir>/u0.bin      vortex state, 16 x 16 quads p=3 on [0,10]^2
//          <dir>/bad.bin     a state with inadmissible nodes (same mesh)
//   writes <dir>/dudt.bin    Mult(u0)
//          <dir>/ssp.bin     u0 after 10 SSPRK3Solver steps of dt
//          <dir>/lsrk.bin    u0 after 4 LSRK45Solver steps of dt (one Run call)
//          <dir>/hooks.bin   bad after EntropyFilter (constraints from u0) and PositivityLimiter
//          <dir>/blend.bin   Mult(u0) of an operator with subcell finite-volume blending
//          <dir>/alpha.bin   its blending coefficients (one per element)
//   general mesh (when <dir>/cavity.msh exists: the reference's lid-driven cavity mesh, read and
//   built by include/prandtl_amd_mesh.hpp; NS, no-slip walls and a moving lid, blending):
//   reads  <dir>/cav_u0.bin
//   writes <dir>/cav_dudt.bin Mult(cav_u0)
//          <dir>/cav_rk4.bin  cav_u0 after 3 RK4Solver steps of dt = 1e-3
//          <dir>/cav_ng.bin, cav_mesh.bin  the mesh as built (per-node geometry; neighbours, codes)
#include <cstdio>
#include <string>
#include <vector>

#include "prandtl_amd.hpp"
#include "prandtl_amd_mesh.hpp"

static std::vector<double> load(const std::string &path, size_t n) {
    std::vector<double> v(n);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(double), n, f) != n) throw std::runtime_error("cannot read " + path);
    std::fclose(f);
    return v;
}

static void save(const std::string &path, const std::vector<double> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    try {
        const prandtl::Mesh mesh = prandtl::Mesh::PeriodicLattice(2, 3, {16, 16, 1}, {10.0, 10.0, 1.0});
        prandtl::Physics phys;   // Euler, gamma = 1.4
        prandtl::DGSEMOperator op(mesh, phys);
        const size_t n = mesh.StateSize();

        prandtl::Vector u(n), dudt(n);
        u.CopyFromHost(load(dir + "/u0.bin", n).data());
        op.Mult(u, dudt);
        save(dir + "/dudt.bin", dudt.HostCopy());
        const double lam = op.GetMaxCharSpeed();
        const double dt = 0.5 * (10.0 / 16) / (lam * 16.0);   // Simulation.cpp:906, CFL 0.5
        std::printf("lambda_max %.15e dt %.15e\n", lam, dt);

        prandtl::SSPRK3Solver ssp;
        ssp.Init(op);
        double t = 0.0, h = dt;
        for (int s = 0; s < 10; s++) ssp.Step(u, t, h);
        save(dir + "/ssp.bin", u.HostCopy());

        u.CopyFromHost(load(dir + "/u0.bin", n).data());
        prandtl::LSRK45Solver lsrk;
        lsrk.Init(op);
        t = 0.0;
        lsrk.Run(u, t, dt, 4);
        save(dir + "/lsrk.bin", u.HostCopy());
        std::printf("t after LSRK45 run %.15e\n", t);

        prandtl::EntropyFilter filter(op);
        prandtl::PositivityLimiter limiter(op);
        u.CopyFromHost(load(dir + "/u0.bin", n).data());
        filter.GetFilterConstraints(u);
        prandtl::Vector w(n);
        w.CopyFromHost(load(dir + "/bad.bin", n).data());
        filter.FilterSolution(w);
        limiter.LimitSolution(w);
        prandtl::Synchronize(op);
        save(dir + "/hooks.bin", w.HostCopy());

        // subcell finite-volume blending (the reference's SUBCELL_FV_BLENDING build)
        prandtl::Options bo;
        bo.blending = true;
        prandtl::DGSEMOperator bop(mesh, phys, bo);
        prandtl::Vector alpha((size_t)mesh.num_elements), eta((size_t)mesh.num_elements);
        w.CopyFromHost(load(dir + "/u0.bin", n).data());
        bop.Mult(w, dudt);
        bop.ComputeBlendingCoefficient(w, alpha, eta);
        save(dir + "/blend.bin", dudt.HostCopy());
        save(dir + "/alpha.bin", alpha.HostCopy());

        // a general mesh from the reference's Gmsh file
        if (FILE *probe = std::fopen((dir + "/cavity.msh").c_str(), "r")) {
            std::fclose(probe);
            prandtl::CellMesh cm = prandtl::ReadGmsh(dir + "/cavity.msh");
            prandtl::Mesh cav = prandtl::BuildMesh(cm, 3, {{"left", 0}, {"right", 0}, {"bottom", 0}, {"lid", 1}});
            cav.bc = {prandtl_bc{PRANDTL_BC_NOSLIP_ADIABATIC, {0.0, 0.0, 0.0}},
                      prandtl_bc{PRANDTL_BC_NOSLIP_ADIABATIC, {1.0, 0.0, 0.0}}};
            prandtl::Physics ns{1.4, 0.72, 0.02, 287.05, true, false};
            prandtl::Options co;
            co.blending = true;
            prandtl::DGSEMOperator cop(cav, ns, co);
            save(dir + "/cav_ng.bin", cav.node_geometry);   // the mesh as built, for the check
            {
                FILE *f = std::fopen((dir + "/cav_mesh.bin").c_str(), "wb");
                std::fwrite(cav.neighbors.data(), sizeof(int), cav.neighbors.size(), f);
                std::fwrite(cav.face_orientation.data(), sizeof(int), cav.face_orientation.size(), f);
                std::fclose(f);
            }
            const size_t cn = cav.StateSize();
            prandtl::Vector cu(cn), cd(cn);
            cu.CopyFromHost(load(dir + "/cav_u0.bin", cn).data());
            cop.Mult(cu, cd);
            save(dir + "/cav_dudt.bin", cd.HostCopy());
            prandtl::RK4Solver rk4;
            rk4.Init(cop);
            double ct = 0.0, ch = 1e-3;
            for (int s = 0; s < 3; s++) rk4.Step(cu, ct, ch);
            save(dir + "/cav_rk4.bin", cu.HostCopy());
            std::printf("cavity: %d elements, t %.15e\n", cav.num_elements, ct);
        }

        // errors surface as exceptions carrying the library's message
        try {
            prandtl::Vector tiny(4);
            op.Mult(tiny, tiny);
            std::printf("aliasing not rejected\n");
            return 1;
        } catch (const prandtl::Error &e) {
            std::printf("rejected as expected (%d): %s\n", e.code(), e.what());
        }
        // a vector of the wrong size never reaches the kernels
        try {
            prandtl::Vector short_dudt((size_t)op.Height() - 1);
            op.Mult(u, short_dudt);
            std::printf("short vector not rejected\n");
            return 1;
        } catch (const prandtl::Error &e) {
            std::printf("rejected as expected (%d): %s\n", e.code(), e.what());
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
