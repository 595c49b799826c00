// Builds a general mesh with the C++ mirror (include/prandtl_amd_mesh.hpp) and writes its arrays
// for comparison with prandtl_amd/mesh.py:  mesh_builder_example <gmsh|mfem> <file> <order>
// <refine> <out_prefix> [group=k ...]
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

#include "prandtl_amd_mesh.hpp"

int main(int argc, char **argv) {
    if (argc < 6) return 2;
    const std::string kind = argv[1], file = argv[2], out = argv[5];
    const int order = std::atoi(argv[3]), refine = std::atoi(argv[4]);
    std::map<std::string, int> bc;
    for (int i = 6; i < argc; i++) {
        const std::string a = argv[i];
        const size_t eq = a.find('=');
        bc[a.substr(0, eq)] = std::atoi(a.c_str() + eq + 1);
    }
    try {
        prandtl::CellMesh cm = kind == "gmsh" ? prandtl::ReadGmsh(file) : prandtl::ReadMfemMesh(file);
        if (refine) prandtl::RefineQuads(cm, refine);
        const prandtl::Mesh m = prandtl::BuildMesh(cm, order, bc);
        FILE *f = std::fopen((out + ".nbr").c_str(), "wb");
        std::fwrite(m.neighbors.data(), sizeof(int), m.neighbors.size(), f);
        std::fclose(f);
        f = std::fopen((out + ".fo").c_str(), "wb");
        std::fwrite(m.face_orientation.data(), sizeof(int), m.face_orientation.size(), f);
        std::fclose(f);
        f = std::fopen((out + ".ng").c_str(), "wb");
        std::fwrite(m.node_geometry.data(), sizeof(double), m.node_geometry.size(), f);
        std::fclose(f);
        std::printf("%d %d\n", m.num_elements, m.dim);
    } catch (const prandtl::Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
