// prandtl_amd_mesh.hpp -- header-only C++17 builder of general meshes for the C ABI: reads the
// reference's mesh files (Gmsh 2.2 ASCII, which the reference reads through MFEM at
// src/Simulation/Simulation.cpp:272, and MFEM mesh v1.0 files with periodic first-order L2 node
// grids), refines quadrilaterals uniformly (ser_ref_levels, Simulation.cpp:293-321), and fills a
// prandtl::Mesh with per-node metrics and face orientation codes (layout: prandtl_mesh_desc).
//
// Conventions (the same as prandtl_amd/mesh.py, which the tests compare against): element vertex
// order counter-clockwise from (0,0) in 2D, bottom face then top in 3D (MFEM's and Gmsh's);
// elements with det J < 0 in 2D are reflected by exchanging vertices 1 and 3 (MFEM's
// CheckElementOrientation); node coordinates are the bilinear / trilinear map at the GLL nodes;
// J[r][c] = dx_r/dxi_c by the GLL derivative matrix; faces matched by their vertex ids; each
// face's code is the one that sends every corner to the neighbour's corner with the same vertex.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "prandtl_amd.hpp"

namespace prandtl {

// A mesh as a file describes it: cells (2^dim vertex ids each, MFEM order), the cells' own corner
// coordinates (periodic meshes: vertex ids wrap, coordinates do not), boundary faces (their
// group: a Gmsh physical name, or an MFEM attribute as a string, and their vertex ids).
struct CellMesh {
    int dim = 0;
    std::vector<long long> cells;     // NE x 2^dim
    std::vector<double> corners;      // NE x 2^dim x dim
    std::vector<std::pair<std::string, std::vector<long long>>> bfaces;
    int num_cells() const { return dim ? (int)(cells.size() >> dim) : 0; }
};

namespace detail {

inline std::vector<std::string> tokens(const std::string &s) {
    std::istringstream is(s);
    std::vector<std::string> t;
    std::string w;
    while (is >> w) t.push_back(w);
    return t;
}

// GLL nodes and weights on [0,1] (Newton on P'_N in long double)
inline void gll(int np, std::vector<double> &x, std::vector<double> &w) {
    const int N = np - 1;
    x.assign(np, 0.0);
    w.assign(np, 0.0);
    auto leg = [N](long double z, long double &P, long double &dP) {
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
            leg(zz, P, dP);
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
            leg(z[j], P, dP);
        w[j] = (double)(1.0L / (N * (N + 1.0L) * P * P));
        x[j] = (double)(0.5L * (z[j] + 1.0L));
    }
    x[0] = 0.0;
    x[N] = 1.0;
}

// D[i][j] = l_j'(x_i) (barycentric; rows sum to zero)
inline std::vector<double> gll_derivative(const std::vector<double> &x) {
    const int np = (int)x.size();
    std::vector<double> c(np, 1.0), D((size_t)np * np, 0.0);
    for (int i = 0; i < np; i++)
        for (int j = 0; j < np; j++)
            if (j != i) c[i] *= x[i] - x[j];
    for (int i = 0; i < np; i++) {
        double s = 0.0;
        for (int j = 0; j < np; j++)
            if (j != i) {
                D[i * np + j] = c[i] / (c[j] * (x[i] - x[j]));
                s += D[i * np + j];
            }
        D[i * np + i] = -s;
    }
    return D;
}

// element vertex index (MFEM order) of the reference corner with coordinate bits b
inline int corner_index(int dim, const int *b) {
    static const int q[4] = {0, 3, 1, 2};   // (0,0) 0, (0,1) 3, (1,0) 1, (1,1) 2
    return q[2 * b[0] + b[1]] + (dim == 3 ? 4 * b[2] : 0);
}

// vertex index at corner (c1, c2) of local face f = 2 d + s: out[c2 * 2 + c1]
inline void face_corner_vertices(int dim, int f, int *out) {
    const int d = f >> 1, s = f & 1;
    int others[2], k = 0;
    for (int a = 0; a < dim; a++)
        if (a != d) others[k++] = a;
    for (int c2 = 0; c2 < (dim == 3 ? 2 : 1); c2++)
        for (int c1 = 0; c1 < 2; c1++) {
            int b[3] = {0, 0, 0};
            b[d] = s;
            b[others[0]] = c1;
            if (dim == 3) b[others[1]] = c2;
            out[c2 * 2 + c1] = corner_index(dim, b);
        }
}

}  // namespace detail

// Gmsh 2.2 ASCII: quadrilaterals (type 3) or hexahedra (type 5); boundary lines (type 1) or
// quadrilaterals (type 3, 3D) named by their physical group's name (or tag).
inline CellMesh ReadGmsh(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw Error(PRANDTL_ERR_ARG, "cannot read " + path);
    std::vector<std::string> lines;
    for (std::string ln; std::getline(in, ln);) lines.push_back(ln);
    std::map<long long, std::string> names;
    std::map<long long, size_t> ids;
    std::vector<std::array<double, 3>> verts;
    std::map<int, std::vector<std::pair<long long, std::vector<long long>>>> cells;
    for (size_t i = 0; i < lines.size(); i++) {
        const std::string key = detail::tokens(lines[i]).empty() ? "" : detail::tokens(lines[i])[0];
        if (key == "$MeshFormat") {
            const auto t = detail::tokens(lines[i + 1]);
            if (t.size() < 2 || t[0].rfind("2.", 0) != 0 || t[1] != "0")
                throw Error(PRANDTL_ERR_ARG, path + ": only the Gmsh 2.x ASCII format is read");
        } else if (key == "$PhysicalNames") {
            const int n = std::stoi(lines[i + 1]);
            for (int k = 0; k < n; k++) {
                std::istringstream is(lines[i + 2 + k]);
                long long d, tag;
                is >> d >> tag;
                std::string rest;
                std::getline(is, rest);
                const size_t a = rest.find('"'), b = rest.rfind('"');
                names[tag] = (a != std::string::npos && b > a) ? rest.substr(a + 1, b - a - 1) : rest;
            }
        } else if (key == "$Nodes") {
            const long long n = std::stoll(lines[i + 1]);
            for (long long k = 0; k < n; k++) {
                std::istringstream is(lines[i + 2 + k]);
                long long id;
                std::array<double, 3> x{};
                is >> id >> x[0] >> x[1] >> x[2];
                ids[id] = verts.size();
                verts.push_back(x);
            }
        } else if (key == "$Elements") {
            const long long n = std::stoll(lines[i + 1]);
            for (long long k = 0; k < n; k++) {
                const auto t = detail::tokens(lines[i + 2 + k]);
                const int type = std::stoi(t[1]), nt = std::stoi(t[2]);
                const long long tag = nt > 0 ? std::stoll(t[3]) : 0;
                std::vector<long long> v;
                for (size_t j = 3 + nt; j < t.size(); j++) v.push_back((long long)ids.at(std::stoll(t[j])));
                cells[type].push_back({tag, v});
            }
        }
    }
    CellMesh m;
    int vol, bnd;
    if (cells.count(5)) {
        m.dim = 3, vol = 5, bnd = 3;
    } else if (cells.count(3)) {
        m.dim = 2, vol = 3, bnd = 1;
    } else {
        throw Error(PRANDTL_ERR_ARG, path + ": no quadrilateral or hexahedral elements");
    }
    for (const auto &c : cells[vol]) {
        for (long long v : c.second) {
            m.cells.push_back(v);
            for (int d = 0; d < m.dim; d++) m.corners.push_back(verts[v][d]);
        }
    }
    for (const auto &c : cells[bnd]) {
        const std::string name = names.count(c.first) ? names[c.first] : std::to_string(c.first);
        m.bfaces.push_back({name, c.second});
    }
    return m;
}

// MFEM mesh v1.0 with quadrilaterals or hexahedra; coordinates from the vertices or from a
// first-order L2 node grid (periodic meshes); boundary groups are the attributes (as strings).
inline CellMesh ReadMfemMesh(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw Error(PRANDTL_ERR_ARG, "cannot read " + path);
    std::vector<std::string> lines;
    for (std::string ln; std::getline(in, ln);) {
        const size_t h = ln.find('#');
        if (h != std::string::npos) ln = ln.substr(0, h);
        const auto t = detail::tokens(ln);
        if (!t.empty()) {
            std::string j;
            for (size_t k = 0; k < t.size(); k++) j += (k ? " " : "") + t[k];
            lines.push_back(j);
        }
    }
    if (lines.empty() || lines[0].rfind("MFEM mesh v1", 0) != 0) throw Error(PRANDTL_ERR_ARG, path + ": not an MFEM mesh");
    CellMesh m;
    std::vector<std::vector<double>> verts;
    std::vector<double> nodes;
    int vdim = 0, ordering = 0;
    bool have_nodes = false;
    for (size_t i = 1; i < lines.size();) {
        const std::string &k = lines[i];
        if (k == "dimension") {
            m.dim = std::stoi(lines[i + 1]);
            i += 2;
        } else if (k == "elements") {
            const int n = std::stoi(lines[i + 1]);
            for (int e = 0; e < n; e++) {
                const auto t = detail::tokens(lines[i + 2 + e]);
                const int geom = std::stoi(t[1]);
                if (geom != 3 && geom != 5) throw Error(PRANDTL_ERR_ARG, path + ": only quadrilaterals and hexahedra");
                for (size_t j = 2; j < t.size(); j++) m.cells.push_back(std::stoll(t[j]));
            }
            i += 2 + n;
        } else if (k == "boundary") {
            const int n = std::stoi(lines[i + 1]);
            for (int e = 0; e < n; e++) {
                const auto t = detail::tokens(lines[i + 2 + e]);
                std::vector<long long> v;
                for (size_t j = 2; j < t.size(); j++) v.push_back(std::stoll(t[j]));
                m.bfaces.push_back({t[0], v});
            }
            i += 2 + n;
        } else if (k == "vertices") {
            const int n = std::stoi(lines[i + 1]);
            if (i + 2 < lines.size() && lines[i + 2] == "nodes") {
                i += 2;
                continue;
            }
            const int vd = std::stoi(lines[i + 2]);
            for (int v = 0; v < n; v++) {
                const auto t = detail::tokens(lines[i + 3 + v]);
                std::vector<double> x;
                for (int d = 0; d < vd; d++) x.push_back(std::stod(t[d]));
                verts.push_back(x);
            }
            i += 3 + n;
        } else if (k == "nodes") {
            size_t j = i + 1;
            std::string fec;
            while (j < lines.size() && (lines[j] == "FiniteElementSpace" || lines[j].find(':') != std::string::npos)) {
                const size_t c = lines[j].find(':');
                if (c != std::string::npos) {
                    const std::string a = lines[j].substr(0, c), b = lines[j].substr(c + 2);
                    if (a == "FiniteElementCollection") fec = b;
                    if (a == "VDim") vdim = std::stoi(b);
                    if (a == "Ordering") ordering = std::stoi(b);
                }
                j++;
            }
            if (fec.rfind("L2_T", 0) != 0 || fec.size() < 3 || fec.substr(fec.size() - 3) != "_P1")
                throw Error(PRANDTL_ERR_ARG, path + ": only first-order L2 node grids are read");
            for (; j < lines.size(); j++)
                for (const auto &t : detail::tokens(lines[j])) nodes.push_back(std::stod(t));
            have_nodes = true;
            i = lines.size();
        } else {
            i++;
        }
    }
    const int nc = 1 << m.dim, NE = (int)(m.cells.size() / nc);
    m.corners.assign((size_t)NE * nc * m.dim, 0.0);
    const int lex2c2[4] = {0, 1, 3, 2}, lex2c3[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    for (int e = 0; e < NE; e++)
        for (int L = 0; L < nc; L++) {
            const int c = m.dim == 2 ? lex2c2[L] : lex2c3[L];
            for (int d = 0; d < m.dim; d++) {
                double x;
                if (have_nodes) {   // node L of the element's lexicographic P1 grid sits at corner c
                    const size_t node = (size_t)e * nc + L, total = (size_t)NE * nc;
                    x = ordering == 1 ? nodes[node * vdim + d] : nodes[d * total + node];
                } else {
                    x = verts[m.cells[(size_t)e * nc + c]][d];
                }
                m.corners[((size_t)e * nc + c) * m.dim + d] = x;
            }
        }
    return m;
}

// Uniform refinement of quadrilaterals (MFEM's UniformRefinement): each cell into four, new
// vertices at the edge midpoints (shared through the vertex ids) and the centre; the sub-cells'
// corners are the bilinear images of the quarter squares; boundary edges split in two.
inline void RefineQuads(CellMesh &m, int levels) {
    if (m.dim != 2) throw Error(PRANDTL_ERR_UNSUPPORTED, "uniform refinement is implemented for quadrilaterals");
    long long nv = 0;
    for (long long v : m.cells) nv = std::max(nv, v + 1);
    for (const auto &b : m.bfaces)
        for (long long v : b.second) nv = std::max(nv, v + 1);
    for (int l = 0; l < levels; l++) {
        std::map<std::pair<long long, long long>, long long> mids;
        auto mid = [&](long long a, long long b) {
            const auto k = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
            auto it = mids.find(k);
            if (it != mids.end()) return it->second;
            mids[k] = nv;
            return nv++;
        };
        const int NE = m.num_cells();
        std::vector<long long> cells;
        std::vector<double> corners;
        for (int e = 0; e < NE; e++) {
            const long long *v = &m.cells[(size_t)e * 4];
            const double *C = &m.corners[(size_t)e * 8];
            auto P = [&](int i, int j, int d) {
                const double x = 0.5 * i, y = 0.5 * j;
                return ((1 - x) * (1 - y)) * C[0 + d] + (x * (1 - y)) * C[2 + d] + (x * y) * C[4 + d] + ((1 - x) * y) * C[6 + d];
            };
            const long long m01 = mid(v[0], v[1]), m12 = mid(v[1], v[2]), m23 = mid(v[2], v[3]), m30 = mid(v[3], v[0]);
            const long long c = nv++;
            const long long ids[3][3] = {{v[0], m30, v[3]}, {m01, c, m23}, {v[1], m12, v[2]}};   // [i][j]
            const int q0[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            for (int q = 0; q < 4; q++) {
                const int i0 = q0[q][0], j0 = q0[q][1];
                const int quad[4][2] = {{i0, j0}, {i0 + 1, j0}, {i0 + 1, j0 + 1}, {i0, j0 + 1}};
                for (int k = 0; k < 4; k++) {
                    cells.push_back(ids[quad[k][0]][quad[k][1]]);
                    for (int d = 0; d < 2; d++) corners.push_back(P(quad[k][0], quad[k][1], d));
                }
            }
        }
        std::vector<std::pair<std::string, std::vector<long long>>> bf;
        for (const auto &b : m.bfaces) {
            const long long a = b.second[0], c = b.second[1], md = mid(a, c);
            bf.push_back({b.first, {a, md}});
            bf.push_back({b.first, {md, c}});
        }
        m.cells.swap(cells);
        m.corners.swap(corners);
        m.bfaces.swap(bf);
    }
}

// The general mesh of a CellMesh at polynomial order `order`: boundary faces take the condition
// index bc_index[group]; conditions are the caller's (mesh.bc).
inline Mesh BuildMesh(const CellMesh &cm, int order, const std::map<std::string, int> &bc_index) {
    const int dim = cm.dim, np = order + 1, nc = 1 << dim, NE = cm.num_cells(), nf = 2 * dim, G = 1 + dim * dim;
    int n = 1;
    for (int d = 0; d < dim; d++) n *= np;
    std::vector<long long> cells(cm.cells);
    std::vector<double> corners(cm.corners);
    if (dim == 2)
        for (int e = 0; e < NE; e++) {   // reflect clockwise cells (MFEM exchanges vertices 1 and 3)
            const double *C = &corners[(size_t)e * 8];
            const double ax = C[2] - C[0], ay = C[3] - C[1], bx = C[6] - C[0], by = C[7] - C[1];
            if (ax * by - ay * bx < 0) {
                std::swap(cells[(size_t)e * 4 + 1], cells[(size_t)e * 4 + 3]);
                for (int d = 0; d < 2; d++) std::swap(corners[(size_t)e * 8 + 2 + d], corners[(size_t)e * 8 + 6 + d]);
            }
        }
    std::vector<double> xg, wg;
    detail::gll(np, xg, wg);
    const std::vector<double> D = detail::gll_derivative(xg);
    // node coordinates X[d][e][node] from the multilinear map
    std::vector<double> X((size_t)dim * NE * n);
    for (int i = 0; i < n; i++) {
        const int a[3] = {i % np, (i / np) % np, i / (np * np)};
        double N[8];
        const double x = xg[a[0]], y = xg[a[1]], z = dim == 3 ? xg[a[2]] : 0.0;
        if (dim == 2) {
            N[0] = (1 - x) * (1 - y), N[1] = x * (1 - y), N[2] = x * y, N[3] = (1 - x) * y;
        } else {
            N[0] = (1 - x) * (1 - y) * (1 - z), N[1] = x * (1 - y) * (1 - z), N[2] = x * y * (1 - z);
            N[3] = (1 - x) * y * (1 - z), N[4] = (1 - x) * (1 - y) * z, N[5] = x * (1 - y) * z, N[6] = x * y * z;
            N[7] = (1 - x) * y * z;
        }
        for (int e = 0; e < NE; e++)
            for (int d = 0; d < dim; d++) {
                double s = 0.0;
                for (int v = 0; v < nc; v++) s += N[v] * corners[((size_t)e * nc + v) * dim + d];
                X[((size_t)d * NE + e) * n + i] = s;
            }
    }
    Mesh m;
    m.dim = dim;
    m.order = order;
    m.num_elements = NE;
    m.node_geometry.assign((size_t)NE * n * G, 0.0);
    const int stride[3] = {1, np, np * np};
    for (int e = 0; e < NE; e++)
        for (int i = 0; i < n; i++) {
            const int a[3] = {i % np, (i / np) % np, i / (np * np)};
            double J[3][3];
            for (int r = 0; r < dim; r++)
                for (int c = 0; c < dim; c++) {
                    double s = 0.0;
                    const int base = i - a[c] * stride[c];
                    for (int b = 0; b < np; b++) s += D[a[c] * np + b] * X[((size_t)r * NE + e) * n + base + b * stride[c]];
                    J[r][c] = s;
                }
            double *g = &m.node_geometry[((size_t)e * n + i) * G];
            if (dim == 2) {
                g[0] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                g[1] = J[1][1], g[2] = -J[0][1], g[3] = -J[1][0], g[4] = J[0][0];
            } else {
                double adj[3][3];
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) {
                        const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                        adj[r][c] = J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1];
                    }
                g[0] = J[0][0] * adj[0][0] + J[0][1] * adj[1][0] + J[0][2] * adj[2][0];
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) g[1 + r * 3 + c] = adj[r][c];
            }
            if (!(g[0] > 0.0)) throw Error(PRANDTL_ERR_MESH, "element " + std::to_string(e) + ": det J <= 0 at a node");
        }
    // faces by vertex ids
    std::map<std::vector<long long>, std::vector<std::pair<int, int>>> table;
    int fcv[6][4];
    for (int f = 0; f < nf; f++) detail::face_corner_vertices(dim, f, fcv[f]);
    const int ncorner = dim == 3 ? 4 : 2;
    auto key_of = [&](int e, int f) {
        std::vector<long long> k(ncorner);
        for (int c = 0; c < ncorner; c++) k[c] = cells[(size_t)e * nc + fcv[f][c]];
        std::sort(k.begin(), k.end());
        return k;
    };
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < nf; f++) table[key_of(e, f)].push_back({e, f});
    std::map<std::vector<long long>, std::string> bgroup;
    for (const auto &b : cm.bfaces) {
        std::vector<long long> k(b.second);
        std::sort(k.begin(), k.end());
        bgroup[k] = b.first;
    }
    m.neighbors.assign((size_t)NE * nf, 0);
    m.face_orientation.assign((size_t)NE * nf, 0);
    for (const auto &kv : table) {
        const auto &lst = kv.second;
        if (lst.size() > 2) throw Error(PRANDTL_ERR_MESH, "more than two faces share their vertices");
        if (lst.size() == 1) {
            const int e = lst[0].first, f = lst[0].second;
            auto g = bgroup.find(kv.first);
            if (g == bgroup.end()) throw Error(PRANDTL_ERR_MESH, "a boundary face has no boundary element");
            auto k = bc_index.find(g->second);
            if (k == bc_index.end()) throw Error(PRANDTL_ERR_MESH, "no boundary condition for the group " + g->second);
            m.neighbors[(size_t)e * nf + f] = -(k->second + 1);
            m.face_orientation[(size_t)e * nf + f] = f ^ 1;
            continue;
        }
        for (int side = 0; side < 2; side++) {
            const int e = lst[side].first, f = lst[side].second, e2 = lst[1 - side].first, f2 = lst[1 - side].second;
            int found = -1;
            for (int code = 0; code < 64 && found < 0; code += 8) {
                if (dim == 2 && code > 8) break;
                bool ok = true;
                for (int c2 = 0; c2 < (dim == 3 ? 2 : 1); c2++)
                    for (int c1 = 0; c1 < 2; c1++) {
                        const int a = (code & 32) ? c2 : c1, b = (code & 32) ? c1 : c2;
                        const int u1 = (code & 8) ? 1 - a : a;
                        const int u2 = dim == 3 ? ((code & 16) ? 1 - b : b) : 0;
                        ok = ok && cells[(size_t)e2 * nc + fcv[f2][u2 * 2 + u1]] == cells[(size_t)e * nc + fcv[f][c2 * 2 + c1]];
                    }
                if (ok) found = code;
            }
            if (found < 0) throw Error(PRANDTL_ERR_MESH, "faces share vertices but not their corners' order");
            m.neighbors[(size_t)e * nf + f] = e2;
            m.face_orientation[(size_t)e * nf + f] = f2 | found;
        }
    }
    m.geometry.assign((size_t)NE * G, 0.0);
    for (int e = 0; e < NE; e++)
        std::copy(&m.node_geometry[(size_t)e * n * G], &m.node_geometry[(size_t)e * n * G] + G, &m.geometry[(size_t)e * G]);
    return m;
}

}  // namespace prandtl
