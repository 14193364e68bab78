// Result files in the reference's VTK legacy layout: the whole mesh, nodal U,
// the nodal stress tensor and the nodal damage phi -- writeReslutsTOTAL_vtk_Binary
// and writeReslutsTOTAL_vtk_ASCII (fioFile.cpp:562-769), byte for byte.
//   * binary: big-endian float / int32 payloads, a newline after each payload;
//   * ASCII: iostream scientific notation with 4 digits, the reference's line breaks.
// Cells are written for every element (PD and FE) with the element's own node count
// and VTK type (12 hexahedron, 10 tetrahedron, 9 quad, 5 triangle).  Also the PD-only and
// FE-only ASCII files and the node tables of an ASCII result (fioFile.cpp:486-538, :771-1001).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "model.hpp"

namespace pdx {

namespace {

int vtk_cell_type(int dim, int nen) {
    if (dim == 3) return nen == 8 ? 12 : 10;
    return nen == 4 ? 9 : 5;
}

void put_be32(std::string &out, uint32_t v) {
    const char b[4] = {char(v >> 24), char(v >> 16), char(v >> 8), char(v)};
    out.append(b, 4);
}

void put_be_float(std::string &out, float f) {
    uint32_t v;
    std::memcpy(&v, &f, 4);
    put_be32(out, v);
}

// operator<< of a stream in scientific notation with precision 4
void put_sci(std::string &out, double v) {
    char b[48];
    std::snprintf(b, sizeof b, "%.4e", v);
    out += b;
}

// one ASCII cell as the element classes print it (print_vtk): the node count and the 0-based
// node ids, each followed by a space; hexahedra, tetrahedra and triangles end the line, quads do
// not (pdfemEleQuad4N::print_vtk, pdfemEleQuad4N.cpp:201-215), so the quads of a 2D mesh share one
// line
void put_cell_ascii(std::string &out, int dim, int n, const int *ids) {
    out += std::to_string(n) + " ";
    for (int a = 0; a < n; a++) out += std::to_string(ids[a]) + " ";
    if (!(dim == 2 && n == 4)) out += "\n";
}

int write_file(const std::string &path, const std::string &out, std::string &err) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot open \"" + path + "\" for writing";
        return PDX_ERR_IO;
    }
    const size_t w = std::fwrite(out.data(), 1, out.size(), f);
    const int rc = std::fclose(f);
    if (w != out.size() || rc != 0) {
        err = "write to \"" + path + "\" failed";
        return PDX_ERR_IO;
    }
    return PDX_OK;
}

}  // namespace

// u: [nn*3] DOF values (UZ 0 in 2D), sigma: [nn*6] (xx, yy, zz, xy, yz, zx), phi: [nn];
// a null array writes zeros
int write_vtk(const Model &m, const std::string &path, const std::string &title, const double *u,
              const double *sigma, const double *phi, bool binary, std::string &err) {
    const int nn = m.nn, ne = m.ne;
    auto U = [&](int i, int j) { return u ? u[(size_t)i * 3 + j] : 0.0; };
    auto S = [&](int i, int j) { return sigma ? sigma[(size_t)i * 6 + j] : 0.0; };
    auto P = [&](int i) { return phi ? phi[i] : 0.0; };
    long long size_list = ne;
    for (int e = 0; e < ne; e++) size_list += m.nen[e];
    std::string out;
    out.reserve((size_t)nn * 80 + (size_t)size_list * 4 + 1024);
    out += "# vtk DataFile Version 3.0\n" + title + "\n" + (binary ? "BINARY\n" : "ASCII\n");
    out += "DATASET UNSTRUCTURED_GRID\nPOINTS " + std::to_string(nn) + " float\n";
    for (int i = 0; i < nn; i++)
        for (int j = 0; j < 3; j++) {
            if (binary) {
                put_be_float(out, float(m.x[(size_t)i * 3 + j]));
            } else {
                put_sci(out, m.x[(size_t)i * 3 + j]);
                out += j < 2 ? ' ' : '\n';
            }
        }
    out += "\nCELLS " + std::to_string(ne) + " " + std::to_string(size_list) + "\n";
    for (int e = 0; e < ne; e++) {
        const int n = m.nen[e];
        if (binary) {
            put_be32(out, uint32_t(n));
            for (int a = 0; a < n; a++) put_be32(out, uint32_t(m.econ[(size_t)e * 8 + a]));
        } else {
            put_cell_ascii(out, m.dim, n, &m.econ[(size_t)e * 8]);
        }
    }
    out += binary ? "\nCELL_TYPES " : "\nCELL_TYPES ";
    out += std::to_string(ne) + "\n";
    for (int e = 0; e < ne; e++) {
        const int t = vtk_cell_type(m.dim, m.nen[e]);
        if (binary) put_be32(out, uint32_t(t));
        else out += std::to_string(t) + "\n";
    }
    if (binary) out += "\n";
    out += "POINT_DATA " + std::to_string(nn) + "\nVECTORS U float\n";
    for (int i = 0; i < nn; i++) {
        for (int j = 0; j < 3; j++) {
            if (binary) {
                put_be_float(out, float(U(i, j)));
            } else {
                put_sci(out, float(U(i, j)));
                out += ' ';
            }
        }
        if (!binary) out += "\n";
    }
    if (binary) out += "\n";
    out += "TENSORS sigma float\n";
    static const int idx[9] = {0, 3, 5, 3, 1, 4, 5, 4, 2};  // rows of the symmetric tensor
    for (int i = 0; i < nn; i++) {
        for (int k = 0; k < 9; k++) {
            const float s = float(S(i, idx[k]));
            if (binary) {
                put_be_float(out, s);
            } else {
                put_sci(out, s);
                out += k % 3 == 2 ? '\n' : ' ';
            }
        }
        if (!binary) out += "\n";
    }
    if (binary) out += "\n";
    out += "SCALARS phi float 1\nLOOKUP_TABLE  default  \n";
    for (int i = 0; i < nn; i++) {
        if (binary) {
            put_be_float(out, float(P(i)));
        } else {
            put_sci(out, float(P(i)));
            out += ' ';
            if (i % 20 == 0) out += "\n";
        }
    }
    if (binary) out += "\n";
    return write_file(path, out, err);
}

// writeResultsPD_vtk / writeResultsFEM_vtk (fioFile.cpp:771-1001): ASCII files of one part of
// the mesh.  part 0: the PD nodes (type != 0, in node order) and the PD elements (algorithm 1),
// with U, sigma and phi; part 1: the nodes that are not pure PD (type != 2) and the FE elements
// (algorithm 2), with U and sigma.  Cells index the part's own point list.
int write_vtk_part(const Model &m, const std::string &path, const std::string &title, const double *u,
                   const double *sigma, const double *phi, int part, std::string &err) {
    if (part != 0 && part != 1) {
        err = "part must be 0 (PD) or 1 (FEM)";
        return PDX_ERR_ARG;
    }
    const int nn = m.nn, ne = m.ne;
    auto keep_node = [&](int i) { return part == 0 ? m.ntype[i] != 0 : m.ntype[i] != 2; };
    auto keep_ele = [&](int e) { return m.algo[e] == (part == 0 ? 1 : 2); };
    auto U = [&](int i, int j) { return u ? u[(size_t)i * 3 + j] : 0.0; };
    auto S = [&](int i, int j) { return sigma ? sigma[(size_t)i * 6 + j] : 0.0; };
    auto P = [&](int i) { return phi ? phi[i] : 0.0; };
    std::vector<int> map(nn, -1);
    int np = 0;
    for (int i = 0; i < nn; i++)
        if (keep_node(i)) map[i] = np++;
    std::string out;
    out += "# vtk DataFile Version 3.0\n" + title + "\nASCII\n";
    out += "DATASET UNSTRUCTURED_GRID\nPOINTS " + std::to_string(np) + " float\n";
    for (int i = 0; i < nn; i++) {
        if (!keep_node(i)) continue;
        for (int j = 0; j < 3; j++) {
            put_sci(out, m.x[(size_t)i * 3 + j]);
            out += j < 2 ? ' ' : '\n';
        }
    }
    long long size_list = 0;
    int nc = 0;
    for (int e = 0; e < ne; e++)
        if (keep_ele(e)) {
            size_list += m.nen[e];
            nc++;
        }
    size_list += nc;
    out += "\nCELLS " + std::to_string(nc) + " " + std::to_string(size_list) + "\n";
    for (int e = 0; e < ne; e++) {
        if (!keep_ele(e)) continue;
        const int n = m.nen[e];
        int ids[8];
        for (int a = 0; a < n; a++) ids[a] = map[m.econ[(size_t)e * 8 + a]];
        put_cell_ascii(out, m.dim, n, ids);
    }
    out += "\nCELL_TYPES " + std::to_string(nc) + "\n";
    for (int e = 0; e < ne; e++)
        if (keep_ele(e)) out += std::to_string(vtk_cell_type(m.dim, m.nen[e])) + "\n";
    out += "POINT_DATA " + std::to_string(np) + "\nVECTORS U float\n";
    for (int i = 0; i < nn; i++) {
        if (!keep_node(i)) continue;
        for (int j = 0; j < 3; j++) {
            put_sci(out, float(U(i, j)));
            out += ' ';
        }
        out += "\n";
    }
    out += "TENSORS sigma float\n";
    static const int idx[9] = {0, 3, 5, 3, 1, 4, 5, 4, 2};
    for (int i = 0; i < nn; i++) {
        if (!keep_node(i)) continue;
        for (int k = 0; k < 9; k++) {
            put_sci(out, float(S(i, idx[k])));
            out += k % 3 == 2 ? '\n' : ' ';
        }
        out += "\n";
    }
    if (part == 0) {
        out += "SCALARS phi float 1\nLOOKUP_TABLE  default  \n";
        int count = 0;
        for (int i = 0; i < nn; i++) {
            if (!keep_node(i)) continue;
            put_sci(out, float(P(i)));
            out += ' ';
            if (count % 20 == 0) out += "\n";
            count++;
        }
    }
    return write_file(path, out, err);
}

// writeUofNode / writeSigofNode (fioFile.cpp:486-538), the text tables the reference writes beside
// an ASCII result file: kind 0 ".disp" (node id, x, y, z, Ux, Uy, Uz; printFinalU, pdNode.cpp:120-130),
// kind 1 ".sig" (node id, x, y, z and the six stresses; printStress, pdNode.cpp:92-103)
int write_node_table(const Model &m, const std::string &path, const double *u, const double *sigma, int kind,
                     std::string &err) {
    if (kind != 0 && kind != 1) {
        err = "kind must be 0 (displacements) or 1 (stresses)";
        return PDX_ERR_ARG;
    }
    std::string out;
    char b[64];
    if (kind == 0) out += "*****The final displacement of  point*****\nID\tx\ty\tz\tUx\tUy\tUz\n";
    else out += "*****The stress of  point*****\nNID\tx\ty\tz\tSxx\tSyy\tSzz\tSxy\tSyz\tSzx\n";
    const char *fmt = kind == 0 ? "%.4e" : "%.5e";
    for (int i = 0; i < m.nn; i++) {
        out += std::to_string(i + 1);
        for (int j = 0; j < 3; j++) {
            std::snprintf(b, sizeof b, fmt, m.x[(size_t)i * 3 + j]);
            out += '\t';
            out += b;
        }
        for (int j = 0; j < (kind == 0 ? 3 : 6); j++) {
            const double v = kind == 0 ? (u ? u[(size_t)i * 3 + j] : 0.0) : (sigma ? sigma[(size_t)i * 6 + j] : 0.0);
            std::snprintf(b, sizeof b, fmt, v);
            out += '\t';
            out += b;
        }
        out += '\n';
    }
    return write_file(path, out, err);
}

}  // namespace pdx
