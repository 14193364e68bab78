// Result files in the reference's VTK legacy layout: the whole mesh, nodal U,
// the nodal stress tensor and the nodal damage phi -- writeReslutsTOTAL_vtk_Binary
// and writeReslutsTOTAL_vtk_ASCII (fioFile.cpp:562-769), byte for byte.
//   * binary: big-endian float / int32 payloads, a newline after each payload;
//   * ASCII: iostream scientific notation with 4 digits, the reference's line breaks.
// Cells are written for every element (PD and FE) with the element's own node count
// and VTK type (12 hexahedron, 10 tetrahedron, 9 quad, 5 triangle).
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
            out += std::to_string(n) + " ";
            for (int a = 0; a < n; a++) out += std::to_string(m.econ[(size_t)e * 8 + a]) + " ";
            out += "\n";
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

}  // namespace pdx
