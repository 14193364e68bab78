// Host set-up of the PDLSM-FEM data model.  Follows the reference's set-up
// sequence pdsolve::setDatModel (SourceFiles/Solver/pdsolve.cpp:114-151) with
// the same rules and the same orderings (family member order, boundary entity
// orientation, CSR column order), restated with cell lists built by counting
// sort and OpenMP-parallel family / pattern loops.
#include "model.hpp"
#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <sstream>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace pdx {

double det33(const double m[3][3]) {
    // pdfemEles::detMat33 (pdfemEles.cpp:155-172): cofactor form in this product order
    double det = 0.0;
    for (int i = 0; i < 3; i++) {
        double a1 = 1., a2 = 1.;
        for (int j = 0; j < 3; j++) {
            a1 *= m[j][(i + j) % 3];
            a2 *= m[j][(i - j + 3) % 3];
        }
        det = det + a1 - a2;
    }
    return det;
}

namespace {

const double kGP = 0.57735026918962573;  // 1/sqrt(3), pdGaussPt.cpp:31-38
const double kHP[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
const double kHQ[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
const double kHR[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};

struct EleGeom {
    double xN[8][3];
};

EleGeom gather(const Model &m, int e) {
    EleGeom g{};
    for (int a = 0; a < m.nen[e]; a++)
        for (int c = 0; c < 3; c++) g.xN[a][c] = m.x[size_t(m.econ[e * 8 + a]) * 3 + c];
    return g;
}

double hex_detJ(const double xN[8][3], double p, double q, double r) {
    // pdfemEleBrick8N::JacobiMat + detJacobi (pdfemEleBrick8N.cpp:93-117, :324-348)
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double dN[8][3];
    for (int i = 0; i < 8; i++) {
        dN[i][0] = 1.0 / 8 * kHP[i] * (1 + q * kHQ[i]) * (1 + r * kHR[i]);
        dN[i][1] = 1.0 / 8 * kHQ[i] * (1 + p * kHP[i]) * (1 + r * kHR[i]);
        dN[i][2] = 1.0 / 8 * kHR[i] * (1 + p * kHP[i]) * (1 + q * kHQ[i]);
    }
    for (int cc = 0; cc < 3; cc++)
        for (int i = 0; i < 8; i++)
            for (int rr = 0; rr < 3; rr++) J[rr][cc] = dN[i][rr] * xN[i][cc] + J[rr][cc];
    return det33(J);
}

double element_volume(const Model &m, int e) {
    EleGeom g = gather(m, e);
    auto &xN = g.xN;
    int n = m.nen[e];
    if (m.dim == 2) {
        // quad4 (pdfemEleQuad4N.cpp:212-237) = two triangles; tri3 (pdfemEleTri3N.cpp:176-183)
        double t[3][3] = {{1, xN[0][0], xN[0][1]}, {1, xN[1][0], xN[1][1]}, {1, xN[2][0], xN[2][1]}};
        double v = 0.5 * std::fabs(det33(t));
        if (n == 3) return v;
        double t2[3][3] = {{1, xN[2][0], xN[2][1]}, {1, xN[3][0], xN[3][1]}, {1, xN[0][0], xN[0][1]}};
        return v + 0.5 * std::fabs(det33(t2));
    }
    if (n == 8) {
        // pdfemEleBrick8N::eleVolume (pdfemEleBrick8N.cpp:273-296)
        const double gp[2] = {-kGP, kGP};
        double v = 0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++) v = v + 1.0 * hex_detJ(xN, gp[a], gp[b], gp[c]);
        return v;
    }
    // pdfemEleTetra4N::eleVolume (pdfemEleTetra4N.cpp:210-217)
    double t[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) t[r][c] = xN[r + 1][c] - xN[0][c];
    return std::fabs(det33(t)) / 6.0;
}

void element_center(const Model &m, int e, double xc[3]) {
    EleGeom g = gather(m, e);
    int n = m.nen[e];
    xc[0] = xc[1] = xc[2] = 0;
    if (m.dim == 2 && n == 3) {
        for (int i = 0; i < 3; i++) xc[i] = (g.xN[0][i] + g.xN[1][i] + g.xN[2][i]) / 3.0;
        return;
    }
    const double w = (m.dim == 3 && n == 8) ? 1.0 / 8 : 0.25;
    for (int a = 0; a < n; a++)
        for (int i = 0; i < 3; i++) xc[i] = xc[i] + g.xN[a][i] * w;
}

bool seg_seg(const double *a1, const double *a2, const double *b1, const double *b2) {
    // pdsolve::intersection (pdsolve.cpp:9058-9079)
    auto mx = [](double u, double v) { return u > v ? u : v; };
    auto mn = [](double u, double v) { return u < v ? u : v; };
    if (mx(a1[0], a2[0]) < mn(b1[0], b2[0]) || mx(a1[1], a2[1]) < mn(b1[1], b2[1]) ||
        mx(b1[0], b2[0]) < mn(a1[0], a2[0]) || mx(b1[1], b2[1]) < mn(a1[1], a2[1]))
        return false;
    double s1 = ((a1[0] - b1[0]) * (b2[1] - b1[1]) - (a1[1] - b1[1]) * (b2[0] - b1[0])) *
                ((a2[0] - b1[0]) * (b2[1] - b1[1]) - (a2[1] - b1[1]) * (b2[0] - b1[0]));
    double s2 = ((b1[0] - a1[0]) * (a2[1] - a1[1]) - (b1[1] - a1[1]) * (a2[0] - a1[0])) *
                ((b2[0] - a1[0]) * (a2[1] - a1[1]) - (b2[1] - a1[1]) * (a2[0] - a1[0]));
    return !(s1 > 0 || s2 > 0);
}

void cross(const double *u, const double *v, double *o) {
    o[0] = u[1] * v[2] - u[2] * v[1];
    o[1] = u[2] * v[0] - u[0] * v[2];
    o[2] = u[0] * v[1] - u[1] * v[0];
}

bool seg_tri_plane(const double *p1, const double *p2, const double *T) {
    // pdsolve::segmentPlaneIntersection (pdsolve.cpp:1123-1209): side test against the
    // triangle's plane, then an area test of the crossing point against the triangle
    double v1[3], v2[3], n[3];
    for (int i = 0; i < 3; i++) { v1[i] = T[3 + i] - T[i]; v2[i] = T[6 + i] - T[i]; }
    cross(v1, v2, n);
    for (int i = 0; i < 3; i++) { v1[i] = p1[i] - T[i]; v2[i] = p2[i] - T[i]; }
    double t1 = v1[0] * n[0] + v1[1] * n[1] + v1[2] * n[2];
    double t2 = v2[0] * n[0] + v2[1] * n[1] + v2[2] * n[2];
    if (t1 * t2 > 0) return false;
    if (t1 * t2 == 0) return true;
    double a[3], b[3];
    for (int i = 0; i < 3; i++) { a[i] = T[i] - p1[i]; b[i] = p2[i] - p1[i]; }
    double d = (a[0] * n[0] + a[1] * n[1] + a[2] * n[2]) / (b[0] * n[0] + b[1] * n[1] + b[2] * n[2]);
    double X[3];
    for (int i = 0; i < 3; i++) X[i] = p1[i] + d * b[i];
    double A = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    auto area2 = [&](const double *u, const double *v) {
        double c[3];
        cross(u, v, c);
        return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    };
    for (int i = 0; i < 3; i++) { v1[i] = T[3 + i] - X[i]; v2[i] = T[6 + i] - X[i]; }
    double A0 = area2(v1, v2);
    for (int i = 0; i < 3; i++) { v1[i] = X[i] - T[i]; v2[i] = T[6 + i] - T[i]; }
    double A1 = area2(v1, v2);
    for (int i = 0; i < 3; i++) { v1[i] = T[3 + i] - T[i]; v2[i] = X[i] - T[i]; }
    double A2 = area2(v1, v2);
    return !(std::fabs(A0 + A1 + A2 - A) > A * 1.0e-14);
}

bool element_cut(const Model &m, int e, const double *crack) {
    // pdsolve::b_eleIntersectCrack (pdsolve.cpp:839-930)
    const int *c = &m.econ[size_t(e) * 8];
    int n = m.nen[e];
    auto X = [&](int a) { return &m.x[size_t(c[a]) * 3]; };
    if (m.dim == 2) {
        for (int i = 0; i < n; i++)
            if (seg_seg(X(i), X((i + 1) % n), crack, crack + 3)) return true;
        return false;
    }
    static const int hexE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                    {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    static const int tetE[6][2] = {{3, 0}, {3, 1}, {3, 2}, {0, 1}, {1, 2}, {2, 0}};
    if (n == 8) {
        for (auto &ed : hexE)
            if (seg_tri_plane(X(ed[0]), X(ed[1]), crack)) return true;
    } else {
        for (auto &ed : tetE)
            if (seg_tri_plane(X(ed[0]), X(ed[1]), crack)) return true;
    }
    return false;
}

template <class F>
void for_cells(const Model &m, const int lo[3], const int hi[3], F &&f) {
    // the reference's x-outer / y / z-inner scan over a block range, clipped to the grid
    for (int xi = lo[0]; xi <= hi[0]; xi++)
        for (int yi = lo[1]; yi <= hi[1]; yi++)
            for (int zi = lo[2]; zi <= hi[2]; zi++)
                if (xi >= 0 && xi < m.nb[0] && yi >= 0 && yi < m.nb[1] && zi >= 0 && zi < m.nb[2])
                    f(xi + yi * m.nb[0] + zi * m.nb[0] * m.nb[1]);
}

void crack_box(const Model &m, const double *crack, int lo[3], int hi[3]) {
    int fn = m.dim == 3 ? 3 : 2;
    for (int k = 0; k < 3; k++) {
        double a = crack[k], b = crack[k];
        for (int p = 1; p < fn; p++) { a = std::min(a, crack[p * 3 + k]); b = std::max(b, crack[p * 3 + k]); }
        lo[k] = int((a - m.lbc[k]) / m.bsize) - 1;
        hi[k] = int((b - m.lbc[k]) / m.bsize) + 1;
    }
    if (m.dim == 2) lo[2] = hi[2] = 0;
}

void convert_by_crack(Model &m, const double *crack) {
    // pdsolve::updatePDEbyCrack (pdsolve.cpp:533-727)
    int lo[3], hi[3];
    crack_box(m, crack, lo, hi);
    if (m.dim == 2) { lo[2] = -1; hi[2] = 1; }
    size_t old = m.crac_ele.size();
    for_cells(m, lo, hi, [&](int b) {
        for (int t = m.cell_ele_off[b]; t < m.cell_ele_off[b + 1]; t++) {
            int e = m.cell_ele[t];
            if (element_cut(m, e, crack) && std::find(m.crac_ele.begin(), m.crac_ele.end(), e) == m.crac_ele.end())
                m.crac_ele.push_back(e);
        }
    });
    size_t sz = m.crac_ele.size() - old;
    if (sz == 0) return;
    std::vector<double> C(sz * 3);
    double mn[3], mx[3];
    for (size_t i = 0; i < sz; i++) {
        element_center(m, m.crac_ele[old + i], &C[i * 3]);
        for (int k = 0; k < 3; k++) {
            mn[k] = i ? std::min(mn[k], C[i * 3 + k]) : C[i * 3 + k];
            mx[k] = i ? std::max(mx[k], C[i * 3 + k]) : C[i * 3 + k];
        }
    }
    for (int k = 0; k < 3; k++) {
        lo[k] = int((mn[k] - m.lbc[k]) / m.bsize) - 1;
        hi[k] = int((mx[k] - m.lbc[k]) / m.bsize) + 1;
    }
    double f = m.me + 1e-15;
    double R2 = m.minDelta * m.minDelta * f * f;
    for_cells(m, lo, hi, [&](int b) {
        for (int t = m.cell_ele_off[b]; t < m.cell_ele_off[b + 1]; t++) {
            int e = m.cell_ele[t];
            if (m.algo[e] != 2) continue;
            double xc[3];
            element_center(m, e, xc);
            for (size_t i = 0; i < sz; i++) {
                double d = (xc[0] - C[i * 3]) * (xc[0] - C[i * 3]) + (xc[1] - C[i * 3 + 1]) * (xc[1] - C[i * 3 + 1]) +
                           (xc[2] - C[i * 3 + 2]) * (xc[2] - C[i * 3 + 2]);
                if (d < R2) {
                    m.algo[e] = 1;
                    m.pde.push_back(e);
                    break;
                }
            }
        }
    });
}

struct Ent {
    int v[4];
    int flag;
};

void boundary_entities(Model &m) {
    // pdsolve::updatePDBE (pdsolve.cpp:1221-1413) + assignPDBE (:1703-1771)
    std::vector<Ent> ents;
    ents.reserve(m.pde.size() * 6);
    for (int e : m.pde) {
        const int *c = &m.econ[size_t(e) * 8];  // 0-based here; ordering compares are id-monotone
        int n = m.nen[e];
        if (m.dim == 2) {
            for (int j = 0; j < n; j++) {
                int a = c[j], b = c[(j + 1) % n];
                ents.push_back(a < b ? Ent{{a, b, 0, 0}, 1} : Ent{{b, a, 0, 0}, -1});
            }
        } else if (n == 8) {
            const int f[6][4] = {{c[0], c[1], c[2], c[3]}, {c[4], c[5], c[6], c[7]}, {c[0], c[3], c[7], c[4]},
                                 {c[1], c[2], c[6], c[5]}, {c[0], c[4], c[5], c[1]}, {c[3], c[7], c[6], c[2]}};
            const int sgn[6] = {-1, 1, -1, 1, -1, 1};
            for (int j = 0; j < 6; j++) {
                int p = int(std::min_element(f[j], f[j] + 4) - f[j]);
                Ent t{{f[j][p], f[j][(p + 1) % 4], f[j][(p + 2) % 4], f[j][(p + 3) % 4]}, sgn[j]};
                ents.push_back(t);
            }
        } else {
            const int f[4][3] = {{c[0], c[1], c[3]}, {c[1], c[2], c[3]}, {c[2], c[0], c[3]}, {c[0], c[2], c[1]}};
            for (int j = 0; j < 4; j++) {
                int p = int(std::min_element(f[j], f[j] + 3) - f[j]);
                Ent t{{f[j][p], f[j][(p + 1) % 3], f[j][(p + 2) % 3], f[j][(p + 2) % 3]}, 0};
                ents.push_back(t);
            }
        }
    }
    std::stable_sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) { return a.v[0] < b.v[0]; });
    std::vector<char> dead(ents.size(), 0);
    for (size_t i = 0; i < ents.size(); i++) {
        if (dead[i]) continue;
        for (size_t j = i + 1; j < ents.size() && ents[j].v[0] == ents[i].v[0]; j++) {
            if (dead[j]) continue;
            const int *a = ents[i].v, *b = ents[j].v;
            bool hit;
            if (m.dim == 2) hit = a[1] == b[1];
            else if (a[2] != a[3]) hit = a[1] == b[1] && a[2] == b[2] && a[3] == b[3];
            else hit = b[2] == b[3] && a[1] == b[2] && a[2] == b[1];
            if (hit) { dead[i] = dead[j] = 1; break; }
        }
    }
    m.pdbe.clear();
    for (size_t i = 0; i < ents.size(); i++) {
        if (dead[i]) continue;
        const Ent &t = ents[i];
        int o[4];
        if (m.dim == 2) {
            o[0] = t.flag == 1 ? t.v[0] : t.v[1];
            o[1] = t.flag == 1 ? t.v[1] : t.v[0];
            o[2] = o[3] = -1;
        } else if (t.v[2] == t.v[3]) {
            o[0] = t.v[0]; o[1] = t.v[1]; o[2] = t.v[2]; o[3] = t.v[2];
        } else if (t.flag == 1) {
            for (int q = 0; q < 4; q++) o[q] = t.v[q];
        } else {
            for (int q = 0; q < 4; q++) o[3 - q] = t.v[q];
        }
        m.pdbe.insert(m.pdbe.end(), o, o + 4);
    }
    m.npdbe = int(m.pdbe.size() / 4);
}

}  // namespace

// updateBondstate (pdsolve.cpp:8910-9004): every intact bond of a family whose node lies in the
// cells of the crack's bounding box (one cell of margin) and whose segment crosses the crack (2D:
// the segment crack[0..1]; 3D: the triangle crack[0..2]) is broken; returns how many
int break_bonds_across(const Model &m, const double *crack, std::vector<int> &bond) {
    int lo[3], hi[3], nb = 0;
    crack_box(m, crack, lo, hi);
    // the families of the nodes in the cells around the crack; each bond is tested on its own,
    // so the families are independent
    std::vector<int> fams;
    for_cells(m, lo, hi, [&](int b) {
        for (int t = m.cell_node_off[b]; t < m.cell_node_off[b + 1]; t++) {
            const int f = m.famid[m.cell_node[t]];
            if (f >= 0) fams.push_back(f);
        }
    });
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nb)
    for (size_t i = 0; i < fams.size(); i++) {
        const int f = fams[i];
        const double *xk = &m.x[size_t(m.fam_node[f]) * 3];
        for (int q = m.fam_off[f] + 1; q < m.fam_off[f + 1]; q++) {
            if (bond[q] != 1) continue;
            const double *xm = &m.x[size_t(m.fam_mem[q]) * 3];
            const bool cut = m.dim == 3 ? seg_tri_plane(xk, xm, crack) : seg_seg(xk, xm, crack, crack + 3);
            if (cut) {
                bond[q] = 0;
                nb++;
            }
        }
    }
    return nb;
}

// initialPDN_Fami's search on the host: per family its node, then the nodes of the 3 x 3 x 3 cells
// around the node's cell (for_cells order, each cell in node order) within either horizon
static void host_families(Model &m, const std::vector<double> &hz, double tiny) {
    const int nfam = m.nfam;
    std::vector<std::vector<int>> mem(nfam);
#pragma omp parallel for schedule(dynamic, 64)
    for (int f = 0; f < nfam; f++) {
        int k = m.fam_node[f];
        const double *cx = &m.x[size_t(k) * 3];
        double dk = m.fam_delta[f];
        std::vector<int> &out = mem[f];
        out.push_back(k);
        int lo[3], hi[3];
        for (int i = 0; i < 3; i++) {
            int c = int((cx[i] - m.lbc[i]) / m.bsize);
            lo[i] = c - 1;
            hi[i] = c + 1;
        }
        for_cells(m, lo, hi, [&](int b) {
            for (int t = m.cell_node_off[b]; t < m.cell_node_off[b + 1]; t++) {
                int j = m.cell_node[t];
                const double *mx = &m.x[size_t(j) * 3];
                double dist = ((mx[0] - cx[0]) * (mx[0] - cx[0]) + (mx[1] - cx[1]) * (mx[1] - cx[1]) +
                               (mx[2] - cx[2]) * (mx[2] - cx[2]));
                double dj = hz[j];
                if (dist > tiny && (dist < (dk * dk + tiny) || dist < (dj * dj + tiny))) out.push_back(j);
            }
        });
    }
    m.fam_off.assign(nfam + 1, 0);
    m.max_family = 0;
    for (int f = 0; f < nfam; f++) {
        m.fam_off[f + 1] = m.fam_off[f] + int(mem[f].size());
        m.max_family = std::max(m.max_family, int(mem[f].size()));
    }
    m.fam_mem.resize(m.fam_off[nfam]);
    for (int f = 0; f < nfam; f++) std::copy(mem[f].begin(), mem[f].end(), m.fam_mem.begin() + m.fam_off[f]);
}

// setCSRIndexes_gloStiffMat's node graph on the host: per node the sorted, unique union of its
// family's members and the nodes of its FE elements (ne_off / ne_item: FE elements per node)
static void host_node_graph(Model &m, const std::vector<int> &ne_off, const std::vector<int> &ne_item) {
    const int nn = m.nn;
    std::vector<std::vector<int>> nbr(nn);
#pragma omp parallel for schedule(dynamic, 256)
    for (int k = 0; k < nn; k++) {
        std::vector<int> &v = nbr[k];
        const int f = m.famid[k];
        if (f >= 0) v.assign(m.fam_mem.begin() + m.fam_off[f], m.fam_mem.begin() + m.fam_off[f + 1]);
        for (int t = ne_off[k]; t < ne_off[k + 1]; t++) {
            const int e = ne_item[t];
            for (int b = 0; b < m.nen[e]; b++) v.push_back(m.econ[size_t(e) * 8 + b]);
        }
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    m.adj_off.assign(nn + 1, 0);
    for (int k = 0; k < nn; k++) m.adj_off[k + 1] = m.adj_off[k] + int(nbr[k].size());
    m.adj.resize(m.adj_off[nn]);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nn; k++) std::copy(nbr[k].begin(), nbr[k].end(), m.adj.begin() + m.adj_off[k]);
}

int Model::setup(std::string &err) {
    // ---- D matrix and Lame constants (pdsolve.cpp:37-94) ----
    std::fill(D, D + 36, 0.0);
    mu = 0.5 * E / (1 + nu);
    lambda = 0;
    if (dim == 2) {
        if (pro_type == 1) {
            lambda = E * nu / (1 - nu * nu);
            D[0] = E / (1 - nu * nu); D[1] = E * nu / (1 - nu * nu);
            D[3] = E * nu / (1 - nu * nu); D[4] = E / (1 - nu * nu); D[8] = 0.5 * E / (1 + nu);
        } else if (pro_type == 2) {
            lambda = E * nu / ((1 - 2 * nu) * (1 + nu));
            double mp = E / ((1 - 2 * nu) * (1 + nu));
            D[0] = mp * (1 - nu); D[1] = mp * nu; D[3] = mp * nu; D[4] = mp * (1 - nu); D[8] = mp * 0.5 * (1 - 2 * nu);
        } else {
            err = "problem type should be 1 or 2";
            return PDX_ERR_ARG;
        }
    } else {
        lambda = E * nu / ((1 - 2 * nu) * (1 + nu));
        double mp = E / ((1.0 - 2.0 * nu) * (1.0 + nu));
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) D[i * 6 + j] = i == j ? mp * (1 - nu) : mp * nu;
        for (int i = 3; i < 6; i++) D[i * 6 + i] = mp * 0.5 * (1 - 2 * nu);
    }

    // ---- prescribed DOFs (datModel.cpp:186-216) and equation numbering (pdsolve.cpp:241-268) ----
    dofval.assign(size_t(nn) * 3, 0.0);
    presc_set.assign(size_t(nn) * 3, -1);
    for (size_t s = 0; s < ebc.size(); s++)
        for (int k : ebc[s].nid) {
            int d = ebc[s].dof;
            if (k < 0 || k >= nn || d < 0 || d >= dim) { err = "EBC node or DOF out of range"; return PDX_ERR_MESH; }
            if (presc_set[size_t(k) * 3 + d] != -1) {
                err = "The DOF of node " + std::to_string(k + 1) + " is repeatly defined";
                return PDX_ERR_MESH;
            }
            presc_set[size_t(k) * 3 + d] = int(s);
            dofval[size_t(k) * 3 + d] = ebc[s].val;
        }
    eq.assign(size_t(nn) * 3, -1);
    neq = 0;
    for (int k = 0; k < nn; k++)
        for (int i = 0; i < dim; i++)
            if (presc_set[size_t(k) * 3 + i] == -1) eq[size_t(k) * 3 + i] = neq++;

    // ---- volumes and min/max delta (pdsolve.cpp:367-439) ----
    ev.assign(ne, 0.0);
    nv.assign(nn, 0.0);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; e++) ev[e] = element_volume(*this, e);
    double vmax = -std::numeric_limits<double>::max(), vmin = std::numeric_limits<double>::max();
    for (int e = 0; e < ne; e++) {
        for (int a = 0; a < nen[e]; a++) nv[econ[size_t(e) * 8 + a]] += ev[e] / nen[e];
        vmax = std::max(vmax, ev[e]);
        vmin = std::min(vmin, ev[e]);
    }
    minDelta = std::pow(vmin, 1.0 / dim);
    maxDelta = std::pow(vmax, 1.0 / dim);

    // ---- bounding box (pdsolve.cpp:153-195) ----
    for (int c = 0; c < 3; c++) lbc[c] = rtc[c] = x[c];
    for (int i = 1; i < nn; i++)
        for (int j = 0; j < dim; j++) {
            lbc[j] = std::min(lbc[j], x[size_t(i) * 3 + j]);
            rtc[j] = std::max(rtc[j], x[size_t(i) * 3 + j]);
        }
    for (int i = 0; i < dim; i++) {
        lbc[i] = lbc[i] - 1.0E-14 * std::fabs(lbc[i]);
        rtc[i] = rtc[i] + 1.0E-14 * std::fabs(rtc[i]);
    }

    // ---- cells (pdsolve.cpp:441-502): counting sort keeps node / element order ----
    bsize = 2.5 * fac * minDelta;
    for (int i = 0; i < 3; i++) nb[i] = int((rtc[i] - lbc[i]) / bsize + 1);
    if (dim == 2) nb[2] = 1;
    long long ncell = (long long)nb[0] * nb[1] * nb[2];
    if (ncell <= 0 || ncell > 400000000LL) { err = "cell grid too large"; return PDX_ERR_MESH; }
    auto cell_of = [&](const double *p) {
        int ib[3];
        for (int i = 0; i < 3; i++) ib[i] = int((p[i] - lbc[i]) / bsize);
        return ib[0] + ib[1] * nb[0] + ib[2] * nb[0] * nb[1];
    };
    std::vector<int> ecell(ne), ncellv(nn);
    cell_ele_off.assign(ncell + 1, 0);
    cell_node_off.assign(ncell + 1, 0);
    for (int e = 0; e < ne; e++) {
        double xc[3];
        element_center(*this, e, xc);
        ecell[e] = cell_of(xc);
        if (ecell[e] < 0 || ecell[e] >= ncell) { err = "element outside the cell grid"; return PDX_ERR_MESH; }
        cell_ele_off[ecell[e] + 1]++;
    }
    for (int k = 0; k < nn; k++) {
        ncellv[k] = cell_of(&x[size_t(k) * 3]);
        if (ncellv[k] < 0 || ncellv[k] >= ncell) { err = "node outside the cell grid"; return PDX_ERR_MESH; }
        cell_node_off[ncellv[k] + 1]++;
    }
    std::partial_sum(cell_ele_off.begin(), cell_ele_off.end(), cell_ele_off.begin());
    std::partial_sum(cell_node_off.begin(), cell_node_off.end(), cell_node_off.begin());
    cell_ele.resize(ne);
    cell_node.resize(nn);
    {
        std::vector<int> fe_fill(cell_ele_off.begin(), cell_ele_off.end() - 1);
        for (int e = 0; e < ne; e++) cell_ele[fe_fill[ecell[e]]++] = e;
        std::vector<int> fn_fill(cell_node_off.begin(), cell_node_off.end() - 1);
        for (int k = 0; k < nn; k++) cell_node[fn_fill[ncellv[k]]++] = k;
    }

    // ---- PD / FE split with crack conversion (pdsolve.cpp:504-531) ----
    pde.clear();
    for (int e = 0; e < ne; e++)
        if (algo[e] == 1) pde.push_back(e);
    size_t ncr = crack.size() / 9;
    for (size_t c = 0; c < ncr; c++) convert_by_crack(*this, &crack[c * 9]);
    fe.clear();
    for (int e = 0; e < ne; e++)
        if (algo[e] == 2) fe.push_back(e);

    // ---- PD-region boundary + node types (pdsolve.cpp:1221-1440) ----
    boundary_entities(*this);
    ntype.assign(nn, 0);
    for (int e : pde)
        for (int a = 0; a < nen[e]; a++) ntype[econ[size_t(e) * 8 + a]] = 2;
    int nper = dim == 2 ? 2 : 4;
    for (int p = 0; p < npdbe; p++)
        for (int a = 0; a < nper; a++) ntype[pdbe[p * 4 + a]] = 1;

    // ---- families (pdsolve.cpp:1442-1568) ----
    fam_node.clear();
    famid.assign(nn, -1);
    for (int k = 0; k < nn; k++)
        if (ntype[k]) { famid[k] = int(fam_node.size()); fam_node.push_back(k); }
    nfam = int(fam_node.size());
    std::vector<double> hz(nn);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nn; k++) hz[k] = fac * std::pow(nv[k], 1.0 / dim);
    fam_delta.resize(nfam);
    fam_v.assign(nfam, 0.0);
    for (int f = 0; f < nfam; f++) fam_delta[f] = hz[fam_node[f]];
    for (int e : pde)
        for (int a = 0; a < nen[e]; a++) {
            int f = famid[econ[size_t(e) * 8 + a]];
            fam_v[f] = fam_v[f] + ev[e] / nen[e];
        }
    const double tiny = minDelta * minDelta * 1.0e-16;
    const int sdev = setup_device();
    if (sdev >= 0) {
        // the same search on the GPU (setup.hip), lists identical to the host's
        CellGrid g{};
        for (int i = 0; i < 3; i++) {
            g.lo[i] = lbc[i];
            g.nb[i] = nb[i];
        }
        g.cs = bsize;
        g.dim = dim;
        g.clamp = false;
        std::vector<long long> off;
        int rc = dev_cell_search(sdev, nn, x.data(), hz.data(), cell_node_off, cell_node, g, nfam, fam_node.data(),
                                 tiny, off, &fam_mem, max_family, nullptr, err);
        if (rc) return rc;
        if (off[nfam] > (long long)INT32_MAX) {
            err = "more than 2^31 family members";
            return PDX_ERR_LIMIT;
        }
        fam_off.assign(nfam + 1, 0);
        for (int f = 0; f <= nfam; f++) fam_off[f] = int(off[f]);
    } else {
        host_families(*this, hz, tiny);
    }
    bond.assign(fam_mem.size(), 1);

    // ---- bond states across pre-existing cracks (pdsolve.cpp:1823-1915, :8910-9004) ----
    for (size_t c = 0; c < ncr; c++) break_bonds_across(*this, &crack[c * 9], bond);

    // ---- node graph and CSR (pdsolve.cpp:959-1121) ----
    // per node: the family members and the nodes of its FE elements, sorted and unique; built
    // node by node in parallel from the node -> FE element incidence
    {
        std::vector<int> ne_off(nn + 1, 0), ne_item;
        for (int e : fe)
            for (int a = 0; a < nen[e]; a++) ne_off[econ[size_t(e) * 8 + a] + 1]++;
        for (int k = 0; k < nn; k++) ne_off[k + 1] += ne_off[k];
        ne_item.resize(ne_off[nn]);
        {
            std::vector<int> fill(ne_off.begin(), ne_off.end() - 1);
            for (int e : fe)
                for (int a = 0; a < nen[e]; a++) ne_item[fill[econ[size_t(e) * 8 + a]]++] = e;
        }
        if (sdev >= 0) {
            if (int rc = dev_node_graph(sdev, *this, ne_off, ne_item, adj_off, adj, err)) return rc;
        } else {
            host_node_graph(*this, ne_off, ne_item);
        }
    }
    // free-DOF count per node -> row length of every row of that node
    std::vector<int> nfree(nn, 0);
    for (int k = 0; k < nn; k++)
        for (int i = 0; i < dim; i++) nfree[k] += eq[size_t(k) * 3 + i] >= 0;
    ia.assign(size_t(neq) + 1, 0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nn; k++) {
        int64_t len = 0;
        for (int t = adj_off[k]; t < adj_off[k + 1]; t++) len += nfree[adj[t]];
        for (int i = 0; i < dim; i++) {
            int r = eq[size_t(k) * 3 + i];
            if (r >= 0) ia[r + 1] = len;
        }
    }
    for (int r = 0; r < neq; r++) ia[r + 1] += ia[r];
    nnz = ia[neq];
    // the column array (8 bytes per entry: 2 GB for the 100^3 box) is only built when asked for
    ja_built = false;
    ja_store.clear();
    return PDX_OK;
}

const std::vector<int64_t> &Model::csr_ja() const {
    std::lock_guard<std::mutex> lk(ja_mu);
    if (ja_built) return ja_store;
    ja_store.resize(nnz);
#pragma omp parallel for schedule(dynamic, 256)
    for (int k = 0; k < nn; k++)
        for (int i = 0; i < dim; i++) {
            int r = eq[size_t(k) * 3 + i];
            if (r < 0) continue;
            int64_t p = ia[r];
            for (int t = adj_off[k]; t < adj_off[k + 1]; t++)
                for (int j = 0; j < dim; j++) {
                    int c = eq[size_t(adj[t]) * 3 + j];
                    if (c >= 0) ja_store[p++] = c;
                }
        }
    ja_built = true;
    return ja_store;
}

// setCSRIndexes_gloMassMat (pdsolve.cpp:5674-5817): per node the family's own node and the nodes
// of its FE elements (algo 2, every element in element order), deduplicated and sorted; rows of
// pure PD nodes (type 2) hold the diagonal only, other rows every free column of those nodes
void mass_pattern(const Model &m, std::vector<int64_t> &ia, std::vector<int64_t> &ja) {
    const int nn = m.nn, dim = m.dim;
    std::vector<std::vector<int>> nbr(nn);
    for (int f = 0; f < m.nfam; f++) nbr[m.fam_node[f]].push_back(m.fam_node[f]);
    for (int e = 0; e < m.ne; e++) {
        if (m.algo[e] != 2) continue;
        for (int a = 0; a < m.nen[e]; a++) {
            auto &v = nbr[m.econ[size_t(e) * 8 + a]];
            for (int b = 0; b < m.nen[e]; b++) v.push_back(m.econ[size_t(e) * 8 + b]);
        }
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int k = 0; k < nn; k++) {
        std::sort(nbr[k].begin(), nbr[k].end());
        nbr[k].erase(std::unique(nbr[k].begin(), nbr[k].end()), nbr[k].end());
    }
    ia.assign(size_t(m.neq) + 1, 0);
    for (int k = 0; k < nn; k++)
        for (int i = 0; i < dim; i++) {
            const int r = m.eq[size_t(k) * 3 + i];
            if (r < 0) continue;
            int64_t len = 0;
            if (m.ntype[k] == 2) len = 1;
            else
                for (int q : nbr[k])
                    for (int j = 0; j < dim; j++) len += m.eq[size_t(q) * 3 + j] >= 0;
            ia[r + 1] = len;
        }
    for (int r = 0; r < m.neq; r++) ia[r + 1] += ia[r];
    ja.resize(ia[m.neq]);
#pragma omp parallel for schedule(dynamic, 256)
    for (int k = 0; k < nn; k++)
        for (int i = 0; i < dim; i++) {
            const int r = m.eq[size_t(k) * 3 + i];
            if (r < 0) continue;
            int64_t p = ia[r];
            if (m.ntype[k] == 2) {
                ja[p] = r;
                continue;
            }
            for (int q : nbr[k])
                for (int j = 0; j < dim; j++) {
                    const int c = m.eq[size_t(q) * 3 + j];
                    if (c >= 0) ja[p++] = c;
                }
        }
}

int model_from_desc(const pdx_mesh_desc &d, Model &m, std::string &err) {
    if ((d.dim != 2 && d.dim != 3) || d.n_node <= 0 || d.n_ele <= 0 || !d.xyz || !d.ele_nodes || !d.ele_nen ||
        !d.ele_algo) {
        err = "invalid mesh description";
        return PDX_ERR_ARG;
    }
    m.dim = d.dim;
    m.pro_type = d.pro_type;
    m.E = d.E; m.nu = d.nu; m.rho = d.rho;
    m.nn = d.n_node; m.ne = d.n_ele;
    m.x.assign(d.xyz, d.xyz + size_t(d.n_node) * 3);
    m.econ.resize(size_t(d.n_ele) * 8);
    m.nen.assign(d.ele_nen, d.ele_nen + d.n_ele);
    m.algo.assign(d.ele_algo, d.ele_algo + d.n_ele);
    for (int e = 0; e < d.n_ele; e++) {
        int n = m.nen[e];
        bool ok = (d.dim == 2 && (n == 3 || n == 4)) || (d.dim == 3 && (n == 4 || n == 8));
        if (!ok || (m.algo[e] != 1 && m.algo[e] != 2)) { err = "bad element " + std::to_string(e + 1); return PDX_ERR_MESH; }
        for (int a = 0; a < 8; a++) {
            int id = d.ele_nodes[size_t(e) * 8 + a];
            if (a < n && (id < 1 || id > d.n_node)) { err = "element node out of range"; return PDX_ERR_MESH; }
            m.econ[size_t(e) * 8 + a] = a < n ? id - 1 : -1;
        }
    }
    m.ebc.resize(d.n_ebc);
    for (int s = 0; s < d.n_ebc; s++) {
        auto &E = m.ebc[s];
        for (int t = d.ebc_off[s]; t < d.ebc_off[s + 1]; t++) E.nid.push_back(d.ebc_nid[t] - 1);
        E.dof = d.ebc_dof[s];
        E.val = d.ebc_val[s];
        E.vel = d.ebc_vel ? d.ebc_vel[s] : 0.0;
        E.vary = d.ebc_vary ? d.ebc_vary[s] != 0 : false;
    }
    int tn = d.dim == 3 ? 4 : 2;
    m.nbc.resize(d.n_nbc);
    for (int s = 0; s < d.n_nbc; s++) {
        auto &N = m.nbc[s];
        for (int t = d.nbc_off[s] * tn; t < d.nbc_off[s + 1] * tn; t++) {
            int id = d.nbc_con[t];
            if (id < 1 || id > d.n_node) { err = "natural BC node out of range"; return PDX_ERR_MESH; }
            N.con.push_back(id - 1);
        }
        N.val = d.nbc_val[s];
        N.vel = d.nbc_vel ? d.nbc_vel[s] : 0.0;
        N.vary = d.nbc_vary ? d.nbc_vary[s] != 0 : false;
    }
    m.crack.assign(d.crack, d.crack + size_t(d.n_crack) * 9);
    m.fac = d.horizon_factor;
    m.nlf = d.nlf;
    m.me = d.me;
    m.dt = d.dt;
    m.beta = d.beta;
    m.gamma = d.gamma;
    m.sig_ult = d.sigma_ult;
    m.kic = d.kic;
    m.fc = d.fc;
    m.topk = d.topk;
    m.lumped = d.lumped_mass != 0;
    m.newmark = d.newmark != 0;
    m.save_freq = d.save_freq > 0 ? d.save_freq : 1;
    m.vtk_binary = d.vtk_ascii == 0;
    return m.setup(err);
}

// ---- .lis / .in readers: datModel::readdata (datModel.cpp:58-326) and
// fioFiles::excuteCMD (fioFile.cpp:69-462) ----
int model_read_files(const std::string &cmd, Model &m, int &solver_flag, int &num_steps, std::string &err) {
    std::ifstream fc(cmd);
    if (!fc.is_open()) { err = "File \"" + cmd + "\" is not exist."; return PDX_ERR_IO; }
    std::string dir = cmd.substr(0, cmd.find_last_of('/') == std::string::npos ? 0 : cmd.find_last_of('/') + 1);
    std::vector<std::vector<std::string>> cmds;
    std::string line;
    bool have_mesh = false;
    while (std::getline(fc, line)) {
        std::istringstream iss(line);
        std::vector<std::string> tk;
        std::string s;
        while (iss >> s) tk.push_back(s);
        if (tk.empty() || tk[0][0] == '#') continue;
        cmds.push_back(tk);
    }
    solver_flag = -1;
    num_steps = 1;
    std::vector<int> ecount;
    for (auto &tk : cmds) {
        if (tk[0] == "MSHFILE" && tk.size() > 1) {
            std::ifstream fin(dir + tk[1]);
            if (!fin.is_open()) { err = "File \"" + tk[1] + "\" is not exist."; return PDX_ERR_IO; }
            std::string title, label, sdim;
            fin >> title;
            m.title = title;
            std::getline(fin, label);
            std::getline(fin, label);
            int pro;
            fin >> sdim >> pro;
            double E, nu, rho, k1, su;
            fin >> E >> nu >> rho >> k1 >> su;
            m.dim = sdim == "3D" ? 3 : (sdim == "2D" ? 2 : 0);
            if (!m.dim) { err = "The problem dimension is not 3D or 2D"; return PDX_ERR_MESH; }
            m.pro_type = pro; m.E = E; m.nu = nu; m.rho = rho;
            m.kic = k1;
            m.sig_ult = su;
            fin >> m.nn >> m.ne;
            m.x.assign(size_t(m.nn) * 3, 0.0);
            for (int i = 0; i < m.nn; i++) {
                int id;
                fin >> id >> m.x[size_t(i) * 3] >> m.x[size_t(i) * 3 + 1] >> m.x[size_t(i) * 3 + 2];
            }
            m.econ.assign(size_t(m.ne) * 8, -1);
            m.nen.assign(m.ne, 0);
            m.algo.assign(m.ne, 0);
            for (int i = 0; i < m.ne; i++) {
                int id, c[8];
                fin >> id >> m.algo[i];
                if (m.dim == 3) {
                    for (int j = 0; j < 8; j++) fin >> c[j];
                    if (c[2] != c[3] && c[4] != c[5]) {
                        m.nen[i] = 8;
                        for (int j = 0; j < 8; j++) m.econ[size_t(i) * 8 + j] = c[j] - 1;
                    } else {
                        m.nen[i] = 4;
                        int t[4] = {c[0], c[1], c[2], c[4]};
                        for (int j = 0; j < 4; j++) m.econ[size_t(i) * 8 + j] = t[j] - 1;
                    }
                } else {
                    for (int j = 0; j < 4; j++) fin >> c[j];
                    m.nen[i] = c[3] != c[2] ? 4 : 3;
                    for (int j = 0; j < m.nen[i]; j++) m.econ[size_t(i) * 8 + j] = c[j] - 1;
                }
            }
            if (!fin) { err = "mesh file truncated"; return PDX_ERR_IO; }
            std::string blank;
            std::getline(fin, blank);
            std::getline(fin, blank);
            int nebc;
            fin >> nebc;
            m.ebc.assign(nebc, EbcSet());
            ecount.assign(nebc, 0);
            for (int s = 0; s < nebc; s++) {
                std::string dof;
                fin >> ecount[s] >> dof >> m.ebc[s].val;
                if (dof == "UX") m.ebc[s].dof = 0;
                else if (dof == "UY") m.ebc[s].dof = 1;
                else if (dof == "UZ") m.ebc[s].dof = 2;
                else { err = "DOFs must be UX, UY, or UZ"; return PDX_ERR_MESH; }
            }
            for (int s = 0; s < nebc; s++) {
                m.ebc[s].nid.resize(ecount[s]);
                for (int j = 0; j < ecount[s]; j++) { fin >> m.ebc[s].nid[j]; m.ebc[s].nid[j] -= 1; }
            }
            std::getline(fin, blank);
            std::getline(fin, blank);
            int nnbc;
            fin >> nnbc;
            m.nbc.assign(nnbc, NbcSet());
            std::vector<int> ncount(nnbc);
            for (int s = 0; s < nnbc; s++) fin >> ncount[s] >> m.nbc[s].val;
            int tn = m.dim == 3 ? 4 : 2;
            for (int s = 0; s < nnbc; s++) {
                m.nbc[s].con.resize(size_t(ncount[s]) * tn);
                for (auto &v : m.nbc[s].con) { fin >> v; v -= 1; }
            }
            std::getline(fin, blank);
            std::getline(fin, blank);
            int nnf;
            fin >> nnf;
            for (int i = 0; i < nnf; i++) { int t; fin >> t; }
            std::getline(fin, blank);
            std::getline(fin, blank);
            int ncr = 0;
            fin >> ncr;
            m.crack.assign(size_t(ncr) * 9, 0.0);
            int np = m.dim == 3 ? 3 : 2;
            for (int i = 0; i < ncr; i++)
                for (int j = 0; j < np; j++)
                    for (int k = 0; k < 3; k++) fin >> m.crack[size_t(i) * 9 + j * 3 + k];
            if (!fin) { err = "mesh file truncated in BC sections"; return PDX_ERR_IO; }
            have_mesh = true;
        } else if (tk[0] == "SOLVER" && tk.size() > 1) {
            solver_flag = tk[1] == "DYNAMIC" ? 0 : tk[1] == "STATIC" ? 1 : tk[1] == "QUASI-STATIC" ? 2 : -1;
            if (solver_flag < 0) { err = "Don't exist command of \"" + tk[1] + "\""; return PDX_ERR_ARG; }
        } else if (tk[0] == "SETSOLVING" && tk.size() >= 6) {
            m.dt = solver_flag == 2 ? 1.0 : std::atof(tk[1].c_str());
            num_steps = std::atoi(tk[2].c_str());
            m.save_freq = std::atoi(tk[3].c_str());
            m.fac = std::atof(tk[4].c_str());
            m.nlf = std::atof(tk[5].c_str());
            if (tk.size() > 6) m.me = std::atof(tk[6].c_str());
        } else if (tk[0] == "FC" && tk.size() > 1) {
            // fioFiles::excuteCMD (fioFile.cpp:340-386): FC flag [k for flag 2]
            m.fc = std::atoi(tk[1].c_str());
            if (m.fc == 2 && tk.size() > 2) m.topk = std::atoi(tk[2].c_str());
        } else if (tk[0] == "TOPK" && tk.size() > 1) {
            m.topk = std::atoi(tk[1].c_str());
        } else if (tk[0] == "NEWMARK" && tk.size() > 1) {
            // fioFiles::excuteCMD (fioFile.cpp:151-181): ON / OFF, then optional gamma and beta
            if (tk[1] == "ON") m.newmark = true;
            else if (tk[1] == "OFF") m.newmark = false;
            if (tk.size() > 3) {
                m.gamma = std::atof(tk[2].c_str());
                m.beta = std::atof(tk[3].c_str());
            }
        } else if (tk[0] == "LUMPMASS" && tk.size() > 1) {
            // fioFile.cpp:208-230; anything but ON / OFF is ignored
            if (tk[1] == "ON") m.lumped = true;
            else if (tk[1] == "OFF") m.lumped = false;
        } else if (tk[0] == "VTKFORMAT" && tk.size() > 1) {
            // fioFile.cpp:422-444
            if (tk[1] == "ASCII") m.vtk_binary = false;
            else if (tk[1] == "BINARY") m.vtk_binary = true;
        } else if (tk[0] == "RF" && tk.size() > 1) {
            // fioFile.cpp:402-421: EBC set ids whose reaction force the reference reports
            for (size_t i = 1; i < tk.size(); i++) m.rf_ids.push_back(std::atoi(tk[i].c_str()));
        } else if (!have_mesh) {
            continue;
        } else if ((tk[0] == "EBC" || tk[0] == "VEBC") && tk.size() > 2) {
            int id = std::atoi(tk[1].c_str());
            if (id < 0 || id >= int(m.ebc.size())) continue;
            if (tk[0] == "EBC") m.ebc[id].val = std::atof(tk[2].c_str());
            else { m.ebc[id].vary = true; m.ebc[id].vel = std::atof(tk[2].c_str()); }
        } else if ((tk[0] == "NBC" || tk[0] == "VNBC") && tk.size() > 2) {
            int id = std::atoi(tk[1].c_str());
            if (id < 0 || id >= int(m.nbc.size())) continue;
            if (tk[0] == "NBC") m.nbc[id].val = std::atof(tk[2].c_str());
            else { m.nbc[id].vary = true; m.nbc[id].vel = std::atof(tk[2].c_str()); }
        }
    }
    if (!have_mesh) { err = "the command file names no MSHFILE"; return PDX_ERR_IO; }
    m.solver_flag = solver_flag;
    m.num_steps = num_steps;
    if (m.save_freq <= 0) m.save_freq = 1;
    for (auto &s : m.ebc)
        for (int k : s.nid)
            if (k < 0 || k >= m.nn) { err = "EBC node out of range"; return PDX_ERR_MESH; }
    return m.setup(err);
}

}  // namespace pdx
