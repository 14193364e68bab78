/*
 * pdlsm_oracle.c -- CPU restatement of the PDLSM-FEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library in pdlsm-fem_mpi_amd/ is tested against
 * this file: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links or calls it.
 *
 * It restates, in plain single-threaded C, the algorithms of
 * SourceFiles/Solver/pdsolve.cpp and the element files of QibangLiu/PDLSM-FEM_mpi
 * exactly as the reference runs them: node-major DOF numbering, block (cell)
 * binning, families in the reference member order, one LU factorisation PER
 * BOND of the shape tensor (the reference calls LAPACKE_dgesv per bond), a
 * binary search per scattered CSR entry, and the same loop nests and
 * floating-point expression order wherever the reference fixes them.
 * Each function cites the reference file:line it follows.
 *
 * Deviations (documented, rounding-level only):
 *  - the reference sorts boundary edges/faces with an unstable std::sort keyed
 *    on the first node (pdsolve.cpp:1274, :1347); ties here keep insertion
 *    order.  Only the summation order of boundary terms into one K entry can
 *    differ.
 *  - LAPACKE_dgesv (MKL, closed) is replaced by a textbook partial-pivot LU
 *    (first maximum pivot, division by the pivot, row-oriented update).
 *  - MPI reductions are single-rank (rank 0 of 1), which is what the
 *    reference computes with one process.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OM_OK 0
#define OM_ERR_ARG (-1)
#define OM_ERR_SINGULAR (-2)
#define OM_ERR_PATTERN (-3)
#define OM_ERR_MESH (-4)

typedef struct {
    /* ---- inputs (copied at om_create) ---- */
    int dim, proType;
    double E, nu, rho;
    int nn;
    double *x;          /* [nn*3] */
    int ne;
    int *econ;          /* [ne*8] 1-based node ids */
    int *nen;           /* nodes per element: 3,4 (2D) / 4,8 (3D) */
    int *algo;          /* 1 = PD element, 2 = FE element (mutated by cracks) */
    int n_ebc;
    int *ebc_off, *ebc_nid, *ebc_dof;
    double *ebc_val, *ebc_vel;
    int *ebc_vary;
    int n_nbc, tn;
    int *nbc_off, *nbc_con;
    double *nbc_val, *nbc_vel;
    int *nbc_vary;
    int n_crack;
    double *crack;      /* [n_crack*9] */
    double fac, nlf, me, dt, beta, gamma;
    /* ---- DOF state (pdDof) ---- */
    int *active;        /* [nn*3] */
    double *val;        /* [nn*3] prescribed / current values */
    int *eq;            /* [nn*3] equation index or -1 */
    int neq;
    /* ---- geometry / setup ---- */
    double *ev, *nv;
    double minDelta, maxDelta, lbc[3], rtc[3], bsize;
    int nb[3];
    int *bn_off, *bn;   /* nodes per block (1-based ids, insertion order) */
    int *be_off, *be;   /* elements per block (1-based ids) */
    int *pde; int npde; /* PD element list (civ_pdeIDX) */
    int *fe; int nfe;   /* FE element list (civ_feIDX) */
    int *crac_ele; int ncrac_ele, cap_crac_ele;
    int *pdbe; int npdbe;  /* [npdbe*4] 1-based */
    int *ntype, *famid;
    int nfam;
    int *fam_node;      /* 0-based node of family k */
    int *fam_off, *fam_mem;  /* 1-based node ids, self first */
    int *bond;
    double *fam_v, *fam_delta;
    long long *ia, *ja, nnz;
    double D[36], lambda, mu;
} om_model;

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

/* pdfemEles::detMat33 (pdfemEles.cpp:155-172): cofactor form, product order kept */
static double det33(double m[3][3]) {
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

/* PLUSolve (calMatrixOperations.cpp:587-638) semantics: A x = b, n <= 19,
 * partial pivoting (first maximum), row-major A copied (A is not modified). */
static int lu_solve(int n, const double *Ain, const double *bin, double *xout) {
    double a[19 * 19], b[19];
    memcpy(a, Ain, sizeof(double) * n * n);
    memcpy(b, bin, sizeof(double) * n);
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; i++) {
            double v = fabs(a[i * n + k]);
            if (v > mx) { mx = v; p = i; }
        }
        if (a[p * n + k] == 0.0) return OM_ERR_SINGULAR;
        if (p != k) {
            for (int c = 0; c < n; c++) { double t = a[k * n + c]; a[k * n + c] = a[p * n + c]; a[p * n + c] = t; }
            double t = b[k]; b[k] = b[p]; b[p] = t;
        }
        for (int i = k + 1; i < n; i++) {
            double l = a[i * n + k] / a[k * n + k];
            a[i * n + k] = l;
            for (int c = k + 1; c < n; c++) a[i * n + c] = a[i * n + c] - l * a[k * n + c];
            b[i] = b[i] - l * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int c = i + 1; c < n; c++) s = s - a[i * n + c] * xout[c];
        xout[i] = s / a[i * n + i];
    }
    return OM_OK;
}

/* matMultiply (calMatrixOperations.cpp:720-743): temp = a_ik*b_kj + temp, k ascending */
static void matmul(int ar, int ac, int bc, const double *A, const double *B, double *C) {
    for (int i = 0; i < ar; i++)
        for (int j = 0; j < bc; j++) {
            double t = 0.0;
            for (int k = 0; k < ac; k++) t = A[i * ac + k] * B[k * bc + j] + t;
            C[i * bc + j] = t;
        }
}

static const double GP2[2] = {-0.57735026918962573, 0.57735026918962573};

/* ------------------------------------------------------------------------ */
/* elements (pdfemEleQuad4N.cpp, pdfemEleTri3N.cpp, pdfemEleBrick8N.cpp,     */
/* pdfemEleTetra4N.cpp, pdfemEleLine2N.cpp)                                 */
/* ------------------------------------------------------------------------ */

static void ele_xyz(const om_model *M, int e, double xN[8][3]) {
    for (int a = 0; a < M->nen[e]; a++)
        for (int c = 0; c < 3; c++) xN[a][c] = M->x[(M->econ[e * 8 + a] - 1) * 3 + c];
}

/* pdfemEleQuad4N::eleVolume (pdfemEleQuad4N.cpp:212-237): two triangles */
static double quad4_volume(double xN[][3]) {
    double m[3][3];
    m[0][0] = 1, m[0][1] = xN[0][0], m[0][2] = xN[0][1];
    m[1][0] = 1, m[1][1] = xN[1][0], m[1][2] = xN[1][1];
    m[2][0] = 1, m[2][1] = xN[2][0], m[2][2] = xN[2][1];
    double v = 0.5 * fabs(det33(m));
    m[0][0] = 1, m[0][1] = xN[2][0], m[0][2] = xN[2][1];
    m[1][0] = 1, m[1][1] = xN[3][0], m[1][2] = xN[3][1];
    m[2][0] = 1, m[2][1] = xN[0][0], m[2][2] = xN[0][1];
    return v + 0.5 * fabs(det33(m));
}

/* pdfemEleTri3N::eleVolume (pdfemEleTri3N.cpp:176-183) */
static double tri3_volume(double xN[][3]) {
    double m[3][3];
    m[0][0] = 1, m[0][1] = xN[0][0], m[0][2] = xN[0][1];
    m[1][0] = 1, m[1][1] = xN[1][0], m[1][2] = xN[1][1];
    m[2][0] = 1, m[2][1] = xN[2][0], m[2][2] = xN[2][1];
    return 0.5 * fabs(det33(m));
}

/* pdfemEleTetra4N::eleVolume (pdfemEleTetra4N.cpp:210-217) */
static double tet4_volume(double xN[][3]) {
    double m[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) m[r][c] = xN[r + 1][c] - xN[0][c];
    return fabs(det33(m)) / 6.0;
}

static const double HP[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
static const double HQ[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
static const double HR[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};

/* pdfemEleBrick8N::JacobiMat (pdfemEleBrick8N.cpp:324-348) */
static void hex8_jac(double xN[][3], double p, double q, double r, double J[9], double dN[8][3]) {
    for (int i = 0; i < 8; i++) {
        dN[i][0] = 1.0 / 8 * HP[i] * (1 + q * HQ[i]) * (1 + r * HR[i]);
        dN[i][1] = 1.0 / 8 * HQ[i] * (1 + p * HP[i]) * (1 + r * HR[i]);
        dN[i][2] = 1.0 / 8 * HR[i] * (1 + p * HP[i]) * (1 + q * HQ[i]);
    }
    for (int k = 0; k < 9; k++) J[k] = 0;
    for (int cc = 0; cc < 3; cc++)
        for (int i = 0; i < 8; i++) {
            J[0 * 3 + cc] = dN[i][0] * xN[i][cc] + J[0 * 3 + cc];
            J[1 * 3 + cc] = dN[i][1] * xN[i][cc] + J[1 * 3 + cc];
            J[2 * 3 + cc] = dN[i][2] * xN[i][cc] + J[2 * 3 + cc];
        }
}

/* pdfemEleBrick8N::detJacobi (pdfemEleBrick8N.cpp:93-117) */
static double hex8_detJ(double xN[][3], double p, double q, double r) {
    double J[9], dN[8][3], m[3][3];
    hex8_jac(xN, p, q, r, J, dN);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) m[a][b] = J[a * 3 + b];
    return det33(m);
}

/* pdfemEleBrick8N::eleVolume (pdfemEleBrick8N.cpp:273-296): 2x2x2 Gauss, p outer */
static double hex8_volume(double xN[][3]) {
    double v = 0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) v = v + 1.0 * 1.0 * 1.0 * hex8_detJ(xN, GP2[a], GP2[b], GP2[c]);
    return v;
}

static double ele_volume(const om_model *M, int e) {
    double xN[8][3];
    ele_xyz(M, e, xN);
    int n = M->nen[e];
    if (M->dim == 2) return n == 4 ? quad4_volume(xN) : tri3_volume(xN);
    return n == 8 ? hex8_volume(xN) : tet4_volume(xN);
}

/* eleCenter of each element type (pdfemEleQuad4N.cpp:28-40, pdfemEleTri3N.cpp:24-30,
 * pdfemEleBrick8N.cpp:37-49, pdfemEleTetra4N.cpp:27-38) */
static void ele_center(const om_model *M, int e, double xc[3]) {
    double xN[8][3];
    ele_xyz(M, e, xN);
    int n = M->nen[e];
    xc[0] = xc[1] = xc[2] = 0;
    if (M->dim == 2 && n == 3) {
        for (int i = 0; i < 3; i++) xc[i] = (xN[0][i] + xN[1][i] + xN[2][i]) / 3.0;
        return;
    }
    double N[8];
    if (M->dim == 2) for (int a = 0; a < 4; a++) N[a] = 0.25;  /* 0.25*(1-0)*(1-0) */
    else if (n == 8) for (int a = 0; a < 8; a++) N[a] = 1.0 / 8;
    else for (int a = 0; a < 4; a++) N[a] = 0.25;
    for (int a = 0; a < n; a++)
        for (int i = 0; i < 3; i++) xc[i] = xc[i] + xN[a][i] * N[a];
}

/* B matrices ---------------------------------------------------------- */

/* pdfemEleQuad4N::BmatFEM (pdfemEleQuad4N.cpp:276-318) and detJacobi (:50-71) */
static void quad4_B(double xN[][3], double p, double q, double B[3 * 8], double *detJ_out) {
    double dNdp[4] = {0.25 * (-1 + q), 0.25 * (1 - q), 0.25 * (1 + q), 0.25 * (-1 - q)};
    double dNdq[4] = {0.25 * (-1 + p), 0.25 * (-1 - p), 0.25 * (1 + p), 0.25 * (1 - p)};
    double J[2][2] = {{0, 0}, {0, 0}};
    for (int i = 0; i < 4; i++) {
        J[0][0] = dNdp[i] * xN[i][0] + J[0][0];
        J[0][1] = dNdp[i] * xN[i][1] + J[0][1];
        J[1][0] = dNdq[i] * xN[i][0] + J[1][0];
        J[1][1] = dNdq[i] * xN[i][1] + J[1][1];
    }
    double detJ = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double iJ[2][2] = {{J[1][1] / detJ, -J[0][1] / detJ}, {-J[1][0] / detJ, J[0][0] / detJ}};
    memset(B, 0, sizeof(double) * 24);
    for (int i = 0; i < 4; i++) {
        double b1 = iJ[0][0] * dNdp[i] + iJ[0][1] * dNdq[i];
        double b2 = iJ[1][0] * dNdp[i] + iJ[1][1] * dNdq[i];
        B[0 * 8 + 2 * i] = b1;
        B[1 * 8 + 2 * i + 1] = b2;
        B[2 * 8 + 2 * i] = b2;
        B[2 * 8 + 2 * i + 1] = b1;
    }
    *detJ_out = detJ;
}

/* pdfemEleBrick8N::BmatFEM (pdfemEleBrick8N.cpp:350-406): dN/dX by an LU solve
 * J * dNdX = dNdpqr per shape function */
static int hex8_B(double xN[][3], double p, double q, double r, double B[6 * 24]) {
    double J[9], dN[8][3];
    hex8_jac(xN, p, q, r, J, dN);
    memset(B, 0, sizeof(double) * 144);
    for (int i = 0; i < 8; i++) {
        double g[3];
        if (lu_solve(3, J, dN[i], g)) return OM_ERR_SINGULAR;
        B[0 * 24 + 3 * i] = g[0];
        B[1 * 24 + 1 + 3 * i] = g[1];
        B[2 * 24 + 2 + 3 * i] = g[2];
        B[3 * 24 + 3 * i] = g[1];
        B[3 * 24 + 1 + 3 * i] = g[0];
        B[4 * 24 + 1 + 3 * i] = g[2];
        B[4 * 24 + 2 + 3 * i] = g[1];
        B[5 * 24 + 3 * i] = g[2];
        B[5 * 24 + 2 + 3 * i] = g[0];
    }
    return OM_OK;
}

/* Ke += (B^T D B) * w  via BtD = Bt*D, BtDB = BtD*B, scaled, then added
 * (pdfemEleQuad4N.cpp:320-359, pdfemEleBrick8N.cpp:436-552) */
static void add_btdb(int nv, int nd, const double *B, const double *D, double w, double *Ke) {
    double Bt[24 * 6], BtD[24 * 6], BtDB[24 * 24];
    for (int i = 0; i < nv; i++)
        for (int j = 0; j < nd; j++) Bt[j * nv + i] = B[i * nd + j];
    matmul(nd, nv, nv, Bt, D, BtD);
    matmul(nd, nv, nd, BtD, B, BtDB);
    for (int k = 0; k < nd * nd; k++) BtDB[k] = BtDB[k] * w;
    for (int k = 0; k < nd * nd; k++) Ke[k] = Ke[k] + BtDB[k];
}

/* element stiffness of every element type; nd = dofs per element */
/* pdfemEleTri3N::BmatFEM (pdfemEleTri3N.cpp:185-210): B (3 x 6) */
static void tri3_B(double xN[][3], double *B) {
    double V2 = 2.0 * tri3_volume(xN);
    double b[3] = {xN[1][1] - xN[2][1], xN[2][1] - xN[0][1], xN[0][1] - xN[1][1]};
    double c[3] = {-xN[1][0] + xN[2][0], -xN[2][0] + xN[0][0], -xN[0][0] + xN[1][0]};
    memset(B, 0, sizeof(double) * 18);
    for (int i = 0; i < 3; i++) {
        B[0 * 6 + 2 * i] = b[i] / V2;
        B[1 * 6 + 2 * i + 1] = c[i] / V2;
        B[2 * 6 + 2 * i] = c[i] / V2;
        B[2 * 6 + 2 * i + 1] = b[i] / V2;
    }
}

/* pdfemEleTetra4N::BmatFEM (pdfemEleTetra4N.cpp:84-170): B (6 x 12) */
static void tet4_B(double xN[][3], double *B) {
    double m[3][3], b[4], c[4], d[4];
#define SETM(r0, r1, r2, c0, c1) \
    m[0][0] = 1, m[0][1] = xN[r0][c0], m[0][2] = xN[r0][c1]; \
    m[1][0] = 1, m[1][1] = xN[r1][c0], m[1][2] = xN[r1][c1]; \
    m[2][0] = 1, m[2][1] = xN[r2][c0], m[2][2] = xN[r2][c1];
    SETM(1, 2, 3, 1, 2) b[0] = -det33(m);
    SETM(0, 2, 3, 1, 2) b[1] = det33(m);
    SETM(0, 1, 3, 1, 2) b[2] = -det33(m);
    SETM(0, 1, 2, 1, 2) b[3] = det33(m);
    SETM(1, 2, 3, 0, 2) c[0] = det33(m);
    SETM(0, 2, 3, 0, 2) c[1] = -det33(m);
    SETM(0, 1, 3, 0, 2) c[2] = det33(m);
    SETM(0, 1, 2, 0, 2) c[3] = -det33(m);
    SETM(1, 2, 3, 0, 1) d[0] = -det33(m);
    SETM(0, 2, 3, 0, 1) d[1] = det33(m);
    SETM(0, 1, 3, 0, 1) d[2] = -det33(m);
    SETM(0, 1, 2, 0, 1) d[3] = det33(m);
#undef SETM
    double V6 = 6.0 * tet4_volume(xN);
    memset(B, 0, sizeof(double) * 72);
    for (int i = 0; i < 4; i++) {
        B[0 * 12 + 3 * i] = b[i] / V6;
        B[1 * 12 + 3 * i + 1] = c[i] / V6;
        B[2 * 12 + 3 * i + 2] = d[i] / V6;
        B[3 * 12 + 3 * i] = c[i] / V6;
        B[3 * 12 + 3 * i + 1] = b[i] / V6;
        B[4 * 12 + 3 * i + 1] = d[i] / V6;
        B[4 * 12 + 3 * i + 2] = c[i] / V6;
        B[5 * 12 + 3 * i] = d[i] / V6;
        B[5 * 12 + 3 * i + 2] = b[i] / V6;
    }
}

static int ele_stiff(const om_model *M, int e, double *Ke, int *nd_out) {
    double xN[8][3];
    ele_xyz(M, e, xN);
    int n = M->nen[e];
    if (M->dim == 2 && n == 4) {
        memset(Ke, 0, sizeof(double) * 64);
        for (int m = 0; m < 2; m++)
            for (int k = 0; k < 2; k++) {
                double B[24], detJ;
                quad4_B(xN, GP2[m], GP2[k], B, &detJ);
                add_btdb(3, 8, B, M->D, 1.0 * 1.0 * detJ, Ke);
            }
        *nd_out = 8;
    } else if (M->dim == 2) {
        /* pdfemEleTri3N::eleStiffMatFEM / BmatFEM (pdfemEleTri3N.cpp:33-47, :185-210) */
        double V = tri3_volume(xN), B[18];
        tri3_B(xN, B);
        double Bt[18], BtD[18];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 6; j++) Bt[j * 3 + i] = B[i * 6 + j];
        matmul(6, 3, 3, Bt, M->D, BtD);
        matmul(6, 3, 6, BtD, B, Ke);
        for (int k = 0; k < 36; k++) Ke[k] = Ke[k] * V;
        *nd_out = 6;
    } else if (n == 8) {
        memset(Ke, 0, sizeof(double) * 576);
        for (int mr = 0; mr < 2; mr++)
            for (int mq = 0; mq < 2; mq++)
                for (int mp = 0; mp < 2; mp++) {
                    double p = GP2[mp], q = GP2[mq], r = GP2[mr], B[144];
                    double detJ = hex8_detJ(xN, p, q, r);
                    if (hex8_B(xN, p, q, r, B)) return OM_ERR_SINGULAR;
                    add_btdb(6, 24, B, M->D, 1.0 * 1.0 * 1.0 * detJ, Ke);
                }
        *nd_out = 24;
    } else {
        /* pdfemEleTetra4N::BmatFEM (pdfemEleTetra4N.cpp:84-170) */
        double V = tet4_volume(xN), B[72];
        tet4_B(xN, B);
        double Bt[72], BtD[72];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 12; j++) Bt[j * 6 + i] = B[i * 12 + j];
        matmul(12, 6, 6, Bt, M->D, BtD);
        matmul(12, 6, 12, BtD, B, Ke);
        for (int k = 0; k < 144; k++) Ke[k] = Ke[k] * V;
        *nd_out = 12;
    }
    return OM_OK;
}

/* consistent element mass (pdfemEleQuad4N.cpp:361-387 + :458-482,
 * pdfemEleBrick8N.cpp:554-582 + :66-91, pdfemEleTri3N.cpp:49-70,
 * pdfemEleTetra4N.cpp:55-76) */
static void ele_mass(const om_model *M, int e, double *Me, int *nd_out) {
    double xN[8][3];
    ele_xyz(M, e, xN);
    int n = M->nen[e];
    double rho = M->rho;
    if (M->dim == 2 && n == 4) {
        memset(Me, 0, sizeof(double) * 64);
        for (int ip = 0; ip < 2; ip++)
            for (int iq = 0; iq < 2; iq++) {
                double p = GP2[ip], q = GP2[iq], Nn[4], NN[16], NNt[16], NtN[64];
                Nn[0] = 0.25 * (1 - p) * (1 - q);
                Nn[1] = 0.25 * (1 + p) * (1 - q);
                Nn[2] = 0.25 * (1 + p) * (1 + q);
                Nn[3] = 0.25 * (1 - p) * (1 + q);
                double dNdp[4] = {0.25 * (-1 + q), 0.25 * (1 - q), 0.25 * (1 + q), 0.25 * (-1 - q)};
                double dNdq[4] = {0.25 * (-1 + p), 0.25 * (-1 - p), 0.25 * (1 + p), 0.25 * (1 - p)};
                double J[2][2] = {{0, 0}, {0, 0}};
                for (int i = 0; i < 4; i++) {
                    J[0][0] = J[0][0] + dNdp[i] * xN[i][0];
                    J[0][1] = J[0][1] + dNdp[i] * xN[i][1];
                    J[1][0] = J[1][0] + dNdq[i] * xN[i][0];
                    J[1][1] = J[1][1] + dNdq[i] * xN[i][1];
                }
                double detJ = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                memset(NN, 0, sizeof(NN));
                for (int i = 0; i < 2; i++)
                    for (int j = 0; j < 4; j++) NN[i * 8 + 2 * j + i] = Nn[j];
                for (int i = 0; i < 2; i++)
                    for (int j = 0; j < 8; j++) NNt[j * 2 + i] = NN[i * 8 + j] * rho;
                matmul(8, 2, 8, NNt, NN, NtN);
                for (int k = 0; k < 64; k++) NtN[k] = NtN[k] * (detJ * 1.0 * 1.0);
                for (int k = 0; k < 64; k++) Me[k] = Me[k] + NtN[k];
            }
        *nd_out = 8;
    } else if (M->dim == 2) {
        double fac = rho * tri3_volume(xN) / 12.0;
        memset(Me, 0, sizeof(double) * 36);
        for (int i = 0; i < 6; i++) Me[i * 6 + i] = fac * 2;
        for (int r = 0; r < 6; r++)
            for (int c = r + 2; c < 6; c += 2) { Me[r * 6 + c] = fac; Me[c * 6 + r] = fac; }
        *nd_out = 6;
    } else if (n == 8) {
        memset(Me, 0, sizeof(double) * 576);
        for (int ip = 0; ip < 2; ip++)
            for (int iq = 0; iq < 2; iq++)
                for (int ir = 0; ir < 2; ir++) {
                    double p = GP2[ip], q = GP2[iq], r = GP2[ir], Nn[8], NN[72], NNt[72], NtN[576];
                    double detJ = hex8_detJ(xN, p, q, r);
                    for (int i = 0; i < 8; i++) Nn[i] = 1.0 / 8 * (1 + p * HP[i]) * (1 + q * HQ[i]) * (1 + r * HR[i]);
                    memset(NN, 0, sizeof(NN));
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 8; j++) NN[i * 24 + 3 * j + i] = Nn[j];
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 24; j++) NNt[j * 3 + i] = NN[i * 24 + j];
                    matmul(24, 3, 24, NNt, NN, NtN);
                    double s = rho * detJ * 1.0 * 1.0 * 1.0;
                    for (int k = 0; k < 576; k++) NtN[k] = NtN[k] * s;
                    for (int k = 0; k < 576; k++) Me[k] = Me[k] + NtN[k];
                }
        *nd_out = 24;
    } else {
        double fac = rho * tet4_volume(xN) / 20.0;
        memset(Me, 0, sizeof(double) * 144);
        for (int i = 0; i < 12; i++) Me[i * 12 + i] = fac * 2;
        for (int r = 0; r < 12; r++)
            for (int c = r + 3; c < 12; c += 3) { Me[r * 12 + c] = fac; Me[c * 12 + r] = fac; }
        *nd_out = 12;
    }
}

/* ------------------------------------------------------------------------ */
/* model setup (pdsolve.cpp:20-151 and the functions it calls)               */
/* ------------------------------------------------------------------------ */

/* pdsolve::pdsolve, D matrix (pdsolve.cpp:37-94) */
static int set_D(om_model *M) {
    double E = M->E, nu = M->nu;
    memset(M->D, 0, sizeof(M->D));
    M->mu = 0.5 * E / (1 + nu);
    M->lambda = 0;
    if (M->dim == 2) {
        double *D = M->D; /* 3x3 stored with stride 3 */
        if (M->proType == 1) {
            M->lambda = E * nu / (1 - nu * nu);
            D[0] = E / (1 - nu * nu); D[1] = E * nu / (1 - nu * nu);
            D[3] = E * nu / (1 - nu * nu); D[4] = E / (1 - nu * nu);
            D[8] = 0.5 * E / (1 + nu);
        } else if (M->proType == 2) {
            M->lambda = E * nu / ((1 - 2 * nu) * (1 + nu));
            double mp = E / ((1 - 2 * nu) * (1 + nu));
            D[0] = mp * (1 - nu); D[1] = mp * nu; D[3] = mp * nu; D[4] = mp * (1 - nu);
            D[8] = mp * 0.5 * (1 - 2 * nu);
        } else return OM_ERR_ARG;
    } else {
        M->lambda = E * nu / ((1 - 2 * nu) * (1 + nu));
        double mp = E / ((1.0 - 2.0 * nu) * (1.0 + nu));
        double *D = M->D; /* 6x6 stride 6 */
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) D[i * 6 + j] = (i == j) ? mp * (1 - nu) : mp * nu;
        for (int i = 3; i < 6; i++) D[i * 6 + i] = mp * 0.5 * (1 - 2 * nu);
    }
    return OM_OK;
}

/* pdsolve::Setdof_Index (pdsolve.cpp:241-268) */
static void setdof_index(om_model *M) {
    int c = 0;
    for (int k = 0; k < M->nn; k++)
        for (int i = 0; i < M->dim; i++) {
            if (M->active[k * 3 + i]) M->eq[k * 3 + i] = c++;
            else M->eq[k * 3 + i] = -1;
        }
    M->neq = c;
}

/* pdsolve::setDeltaMaxMin_EleV_NV (pdsolve.cpp:367-439) */
static void set_volumes(om_model *M) {
    double vmax = -1.79769313486231570e308, vmin = 1.79769313486231570e308;
    M->ev = xcalloc(M->ne, sizeof(double));
    M->nv = xcalloc(M->nn, sizeof(double));
    for (int e = 0; e < M->ne; e++) {
        double dv = ele_volume(M, e);
        M->ev[e] = dv;
        int n = M->nen[e];
        for (int a = 0; a < n; a++) {
            int nid = M->econ[e * 8 + a] - 1;
            M->nv[nid] = M->nv[nid] + dv / n;
        }
        if (dv > vmax) vmax = dv;
        if (dv < vmin) vmin = dv;
    }
    M->minDelta = pow(vmin, 1.0 / M->dim);
    M->maxDelta = pow(vmax, 1.0 / M->dim);
}

/* pdsolve::findDomainDimen (pdsolve.cpp:153-195) */
static void domain_dimen(om_model *M) {
    double lbc[3] = {0, 0, 0}, rtc[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++) lbc[c] = rtc[c] = M->x[c];
    for (int i = 1; i < M->nn; i++)
        for (int j = 0; j < M->dim; j++) {
            double v = M->x[i * 3 + j];
            if (v < lbc[j]) lbc[j] = v;
            if (v > rtc[j]) rtc[j] = v;
        }
    for (int i = 0; i < M->dim; i++) {
        lbc[i] = lbc[i] - 1.0E-14 * fabs(lbc[i]);
        rtc[i] = rtc[i] + 1.0E-14 * fabs(rtc[i]);
    }
    memcpy(M->lbc, lbc, sizeof(lbc));
    memcpy(M->rtc, rtc, sizeof(rtc));
}

static int block_of(const om_model *M, const double *xc) {
    int ib[3];
    for (int i = 0; i < 3; i++) ib[i] = (int)((xc[i] - M->lbc[i]) / M->bsize);
    return ib[0] + ib[1] * M->nb[0] + ib[2] * M->nb[0] * M->nb[1];
}

/* pdsolve::setBlock (pdsolve.cpp:441-502) */
static int set_block(om_model *M) {
    M->bsize = 2.5 * M->fac * M->minDelta;
    for (int i = 0; i < 3; i++) M->nb[i] = (int)((M->rtc[i] - M->lbc[i]) / M->bsize + 1);
    if (M->dim == 2) M->nb[2] = 1;
    long long nbt = (long long)M->nb[0] * M->nb[1] * M->nb[2];
    if (nbt <= 0 || nbt > 400000000LL) return OM_ERR_MESH;
    int *eb = xcalloc(M->ne, sizeof(int)), *nbk = xcalloc(M->nn, sizeof(int));
    M->be_off = xcalloc(nbt + 1, sizeof(int));
    M->bn_off = xcalloc(nbt + 1, sizeof(int));
    for (int e = 0; e < M->ne; e++) {
        double xc[3];
        ele_center(M, e, xc);
        eb[e] = block_of(M, xc);
        if (eb[e] < 0 || eb[e] >= nbt) return OM_ERR_MESH;
        M->be_off[eb[e] + 1]++;
    }
    for (int k = 0; k < M->nn; k++) {
        nbk[k] = block_of(M, &M->x[k * 3]);
        if (nbk[k] < 0 || nbk[k] >= nbt) return OM_ERR_MESH;
        M->bn_off[nbk[k] + 1]++;
    }
    for (long long b = 0; b < nbt; b++) {
        M->be_off[b + 1] += M->be_off[b];
        M->bn_off[b + 1] += M->bn_off[b];
    }
    M->be = xcalloc(M->ne, sizeof(int));
    M->bn = xcalloc(M->nn, sizeof(int));
    int *fill = xcalloc(nbt, sizeof(int));
    for (int e = 0; e < M->ne; e++) M->be[M->be_off[eb[e]] + fill[eb[e]]++] = e + 1;
    memset(fill, 0, sizeof(int) * nbt);
    for (int k = 0; k < M->nn; k++) M->bn[M->bn_off[nbk[k]] + fill[nbk[k]]++] = k + 1;
    free(fill); free(eb); free(nbk);
    return OM_OK;
}

/* pdsolve::intersection (pdsolve.cpp:9058-9079) */
static int seg_intersect(const double *a1, const double *a2, const double *b1, const double *b2) {
    if ((a1[0] > a2[0] ? a1[0] : a2[0]) < (b1[0] < b2[0] ? b1[0] : b2[0]) ||
        (a1[1] > a2[1] ? a1[1] : a2[1]) < (b1[1] < b2[1] ? b1[1] : b2[1]) ||
        (b1[0] > b2[0] ? b1[0] : b2[0]) < (a1[0] < a2[0] ? a1[0] : a2[0]) ||
        (b1[1] > b2[1] ? b1[1] : b2[1]) < (a1[1] < a2[1] ? a1[1] : a2[1]))
        return 0;
    if ((((a1[0] - b1[0]) * (b2[1] - b1[1]) - (a1[1] - b1[1]) * (b2[0] - b1[0])) *
         ((a2[0] - b1[0]) * (b2[1] - b1[1]) - (a2[1] - b1[1]) * (b2[0] - b1[0]))) > 0 ||
        (((b1[0] - a1[0]) * (a2[1] - a1[1]) - (b1[1] - a1[1]) * (a2[0] - a1[0])) *
         ((b2[0] - a1[0]) * (a2[1] - a1[1]) - (b2[1] - a1[1]) * (a2[0] - a1[0]))) > 0)
        return 0;
    return 1;
}

/* pdsolve::segmentPlaneIntersection (pdsolve.cpp:1123-1209); crack = 3 points */
static int seg_plane(const double *p1, const double *p2, const double *xN) {
    double nrm[3], v1[3], v2[3];
    for (int i = 0; i < 3; i++) { v1[i] = xN[3 + i] - xN[i]; v2[i] = xN[6 + i] - xN[i]; }
    nrm[0] = v1[1] * v2[2] - v1[2] * v2[1];
    nrm[1] = v1[2] * v2[0] - v1[0] * v2[2];
    nrm[2] = v1[0] * v2[1] - v1[1] * v2[0];
    for (int i = 0; i < 3; i++) { v1[i] = p1[i] - xN[i]; v2[i] = p2[i] - xN[i]; }
    double t1 = v1[0] * nrm[0] + v1[1] * nrm[1] + v1[2] * nrm[2];
    double t2 = v2[0] * nrm[0] + v2[1] * nrm[1] + v2[2] * nrm[2];
    if (t1 * t2 > 0) return 0;
    if (t1 * t2 == 0) return 1;
    double a[3], b[3];
    for (int i = 0; i < 3; i++) { a[i] = xN[i] - p1[i]; b[i] = p2[i] - p1[i]; }
    double d = (a[0] * nrm[0] + a[1] * nrm[1] + a[2] * nrm[2]) / (b[0] * nrm[0] + b[1] * nrm[1] + b[2] * nrm[2]);
    double X[3];
    for (int i = 0; i < 3; i++) X[i] = p1[i] + d * b[i];
    double A = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    double n2[3];
#define CROSS_NORM(out)                                   \
    n2[0] = v1[1] * v2[2] - v1[2] * v2[1];                \
    n2[1] = v1[2] * v2[0] - v1[0] * v2[2];                \
    n2[2] = v1[0] * v2[1] - v1[1] * v2[0];                \
    out = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
    double A0, A1, A2;
    for (int i = 0; i < 3; i++) { v1[i] = xN[3 + i] - X[i]; v2[i] = xN[6 + i] - X[i]; }
    CROSS_NORM(A0)
    for (int i = 0; i < 3; i++) { v1[i] = X[i] - xN[i]; v2[i] = xN[6 + i] - xN[i]; }
    CROSS_NORM(A1)
    for (int i = 0; i < 3; i++) { v1[i] = xN[3 + i] - xN[i]; v2[i] = X[i] - xN[i]; }
    CROSS_NORM(A2)
#undef CROSS_NORM
    return fabs(A0 + A1 + A2 - A) > A * 1.0e-14 ? 0 : 1;
}

/* pdsolve::b_eleIntersectCrack (pdsolve.cpp:839-930) */
static int ele_cut(const om_model *M, int e, const double *crack) {
    int n = M->nen[e];
    const int *cn = &M->econ[e * 8];
    if (M->dim == 2) {
        for (int i = 0; i < n; i++) {
            int a = cn[i] - 1, b = cn[(i + 1) % n] - 1;
            if (seg_intersect(&M->x[a * 3], &M->x[b * 3], &crack[0], &crack[3])) return 1;
        }
        return 0;
    }
    int edge[12][2], ne = 0;
    if (n == 8) {
        ne = 12;
        for (int i = 0; i < 3; i++) { edge[i][0] = i; edge[i][1] = i + 1; edge[i + 4][0] = i + 4; edge[i + 4][1] = i + 5; }
        edge[3][0] = 3; edge[3][1] = 0; edge[7][0] = 7; edge[7][1] = 4;
        for (int i = 0; i < 4; i++) { edge[8 + i][0] = i; edge[8 + i][1] = i + 4; }
    } else {
        ne = 6;
        for (int i = 0; i < 3; i++) { edge[i][0] = 3; edge[i][1] = i; }
        edge[3][0] = 0; edge[3][1] = 1; edge[4][0] = 1; edge[4][1] = 2; edge[5][0] = 2; edge[5][1] = 0;
    }
    for (int i = 0; i < ne; i++)
        if (seg_plane(&M->x[(cn[edge[i][0]] - 1) * 3], &M->x[(cn[edge[i][1]] - 1) * 3], crack)) return 1;
    return 0;
}

static void push_int(int **arr, int *n, int *cap, int v) {
    if (*n == *cap) {
        *cap = *cap ? *cap * 2 : 64;
        *arr = realloc(*arr, sizeof(int) * (*cap));
        if (!*arr) abort();
    }
    (*arr)[(*n)++] = v;
}

/* pdsolve::updatePDEbyCrack (pdsolve.cpp:533-727) */
static void pde_by_crack(om_model *M, const double *crack, int *pde_cap) {
    int fn = M->dim == 3 ? 3 : 2;
    double Xmin[3], Xmax[3];
    for (int kk = 0; kk < 3; kk++) {
        Xmin[kk] = crack[kk]; Xmax[kk] = crack[kk];
        for (int p = 1; p < fn; p++) {
            double v = crack[p * 3 + kk];
            if (v < Xmin[kk]) Xmin[kk] = v;
            if (v > Xmax[kk]) Xmax[kk] = v;
        }
    }
    int imin[3], imax[3];
    for (int i = 0; i < 3; i++) {
        imin[i] = (int)((Xmin[i] - M->lbc[i]) / M->bsize);
        imax[i] = (int)((Xmax[i] - M->lbc[i]) / M->bsize);
    }
    if (M->dim == 2) imin[2] = imax[2] = 0;
    int old = M->ncrac_ele;
    for (int xi = imin[0] - 1; xi < imax[0] + 2; xi++)
        for (int yi = imin[1] - 1; yi < imax[1] + 2; yi++)
            for (int zi = imin[2] - 1; zi < imax[2] + 2; zi++) {
                if (!(xi >= 0 && xi < M->nb[0] && yi >= 0 && yi < M->nb[1] && zi >= 0 && zi < M->nb[2])) continue;
                int b = xi + yi * M->nb[0] + zi * M->nb[0] * M->nb[1];
                for (int t = M->be_off[b]; t < M->be_off[b + 1]; t++) {
                    int eid = M->be[t];
                    if (!ele_cut(M, eid - 1, crack)) continue;
                    int dup = 0;
                    for (int q = 0; q < M->ncrac_ele; q++) if (M->crac_ele[q] == eid - 1) { dup = 1; break; }
                    if (!dup) push_int(&M->crac_ele, &M->ncrac_ele, &M->cap_crac_ele, eid - 1);
                }
            }
    int sz = M->ncrac_ele - old;
    if (sz <= 0) return;  /* the reference would take min/max of an empty array here */
    double *X = xcalloc(sz, sizeof(double)), *Y = xcalloc(sz, sizeof(double)), *Z = xcalloc(sz, sizeof(double));
    for (int i = 0; i < sz; i++) {
        double xc[3];
        ele_center(M, M->crac_ele[old + i], xc);
        X[i] = xc[0]; Y[i] = xc[1]; Z[i] = xc[2];
    }
    Xmin[0] = Xmax[0] = X[0]; Xmin[1] = Xmax[1] = Y[0]; Xmin[2] = Xmax[2] = Z[0];
    for (int i = 1; i < sz; i++) {
        if (X[i] < Xmin[0]) Xmin[0] = X[i];
        if (Y[i] < Xmin[1]) Xmin[1] = Y[i];
        if (Z[i] < Xmin[2]) Xmin[2] = Z[i];
        if (X[i] > Xmax[0]) Xmax[0] = X[i];
        if (Y[i] > Xmax[1]) Xmax[1] = Y[i];
        if (Z[i] > Xmax[2]) Xmax[2] = Z[i];
    }
    for (int i = 0; i < 3; i++) {
        imin[i] = (int)((Xmin[i] - M->lbc[i]) / M->bsize);
        imax[i] = (int)((Xmax[i] - M->lbc[i]) / M->bsize);
    }
    double m = M->me + 1e-15;
    double Rsq = M->minDelta * M->minDelta * m * m;
    for (int xi = imin[0] - 1; xi < imax[0] + 2; xi++)
        for (int yi = imin[1] - 1; yi < imax[1] + 2; yi++)
            for (int zi = imin[2] - 1; zi < imax[2] + 2; zi++) {
                if (!(xi >= 0 && xi < M->nb[0] && yi >= 0 && yi < M->nb[1] && zi >= 0 && zi < M->nb[2])) continue;
                int b = xi + yi * M->nb[0] + zi * M->nb[0] * M->nb[1];
                for (int t = M->be_off[b]; t < M->be_off[b + 1]; t++) {
                    int e = M->be[t] - 1;
                    if (M->algo[e] != 2) continue;
                    double xc[3];
                    ele_center(M, e, xc);
                    for (int i = 0; i < sz; i++) {
                        double dis = (xc[0] - X[i]) * (xc[0] - X[i]) + (xc[1] - Y[i]) * (xc[1] - Y[i]) + (xc[2] - Z[i]) * (xc[2] - Z[i]);
                        if (dis < Rsq) {
                            M->algo[e] = 1;
                            push_int(&M->pde, &M->npde, pde_cap, e);
                            break;
                        }
                    }
                }
            }
    free(X); free(Y); free(Z);
}

/* pdsolve::initializePDE_FE (pdsolve.cpp:504-531) */
static void init_pde_fe(om_model *M) {
    int cap = 0;
    M->npde = 0;
    for (int e = 0; e < M->ne; e++)
        if (M->algo[e] == 1) push_int(&M->pde, &M->npde, &cap, e);
    for (int c = 0; c < M->n_crack; c++) pde_by_crack(M, &M->crack[c * 9], &cap);
    int fcap = 0;
    M->nfe = 0;
    for (int e = 0; e < M->ne; e++)
        if (M->algo[e] == 2) push_int(&M->fe, &M->nfe, &fcap, e);
}

/* stable merge sort of rows keyed on column 0 (std::sort in the reference is
 * unstable; ties only change the order of boundary terms) */
static void sort_rows_key0(int *rows, int n, int w) {
    if (n < 2) return;
    int *tmp = xcalloc((size_t)n * w, sizeof(int));
    for (int width = 1; width < n; width *= 2) {
        for (int lo = 0; lo < n; lo += 2 * width) {
            int mid = lo + width < n ? lo + width : n, hi = lo + 2 * width < n ? lo + 2 * width : n;
            int i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (rows[j * w] < rows[i * w]) memcpy(&tmp[(k++) * w], &rows[(j++) * w], sizeof(int) * w);
                else memcpy(&tmp[(k++) * w], &rows[(i++) * w], sizeof(int) * w);
            }
            while (i < mid) memcpy(&tmp[(k++) * w], &rows[(i++) * w], sizeof(int) * w);
            while (j < hi) memcpy(&tmp[(k++) * w], &rows[(j++) * w], sizeof(int) * w);
        }
        memcpy(rows, tmp, sizeof(int) * n * w);
    }
    free(tmp);
}

/* pdsolve::updatePDBE (pdsolve.cpp:1221-1413) + assignPDBE (:1703-1771) */
static void update_pdbe(om_model *M) {
    int w = 5, cnt = 0;
    int cap = M->npde * 6 + 1;
    int *rows = xcalloc((size_t)cap * w, sizeof(int));
    char *dead = NULL;
    if (M->dim == 2) {
        for (int i = 0; i < M->npde; i++) {
            int e = M->pde[i], n = M->nen[e];
            const int *c = &M->econ[e * 8];
            for (int j = 0; j < n; j++) {
                int a = c[j], b = c[(j + 1) % n];
                int *r = &rows[(cnt++) * w];
                if (a < b) { r[0] = a; r[1] = b; r[2] = 1; }
                else { r[1] = a; r[0] = b; r[2] = -1; }
                r[3] = r[4] = 0;
            }
        }
    } else {
        static const int bE[7] = {0, 1, 2, 3, 0, 1, 2};
        static const int tE[5] = {0, 1, 2, 0, 1};
        static const int dbE[6] = {-1, 1, -1, 1, -1, 1};
        for (int i = 0; i < M->npde; i++) {
            int e = M->pde[i];
            const int *c = &M->econ[e * 8];
            if (M->nen[e] == 8) {
                int f[6][4] = {{c[0], c[1], c[2], c[3]}, {c[4], c[5], c[6], c[7]}, {c[0], c[3], c[7], c[4]},
                               {c[1], c[2], c[6], c[5]}, {c[0], c[4], c[5], c[1]}, {c[3], c[7], c[6], c[2]}};
                for (int j = 0; j < 6; j++) {
                    int pos = 0;
                    for (int t = 1; t < 4; t++) if (f[j][t] < f[j][pos]) pos = t;
                    int *r = &rows[(cnt++) * w];
                    r[0] = f[j][bE[pos]]; r[1] = f[j][bE[pos + 1]]; r[2] = f[j][bE[pos + 2]]; r[3] = f[j][bE[pos + 3]];
                    r[4] = dbE[j];
                }
            } else {
                int f[4][3] = {{c[0], c[1], c[3]}, {c[1], c[2], c[3]}, {c[2], c[0], c[3]}, {c[0], c[2], c[1]}};
                for (int j = 0; j < 4; j++) {
                    int pos = 0;
                    for (int t = 1; t < 3; t++) if (f[j][t] < f[j][pos]) pos = t;
                    int *r = &rows[(cnt++) * w];
                    r[0] = f[j][tE[pos]]; r[1] = f[j][tE[pos + 1]]; r[2] = f[j][tE[pos + 2]]; r[3] = f[j][tE[pos + 2]];
                    r[4] = 0;
                }
            }
        }
    }
    sort_rows_key0(rows, cnt, w);
    dead = xcalloc(cnt, 1);
    for (int i = 0; i < cnt; i++) {
        if (dead[i]) continue;
        const int *r1 = &rows[i * w];
        for (int j = i + 1; j < cnt; j++) {
            if (dead[j]) continue;
            const int *r2 = &rows[j * w];
            if (r1[0] != r2[0]) break;
            int hit;
            if (M->dim == 2) hit = (r1[1] == r2[1]);
            else if (r1[2] != r1[3]) hit = (r1[1] == r2[1] && r1[2] == r2[2] && r1[3] == r2[3]);
            else hit = (r2[2] == r2[3] && r1[1] == r2[2] && r1[2] == r2[1]);
            if (hit) { dead[i] = dead[j] = 1; break; }
        }
    }
    int np = 0;
    for (int i = 0; i < cnt; i++) np += !dead[i];
    M->npdbe = np;
    M->pdbe = xcalloc((size_t)np * 4 + 1, sizeof(int));
    int k = 0;
    for (int i = 0; i < cnt; i++) {
        if (dead[i]) continue;
        const int *r = &rows[i * w];
        int *o = &M->pdbe[(k++) * 4];
        if (M->dim == 2) {
            if (r[2] == 1) { o[0] = r[0]; o[1] = r[1]; }
            else { o[0] = r[1]; o[1] = r[0]; }
            o[2] = o[3] = 0;
        } else if (r[2] == r[3]) {
            o[0] = r[0]; o[1] = r[1]; o[2] = r[2]; o[3] = r[2];
        } else if (r[4] == 1) {
            for (int t = 0; t < 4; t++) o[t] = r[t];
        } else {
            for (int t = 0; t < 4; t++) o[3 - t] = r[t];
        }
    }
    free(rows); free(dead);
}

/* pdsolve::UpdateNodeType (pdsolve.cpp:1415-1440) */
static void update_node_type(om_model *M) {
    for (int i = 0; i < M->npde; i++) {
        int e = M->pde[i];
        for (int a = 0; a < M->nen[e]; a++) M->ntype[M->econ[e * 8 + a] - 1] = 2;
    }
    int nper = M->dim == 2 ? 2 : 4;
    for (int p = 0; p < M->npdbe; p++)
        for (int a = 0; a < nper; a++) M->ntype[M->pdbe[p * 4 + a] - 1] = 1;
}

/* pdsolve::initialPDN_Fami (pdsolve.cpp:1442-1568) */
static void init_families(om_model *M) {
    int nfam = 0;
    for (int k = 0; k < M->nn; k++) if (M->ntype[k]) nfam++;
    M->nfam = nfam;
    M->fam_node = xcalloc(nfam, sizeof(int));
    M->fam_v = xcalloc(nfam, sizeof(double));
    M->fam_delta = xcalloc(nfam, sizeof(double));
    nfam = 0;
    for (int k = 0; k < M->nn; k++) if (M->ntype[k]) M->fam_node[nfam++] = k;
    double fac = M->fac;
    for (int k = 0; k < M->nfam; k++) {
        int nd = M->fam_node[k];
        M->fam_delta[k] = fac * pow(M->nv[nd], 1.0 / M->dim);
        M->famid[nd] = k + 1;
    }
    for (int i = 0; i < M->npde; i++) {
        int e = M->pde[i], n = M->nen[e];
        double dv = M->ev[e];
        for (int a = 0; a < n; a++) {
            int f = M->famid[M->econ[e * 8 + a] - 1] - 1;
            M->fam_v[f] = M->fam_v[f] + dv / n;
        }
    }
    /* member search: two passes (count, then fill) over the 27 neighbouring blocks */
    double md2 = M->minDelta * M->minDelta * 1.0e-16;
    /* delta_j = fac * V_j^(1/dim) (pdsolve.cpp:1523), hoisted out of the search loop */
    double *dj_all = xcalloc(M->nn, sizeof(double));
    for (int j = 0; j < M->nn; j++) dj_all[j] = fac * pow(M->nv[j], 1.0 / M->dim);
    M->fam_off = xcalloc(M->nfam + 1, sizeof(int));
    for (int pass = 0; pass < 2; pass++) {
        for (int k = 0; k < M->nfam; k++) {
            int nd = M->fam_node[k], cnt = 0;
            double dk = M->fam_delta[k];
            const double *cx = &M->x[nd * 3];
            int base = M->fam_off[k];
            if (pass) M->fam_mem[base] = nd + 1;
            cnt = 1;
            int ib[3];
            for (int i = 0; i < 3; i++) ib[i] = (int)((cx[i] - M->lbc[i]) / M->bsize);
            for (int xi = ib[0] - 1; xi < ib[0] + 2; xi++)
                for (int yi = ib[1] - 1; yi < ib[1] + 2; yi++)
                    for (int zi = ib[2] - 1; zi < ib[2] + 2; zi++) {
                        if (!(xi >= 0 && xi < M->nb[0] && yi >= 0 && yi < M->nb[1] && zi >= 0 && zi < M->nb[2])) continue;
                        int b = xi + yi * M->nb[0] + zi * M->nb[0] * M->nb[1];
                        for (int t = M->bn_off[b]; t < M->bn_off[b + 1]; t++) {
                            int j = M->bn[t] - 1;
                            const double *mx = &M->x[j * 3];
                            double dj = dj_all[j];
                            double dist = ((mx[0] - cx[0]) * (mx[0] - cx[0]) + (mx[1] - cx[1]) * (mx[1] - cx[1]) +
                                           (mx[2] - cx[2]) * (mx[2] - cx[2]));
                            if (dist > md2 && (dist < (dk * dk + md2) || dist < (dj * dj + md2))) {
                                if (pass) M->fam_mem[base + cnt] = j + 1;
                                cnt++;
                            }
                        }
                    }
            if (!pass) M->fam_off[k + 1] = cnt;
        }
        if (!pass) {
            for (int k = 0; k < M->nfam; k++) M->fam_off[k + 1] += M->fam_off[k];
            M->fam_mem = xcalloc(M->fam_off[M->nfam] + 1, sizeof(int));
        }
    }
    free(dj_all);
    M->bond = xcalloc(M->fam_off[M->nfam] + 1, sizeof(int));
    for (int i = 0; i < M->fam_off[M->nfam]; i++) M->bond[i] = 1;
}

/* pdsolve::updateBondstate (pdsolve.cpp:8910-9004) */
static void update_bondstate(om_model *M, const double *crack) {
    int fn = M->dim == 3 ? 3 : 2;
    double Xmin[3], Xmax[3];
    for (int kk = 0; kk < 3; kk++) {
        Xmin[kk] = crack[kk]; Xmax[kk] = crack[kk];
        for (int p = 1; p < fn; p++) {
            double v = crack[p * 3 + kk];
            if (v < Xmin[kk]) Xmin[kk] = v;
            if (v > Xmax[kk]) Xmax[kk] = v;
        }
    }
    int imin[3], imax[3];
    for (int i = 0; i < 3; i++) {
        imin[i] = (int)((Xmin[i] - M->lbc[i]) / M->bsize);
        imax[i] = (int)((Xmax[i] - M->lbc[i]) / M->bsize);
    }
    if (M->dim == 2) imin[2] = imax[2] = 0;
    for (int xi = imin[0] - 1; xi < imax[0] + 2; xi++)
        for (int yi = imin[1] - 1; yi < imax[1] + 2; yi++)
            for (int zi = imin[2] - 1; zi < imax[2] + 2; zi++) {
                if (!(xi >= 0 && xi < M->nb[0] && yi >= 0 && yi < M->nb[1] && zi >= 0 && zi < M->nb[2])) continue;
                int b = xi + yi * M->nb[0] + zi * M->nb[0] * M->nb[1];
                for (int t = M->bn_off[b]; t < M->bn_off[b + 1]; t++) {
                    int f = M->famid[M->bn[t] - 1];
                    if (f == -1) continue;
                    f -= 1;
                    int nk = M->fam_mem[M->fam_off[f]] - 1;
                    const double *xk = &M->x[nk * 3];
                    for (int m = 1; m < M->fam_off[f + 1] - M->fam_off[f]; m++) {
                        int idx = M->fam_off[f] + m;
                        if (M->bond[idx] != 1) continue;
                        const double *xm = &M->x[(M->fam_mem[idx] - 1) * 3];
                        int cross = M->dim == 3 ? seg_plane(xk, xm, crack) : seg_intersect(xk, xm, &crack[0], &crack[3]);
                        if (cross) M->bond[idx] = 0;
                    }
                }
            }
}

/* pdsolve::setCSRIndexes_gloStiffMat (pdsolve.cpp:959-1121) */
static int icmp(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static void set_csr(om_model *M) {
    int nn = M->nn, d = M->dim;
    int *cnt = xcalloc(nn + 1, sizeof(int));
    for (int f = 0; f < M->nfam; f++) cnt[M->fam_node[f]] += M->fam_off[f + 1] - M->fam_off[f];
    for (int i = 0; i < M->nfe; i++) {
        int e = M->fe[i], n = M->nen[e];
        for (int a = 0; a < n; a++) cnt[M->econ[e * 8 + a] - 1] += n;
    }
    int *off = xcalloc(nn + 1, sizeof(int));
    for (int k = 0; k < nn; k++) off[k + 1] = off[k] + cnt[k];
    int *adj = xcalloc((size_t)off[nn] + 1, sizeof(int)), *fill = xcalloc(nn, sizeof(int));
    for (int f = 0; f < M->nfam; f++) {
        int k = M->fam_node[f];
        for (int t = M->fam_off[f]; t < M->fam_off[f + 1]; t++) adj[off[k] + fill[k]++] = M->fam_mem[t];
    }
    for (int i = 0; i < M->nfe; i++) {
        int e = M->fe[i], n = M->nen[e];
        for (int a = 0; a < n; a++) {
            int k = M->econ[e * 8 + a] - 1;
            for (int b = 0; b < n; b++) adj[off[k] + fill[k]++] = M->econ[e * 8 + b];
        }
    }
    /* unique + sort per node */
    int *uoff = xcalloc(nn + 1, sizeof(int));
    for (int k = 0; k < nn; k++) {
        int *a = &adj[off[k]], n = cnt[k];
        qsort(a, n, sizeof(int), icmp);
        int u = 0;
        for (int i = 0; i < n; i++) if (i == 0 || a[i] != a[i - 1]) a[u++] = a[i];
        uoff[k] = u;
    }
    M->ia = xcalloc(M->neq + 1, sizeof(long long));
    long long nz = 0;
    for (int k = 0; k < nn; k++)
        for (int i = 0; i < d; i++) {
            int r = M->eq[k * 3 + i];
            if (r == -1) continue;
            for (int t = 0; t < uoff[k]; t++)
                for (int j = 0; j < d; j++)
                    if (M->eq[(adj[off[k] + t] - 1) * 3 + j] != -1) { M->ia[r + 1]++; nz++; }
        }
    for (int i = 0; i < M->neq; i++) M->ia[i + 1] = M->ia[i] + M->ia[i + 1];
    M->nnz = nz;
    M->ja = xcalloc(nz + 1, sizeof(long long));
    long long p = 0;
    for (int k = 0; k < nn; k++)
        for (int i = 0; i < d; i++) {
            int r = M->eq[k * 3 + i];
            if (r == -1) continue;
            for (int t = 0; t < uoff[k]; t++)
                for (int j = 0; j < d; j++) {
                    int c = M->eq[(adj[off[k] + t] - 1) * 3 + j];
                    if (c != -1) M->ja[p++] = c;
                }
        }
    free(cnt); free(off); free(adj); free(fill); free(uoff);
}

/* ------------------------------------------------------------------------ */
/* public: creation and setup                                               */
/* ------------------------------------------------------------------------ */

#define DUPI(dst, src, n) do { dst = xcalloc((n) + 1, sizeof(int)); if (n) memcpy(dst, src, sizeof(int) * (size_t)(n)); } while (0)
#define DUPD(dst, src, n) do { dst = xcalloc((n) + 1, sizeof(double)); if (n) memcpy(dst, src, sizeof(double) * (size_t)(n)); } while (0)

/* Build the model from raw mesh arrays (what datModel::readdata, datModel.cpp:58-326,
 * plus the SETSOLVING/EBC/NBC/VEBC/VNBC commands, fioFile.cpp:69-462, produce) and run
 * pdsolve::setDatModel (pdsolve.cpp:114-151).  Returns NULL on error (*err set). */
om_model *om_create(int dim, int proType, double E, double nu, double rho,
                    int nn, const double *x, int ne, const int *econ, const int *nen, const int *algo,
                    int n_ebc, const int *ebc_off, const int *ebc_nid, const int *ebc_dof,
                    const double *ebc_val, const double *ebc_vel, const int *ebc_vary,
                    int n_nbc, const int *nbc_off, const int *nbc_con, const double *nbc_val,
                    const double *nbc_vel, const int *nbc_vary,
                    int n_crack, const double *crack,
                    double fac, double nlf, double me, double dt, double beta, double gamma, int *err) {
    *err = OM_OK;
    if ((dim != 2 && dim != 3) || nn <= 0 || ne <= 0) { *err = OM_ERR_ARG; return NULL; }
    om_model *M = xcalloc(1, sizeof(om_model));
    M->dim = dim; M->proType = proType; M->E = E; M->nu = nu; M->rho = rho;
    M->nn = nn; M->ne = ne;
    DUPD(M->x, x, nn * 3);
    DUPI(M->econ, econ, ne * 8);
    DUPI(M->nen, nen, ne);
    DUPI(M->algo, algo, ne);
    M->n_ebc = n_ebc;
    DUPI(M->ebc_off, ebc_off, n_ebc + 1);
    DUPI(M->ebc_nid, ebc_nid, n_ebc ? ebc_off[n_ebc] : 0);
    DUPI(M->ebc_dof, ebc_dof, n_ebc);
    DUPD(M->ebc_val, ebc_val, n_ebc);
    DUPD(M->ebc_vel, ebc_vel, n_ebc);
    DUPI(M->ebc_vary, ebc_vary, n_ebc);
    M->n_nbc = n_nbc; M->tn = dim == 3 ? 4 : 2;
    DUPI(M->nbc_off, nbc_off, n_nbc + 1);
    DUPI(M->nbc_con, nbc_con, n_nbc ? nbc_off[n_nbc] * M->tn : 0);
    DUPD(M->nbc_val, nbc_val, n_nbc);
    DUPD(M->nbc_vel, nbc_vel, n_nbc);
    DUPI(M->nbc_vary, nbc_vary, n_nbc);
    M->n_crack = n_crack;
    DUPD(M->crack, crack, n_crack * 9);
    M->fac = fac; M->nlf = nlf; M->me = me; M->dt = dt; M->beta = beta; M->gamma = gamma;
    M->active = xcalloc((size_t)nn * 3, sizeof(int));
    M->val = xcalloc((size_t)nn * 3, sizeof(double));
    M->eq = xcalloc((size_t)nn * 3, sizeof(int));
    M->ntype = xcalloc(nn, sizeof(int));
    M->famid = xcalloc(nn, sizeof(int));
    for (int k = 0; k < nn * 3; k++) M->active[k] = 1;
    for (int k = 0; k < nn; k++) M->famid[k] = -1;
    /* datModel::readdata EBC part (datModel.cpp:186-216) */
    for (int s = 0; s < n_ebc; s++)
        for (int t = ebc_off[s]; t < ebc_off[s + 1]; t++) {
            int k = ebc_nid[t] - 1, dof = ebc_dof[s];
            if (k < 0 || k >= nn || dof < 0 || dof >= dim || !M->active[k * 3 + dof]) { *err = OM_ERR_MESH; goto fail; }
            M->val[k * 3 + dof] = ebc_val[s];
            M->active[k * 3 + dof] = 0;
        }
    if ((*err = set_D(M))) goto fail;
    setdof_index(M);
    set_volumes(M);
    domain_dimen(M);
    if ((*err = set_block(M))) goto fail;
    init_pde_fe(M);
    update_pdbe(M);
    update_node_type(M);
    init_families(M);
    for (int c = 0; c < n_crack; c++) update_bondstate(M, &M->crack[c * 9]);
    set_csr(M);
    return M;
fail:
    return NULL;  /* leaked on error paths; test code only */
}

void om_destroy(om_model *M) {
    if (!M) return;
    void *ptrs[] = {M->x, M->econ, M->nen, M->algo, M->ebc_off, M->ebc_nid, M->ebc_dof, M->ebc_val, M->ebc_vel,
                    M->ebc_vary, M->nbc_off, M->nbc_con, M->nbc_val, M->nbc_vel, M->nbc_vary, M->crack, M->active,
                    M->val, M->eq, M->ev, M->nv, M->bn_off, M->bn, M->be_off, M->be, M->pde, M->fe, M->crac_ele,
                    M->pdbe, M->ntype, M->famid, M->fam_node, M->fam_off, M->fam_mem, M->bond, M->fam_v,
                    M->fam_delta, M->ia, M->ja};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(M);
}

/* sizes: [neq, nnz, nfam, total members, npdbe, npde, nfe, nb0, nb1, nb2] */
void om_sizes(const om_model *M, long long *out) {
    out[0] = M->neq; out[1] = M->nnz; out[2] = M->nfam; out[3] = M->fam_off[M->nfam];
    out[4] = M->npdbe; out[5] = M->npde; out[6] = M->nfe;
    out[7] = M->nb[0]; out[8] = M->nb[1]; out[9] = M->nb[2];
}

void om_get_csr(const om_model *M, long long *ia, long long *ja) {
    memcpy(ia, M->ia, sizeof(long long) * (M->neq + 1));
    memcpy(ja, M->ja, sizeof(long long) * M->nnz);
}

void om_get_families(const om_model *M, int *fam_node, int *fam_off, int *fam_mem, int *bond,
                     double *fam_v, double *fam_delta) {
    memcpy(fam_node, M->fam_node, sizeof(int) * M->nfam);
    memcpy(fam_off, M->fam_off, sizeof(int) * (M->nfam + 1));
    memcpy(fam_mem, M->fam_mem, sizeof(int) * M->fam_off[M->nfam]);
    memcpy(bond, M->bond, sizeof(int) * M->fam_off[M->nfam]);
    memcpy(fam_v, M->fam_v, sizeof(double) * M->nfam);
    memcpy(fam_delta, M->fam_delta, sizeof(double) * M->nfam);
}

void om_get_nodes(const om_model *M, int *eq, int *ntype, double *nv, double *ev, int *pdbe, int *algo) {
    memcpy(eq, M->eq, sizeof(int) * M->nn * 3);
    memcpy(ntype, M->ntype, sizeof(int) * M->nn);
    memcpy(nv, M->nv, sizeof(double) * M->nn);
    memcpy(ev, M->ev, sizeof(double) * M->ne);
    memcpy(pdbe, M->pdbe, sizeof(int) * M->npdbe * 4);
    memcpy(algo, M->algo, sizeof(int) * M->ne);
}

void om_get_geom(const om_model *M, double *out) {
    out[0] = M->minDelta; out[1] = M->maxDelta; out[2] = M->bsize;
    for (int i = 0; i < 3; i++) { out[3 + i] = M->lbc[i]; out[6 + i] = M->rtc[i]; }
    out[9] = M->lambda; out[10] = M->mu;
}

/* ------------------------------------------------------------------------ */
/* PDLSM family kernel (pdsolve.cpp:1929-2416, :3212-3398)                  */
/* ------------------------------------------------------------------------ */

/* pdsolve::inflFunc (pdsolve.cpp:1929-1935) */
static double infl(const om_model *M, const double xi[3], double delt) {
    double Aa = M->nlf * M->nlf;
    return exp(-(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]) / Aa / delt / delt);
}

static void fam_xi(const om_model *M, int f, int m, double xi[3]) {
    int k = M->fam_mem[M->fam_off[f]] - 1, j = M->fam_mem[M->fam_off[f] + m] - 1;
    for (int i = 0; i < 3; i++) xi[i] = M->x[j * 3 + i] - M->x[k * 3 + i];
}

/* one member's term of pdsolve::shapTens2D / shapTens3D (pdsolve.cpp:1938-2064),
 * in the reference's expression order for every entry */
static void shape_term(int dim, const double xi[3], double om, double dv, double *A) {
    double x = xi[0], y = xi[1], z = xi[2];
    if (dim == 2) {
        double t[25] = {
            om * x * x * dv, om * dv * x * y, om * dv * x * x * x * 0.5, om * dv * x * y * y * 0.5, om * dv * x * x * y,
            om * dv * x * y, om * dv * y * y, om * dv * x * x * y * 0.5, om * dv * y * y * y * 0.5, om * dv * x * y * y,
            om * dv * x * x * x, om * dv * x * x * y, om * dv * x * x * x * x * 0.5, om * dv * x * x * y * y * 0.5, om * dv * x * x * x * y,
            om * dv * x * y * y, om * dv * y * y * y, om * dv * x * x * y * y * 0.5, om * dv * y * y * y * y * 0.5, om * dv * x * y * y * y,
            om * dv * x * x * y, om * dv * x * y * y, om * dv * x * x * x * y * 0.5, om * dv * x * y * y * y * 0.5, om * dv * x * x * y * y};
        for (int e = 0; e < 25; e++) A[e] = t[e] + A[e];
    } else {
        double t[81] = {
            x * x, x * y, x * z, (x * x * x) * 0.5, (x * y * y) * 0.5, (x * z * z) * 0.5, x * x * y, x * y * z, x * x * z,
            x * y, y * y, y * z, (x * x * y) * 0.5, (y * y * y) * 0.5, (y * z * z) * 0.5, x * y * y, y * y * z, x * y * z,
            x * z, y * z, z * z, (x * x * z) * 0.5, (y * y * z) * 0.5, (z * z * z) * 0.5, x * y * z, y * z * z, x * z * z,
            x * x * x, x * x * y, x * x * z, (x * x * x * x) * 0.5, (x * x * y * y) * 0.5, (x * x * z * z) * 0.5, x * x * x * y, x * x * y * z, x * x * x * z,
            x * y * y, y * y * y, y * y * z, (x * x * y * y) * 0.5, (y * y * y * y) * 0.5, (y * y * z * z) * 0.5, x * y * y * y, y * y * y * z, x * y * y * z,
            x * z * z, y * z * z, z * z * z, (x * x * z * z) * 0.5, (y * y * z * z) * 0.5, (z * z * z * z) * 0.5, x * y * z * z, y * z * z * z, x * z * z * z,
            x * x * y, x * y * y, x * y * z, (x * x * x * y) * 0.5, (x * y * y * y) * 0.5, (x * y * z * z) * 0.5, x * x * y * y, x * y * y * z, x * x * y * z,
            x * y * z, y * y * z, y * z * z, (x * x * y * z) * 0.5, (y * y * y * z) * 0.5, (y * z * z * z) * 0.5, x * y * y * z, y * y * z * z, x * y * z * z,
            x * x * z, x * y * z, x * z * z, (x * x * x * z) * 0.5, (x * y * y * z) * 0.5, (x * z * z * z) * 0.5, x * x * y * z, x * y * z * z, x * x * z * z};
        for (int e = 0; e < 81; e++) A[e] = om * dv * t[e] + A[e];
    }
}

/* pdsolve::shapTens2D / shapTens3D (pdsolve.cpp:1938-2064) */
static void shape_tensor(const om_model *M, int f, double *A) {
    int n = M->fam_off[f + 1] - M->fam_off[f];
    double delt = M->fam_delta[f];
    int p = M->dim == 2 ? 5 : 9;
    memset(A, 0, sizeof(double) * p * p);
    for (int j = 1; j < n; j++) {
        double xi[3];
        fam_xi(M, f, j, xi);
        double dv = M->nv[M->fam_mem[M->fam_off[f] + j] - 1];
        shape_term(M->dim, xi, infl(M, xi, delt), dv, A);
    }
}

/* pdsolve::vec_gd2D / vec_gd3D (pdsolve.cpp:2174-2232): one dgesv per bond */
static int vec_gd(const om_model *M, const double *A, const double xi[3], double *g, double *d) {
    double b[9], s[9];
    if (M->dim == 2) {
        b[0] = xi[0]; b[1] = xi[1]; b[2] = xi[0] * xi[0]; b[3] = xi[1] * xi[1]; b[4] = xi[0] * xi[1];
        if (lu_solve(5, A, b, s)) return OM_ERR_SINGULAR;
        g[0] = s[0]; g[1] = s[1]; d[0] = s[2]; d[1] = s[3]; d[2] = s[4];
    } else {
        b[0] = xi[0]; b[1] = xi[1]; b[2] = xi[2];
        b[3] = xi[0] * xi[0]; b[4] = xi[1] * xi[1]; b[5] = xi[2] * xi[2];
        b[6] = xi[0] * xi[1]; b[7] = xi[1] * xi[2]; b[8] = xi[2] * xi[0];
        if (lu_solve(9, A, b, s)) return OM_ERR_SINGULAR;
        for (int i = 0; i < 3; i++) g[i] = s[i];
        for (int i = 0; i < 6; i++) d[i] = s[3 + i];
    }
    return OM_OK;
}

/* pdsolve::matG2D / matG3D (pdsolve.cpp:2277-2347): bond stiffness from the d-weights */
static void bond_G(int d, double lam, double mu, double om, double mukm, const double *dd, double *G) {
    if (d == 2) {
        double t = mukm * om * ((2 * mu + lam) * dd[0] + mu * dd[1]);
        G[0] = t;
        t = mukm * om * (lam + mu) * dd[2];
        G[1] = t; G[2] = t;
        G[3] = mukm * om * ((2 * mu + lam) * dd[1] + mu * dd[0]);
    } else {
        double tr = dd[0] + dd[1] + dd[2], t;
        G[0] = mukm * om * ((mu + lam) * dd[0] + mu * tr);
        t = mukm * om * (lam + mu) * dd[3]; G[1] = t; G[3] = t;
        t = mukm * om * (lam + mu) * dd[5]; G[2] = t; G[6] = t;
        G[4] = mukm * om * ((mu + lam) * dd[1] + mu * tr);
        t = mukm * om * (lam + mu) * dd[4]; G[5] = t; G[7] = t;
        G[8] = mukm * om * ((mu + lam) * dd[2] + mu * tr);
    }
}

/* pdsolve::matH2D / matH3D (pdsolve.cpp:2349-2416) with matG2D/3D;
 * H is d x (d*n) row-major */
static int mat_H(const om_model *M, int f, double *H) {
    int n = M->fam_off[f + 1] - M->fam_off[f], d = M->dim, w = d * n;
    double A[81];
    memset(H, 0, sizeof(double) * d * w);
    shape_tensor(M, f, A);
    for (int m = 1; m < n; m++) {
        double xi[3], g[3], dd[6], G[9];
        int idx = M->fam_off[f] + m;
        double dvm = M->nv[M->fam_mem[idx] - 1];
        fam_xi(M, f, m, xi);
        double mukm = M->bond[idx];
        if (vec_gd(M, A, xi, g, dd)) return OM_ERR_SINGULAR;
        bond_G(d, M->lambda, M->mu, infl(M, xi, M->fam_delta[f]), mukm, dd, G);
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) {
                double t = G[i * d + j] * dvm;
                H[i * w + j] = -t + H[i * w + j];
                H[i * w + j + d * m] = t;
            }
    }
    return OM_OK;
}

/* pdsolve::matC2D / matC3D (pdsolve.cpp:3212-3252, :3324-3398); C is nv x (d*n) */
static int mat_C(const om_model *M, int f, double *C) {
    int n = M->fam_off[f + 1] - M->fam_off[f], d = M->dim, w = d * n, nv = d == 2 ? 3 : 6;
    double A[81];
    memset(C, 0, sizeof(double) * nv * w);
    shape_tensor(M, f, A);
    for (int m = 1; m < n; m++) {
        double xi[3], g[3], dd[6];
        int idx = M->fam_off[f] + m;
        double dvm = M->nv[M->fam_mem[idx] - 1];
        fam_xi(M, f, m, xi);
        double mukm = M->bond[idx];
        double om = infl(M, xi, M->fam_delta[f]);
        if (vec_gd(M, A, xi, g, dd)) return OM_ERR_SINGULAR;
        if (d == 2) {
            double t = mukm * om * g[0] * dvm;
            C[0 * w + 2 * m] = t; C[2 * w + 2 * m + 1] = t;
            C[0 * w + 0] = -t + C[0 * w + 0]; C[2 * w + 1] = -t + C[2 * w + 1];
            t = mukm * om * g[1] * dvm;
            C[1 * w + 2 * m + 1] = t; C[2 * w + 2 * m] = t;
            C[1 * w + 1] = -t + C[1 * w + 1]; C[2 * w + 0] = -t + C[2 * w + 0];
        } else {
            double c[3];
            for (int i = 0; i < 3; i++) c[i] = mukm * om * g[i] * dvm;
            C[0 * w + 3 * m] = c[0]; C[1 * w + 3 * m + 1] = c[1]; C[2 * w + 3 * m + 2] = c[2];
            C[3 * w + 3 * m] = c[1]; C[3 * w + 3 * m + 1] = c[0];
            C[4 * w + 3 * m + 1] = c[2]; C[4 * w + 3 * m + 2] = c[1];
            C[5 * w + 3 * m] = c[2]; C[5 * w + 3 * m + 2] = c[0];
            C[0 * w + 0] = -c[0] + C[0 * w + 0]; C[1 * w + 1] = -c[1] + C[1 * w + 1]; C[2 * w + 2] = -c[2] + C[2 * w + 2];
            C[3 * w + 0] = -c[1] + C[3 * w + 0]; C[3 * w + 1] = -c[0] + C[3 * w + 1];
            C[4 * w + 1] = -c[2] + C[4 * w + 1]; C[4 * w + 2] = -c[1] + C[4 * w + 2];
            C[5 * w + 0] = -c[2] + C[5 * w + 0]; C[5 * w + 2] = -c[0] + C[5 * w + 2];
        }
    }
    return OM_OK;
}

/* ------------------------------------------------------------------------ */
/* assembly (pdsolve.cpp:2524-2915, :4049-4253, :4882-5071)                  */
/* ------------------------------------------------------------------------ */

/* pdsolve::findCSRIndexOfMat (pdsolve.cpp:1773-1802) */
static long long find_csr(const om_model *M, int row, int col) {
    long long lo = M->ia[row], hi = M->ia[row + 1];
    while (lo < hi) {
        long long mid = (lo + hi) / 2;
        if (col == M->ja[mid]) return mid;
        else if (col < M->ja[mid]) hi = mid;
        else lo = mid + 1;
    }
    return -1;
}

typedef struct {
    int mode;          /* 1 static, 2 quasi-static, 0 dynamic */
    int newmark;
    const double *UN;  /* [nn*d] for Newmark */
    double *Ku, *F;
    int err;
} scat_ctx;

/* the per-entry branch shared by every assembler (e.g. pdsolve.cpp:2563-2621) */
static void scatter(const om_model *M, scat_ctx *S, int row, int node_c, int j, double temp) {
    int col = M->eq[node_c * 3 + j];
    double u = M->val[node_c * 3 + j];
    if (S->mode == 1) {
        if (col != -1) {
            long long s = find_csr(M, row, col);
            if (s < 0) { S->err = OM_ERR_PATTERN; return; }
            S->Ku[s] = S->Ku[s] + temp;
        } else S->F[row] = S->F[row] - temp * u;
    } else if (S->mode == 2) {
        S->F[row] = S->F[row] - temp * u;
        if (col != -1) {
            long long s = find_csr(M, row, col);
            if (s < 0) { S->err = OM_ERR_PATTERN; return; }
            S->Ku[s] = S->Ku[s] + temp;
        }
    } else if (!S->newmark) {
        S->F[row] = S->F[row] - temp * u;
    } else {
        u = S->UN[node_c * M->dim + j];
        S->F[row] = S->F[row] - temp * u;
        if (col != -1) {
            long long s = find_csr(M, row, col);
            if (s < 0) { S->err = OM_ERR_PATTERN; return; }
            S->Ku[s] = S->Ku[s] + temp;
        }
    }
}

/* pdsolve::assembleInterWorkPD_CSRformat (pdsolve.cpp:2524-2730) */
static int assemble_pd_range(const om_model *M, scat_ctx *S, int f0, int f1) {
    int d = M->dim;
    for (int f = f0; f < f1; f++) {
        int n = M->fam_off[f + 1] - M->fam_off[f], w = d * n;
        int k = M->fam_node[f];
        double dvk = M->fam_v[f];
        double *H = xcalloc((size_t)d * w, sizeof(double));
        int e = mat_H(M, f, H);
        if (e) { free(H); return e; }
        for (int i = 0; i < d; i++) {
            int row = M->eq[k * 3 + i];
            if (row == -1) continue;
            for (int m = 0; m < n; m++) {
                int nm = M->fam_mem[M->fam_off[f] + m] - 1;
                for (int j = 0; j < d; j++) scatter(M, S, row, nm, j, -(H[i * w + d * m + j] * dvk));
            }
        }
        free(H);
        if (S->err) return S->err;
    }
    return OM_OK;
}

/* pdsolve::normMatArea (pdsolve.cpp:2917-2967) */
static int norm_mat_area(const int con[4], double xN[4][3], double *N /*3x6*/) {
    double v1[3], v2[3], w = 1.0 / 3;
    for (int i = 0; i < 3; i++) { v1[i] = xN[1][i] - xN[0][i]; v2[i] = xN[2][i] - xN[0][i]; }
    double N1 = v1[1] * v2[2] - v1[2] * v2[1];
    double N2 = v2[0] * v1[2] - v1[0] * v2[2];
    double N3 = v1[0] * v2[1] - v2[0] * v1[1];
    double A1 = sqrt(N1 * N1 + N2 * N2 + N3 * N3), A2 = 0;
    if (A1 == 0) return OM_ERR_MESH;
    if (con[3] != con[2]) {
        for (int i = 0; i < 3; i++) { v1[i] = xN[3][i] - xN[0][i]; v2[i] = xN[2][i] - xN[0][i]; }
        double t1 = v1[1] * v2[2] - v1[2] * v2[1];
        double t2 = v2[0] * v1[2] - v1[0] * v2[2];
        double t3 = v1[0] * v2[1] - v2[0] * v1[1];
        A2 = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
        w = 0.25;
    }
    double area = (A1 + A2) * 0.5;
    N1 = N1 / A1 * area * w;
    N2 = N2 / A1 * area * w;
    N3 = N3 / A1 * area * w;
    memset(N, 0, sizeof(double) * 18);
    N[0 * 6 + 0] = N1; N[1 * 6 + 1] = N2; N[2 * 6 + 2] = N3;
    N[0 * 6 + 3] = N2; N[0 * 6 + 5] = N3; N[1 * 6 + 3] = N1;
    N[1 * 6 + 4] = N3; N[2 * 6 + 4] = N2; N[2 * 6 + 5] = N1;
    return OM_OK;
}

/* pdsolve::assemblePDBEwork_CSRformat 2D part (pdsolve.cpp:4049-4253) and
 * assemblePDBEwork3D_CSRformat (:2732-2915) */
static int assemble_pdbe_range(const om_model *M, scat_ctx *S, int p0, int p1) {
    int d = M->dim;
    for (int p = p0; p < p1; p++) {
        const int *con = &M->pdbe[p * 4];
        if (d == 2) {
            int a = con[0] - 1, b = con[1] - 1;
            if (!M->ntype[a] || !M->ntype[b]) return OM_ERR_MESH;
            double nx = M->x[b * 3 + 1] - M->x[a * 3 + 1];
            double ny = M->x[a * 3 + 0] - M->x[b * 3 + 0];
            double N[6] = {nx, 0, ny, 0, ny, nx};
            for (int side = 0; side < 2; side++) {
                int node = con[side] - 1, f = M->famid[node] - 1;
                int n = M->fam_off[f + 1] - M->fam_off[f], w = 2 * n;
                double *C = xcalloc((size_t)3 * w, sizeof(double)), *DC = xcalloc((size_t)3 * w, sizeof(double));
                double *NDC = xcalloc((size_t)2 * w, sizeof(double));
                int e = mat_C(M, f, C);
                if (e) { free(C); free(DC); free(NDC); return e; }
                matmul(3, 3, w, M->D, C, DC);
                matmul(2, 3, w, N, DC, NDC);
                for (int i = 0; i < 2; i++) {
                    int row = M->eq[node * 3 + i];
                    if (row == -1) continue;
                    for (int m = 0; m < n; m++) {
                        int nm = M->fam_mem[M->fam_off[f] + m] - 1;
                        for (int j = 0; j < 2; j++) scatter(M, S, row, nm, j, 1.0 / 2 * NDC[i * w + 2 * m + j]);
                    }
                }
                free(C); free(DC); free(NDC);
            }
        } else {
            double xN[4][3], Nm[18], ND[18];
            for (int i = 0; i < 4; i++)
                for (int c = 0; c < 3; c++) xN[i][c] = M->x[(con[i] - 1) * 3 + c];
            if (norm_mat_area(con, xN, Nm)) return OM_ERR_MESH;
            matmul(3, 6, 6, Nm, M->D, ND);
            int nn = con[3] == con[2] ? 3 : 4;
            for (int ei = 0; ei < nn; ei++) {
                int node = con[ei] - 1, f = M->famid[node] - 1;
                int n = M->fam_off[f + 1] - M->fam_off[f], w = 3 * n;
                double *C = xcalloc((size_t)6 * w, sizeof(double)), *fin = xcalloc((size_t)3 * w, sizeof(double));
                int e = mat_C(M, f, C);
                if (e) { free(C); free(fin); return e; }
                matmul(3, 6, w, ND, C, fin);
                for (int j = 0; j < 3; j++) {
                    int row = M->eq[node * 3 + j];
                    if (row == -1) continue;
                    for (int m = 0; m < n; m++) {
                        int nm = M->fam_mem[M->fam_off[f] + m] - 1;
                        for (int jj = 0; jj < 3; jj++) scatter(M, S, row, nm, jj, fin[j * w + 3 * m + jj]);
                    }
                }
                free(C); free(fin);
            }
        }
        if (S->err) return S->err;
    }
    return OM_OK;
}

/* pdsolve::assembleSEDbyFEM_CSRformat (pdsolve.cpp:4882-5007) */
static int assemble_fe_range(const om_model *M, scat_ctx *S, int t0, int t1) {
    int d = M->dim;
    double Ke[576];
    for (int t = t0; t < t1; t++) {
        int e = M->fe[t], n = M->nen[e], nd;
        int r = ele_stiff(M, e, Ke, &nd);
        if (r) return r;
        for (int a = 0; a < n; a++) {
            int na = M->econ[e * 8 + a] - 1;
            for (int ii = 0; ii < d; ii++) {
                int row = M->eq[na * 3 + ii];
                if (row == -1) continue;
                for (int b = 0; b < n; b++) {
                    int nb = M->econ[e * 8 + b] - 1;
                    for (int jj = 0; jj < d; jj++) scatter(M, S, row, nb, jj, Ke[(d * a + ii) * nd + d * b + jj]);
                }
            }
        }
        if (S->err) return S->err;
    }
    return OM_OK;
}

/* pdfemEleLine2N::eleEquivNodalForce (pdfemEleLine2N.cpp:45-80) and
 * pdfemEleQuad4N::eleEquivNodalForce + vec_mNtN_NBC (pdfemEleQuad4N.cpp:389-456) */
static void nbc_element_force(const om_model *M, const int *con, double t, double *Fe) {
    if (M->dim == 2) {
        double xN[2][3];
        for (int a = 0; a < 2; a++)
            for (int c = 0; c < 3; c++) xN[a][c] = M->x[(con[a] - 1) * 3 + c];
        for (int k = 0; k < 4; k++) Fe[k] = 0;
        for (int m = 0; m < 2; m++) {
            double p = GP2[m], wp = 1.0, N[2] = {0.5 * (1 - p), 0.5 * (1 + p)}, dNdp[2] = {-0.5, 0.5};
            for (int ii = 0; ii < 2; ii++) {
                double t1 = 0, t2 = 0;
                for (int jj = 0; jj < 2; jj++) { t1 = t1 + dNdp[jj] * xN[jj][1]; t2 = t2 + dNdp[jj] * xN[jj][0]; }
                double v = wp * N[ii] * (t * t1);
                Fe[ii * 2] = v + Fe[ii * 2];
                v = wp * N[ii] * (-t * t2);
                Fe[ii * 2 + 1] = v + Fe[ii * 2 + 1];
            }
        }
    } else {
        double xN[4][3];
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 3; c++) xN[a][c] = M->x[(con[a] - 1) * 3 + c];
        for (int k = 0; k < 12; k++) Fe[k] = 0;
        for (int mp = 0; mp < 2; mp++)
            for (int mq = 0; mq < 2; mq++) {
                double p = GP2[mp], q = GP2[mq], N[4];
                N[0] = 0.25 * (1 - p) * (1 - q); N[1] = 0.25 * (1 + p) * (1 - q);
                N[2] = 0.25 * (1 + p) * (1 + q); N[3] = 0.25 * (1 - p) * (1 + q);
                double dNdp[4] = {0.25 * (-1 + q), 0.25 * (1 - q), 0.25 * (1 + q), 0.25 * (-1 - q)};
                double dNdq[4] = {0.25 * (-1 + p), 0.25 * (-1 - p), 0.25 * (1 + p), 0.25 * (1 - p)};
                double dXdp[3] = {0, 0, 0}, dXdq[3] = {0, 0, 0};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 4; j++) { dXdp[i] = dXdp[i] + dNdp[j] * xN[j][i]; dXdq[i] = dXdq[i] + dNdq[j] * xN[j][i]; }
                double nr[3] = {dXdp[1] * dXdq[2] - dXdp[2] * dXdq[1], dXdq[0] * dXdp[2] - dXdp[0] * dXdq[2],
                                dXdp[0] * dXdq[1] - dXdq[0] * dXdp[1]};
                /* mNt (12x3) * normN: row i+3k has N[k] in column i only */
                for (int k = 0; k < 4; k++)
                    for (int i = 0; i < 3; i++) {
                        double v = 0.0;
                        for (int c = 0; c < 3; c++) v = (c == i ? N[k] : 0.0) * nr[c] + v;
                        Fe[3 * k + i] = Fe[3 * k + i] + v;
                    }
            }
        for (int k = 0; k < 12; k++) Fe[k] = Fe[k] * t;
    }
}

/* pdsolve::calExternalForce_CSRformat (pdsolve.cpp:5009-5071); point BCs are
 * disabled in the reference (datModel.cpp:218) */
static void external_force(const om_model *M, double *F) {
    int d = M->dim, tn = M->tn;
    for (int s = 0; s < M->n_nbc; s++) {
        double t = M->nbc_val[s];
        for (int e = M->nbc_off[s]; e < M->nbc_off[s + 1]; e++) {
            const int *con = &M->nbc_con[e * tn];
            double Fe[12];
            nbc_element_force(M, con, t, Fe);
            for (int a = 0; a < tn; a++)
                for (int ii = 0; ii < d; ii++) {
                    int r = M->eq[(con[a] - 1) * 3 + ii];
                    if (r != -1) F[r] = F[r] + Fe[d * a + ii];
                }
        }
    }
}

/* Assemble K (CSR values on the om pattern) and F.  mode: 1 static (K + F with
 * prescribed columns moved to F), 2 quasi-static, 0 dynamic (explicit: F = -K u
 * only; with newmark=1: F = -K U_N and K assembled).  Ku may be NULL in explicit
 * mode.  F includes the external force when with_ext != 0. */
/* test hook: bit 1 drops the PD-boundary terms, bit 2 the FE terms, bit 4 the PD
 * terms -- used only by tests that check a planted mistake is detected */
static int g_skip = 0;
void om_set_skip(int flags) { g_skip = flags; }

/* Threads of the explicit internal force (om_set_threads; 1 = the serial restatement).
 * With P > 1 threads the explicit force (no K) is computed as the reference's P MPI ranks
 * do (pdsolve.cpp:2529-2531, :4887-4889, :5831-5840): thread r takes the contiguous slices
 * [r N / P, (r + 1) N / P) of the families, the PD-boundary entities and the FE elements
 * into a private F, and the P vectors are summed in rank order (the MPI_Allreduce).  Used
 * for the timed CPU baseline on all cores; the checks use the serial path. */
static int g_threads = 1;
void om_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

static int assemble_pd(const om_model *M, scat_ctx *S) { return assemble_pd_range(M, S, 0, M->nfam); }
static int assemble_pdbe(const om_model *M, scat_ctx *S) { return assemble_pdbe_range(M, S, 0, M->npdbe); }
static int assemble_fe(const om_model *M, scat_ctx *S) { return assemble_fe_range(M, S, 0, M->nfe); }

static int explicit_force_threads(const om_model *M, double *F) {
    const int P = g_threads, neq = M->neq;
    double *part = (double *)xcalloc((size_t)P * (neq > 0 ? neq : 1), sizeof(double));
    int err = OM_OK;
#pragma omp parallel for num_threads(P) schedule(static, 1)
    for (int r = 0; r < P; r++) {
        scat_ctx S = {0, 0, NULL, NULL, part + (size_t)r * neq, 0};
        int e = OM_OK;
        long long P_ = P;
        if (!(g_skip & 4)) e = assemble_pd_range(M, &S, (int)(r * (long long)M->nfam / P_), (int)((r + 1) * (long long)M->nfam / P_));
        if (!e && !(g_skip & 1))
            e = assemble_pdbe_range(M, &S, (int)(r * (long long)M->npdbe / P_), (int)((r + 1) * (long long)M->npdbe / P_));
        if (!e && !(g_skip & 2)) e = assemble_fe_range(M, &S, (int)(r * (long long)M->nfe / P_), (int)((r + 1) * (long long)M->nfe / P_));
        if (e) {
#pragma omp critical
            err = e;
        }
    }
    for (int i = 0; i < neq; i++) {
        double t = 0.0;
        for (int r = 0; r < P; r++) t = t + part[(size_t)r * neq + i];
        F[i] = t;
    }
    free(part);
    return err;
}

int om_assemble(om_model *M, int mode, int newmark, const double *UN, double *Ku, double *F, int with_ext) {
    scat_ctx S = {mode, newmark, UN, Ku, F, 0};
    if (Ku) memset(Ku, 0, sizeof(double) * M->nnz);
    memset(F, 0, sizeof(double) * M->neq);
    if (!Ku && (mode != 0 || newmark)) return OM_ERR_ARG;
    int e;
    if (g_threads > 1 && mode == 0 && !newmark && !Ku) {
        if ((e = explicit_force_threads(M, F))) return e;
    } else {
        if (!(g_skip & 4) && (e = assemble_pd(M, &S))) return e;
        if (!(g_skip & 1) && (e = assemble_pdbe(M, &S))) return e;
        if (!(g_skip & 2) && (e = assemble_fe(M, &S))) return e;
    }
    if (with_ext) external_force(M, F);
    return OM_OK;
}

/* per-family H (d x d*n) for tests: writes family f's H into out */
int om_family_H(om_model *M, int f, double *out) { return mat_H(M, f, out); }
int om_family_C(om_model *M, int f, double *out) { return mat_C(M, f, out); }
int om_family_A(om_model *M, int f, double *out) { shape_tensor(M, f, out); return OM_OK; }
int om_element_K(om_model *M, int e, double *out, int *nd) { return ele_stiff(M, e, out, nd); }
void om_element_M(om_model *M, int e, double *out, int *nd) { ele_mass(M, e, out, nd); }

/* pdsolve::assembleLumpedMass (pdsolve.cpp:4618-4669) */
void om_lumped_mass(const om_model *M, double *m) {
    memset(m, 0, sizeof(double) * M->neq);
    for (int e = 0; e < M->ne; e++) {
        int n = M->nen[e];
        double dv = M->ev[e] / n;
        for (int a = 0; a < n; a++)
            for (int i = 0; i < M->dim; i++) {
                int r = M->eq[(M->econ[e * 8 + a] - 1) * 3 + i];
                if (r != -1) m[r] = m[r] + dv * M->rho;
            }
    }
}

/* consistent FE mass + PD lumped mass on the K pattern
 * (assembleElemassMatFEM_CSRformat pdsolve.cpp:5154-5210, assembleMassMatPD_CSRformat :4568-4595) */
int om_mass_on_K(const om_model *M, double *Mv) {
    int d = M->dim;
    double Me[576];
    memset(Mv, 0, sizeof(double) * M->nnz);
    for (int t = 0; t < M->nfe; t++) {
        int e = M->fe[t], n = M->nen[e], nd;
        ele_mass(M, e, Me, &nd);
        for (int a = 0; a < n; a++)
            for (int ii = 0; ii < d; ii++) {
                int row = M->eq[(M->econ[e * 8 + a] - 1) * 3 + ii];
                if (row == -1) continue;
                for (int b = 0; b < n; b++)
                    for (int jj = 0; jj < d; jj++) {
                        int col = M->eq[(M->econ[e * 8 + b] - 1) * 3 + jj];
                        if (col == -1) continue;
                        long long s = find_csr(M, row, col);
                        if (s < 0) return OM_ERR_PATTERN;
                        Mv[s] = Mv[s] + Me[(d * a + ii) * nd + d * b + jj];
                    }
            }
    }
    for (int f = 0; f < M->nfam; f++) {
        int k = M->fam_node[f];
        double dv = M->fam_v[f];
        for (int i = 0; i < d; i++) {
            int r = M->eq[k * 3 + i];
            if (r == -1) continue;
            long long s = find_csr(M, r, r);
            if (s < 0) return OM_ERR_PATTERN;
            Mv[s] = Mv[s] + dv * M->rho;
        }
    }
    return OM_OK;
}

/* Consistent mass of the explicit driver on its own pattern: pdsolve::setCSRIndexes_gloMassMat
 * (pdsolve.cpp:5674-5817) -- per node, the family's own node (:5698-5703) and the nodes of
 * every FE element (algo 2, element order, :5705-5723), deduplicated and sorted (:5725-5737);
 * rows of pure PD nodes (type 2) hold their diagonal only (:5764-5781), other rows every
 * free column of those nodes (:5742-5762); zero-based int64 ia / ja -- then
 * assembleElemassMatFEM_CSRformat (:5154-5210) and assembleMassMatPD_CSRformat (:4568-4595)
 * with the slot search of findCSRIndexOfMat (:1773-1802) on that pattern.
 * Call with ia == NULL for the size: *nnz_out gets the entry count.  vals may be NULL. */
int om_mass_csr(const om_model *M, long long *ia, long long *ja, double *vals, long long *nnz_out) {
    int nn = M->nn, d = M->dim;
    int *cnt = xcalloc(nn + 1, sizeof(int));
    for (int f = 0; f < M->nfam; f++) cnt[M->fam_node[f]] += 1;
    for (int e = 0; e < M->ne; e++)
        if (M->algo[e] == 2)
            for (int a = 0; a < M->nen[e]; a++) cnt[M->econ[e * 8 + a] - 1] += M->nen[e];
    int *off = xcalloc(nn + 1, sizeof(int));
    for (int k = 0; k < nn; k++) off[k + 1] = off[k] + cnt[k];
    int *adj = xcalloc((size_t)off[nn] + 1, sizeof(int)), *fill = xcalloc(nn + 1, sizeof(int));
    for (int f = 0; f < M->nfam; f++) {
        int k = M->fam_node[f];
        adj[off[k] + fill[k]++] = k + 1;
    }
    for (int e = 0; e < M->ne; e++) {
        if (M->algo[e] != 2) continue;
        for (int a = 0; a < M->nen[e]; a++) {
            int k = M->econ[e * 8 + a] - 1;
            for (int b = 0; b < M->nen[e]; b++) adj[off[k] + fill[k]++] = M->econ[e * 8 + b];
        }
    }
    for (int k = 0; k < nn; k++) {
        int *a = &adj[off[k]], n = cnt[k], u = 0;
        qsort(a, n, sizeof(int), icmp);
        for (int i = 0; i < n; i++) if (i == 0 || a[i] != a[i - 1]) a[u++] = a[i];
        fill[k] = u;
    }
    long long *lia = xcalloc(M->neq + 2, sizeof(long long)), nz = 0;
    for (int pass = 0; pass < 2; pass++) {
        long long p = 0;
        for (int k = 0; k < nn; k++)
            for (int i = 0; i < d; i++) {
                int r = M->eq[k * 3 + i];
                if (r == -1) continue;
                if (M->ntype[k] == 2) {
                    if (pass) ja[p] = r;
                    p++;
                    if (!pass) lia[r + 1]++;
                    continue;
                }
                for (int t = 0; t < fill[k]; t++)
                    for (int j = 0; j < d; j++) {
                        int c = M->eq[(adj[off[k] + t] - 1) * 3 + j];
                        if (c == -1) continue;
                        if (pass) ja[p] = c;
                        p++;
                        if (!pass) lia[r + 1]++;
                    }
            }
        if (!pass) {
            nz = p;
            for (int i = 0; i < M->neq; i++) lia[i + 1] = lia[i] + lia[i + 1];
            if (nnz_out) *nnz_out = nz;
            if (!ia) break;
            memcpy(ia, lia, sizeof(long long) * (M->neq + 1));
        }
    }
    int rc = OM_OK;
    if (ia && vals) {
        double Me[576];
        memset(vals, 0, sizeof(double) * nz);
        for (int t = 0; t < M->nfe && rc == OM_OK; t++) {
            int e = M->fe[t], n = M->nen[e], nd;
            ele_mass(M, e, Me, &nd);
            for (int a = 0; a < n; a++)
                for (int ii = 0; ii < d; ii++) {
                    int row = M->eq[(M->econ[e * 8 + a] - 1) * 3 + ii];
                    if (row == -1) continue;
                    for (int b = 0; b < n; b++)
                        for (int jj = 0; jj < d; jj++) {
                            int col = M->eq[(M->econ[e * 8 + b] - 1) * 3 + jj];
                            if (col == -1) continue;
                            long long lo = ia[row], hi = ia[row + 1], s = -1;
                            while (lo < hi) {
                                long long mid = (lo + hi) / 2;
                                if (col == ja[mid]) { s = mid; break; }
                                else if (col < ja[mid]) hi = mid;
                                else lo = mid + 1;
                            }
                            if (s < 0) { rc = OM_ERR_PATTERN; continue; }
                            vals[s] = vals[s] + Me[(d * a + ii) * nd + d * b + jj];
                        }
                }
        }
        for (int f = 0; f < M->nfam && rc == OM_OK; f++) {
            int k = M->fam_node[f];
            double dv = M->fam_v[f];
            for (int i = 0; i < d; i++) {
                int r = M->eq[k * 3 + i];
                if (r == -1) continue;
                long long lo = ia[r], hi = ia[r + 1], s = -1;
                while (lo < hi) {
                    long long mid = (lo + hi) / 2;
                    if (r == ja[mid]) { s = mid; break; }
                    else if (r < ja[mid]) hi = mid;
                    else lo = mid + 1;
                }
                if (s < 0) { rc = OM_ERR_PATTERN; break; }
                vals[s] = vals[s] + dv * M->rho;
            }
        }
    }
    free(cnt); free(off); free(adj); free(fill); free(lia);
    return rc;
}

/* pdsolve::setVaryEssentialBC (rati = 1, non-quasi-static branch, pdsolve.cpp:5263-5312)
 * and setVaryNaturalBC (:5314-5329) */
void om_advance_bc(om_model *M) {
    for (int s = 0; s < M->n_ebc; s++) {
        if (!M->ebc_vary[s]) continue;
        double U = M->ebc_val[s] + M->dt * M->ebc_vel[s];
        M->ebc_val[s] = U;
        for (int t = M->ebc_off[s]; t < M->ebc_off[s + 1]; t++) M->val[(M->ebc_nid[t] - 1) * 3 + M->ebc_dof[s]] = U;
    }
    for (int s = 0; s < M->n_nbc; s++)
        if (M->nbc_vary[s]) M->nbc_val[s] = M->nbc_val[s] + M->dt * M->nbc_vel[s];
}

/* pdsolve::storeDisplacementResult (pdsolve.cpp:5656-5672) */
void om_store_u(om_model *M, const double *U) {
    for (int k = 0; k < M->nn; k++)
        for (int i = 0; i < M->dim; i++) {
            int r = M->eq[k * 3 + i];
            if (r != -1) M->val[k * 3 + i] = U[r];
        }
}

void om_get_values(const om_model *M, double *out /*[nn*dim]*/) {
    for (int k = 0; k < M->nn; k++)
        for (int i = 0; i < M->dim; i++) out[k * M->dim + i] = M->val[k * 3 + i];
}

/* explicit central difference with lumped mass: pdsolve::pdfemDynamicSolver_CSRformat
 * (pdsolve.cpp:5842-6090) with timeIntegration_LM (:5113-5132), stress recovery and
 * failure (no-ops for FC 0 as far as u is concerned) left out.  Runs nsteps loop
 * iterations after the start-up step; u_out receives u_{nsteps+1} (free DOFs),
 * u_prev receives u_{nsteps}. */
int om_explicit_run(om_model *M, int nsteps, double *u_out, double *u_prev) {
    int neq = M->neq;
    double *m = xcalloc(neq, sizeof(double)), *F = xcalloc(neq, sizeof(double));
    double *un = xcalloc(neq, sizeof(double)), *unm1 = xcalloc(neq, sizeof(double)), *unp1 = xcalloc(neq, sizeof(double));
    om_lumped_mass(M, m);
    double dt = M->dt;
    for (int k = 0; k < M->nn; k++)
        for (int i = 0; i < M->dim; i++) {
            int r = M->eq[k * 3 + i];
            if (r != -1) un[r] = M->val[k * 3 + i];
        }
    om_advance_bc(M);
    int e = om_assemble(M, 0, 0, NULL, NULL, F, 1);
    if (e) goto out;
    for (int i = 0; i < neq; i++) {
        double a = F[i] / m[i];
        unp1[i] = a * (0.5 * dt * dt);
        unp1[i] = unp1[i] + un[i];
    }
    om_store_u(M, unp1);
    for (int n = 1; n < nsteps + 1; n++) {
        memcpy(unm1, un, sizeof(double) * neq);
        memcpy(un, unp1, sizeof(double) * neq);
        om_advance_bc(M);
        if ((e = om_assemble(M, 0, 0, NULL, NULL, F, 1))) goto out;
        for (int i = 0; i < neq; i++) {
            double tA = dt * dt * F[i] / m[i];
            unp1[i] = un[i] * 2.0 + tA - unm1[i];
        }
        om_store_u(M, unp1);
    }
    memcpy(u_out, unp1, sizeof(double) * neq);
    if (u_prev) memcpy(u_prev, un, sizeof(double) * neq);
out:
    free(m); free(F); free(un); free(unm1); free(unp1);
    return e;
}

/* Newmark predictor U_N (pdsolve.cpp:6333-6366) over all nodes, full DOF layout [nn*dim] */
void om_newmark_UN(const om_model *M, const double *v, const double *a, double *UN) {
    int d = M->dim;
    double dt = M->dt, beta = M->beta;
    for (int i = 0; i < M->nn; i++)
        for (int j = 0; j < d; j++) {
            int r = M->eq[i * 3 + j];
            if (r != -1) UN[i * d + j] = M->val[i * 3 + j] + dt * v[r] + dt * dt * (0.5 - beta) * a[r];
            else UN[i * d + j] = M->val[i * 3 + j];
        }
    for (int s = 0; s < M->n_ebc; s++) {
        if (!M->ebc_vary[s]) continue;
        double dU = dt * M->ebc_vel[s];
        for (int t = M->ebc_off[s]; t < M->ebc_off[s + 1]; t++) {
            int k = M->ebc_nid[t] - 1;
            UN[k * d + M->ebc_dof[s]] = UN[k * d + M->ebc_dof[s]] + dU;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* PD family batch: a node cloud in which every node is a family            */
/* ------------------------------------------------------------------------ */

/* Members by the rule of pdsolve::initialPDN_Fami (pdsolve.cpp:1442-1568):
 * j belongs to the family of k iff |x_j - x_k|^2 > md2 and
 * (|x_j - x_k|^2 < delta_k^2 + md2 or |x_j - x_k|^2 < delta_j^2 + md2),
 * md2 = (min delta)^2 * 1e-16; self first, then the others in node order
 * (brute force, for small clouds).  off [n+1]; returns -1 if cap is exceeded. */
int om_fb_members(int dim, int n, const double *x, const double *delta, int *off, int *mem, int cap) {
    double mind = delta[0];
    for (int k = 1; k < n; k++) mind = delta[k] < mind ? delta[k] : mind;
    double md2 = mind * mind * 1.0e-16;
    int c = 0;
    off[0] = 0;
    for (int k = 0; k < n; k++) {
        if (c >= cap) return -1;
        mem[c++] = k;
        for (int j = 0; j < n; j++) {
            double dist = 0.0;
            for (int i = 0; i < 3; i++) dist += (x[j * 3 + i] - x[k * 3 + i]) * (x[j * 3 + i] - x[k * 3 + i]);
            if (dist > md2 && (dist < delta[k] * delta[k] + md2 || dist < delta[j] * delta[j] + md2)) {
                if (c >= cap) return -1;
                mem[c++] = j;
            }
        }
        off[k + 1] = c;
    }
    (void)dim;
    return c;
}

/* Shape tensor A (p x p) and the H blocks of one family: d x d per member,
 * H_m = G_m V_m (matG/matH, pdsolve.cpp:2277-2416, vec_gd :2174-2232 with one
 * dgesv per bond), self block H_0 = -sum_m H_m; all bonds intact. */
int om_fb_family(int dim, const double *x, const double *vol, const double *delta, double nlf, double lam, double mu,
                 int n, const int *mem, double *A, double *H) {
    int p = dim == 2 ? 5 : 9, dd2 = dim * dim, k = mem[0];
    double delt = delta[k], Aa = nlf * nlf;
    memset(A, 0, sizeof(double) * p * p);
    memset(H, 0, sizeof(double) * dd2 * n);
    for (int m = 1; m < n; m++) {
        double xi[3];
        for (int i = 0; i < 3; i++) xi[i] = x[mem[m] * 3 + i] - x[k * 3 + i];
        double om = exp(-(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]) / Aa / delt / delt);
        shape_term(dim, xi, om, vol[mem[m]], A);
    }
    for (int m = 1; m < n; m++) {
        double xi[3], b[9], sl[9], G[9];
        for (int i = 0; i < 3; i++) xi[i] = x[mem[m] * 3 + i] - x[k * 3 + i];
        if (dim == 2) {
            b[0] = xi[0]; b[1] = xi[1]; b[2] = xi[0] * xi[0]; b[3] = xi[1] * xi[1]; b[4] = xi[0] * xi[1];
        } else {
            b[0] = xi[0]; b[1] = xi[1]; b[2] = xi[2];
            b[3] = xi[0] * xi[0]; b[4] = xi[1] * xi[1]; b[5] = xi[2] * xi[2];
            b[6] = xi[0] * xi[1]; b[7] = xi[1] * xi[2]; b[8] = xi[2] * xi[0];
        }
        if (lu_solve(p, A, b, sl)) return OM_ERR_SINGULAR;
        double om = exp(-(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]) / Aa / delt / delt);
        bond_G(dim, lam, mu, om, 1.0, sl + dim, G);
        for (int e = 0; e < dd2; e++) {
            double t = G[e] * vol[mem[m]];
            H[m * dd2 + e] = t;
            H[e] = -t + H[e];
        }
    }
    return OM_OK;
}

/* ------------------------------------------------------------------------ */
/* nodal stresses                                                           */
/* ------------------------------------------------------------------------ */

/* pdsolve::calGlobalNodeStresses (pdsolve.cpp:9540-9604) for the DOF values
 * u_all [nn*dim]: every PD family node first (calPDNodeStresses :9081-9166,
 * sigma = D (C u_family) stored, count 1), then every FE element in FE order
 * (calFEMNodeStresses_EXP :9168-9335 with eleFitStresses flag 1: Gauss-point
 * stresses D B u_e extrapolated to the corner nodes with L, added, count + 1
 * per element node), then sigma / count and, for 2D plane strain,
 * sigma_zz = nu (sigma_xx + sigma_yy) (pdNode::calSigzz).  sigma [nn*6]; the
 * 2D components sit at 0 (xx), 1 (yy), 3 (xy).  A node with count 0 keeps 0
 * (the reference would divide by zero). */
int om_node_stresses(const om_model *M, const double *u_all, double *sigma) {
    int d = M->dim, nv = d == 2 ? 3 : 6, nn = M->nn;
    int *count = xcalloc(nn, sizeof(int));
    memset(sigma, 0, sizeof(double) * 6 * nn);
    for (int f = 0; f < M->nfam; f++) {
        int n = M->fam_off[f + 1] - M->fam_off[f], w = d * n, k = M->fam_mem[M->fam_off[f]] - 1;
        double *C = xcalloc((size_t)nv * w, sizeof(double)), *uk = xcalloc(w, sizeof(double));
        double eps[6], sig[6];
        count[k]++;
        if (mat_C(M, f, C)) { free(C); free(uk); free(count); return OM_ERR_SINGULAR; }
        for (int m = 0; m < n; m++) {
            int j = M->fam_mem[M->fam_off[f] + m] - 1;
            for (int i = 0; i < d; i++) uk[d * m + i] = u_all[(size_t)j * d + i];
        }
        matmul(nv, w, 1, C, uk, eps);
        matmul(nv, nv, 1, M->D, eps, sig);
        if (d == 2) {
            sigma[6 * k + 0] = sig[0]; sigma[6 * k + 1] = sig[1]; sigma[6 * k + 3] = sig[2];
        } else {
            for (int i = 0; i < 6; i++) sigma[6 * k + i] = sig[i];
        }
        free(C); free(uk);
    }
    /* extrapolation matrices L (pdsolve.cpp:9176-9196 for 3D, :9254-9277 for 2D) */
    double L8[64], L4[16];
    {
        double pI[8] = {-1.0, 1., 1., -1., -1., 1., 1., -1.}, qI[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0},
               rI[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0}, p[8], q[8], r[8];
        for (int i = 0; i < 8; i++) { p[i] = sqrt(3.0) * pI[i]; q[i] = sqrt(3.0) * qI[i]; r[i] = sqrt(3.0) * rI[i]; }
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++)
                L8[i * 8 + j] = 1.0 / 8 * (1 + pI[j] * p[i]) * (1 + qI[j] * q[i]) * (1 + rI[j] * r[i]);
        double a = 1 + 0.5 * sqrt(3.0), b = -0.5, c = 1 - 0.5 * sqrt(3.0);
        double row0[4] = {a, b, c, b};
        memset(L4, 0, sizeof(L4));
        for (int i = 0; i < 4; i++) L4[i] = row0[i];
        L4[1 * 4 + 1] = a; L4[1 * 4 + 2] = b; L4[1 * 4 + 3] = c;
        L4[2 * 4 + 2] = a; L4[2 * 4 + 3] = b;
        L4[3 * 4 + 3] = a;
        for (int i = 1; i < 4; i++)
            for (int j = 0; j < i; j++) L4[i * 4 + j] = L4[j * 4 + i];
    }
    for (int t = 0; t < M->nfe; t++) {
        int e = M->fe[t], ne = M->nen[e];
        double xN[8][3], Ue[24], Ns[6][8];
        ele_xyz(M, e, xN);
        for (int a = 0; a < ne; a++) {
            int j = M->econ[e * 8 + a] - 1;
            for (int i = 0; i < d; i++) Ue[d * a + i] = u_all[(size_t)j * d + i];
            count[j]++;
        }
        if (d == 2 && ne == 4) {
            /* pdfemEleQuad4N::eleFitStresses (pdfemEleQuad4N.cpp:73-124): q outer, p inner */
            double gp[3][4];
            int cnt = 0;
            for (int nq = 0; nq < 2; nq++)
                for (int np = 0; np < 2; np++) {
                    double B[24], detJ, eps[3], sig[3];
                    quad4_B(xN, GP2[np], GP2[nq], B, &detJ);
                    matmul(3, 8, 1, B, Ue, eps);
                    matmul(3, 3, 1, M->D, eps, sig);
                    for (int i = 0; i < 3; i++) gp[i][cnt] = sig[i];
                    cnt++;
                }
            for (int i = 0; i < 3; i++) matmul(4, 4, 1, L4, gp[i], Ns[i]);
        } else if (d == 2) {
            /* pdfemEleTri3N::eleFitStresses (pdfemEleTri3N.cpp:82-103): constant */
            double B[18], eps[3], sig[3];
            tri3_B(xN, B);
            matmul(3, 6, 1, B, Ue, eps);
            matmul(3, 3, 1, M->D, eps, sig);
            for (int a = 0; a < 3; a++)
                for (int i = 0; i < 3; i++) Ns[i][a] = sig[i];
        } else if (ne == 8) {
            /* pdfemEleBrick8N::eleFitStresses (pdfemEleBrick8N.cpp:119-166) */
            double g = 1.0 / sqrt(3), p[8] = {-g, g, g, -g, -g, g, g, -g}, q[8] = {-g, -g, g, g, -g, -g, g, g},
                   r[8] = {-g, -g, -g, -g, g, g, g, g}, gp[6][8];
            for (int i = 0; i < 8; i++) {
                double B[144], eps[6], sig[6];
                if (hex8_B(xN, p[i], q[i], r[i], B)) { free(count); return OM_ERR_SINGULAR; }
                matmul(6, 24, 1, B, Ue, eps);
                matmul(6, 6, 1, M->D, eps, sig);
                for (int j = 0; j < 6; j++) gp[j][i] = sig[j];
            }
            for (int j = 0; j < 6; j++) matmul(8, 8, 1, L8, gp[j], Ns[j]);
        } else {
            /* pdfemEleTetra4N::eleFitStresses: constant */
            double B[72], eps[6], sig[6];
            tet4_B(xN, B);
            matmul(6, 12, 1, B, Ue, eps);
            matmul(6, 6, 1, M->D, eps, sig);
            for (int a = 0; a < 4; a++)
                for (int i = 0; i < 6; i++) Ns[i][a] = sig[i];
        }
        for (int a = 0; a < ne; a++) {
            int j = M->econ[e * 8 + a] - 1;
            if (d == 2) {
                sigma[6 * j + 0] = sigma[6 * j + 0] + Ns[0][a];
                sigma[6 * j + 1] = sigma[6 * j + 1] + Ns[1][a];
                sigma[6 * j + 3] = sigma[6 * j + 3] + Ns[2][a];
            } else {
                for (int i = 0; i < 6; i++) sigma[6 * j + i] = sigma[6 * j + i] + Ns[i][a];
            }
        }
    }
    for (int j = 0; j < nn; j++) {
        if (count[j] == 0) continue;
        for (int i = 0; i < 6; i++) sigma[6 * j + i] = sigma[6 * j + i] / count[j];
    }
    if (d == 2 && M->proType == 2)
        for (int j = 0; j < nn; j++) sigma[6 * j + 2] = M->nu * (sigma[6 * j + 0] + sigma[6 * j + 1]);
    free(count);
    return OM_OK;
}

/* ---- quasi-static failure (2D, FC 2) -------------------------------------- */

/* DOF values [nn*dim] into the pdDof values (storeDisplacementResult for every DOF) */
void om_set_values(om_model *M, const double *u_all) {
    for (int k = 0; k < M->nn; k++)
        for (int i = 0; i < M->dim; i++) M->val[k * 3 + i] = u_all[k * M->dim + i];
}

void om_set_bonds(om_model *M, const int *bond) { memcpy(M->bond, bond, sizeof(int) * M->fam_off[M->nfam]); }

/* setVaryEssentialBC with the quasi-static ratio branch (pdsolve.cpp:5263-5312, solvFlag 2)
 * and setVaryNaturalBC (:5314-5329) */
void om_advance_bc_ratio(om_model *M, double rati) {
    for (int s = 0; s < M->n_ebc; s++) {
        if (!M->ebc_vary[s]) continue;
        double dv;
        if (M->ebc_val[s] == 0) {
            dv = M->ebc_vel[s];
        } else {
            double ndv = (rati - 1) * M->ebc_val[s];
            if (fabs(ndv) > fabs(M->ebc_vel[s])) dv = M->ebc_vel[s];
            else if (rati < 1) dv = 0;
            else dv = ndv;
        }
        double U = M->ebc_val[s] + M->dt * dv;
        M->ebc_val[s] = U;
        for (int t = M->ebc_off[s]; t < M->ebc_off[s + 1]; t++) M->val[(M->ebc_nid[t] - 1) * 3 + M->ebc_dof[s]] = U;
    }
    for (int s = 0; s < M->n_nbc; s++)
        if (M->nbc_vary[s]) M->nbc_val[s] = M->nbc_val[s] + M->dt * M->nbc_vel[s];
}

typedef struct {
    double v;
    long long order;
    int f, m;
} om_cand;

static int cand_cmp(const void *a, const void *b) {
    const om_cand *x = a, *y = b;
    if (x->v > y->v) return -1;
    if (x->v < y->v) return 1;
    return (x->order > y->order) - (x->order < y->order);
}

/* failureCriterion_maxPriSig + Finally_TopK_andUpdate_bondStaus with flag 2 (pdsolve.cpp:7184-7289,
 * :6981-7182), single rank, 2D.  sigma [nn*6] (xx, yy, zz, xy, ...).  For every intact bond the
 * larger eigenvalue s1 of the mean of the two nodes' in-plane stresses, computed in closed form
 * (the reference calls dsyev; the two agree to rounding); the bonds with s1 > sigult are kept
 * as the topk largest (0: all; among equal values the earlier in family / member order, as the
 * reference's min-heap replaces only on a larger value); each kept bond puts a crack segment of
 * length minDelta through its midpoint, normal to the principal direction, and updateBondstate
 * breaks the bonds crossing it.  out: [0] bonds kept, [1] bonds broken, [2] largest s1 (>= 0). */
int om_fc2_step(om_model *M, const double *sigma, double sigult, int topk, double *out) {
    if (M->dim != 2) return OM_ERR_ARG;
    int nb = M->fam_off[M->nfam];
    om_cand *cand = xcalloc(nb + 1, sizeof(om_cand));
    int nc = 0;
    long long order = 0;
    double smax = 0;
    for (int f = 0; f < M->nfam; f++) {
        int nk = M->fam_mem[M->fam_off[f]] - 1;
        const double *sk = &sigma[nk * 6];
        for (int m = 1; m < M->fam_off[f + 1] - M->fam_off[f]; m++, order++) {
            int idx = M->fam_off[f] + m;
            if (M->bond[idx] != 1) continue;
            const double *sm = &sigma[(M->fam_mem[idx] - 1) * 6];
            double a = 0.5 * (sk[0] + sm[0]), b = 0.5 * (sk[1] + sm[1]), c = 0.5 * (sk[3] + sm[3]);
            double h = 0.5 * (a - b);
            double s1 = 0.5 * (a + b) + sqrt(h * h + c * c);
            if (s1 > smax) smax = s1;
            if (s1 > sigult) {
                cand[nc].v = s1;
                cand[nc].order = order;
                cand[nc].f = f;
                cand[nc].m = m;
                nc++;
            }
        }
    }
    qsort(cand, nc, sizeof(om_cand), cand_cmp);
    if (topk > 0 && nc > topk) nc = topk;
    int before = 0;
    for (int i = 0; i < nb; i++) before += M->bond[i];
    double minD = 0.5 * M->minDelta;
    for (int i = 0; i < nc; i++) {
        int f = cand[i].f, idx = M->fam_off[f] + cand[i].m;
        int nk = M->fam_mem[M->fam_off[f]] - 1, nm = M->fam_mem[idx] - 1;
        const double *xk = &M->x[nk * 3], *xm = &M->x[nm * 3];
        double xc[3];
        for (int q = 0; q < 3; q++) xc[q] = 0.5 * (xk[q] + xm[q]);
        const double *sk = &sigma[nk * 6], *sm = &sigma[nm * 6];
        double a = 0.5 * (sk[0] + sm[0]), b = 0.5 * (sk[1] + sm[1]), c = 0.5 * (sk[3] + sm[3]);
        double h = 0.5 * (a - b), l1 = 0.5 * (a + b) + sqrt(h * h + c * c);
        double e0, e1;
        if (c != 0.0) { e0 = l1 - b; e1 = c; }
        else if (a >= b) { e0 = 1.0; e1 = 0.0; }
        else { e0 = 0.0; e1 = 1.0; }
        double nrm = sqrt(e0 * e0 + e1 * e1);
        e0 /= nrm;
        e1 /= nrm;
        double crack[9] = {xc[0] - minD * e1, xc[1] + minD * e0, 0, xc[0] + minD * e1, xc[1] - minD * e0, 0, 0, 0, 0};
        update_bondstate(M, crack);
    }
    int after = 0;
    for (int i = 0; i < nb; i++) after += M->bond[i];
    out[0] = nc;
    out[1] = before - after;
    out[2] = smax;
    free(cand);
    return OM_OK;
}

/* calLocalDamage (pdsolve.cpp:7338-7373) */
void om_local_damage(const om_model *M, double *phi) {
    for (int k = 0; k < M->nn; k++) phi[k] = 0.0;
    for (int f = 0; f < M->nfam; f++) {
        double t = 0;
        int n = M->fam_off[f + 1] - M->fam_off[f];
        for (int m = 1; m < n; m++) t = t + M->bond[M->fam_off[f] + m];
        phi[M->fam_mem[M->fam_off[f]] - 1] = 1.0 - t / (n - 1.0);
    }
}

/* ------------------------------------------------------------------------ */
/* Block ILU(0) preconditioner.  The reference has no counterpart (it calls   */
/* MKL PARDISO, calMatrixOperations.cpp:378-498); these are textbook          */
/* restatements that pin the library's ILU(0) path:                            */
/*  - greedy colouring of the node graph in node order;                        */
/*  - ILU(0), IKJ form (Saad, Iterative Methods for Sparse Linear Systems,     */
/*    2nd ed., Alg. 10.4) over d x d node blocks in a given elimination order; */
/*  - the forward / backward solves with its factors.                          */
/* The node-block matrix uses the node adjacency (off, adj: sorted columns,    */
/* self included), values K[slot*d*d + i*d + j] row-major per block.           */
/* ------------------------------------------------------------------------ */

/* greedy colouring of the symmetrised node graph: node a takes the smallest
   colour no lower-numbered neighbour holds (b < a with b in row a or a in
   row b); returns the number of colours */
int om_greedy_colors(int nn, const int *off, const int *adj, int *color) {
    int *toff = (int *)xcalloc((size_t)nn + 1, sizeof(int));
    int *tadj = (int *)xcalloc((size_t)off[nn] + 1, sizeof(int));
    for (int a = 0; a < nn; a++)
        for (int s = off[a]; s < off[a + 1]; s++) toff[adj[s] + 1]++;
    for (int a = 0; a < nn; a++) toff[a + 1] += toff[a];
    int *pos = (int *)xcalloc((size_t)nn + 1, sizeof(int));
    memcpy(pos, toff, sizeof(int) * (size_t)nn);
    for (int a = 0; a < nn; a++)
        for (int s = off[a]; s < off[a + 1]; s++) tadj[pos[adj[s]]++] = a;
    free(pos);
    int nc = 0;
    int *mark = (int *)xcalloc((size_t)nn + 1, sizeof(int));
    for (int a = 0; a < nn; a++) {
        for (int s = off[a]; s < off[a + 1]; s++)
            if (adj[s] < a) mark[color[adj[s]]] = a + 1;
        for (int s = toff[a]; s < toff[a + 1]; s++)
            if (tadj[s] < a) mark[color[tadj[s]]] = a + 1;
        int c = 0;
        while (mark[c] == a + 1) c++;
        color[a] = c;
        if (c + 1 > nc) nc = c + 1;
    }
    free(mark);
    free(toff);
    free(tadj);
    return nc;
}

/* inverse of a d x d block by cofactors (d <= 3); returns 0 when singular */
static int block_inverse(int d, const double *A, double *X) {
    if (d == 1) {
        if (A[0] == 0.0) return 0;
        X[0] = 1.0 / A[0];
        return 1;
    }
    if (d == 2) {
        double det = A[0] * A[3] - A[1] * A[2];
        if (det == 0.0 || !isfinite(det)) return 0;
        X[0] = A[3] / det; X[1] = -A[1] / det; X[2] = -A[2] / det; X[3] = A[0] / det;
        return 1;
    }
    double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8], c02 = A[3] * A[7] - A[4] * A[6];
    double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    if (det == 0.0 || !isfinite(det)) return 0;
    X[0] = c00 / det;
    X[1] = (A[2] * A[7] - A[1] * A[8]) / det;
    X[2] = (A[1] * A[5] - A[2] * A[4]) / det;
    X[3] = c01 / det;
    X[4] = (A[0] * A[8] - A[2] * A[6]) / det;
    X[5] = (A[2] * A[3] - A[0] * A[5]) / det;
    X[6] = c02 / det;
    X[7] = (A[1] * A[6] - A[0] * A[7]) / det;
    X[8] = (A[0] * A[4] - A[1] * A[3]) / det;
    return 1;
}

typedef struct { int rank, slot; } om_rs;
static int rs_cmp(const void *a, const void *b) {
    const om_rs *x = (const om_rs *)a, *y = (const om_rs *)b;
    return (x->rank > y->rank) - (x->rank < y->rank);
}

/* Block ILU(0) in the elimination order rank[] (a permutation of 0..nn-1).
   F (same layout as K) receives L's strict lower blocks (unit block diagonal
   implied) and U's diagonal and upper blocks; Dinv[a*d*d] the inverses of
   U's diagonal blocks.  For node i in increasing rank, for each lower
   neighbour k in increasing rank: F_ik <- F_ik * Dinv_k, then F_ij -= F_ik * F_kj
   for every j of row i with rank[j] > rank[k] and (k, j) in the pattern. */
int om_bilu0(int nn, int d, const int *off, const int *adj, const double *K, const int *rank, double *F,
             double *Dinv) {
    const int dd = d * d;
    memcpy(F, K, sizeof(double) * (size_t)off[nn] * dd);
    int *node_of = (int *)xcalloc((size_t)nn, sizeof(int));
    for (int a = 0; a < nn; a++) node_of[rank[a]] = a;
    int maxrow = 0;
    for (int a = 0; a < nn; a++)
        if (off[a + 1] - off[a] > maxrow) maxrow = off[a + 1] - off[a];
    om_rs *lo = (om_rs *)xcalloc((size_t)maxrow + 1, sizeof(om_rs));
    int rc = OM_OK;
    for (int t = 0; t < nn && rc == OM_OK; t++) {
        int i = node_of[t], nl = 0, di = -1;
        for (int s = off[i]; s < off[i + 1]; s++) {
            if (adj[s] == i) di = s;
            else if (rank[adj[s]] < t) { lo[nl].rank = rank[adj[s]]; lo[nl].slot = s; nl++; }
        }
        if (di < 0) { rc = OM_ERR_PATTERN; break; }
        qsort(lo, (size_t)nl, sizeof(om_rs), rs_cmp);
        for (int q = 0; q < nl; q++) {
            int sk = lo[q].slot, k = adj[sk], rk = lo[q].rank;
            double X[9], T[9];
            const double *Dk = &Dinv[(size_t)k * dd];
            double *Fik = &F[(size_t)sk * dd];
            for (int r = 0; r < d; r++)
                for (int c = 0; c < d; c++) {
                    double acc = 0.0;
                    for (int e = 0; e < d; e++) acc += Fik[r * d + e] * Dk[e * d + c];
                    X[r * d + c] = acc;
                }
            memcpy(Fik, X, sizeof(double) * dd);
            /* merge row i and row k (both sorted by node) */
            int s = off[i], u = off[k];
            while (s < off[i + 1] && u < off[k + 1]) {
                if (adj[s] < adj[u]) { s++; continue; }
                if (adj[s] > adj[u]) { u++; continue; }
                if (rank[adj[s]] > rk) {
                    const double *Fkj = &F[(size_t)u * dd];
                    double *Fij = &F[(size_t)s * dd];
                    for (int r = 0; r < d; r++)
                        for (int c = 0; c < d; c++) {
                            double acc = 0.0;
                            for (int e = 0; e < d; e++) acc += X[r * d + e] * Fkj[e * d + c];
                            T[r * d + c] = acc;
                        }
                    for (int e = 0; e < dd; e++) Fij[e] -= T[e];
                }
                s++;
                u++;
            }
        }
        if (!block_inverse(d, &F[(size_t)di * dd], &Dinv[(size_t)i * dd])) rc = OM_ERR_SINGULAR;
    }
    free(lo);
    free(node_of);
    return rc;
}

/* z = (L U)^-1 r with the factors of om_bilu0: forward solve in increasing
   rank (y_i = r_i - sum_{rank k < rank i} F_ik y_k), backward in decreasing
   rank (z_i = Dinv_i (y_i - sum_{rank j > rank i} F_ij z_j)); row sums in slot
   order.  r, z: [nn*d]; z may not alias r. */
void om_bilu0_apply(int nn, int d, const int *off, const int *adj, const double *F, const double *Dinv,
                    const int *rank, const double *r, double *z) {
    const int dd = d * d;
    int *node_of = (int *)xcalloc((size_t)nn, sizeof(int));
    for (int a = 0; a < nn; a++) node_of[rank[a]] = a;
    for (int t = 0; t < nn; t++) {
        int i = node_of[t];
        double acc[3];
        for (int r0 = 0; r0 < d; r0++) acc[r0] = r[(size_t)i * d + r0];
        for (int s = off[i]; s < off[i + 1]; s++) {
            int k = adj[s];
            if (k == i || rank[k] > t) continue;
            for (int r0 = 0; r0 < d; r0++)
                for (int c = 0; c < d; c++) acc[r0] -= F[(size_t)s * dd + r0 * d + c] * z[(size_t)k * d + c];
        }
        for (int r0 = 0; r0 < d; r0++) z[(size_t)i * d + r0] = acc[r0];
    }
    for (int t = nn - 1; t >= 0; t--) {
        int i = node_of[t];
        double acc[3];
        for (int r0 = 0; r0 < d; r0++) acc[r0] = z[(size_t)i * d + r0];
        for (int s = off[i]; s < off[i + 1]; s++) {
            int j = adj[s];
            if (j == i || rank[j] < t) continue;
            for (int r0 = 0; r0 < d; r0++)
                for (int c = 0; c < d; c++) acc[r0] -= F[(size_t)s * dd + r0 * d + c] * z[(size_t)j * d + c];
        }
        for (int r0 = 0; r0 < d; r0++) {
            double v = 0.0;
            for (int c = 0; c < d; c++) v += Dinv[(size_t)i * dd + r0 * d + c] * acc[c];
            z[(size_t)i * d + r0] = v;
        }
    }
    free(node_of);
}
