// Nodal stresses of a displacement field: pdsolve::calGlobalNodeStresses
// (pdsolve.cpp:9540-9604), which the reference's drivers run at every output
// step.
//   * PD family nodes (calPDNodeStresses :9081-9166): sigma = D (C u_family),
//     C the non-local strain operator of matC2D / matC3D (:3212-3398), built
//     from the bond strain weights mu_km omega g V that the family kernel keeps;
//   * FE elements (calFEMNodeStresses_EXP :9168-9335, eleFitStresses with the
//     extrapolation flag): stresses D B u_e at the 2x2 / 2x2x2 Gauss points,
//     mapped to the corner nodes with the reference's matrix L; tri3 / tet4
//     give their constant stress to every node;
//   * every node: (PD stress + the corner values of its FE elements, in FE
//     order) / count, and sigma_zz = nu (sigma_xx + sigma_yy) in 2D plane
//     strain (pdNode::calSigzz).  2D components sit at 0 (xx), 1 (yy), 3 (xy).
// Products and sums follow the reference's order (matMultiply: t = a*b + t),
// without FMA contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device.hpp"
#include "element.hpp"

namespace pdx {

// ---- PD family nodes ------------------------------------------------------
template <int D>
__global__ void k_stress_pd(int nfam, const int *fam_off, const int *fam_mem, const double *fam_c, const double *u,
                            const double *Dm, double *sig_fam) {
#pragma clang fp contract(off)
    constexpr int NV = D == 3 ? 6 : 3;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfam) return;
    const int off = fam_off[f], n = fam_off[f + 1] - off;
    // block 0 of C: the negative sums of the member weights, in member order
    double c0[D];
#pragma unroll
    for (int i = 0; i < D; i++) c0[i] = 0.0;
    for (int m = 1; m < n; m++)
#pragma unroll
        for (int i = 0; i < D; i++) c0[i] = -fam_c[(size_t)(off + m) * D + i] + c0[i];
    double eps[NV];
#pragma unroll
    for (int r = 0; r < NV; r++) eps[r] = 0.0;
    for (int m = 0; m < n; m++) {
        const int j = fam_mem[off + m];
        double c[D], uj[D];
#pragma unroll
        for (int i = 0; i < D; i++) {
            c[i] = m == 0 ? c0[i] : fam_c[(size_t)(off + m) * D + i];
            uj[i] = u[(size_t)j * D + i];
        }
        // row r of C times the member's DOFs, DOF-minor (zero entries skipped)
        if constexpr (D == 2) {
            eps[0] = c[0] * uj[0] + eps[0];
            eps[1] = c[1] * uj[1] + eps[1];
            eps[2] = c[1] * uj[0] + eps[2];
            eps[2] = c[0] * uj[1] + eps[2];
        } else {
            eps[0] = c[0] * uj[0] + eps[0];
            eps[1] = c[1] * uj[1] + eps[1];
            eps[2] = c[2] * uj[2] + eps[2];
            eps[3] = c[1] * uj[0] + eps[3];
            eps[3] = c[0] * uj[1] + eps[3];
            eps[4] = c[2] * uj[1] + eps[4];
            eps[4] = c[1] * uj[2] + eps[4];
            eps[5] = c[2] * uj[0] + eps[5];
            eps[5] = c[0] * uj[2] + eps[5];
        }
    }
    double *out = sig_fam + (size_t)f * 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < NV; k++) t = Dm[i * NV + k] * eps[k] + t;
        out[i] = t;
    }
}

// ---- FE elements ----------------------------------------------------------
__constant__ double c_shp[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
__constant__ double c_shq[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
__constant__ double c_shr[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};

struct StressFeArgs {
    const double *x, *u, *Dm;
    const int *econ, *nen, *fe_elem, *list;
    int nlist;
    double L8[64], L4[16];
    double *stage;  // [nfe][8 nodes][6]
    int *status;
};

// hex8: 32 elements per block; phase 1 one thread per (element, Gauss point)
// (pdfemEleBrick8N::eleFitStresses points p[i], q[i], r[i], :136-151), phase 2
// one thread per (element, corner): Nsigma = L sigma_gp
constexpr int SH_EPB = 32;
__global__ void __launch_bounds__(SH_EPB * 8) k_stress_hex8(StressFeArgs A) {
#pragma clang fp contract(off)
    __shared__ double sgp[SH_EPB][8][6];
    const int el = threadIdx.x / 8, g = threadIdx.x % 8, idx = blockIdx.x * SH_EPB + el;
    const double a = 1.0 / sqrt(3.0);
    if (idx < A.nlist) {
        const int e = A.fe_elem[A.list[idx]];
        double xN[8][3], Ue[24];
        for (int b = 0; b < 8; b++) {
            const int j = A.econ[(size_t)e * 8 + b];
            for (int c = 0; c < 3; c++) {
                xN[b][c] = A.x[(size_t)j * 3 + c];
                Ue[3 * b + c] = A.u[(size_t)j * 3 + c];
            }
        }
        const double p = a * c_shp[g], q = a * c_shq[g], r = a * c_shr[g];
        double dN[8][3], J[9];
        for (int i = 0; i < 8; i++) {
            dN[i][0] = 1.0 / 8 * c_shp[i] * (1 + q * c_shq[i]) * (1 + r * c_shr[i]);
            dN[i][1] = 1.0 / 8 * c_shq[i] * (1 + p * c_shp[i]) * (1 + r * c_shr[i]);
            dN[i][2] = 1.0 / 8 * c_shr[i] * (1 + p * c_shp[i]) * (1 + q * c_shq[i]);
        }
        for (int i = 0; i < 9; i++) J[i] = 0.0;
        for (int cc = 0; cc < 3; cc++)
            for (int i = 0; i < 8; i++) {
                J[0 * 3 + cc] = dN[i][0] * xN[i][cc] + J[0 * 3 + cc];
                J[1 * 3 + cc] = dN[i][1] * xN[i][cc] + J[1 * 3 + cc];
                J[2 * 3 + cc] = dN[i][2] * xN[i][cc] + J[2 * 3 + cc];
            }
        double gr[8][3];
        if (!lu3_solve8(J, dN, gr)) atomicOr(A.status, 2);
        // eps = B Ue over the 24 columns (node-major, DOF-minor; zeros skipped)
        double eps[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 8; i++) {
            const double ux = Ue[3 * i], uy = Ue[3 * i + 1], uz = Ue[3 * i + 2];
            eps[0] = gr[i][0] * ux + eps[0];
            eps[1] = gr[i][1] * uy + eps[1];
            eps[2] = gr[i][2] * uz + eps[2];
            eps[3] = gr[i][1] * ux + eps[3];
            eps[3] = gr[i][0] * uy + eps[3];
            eps[4] = gr[i][2] * uy + eps[4];
            eps[4] = gr[i][1] * uz + eps[4];
            eps[5] = gr[i][2] * ux + eps[5];
            eps[5] = gr[i][0] * uz + eps[5];
        }
        for (int i = 0; i < 6; i++) {
            double t = 0.0;
            for (int k = 0; k < 6; k++) t = A.Dm[i * 6 + k] * eps[k] + t;
            sgp[el][g][i] = t;
        }
    }
    __syncthreads();
    if (idx >= A.nlist) return;
    const int node = g;  // phase 2: thread (el, corner)
    double *out = A.stage + ((size_t)A.list[idx] * 8 + node) * 6;
    for (int i = 0; i < 6; i++) {
        double t = 0.0;
        for (int k = 0; k < 8; k++) t = A.L8[node * 8 + k] * sgp[el][k][i] + t;
        out[i] = t;
    }
}

// quad4: 64 elements per block; Gauss points q outer, p inner (pdfemEleQuad4N.cpp:93-110)
constexpr int SQ_EPB = 64;
__global__ void __launch_bounds__(SQ_EPB * 4) k_stress_quad4(StressFeArgs A) {
#pragma clang fp contract(off)
    __shared__ double sgp[SQ_EPB][4][3];
    const int el = threadIdx.x / 4, g = threadIdx.x % 4, idx = blockIdx.x * SQ_EPB + el;
    const double a = 0.57735026918962573;
    const double gpt[2] = {-a, a};
    if (idx < A.nlist) {
        const int e = A.fe_elem[A.list[idx]];
        double xN[4][2], Ue[8];
        for (int b = 0; b < 4; b++) {
            const int j = A.econ[(size_t)e * 8 + b];
            xN[b][0] = A.x[(size_t)j * 3];
            xN[b][1] = A.x[(size_t)j * 3 + 1];
            Ue[2 * b] = A.u[(size_t)j * 2];
            Ue[2 * b + 1] = A.u[(size_t)j * 2 + 1];
        }
        const double p = gpt[g & 1], q = gpt[g >> 1];
        // pdfemEleQuad4N::JacobiMat / BmatFEM (pdfemEleQuad4N.cpp:252-318)
        const double dNdp[4] = {0.25 * (-1 + q), 0.25 * (1 - q), 0.25 * (1 + q), 0.25 * (-1 - q)};
        const double dNdq[4] = {0.25 * (-1 + p), 0.25 * (-1 - p), 0.25 * (1 + p), 0.25 * (1 - p)};
        double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
        for (int i = 0; i < 4; i++) {
            J00 = dNdp[i] * xN[i][0] + J00;
            J01 = dNdp[i] * xN[i][1] + J01;
            J10 = dNdq[i] * xN[i][0] + J10;
            J11 = dNdq[i] * xN[i][1] + J11;
        }
        const double detJ = J00 * J11 - J01 * J10;
        const double i00 = J11 / detJ, i01 = -J01 / detJ, i10 = -J10 / detJ, i11 = J00 / detJ;
        double eps[3] = {0, 0, 0};
        for (int i = 0; i < 4; i++) {
            const double b1 = i00 * dNdp[i] + i01 * dNdq[i];
            const double b2 = i10 * dNdp[i] + i11 * dNdq[i];
            eps[0] = b1 * Ue[2 * i] + eps[0];
            eps[1] = b2 * Ue[2 * i + 1] + eps[1];
            eps[2] = b2 * Ue[2 * i] + eps[2];
            eps[2] = b1 * Ue[2 * i + 1] + eps[2];
        }
        for (int i = 0; i < 3; i++) {
            double t = 0.0;
            for (int k = 0; k < 3; k++) t = A.Dm[i * 3 + k] * eps[k] + t;
            sgp[el][g][i] = t;
        }
    }
    __syncthreads();
    if (idx >= A.nlist) return;
    const int node = g;
    double *out = A.stage + ((size_t)A.list[idx] * 8 + node) * 6;
    const int slot[3] = {0, 1, 3};
    for (int i = 0; i < 3; i++) {
        double t = 0.0;
        for (int k = 0; k < 4; k++) t = A.L4[node * 4 + k] * sgp[el][k][i] + t;
        out[slot[i]] = t;
    }
}

// tri3 / tet4: one thread per element, constant stress to every node
template <int D>
__global__ void k_stress_simplex(StressFeArgs A) {
#pragma clang fp contract(off)
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= A.nlist) return;
    const int e = A.fe_elem[A.list[idx]];
    double xN[4][3], Ue[12];
    constexpr int NE = D + 1;
    for (int b = 0; b < NE; b++) {
        const int j = A.econ[(size_t)e * 8 + b];
        for (int c = 0; c < 3; c++) xN[b][c] = A.x[(size_t)j * 3 + c];
        for (int c = 0; c < D; c++) Ue[D * b + c] = A.u[(size_t)j * D + c];
    }
    double sig[6] = {0, 0, 0, 0, 0, 0};
    if constexpr (D == 2) {
        // pdfemEleTri3N::BmatFEM (pdfemEleTri3N.cpp:185-210), volume 0.5 |det [1 x y]|
        double m3[9] = {1, xN[0][0], xN[0][1], 1, xN[1][0], xN[1][1], 1, xN[2][0], xN[2][1]};
        const double V2 = 2.0 * (0.5 * fabs(det3(m3)));
        const double b[3] = {xN[1][1] - xN[2][1], xN[2][1] - xN[0][1], xN[0][1] - xN[1][1]};
        const double c[3] = {-xN[1][0] + xN[2][0], -xN[2][0] + xN[0][0], -xN[0][0] + xN[1][0]};
        double eps[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++) {
            const double bi = b[i] / V2, ci = c[i] / V2;
            eps[0] = bi * Ue[2 * i] + eps[0];
            eps[1] = ci * Ue[2 * i + 1] + eps[1];
            eps[2] = ci * Ue[2 * i] + eps[2];
            eps[2] = bi * Ue[2 * i + 1] + eps[2];
        }
        double s3[3];
        for (int i = 0; i < 3; i++) {
            double t = 0.0;
            for (int k = 0; k < 3; k++) t = A.Dm[i * 3 + k] * eps[k] + t;
            s3[i] = t;
        }
        sig[0] = s3[0]; sig[1] = s3[1]; sig[3] = s3[2];
    } else {
        // pdfemEleTetra4N::BmatFEM (pdfemEleTetra4N.cpp:84-170): cofactors of [1 x y z]
        double m3[9];
        auto det_rows = [&](int r0, int r1, int r2, int c0, int c1) {
            m3[0] = 1; m3[1] = xN[r0][c0]; m3[2] = xN[r0][c1];
            m3[3] = 1; m3[4] = xN[r1][c0]; m3[5] = xN[r1][c1];
            m3[6] = 1; m3[7] = xN[r2][c0]; m3[8] = xN[r2][c1];
            return det3(m3);
        };
        double bb[4], cc[4], dd[4];
        bb[0] = -det_rows(1, 2, 3, 1, 2); bb[1] = det_rows(0, 2, 3, 1, 2);
        bb[2] = -det_rows(0, 1, 3, 1, 2); bb[3] = det_rows(0, 1, 2, 1, 2);
        cc[0] = det_rows(1, 2, 3, 0, 2); cc[1] = -det_rows(0, 2, 3, 0, 2);
        cc[2] = det_rows(0, 1, 3, 0, 2); cc[3] = -det_rows(0, 1, 2, 0, 2);
        dd[0] = -det_rows(1, 2, 3, 0, 1); dd[1] = det_rows(0, 2, 3, 0, 1);
        dd[2] = -det_rows(0, 1, 3, 0, 1); dd[3] = det_rows(0, 1, 2, 0, 1);
        double Jv[9];
        for (int r = 0; r < 3; r++)
            for (int k = 0; k < 3; k++) Jv[r * 3 + k] = xN[r + 1][k] - xN[0][k];
        const double V6 = 6.0 * (fabs(det3(Jv)) / 6.0);  // tet4 volume (pdfemEleTetra4N.cpp:210-217)
        double eps[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 4; i++) {
            const double b = bb[i] / V6, c = cc[i] / V6, d = dd[i] / V6;
            const double ux = Ue[3 * i], uy = Ue[3 * i + 1], uz = Ue[3 * i + 2];
            eps[0] = b * ux + eps[0];
            eps[1] = c * uy + eps[1];
            eps[2] = d * uz + eps[2];
            eps[3] = c * ux + eps[3];
            eps[3] = b * uy + eps[3];
            eps[4] = d * uy + eps[4];
            eps[4] = c * uz + eps[4];
            eps[5] = d * ux + eps[5];
            eps[5] = b * uz + eps[5];
        }
        for (int i = 0; i < 6; i++) {
            double t = 0.0;
            for (int k = 0; k < 6; k++) t = A.Dm[i * 6 + k] * eps[k] + t;
            sig[i] = t;
        }
    }
    for (int node = 0; node < NE; node++) {
        double *out = A.stage + ((size_t)A.list[idx] * 8 + node) * 6;
        for (int i = 0; i < 6; i++) out[i] = sig[i];
    }
}

// ---- nodes: PD value, then the FE corner values in FE order, then the average
template <int D>
__global__ void k_stress_nodes(int nn, const int *famid, const double *sig_fam, const int *nfe_off,
                               const int *nfe_item, const double *stage, double nu, int plane_strain, double *sigma) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    int count = 0;
    const int f = famid[a];
    if (f >= 0) {
        count = 1;
        if constexpr (D == 2) {
            s[0] = sig_fam[(size_t)f * 6 + 0];
            s[1] = sig_fam[(size_t)f * 6 + 1];
            s[3] = sig_fam[(size_t)f * 6 + 2];
        } else {
            for (int i = 0; i < 6; i++) s[i] = sig_fam[(size_t)f * 6 + i];
        }
    }
    for (int q = nfe_off[a]; q < nfe_off[a + 1]; q++) {
        const int item = nfe_item[q];
        const double *v = stage + ((size_t)(item >> 3) * 8 + (item & 7)) * 6;
        if constexpr (D == 2) {
            s[0] = s[0] + v[0];
            s[1] = s[1] + v[1];
            s[3] = s[3] + v[3];
        } else {
            for (int i = 0; i < 6; i++) s[i] = s[i] + v[i];
        }
        count++;
    }
    if (count > 0)
        for (int i = 0; i < 6; i++) s[i] = s[i] / count;
    if (D == 2 && plane_strain) s[2] = nu * (s[0] + s[1]);
    for (int i = 0; i < 6; i++) sigma[(size_t)a * 6 + i] = s[i];
}

// sigma [nn*6] (device) of the DOF values u_dev [nn*d] (device)
// Largest principal stress of every intact bond of a 2D family: the eigenvalue of the 2x2
// mean of the two nodes' stresses, (a + b) / 2 + sqrt(((a - b) / 2)^2 + c^2), as
// failureCriterion_maxPriSig computes it with dsyev (pdsolve.cpp:7184-7247).  One thread per
// family; broken bonds and the self member get -inf.
__global__ void k_bond_sigma1_2d(int nfam, const int *__restrict__ fam_off, const int *__restrict__ fam_mem,
                                 const int *__restrict__ bond, const double *__restrict__ sig, double *s1) {
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfam) return;
    const int o0 = fam_off[f], o1 = fam_off[f + 1];
    const double *sk = sig + (size_t)fam_mem[o0] * 6;
    s1[o0] = -INFINITY;
    for (int q = o0 + 1; q < o1; q++) {
        if (bond[q] != 1) {
            s1[q] = -INFINITY;
            continue;
        }
        const double *sm = sig + (size_t)fam_mem[q] * 6;
        const double a = 0.5 * (sk[0] + sm[0]), b = 0.5 * (sk[1] + sm[1]), c = 0.5 * (sk[3] + sm[3]);
        const double h = 0.5 * (a - b);
        s1[q] = 0.5 * (a + b) + sqrt(h * h + c * c);
    }
}

int dev_bond_sigma1(Ctx &c, const int *bond_dev, const double *sigma_dev, double *s1_dev, std::string &err) {
    if (c.dim != 2) {
        err = "the principal-stress bond criterion is two-dimensional";
        return PDX_ERR_ARG;
    }
    if (c.nfam > 0) k_bond_sigma1_2d<<<(c.nfam + 127) / 128, 128, 0, c.stream>>>(c.nfam, c.fam_off, c.fam_mem, bond_dev,
                                                                                sigma_dev, s1_dev);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

int dev_node_stresses(Ctx &c, const double *u_dev, double *sigma_dev, std::string &err) {
    if (!c.K_valid) {
        int rc = dev_assemble_K(c, err);  // the family kernel leaves the bond strain weights
        if (rc) return rc;
    }
    const Model &m = *c.host;
    const int D = c.dim;
    if (!c.stress_stage)
        PDX_HIP(hipMalloc((void **)&c.stress_stage, ((size_t)std::max(c.nfe, 1) * 48 + (size_t)std::max(c.nfam, 1) * 6) *
                                                        sizeof(double)));
    double *stage = c.stress_stage, *sig_fam = c.stress_stage + (size_t)std::max(c.nfe, 1) * 48;
    int *status;
    PDX_HIP(hipMalloc((void **)&status, sizeof(int)));
    PDX_HIP(hipMemsetAsync(status, 0, sizeof(int), c.stream));
    if (c.nfam > 0) {
        if (D == 3) k_stress_pd<3><<<(c.nfam + 127) / 128, 128, 0, c.stream>>>(c.nfam, c.fam_off, c.fam_mem, c.fam_c,
                                                                             u_dev, c.Dm, sig_fam);
        else k_stress_pd<2><<<(c.nfam + 127) / 128, 128, 0, c.stream>>>(c.nfam, c.fam_off, c.fam_mem, c.fam_c, u_dev,
                                                                       c.Dm, sig_fam);
    }
    if (c.nfe > 0) {
        StressFeArgs A;
        A.x = c.x; A.u = u_dev; A.Dm = c.Dm; A.econ = c.econ; A.nen = c.nen; A.fe_elem = c.fe_elem;
        A.stage = stage; A.status = status;
        // extrapolation matrices L (pdsolve.cpp:9176-9196, :9254-9277)
        {
            const double pI[8] = {-1.0, 1., 1., -1., -1., 1., 1., -1.}, qI[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0},
                         rI[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};
            double p[8], q[8], r[8];
            for (int i = 0; i < 8; i++) { p[i] = std::sqrt(3.0) * pI[i]; q[i] = std::sqrt(3.0) * qI[i]; r[i] = std::sqrt(3.0) * rI[i]; }
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++)
                    A.L8[i * 8 + j] = 1.0 / 8 * (1 + pI[j] * p[i]) * (1 + qI[j] * q[i]) * (1 + rI[j] * r[i]);
            const double a = 1 + 0.5 * std::sqrt(3.0), b = -0.5, cc = 1 - 0.5 * std::sqrt(3.0);
            const double L4[16] = {a, b, cc, b, b, a, b, cc, cc, b, a, b, b, cc, b, a};
            for (int i = 0; i < 16; i++) A.L4[i] = L4[i];
        }
        if (c.n_fe_a > 0) {
            A.list = c.fe_list_a;
            A.nlist = c.n_fe_a;
            if (D == 3) k_stress_hex8<<<(c.n_fe_a + SH_EPB - 1) / SH_EPB, SH_EPB * 8, 0, c.stream>>>(A);
            else k_stress_quad4<<<(c.n_fe_a + SQ_EPB - 1) / SQ_EPB, SQ_EPB * 4, 0, c.stream>>>(A);
        }
        if (c.n_fe_b > 0) {
            A.list = c.fe_list_b;
            A.nlist = c.n_fe_b;
            if (D == 3) k_stress_simplex<3><<<(c.n_fe_b + 127) / 128, 128, 0, c.stream>>>(A);
            else k_stress_simplex<2><<<(c.n_fe_b + 127) / 128, 128, 0, c.stream>>>(A);
        }
    }
    int *famid;
    PDX_HIP(hipMalloc((void **)&famid, (size_t)c.nn * sizeof(int)));
    {
        std::vector<int> h(c.nn);
        for (int k = 0; k < c.nn; k++) h[k] = m.famid[k];
        PDX_HIP(hipMemcpyAsync(famid, h.data(), (size_t)c.nn * sizeof(int), hipMemcpyHostToDevice, c.stream));
        const int nb = (c.nn + 255) / 256;
        if (D == 3) k_stress_nodes<3><<<nb, 256, 0, c.stream>>>(c.nn, famid, sig_fam, c.nfe_off, c.nfe_item, stage, m.nu,
                                                               0, sigma_dev);
        else k_stress_nodes<2><<<nb, 256, 0, c.stream>>>(c.nn, famid, sig_fam, c.nfe_off, c.nfe_item, stage, m.nu,
                                                         m.pro_type == 2, sigma_dev);
        PDX_HIP(hipGetLastError());
        int hs = 0;
        PDX_HIP(hipMemcpyAsync(&hs, status, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        PDX_HIP(hipStreamSynchronize(c.stream));
        hipFree(famid);
        hipFree(status);
        if (hs) { err = "singular element Jacobian in the stress recovery"; return PDX_ERR_SINGULAR; }
    }
    return PDX_OK;
}

}  // namespace pdx
