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
#include <map>
#include <mutex>

#include "device.hpp"
#include "element.hpp"

namespace pdx {

// ---- PD family nodes ------------------------------------------------------
// One wavefront per family: the lanes stride over the members m >= 1 accumulating the strain
// terms C_m u_m and the weight sum; after a wave reduction the self block C_0 = -sum_m C_m
// (calPDNodeStresses, pdsolve.cpp:9081-9166) adds C_0 u_self, and sigma = D eps.
template <int D>
__global__ void __launch_bounds__(256) k_stress_pd(int nfam, const int *fam_off, const int *fam_mem,
                                                   const double *fam_c, const double *u, const double *Dm,
                                                   double *sig_fam) {
#pragma clang fp contract(off)
    constexpr int NV = D == 3 ? 6 : 3;
    const int f = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
    if (f >= nfam) return;  // whole waves leave together
    const int off = fam_off[f], n = fam_off[f + 1] - off;
    double eps[NV], cs[D];
#pragma unroll
    for (int r = 0; r < NV; r++) eps[r] = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) cs[i] = 0.0;
    auto add = [&](const double *c, const double *uj) {
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
    };
    for (int m = 1 + lane; m < n; m += 64) {
        const int j = fam_mem[off + m];
        double c[D], uj[D];
#pragma unroll
        for (int i = 0; i < D; i++) {
            c[i] = fam_c[(size_t)(off + m) * D + i];
            uj[i] = u[(size_t)j * D + i];
            cs[i] = cs[i] + c[i];
        }
        add(c, uj);
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
#pragma unroll
        for (int r = 0; r < NV; r++) eps[r] += __shfl_xor(eps[r], s, 64);
#pragma unroll
        for (int i = 0; i < D; i++) cs[i] += __shfl_xor(cs[i], s, 64);
    }
    {
        const int j = fam_mem[off];
        double c[D], uj[D];
#pragma unroll
        for (int i = 0; i < D; i++) {
            c[i] = -cs[i];
            uj[i] = u[(size_t)j * D + i];
        }
        add(c, uj);
    }
    if (lane < NV) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < NV; k++) t = Dm[lane * NV + k] * eps[k] + t;
        sig_fam[(size_t)f * 6 + lane] = t;
    }
}

// ---- FE elements ----------------------------------------------------------
__constant__ double c_shp[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
__constant__ double c_shq[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
__constant__ double c_shr[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};

struct StressFeArgs {
    const double *x, *u, *Dm;
    double lam, mu;  // isotropic 3D D: D_01 and D_33 (iso != 0)
    int iso;
    const int *econ, *nen, *fe_elem, *list;
    int nlist;
    double L8[64], L4[16];
    double *stage;  // [nfe][8 nodes][6]
    int *status;
};

// hex8: 32 elements per block.  Phase 0, one thread per (element, node): the node's
// coordinates and DOF values into LDS.  Phase 1, one thread per (element, Gauss point)
// (pdfemEleBrick8N::eleFitStresses points p[i], q[i], r[i], :136-151): J = sum_i x_i dN_i^T
// and H = sum_i u_i dN_i^T over the parent derivatives, the displacement gradient from
// J^T grad(u) = H (one partial-pivot LU, three right-hand sides, instead of a solve per
// shape function), the engineering strains and sigma = D eps (lam tr(eps) + 2 mu eps,
// mu gamma for an isotropic D).  Phase 2, one thread per (element, corner):
// Nsigma = L sigma_gp into LDS; the block then writes its elements' 8 x 6 values with
// 16-byte stores over one contiguous range when its list slots are consecutive (the list
// is increasing, so a span of count - 1 means consecutive), else one corner per thread.
constexpr int SH_EPB = 32;

// The hardware deals workgroups to the 8 XCDs round-robin; with a grid that is a multiple of 8,
// logical block (b % 8) * (grid / 8) + b / 8 gives each XCD (and its L2) one contiguous eighth
// of the range, so corners and nodes shared with the next layers of the mesh stay in one L2.
__device__ inline int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    return nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
}
inline int xcd_grid(int blocks) { return (blocks + 7) / 8 * 8; }

// at least 5 waves per SIMD: the register budget falls from 130 to 96 VGPRs without spills and the
// recovery takes ~102 instead of ~109 us on the 60^3 box (6 waves spill)
__global__ void __launch_bounds__(SH_EPB * 8, 5) k_stress_hex8(StressFeArgs A) {
#pragma clang fp contract(off)
    __shared__ __align__(16) double sxu[SH_EPB][8][6];
    __shared__ double sgp[SH_EPB][8][6];
    const int lb = xcd_block();
    if (lb * SH_EPB >= A.nlist) return;  // padding blocks of the XCD-rounded grid (whole block)
    const int el = threadIdx.x / 8, g = threadIdx.x % 8, idx = lb * SH_EPB + el;
    const double a = 1.0 / sqrt(3.0);
    if (idx < A.nlist) {
        const int e = A.fe_elem[A.list[idx]];
        const int j = A.econ[(size_t)e * 8 + g];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sxu[el][g][c] = A.x[(size_t)j * 3 + c];
            sxu[el][g][3 + c] = A.u[(size_t)j * 3 + c];
        }
    }
    __syncthreads();
    if (idx < A.nlist) {
        const double p = a * c_shp[g], q = a * c_shq[g], r = a * c_shr[g];
        // M = [J | H]: row k (parent direction), columns x, y, z then ux, uy, uz
        double M[3][6];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int c = 0; c < 6; c++) M[k][c] = 0.0;
#pragma unroll 2
        for (int i = 0; i < 8; i++) {
            const double d0 = 1.0 / 8 * c_shp[i] * (1 + q * c_shq[i]) * (1 + r * c_shr[i]);
            const double d1 = 1.0 / 8 * c_shq[i] * (1 + p * c_shp[i]) * (1 + r * c_shr[i]);
            const double d2 = 1.0 / 8 * c_shr[i] * (1 + p * c_shp[i]) * (1 + q * c_shq[i]);
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const double v = sxu[el][i][c];
                M[0][c] = d0 * v + M[0][c];
                M[1][c] = d1 * v + M[1][c];
                M[2][c] = d2 * v + M[2][c];
            }
        }
        // J G = H with G[d][c] = d u_c / d x_d: partial-pivot elimination on [J | H]
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            int pv = k;
            double mx = fabs(M[k][k]);
#pragma unroll
            for (int i = k + 1; i < 3; i++)
                if (fabs(M[i][k]) > mx) { mx = fabs(M[i][k]); pv = i; }
#pragma unroll
            for (int i = k + 1; i < 3; i++)
                if (pv == i)
#pragma unroll
                    for (int c = 0; c < 6; c++) { const double t = M[k][c]; M[k][c] = M[i][c]; M[i][c] = t; }
            if (M[k][k] == 0.0) ok = false;
#pragma unroll
            for (int i = k + 1; i < 3; i++) {
                const double l = M[i][k] / M[k][k];
#pragma unroll
                for (int c = k; c < 6; c++) M[i][c] = M[i][c] - l * M[k][c];
            }
        }
        if (!ok) atomicOr(A.status, 2);
        double G[3][3];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int i = 2; i >= 0; i--) {
                double t = M[i][3 + c];
#pragma unroll
                for (int k = i + 1; k < 3; k++) t = t - M[i][k] * G[k][c];
                G[i][c] = t / M[i][i];
            }
        const double eps[6] = {G[0][0], G[1][1], G[2][2], G[1][0] + G[0][1], G[2][1] + G[1][2], G[2][0] + G[0][2]};
        if (A.iso) {
            const double lt = A.lam * (eps[0] + eps[1] + eps[2]);
#pragma unroll
            for (int i = 0; i < 3; i++) sgp[el][g][i] = lt + 2.0 * A.mu * eps[i];
#pragma unroll
            for (int i = 3; i < 6; i++) sgp[el][g][i] = A.mu * eps[i];
        } else {
            for (int i = 0; i < 6; i++) {
                double t = 0.0;
                for (int k = 0; k < 6; k++) t = A.Dm[i * 6 + k] * eps[k] + t;
                sgp[el][g][i] = t;
            }
        }
    }
    __syncthreads();
    const int node = g;  // phase 2: thread (el, corner); sxu is free again
    if (idx < A.nlist)
        for (int i = 0; i < 6; i++) {
            double t = 0.0;
            for (int k = 0; k < 8; k++) t = A.L8[node * 8 + k] * sgp[el][k][i] + t;
            sxu[el][node][i] = t;
        }
    __syncthreads();
    const int idx0 = lb * SH_EPB, nel = min(SH_EPB, A.nlist - idx0);
    const int first = A.list[idx0];
    if (A.list[idx0 + nel - 1] - first == nel - 1) {
        double2 *dst = reinterpret_cast<double2 *>(A.stage + (size_t)first * 48);
        const double2 *src = reinterpret_cast<const double2 *>(&sxu[0][0][0]);
        for (int v = threadIdx.x; v < nel * 24; v += blockDim.x) dst[v] = src[v];
    } else if (idx < A.nlist) {
        double *out = A.stage + ((size_t)A.list[idx] * 8 + node) * 6;
        for (int i = 0; i < 6; i++) out[i] = sxu[el][node][i];
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
    const int a = xcd_block() * blockDim.x + threadIdx.x;
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
    // incident element corners in FE order, four at a time: all loads of a batch are issued
    // before the batch is summed (same order of the sum)
    const int q1 = nfe_off[a + 1];
    for (int q0 = nfe_off[a]; q0 < q1; q0 += 4) {
        double v[4][6];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (q0 + k < q1) {
                const int item = nfe_item[q0 + k];
                // 48-byte corner records: 16-byte aligned
                const double2 *src = reinterpret_cast<const double2 *>(stage + ((size_t)(item >> 3) * 8 + (item & 7)) * 6);
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const double2 w = src[i];
                    v[k][2 * i] = w.x;
                    v[k][2 * i + 1] = w.y;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (q0 + k < q1) {
                if constexpr (D == 2) {
                    s[0] = s[0] + v[k][0];
                    s[1] = s[1] + v[k][1];
                    s[3] = s[3] + v[k][3];
                } else {
#pragma unroll
                    for (int i = 0; i < 6; i++) s[i] = s[i] + v[k][i];
                }
                count++;
            }
        }
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

// the singular-Jacobian flag of the last recovery (synchronises the stream)
int stress_status_check(Ctx &c, std::string &err) {
    int hs = 0;
    PDX_HIP(hipMemcpyAsync(&hs, c.stress_status, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    if (hs) { err = "singular element Jacobian in the stress recovery"; return PDX_ERR_SINGULAR; }
    return PDX_OK;
}

// doubles of scratch one recovery uses: FE corner stresses (8 corners x 6 per element), family
// stresses (6 per family) and one slot for the singular-Jacobian flag
static size_t stress_scratch_doubles(const Ctx &c) {
    return (size_t)std::max(c.nfe, 1) * 48 + (size_t)std::max(c.nfam, 1) * 6 + 1;
}

__global__ void k_clear_flag(int *flag) { *flag = 0; }

// The launches of one recovery (no allocation, no synchronisation) with the scratch at base
// (stress_scratch_doubles(c) doubles); the flag is cleared by a kernel, not a memset, so the
// launches may be captured in a graph.
static int stress_launch(Ctx &c, const double *u_dev, double *sigma_dev, double *base, std::string &err) {
    const Model &m = *c.host;
    const int D = c.dim;
    double *stage = base, *sig_fam = base + (size_t)std::max(c.nfe, 1) * 48;
    int *status = (int *)(sig_fam + (size_t)std::max(c.nfam, 1) * 6);
    const int *famid = c.stress_famid;
    k_clear_flag<<<1, 1, 0, c.stream>>>(status);
    if (c.nfam > 0) {
        if (D == 3) k_stress_pd<3><<<(c.nfam + 3) / 4, 256, 0, c.stream>>>(c.nfam, c.fam_off, c.fam_mem, c.fam_c,
                                                                           u_dev, c.Dm, sig_fam);
        else k_stress_pd<2><<<(c.nfam + 3) / 4, 256, 0, c.stream>>>(c.nfam, c.fam_off, c.fam_mem, c.fam_c, u_dev,
                                                                     c.Dm, sig_fam);
    }
    if (c.nfe > 0) {
        StressFeArgs A;
        A.x = c.x; A.u = u_dev; A.Dm = c.Dm; A.econ = c.econ; A.nen = c.nen; A.fe_elem = c.fe_elem;
        A.stage = stage; A.status = status;
        A.iso = 0;
        A.lam = A.mu = 0.0;
        if (D == 3) {
            const double *Dh = m.D;
            const double lam = Dh[1], mu = Dh[21], tol = 1e-14 * (std::fabs(lam) + std::fabs(mu));
            A.iso = 1;
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    const double want = i < 3 && j < 3 ? lam + (i == j ? 2 * mu : 0.0) : (i == j ? mu : 0.0);
                    if (std::fabs(Dh[i * 6 + j] - want) > tol) A.iso = 0;
                }
            A.lam = lam;
            A.mu = mu;
        }
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
            if (D == 3) k_stress_hex8<<<xcd_grid((c.n_fe_a + SH_EPB - 1) / SH_EPB), SH_EPB * 8, 0, c.stream>>>(A);
            else k_stress_quad4<<<(c.n_fe_a + SQ_EPB - 1) / SQ_EPB, SQ_EPB * 4, 0, c.stream>>>(A);
        }
        if (c.n_fe_b > 0) {
            A.list = c.fe_list_b;
            A.nlist = c.n_fe_b;
            if (D == 3) k_stress_simplex<3><<<(c.n_fe_b + 127) / 128, 128, 0, c.stream>>>(A);
            else k_stress_simplex<2><<<(c.n_fe_b + 127) / 128, 128, 0, c.stream>>>(A);
        }
    }
    {
        const int nb = xcd_grid((c.nn + 255) / 256);
        if (D == 3) k_stress_nodes<3><<<nb, 256, 0, c.stream>>>(c.nn, famid, sig_fam, c.nfe_off, c.nfe_item, stage, m.nu,
                                                               0, sigma_dev);
        else k_stress_nodes<2><<<nb, 256, 0, c.stream>>>(c.nn, famid, sig_fam, c.nfe_off, c.nfe_item, stage, m.nu,
                                                         m.pro_type == 2, sigma_dev);
        PDX_HIP(hipGetLastError());
    }
    return PDX_OK;
}

int dev_node_stresses(Ctx &c, const double *u_dev, double *sigma_dev, std::string &err, bool check) {
    if (!c.K_valid) {
        int rc = dev_assemble_K(c, err);  // the family kernel leaves the bond strain weights
        if (rc) return rc;
    }
    if (!c.stress_stage) PDX_HIP(hipMalloc((void **)&c.stress_stage, stress_scratch_doubles(c) * sizeof(double)));
    c.stress_status = (int *)(c.stress_stage + stress_scratch_doubles(c) - 1);
    if (int rc = stress_launch(c, u_dev, sigma_dev, c.stress_stage, err)) return rc;
    return check ? stress_status_check(c, err) : PDX_OK;
}

// ---- the recovery without allocation ----------------------------------------------------
// Scratch of pdx_node_stresses_noalloc in a file-scope device array: every GPU the library runs
// on holds its own copy until the process ends.  Host code never passes the array's name to a
// kernel (it is a host-side stand-in there): the address on each device comes from
// hipGetSymbolAddress, taken once per device at context creation with that device current.
constexpr size_t kNoAllocStressDoubles = size_t(12) << 20;  // 96 MiB
static __device__ __align__(256) double scratch[kNoAllocStressDoubles];
static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;  // device -> address of that device's scratch

int stress_scratch_init(int device, std::string &err) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return PDX_OK;
    double *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return PDX_ERR_HIP;
    }
    g_scratch_ptr[device] = ptr;
    return PDX_OK;
}

int dev_node_stresses_noalloc(Ctx &c, const double *u_dev, double *sigma_dev, std::string &err) {
    if (!c.K_valid) {
        err = "K is not assembled: the recovery reads the bond strain weights of the last assembly "
              "(call pdx_assemble or a driver first)";
        return PDX_ERR_ARG;
    }
    const size_t need = stress_scratch_doubles(c);
    if (need > kNoAllocStressDoubles) {
        err = "model needs " + std::to_string(need * 8) + " bytes of stress scratch, more than the " +
              std::to_string(kNoAllocStressDoubles * 8) + " bytes of pdx_node_stresses_noalloc";
        return PDX_ERR_LIMIT;
    }
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = g_scratch_ptr.find(c.device);
        if (it != g_scratch_ptr.end()) base = it->second;
    }
    if (!base) {
        err = "stress scratch not set up on this device";
        return PDX_ERR_ARG;
    }
    if (int rc = stress_launch(c, u_dev, sigma_dev, base, err)) return rc;
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

}  // namespace pdx
