// Assembly kernels of the PDLSM-FEM hot path for CDNA4 (gfx950).
//
//  k_family   one wavefront per PD family: member staging in LDS, shape tensor
//             (5x5 / 9x9) accumulated in the reference member order, one
//             partial-pivot LU per family, one triangular solve per bond
//             (lane-parallel), bond stiffness G, row block H, and the
//             PD-region boundary traction terms N.D.C of the same family, all
//             written to the family node's block row through the scatter map
//             (plain stores: each family owns its row).
//  k_ke_*     element stiffness matrices (B per element and Gauss point in LDS)
//             into a staging array; k_fe_gather adds them node-centrically in
//             element order.
//  k_lumped   lumped mass rho * sum V_e / n_e, node-centric.
//  k_nbc_*    equivalent nodal forces of natural BCs for a unit traction.
//  k_spmv_*   node-block SpMV, short rows on 8-lane groups, long rows on a wave.
//  k_csr_extract  block rows -> the reference CSR (K_ff values).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device.hpp"
#include "element.hpp"
#include "rowblock.hpp"

namespace pdx {

template <int D>
struct Dm;
template <>
struct Dm<2> {
    static constexpr int P = 5, NV = 3, DD = 4;
};
template <>
struct Dm<3> {
    static constexpr int P = 9, NV = 6, DD = 9;
};

// exponents of the basis p = [x,y,z,x^2,y^2,z^2,xy,yz,zx] (3D) / [x,y,x^2,y^2,xy] (2D)
__constant__ int c_pe3[9][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0},
                                {0, 0, 2}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
__constant__ int c_pe2[5][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}, {1, 1}};

// Shape-tensor term of entry e for one bond, in the reference's expression order
// (pdsolve.cpp:1938-2013 for 2D, :2015-2064 for 3D).
template <int D>
__device__ inline double a_term(int e, double x, double y, double z, double om, double vm);

// the products of a_term with the exponents of entry e fixed per lane: multiplications by a
// selected factor, 1.0 where a_term's loops would stop (exact), so the bits are the same and the
// lanes do not diverge
struct ATermExp {
    int ex, ey, ez;
    double half;
};
template <int D>
__device__ inline ATermExp a_term_exp(int e) {
    ATermExp t;
    if constexpr (D == 3) {
        const int r = e / 9, c = e % 9;
        t.ex = c_pe3[r][0] + c_pe3[c][0];
        t.ey = c_pe3[r][1] + c_pe3[c][1];
        t.ez = c_pe3[r][2] + c_pe3[c][2];
        t.half = (c >= 3 && c < 6) ? 0.5 : 1.0;
    } else {
        const int r = e / 5, c = e % 5;
        t.ex = c_pe2[r][0] + c_pe2[c][0];
        t.ey = c_pe2[r][1] + c_pe2[c][1];
        t.ez = 0;
        t.half = (c == 2 || c == 3) ? 0.5 : 1.0;
    }
    return t;
}

template <int D>
__device__ inline double a_term_sel(int e, const ATermExp &q, double x, double y, double z, double om, double vm) {
#pragma clang fp contract(off)
    if constexpr (D == 3) {
        double t = 1.0;
#pragma unroll
        for (int i = 0; i < 4; i++) t = t * (q.ex > i ? x : 1.0);
#pragma unroll
        for (int i = 0; i < 4; i++) t = t * (q.ey > i ? y : 1.0);
#pragma unroll
        for (int i = 0; i < 4; i++) t = t * (q.ez > i ? z : 1.0);
        t = t * q.half;
        return (om * vm) * t;
    } else {
        if (e == 0) return ((om * x) * x) * vm;
        double t = om * vm;
#pragma unroll
        for (int i = 0; i < 4; i++) t = t * (q.ex > i ? x : 1.0);
#pragma unroll
        for (int i = 0; i < 4; i++) t = t * (q.ey > i ? y : 1.0);
        return t * q.half;
    }
}

template <>
__device__ inline double a_term<3>(int e, double x, double y, double z, double om, double vm) {
#pragma clang fp contract(off)
    int r = e / 9, c = e % 9;
    int ex = c_pe3[r][0] + c_pe3[c][0], ey = c_pe3[r][1] + c_pe3[c][1], ez = c_pe3[r][2] + c_pe3[c][2];
    double t = 1.0;
    for (int i = 0; i < ex; i++) t = t * x;
    for (int i = 0; i < ey; i++) t = t * y;
    for (int i = 0; i < ez; i++) t = t * z;
    if (c >= 3 && c < 6) t = t * 0.5;
    return (om * vm) * t;
}

template <>
__device__ inline double a_term<2>(int e, double x, double y, double, double om, double vm) {
#pragma clang fp contract(off)
    if (e == 0) return ((om * x) * x) * vm;
    int r = e / 5, c = e % 5;
    int ex = c_pe2[r][0] + c_pe2[c][0], ey = c_pe2[r][1] + c_pe2[c][1];
    double t = om * vm;
    for (int i = 0; i < ex; i++) t = t * x;
    for (int i = 0; i < ey; i++) t = t * y;
    if (c == 2 || c == 3) t = t * 0.5;
    return t;
}

struct FamArgs {
    const double *x, *nv;
    const int *fam_node, *fam_off, *fam_mem, *bond, *fam_blk;
    const double *fam_v, *fam_delta;
    const int *fp_off, *fp_list, *pdbe;
    const double *Dm;
    double lam, mu, nlf;
    double *K;
    double *fam_c;  // [members * d] strain weights mu_km omega g V of every bond (kept for stresses)
    long long nblk;
    int *status;
    int mf;     // LDS capacity in members
    int maxpb;  // LDS capacity in PD-boundary entities per family (3D N.D cache)
    const int *flist = nullptr;  // families of the blocks (null: block f is family f)
};

// normMatArea (pdsolve.cpp:2917-2967): area-weighted face normal matrix 3x6
__device__ inline bool norm_mat_area(const double *x, const int *con, double Nm[18]) {
#pragma clang fp contract(off)
    double xN[4][3];
    for (int a = 0; a < 4; a++)
        for (int c = 0; c < 3; c++) xN[a][c] = x[(size_t)con[a] * 3 + c];
    double v1[3], v2[3], w = 1.0 / 3;
    for (int i = 0; i < 3; i++) { v1[i] = xN[1][i] - xN[0][i]; v2[i] = xN[2][i] - xN[0][i]; }
    double N1 = v1[1] * v2[2] - v1[2] * v2[1];
    double N2 = v2[0] * v1[2] - v1[0] * v2[2];
    double N3 = v1[0] * v2[1] - v2[0] * v1[1];
    double A1 = sqrt(N1 * N1 + N2 * N2 + N3 * N3), A2 = 0;
    if (A1 == 0) return false;
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
    for (int i = 0; i < 18; i++) Nm[i] = 0;
    Nm[0] = N1; Nm[7] = N2; Nm[14] = N3;
    Nm[3] = N2; Nm[5] = N3; Nm[9] = N1;
    Nm[10] = N3; Nm[16] = N2; Nm[17] = N1;
    return true;
}

// C-block of one bond (strain operator columns of member m, matC2D/3D
// pdsolve.cpp:3212-3252, :3324-3398): NV x D
template <int D>
__device__ inline void c_block(const double c[D], double Cb[Dm<D>::NV * D]) {
#pragma clang fp contract(off)
    if constexpr (D == 2) {
        Cb[0] = c[0]; Cb[1] = 0;
        Cb[2] = 0; Cb[3] = c[1];
        Cb[4] = c[1]; Cb[5] = c[0];
    } else {
        for (int i = 0; i < 18; i++) Cb[i] = 0;
        Cb[0 * 3 + 0] = c[0]; Cb[1 * 3 + 1] = c[1]; Cb[2 * 3 + 2] = c[2];
        Cb[3 * 3 + 0] = c[1]; Cb[3 * 3 + 1] = c[0];
        Cb[4 * 3 + 1] = c[2]; Cb[4 * 3 + 2] = c[1];
        Cb[5 * 3 + 0] = c[2]; Cb[5 * 3 + 2] = c[0];
    }
}

// 3D: Nm . D (3 x 6) of PD-boundary entity p into ND; a degenerate face sets the status and
// leaves ND zero (the caller reports the error)
__device__ inline void pdbe_nd3(const FamArgs &A, int p, double *ND) {
#pragma clang fp contract(off)
    double Nm[18];
    if (!norm_mat_area(A.x, &A.pdbe[p * 4], Nm)) {
        atomicOr(A.status, 4);
        for (int i = 0; i < 18; i++) ND[i] = 0.0;
        return;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 6; j++) {
            double t = 0.0;
            for (int k = 0; k < 6; k++) t = Nm[i * 6 + k] * A.Dm[k * 6 + j] + t;
            ND[i * 6 + j] = t;
        }
}

// 3D: (Nm . D) . C with the cached Nm . D, added to val (3 x 3)
__device__ inline void add_pdbe_nd3(const double *ND, const double *Cb, double *val) {
#pragma clang fp contract(off)
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double t = 0.0;
            for (int k = 0; k < 6; k++) t = ND[i * 6 + k] * Cb[k * 3 + j] + t;
            val[i * 3 + j] = val[i * 3 + j] + t;
        }
}

// boundary traction term of one PD-boundary entity for one C block, added to val (D x D)
// 2D: 1/2 * N . (D . C)  (pdsolve.cpp:4049-4253); 3D: (Nm . D) . C  (pdsolve.cpp:2785-2915)
template <int D>
__device__ inline void add_pdbe_term(const FamArgs &A, int p, const double *Cb, double *val) {
#pragma clang fp contract(off)
    constexpr int NV = Dm<D>::NV;
    const double *Dg = A.Dm;
    if constexpr (D == 2) {
        int a = A.pdbe[p * 4], b = A.pdbe[p * 4 + 1];
        double nx = A.x[(size_t)b * 3 + 1] - A.x[(size_t)a * 3 + 1];
        double ny = A.x[(size_t)a * 3 + 0] - A.x[(size_t)b * 3 + 0];
        double N[6] = {nx, 0, ny, 0, ny, nx};
        double DC[6];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 2; j++) {
                double t = 0.0;
                for (int k = 0; k < 3; k++) t = Dg[i * 3 + k] * Cb[k * 2 + j] + t;
                DC[i * 2 + j] = t;
            }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                double t = 0.0;
                for (int k = 0; k < 3; k++) t = N[i * 3 + k] * DC[k * 2 + j] + t;
                val[i * 2 + j] = val[i * 2 + j] + 1.0 / 2 * t;
            }
    } else {
        double Nm[18], ND[18];
        if (!norm_mat_area(A.x, &A.pdbe[p * 4], Nm)) {
            atomicOr(A.status, 4);
            return;
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 6; j++) {
                double t = 0.0;
                for (int k = 0; k < 6; k++) t = Nm[i * 6 + k] * Dg[k * 6 + j] + t;
                ND[i * 6 + j] = t;
            }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double t = 0.0;
                for (int k = 0; k < NV; k++) t = ND[i * 6 + k] * Cb[k * 3 + j] + t;
                val[i * 3 + j] = val[i * 3 + j] + t;
            }
    }
}

template <int D>
__global__ void __launch_bounds__(64) k_family(FamArgs A) {
    // no FMA contraction: the shape tensor of a one-sided family is ill-conditioned, and the
    // reference's arithmetic (separate multiply and add) must be followed to keep K to rounding
#pragma clang fp contract(off)
    constexpr int P = Dm<D>::P, DD = Dm<D>::DD;
    extern __shared__ double lds[];
    const int f = A.flist ? A.flist[blockIdx.x] : (int)blockIdx.x, lane = threadIdx.x;
    const int off = A.fam_off[f], n = A.fam_off[f + 1] - off;
    const int mf = A.mf;
    double *sx = lds;               // [mf]
    double *sy = sx + mf;           // [mf]
    double *sz = sy + mf;           // [mf]
    double *som = sz + mf;          // [mf] influence function
    double *svm = som + mf;         // [mf] member volume
    double *sgv = svm + mf;         // [mf*DD] G_m * V_m
    double *sc = sgv + (size_t)mf * DD;  // [mf*D]  mu_km * om * g * V_m
    double *sND = sc + (size_t)mf * D;   // [maxpb*18] 3D: Nm . D of the family's PD-boundary entities
    double *sA = sND + (D == 3 ? (size_t)A.maxpb * 18 : 0);  // [P*P]
    int *spiv = (int *)(sA + P * P);     // [P] + flag
    if (n > mf) {
        if (lane == 0) atomicOr(A.status, 8);
        return;
    }
    const int k = A.fam_node[f];
    const double xk0 = A.x[(size_t)k * 3], xk1 = A.x[(size_t)k * 3 + 1], xk2 = A.x[(size_t)k * 3 + 2];
    const double delt = A.fam_delta[f];
    const double Aa = A.nlf * A.nlf;
    // stage members (inflFunc pdsolve.cpp:1929-1935)
    for (int m = lane; m < n; m += 64) {
        int j = A.fam_mem[off + m];
        double x0 = A.x[(size_t)j * 3] - xk0, x1 = A.x[(size_t)j * 3 + 1] - xk1, x2 = A.x[(size_t)j * 3 + 2] - xk2;
        sx[m] = x0; sy[m] = x1; sz[m] = x2;
        som[m] = exp(-(x0 * x0 + x1 * x1 + x2 * x2) / Aa / delt / delt);
        svm[m] = A.nv[j];
    }
    __syncthreads();
    // shape tensor, members in reference order.  3D: A[r][c] = half_c * U_rc with U symmetric
    // (the term's half factor is the last, exact, power-of-two scaling, which commutes with every
    // rounding of the sum), so one lane sums each of the 45 entries r <= c and writes both
    // orientations: the same bits as 81 separate sums.  2D: one lane per entry.
    if constexpr (D == 3) {
        if (lane < 45) {
            int r = 0, rem = lane;
            while (rem >= P - r) {
                rem -= P - r;
                r++;
            }
            const int c = r + rem;
            const int e = r * P + c;
            ATermExp q = a_term_exp<D>(e);
            q.half = 1.0;
            double acc = 0.0;
            for (int m = 1; m < n; m++) acc = a_term_sel<D>(e, q, sx[m], sy[m], sz[m], som[m], svm[m]) + acc;
            const double hc = (c >= 3 && c < 6) ? 0.5 : 1.0, hr = (r >= 3 && r < 6) ? 0.5 : 1.0;
            sA[r * P + c] = acc * hc;
            sA[c * P + r] = acc * hr;
        }
    } else {
        for (int e = lane; e < P * P; e += 64) {
            const ATermExp q = a_term_exp<D>(e);
            double acc = 0.0;
            for (int m = 1; m < n; m++) acc = a_term_sel<D>(e, q, sx[m], sy[m], sz[m], som[m], svm[m]) + acc;
            sA[e] = acc;
        }
    }
    __syncthreads();
    // partial-pivot LU (first maximum pivot) in LDS, lane-parallel with the operations of the
    // sequential loop on every entry (so the same bits): per column the pivot row is an argmax
    // over lanes with ties to the lower row (the sequential scan replaces only on '>'), then the
    // row swap, the multipliers l = a_ik / a_kk, and the update a_ic - l a_kc
    {
        int sing = 0;
        for (int kk = 0; kk < P; kk++) {
            double v = -1.0;
            int idx = P;
            if (lane >= kk && lane < P) {
                v = fabs(sA[lane * P + kk]);
                idx = lane;
            }
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) {
                const double ov = __shfl_down(v, o, 16);
                const int oi = __shfl_down(idx, o, 16);
                if (ov > v || (ov == v && oi < idx)) {
                    v = ov;
                    idx = oi;
                }
            }
            const int p = __shfl(idx, 0);
            if (lane == 0) spiv[kk] = p;
            if (sA[p * P + kk] == 0.0) {
                sing = 1;
                break;
            }
            if (p != kk && lane < P) {
                const double t = sA[kk * P + lane];
                sA[kk * P + lane] = sA[p * P + lane];
                sA[p * P + lane] = t;
            }
            __syncthreads();
            if (lane > kk && lane < P) sA[lane * P + kk] = sA[lane * P + kk] / sA[kk * P + kk];
            __syncthreads();
            const int mm = P - kk - 1;
            if (lane < mm * mm) {
                const int i = kk + 1 + lane / mm, c = kk + 1 + lane % mm;
                sA[i * P + c] = sA[i * P + c] - sA[i * P + kk] * sA[kk * P + c];
            }
            __syncthreads();
        }
        if (lane == 0) spiv[P] = sing;
    }
    __syncthreads();
    if (spiv[P]) {
        if (lane == 0) atomicOr(A.status, 1);
        return;
    }
    const double cdv = A.fam_v[f];
    const double lam = A.lam, mu = A.mu;
    const int p0 = A.fp_off[f], p1 = A.fp_off[f + 1];
    if constexpr (D == 3) {
        // Nm . D of every incident PD-boundary entity, once per family instead of once per member
        // (the same products in the same order, so the same bits)
        for (int q = p0 + lane; q < p1; q += 64) pdbe_nd3(A, A.fp_list[q], sND + (size_t)(q - p0) * 18);
        __syncthreads();
    }
    const size_t nb = (size_t)A.nblk;
    for (int m = 1 + lane; m < n; m += 64) {
        const double x0 = sx[m], x1 = sy[m], x2 = sz[m];
        double b[P];
        if constexpr (D == 2) {
            b[0] = x0; b[1] = x1; b[2] = x0 * x0; b[3] = x1 * x1; b[4] = x0 * x1;
        } else {
            b[0] = x0; b[1] = x1; b[2] = x2;
            b[3] = x0 * x0; b[4] = x1 * x1; b[5] = x2 * x2;
            b[6] = x0 * x1; b[7] = x1 * x2; b[8] = x2 * x0;
        }
        // the LU swapped whole rows (multipliers included): permute b completely, then eliminate
#pragma unroll
        for (int kk = 0; kk < P; kk++) {
            const int p = spiv[kk];
#pragma unroll
            for (int i = kk + 1; i < P; i++)
                if (p == i) { double t = b[kk]; b[kk] = b[i]; b[i] = t; }
        }
#pragma unroll
        for (int kk = 0; kk < P; kk++)
#pragma unroll
            for (int i = kk + 1; i < P; i++) b[i] = b[i] - sA[i * P + kk] * b[kk];
        double s[P];
#pragma unroll
        for (int i = P - 1; i >= 0; i--) {
            double t = b[i];
#pragma unroll
            for (int c = i + 1; c < P; c++) t = t - sA[i * P + c] * s[c];
            s[i] = t / sA[i * P + i];
        }
        const double om = som[m], vm = svm[m];
        const double mukm = (double)A.bond[off + m];
        double G[DD];
        // matG2D / matG3D (pdsolve.cpp:2277-2347)
        if constexpr (D == 2) {
            const double *dd = s + 2;
            G[0] = mukm * om * ((2 * mu + lam) * dd[0] + mu * dd[1]);
            G[1] = G[2] = mukm * om * (lam + mu) * dd[2];
            G[3] = mukm * om * ((2 * mu + lam) * dd[1] + mu * dd[0]);
        } else {
            const double *dd = s + 3;
            double tr = dd[0] + dd[1] + dd[2];
            G[0] = mukm * om * ((mu + lam) * dd[0] + mu * tr);
            G[1] = G[3] = mukm * om * (lam + mu) * dd[3];
            G[2] = G[6] = mukm * om * (lam + mu) * dd[5];
            G[4] = mukm * om * ((mu + lam) * dd[1] + mu * tr);
            G[5] = G[7] = mukm * om * (lam + mu) * dd[4];
            G[8] = mukm * om * ((mu + lam) * dd[2] + mu * tr);
        }
        double val[DD];
#pragma unroll
        for (int e = 0; e < DD; e++) {
            double h = G[e] * vm;  // H[i][j + d*m]
            sgv[(size_t)m * DD + e] = h;
            val[e] = -(h * cdv);   // -(H * dv_k), pdsolve.cpp:2562
        }
        double c[D];
#pragma unroll
        for (int i = 0; i < D; i++) {
            c[i] = mukm * om * s[i] * vm;  // matC2D / matC3D (pdsolve.cpp:3212-3252, :3324-3398)
            sc[(size_t)m * D + i] = c[i];
            A.fam_c[(size_t)(off + m) * D + i] = c[i];
        }
        if (p1 > p0) {
            double Cb[Dm<D>::NV * D];
            c_block<D>(c, Cb);
            for (int q = p0; q < p1; q++) {
                if constexpr (D == 3) add_pdbe_nd3(sND + (size_t)(q - p0) * 18, Cb, val);
                else add_pdbe_term<D>(A, A.fp_list[q], Cb, val);
            }
        }
        const size_t blk = (size_t)A.fam_blk[off + m];
#pragma unroll
        for (int e = 0; e < DD; e++) A.K[blk * DD + e] = val[e];
    }
    __syncthreads();
    // self block: sequential sums over members in reference order
    if (lane < DD) {
        const int e = lane;
        double hs = 0.0;
        for (int m = 1; m < n; m++) hs = -sgv[(size_t)m * DD + e] + hs;
        double val[DD];
        for (int q = 0; q < DD; q++) val[q] = 0;
        val[e] = -(hs * cdv);
        if (p1 > p0) {
            double cs[D];
            for (int i = 0; i < D; i++) {
                double t = 0.0;
                for (int m = 1; m < n; m++) t = -sc[(size_t)m * D + i] + t;
                cs[i] = t;
            }
            double Cb[Dm<D>::NV * D];
            c_block<D>(cs, Cb);
            for (int q = p0; q < p1; q++) {
                if constexpr (D == 3) add_pdbe_nd3(sND + (size_t)(q - p0) * 18, Cb, val);
                else add_pdbe_term<D>(A, A.fp_list[q], Cb, val);
            }
        }
        A.K[(size_t)A.fam_blk[off] * DD + e] = val[e];
    }
}

// ------------------------------------------------------------------------
// FE element rows (pdfemEleQuad4N.cpp:252-359, pdfemEleTri3N.cpp:33-47,
// pdfemEleBrick8N.cpp:324-552, pdfemEleTetra4N.cpp:38-170), node-centric
// ------------------------------------------------------------------------



__constant__ double c_hp[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
__constant__ double c_hq[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
__constant__ double c_hr[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};

// row `row` of B^T D B for a 3D element whose B columns are given by dNdX (nn nodes)
// BtD[row][k] = sum_l B[l][row] D[l][k];  (BtD B)[row][c] = sum_k BtD[row][k] B[k][c]
__device__ inline void btdb_row3(const double (*g)[3], int nnod, int la, int ii, const double *Dg, double w,
                                 double *Krow) {
    double Bc[6];  // column `row` of B
    const double *ga = g[la];
    for (int l = 0; l < 6; l++) Bc[l] = 0.0;
    if (ii == 0) { Bc[0] = ga[0]; Bc[3] = ga[1]; Bc[5] = ga[2]; }
    else if (ii == 1) { Bc[1] = ga[1]; Bc[3] = ga[0]; Bc[4] = ga[2]; }
    else { Bc[2] = ga[2]; Bc[4] = ga[1]; Bc[5] = ga[0]; }
    double btd[6];
    for (int k = 0; k < 6; k++) {
        double t = 0.0;
        for (int l = 0; l < 6; l++) t = Bc[l] * Dg[l * 6 + k] + t;
        btd[k] = t;
    }
    for (int b = 0; b < nnod; b++) {
        const double *gb = g[b];
        // B[:, 3b+0] = (gx,0,0,gy,0,gz); B[:,3b+1] = (0,gy,0,gx,gz,0); B[:,3b+2] = (0,0,gz,0,gy,gx)
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        t0 = btd[0] * gb[0] + t0; t0 = btd[3] * gb[1] + t0; t0 = btd[5] * gb[2] + t0;
        t1 = btd[1] * gb[1] + t1; t1 = btd[3] * gb[0] + t1; t1 = btd[4] * gb[2] + t1;
        t2 = btd[2] * gb[2] + t2; t2 = btd[4] * gb[1] + t2; t2 = btd[5] * gb[0] + t2;
        Krow[3 * b + 0] = Krow[3 * b + 0] + t0 * w;
        Krow[3 * b + 1] = Krow[3 * b + 1] + t1 * w;
        Krow[3 * b + 2] = Krow[3 * b + 2] + t2 * w;
    }
}

__device__ inline void btdb_row2(const double (*g)[2], int nnod, int la, int ii, const double *Dg, double w,
                                 double *Krow) {
    double Bc[3];
    const double *ga = g[la];
    if (ii == 0) { Bc[0] = ga[0]; Bc[1] = 0; Bc[2] = ga[1]; }
    else { Bc[0] = 0; Bc[1] = ga[1]; Bc[2] = ga[0]; }
    double btd[3];
    for (int k = 0; k < 3; k++) {
        double t = 0.0;
        for (int l = 0; l < 3; l++) t = Bc[l] * Dg[l * 3 + k] + t;
        btd[k] = t;
    }
    for (int b = 0; b < nnod; b++) {
        const double *gb = g[b];
        double t0 = 0.0, t1 = 0.0;
        t0 = btd[0] * gb[0] + t0; t0 = btd[2] * gb[1] + t0;
        t1 = btd[1] * gb[1] + t1; t1 = btd[2] * gb[0] + t1;
        Krow[2 * b + 0] = Krow[2 * b + 0] + t0 * w;
        Krow[2 * b + 1] = Krow[2 * b + 1] + t1 * w;
    }
}

// Element stiffness of every FE element into a staging array S (one Ke per
// element, row stride ND = 24 in 3D / 8 in 2D).  Hex8 / quad4: a workgroup
// takes EPB elements; first one thread per (element, Gauss point) builds the
// Jacobian, its partial-pivot LU and dN/dX of all nodes into LDS, then one
// thread per (element, Ke row) accumulates B^T D B w detJ over the Gauss
// points in the reference loop order.
// Staged element matrices keep the upper triangle of node blocks only (Ke is symmetric): node
// pair a <= b of an element with at most NB nodes is pair = a NB - a (a - 1) / 2 + b - a of
// NP = NB (NB + 1) / 2, and its d x d block goes to the staging slot the context assigned it:
// the slots of one K node pair (r, c) follow each other in element order, every block in the
// (r, c) orientation, so the gather reads each pair's run front to back.
__host__ __device__ inline int ke_pair(int NB, int a, int b) { return a * NB - a * (a - 1) / 2 + (b - a); }

struct KeArgs {
    const double *x, *Dm;
    double lam, mu;  // isotropic D: D_01 and D_33
    const int *econ, *nen, *fe_elem;
    const int *list;  // FE-list positions of the elements of this type
    int nlist;
    double *S;        // staging: one d x d block per slot
    const int *slot;  // (slot << 1 | transposed) of every (FE-list position, upper node pair)
    int *status;
};

// entry (i, j) of node-pair block (a, b), a <= b, of FE element t into its staging slot, in the
// orientation of the K block the slot feeds
template <int D>
__device__ inline void ke_put(const KeArgs &A, int t, int a, int b, int i, int j, double v) {
    constexpr int NB = D == 3 ? 8 : 4, NP = NB * (NB + 1) / 2;
    const int code = A.slot[(size_t)t * NP + ke_pair(NB, a, b)];
    A.S[(size_t)(code >> 1) * (D * D) + ((code & 1) ? j * D + i : i * D + j)] = v;
}

// Parent derivatives of hex8 node i at Gauss point g (pdfemEleBrick8N::eleStiffMatFEM's order:
// r outer, q, p inner)
__device__ inline void hex8_dN(int g, int i, double d[3]) {
    const double gp = 0.57735026918962573;
    const double gpt[2] = {-gp, gp};
    const double p = gpt[g & 1], qq = gpt[(g >> 1) & 1], r = gpt[g >> 2];
    d[0] = 1.0 / 8 * c_hp[i] * (1 + qq * c_hq[i]) * (1 + r * c_hr[i]);
    d[1] = 1.0 / 8 * c_hq[i] * (1 + p * c_hp[i]) * (1 + r * c_hr[i]);
    d[2] = 1.0 / 8 * c_hr[i] * (1 + p * c_hp[i]) * (1 + qq * c_hq[i]);
}

// Jacobian J at Gauss point g of the element with node coordinates xn, its partial-pivot LU
// (J, piv) and the weight w detJ
template <class XN>
__device__ inline void hex8_gp_lu(const XN &xn, int g, double J[9], int piv[3], double *wdet, int *status) {
#pragma unroll
    for (int i = 0; i < 9; i++) J[i] = 0.0;
    // every J entry sums over the nodes in node order, as the reference's cc-outer loop does
#pragma unroll 2
    for (int i = 0; i < 8; i++) {
        double d[3], xi[3];
        hex8_dN(g, i, d);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) xi[cc] = xn(i, cc);
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {
            J[0 * 3 + cc] = d[0] * xi[cc] + J[0 * 3 + cc];
            J[1 * 3 + cc] = d[1] * xi[cc] + J[1 * 3 + cc];
            J[2 * 3 + cc] = d[2] * xi[cc] + J[2 * 3 + cc];
        }
    }
    const double detJ = det3(J);
    if (!lu3_factor(J, piv)) atomicOr(status, 2);
    *wdet = 1.0 * 1.0 * 1.0 * detJ;
}

// dN/dX of the 8 nodes and detJ at Gauss point g of the element with node coordinates xn,
// into g8o / *wdet; the parent derivatives are formed again per node for the solves instead of
// being kept live
template <class XN>
__device__ inline void hex8_gp(const XN &xn, int g, double (*g8o)[3], double *wdet, int *status) {
    double J[9];
    int piv[3];
    hex8_gp_lu(xn, g, J, piv, wdet, status);
#pragma unroll 2
    for (int b = 0; b < 8; b++) {
        double d[3], o[3];
        hex8_dN(g, b, d);
        lu3_solve(J, piv, d, o);
#pragma unroll
        for (int c = 0; c < 3; c++) g8o[b][c] = o[c];
    }
}

// the same at Gauss point g of element list entry idx, coordinates read from global memory
__device__ inline void hex8_gauss_point(const KeArgs &A, int idx, int g, double (*g8o)[3], double *wdet) {
    const int e = A.fe_elem[A.list[idx]];
    auto xn = [&](int i, int cc) { return A.x[(size_t)A.econ[(size_t)e * 8 + i] * 3 + cc]; };
    hex8_gp(xn, g, g8o, wdet, A.status);
}

constexpr int HEX_EPB = 8;  // 8 elements x 24 rows = 192 threads
__global__ void __launch_bounds__(HEX_EPB * 24) k_ke_hex8(KeArgs A) {
    __shared__ double sg[HEX_EPB][8][8][3];
    __shared__ double sw[HEX_EPB][8];
    const int t = threadIdx.x;
    const int e0 = blockIdx.x * HEX_EPB;
    if (t < HEX_EPB * 8) {
        const int el = t / 8, g = t % 8, idx = e0 + el;
        if (idx < A.nlist) hex8_gauss_point(A, idx, g, sg[el][g], &sw[el][g]);
    }
    __syncthreads();
    const int el = t / 24, row = t % 24, idx = e0 + el;
    if (idx >= A.nlist) return;
    const int la = row / 3, ii = row % 3;
    double Krow[24];
    for (int c = 0; c < 24; c++) Krow[c] = 0.0;
    for (int g = 0; g < 8; g++) btdb_row3(sg[el][g], 8, la, ii, A.Dm, sw[el][g], Krow);
    for (int c = 3 * la; c < 24; c++) ke_put<3>(A, A.list[idx], la, c / 3, ii, c % 3, Krow[c]);
}

// Hex8 with an isotropic D (the reference's D, pdsolve.cpp:37-94, always is): per Gauss point
// the node block B_a^T D B_b w detJ = lam (w g_a) g_b^T + mu g_b (w g_a)^T + mu (w g_a . g_b) I,
// so one thread per (element, node pair a <= b) accumulates M = sum_g (w g_a) g_b^T over the 8
// Gauss points (9 FMA each instead of the 324 of three full B^T D B rows), forms
// lam M + mu M^T + mu tr(M) I once, and writes the block (Ke being symmetric, only a <= b).
constexpr int HEXI_EPB = 4;  // 4 elements: 32 set-up threads, 4 x 36 = 144 block threads
// The element's node coordinates are loaded into LDS by the whole block first (one coordinate
// per thread, so the gathers are in flight together) and the Gauss-point set-up reads them there.
// Each block goes to its staging slot, transposed when the slot's K block is the other
// orientation.  Four elements per 144-thread block: the set-up runs on 32 threads while the
// rest wait, and small blocks let more of them share a CU (one thread per Gauss point for J and
// its LU and one per Gauss point and node for the solves, in 256-thread blocks, was slower).
__global__ void __launch_bounds__(HEXI_EPB * 36) k_ke_hex8_iso(KeArgs A) {
    __shared__ double sx[HEXI_EPB][8][3];
    __shared__ double sw[HEXI_EPB][8];
    __shared__ double sg[HEXI_EPB][8][8][3];
    const int t = threadIdx.x;
    const int e0 = blockIdx.x * HEXI_EPB;
    const int nel = min(HEXI_EPB, A.nlist - e0);
    for (int q = t; q < nel * 24; q += blockDim.x) {
        const int el = q / 24, a = (q % 24) / 3, cc = q % 3;
        const int e = A.fe_elem[A.list[e0 + el]];
        sx[el][a][cc] = A.x[(size_t)A.econ[(size_t)e * 8 + a] * 3 + cc];
    }
    __syncthreads();
    if (t < nel * 8) {
        const int el = t / 8, g = t % 8;
        auto xn = [&](int i, int cc) { return sx[el][i][cc]; };
        hex8_gp(xn, g, sg[el][g], &sw[el][g], A.status);
    }
    __syncthreads();
    const int el = t / 36;
    const bool act = el < nel;
    int a = 0, rem = t % 36;
    while (rem >= 8 - a) {
        rem -= 8 - a;
        a++;
    }
    const int b = a + rem;
    const double lam = A.lam, mu = A.mu;
    // M = sum_g w_g g_a g_b^T over the Gauss points (9 FMA each), then the block
    // lam M + mu M^T + mu tr(M) I
    double M[9];
#pragma unroll
    for (int e = 0; e < 9; e++) M[e] = 0.0;
    if (act) {
#pragma unroll 2
        for (int g = 0; g < 8; g++) {
            const double w = sw[el][g];
            const double wa[3] = {w * sg[el][g][a][0], w * sg[el][g][a][1], w * sg[el][g][a][2]};
            const double gb[3] = {sg[el][g][b][0], sg[el][g][b][1], sg[el][g][b][2]};
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) M[i * 3 + j] = wa[i] * gb[j] + M[i * 3 + j];
        }
    }
    const double tr = M[0] + M[4] + M[8];
    double k[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) k[i * 3 + j] = lam * M[i * 3 + j] + mu * M[j * 3 + i] + (i == j ? mu * tr : 0.0);
    if (act) {
        // pair index t % 36 enumerates a <= b in ke_pair order
        const int code = A.slot[(size_t)A.list[e0 + el] * 36 + t % 36];
        double *out = A.S + (size_t)(code >> 1) * 9;
        if (code & 1) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) out[j * 3 + i] = k[i * 3 + j];
        } else {
#pragma unroll
            for (int e = 0; e < 9; e++) out[e] = k[e];
        }
    }
}

constexpr int QUAD_EPB = 32;  // 32 elements x 8 rows = 256 threads
__global__ void __launch_bounds__(QUAD_EPB * 8) k_ke_quad4(KeArgs A) {
    __shared__ double sg[QUAD_EPB][4][4][2];
    __shared__ double sw[QUAD_EPB][4];
    const int t = threadIdx.x;
    const int e0 = blockIdx.x * QUAD_EPB;
    const double gp = 0.57735026918962573;
    const double gpt[2] = {-gp, gp};
    if (t < QUAD_EPB * 4) {
        const int el = t / 4, g = t % 4, idx = e0 + el;
        if (idx < A.nlist) {
            const int e = A.fe_elem[A.list[idx]];
            double xN[4][2];
            for (int b = 0; b < 4; b++)
                for (int c = 0; c < 2; c++) xN[b][c] = A.x[(size_t)A.econ[(size_t)e * 8 + b] * 3 + c];
            // pdfemEleQuad4N::eleStiffMatFEM: p outer, q inner
            const double p = gpt[g >> 1], qq = gpt[g & 1];
            const double dNdp[4] = {0.25 * (-1 + qq), 0.25 * (1 - qq), 0.25 * (1 + qq), 0.25 * (-1 - qq)};
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
            for (int i = 0; i < 4; i++) {
                sg[el][g][i][0] = i00 * dNdp[i] + i01 * dNdq[i];
                sg[el][g][i][1] = i10 * dNdp[i] + i11 * dNdq[i];
            }
            sw[el][g] = 1.0 * 1.0 * detJ;
        }
    }
    __syncthreads();
    const int el = t / 8, row = t % 8, idx = e0 + el;
    if (idx >= A.nlist) return;
    const int la = row / 2, ii = row % 2;
    double Krow[8];
    for (int c = 0; c < 8; c++) Krow[c] = 0.0;
    for (int g = 0; g < 4; g++) btdb_row2(sg[el][g], 4, la, ii, A.Dm, sw[el][g], Krow);
    for (int c = 2 * la; c < 8; c++) ke_put<2>(A, A.list[idx], la, c / 2, ii, c % 2, Krow[c]);
}

// tri3 / tet4: constant B, Ke = (Bt D B) V (pdfemEleTri3N.cpp:33-47, :185-210;
// pdfemEleTetra4N.cpp:38-170); one thread per (element, row)
template <int D>
__global__ void k_ke_simplex(KeArgs A) {
    constexpr int NEN = D + 1, ND = D == 3 ? 24 : 8;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nlist * NEN * D) return;
    const int idx = t / (NEN * D), row = t % (NEN * D);
    const int e = A.fe_elem[A.list[idx]];
    const int la = row / D, ii = row % D;
    double xN[4][3];
    for (int b = 0; b < NEN; b++)
        for (int c = 0; c < 3; c++) xN[b][c] = A.x[(size_t)A.econ[(size_t)e * 8 + b] * 3 + c];
    double Krow[24];
    for (int c = 0; c < 24; c++) Krow[c] = 0.0;
    double V;
    if constexpr (D == 3) {
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
            for (int c = 0; c < 3; c++) Jv[r * 3 + c] = xN[r + 1][c] - xN[0][c];
        V = fabs(det3(Jv)) / 6.0;
        const double V6 = 6.0 * (fabs(det3(Jv)) / 6.0);
        double g[4][3];
        for (int i = 0; i < 4; i++) { g[i][0] = bb[i] / V6; g[i][1] = cc[i] / V6; g[i][2] = dd[i] / V6; }
        btdb_row3(g, 4, la, ii, A.Dm, 1.0, Krow);
    } else {
        double m3[9] = {1, xN[0][0], xN[0][1], 1, xN[1][0], xN[1][1], 1, xN[2][0], xN[2][1]};
        V = 0.5 * fabs(det3(m3));
        const double V2 = 2.0 * (0.5 * fabs(det3(m3)));
        const double bb[3] = {xN[1][1] - xN[2][1], xN[2][1] - xN[0][1], xN[0][1] - xN[1][1]};
        const double cc[3] = {-xN[1][0] + xN[2][0], -xN[2][0] + xN[0][0], -xN[0][0] + xN[1][0]};
        double g[3][2];
        for (int i = 0; i < 3; i++) { g[i][0] = bb[i] / V2; g[i][1] = cc[i] / V2; }
        btdb_row2(g, 3, la, ii, A.Dm, 1.0, Krow);
    }
    for (int c = D * la; c < NEN * D; c++) ke_put<D>(A, A.list[idx], la, c / D, ii, c % D, Krow[c] * V);
}

// Block-owned add of the staged element blocks, one node pair at a time: one thread per (pair,
// value e) takes value e of an upper block (r, c), r <= c, and the transposed value of its lower
// partner (c, r), starts each from its PD part and adds value e of the pair's element blocks in
// element order (the reference's summation order of FE terms into K), then writes both values
// once.  The element blocks of a pair sit consecutively in the staging, in the (r, c)
// orientation, so the d x d threads of consecutive pairs read the staging front to back, each
// staged value once.
template <int D>
__global__ void k_fe_gather(int npair, const int *__restrict__ pair_up, const int *__restrict__ pair_lo,
                            const int *__restrict__ pair_off, const unsigned char *__restrict__ pd_bits,
                            const double *__restrict__ S, double *K) {
    constexpr int DD = D * D;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)npair * DD) return;
    const int pr = int(t / DD), e = int(t % DD), et = (e % D) * D + e / D;
    const int bu = pair_up[pr], bl = pair_lo[pr];
    const int c0 = pair_off[pr], c1 = pair_off[pr + 1];
    auto pd = [&](int q) { return (pd_bits[q >> 3] >> (q & 7)) & 1; };
    // blocks without a PD part start from zero instead of reading K
    double au = pd(bu) ? K[(size_t)bu * DD + e] : 0.0;
    double al = bl >= 0 && pd(bl) ? K[(size_t)bl * DD + et] : 0.0;
    // in chunks of 4 blocks: the values are loaded before any is added; the sums stay in element order
    for (int cb = c0; cb < c1; cb += 4) {
        const int cnt = min(4, c1 - cb);
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < cnt) v[k] = S[(size_t)(cb + k) * DD + e];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < cnt) {
                au = au + v[k];
                al = al + v[k];
            }
    }
    K[(size_t)bu * DD + e] = au;
    if (bl >= 0) K[(size_t)bl * DD + et] = al;
}

// The same sums for the node pairs with a PD part whose row r (upper block) or column c (lower
// block) belongs to a family recomputed after bond changes (rowflag), writing only those blocks:
// their PD parts are the family kernel's new values, the FE terms the staged element blocks of
// the last whole assembly, added in the same order, so K equals a whole assembly bit for bit.
template <int D>
__global__ void k_fe_gather_rows(int npair, const int *__restrict__ pair_up, const int *__restrict__ pair_lo,
                                 const int *__restrict__ pair_off, const int *__restrict__ pair_row,
                                 const int *__restrict__ adj, const unsigned char *__restrict__ rowflag,
                                 const unsigned char *__restrict__ pd_bits, const double *__restrict__ S, double *K) {
    constexpr int DD = D * D;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)npair * DD) return;
    const int pr = int(t / DD), e = int(t % DD), et = (e % D) * D + e / D;
    const int bu = pair_up[pr], bl = pair_lo[pr];
    const bool du = rowflag[pair_row[pr]], dl = bl >= 0 && rowflag[adj[bu]];
    if (!du && !dl) return;
    const int c0 = pair_off[pr], c1 = pair_off[pr + 1];
    auto pd = [&](int q) { return (pd_bits[q >> 3] >> (q & 7)) & 1; };
    double au = pd(bu) ? K[(size_t)bu * DD + e] : 0.0;
    double al = bl >= 0 && pd(bl) ? K[(size_t)bl * DD + et] : 0.0;
    for (int cb = c0; cb < c1; cb += 4) {
        const int cnt = min(4, c1 - cb);
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < cnt) v[k] = S[(size_t)(cb + k) * DD + e];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < cnt) {
                au = au + v[k];
                al = al + v[k];
            }
    }
    if (du) K[(size_t)bu * DD + e] = au;
    if (dl) K[(size_t)bl * DD + et] = al;
}

// ------------------------------------------------------------------------
// lumped mass (pdsolve.cpp:4618-4669), node-centric over all elements
// ------------------------------------------------------------------------
__global__ void k_lumped(const int *nall_off, const int *nall_ele, const double *ev, const int *nen,
                         const int *eq, double rho, int nn, int D, double *m) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * D) return;
    const int a = t / D, i = t % D;
    double s = 0.0;
    for (int q = nall_off[a]; q < nall_off[a + 1]; q++) {
        const int e = nall_ele[q];
        const double dv = ev[e] / nen[e];
        s = s + dv * rho;
    }
    m[t] = eq[(size_t)a * 3 + i] >= 0 ? s : 1.0;
}

// ------------------------------------------------------------------------
// natural BCs: element forces for a unit traction
// (pdfemEleLine2N.cpp:45-80; pdfemEleQuad4N.cpp:389-456)
// ------------------------------------------------------------------------
__global__ void k_nbc_elements(const double *x, const int *con, int n_ele, int D, double *fe) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_ele) return;
    const double gp = 0.57735026918962573;
    const double gpt[2] = {-gp, gp};
    double *F = &fe[(size_t)e * 12];
    for (int k = 0; k < 12; k++) F[k] = 0.0;
    if (D == 2) {
        const int a = con[e * 2], b = con[e * 2 + 1];
        const double xN[2][2] = {{x[(size_t)a * 3], x[(size_t)a * 3 + 1]}, {x[(size_t)b * 3], x[(size_t)b * 3 + 1]}};
        for (int m = 0; m < 2; m++) {
            const double p = gpt[m], wp = 1.0;
            const double N[2] = {0.5 * (1 - p), 0.5 * (1 + p)}, dNdp[2] = {-0.5, 0.5};
            for (int ii = 0; ii < 2; ii++) {
                double t1 = 0, t2 = 0;
                for (int jj = 0; jj < 2; jj++) { t1 = t1 + dNdp[jj] * xN[jj][1]; t2 = t2 + dNdp[jj] * xN[jj][0]; }
                F[ii * 2] = wp * N[ii] * (1.0 * t1) + F[ii * 2];
                F[ii * 2 + 1] = wp * N[ii] * (-1.0 * t2) + F[ii * 2 + 1];
            }
        }
    } else {
        double xN[4][3];
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 3; c++) xN[a][c] = x[(size_t)con[e * 4 + a] * 3 + c];
        for (int mp = 0; mp < 2; mp++)
            for (int mq = 0; mq < 2; mq++) {
                const double p = gpt[mp], q = gpt[mq];
                double N[4] = {0.25 * (1 - p) * (1 - q), 0.25 * (1 + p) * (1 - q), 0.25 * (1 + p) * (1 + q),
                               0.25 * (1 - p) * (1 + q)};
                double dNdp[4] = {0.25 * (-1 + q), 0.25 * (1 - q), 0.25 * (1 + q), 0.25 * (-1 - q)};
                double dNdq[4] = {0.25 * (-1 + p), 0.25 * (-1 - p), 0.25 * (1 + p), 0.25 * (1 - p)};
                double dXdp[3] = {0, 0, 0}, dXdq[3] = {0, 0, 0};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 4; j++) {
                        dXdp[i] = dXdp[i] + dNdp[j] * xN[j][i];
                        dXdq[i] = dXdq[i] + dNdq[j] * xN[j][i];
                    }
                const double nr[3] = {dXdp[1] * dXdq[2] - dXdp[2] * dXdq[1], dXdq[0] * dXdp[2] - dXdp[0] * dXdq[2],
                                      dXdp[0] * dXdq[1] - dXdq[0] * dXdp[1]};
                for (int k = 0; k < 4; k++)
                    for (int i = 0; i < 3; i++) F[3 * k + i] = F[3 * k + i] + N[k] * nr[i];
            }
    }
}

// node-centric gather of unit element forces into f_unit[set][node*D+i] (element order)
__global__ void k_nbc_gather(const int *off, const int *item, const int *nbc_set, const double *fe, int nn, int D,
                             int n_set, double *f_unit) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * D) return;
    const int a = t / D, i = t % D;
    const int tn = D == 3 ? 4 : 2;
    for (int s = 0; s < n_set; s++) f_unit[(size_t)s * nn * D + t] = 0.0;
    for (int q = off[a]; q < off[a + 1]; q++) {
        const int it = item[q];
        const int e = it / tn, la = it % tn;
        const int s = nbc_set[e];
        f_unit[(size_t)s * nn * D + t] = f_unit[(size_t)s * nn * D + t] + fe[(size_t)e * 12 + D * la + i];
    }
}

// ------------------------------------------------------------------------
// node-block SpMV: y = K x (full DOF vectors), optional mask of prescribed rows
// ------------------------------------------------------------------------
// one launch for both row classes: blocks [0, nbs) take short rows on groups of
// 8 lanes, the remaining blocks take long (PD) rows on whole wavefronts
template <int D, typename T>
__global__ void __launch_bounds__(256) k_spmv(const int *rows_s, int ns, int nbs, const int *rows_l, int nl,
                                              const int *adj_off, const int *adj, const T *K, const double *x,
                                              double *y, const int *eq, int mask) {
    T acc[D];
    int r, lane;
    if ((int)blockIdx.x < nbs) {
        const int gid = (blockIdx.x * blockDim.x + threadIdx.x) / 8;
        lane = threadIdx.x % 8;
        if (gid >= ns) return;
        r = rows_s[gid];
        row_matvec<D, T, double, 8, false>(r, lane, adj_off, adj, K, x, acc);
    } else {
        const int gid = ((blockIdx.x - nbs) * blockDim.x + threadIdx.x) / 64;
        lane = threadIdx.x % 64;
        if (gid >= nl) return;
        r = rows_l[gid];
        row_matvec<D, T, double, 64, false>(r, lane, adj_off, adj, K, x, acc);
    }
    if (lane < D) {
        const T v = pick<D, T>(acc, lane);
        const size_t o = (size_t)r * D + lane;
        y[o] = (mask && eq[(size_t)r * 3 + lane] < 0) ? 0.0 : double(v);
    }
}

// ------------------------------------------------------------------------
// block rows -> reference CSR values of K_ff
// ------------------------------------------------------------------------
__global__ void k_csr_extract(const int *row_node, int neq, const int64_t *ia, const int *adj_off, const int *adj,
                              const int *eq, const double *K, long long nblk, int D, double *Ku) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= neq) return;
    const int code = row_node[r];
    const int k = code / 3, ii = code % 3;
    int64_t p = ia[r];
    const size_t nb = (size_t)nblk;
    for (int b = adj_off[k]; b < adj_off[k + 1]; b++) {
        const int c = adj[b];
        for (int jj = 0; jj < D; jj++)
            if (eq[(size_t)c * 3 + jj] >= 0) Ku[p++] = K[(size_t)b * D * D + ii * D + jj];
    }
}

__global__ void k_fill(double *p, double v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------
int dev_fill(Ctx &c, double *p, double v, size_t n) {
    if (!n) return PDX_OK;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    k_fill<<<blocks, 256, 0, c.stream>>>(p, v, n);
    return hipGetLastError() == hipSuccess ? PDX_OK : PDX_ERR_HIP;
}

static size_t family_lds_bytes(int D, int mf, int maxpb) {
    const int P = D == 2 ? 5 : 9, DD = D * D;
    return (size_t)mf * (5 + DD + D) * sizeof(double) + (D == 3 ? (size_t)maxpb * 18 * sizeof(double) : 0) +
           (size_t)P * P * sizeof(double) + (P + 2) * sizeof(int);
}

// arguments and LDS bytes of the family kernel over all families (A.flist null)
static int family_args(Ctx &c, int *status, FamArgs &A, size_t &lds, std::string &err) {
    A.x = c.x; A.nv = c.nv; A.fam_node = c.fam_node; A.fam_off = c.fam_off; A.fam_mem = c.fam_mem;
    A.bond = c.bond; A.fam_blk = c.fam_blk; A.fam_v = c.fam_v; A.fam_delta = c.fam_delta;
    A.fp_off = c.fp_off; A.fp_list = c.fp_list; A.pdbe = c.pdbe; A.Dm = c.Dm;
    A.lam = c.lambda; A.mu = c.mu; A.nlf = c.nlf; A.K = c.K; A.nblk = c.nblk; A.status = status;
    A.mf = c.max_family;
    A.maxpb = c.max_pdbe_per_family;
    A.flist = nullptr;
    if (!c.fam_c) {
        int nmem = 0;
        PDX_HIP(hipMemcpy(&nmem, c.fam_off + c.nfam, sizeof(int), hipMemcpyDeviceToHost));
        PDX_HIP(hipMalloc((void **)&c.fam_c, std::max<size_t>((size_t)nmem * c.dim, 1) * sizeof(double)));
        PDX_HIP(hipMemsetAsync(c.fam_c, 0, std::max<size_t>((size_t)nmem * c.dim, 1) * sizeof(double), c.stream));
    }
    A.fam_c = c.fam_c;
    lds = family_lds_bytes(c.dim, A.mf, A.maxpb);
    if (lds > 160 * 1024) {
        err = "family too large for LDS (" + std::to_string(c.max_family) + " members)";
        return PDX_ERR_LIMIT;
    }
    return PDX_OK;
}

template <int D>
static int launch_family(const FamArgs &A, int nblocks, size_t lds, hipStream_t s, std::string &err) {
    PDX_HIP(hipFuncSetAttribute((const void *)k_family<D>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_family<D><<<nblocks, 64, lds, s>>>(A);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

// Bond changes applied to a whole assembly of K: the family kernel recomputes the rows of the
// changed families (c.dirty_fams), then the node pairs with a PD part in those rows add their
// staged element blocks again (the staging of the last whole assembly; elements do not change).
static int update_K_bonds(Ctx &c, std::string &err) {
    std::vector<int> fams(c.dirty_fams);
    std::sort(fams.begin(), fams.end());
    fams.erase(std::unique(fams.begin(), fams.end()), fams.end());
    const int nf = (int)fams.size();
    int *status = c.asm_status;
    PDX_HIP(hipMemsetAsync(status, 0, sizeof(int), c.stream));
    if (!c.d_flist) PDX_HIP(hipMalloc((void **)&c.d_flist, std::max(c.nfam, 1) * sizeof(int)));
    if (!c.d_rowflag) PDX_HIP(hipMalloc((void **)&c.d_rowflag, std::max(c.nn, 1)));
    const Model &m = *c.host;
    std::vector<unsigned char> flag(c.nn, 0);
    for (int f : fams) flag[m.fam_node[f]] = 1;
    PDX_HIP(hipMemcpyAsync(c.d_flist, fams.data(), nf * sizeof(int), hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemcpyAsync(c.d_rowflag, flag.data(), c.nn, hipMemcpyHostToDevice, c.stream));
    FamArgs A;
    size_t lds = 0;
    if (int rc = family_args(c, status, A, lds, err)) return rc;
    A.flist = c.d_flist;
    if (int rc = c.dim == 3 ? launch_family<3>(A, nf, lds, c.stream, err) : launch_family<2>(A, nf, lds, c.stream, err))
        return rc;
    const int p0 = c.n_fe_pair_fe, np = c.n_fe_pair - p0;
    if (np > 0) {
        if (c.dim == 3)
            k_fe_gather_rows<3><<<(unsigned)(((long long)np * 9 + 255) / 256), 256, 0, c.stream>>>(
                np, c.fe_pair_up + p0, c.fe_pair_lo + p0, c.pair_off + p0, c.fe_pair_row, c.adj, c.d_rowflag,
                c.pd_bits, c.Ke_stage, c.K);
        else
            k_fe_gather_rows<2><<<(unsigned)(((long long)np * 4 + 255) / 256), 256, 0, c.stream>>>(
                np, c.fe_pair_up + p0, c.fe_pair_lo + p0, c.pair_off + p0, c.fe_pair_row, c.adj, c.d_rowflag,
                c.pd_bits, c.Ke_stage, c.K);
        PDX_HIP(hipGetLastError());
    }
    // the host copies of fams / flag must outlive the copies above
    PDX_HIP(hipStreamSynchronize(c.stream));
    return PDX_OK;
}

// The gather on the context's stream: the pairs without a PD part right away, the others after
// the family kernel on stream fs (when that is the side stream) has finished.
template <int D>
static int gather_pairs(Ctx &c, hipStream_t fs, std::string &err) {
    constexpr int DD = D * D;
    auto launch = [&](int p0, int p1) {
        if (p1 > p0)
            k_fe_gather<D><<<(unsigned)(((long long)(p1 - p0) * DD + 255) / 256), 256, 0, c.stream>>>(
                p1 - p0, c.fe_pair_up + p0, c.fe_pair_lo + p0, c.pair_off + p0, c.pd_bits, c.Ke_stage, c.K);
    };
    if (fs != c.stream) {
        launch(0, c.n_fe_pair_fe);
        PDX_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
        launch(c.n_fe_pair_fe, c.n_fe_pair);
    } else {
        launch(0, c.n_fe_pair);
    }
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

// Enqueue one assembly of K on the context's stream (the status word is read by
// dev_assemble_K_finish).
int dev_assemble_K_launch(Ctx &c, std::string &err) {
    if (!c.asm_status) PDX_HIP(hipMalloc((void **)&c.asm_status, sizeof(int)));
    int *status = c.asm_status;
    PDX_HIP(hipMemsetAsync(status, 0, sizeof(int), c.stream));
    if (c.k_needs_fill && dev_fill(c, c.K, 0.0, (size_t)c.dd * c.nblk)) { err = "fill failed"; return PDX_ERR_HIP; }
    hipStream_t fs = c.stream;  // stream of the family kernel
    if (c.nfam > 0) {
        FamArgs A;
        size_t lds = 0;
        if (int rc = family_args(c, status, A, lds, err)) return rc;
        // with elements to do as well, the families run on the side stream beside the element
        // kernel (they write disjoint arrays: K's PD blocks and fam_c, the element stage); the
        // gather joins both
        if (c.nfe > 0) {
            if (!c.side) PDX_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
            if (!c.ev_fork) PDX_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
            if (!c.ev_join) PDX_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
            PDX_HIP(hipEventRecord(c.ev_fork, c.stream));
            PDX_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
            fs = c.side;
        }
        if (int rc = c.dim == 3 ? launch_family<3>(A, c.nfam, lds, fs, err) : launch_family<2>(A, c.nfam, lds, fs, err))
            return rc;
        if (fs != c.stream) PDX_HIP(hipEventRecord(c.ev_join, fs));
    }
    if (c.nfe > 0) {
        if (!c.Ke_stage)
            PDX_HIP(hipMalloc((void **)&c.Ke_stage, (size_t)std::max<long long>(c.n_stage, 1) * c.dd * sizeof(double)));
        KeArgs A;
        A.x = c.x; A.Dm = c.Dm; A.econ = c.econ; A.nen = c.nen; A.fe_elem = c.fe_elem; A.S = c.Ke_stage;
        A.slot = c.fe_slot;
        A.status = status;
        // isotropic 3D D (lam + 2 mu on the normal diagonal, lam off it, mu on the shear diagonal)
        bool iso = false;
        A.lam = A.mu = 0.0;
        if (c.dim == 3 && c.host) {
            const double *Dh = c.host->D;
            const double lam = Dh[1], mu = Dh[21], tol = 1e-14 * (std::fabs(lam) + std::fabs(mu));
            iso = true;
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    const double want = i < 3 && j < 3 ? lam + (i == j ? 2 * mu : 0.0) : (i == j ? mu : 0.0);
                    if (std::fabs(Dh[i * 6 + j] - want) > tol) iso = false;
                }
            A.lam = lam;
            A.mu = mu;
        }
        if (c.dim == 3) {
            if (c.n_fe_a) {
                A.list = c.fe_list_a; A.nlist = c.n_fe_a;
                if (iso)
                    k_ke_hex8_iso<<<(c.n_fe_a + HEXI_EPB - 1) / HEXI_EPB, HEXI_EPB * 36, 0, c.stream>>>(A);
                else
                    k_ke_hex8<<<(c.n_fe_a + HEX_EPB - 1) / HEX_EPB, HEX_EPB * 24, 0, c.stream>>>(A);
            }
            if (c.n_fe_b) {
                A.list = c.fe_list_b; A.nlist = c.n_fe_b;
                k_ke_simplex<3><<<(c.n_fe_b * 12 + 127) / 128, 128, 0, c.stream>>>(A);
            }
            if (int rc = gather_pairs<3>(c, fs, err)) return rc;
        } else {
            if (c.n_fe_a) {
                A.list = c.fe_list_a; A.nlist = c.n_fe_a;
                k_ke_quad4<<<(c.n_fe_a + QUAD_EPB - 1) / QUAD_EPB, QUAD_EPB * 8, 0, c.stream>>>(A);
            }
            if (c.n_fe_b) {
                A.list = c.fe_list_b; A.nlist = c.n_fe_b;
                k_ke_simplex<2><<<(c.n_fe_b * 6 + 127) / 128, 128, 0, c.stream>>>(A);
            }
            if (int rc = gather_pairs<2>(c, fs, err)) return rc;
        }
        PDX_HIP(hipGetLastError());
    }
    return PDX_OK;
}

// Wait for the assembly and check what its kernels reported.
int dev_assemble_K_finish(Ctx &c, std::string &err) {
    int hs = 0;
    PDX_HIP(hipMemcpyAsync(&hs, c.asm_status, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    if (hs & 1) { err = "singular shape tensor (family with too few non-collinear members)"; return PDX_ERR_SINGULAR; }
    if (hs & 2) { err = "singular element Jacobian"; return PDX_ERR_SINGULAR; }
    if (hs & 4) { err = "Area is zero (PD boundary face)"; return PDX_ERR_MESH; }
    if (hs & 8) { err = "family exceeds the LDS capacity"; return PDX_ERR_LIMIT; }
    c.K_valid = true;
    c.K_full = true;
    c.dirty_fams.clear();
    c.K_gen++;
    c.K32_valid = false;
    return PDX_OK;
}

int dev_assemble_K(Ctx &c, std::string &err) {
    // bond changes on a whole assembly: only the changed families' rows are recomputed
    if (!c.K_valid && c.K_full && c.incremental_bonds && !c.dirty_fams.empty() && c.asm_status) {
        c.K_full = false;  // until the update below has finished
        if (int rc = update_K_bonds(c, err)) return rc;
        return dev_assemble_K_finish(c, err);
    }
    if (int rc = dev_assemble_K_launch(c, err)) return rc;
    return dev_assemble_K_finish(c, err);
}

int dev_lumped_mass(Ctx &c, std::string &err) {
    int th = c.nn * c.dim;
    k_lumped<<<(th + 255) / 256, 256, 0, c.stream>>>(c.nall_off, c.nall_ele, c.ev, c.nen, c.eq, c.rho, c.nn, c.dim,
                                                      c.mlump);
    PDX_HIP(hipGetLastError());
    c.mlump_valid = true;
    return PDX_OK;
}

int dev_nbc_unit(Ctx &c, std::string &err) {
    if (c.n_nbc == 0) return PDX_OK;
    if (c.n_nbc_ele > 0) {
        k_nbc_elements<<<(c.n_nbc_ele + 127) / 128, 128, 0, c.stream>>>(c.x, c.nbc_con, c.n_nbc_ele, c.dim, c.nbc_fe);
        PDX_HIP(hipGetLastError());
    }
    int th = c.nn * c.dim;
    k_nbc_gather<<<(th + 255) / 256, 256, 0, c.stream>>>(c.nbcn_off, c.nbcn_item, c.nbc_set, c.nbc_fe, c.nn, c.dim,
                                                          c.n_nbc, c.f_unit);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

int dev_extract_csr(Ctx &c, const double *Kblk, double *Ku_dev) {
    if (c.neq == 0) return PDX_OK;
    k_csr_extract<<<(c.neq + 255) / 256, 256, 0, c.stream>>>(c.row_node, c.neq, c.ia, c.adj_off, c.adj, c.eq, Kblk,
                                                             c.nblk, c.dim, Ku_dev);
    return hipGetLastError() == hipSuccess ? PDX_OK : PDX_ERR_HIP;
}

template <int D, typename T>
static void launch_spmv(Ctx &c, const T *K, const double *x, double *y, bool mask) {
    const int nbs = (c.n_short * 8 + 255) / 256, nbl = (c.n_long * 64 + 255) / 256;
    if (nbs + nbl == 0) return;
    k_spmv<D, T><<<nbs + nbl, 256, 0, c.stream>>>(c.rows_short, c.n_short, nbs, c.rows_long, c.n_long, c.adj_off,
                                                   c.adj, K, x, y, c.eq, mask);
}

int dev_spmv_K(Ctx &c, const double *K64, const double *x, double *y) {
    if (c.dim == 3) launch_spmv<3, double>(c, K64, x, y, true);
    else launch_spmv<2, double>(c, K64, x, y, true);
    return hipGetLastError() == hipSuccess ? PDX_OK : PDX_ERR_HIP;
}

int dev_spmv(Ctx &c, const double *x, double *y, int prec, bool mask) {
    if (prec == PDX_FP32) {
        if (c.dim == 3) launch_spmv<3, float>(c, c.K32, x, y, mask);
        else launch_spmv<2, float>(c, c.K32, x, y, mask);
    } else {
        if (c.dim == 3) launch_spmv<3, double>(c, c.K, x, y, mask);
        else launch_spmv<2, double>(c, c.K, x, y, mask);
    }
    return hipGetLastError() == hipSuccess ? PDX_OK : PDX_ERR_HIP;
}

}  // namespace pdx
