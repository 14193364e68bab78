// Small dense helpers of the element kernels (assembly and stress recovery).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace pdx {

__device__ inline double det3(const double J[9]) {
#pragma clang fp contract(off)
    // cofactor form of pdfemEles::detMat33 (pdfemEles.cpp:155-172)
    double det = 0.0;
    for (int i = 0; i < 3; i++) {
        double a1 = 1., a2 = 1.;
        for (int j = 0; j < 3; j++) {
            a1 *= J[j * 3 + (i + j) % 3];
            a2 *= J[j * 3 + (i - j + 3) % 3];
        }
        det = det + a1 - a2;
    }
    return det;
}

// 3x3 partial-pivot LU solve for 8 right-hand sides (PLUSolve per shape function,
// pdfemEleBrick8N.cpp:380)
__device__ inline bool lu3_solve8(double J[9], double dN[8][3], double out[8][3]) {
#pragma clang fp contract(off)
    int piv[3];
    for (int k = 0; k < 3; k++) {
        int p = k;
        double mx = fabs(J[k * 3 + k]);
        for (int i = k + 1; i < 3; i++)
            if (fabs(J[i * 3 + k]) > mx) { mx = fabs(J[i * 3 + k]); p = i; }
        piv[k] = p;
        if (J[p * 3 + k] == 0.0) return false;
        if (p != k)
            for (int c = 0; c < 3; c++) { double t = J[k * 3 + c]; J[k * 3 + c] = J[p * 3 + c]; J[p * 3 + c] = t; }
        for (int i = k + 1; i < 3; i++) {
            double l = J[i * 3 + k] / J[k * 3 + k];
            J[i * 3 + k] = l;
            for (int c = k + 1; c < 3; c++) J[i * 3 + c] = J[i * 3 + c] - l * J[k * 3 + c];
        }
    }
    for (int a = 0; a < 8; a++) {
        double b0 = dN[a][0], b1 = dN[a][1], b2 = dN[a][2];
        double bb[3] = {b0, b1, b2};
        // rows were swapped whole (multipliers included): permute first, then eliminate
        for (int k = 0; k < 3; k++) {
            int p = piv[k];
            if (p != k) { double t = bb[k]; bb[k] = bb[p]; bb[p] = t; }
        }
        for (int k = 0; k < 3; k++)
            for (int i = k + 1; i < 3; i++) bb[i] = bb[i] - J[i * 3 + k] * bb[k];
        for (int i = 2; i >= 0; i--) {
            double t = bb[i];
            for (int c = i + 1; c < 3; c++) t = t - J[i * 3 + c] * out[a][c];
            out[a][i] = t / J[i * 3 + i];
        }
    }
    return true;
}

// The same partial-pivot LU as lu3_solve8, split into a factorisation and one solve per
// right-hand side, with every row swap written with compile-time indices (a selected swap per
// candidate row) so that J and the right-hand side stay in registers.
__device__ inline bool lu3_factor(double J[9], int piv[3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int p = k;
        double mx = fabs(J[k * 3 + k]);
#pragma unroll
        for (int i = k + 1; i < 3; i++)
            if (fabs(J[i * 3 + k]) > mx) { mx = fabs(J[i * 3 + k]); p = i; }
        piv[k] = p;
#pragma unroll
        for (int i = k + 1; i < 3; i++)
            if (p == i)
#pragma unroll
                for (int c = 0; c < 3; c++) { const double t = J[k * 3 + c]; J[k * 3 + c] = J[i * 3 + c]; J[i * 3 + c] = t; }
        if (J[k * 3 + k] == 0.0) return false;
#pragma unroll
        for (int i = k + 1; i < 3; i++) {
            const double l = J[i * 3 + k] / J[k * 3 + k];
            J[i * 3 + k] = l;
#pragma unroll
            for (int c = k + 1; c < 3; c++) J[i * 3 + c] = J[i * 3 + c] - l * J[k * 3 + c];
        }
    }
    return true;
}

__device__ inline void lu3_solve(const double J[9], const int piv[3], const double b[3], double out[3]) {
#pragma clang fp contract(off)
    double bb[3] = {b[0], b[1], b[2]};
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int i = k + 1; i < 3; i++)
            if (piv[k] == i) { const double t = bb[k]; bb[k] = bb[i]; bb[i] = t; }
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int i = k + 1; i < 3; i++) bb[i] = bb[i] - J[i * 3 + k] * bb[k];
#pragma unroll
    for (int i = 2; i >= 0; i--) {
        double t = bb[i];
#pragma unroll
        for (int c = i + 1; c < 3; c++) t = t - J[i * 3 + c] * out[c];
        out[i] = t / J[i * 3 + i];
    }
}

}  // namespace pdx
