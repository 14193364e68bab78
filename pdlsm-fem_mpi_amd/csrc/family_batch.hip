// Batched PDLSM family kernel: every node of a PD node cloud is a family with
// its own horizon; per family the shape tensor A (FP64, p x p), a Cholesky
// factor of its symmetric part, and the bond blocks H_m = G_m V_m (FP32, d x d
// per member, self block H_0 = -sum_m H_m) -- shapTens3D, vec_gd3D, matG3D and
// matH3D of pdsolve.cpp:1938-2416 for a whole batch.
//
// A = S diag(s) with S = sum_m w_m p_m p_m^T symmetric positive definite
// (w_m = omega_m V_m > 0, p = [x,y,z,x^2,y^2,z^2,xy,yz,zx], 2D [x,y,x^2,y^2,xy])
// and s = 1/2 on the squared terms.  The bond stiffness is linear in the
// solution x_m = diag(1/s) S^-1 p_m of the per-bond system, G_m = w_m M x_m|d
// (M: the constant map of matG from the d-part of x to the d x d block), so
// G_m = w_m T p_m with one p x p matrix per family, T = M diag(1/s)|d S^-1|d.
// Per family, on one wave:
//   * S on the matrix cores: v_mfma_f64_16x16x4f64 over the members four at a
//     time (a lane holds w p_i and p_i of one member for the A and B operands;
//     the 9x9 / 5x5 result sits in the 16x16 tile);
//   * S = L L^T in LDS, column by column over the lanes (no pivoting is needed
//     for an SPD matrix; the reference's per-bond dgesv factors A with partial
//     pivoting, which gives the same solution to rounding);
//   * the rows of S^-1 that T needs: lane k solves S y = e_k with L;
//   * every bond block as a second product on the matrix cores,
//     G(e, m) = w_m sum_k T[e][k] p_k(m), sixteen members per tile.
// Only the bond vectors xi_m and the weights w_m are staged in LDS; the basis
// values are formed from them in registers where each product needs them.
// Launch forms of the same per-family code: one family per 64-thread block, or
// persistent waves (1, 2 or 4 per block) taking chunks of 4 or 16 families from
// eight atomic counters, one per contiguous eighth of the batch; the fastest is
// timed once per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pdx.h"
#include "device.hpp"
#include "setup.hpp"

namespace pdx {

template <int D>
struct FbDim;
template <>
struct FbDim<2> {
    static constexpr int P = 5, DD = 4;
};
template <>
struct FbDim<3> {
    static constexpr int P = 9, DD = 9;
};

struct FbArgs {
    long long nf;
    const double *x, *vol, *delta;  // [n*3], [n], [n]
    const long long *off;           // [nf+1]
    const int *mem;                 // members, self first
    double nlf, lam, mu;
    int mf;                         // largest family (LDS slot size)
    double *A, *LU;                 // [nf*P*P]: A; lower Cholesky factor of its symmetric part
    float *H;                       // [off[nf]*D*D]
    int *status;                    // bit 1: singular shape tensor, bit 8: family larger than mf
    unsigned int *counter;          // persistent work counters, FB_GROUPS of them FB_CSTRIDE apart
};

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int D>
__device__ inline bool is_square(int c) {
    return D == 3 ? (c >= 3 && c < 6) : (c == 2 || c == 3);
}

// slot layout (doubles): member table [(mf + 16) * 4] of [xi_0, xi_1, xi_2, w] rows, zero past the
// family so that the 4- and 16-member products need no bounds test; S / L [P*16] (row-major,
// ld 16); T [DD*12] (ld 12); reciprocal diagonal of L [P, rounded up to even]; the constants 0, 1
template <int D>
__host__ __device__ constexpr int fb_slot_doubles(int mf) {
    return (mf + 16) * 4 + FbDim<D>::P * 16 + FbDim<D>::DD * 12 + ((FbDim<D>::P + 1) & ~1) + 2;
}

typedef double fb_d4 __attribute__((ext_vector_type(4)));

// basis function b of member m read from the slot as slot[a0 + 4 m sa] * slot[b0 + 4 m sb]: the
// coordinate (times the constant 1) or the product of two coordinates; 0 * 1 past the basis.
// Offsets in doubles; the products are the same products as p(xi) written out.
struct FbRead {
    int a0, sa, b0, sb;
};

template <int D>
__device__ inline FbRead fb_read(int b, int zero, int one) {
    int ia = -1, ib = -1;
    if constexpr (D == 3) {
        if (b < 3) {
            ia = b;
        } else if (b < 6) {
            ia = ib = b - 3;
        } else if (b < 9) {  // xy, yz, zx
            ia = b - 6;
            ib = b == 8 ? 0 : b - 5;
        }
    } else {
        if (b < 2) {
            ia = b;
        } else if (b < 4) {
            ia = ib = b - 2;
        } else if (b == 4) {  // xy
            ia = 0;
            ib = 1;
        }
    }
    FbRead r;
    r.a0 = ia < 0 ? zero : ia;
    r.sa = ia < 0 ? 0 : 1;
    r.b0 = ib < 0 ? one : ib;
    r.sb = ib < 0 ? 0 : 1;
    return r;
}


// one family on one wave; slot = its LDS region
template <int D>
__device__ void fb_family(const FbArgs &a, long long f, int lane, double *slot) {
    constexpr int P = FbDim<D>::P, PP = P * P, DD = FbDim<D>::DD;
    constexpr int NK = (P + 3) / 4, NE = (DD + 3) / 4;  // k-steps of the bond product, result items used
    const long long o0 = a.off[f];
    const int n = int(a.off[f + 1] - o0);
    if (n > a.mf) {
        if (lane == 0) atomicOr(a.status, 8);
        return;
    }
    double *sx = slot;                 // [(mf+16)*4] xi_m and w_m (w = 0 for the self member)
    double *sS = sx + (a.mf + 16) * 4;  // [P*16] S, then L in the lower triangle
    double *sT = sS + P * 16;          // [DD*12] T
    double *srd = sT + DD * 12;        // [P] 1 / L_ii
    const int zero = int(srd - slot) + ((P + 1) & ~1), one = zero + 1;
    const int k = a.mem[o0];
    const double xk = a.x[(size_t)k * 3], yk = a.x[(size_t)k * 3 + 1], zk = a.x[(size_t)k * 3 + 2];
    // inflFunc (pdsolve.cpp:1929-1935), exp(-|xi|^2 / (NLF delta_k)^2), with one scale per family
    const double nsc = -1.0 / (a.nlf * a.nlf * a.delta[k] * a.delta[k]);
    auto stage = [&](int m, int j, double cx0, double cx1, double cx2, double v) {
        double x0 = 0.0, x1 = 0.0, x2 = 0.0, w = 0.0;
        if (m < n) {
            x0 = cx0 - xk;
            x1 = cx1 - yk;
            x2 = cx2 - zk;
            // times the member volume
            w = m == 0 ? 0.0 : exp((x0 * x0 + x1 * x1 + x2 * x2) * nsc) * v;
        }
        sx[m * 4] = x0;
        sx[m * 4 + 1] = x1;
        sx[m * 4 + 2] = x2;
        sx[m * 4 + 3] = w;
    };
    constexpr int MS = 4;  // rows per lane of the staging fast path
    if (n + 16 <= 64 * MS) {
        // every member index, then every coordinate load, issued before any is used (index clamped
        // into the family, so no load is conditional)
        int jj[MS];
#pragma unroll
        for (int t = 0; t < MS; t++) jj[t] = a.mem[o0 + min(lane + 64 * t, n - 1)];
        double cx[MS][4];
#pragma unroll
        for (int t = 0; t < MS; t++) {
            cx[t][0] = a.x[(size_t)jj[t] * 3];
            cx[t][1] = a.x[(size_t)jj[t] * 3 + 1];
            cx[t][2] = a.x[(size_t)jj[t] * 3 + 2];
            cx[t][3] = a.vol[jj[t]];
        }
#pragma unroll
        for (int t = 0; t < MS; t++) {
            const int m = lane + 64 * t;
            if (m < n + 16) stage(m, jj[t], cx[t][0], cx[t][1], cx[t][2], cx[t][3]);
        }
    } else {
        for (int m = lane; m < n + 16; m += 64) {
            const int j = a.mem[o0 + min(m, n - 1)];
            stage(m, j, a.x[(size_t)j * 3], a.x[(size_t)j * 3 + 1], a.x[(size_t)j * 3 + 2], a.vol[j]);
        }
    }
    if (lane == 0) {
        slot[zero] = 0.0;
        slot[one] = 1.0;
    }
    wave_sync();
    // S = sum_m w_m p_m p_m^T on the matrix cores: A-operand lane (i = lane%16, member k0 + lane/16)
    // holds w p_i, the B-operand lane (member k0 + lane/16, column j = lane%16 = i) holds p_i
    const int i16 = lane & 15, kq = lane >> 4;
    {
        const FbRead ri = fb_read<D>(i16, zero, one);
        // running pointers to member kq's factors and weight, four members (16 doubles) per step
        const double *pa = slot + ri.a0 + 4 * kq * ri.sa, *pb = slot + ri.b0 + 4 * kq * ri.sb;
        const double *pw = sx + 4 * kq + 3;
        const int da = 16 * ri.sa, db = 16 * ri.sb;
        fb_d4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += 4) {  // members past n are zero rows
            const double bi = pa[0] * pb[0];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(pw[0] * bi, bi, acc, 0, 0, 0);
            pa += da;
            pb += db;
            pw += 16;
        }
        // result layout (checked on gfx950): lane l item r = S[l/16 + 4 r][l%16]
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = kq + 4 * r, col = i16;
            if (row < P && col < P) {
                sS[row * 16 + col] = acc[r];
                a.A[f * PP + row * P + col] = is_square<D>(col) ? 0.5 * acc[r] : acc[r];
            }
        }
    }
    wave_sync();
    // S = L L^T, right-looking over the columns; every entry is reduced in the
    // order q = 0, 1, ... of the left-looking formula
    bool sing = false;
#pragma unroll
    for (int c = 0; c < P; c++) {
        const double d = sS[c * 16 + c];
        sing |= !(d > 0.0);
        // 1 / sqrt(d): the hardware estimate and two Newton steps (relative error of the estimate
        // below 2^-22, so ~2^-88 before rounding), then sqrt(d) = d / sqrt(d)
        const double dp = d > 0.0 ? d : 1.0;
        double il = __builtin_amdgcn_rsq(dp);
        il = fma(0.5 * il, fma(-dp * il, il, 1.0), il);
        il = fma(0.5 * il, fma(-dp * il, il, 1.0), il);
        const double lc = dp * il;
        if (lane > c && lane < P) sS[lane * 16 + c] *= il;
        if (lane == c) {
            sS[c * 16 + c] = lc;
            srd[c] = il;
        }
        wave_sync();
        const int r = c + 1 + (lane >> 3), q = c + 1 + (lane & 7);
        if (r < P && q <= r) sS[r * 16 + q] -= sS[r * 16 + c] * sS[q * 16 + c];
        wave_sync();
    }
    for (int e = lane; e < PP; e += 64) a.LU[f * PP + e] = (e % P) <= (e / P) ? sS[(e / P) * 16 + (e % P)] : 0.0;
    if (sing) {  // uniform: every lane read the same diagonal
        if (lane == 0) atomicOr(a.status, 1);
        return;
    }
    const double lam = a.lam, mu = a.mu;
    if (lane < P) {
        // column `lane` of S^-1: L z = e_lane, L^T y = z
        double z[P];
#pragma unroll
        for (int r = 0; r < P; r++) {
            double t = r == lane ? 1.0 : 0.0;
#pragma unroll
            for (int q = 0; q < r; q++) t -= sS[r * 16 + q] * z[q];
            z[r] = t * srd[r];
        }
        double y[P];
#pragma unroll
        for (int r = P - 1; r >= 0; r--) {
            double t = z[r];
#pragma unroll
            for (int q = r + 1; q < P; q++) t -= sS[q * 16 + r] * y[q];
            y[r] = t * srd[r];
        }
        // x = y / s on the d-part; column `lane` of T = M x|d by matG (pdsolve.cpp:2277-2347) with w = 1
        double dd[P - D];
#pragma unroll
        for (int t = 0; t < P - D; t++) dd[t] = is_square<D>(D + t) ? 2.0 * y[D + t] : y[D + t];
        double T[DD];
        if constexpr (D == 2) {
            T[0] = (2 * mu + lam) * dd[0] + mu * dd[1];
            T[1] = T[2] = (lam + mu) * dd[2];
            T[3] = (2 * mu + lam) * dd[1] + mu * dd[0];
        } else {
            const double tr = dd[0] + dd[1] + dd[2];
            T[0] = (mu + lam) * dd[0] + mu * tr;
            T[1] = T[3] = (lam + mu) * dd[3];
            T[2] = T[6] = (lam + mu) * dd[5];
            T[4] = (mu + lam) * dd[1] + mu * tr;
            T[5] = T[7] = (lam + mu) * dd[4];
            T[8] = (mu + lam) * dd[2] + mu * tr;
        }
#pragma unroll
        for (int e = 0; e < DD; e++) sT[e * 12 + lane] = T[e];
    }
    wave_sync();
    // G(e, m) = w_m sum_k T[e][k] p_k(m): A-operand lane (e = lane%16, k = 4 t + lane/16),
    // B-operand lane (k = 4 t + lane/16, member j0 + lane%16)
    double ta[NK];
    FbRead rk[NK];
#pragma unroll
    for (int t = 0; t < NK; t++) {
        const int kk = 4 * t + kq;
        ta[t] = (i16 < DD && kk < P) ? sT[i16 * 12 + kk] : 0.0;
        rk[t] = fb_read<D>(kk, zero, one);
    }
    double hs[NE];
#pragma unroll
    for (int r = 0; r < NE; r++) hs[r] = 0.0;
    // running pointers to member j0 + i16's factors, sixteen members (64 doubles) per tile
    const double *pa[NK], *pb[NK];
    int da[NK], db[NK];
#pragma unroll
    for (int t = 0; t < NK; t++) {
        pa[t] = slot + rk[t].a0 + 4 * (1 + i16) * rk[t].sa;
        pb[t] = slot + rk[t].b0 + 4 * (1 + i16) * rk[t].sb;
        da[t] = 64 * rk[t].sa;
        db[t] = 64 * rk[t].sb;
    }
    for (int j0 = 1; j0 < n; j0 += 16) {  // members past n are zero rows
        const int m = j0 + i16;
        const bool ok = m < n;
        const double w = sx[m * 4 + 3];
        fb_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < NK; t++) {
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ta[t], pa[t][0] * pb[t][0], acc, 0, 0, 0);
            pa[t] += da[t];
            pb[t] += db[t];
        }
        // lane l item r = G(l/16 + 4 r, member j0 + l%16) / w
        float *h = a.H + (size_t)(o0 + m) * DD;
#pragma unroll
        for (int r = 0; r < NE; r++) {
            const int e = kq + 4 * r;
            if (e < DD) {
                const double g = w * acc[r];
                if (ok) h[e] = float(g);
                hs[r] += g;
            }
        }
    }
    // self block H_0 = -sum_m H_m: reduce over the 16 lanes that hold the same entries
#pragma unroll
    for (int r = 0; r < NE; r++) {
        double v = hs[r];
        for (int sft = 8; sft > 0; sft >>= 1) v += __shfl_xor(v, sft, 64);
        hs[r] = v;
    }
    if (i16 == 0) {
        float *h = a.H + (size_t)o0 * DD;
#pragma unroll
        for (int r = 0; r < NE; r++) {
            const int e = kq + 4 * r;
            if (e < DD) h[e] = float(-hs[r]);
        }
    }
}

template <int D>
__global__ void __launch_bounds__(64) k_fb_block(FbArgs a) {
    extern __shared__ double lds[];
    const long long f = blockIdx.x;
    if (f < a.nf) fb_family<D>(a, f, threadIdx.x, lds);
}

// persistent work split: FB_GROUPS contiguous ranges of the batch, each with its own counter on its own
// 128-byte line; block b starts on range b % FB_GROUPS (the XCD the hardware deals it to), so each
// XCD's L2 sees one slab of the node cloud, then helps with the other ranges once its own is drained
constexpr int FB_GROUPS = 8, FB_CSTRIDE = 32;

// (at most 96 VGPRs: five waves per SIMD, above what the LDS slots allow, with no spill)
template <int D, int CH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_fb_persistent(FbArgs a) {
    extern __shared__ double lds[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int slot0 = wave * fb_slot_doubles<D>(a.mf);
    const int g0 = blockIdx.x % FB_GROUPS;
    for (int gi = 0; gi < FB_GROUPS; gi++) {
        const int g = (g0 + gi) % FB_GROUPS;
        const long long lo = a.nf * g / FB_GROUPS, hi = a.nf * (g + 1) / FB_GROUPS;
        for (;;) {
            // chunk start in a scalar register (fewer than 2^31 families per range; a counter
            // overshoots its range by at most CH per wave)
            unsigned int c0 = 0;
            if (lane == 0) c0 = atomicAdd(a.counter + g * FB_CSTRIDE, (unsigned int)CH);
            c0 = __builtin_amdgcn_readfirstlane(c0);
            if (lo + (long long)c0 >= hi) break;  // uniform across the wave: every wave leaves every range
            const long long fe = lo + c0 + CH < hi ? lo + c0 + CH : hi;
            for (long long f = lo + c0; f < fe; f++) {
                // the lane index and the slot offset pass through an empty asm per family, so that the
                // addresses derived from them are formed inside the loop and not held (and spilled)
                // across it; the pointer is formed after the asm so that it stays an LDS pointer
                int ln = lane, so = slot0;
                asm volatile("" : "+v"(ln), "+s"(so));
                fb_family<D>(a, f, ln, lds + so);
                wave_sync();
            }
        }
    }
}

}  // namespace pdx

// ------------------------------------------------------------------------
// host
// ------------------------------------------------------------------------
struct pdx_fbatch {  // handle behind the opaque C-ABI pointer
    int dim = 3, device = 0;
    hipStream_t stream = nullptr;
    long long n = 0, nf = 0, nbond = 0;
    int mf = 0;
    std::vector<long long> off;
    double nlf = 0.33, lam = 0, mu = 0;
    double *x = nullptr, *vol = nullptr, *delta = nullptr, *A = nullptr, *LU = nullptr;
    long long *doff = nullptr;
    int *dmem = nullptr, *status = nullptr;
    float *H = nullptr;
    unsigned int *counter = nullptr;
    int tuned = -1;
    float tune_ms[PDX_FB_VARIANTS] = {0};
    int n_sm = 256;
    size_t lds_max = 65536;  // LDS bytes a block may use on this device
};

namespace pdx {

// launch forms: 0 one family per block; 1..6 persistent with (waves per block, families per
// chunk) = (4, 4), (2, 4), (1, 4), (4, 16), (2, 16), (1, 16)
constexpr int FB_VARIANTS = PDX_FB_VARIANTS;

static FbArgs fb_args(pdx_fbatch &b) {
    FbArgs a;
    a.nf = b.nf;
    a.x = b.x;
    a.vol = b.vol;
    a.delta = b.delta;
    a.off = b.doff;
    a.mem = b.dmem;
    a.nlf = b.nlf;
    a.lam = b.lam;
    a.mu = b.mu;
    a.mf = b.mf;
    a.A = b.A;
    a.LU = b.LU;
    a.H = b.H;
    a.status = b.status;
    a.counter = b.counter;
    return a;
}

template <int D, int WPB, int CH>
static hipError_t launch_persistent(pdx_fbatch &b, FbArgs a) {
    const size_t lds = (size_t)WPB * fb_slot_doubles<D>(b.mf) * sizeof(double);
    if (lds > b.lds_max) return hipErrorInvalidValue;  // this form does not fit the families
    const void *fn = (const void *)k_fb_persistent<D, CH>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_sm = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_sm, fn, WPB * 64, lds);
    if (e != hipSuccess) return e;
    const int grid = std::max(1, per_sm) * b.n_sm;
    e = hipMemsetAsync(b.counter, 0, FB_GROUPS * FB_CSTRIDE * sizeof(unsigned int), b.stream);
    if (e != hipSuccess) return e;
    k_fb_persistent<D, CH><<<grid, WPB * 64, lds, b.stream>>>(a);
    return hipGetLastError();
}

template <int D>
static hipError_t fb_launch(pdx_fbatch &b, int variant) {
    FbArgs a = fb_args(b);
    switch (variant) {
    case 1: return launch_persistent<D, 4, 4>(b, a);
    case 2: return launch_persistent<D, 2, 4>(b, a);
    case 3: return launch_persistent<D, 1, 4>(b, a);
    case 4: return launch_persistent<D, 4, 16>(b, a);
    case 5: return launch_persistent<D, 2, 16>(b, a);
    case 6: return launch_persistent<D, 1, 16>(b, a);
    default: {
        const size_t lds = (size_t)fb_slot_doubles<D>(b.mf) * sizeof(double);
        if (lds > b.lds_max) return hipErrorInvalidValue;
        hipError_t e = hipFuncSetAttribute((const void *)k_fb_block<D>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
        k_fb_block<D><<<(unsigned)b.nf, 64, lds, b.stream>>>(a);
        return hipGetLastError();
    }
    }
}

static hipError_t fb_run(pdx_fbatch &b, int variant) {
    return b.dim == 3 ? fb_launch<3>(b, variant) : fb_launch<2>(b, variant);
}

// Members of every node's family by the rule of initialPDN_Fami
// (pdsolve.cpp:1442-1568): j joins k iff 0 < |x_j - x_k|^2 (beyond 1e-16 of the
// smallest horizon squared) and |x_j - x_k| is within delta_k or within delta_j.
// Cell list of edge max(delta); the 27 (2D: 9) neighbouring cells are scanned
// x-cell outermost, then y, then z, nodes in node order inside a cell (the
// reference's block scan order).  Self first.
// Members of every node's family (self first, then the 3 x 3 x 3 cells around the node's cell,
// x outer, z inner, each cell's nodes in node order; tiny < d2 and (d2 < delta_k^2 + tiny or
// d2 < delta_j^2 + tiny), tiny = 1e-16 delta_min^2) on the GPU.  The cells are delta_max wide
// and clamp to the grid; the cell lists are built here by a counting sort that keeps node order.
// The members stay on the device (*dmem); off comes back to the host.
static int fb_members(int device, int dim, long long n, const double *x, const double *delta,
                      std::vector<long long> &off, int **dmem, int &mf, std::string &err) {
    if (n > (long long)INT32_MAX) {
        err = "pdx_fb_create: more than 2^31 nodes";
        return PDX_ERR_ARG;
    }
    CellGrid g{};
    double hi[3] = {-1e300, -1e300, -1e300}, dmax = 0, dmin = 1e300;
    for (int i = 0; i < 3; i++) g.lo[i] = 1e300;
    for (long long k = 0; k < n; k++) {
        for (int i = 0; i < 3; i++) {
            g.lo[i] = std::min(g.lo[i], x[k * 3 + i]);
            hi[i] = std::max(hi[i], x[k * 3 + i]);
        }
        dmax = std::max(dmax, delta[k]);
        dmin = std::min(dmin, delta[k]);
    }
    const double tiny = dmin * dmin * 1.0e-16;
    g.cs = dmax > 0 ? dmax : 1.0;
    g.dim = dim;
    g.clamp = true;
    for (int i = 0; i < 3; i++) g.nb[i] = i < dim ? std::max(1, (int)std::floor((hi[i] - g.lo[i]) / g.cs) + 1) : 1;
    const long long ncell = (long long)g.nb[0] * g.nb[1] * g.nb[2];
    if (ncell > 400000000LL) {
        err = "pdx_fb_create: cell grid too large";
        return PDX_ERR_ARG;
    }
    auto cell_of = [&](const double *p) {
        long long c = 0, stride = 1;
        for (int i = 0; i < dim; i++) {
            const int ci = std::min(g.nb[i] - 1, std::max(0, (int)std::floor((p[i] - g.lo[i]) / g.cs)));
            c += ci * stride;
            stride *= g.nb[i];
        }
        return c;
    };
    std::vector<long long> cell((size_t)n);
    std::vector<int> coff((size_t)ncell + 1, 0), cnode((size_t)n);
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < n; k++) cell[k] = cell_of(&x[k * 3]);
    for (long long k = 0; k < n; k++) coff[cell[k] + 1]++;
    for (long long c = 0; c < ncell; c++) coff[c + 1] += coff[c];
    {
        std::vector<int> fill(coff.begin(), coff.end() - 1);
        for (long long k = 0; k < n; k++) cnode[fill[cell[k]]++] = int(k);
    }
    return dev_cell_search(device, int(n), x, delta, coff, cnode, g, n, nullptr, tiny, off, nullptr, mf, dmem, err);
}

}  // namespace pdx

using namespace pdx;

extern "C" {

void pdx_fb_destroy(pdx_fbatch *b) {
    if (!b) return;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    void *ptrs[] = {b->x, b->vol, b->delta, b->A, b->LU, b->doff, b->dmem, b->status, b->H, b->counter};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

int pdx_fb_create(int device, int dim, long long n, const double *xyz, const double *vol, const double *delta,
                  double nlf, double lambda, double mu, pdx_fbatch **out) {
    if (!out || !xyz || !vol || !delta || n < 1 || (dim != 2 && dim != 3))
        return set_error(PDX_ERR_ARG, "pdx_fb_create: bad arguments");
    std::string err;
    pdx_fbatch *b = new pdx_fbatch();
    auto run = [&]() -> int {
        b->dim = dim;
        b->device = device;
        b->n = n;
        b->nlf = nlf;
        b->lam = lambda;
        b->mu = mu;
        if (int rc = fb_members(device, dim, n, xyz, delta, b->off, &b->dmem, b->mf, err)) return rc;
        b->nf = n;
        b->nbond = b->off[n];
        if (b->nbond > (long long)INT32_MAX) { err = "pdx_fb_create: more than 2^31 bonds"; return PDX_ERR_ARG; }
        PDX_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        PDX_HIP(hipGetDeviceProperties(&prop, device));
        b->n_sm = prop.multiProcessorCount;
        b->lds_max = prop.sharedMemPerBlock;
        PDX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        const int P = dim == 3 ? 9 : 5, DD = dim * dim;
        PDX_HIP(hipMalloc((void **)&b->x, n * 3 * 8));
        PDX_HIP(hipMalloc((void **)&b->vol, n * 8));
        PDX_HIP(hipMalloc((void **)&b->delta, n * 8));
        PDX_HIP(hipMalloc((void **)&b->doff, (n + 1) * 8));
        PDX_HIP(hipMalloc((void **)&b->A, n * P * P * 8));
        PDX_HIP(hipMalloc((void **)&b->LU, n * P * P * 8));
        PDX_HIP(hipMalloc((void **)&b->H, std::max<long long>(b->nbond, 1) * DD * 4));
        PDX_HIP(hipMalloc((void **)&b->status, 4));
        PDX_HIP(hipMalloc((void **)&b->counter, FB_GROUPS * FB_CSTRIDE * sizeof(unsigned int)));
        PDX_HIP(hipMemcpy(b->x, xyz, n * 3 * 8, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(b->vol, vol, n * 8, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(b->delta, delta, n * 8, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(b->doff, b->off.data(), (n + 1) * 8, hipMemcpyHostToDevice));
        PDX_HIP(hipMemset(b->status, 0, 4));
        return PDX_OK;
    };
    int rc = run();
    if (rc) {
        pdx_fb_destroy(b);
        return set_error(rc, err);
    }
    *out = b;
    return PDX_OK;
}

int pdx_fb_sizes(const pdx_fbatch *b, long long *n_fam, long long *n_bond, int *max_family) {
    if (!b) return set_error(PDX_ERR_ARG, "null argument");
    if (n_fam) *n_fam = b->nf;
    if (n_bond) *n_bond = b->nbond;
    if (max_family) *max_family = b->mf;
    return PDX_OK;
}

int pdx_fb_families(const pdx_fbatch *b, long long *off, int *mem) {
    if (!b) return set_error(PDX_ERR_ARG, "null argument");
    if (off) std::copy(b->off.begin(), b->off.end(), off);
    if (mem && b->nbond) {
        std::string err;
        auto run = [&]() -> int {
            PDX_HIP(hipSetDevice(b->device));
            PDX_HIP(hipMemcpy(mem, b->dmem, (size_t)b->nbond * 4, hipMemcpyDeviceToHost));
            return PDX_OK;
        };
        if (int rc = run()) return set_error(rc, err);
    }
    return PDX_OK;
}

int pdx_fb_run(pdx_fbatch *b, int variant, float *ms) {
    if (!b) return set_error(PDX_ERR_ARG, "null argument");
    std::string err;
    auto run = [&]() -> int {
        PDX_HIP(hipSetDevice(b->device));
        int v = variant;
        if (v < 0) {
            if (b->tuned < 0) {
                // time every launch form on the batch (the results are identical: one family per
                // block and every persistent shape), best of two runs each, keep the fastest
                hipEvent_t e0, e1;
                PDX_HIP(hipEventCreate(&e0));
                PDX_HIP(hipEventCreate(&e1));
                float best = 1e30f;
                for (int k = 0; k < FB_VARIANTS; k++) {
                    b->tune_ms[k] = -1.f;  // form not available (LDS)
                    if (fb_run(*b, k) != hipSuccess) {  // warm
                        (void)hipGetLastError();
                        continue;
                    }
                    for (int rep = 0; rep < 2; rep++) {
                        float t = 0;
                        PDX_HIP(hipEventRecord(e0, b->stream));
                        PDX_HIP(fb_run(*b, k));
                        PDX_HIP(hipEventRecord(e1, b->stream));
                        PDX_HIP(hipEventSynchronize(e1));
                        PDX_HIP(hipEventElapsedTime(&t, e0, e1));
                        if (rep == 0 || t < b->tune_ms[k]) b->tune_ms[k] = t;
                    }
                    if (b->tune_ms[k] < best) {
                        best = b->tune_ms[k];
                        b->tuned = k;
                    }
                }
                hipEventDestroy(e0);
                hipEventDestroy(e1);
                if (b->tuned < 0) { err = "pdx_fb_run: no launch form fits the families"; return PDX_ERR_ARG; }
            }
            v = b->tuned;
        }
        if (v >= FB_VARIANTS) { err = "pdx_fb_run: unknown variant"; return PDX_ERR_ARG; }
        hipEvent_t e0, e1;
        PDX_HIP(hipEventCreate(&e0));
        PDX_HIP(hipEventCreate(&e1));
        PDX_HIP(hipEventRecord(e0, b->stream));
        PDX_HIP(fb_run(*b, v));
        PDX_HIP(hipEventRecord(e1, b->stream));
        PDX_HIP(hipEventSynchronize(e1));
        float t = 0;
        PDX_HIP(hipEventElapsedTime(&t, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        if (ms) *ms = t;
        int st = 0;
        PDX_HIP(hipMemcpy(&st, b->status, 4, hipMemcpyDeviceToHost));
        if (st & 8) { err = "pdx_fb_run: family larger than the LDS slot"; return PDX_ERR_ARG; }
        if (st & 1) { err = "pdx_fb_run: shape tensor not positive definite (degenerate family)"; return PDX_ERR_SINGULAR; }
        return PDX_OK;
    };
    int rc = run();
    return rc ? set_error(rc, err) : PDX_OK;
}

int pdx_fb_tuning(const pdx_fbatch *b, int *best, float *ms) {
    if (!b) return set_error(PDX_ERR_ARG, "null argument");
    if (best) *best = b->tuned;
    if (ms) std::copy(b->tune_ms, b->tune_ms + FB_VARIANTS, ms);
    return PDX_OK;
}

int pdx_fb_results(const pdx_fbatch *b, long long first, long long count, double *A, double *L, float *H) {
    if (!b || first < 0 || count < 0 || first + count > b->nf) return set_error(PDX_ERR_ARG, "pdx_fb_results: range");
    std::string err;
    auto run = [&]() -> int {
        PDX_HIP(hipSetDevice(b->device));
        PDX_HIP(hipStreamSynchronize(b->stream));
        const int P = b->dim == 3 ? 9 : 5, DD = b->dim * b->dim;
        if (A) PDX_HIP(hipMemcpy(A, b->A + first * P * P, count * P * P * 8, hipMemcpyDeviceToHost));
        if (L) PDX_HIP(hipMemcpy(L, b->LU + first * P * P, count * P * P * 8, hipMemcpyDeviceToHost));
        const long long o0 = b->off[first], o1 = b->off[first + count];
        if (H && o1 > o0) PDX_HIP(hipMemcpy(H, b->H + o0 * DD, (o1 - o0) * DD * 4, hipMemcpyDeviceToHost));
        return PDX_OK;
    };
    int rc = run();
    return rc ? set_error(rc, err) : PDX_OK;
}

}  // extern "C"
