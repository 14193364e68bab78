// Batched PDLSM family kernel: every node of a PD node cloud is a family with
// its own horizon; per family the shape tensor A (FP64, p x p), a Cholesky
// factor of its symmetric part, and the bond blocks H_m = G_m V_m (FP32, d x d
// per member, self block H_0 = -sum_m H_m) -- shapTens3D, vec_gd3D, matG3D and
// matH3D of pdsolve.cpp:1938-2416 for a whole batch.
//
// A = S diag(s) with S = sum_m w_m p_m p_m^T symmetric positive definite
// (w_m = omega_m V_m > 0, p = [x,y,z,x^2,y^2,z^2,xy,yz,zx], 2D [x,y,x^2,y^2,xy])
// and s = 1/2 on the squared terms.  So per family:
//   * S on the matrix cores: one wave, v_mfma_f64_16x16x4f64 over the members
//     four at a time (a lane holds w p_i and p_i of one member for the A and B
//     operands; the 9x9 / 5x5 result sits in the 16x16 tile);
//   * S = L L^T by one lane in registers (no pivoting is needed for an SPD
//     matrix; the reference's per-bond dgesv factors A with partial pivoting,
//     which gives the same solution to rounding);
//   * every bond on its own lane: two triangular solves with the shared L
//     (reciprocal diagonal), x = y / s, then G and H.
// Launch forms of the same per-family code: one family per 64-thread block, or
// persistent waves pulling families from an atomic counter with 1, 2 or 4
// waves per block; the fastest is timed once per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pdx.h"
#include "device.hpp"

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
    unsigned long long *counter;    // persistent work counter
};

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int D>
__device__ inline void basis(double x, double y, double z, double p[FbDim<D>::P]) {
    if constexpr (D == 3) {
        p[0] = x; p[1] = y; p[2] = z;
        p[3] = x * x; p[4] = y * y; p[5] = z * z;
        p[6] = x * y; p[7] = y * z; p[8] = z * x;
    } else {
        p[0] = x; p[1] = y; p[2] = x * x; p[3] = y * y; p[4] = x * y;
    }
}

template <int D>
__device__ inline bool is_square(int c) {
    return D == 3 ? (c >= 3 && c < 6) : (c == 2 || c == 3);
}

// slot layout (doubles): member basis values [mf*P], weights w = omega V [mf],
// S / L [16*16], reciprocal diagonal of L [P] (+ singular flag)
template <int D>
__host__ __device__ constexpr int fb_slot_doubles(int mf) {
    return mf * (FbDim<D>::P + 1) + 256 + FbDim<D>::P + 1;
}

typedef double fb_d4 __attribute__((ext_vector_type(4)));

// one family on one wave; slot = its LDS region
template <int D>
__device__ void fb_family(const FbArgs &a, long long f, int lane, double *slot) {
    constexpr int P = FbDim<D>::P, PP = P * P, DD = FbDim<D>::DD;
    const long long o0 = a.off[f];
    const int n = int(a.off[f + 1] - o0);
    if (n > a.mf) {
        if (lane == 0) atomicOr(a.status, 8);
        return;
    }
    double *sb = slot;             // [mf*P] basis values p(xi_m)
    double *sw = slot + a.mf * P;  // [mf] omega_m V_m (0 for the self member)
    double *sS = sw + a.mf;        // [16*16] S, then L (row-major, ld 16)
    double *srd = sS + 256;        // [P] 1 / L_ii, [P]: singular flag
    const int k = a.mem[o0];
    const double xk = a.x[(size_t)k * 3], yk = a.x[(size_t)k * 3 + 1], zk = a.x[(size_t)k * 3 + 2];
    const double delt = a.delta[k], Aa = a.nlf * a.nlf;
    for (int m = lane; m < n; m += 64) {
        const int j = a.mem[o0 + m];
        const double x0 = a.x[(size_t)j * 3] - xk, x1 = a.x[(size_t)j * 3 + 1] - yk, x2 = a.x[(size_t)j * 3 + 2] - zk;
        double p[P];
        basis<D>(x0, x1, x2, p);
#pragma unroll
        for (int q = 0; q < P; q++) sb[m * P + q] = p[q];
        // inflFunc (pdsolve.cpp:1929-1935) times the member volume
        sw[m] = m == 0 ? 0.0 : exp(-(x0 * x0 + x1 * x1 + x2 * x2) / Aa / delt / delt) * a.vol[j];
    }
    wave_sync();
    // S = sum_m w_m p_m p_m^T on the matrix cores: A-operand lane (i = lane%16, member k0 + lane/16)
    // holds w p_i, the B-operand lane (member k0 + lane/16, column j = lane%16 = i) holds p_i
    const int i16 = lane & 15, kq = lane >> 4;
    fb_d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < n; k0 += 4) {
        const int m = k0 + kq;
        const bool ok = m < n && i16 < P;
        const double bi = ok ? sb[m * P + i16] : 0.0;
        const double wm = m < n ? sw[m] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(wm * bi, bi, acc, 0, 0, 0);
    }
    // result layout (checked on gfx950): lane l item r = S[l/16 + 4 r][l%16]
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int row = kq + 4 * r, col = i16;
        sS[row * 16 + col] = acc[r];
        if (row < P && col < P) a.A[f * PP + row * P + col] = is_square<D>(col) ? 0.5 * acc[r] : acc[r];
    }
    wave_sync();
    // Cholesky S = L L^T in registers on one lane (fixed indices, fully unrolled)
    if (lane == 0) {
        double L[PP];
#pragma unroll
        for (int r = 0; r < P; r++)
#pragma unroll
            for (int c = 0; c < P; c++) L[r * P + c] = c <= r ? sS[r * 16 + c] : 0.0;
        int sing = 0;
#pragma unroll
        for (int c = 0; c < P; c++) {
            double d = L[c * P + c];
#pragma unroll
            for (int q = 0; q < c; q++) d -= L[c * P + q] * L[c * P + q];
            sing |= !(d > 0.0);
            const double lc = sqrt(d > 0.0 ? d : 1.0);
            const double il = 1.0 / lc;
            L[c * P + c] = lc;
            srd[c] = il;
#pragma unroll
            for (int r = c + 1; r < P; r++) {
                double t = L[r * P + c];
#pragma unroll
                for (int q = 0; q < c; q++) t -= L[r * P + q] * L[c * P + q];
                L[r * P + c] = t * il;
            }
        }
#pragma unroll
        for (int e = 0; e < PP; e++) sS[(e / P) * 16 + (e % P)] = L[e];
        srd[P] = sing;
    }
    wave_sync();
    for (int e = lane; e < PP; e += 64) a.LU[f * PP + e] = sS[(e / P) * 16 + (e % P)];
    if (srd[P] != 0.0) {
        if (lane == 0) atomicOr(a.status, 1);
        return;
    }
    const double lam = a.lam, mu = a.mu;
    double hs[DD];
#pragma unroll
    for (int e = 0; e < DD; e++) hs[e] = 0.0;
    for (int m = 1 + lane; m < n; m += 64) {
        // S y = p_m: L z = p_m, L^T y = z; then x = y / s (A = S diag(s))
        double z[P];
#pragma unroll
        for (int r = 0; r < P; r++) {
            double t = sb[m * P + r];
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
        double sx[P];
#pragma unroll
        for (int r = 0; r < P; r++) sx[r] = is_square<D>(r) ? 2.0 * y[r] : y[r];
        // matG (pdsolve.cpp:2277-2347) times omega V: H_m = G_m V_m (all bonds intact)
        const double w = sw[m];
        double G[DD];
        if constexpr (D == 2) {
            const double *dd = sx + 2;
            G[0] = w * ((2 * mu + lam) * dd[0] + mu * dd[1]);
            G[1] = G[2] = w * (lam + mu) * dd[2];
            G[3] = w * ((2 * mu + lam) * dd[1] + mu * dd[0]);
        } else {
            const double *dd = sx + 3;
            const double tr = dd[0] + dd[1] + dd[2];
            G[0] = w * ((mu + lam) * dd[0] + mu * tr);
            G[1] = G[3] = w * (lam + mu) * dd[3];
            G[2] = G[6] = w * (lam + mu) * dd[5];
            G[4] = w * ((mu + lam) * dd[1] + mu * tr);
            G[5] = G[7] = w * (lam + mu) * dd[4];
            G[8] = w * ((mu + lam) * dd[2] + mu * tr);
        }
        float *h = a.H + (size_t)(o0 + m) * DD;
#pragma unroll
        for (int e = 0; e < DD; e++) {
            h[e] = float(G[e]);
            hs[e] += G[e];
        }
    }
    // self block: reduction over the wave in a fixed tree
#pragma unroll
    for (int e = 0; e < DD; e++) {
        double v = hs[e];
        for (int sft = 32; sft > 0; sft >>= 1) v += __shfl_xor(v, sft, 64);
        hs[e] = v;
    }
    if (lane == 0) {
        float *h = a.H + (size_t)o0 * DD;
#pragma unroll
        for (int e = 0; e < DD; e++) h[e] = float(-hs[e]);
    }
}

template <int D>
__global__ void __launch_bounds__(64) k_fb_block(FbArgs a) {
    extern __shared__ double lds[];
    const long long f = blockIdx.x;
    if (f < a.nf) fb_family<D>(a, f, threadIdx.x, lds);
}

// persistent: every wave pulls the next family per atomic fetch until the batch is drained
template <int D>
__global__ void __launch_bounds__(256) k_fb_persistent(FbArgs a) {
    extern __shared__ double lds[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *slot = lds + (size_t)wave * fb_slot_doubles<D>(a.mf);
    for (;;) {
        unsigned long long f = 0;
        if (lane == 0) f = atomicAdd(a.counter, 1ull);
        f = __shfl(f, 0, 64);
        if (f >= (unsigned long long)a.nf) break;  // uniform across the wave: every wave exits
        fb_family<D>(a, (long long)f, lane, slot);
        wave_sync();
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
    std::vector<int> mem;
    double nlf = 0.33, lam = 0, mu = 0;
    double *x = nullptr, *vol = nullptr, *delta = nullptr, *A = nullptr, *LU = nullptr;
    long long *doff = nullptr;
    int *dmem = nullptr, *status = nullptr;
    float *H = nullptr;
    unsigned long long *counter = nullptr;
    int tuned = -1;
    float tune_ms[8] = {0};
    int n_sm = 256;
    size_t lds_max = 65536;  // LDS bytes a block may use on this device
};

namespace pdx {

// launch forms: 0 one family per block; 1..3 persistent with 4, 2, 1 waves per block
constexpr int FB_VARIANTS = 4;

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

template <int D, int WPB>
static hipError_t launch_persistent(pdx_fbatch &b, FbArgs a) {
    const size_t lds = (size_t)WPB * fb_slot_doubles<D>(b.mf) * sizeof(double);
    if (lds > b.lds_max) return hipErrorInvalidValue;  // this form does not fit the families
    hipError_t e = hipFuncSetAttribute((const void *)k_fb_persistent<D>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    int per_sm = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_sm, (const void *)k_fb_persistent<D>, WPB * 64, lds);
    if (e != hipSuccess) return e;
    const int grid = std::max(1, per_sm) * b.n_sm;
    e = hipMemsetAsync(b.counter, 0, sizeof(unsigned long long), b.stream);
    if (e != hipSuccess) return e;
    k_fb_persistent<D><<<grid, WPB * 64, lds, b.stream>>>(a);
    return hipGetLastError();
}

template <int D>
static hipError_t fb_launch(pdx_fbatch &b, int variant) {
    FbArgs a = fb_args(b);
    switch (variant) {
    case 1: return launch_persistent<D, 4>(b, a);
    case 2: return launch_persistent<D, 2>(b, a);
    case 3: return launch_persistent<D, 1>(b, a);
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
static void fb_members(int dim, long long n, const double *x, const double *delta, std::vector<long long> &off,
                       std::vector<int> &mem, int &mf) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, dmax = 0, dmin = 1e300;
    for (long long k = 0; k < n; k++) {
        for (int i = 0; i < 3; i++) {
            lo[i] = std::min(lo[i], x[k * 3 + i]);
            hi[i] = std::max(hi[i], x[k * 3 + i]);
        }
        dmax = std::max(dmax, delta[k]);
        dmin = std::min(dmin, delta[k]);
    }
    const double md2 = dmin * dmin * 1.0e-16;
    const double cs = dmax > 0 ? dmax : 1.0;
    int nb[3] = {1, 1, 1};
    for (int i = 0; i < dim; i++) nb[i] = std::max(1, (int)std::floor((hi[i] - lo[i]) / cs) + 1);
    auto cell_of = [&](const double *p, int ci[3]) {
        for (int i = 0; i < 3; i++) {
            ci[i] = i < dim ? std::min(nb[i] - 1, std::max(0, (int)std::floor((p[i] - lo[i]) / cs))) : 0;
        }
    };
    const long long ncell = (long long)nb[0] * nb[1] * nb[2];
    std::vector<long long> coff(ncell + 1, 0);
    std::vector<long long> cnode(n);
    for (long long k = 0; k < n; k++) {
        int ci[3];
        cell_of(&x[k * 3], ci);
        coff[ci[0] + (long long)nb[0] * (ci[1] + (long long)nb[1] * ci[2]) + 1]++;
    }
    for (long long c = 0; c < ncell; c++) coff[c + 1] += coff[c];
    {
        std::vector<long long> fill(coff.begin(), coff.end() - 1);
        for (long long k = 0; k < n; k++) {
            int ci[3];
            cell_of(&x[k * 3], ci);
            cnode[fill[ci[0] + (long long)nb[0] * (ci[1] + (long long)nb[1] * ci[2])]++] = k;
        }
    }
    std::vector<int> cnt(n);
    auto scan = [&](long long k, int *out) {
        int c = 0;
        if (out) out[c] = (int)k;
        c++;
        int ci[3];
        cell_of(&x[k * 3], ci);
        const double *p = &x[k * 3];
        const int r2 = dim == 3 ? 1 : 0;
        for (int xi = ci[0] - 1; xi <= ci[0] + 1; xi++)
            for (int yi = ci[1] - 1; yi <= ci[1] + 1; yi++)
                for (int zi = ci[2] - r2; zi <= ci[2] + r2; zi++) {
                    if (xi < 0 || yi < 0 || zi < 0 || xi >= nb[0] || yi >= nb[1] || zi >= nb[2]) continue;
                    const long long cc = xi + (long long)nb[0] * (yi + (long long)nb[1] * zi);
                    for (long long t = coff[cc]; t < coff[cc + 1]; t++) {
                        const long long j = cnode[t];
                        const double *q = &x[j * 3];
                        const double dist = (q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) +
                                            (q[2] - p[2]) * (q[2] - p[2]);
                        if (dist > md2 && (dist < delta[k] * delta[k] + md2 || dist < delta[j] * delta[j] + md2)) {
                            if (out) out[c] = (int)j;
                            c++;
                        }
                    }
                }
        return c;
    };
#pragma omp parallel for schedule(dynamic, 4096)
    for (long long k = 0; k < n; k++) cnt[k] = scan(k, nullptr);
    off.assign(n + 1, 0);
    mf = 0;
    for (long long k = 0; k < n; k++) {
        off[k + 1] = off[k] + cnt[k];
        mf = std::max(mf, cnt[k]);
    }
    mem.resize(off[n]);
#pragma omp parallel for schedule(dynamic, 4096)
    for (long long k = 0; k < n; k++) scan(k, mem.data() + off[k]);
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
        fb_members(dim, n, xyz, delta, b->off, b->mem, b->mf);
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
        PDX_HIP(hipMalloc((void **)&b->dmem, std::max<long long>(b->nbond, 1) * 4));
        PDX_HIP(hipMalloc((void **)&b->A, n * P * P * 8));
        PDX_HIP(hipMalloc((void **)&b->LU, n * P * P * 8));
        PDX_HIP(hipMalloc((void **)&b->H, std::max<long long>(b->nbond, 1) * DD * 4));
        PDX_HIP(hipMalloc((void **)&b->status, 4));
        PDX_HIP(hipMalloc((void **)&b->counter, 8));
        PDX_HIP(hipMemcpy(b->x, xyz, n * 3 * 8, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(b->vol, vol, n * 8, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(b->delta, delta, n * 8, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(b->doff, b->off.data(), (n + 1) * 8, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(b->dmem, b->mem.data(), b->nbond * 4, hipMemcpyHostToDevice));
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
    if (mem) std::copy(b->mem.begin(), b->mem.end(), mem);
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
                // time every launch form on the batch (the results are identical), keep the fastest
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
                    PDX_HIP(hipEventRecord(e0, b->stream));
                    PDX_HIP(fb_run(*b, k));
                    PDX_HIP(hipEventRecord(e1, b->stream));
                    PDX_HIP(hipEventSynchronize(e1));
                    PDX_HIP(hipEventElapsedTime(&b->tune_ms[k], e0, e1));
                    if (k > 0 && b->tune_ms[k] < best) {
                        best = b->tune_ms[k];
                        b->tuned = k;
                    }
                }
                hipEventDestroy(e0);
                hipEventDestroy(e1);
                if (b->tuned < 0) { err = "pdx_fb_run: no persistent form fits the families"; return PDX_ERR_ARG; }
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
