// Explicit central-difference steps, power iteration and Krylov solvers on
// the node-block matrix, for CDNA4 (gfx950).
//
// Explicit step (pdsolve.cpp:5842-6090, timeIntegration_LM :5113-5132): one
// fused kernel per row class computes y = K u_n for a node's rows and updates
//   u_{n+1} = 2 u_n + dt^2 (F_ext - K u)/m - u_{n-1}   (free DOFs)
//   v_{n+2} = v_{n+1} + dt * velocity                  (prescribed DOFs)
// with three rotating state vectors, so a step is a pure function of device
// state and a run of steps can be replayed from a hipGraph.
//
// BiCGStab keeps every scalar on the device: reducing kernels write per-block
// partial sums, consumer kernels re-reduce them in a fixed order (bitwise
// deterministic), so no host synchronisation is needed inside the iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "device.hpp"
#include "rowblock.hpp"

namespace pdx {

// ------------------------------------------------------------------------
// explicit step
// ------------------------------------------------------------------------
// FP32 values use the difference form (see rowblock.hpp), FP64 the plain form.
// GS = lanes per short row, NT = non-temporal loads of K (chosen by the tuner).
template <int D, typename T, int GS, bool NT>
__global__ void __launch_bounds__(256) k_explicit(const int *rows_s, int ns, int nbs, const int *rows_l, int nl,
                                                  const int *adj_off, const int *adj, const T *K, const double *Un,
                                                  const double *Unm1, double *Unp1, const double *m,
                                                  const double *f_unit, int n_nbc, int nn, const double *nbc_cur,
                                                  double *nbc_next, const double *nbc_inc, const int *eq,
                                                  const int *presc_set, const double *ebc_inc, double dt,
                                                  int startup) {
    constexpr bool DIFF = sizeof(T) == 4;
    if (blockIdx.x == 0 && threadIdx.x < n_nbc) nbc_next[threadIdx.x] = nbc_cur[threadIdx.x] + nbc_inc[threadIdx.x];
    T acc[D];
    int r, lane;
    if ((int)blockIdx.x < nbs) {
        const int gid = (blockIdx.x * blockDim.x + threadIdx.x) / GS;
        lane = threadIdx.x % GS;
        if (gid >= ns) return;
        r = rows_s[gid];
        row_matvec<D, T, double, GS, DIFF, NT>(r, lane, adj_off, adj, K, Un, acc);
    } else {
        const int gid = ((blockIdx.x - nbs) * blockDim.x + threadIdx.x) / 64;
        lane = threadIdx.x % 64;
        if (gid >= nl) return;
        r = rows_l[gid];
        row_matvec<D, T, double, 64, DIFF, NT>(r, lane, adj_off, adj, K, Un, acc);
    }
    if (lane < D) {
        const T v = pick<D, T>(acc, lane);
        const size_t o = (size_t)r * D + lane;
        const int code = r * 3 + lane;
        if (eq[code] >= 0) {
            double fext = 0.0;
            for (int s = 0; s < n_nbc; s++) fext = fext + nbc_cur[s] * f_unit[(size_t)s * nn * D + o];
            const double F = fext - double(v);
            if (startup) {
                const double a = F / m[o];
                Unp1[o] = a * (0.5 * dt * dt) + Un[o];
            } else {
                const double tA = dt * dt * F / m[o];
                Unp1[o] = Un[o] * 2.0 + tA - Unm1[o];
            }
        } else {
            Unp1[o] = Un[o] + ebc_inc[presc_set[code]];
        }
    }
}

// ------------------------------------------------------------------------
// reductions and vector kernels (V = double or float)
// ------------------------------------------------------------------------

// deterministic block reduction of one value per thread (256 threads)
__device__ inline double block_sum(double v, double *sm) {
    const int t = threadIdx.x;
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    if ((t & 63) == 0) sm[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sm[w];
        sm[8] = r;
    }
    __syncthreads();
    r = sm[8];
    __syncthreads();
    return r;
}

// every block re-reduces n partial sums in the same fixed order
__device__ inline double sum_partials(const double *p, int n, double *sm) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += p[i];
    return block_sum(v, sm);
}

constexpr int NPART = 2 * RED_PARTS;  // short-row + long-row SpMV partials

template <typename V>
__global__ void __launch_bounds__(256) k_dot(const V *a, const V *b, size_t n, double *part) {
    __shared__ double sm[16];
    double v = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v += double(a[i]) * double(b[i]);
    v = block_sum(v, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
    if (threadIdx.x == 0 && blockIdx.x == 0)
        for (int i = gridDim.x; i < NPART; i++) part[i] = 0.0;
}

// SpMV with fused dot partials: y = mask(K x); part0 += (w0, y); part1 += (y, y) when w1 set
template <int D, typename T, typename V, int G>
__device__ inline void spmv_dot_rows(const int *rows, int nrows, int first_group, int groups, const int *adj_off,
                                     const int *adj, const T *K, const V *x, V *y, const int *eq, const V *w0,
                                     const V *w1, double &d0, double &d1) {
    const int lane = threadIdx.x % G;
    for (int gid = first_group; gid < nrows; gid += groups) {
        const int r = rows[gid];
        T acc[D];
        row_matvec<D, T, V, G, false>(r, lane, adj_off, adj, K, x, acc);
        if (lane < D) {
            const T v = pick<D, T>(acc, lane);
            const size_t o = (size_t)r * D + lane;
            V yv = eq[r * 3 + lane] >= 0 ? V(v) : V(0);
            y[o] = yv;
            if (w0) d0 += double(w0[o]) * double(yv);
            if (w1) d1 += double(w1[o]) * double(yv);
        }
    }
}

// SpMV with fused dot partials over both row classes (grid-stride, RED_PARTS
// blocks): y = mask(K x); part0 = (w0, y); part1 = (w1, y)
template <int D, typename T, typename V>
__global__ void __launch_bounds__(256) k_spmv_dot(const int *rows_s, int ns, const int *rows_l, int nl,
                                                  const int *adj_off, const int *adj, const T *K, const V *x, V *y,
                                                  const int *eq, const V *w0, double *part0, const V *w1,
                                                  double *part1) {
    __shared__ double sm[16];
    double d0 = 0.0, d1 = 0.0;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    spmv_dot_rows<D, T, V, 8>(rows_s, ns, t / 8, nt / 8, adj_off, adj, K, x, y, eq, w0, w1, d0, d1);
    spmv_dot_rows<D, T, V, 64>(rows_l, nl, t / 64, nt / 64, adj_off, adj, K, x, y, eq, w0, w1, d0, d1);
    if (part0) {
        d0 = block_sum(d0, sm);
        if (threadIdx.x == 0) part0[blockIdx.x] = d0;
    }
    if (part1) {
        d1 = block_sum(d1, sm);
        if (threadIdx.x == 0) part1[blockIdx.x] = d1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int i = gridDim.x; i < NPART; i++) {
            if (part0) part0[i] = 0.0;
            if (part1) part1[i] = 0.0;
        }
}

// apply the node-diagonal block preconditioner z = Minv r (Minv: [nn][D*D])
template <int D, typename V>
__device__ inline void apply_blk(const V *Minv, const V *r, V *z, int a) {
    V rv[D];
#pragma unroll
    for (int j = 0; j < D; j++) rv[j] = r[(size_t)a * D + j];
#pragma unroll
    for (int i = 0; i < D; i++) {
        V s = V(0);
#pragma unroll
        for (int j = 0; j < D; j++) s += Minv[(size_t)a * D * D + i * D + j] * rv[j];
        z[(size_t)a * D + i] = s;
    }
}

// scalar slots in the device array sc[]
enum { SC_RHO_OLD = 0, SC_ALPHA = 1, SC_OMEGA = 2 };
// partial-sum slots (each NPART doubles)
enum { P_RHO0 = 0, P_RHO1 = 1, P_RV = 2, P_TS = 3, P_TT = 4, P_RR = 5, P_BB = 6, P_TMP = 7 };

template <int D, typename V>
__global__ void __launch_bounds__(256) k_bicg_p(int nn, const V *r, V *p, const V *v, V *ph, const V *Minv,
                                                const double *rho_part, double *sc, int first) {
    __shared__ double sm[16];
    const double rho = sum_partials(rho_part, NPART, sm);
    // a vanishing rho_old or omega (breakdown, or exact convergence in low
    // precision) restarts the search direction with p = r
    double beta = 0.0, omg = 0.0;
    if (!first && sc[SC_RHO_OLD] != 0.0 && sc[SC_OMEGA] != 0.0) {
        beta = (rho / sc[SC_RHO_OLD]) * (sc[SC_ALPHA] / sc[SC_OMEGA]);
        omg = sc[SC_OMEGA];
    }
    if (!isfinite(beta)) beta = 0.0;
    const V bv = V(beta), om = V(omg);
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < nn; a += gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            const size_t o = (size_t)a * D + i;
            p[o] = first ? r[o] : V(r[o] + bv * (p[o] - om * v[o]));
        }
        apply_blk<D, V>(Minv, p, ph, a);
    }
}

template <int D, typename V>
__global__ void __launch_bounds__(256) k_bicg_s(int nn, const V *r, const V *v, V *s, V *sh, const V *Minv,
                                                const double *rho_part, const double *rv_part, double *sc) {
    __shared__ double sm[16];
    const double rho = sum_partials(rho_part, NPART, sm);
    const double rv = sum_partials(rv_part, NPART, sm);
    double alpha = rv != 0.0 ? rho / rv : 0.0;
    if (!isfinite(alpha)) alpha = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_ALPHA] = alpha;
    const V al = V(alpha);
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < nn; a += gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            const size_t o = (size_t)a * D + i;
            s[o] = r[o] - al * v[o];
        }
        apply_blk<D, V>(Minv, s, sh, a);
    }
}

template <int D, typename V>
__global__ void __launch_bounds__(256) k_bicg_x(int nn, V *x, V *r, const V *rh, const V *ph, const V *sh,
                                                const V *s, const V *t, const double *rho_part,
                                                const double *ts_part, const double *tt_part, double *sc,
                                                double *rho_next, double *rr_part) {
    __shared__ double sm[16];
    const double rho = sum_partials(rho_part, NPART, sm);
    const double ts = sum_partials(ts_part, NPART, sm);
    const double tt = sum_partials(tt_part, NPART, sm);
    double omega = tt > 0 ? ts / tt : 0.0;
    if (!isfinite(omega)) omega = 0.0;
    const double alpha = sc[SC_ALPHA];
    const V al = V(alpha), om = V(omega);
    double d0 = 0.0, d1 = 0.0;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < nn; a += gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            const size_t o = (size_t)a * D + i;
            x[o] = x[o] + al * ph[o] + om * sh[o];
            const V rn = s[o] - om * t[o];
            r[o] = rn;
            d0 += double(rh[o]) * double(rn);
            d1 += double(rn) * double(rn);
        }
    }
    d0 = block_sum(d0, sm);
    d1 = block_sum(d1, sm);
    if (threadIdx.x == 0) {
        rho_next[blockIdx.x] = d0;
        rr_part[blockIdx.x] = d1;
        if (blockIdx.x == 0) {
            for (int i = gridDim.x; i < NPART; i++) { rho_next[i] = 0.0; rr_part[i] = 0.0; }
            sc[SC_OMEGA] = omega;
            sc[SC_RHO_OLD] = rho;
        }
    }
}

// r = b - y (masked by y already), plus partial (r, r); optional copy to rh
template <typename V, typename B>
__global__ void __launch_bounds__(256) k_resid(size_t n, const B *b, const double *y, V *r, V *rh, double *part) {
    __shared__ double sm[16];
    double v = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double ri = double(b[i]) - y[i];
        r[i] = V(ri);
        if (rh) rh[i] = V(ri);
        v += ri * ri;
    }
    v = block_sum(v, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
    if (threadIdx.x == 0 && blockIdx.x == 0)
        for (int i = gridDim.x; i < NPART; i++) part[i] = 0.0;
}

template <typename A, typename B>
__global__ void k_copy(size_t n, const A *a, B *b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = B(a[i]);
}

__global__ void k_axpy_d(size_t n, double a, const double *x, double *y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = y[i] + a * x[i];
}

__global__ void k_add_f2d(size_t n, const float *d, double *x) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = x[i] + double(d[i]);
}

// block-Jacobi / Jacobi inverse from the diagonal blocks; prescribed rows/cols removed
template <int D, typename V>
__global__ void k_blockjacobi(int nn, const int *adj_off, const int *adj, const double *K, long long nblk,
                              const int *eq, int diag_only, V *Minv, int *status) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    int lo = adj_off[a], hi = adj_off[a + 1] - 1, b = -1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (adj[mid] == a) { b = mid; break; }
        if (adj[mid] < a) lo = mid + 1; else hi = mid - 1;
    }
    double M[D * D], I[D * D];
    int fr[D];
    for (int i = 0; i < D; i++) fr[i] = eq[a * 3 + i] >= 0;
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) {
            double v = (b >= 0) ? K[(size_t)b * D * D + i * D + j] : 0.0;
            if (diag_only && i != j) v = 0.0;
            if (!fr[i] || !fr[j]) v = (i == j) ? 1.0 : 0.0;
            M[i * D + j] = v;
            I[i * D + j] = i == j ? 1.0 : 0.0;
        }
    // Gauss-Jordan with partial pivoting on a D x D block
    for (int k = 0; k < D; k++) {
        int p = k;
        for (int i = k + 1; i < D; i++)
            if (fabs(M[i * D + k]) > fabs(M[p * D + k])) p = i;
        if (M[p * D + k] == 0.0) { atomicOr(status, 1); return; }
        for (int j = 0; j < D; j++) {
            double t = M[k * D + j]; M[k * D + j] = M[p * D + j]; M[p * D + j] = t;
            t = I[k * D + j]; I[k * D + j] = I[p * D + j]; I[p * D + j] = t;
        }
        const double piv = M[k * D + k];
        for (int j = 0; j < D; j++) { M[k * D + j] /= piv; I[k * D + j] /= piv; }
        for (int i = 0; i < D; i++) {
            if (i == k) continue;
            const double f = M[i * D + k];
            for (int j = 0; j < D; j++) { M[i * D + j] -= f * M[k * D + j]; I[i * D + j] -= f * I[k * D + j]; }
        }
    }
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) Minv[(size_t)a * D * D + i * D + j] = (fr[i] && fr[j]) ? V(I[i * D + j]) : V(0);
}

__global__ void k_fill_s(double *p, double v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_to_f32(size_t n, const double *a, float *b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = float(a[i]);
}

// z = M^-1 y (lumped), masked
__global__ void k_scale_mass(size_t n, const double *y, const double *m, const int *free_mask_eq, int D,
                             double *z) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t a = i / D, k = i % D;
        z[i] = free_mask_eq[a * 3 + k] >= 0 ? y[i] / m[i] : 0.0;
    }
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
static double host_sum(Ctx &c, const double *part, int n, std::string &err) {
    std::vector<double> h(n);
    if (hipMemcpyAsync(h.data(), part, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        err = "reduction copy failed";
        return NAN;
    }
    double s = 0.0;
    for (int i = 0; i < n; i++) s += h[i];
    return s;
}

static int ensure_K32(Ctx &c, std::string &err) {
    if (!c.K_valid) { err = "K not assembled"; return PDX_ERR_ARG; }
    if (c.K32_valid) return PDX_OK;
    if (!c.K32) PDX_HIP(hipMalloc((void **)&c.K32, sizeof(float) * c.dd * c.nblk));
    size_t n = (size_t)c.dd * c.nblk;
    k_to_f32<<<4096, 256, 0, c.stream>>>(n, c.K, c.K32);
    PDX_HIP(hipGetLastError());
    c.K32_valid = true;
    return PDX_OK;
}

// explicit-step variants searched by the tuner: (lanes per short row, non-temporal K)
constexpr int N_EXPLICIT_VARIANTS = 6;
static const int kVarGS[N_EXPLICIT_VARIANTS] = {8, 4, 16, 8, 4, 16};
static const int kVarNT[N_EXPLICIT_VARIANTS] = {0, 0, 0, 1, 1, 1};

template <int D, typename T, int GS, bool NT>
static void launch_explicit_v(Ctx &c, const T *K, const double *Un, const double *Unm1, double *Unp1,
                              const double *nbc_cur, double *nbc_next, int startup) {
    const int nbs = (c.n_short * GS + 255) / 256, nbl = (c.n_long * 64 + 255) / 256;
    k_explicit<D, T, GS, NT><<<std::max(nbs + nbl, 1), 256, 0, c.stream>>>(
        c.rows_short, c.n_short, nbs, c.rows_long, c.n_long, c.adj_off, c.adj, K, Un, Unm1, Unp1, c.mlump, c.f_unit,
        c.n_nbc, c.nn, nbc_cur, nbc_next, c.d_nbc_vel, c.eq, c.presc_set, c.d_ebc_vel, c.dt, startup);
}

template <int D, typename T>
static void launch_explicit(Ctx &c, int variant, const T *K, const double *Un, const double *Unm1, double *Unp1,
                            const double *nbc_cur, double *nbc_next, int startup) {
    switch (variant) {
    case 1: launch_explicit_v<D, T, 4, false>(c, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup); break;
    case 2: launch_explicit_v<D, T, 16, false>(c, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup); break;
    case 3: launch_explicit_v<D, T, 8, true>(c, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup); break;
    case 4: launch_explicit_v<D, T, 4, true>(c, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup); break;
    case 5: launch_explicit_v<D, T, 16, true>(c, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup); break;
    default: launch_explicit_v<D, T, 8, false>(c, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup); break;
    }
}

static void launch_explicit_any(Ctx &c, int prec, int variant, const double *Un, const double *Unm1, double *Unp1,
                                const double *nbc_cur, double *nbc_next, int startup) {
    if (prec == PDX_FP32) {
        if (c.dim == 3) launch_explicit<3, float>(c, variant, c.K32, Un, Unm1, Unp1, nbc_cur, nbc_next, startup);
        else launch_explicit<2, float>(c, variant, c.K32, Un, Unm1, Unp1, nbc_cur, nbc_next, startup);
    } else {
        if (c.dim == 3) launch_explicit<3, double>(c, variant, c.K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup);
        else launch_explicit<2, double>(c, variant, c.K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup);
    }
}

// Autotuner: time every variant on the current state with the result written to
// scratch buffers (the simulation state is not advanced), keep the fastest.
static int tune_explicit(Ctx &c, int prec, std::string &err) {
    const size_t n = (size_t)c.nn * c.dim;
    double *scratch, *nscratch;
    PDX_HIP(hipMalloc((void **)&scratch, n * 8));
    PDX_HIP(hipMalloc((void **)&nscratch, 2 * (size_t)std::max(c.n_nbc, 1) * 8));
    hipEvent_t e0, e1;
    PDX_HIP(hipEventCreate(&e0));
    PDX_HIP(hipEventCreate(&e1));
    const int p = c.step_phase % 3;
    float best = 1e30f;
    int bestv = 0;
    for (int v = 0; v < N_EXPLICIT_VARIANTS; v++) {
        launch_explicit_any(c, prec, v, c.U[p], c.U[(p + 2) % 3], scratch, c.nbcv, nscratch, 0);  // warm
        PDX_HIP(hipEventRecord(e0, c.stream));
        for (int r = 0; r < 5; r++)
            launch_explicit_any(c, prec, v, c.U[p], c.U[(p + 2) % 3], scratch, c.nbcv, nscratch, 0);
        PDX_HIP(hipEventRecord(e1, c.stream));
        PDX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        PDX_HIP(hipEventElapsedTime(&ms, e0, e1));
        c.tune_ms[prec][v] = ms / 5;
        if (ms < best) { best = ms; bestv = v; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(scratch);
    hipFree(nscratch);
    c.tune_explicit[prec] = bestv;
    return PDX_OK;
}

// one step on the device state: phase p = index of u_n buffer, s = step index
static void explicit_step(Ctx &c, int prec, int startup) {
    const int p = c.step_phase % 3;
    double *Un = c.U[p], *Unm1 = c.U[(p + 2) % 3], *Unp1 = c.U[(p + 1) % 3];
    const long long s = startup ? 0 : c.steps_done + 1;
    double *nbc_cur = c.nbcv + (s % 2) * std::max(c.n_nbc, 1);
    double *nbc_next = c.nbcv + ((s + 1) % 2) * std::max(c.n_nbc, 1);
    launch_explicit_any(c, prec, std::max(c.tune_explicit[prec == PDX_FP32 ? 1 : 0], 0), Un, Unm1, Unp1, nbc_cur,
                        nbc_next, startup);
    c.step_phase = (p + 1) % 3;
    c.steps_done = s;
}

// advance the host mirrors of the BC values exactly like setVaryEssentialBC /
// setVaryNaturalBC (pdsolve.cpp:5263-5329)
static void host_advance_bc(Ctx &c) {
    for (size_t s = 0; s < c.ebc_val.size(); s++)
        if (c.ebc_vary[s]) c.ebc_val[s] = c.ebc_val[s] + c.dt * c.ebc_vel[s];
    for (size_t s = 0; s < c.nbc_val.size(); s++)
        if (c.nbc_vary[s]) c.nbc_val[s] = c.nbc_val[s] + c.dt * c.nbc_vel[s];
}

int explicit_prepare(Ctx &c, int prec, std::string &err) {
    if (!c.K_valid) {
        int e = dev_assemble_K(c, err);
        if (e) return e;
    }
    if (prec == PDX_FP32) {
        int e = ensure_K32(c, err);
        if (e) return e;
    }
    if (!c.mlump_valid) {
        int e = dev_lumped_mass(c, err);
        if (e) return e;
    }
    if (c.steps_done >= 0) return PDX_OK;
    // start-up: u_0 from DOF values, BCs advanced once before the first force (pdsolve.cpp:5973-5992)
    const Model &m = *c.host;
    const int D = c.dim;
    std::vector<double> u0((size_t)c.nn * D);
    host_advance_bc(c);
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < D; i++) {
            int ps = m.presc_set[(size_t)k * 3 + i];
            u0[(size_t)k * D + i] = ps >= 0 ? (m.ebc[ps].vary ? c.ebc_val[ps] : m.dofval[(size_t)k * 3 + i])
                                            : m.dofval[(size_t)k * 3 + i];
        }
    PDX_HIP(hipMemcpyAsync(c.U[0], u0.data(), sizeof(double) * u0.size(), hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemcpyAsync(c.U[2], u0.data(), sizeof(double) * u0.size(), hipMemcpyHostToDevice, c.stream));
    std::vector<double> inc(std::max<size_t>(c.ebc_val.size(), 1), 0.0), ninc(std::max(c.n_nbc, 1), 0.0);
    for (size_t s = 0; s < c.ebc_val.size(); s++) inc[s] = c.ebc_vary[s] ? c.dt * c.ebc_vel[s] : 0.0;
    for (int s = 0; s < c.n_nbc; s++) ninc[s] = c.nbc_vary[s] ? c.dt * c.nbc_vel[s] : 0.0;
    PDX_HIP(hipMemcpyAsync(c.d_ebc_vel, inc.data(), sizeof(double) * inc.size(), hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemcpyAsync(c.d_nbc_vel, ninc.data(), sizeof(double) * ninc.size(), hipMemcpyHostToDevice, c.stream));
    if (c.n_nbc)
        PDX_HIP(hipMemcpyAsync(c.nbcv, c.nbc_val.data(), sizeof(double) * c.n_nbc, hipMemcpyHostToDevice, c.stream));
    c.step_phase = 0;
    explicit_step(c, prec, 1);
    host_advance_bc(c);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

int explicit_tune_if_needed(Ctx &c, int prec, std::string &err) {
    const int pi = prec == PDX_FP32 ? 1 : 0;
    if (c.tune_explicit[pi] >= 0) return PDX_OK;
    int rc = tune_explicit(c, pi, err);
    for (auto &g : c.graph)
        if (g) { hipGraphExecDestroy(g); g = nullptr; }
    return rc;
}

int dev_explicit_steps(Ctx &c, int nsteps, int prec, int use_graph, std::string &err) {
    int e = explicit_prepare(c, prec, err);
    if (e) return e;
    if ((e = explicit_tune_if_needed(c, prec, err))) return e;
    hipEvent_t e0, e1;
    PDX_HIP(hipEventCreate(&e0));
    PDX_HIP(hipEventCreate(&e1));
    const int period = 6;
    int done = 0;
    PDX_HIP(hipEventRecord(e0, c.stream));
    if (use_graph && nsteps >= period) {
        // align to the buffer / NBC rotation period, then replay a captured 6-step graph
        while ((c.step_phase % 3 != 0 || (c.steps_done + 1) % 2 != 0) && done < nsteps) {
            explicit_step(c, prec, 0);
            done++;
        }
        int gi = prec == PDX_FP32 ? 1 : 0;
        if (!c.graph[gi]) {
            hipGraph_t g;
            const int sp = c.step_phase;
            const long long sd = c.steps_done;
            PDX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < period; i++) explicit_step(c, prec, 0);
            PDX_HIP(hipStreamEndCapture(c.stream, &g));
            PDX_HIP(hipGraphInstantiate(&c.graph[gi], g, nullptr, nullptr, 0));
            PDX_HIP(hipGraphDestroy(g));
            c.step_phase = sp;
            c.steps_done = sd;
        }
        while (nsteps - done >= period) {
            PDX_HIP(hipGraphLaunch(c.graph[gi], c.stream));
            c.steps_done += period;
            done += period;
        }
    }
    while (done < nsteps) {
        explicit_step(c, prec, 0);
        done++;
    }
    PDX_HIP(hipEventRecord(e1, c.stream));
    PDX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    PDX_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c.last_total_ms = ms;
    c.last_step_ms = nsteps ? ms / nsteps : 0;
    for (int i = 0; i < nsteps; i++) host_advance_bc(c);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

// u_{n+1} (current) and u_n of the explicit state, free DOFs in equation order
int dev_explicit_state(Ctx &c, double *u_out, double *u_prev, std::string &err) {
    const int D = c.dim;
    std::vector<double> a((size_t)c.nn * D), b((size_t)c.nn * D);
    const int p = c.step_phase % 3;
    PDX_HIP(hipMemcpyAsync(a.data(), c.U[p], sizeof(double) * a.size(), hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipMemcpyAsync(b.data(), c.U[(p + 2) % 3], sizeof(double) * b.size(), hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    const Model &m = *c.host;
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < D; i++) {
            int r = m.eq[(size_t)k * 3 + i];
            if (r < 0) continue;
            if (u_out) u_out[r] = a[(size_t)k * D + i];
            if (u_prev) u_prev[r] = b[(size_t)k * D + i];
        }
    return PDX_OK;
}

// ------------------------------------------------------------------------
// power iteration for lambda_max(M^-1 K) on free DOFs
// ------------------------------------------------------------------------
int dev_max_eig(Ctx &c, int iters, double *lmax, std::string &err) {
    if (!c.K_valid) {
        int e = dev_assemble_K(c, err);
        if (e) return e;
    }
    if (!c.mlump_valid) {
        int e = dev_lumped_mass(c, err);
        if (e) return e;
    }
    const size_t n = (size_t)c.nn * c.dim;
    double *v, *y, *part;
    PDX_HIP(hipMalloc((void **)&v, n * 8));
    PDX_HIP(hipMalloc((void **)&y, n * 8));
    PDX_HIP(hipMalloc((void **)&part, NPART * 8));
    std::vector<double> h(n);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i < n; i++) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        h[i] = 0.5 + double(s % 1000003) / 1000003.0;
    }
    PDX_HIP(hipMemcpy(v, h.data(), n * 8, hipMemcpyHostToDevice));
    double lam = 0.0;
    for (int it = 0; it < iters; it++) {
        if (dev_spmv(c, v, y, PDX_FP64, true)) { err = "spmv failed"; return PDX_ERR_HIP; }
        k_scale_mass<<<2048, 256, 0, c.stream>>>(n, y, c.mlump, c.eq, c.dim, y);
        k_dot<double><<<RED_PARTS, 256, 0, c.stream>>>(v, y, n, part);
        double vy = host_sum(c, part, NPART, err);
        k_dot<double><<<RED_PARTS, 256, 0, c.stream>>>(v, v, n, part);
        double vv = host_sum(c, part, NPART, err);
        k_dot<double><<<RED_PARTS, 256, 0, c.stream>>>(y, y, n, part);
        double yy = host_sum(c, part, NPART, err);
        lam = vy / vv;
        double sc = 1.0 / std::sqrt(yy);
        PDX_HIP(hipMemsetAsync(v, 0, n * 8, c.stream));
        k_axpy_d<<<2048, 256, 0, c.stream>>>(n, sc, y, v);
    }
    PDX_HIP(hipStreamSynchronize(c.stream));
    hipFree(v);
    hipFree(y);
    hipFree(part);
    *lmax = lam;
    return PDX_OK;
}

// ------------------------------------------------------------------------
// BiCGStab on the masked node-block operator (right block-Jacobi preconditioning)
// ------------------------------------------------------------------------
template <typename V>
struct KrylovWs {
    V *x = nullptr, *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr,
      *ph = nullptr, *sh = nullptr, *Minv = nullptr;
};

template <typename V>
static int ws_alloc(Ctx &c, KrylovWs<V> &w, std::string &err) {
    const size_t n = (size_t)c.nn * c.dim;
    V **ptrs[] = {&w.x, &w.r, &w.rh, &w.p, &w.v, &w.s, &w.t, &w.ph, &w.sh};
    for (auto pp : ptrs) {
        PDX_HIP(hipMalloc((void **)pp, n * sizeof(V)));
        PDX_HIP(hipMemsetAsync(*pp, 0, n * sizeof(V), c.stream));
    }
    PDX_HIP(hipMalloc((void **)&w.Minv, (size_t)c.nn * c.dd * sizeof(V)));
    return PDX_OK;
}

template <typename V>
static void ws_free(KrylovWs<V> &w) {
    V *ptrs[] = {w.x, w.r, w.rh, w.p, w.v, w.s, w.t, w.ph, w.sh, w.Minv};
    for (auto p : ptrs)
        if (p) hipFree(p);
}

template <int D, typename T, typename V>
static void spmv_dot(Ctx &c, const T *K, const V *x, V *y, const V *w0, double *p0, const V *w1, double *p1) {
    k_spmv_dot<D, T, V><<<RED_PARTS, 256, 0, c.stream>>>(c.rows_short, c.n_short, c.rows_long, c.n_long, c.adj_off,
                                                          c.adj, K, x, y, c.eq, w0, p0, w1, p1);
}

// one BiCGStab iteration (it = iteration index, parity selects the rho slot)
template <int D, typename T, typename V>
static void bicg_iter(Ctx &c, const T *K, KrylovWs<V> &w, double *red, double *sc, int it) {
    double *rho = red + (size_t)((it & 1) ? P_RHO1 : P_RHO0) * NPART;
    double *rho_next = red + (size_t)((it & 1) ? P_RHO0 : P_RHO1) * NPART;
    double *rv = red + (size_t)P_RV * NPART, *ts = red + (size_t)P_TS * NPART, *tt = red + (size_t)P_TT * NPART,
           *rr = red + (size_t)P_RR * NPART;
    k_bicg_p<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.r, w.p, w.v, w.ph, w.Minv, rho, sc, it == 0);
    spmv_dot<D, T, V>(c, K, w.ph, w.v, w.rh, rv, (const V *)nullptr, nullptr);
    k_bicg_s<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.r, w.v, w.s, w.sh, w.Minv, rho, rv, sc);
    spmv_dot<D, T, V>(c, K, w.sh, w.t, w.s, ts, w.t, tt);
    k_bicg_x<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.x, w.r, w.rh, w.ph, w.sh, w.s, w.t, rho, ts, tt, sc,
                                                    rho_next, rr);
}

// Solve K_ff x = b (b, x: full DOF vectors, prescribed entries zero) with
// preconditioned BiCGStab in precision V/T to relative residual tol (recurrence
// residual), starting from x (device, type V).  Returns iterations.
template <int D, typename T, typename V>
static int bicgstab(Ctx &c, const T *K, KrylovWs<V> &w, const V *b, double bnorm, double tol, int max_iter,
                    int *iters, double *relres, std::string &err) {
    const size_t n = (size_t)c.nn * c.dim;
    double *red = c.red, *sc = c.red + (size_t)RED_SLOTS * NPART;
    double *y;
    PDX_HIP(hipMalloc((void **)&y, n * sizeof(double)));
    // r = b - K x ; rh = r
    {
        V *tmp = w.t;
        spmv_dot<D, T, V>(c, K, w.x, tmp, (const V *)nullptr, nullptr, (const V *)nullptr, nullptr);
        k_copy<V, double><<<2048, 256, 0, c.stream>>>(n, tmp, y);
        k_resid<V, V><<<RED_PARTS, 256, 0, c.stream>>>(n, b, y, w.r, w.rh, red + (size_t)P_RR * NPART);
        k_dot<V><<<RED_PARTS, 256, 0, c.stream>>>(w.rh, w.r, n, red + (size_t)P_RHO0 * NPART);
        PDX_HIP(hipMemsetAsync(w.p, 0, n * sizeof(V), c.stream));
        PDX_HIP(hipMemsetAsync(w.v, 0, n * sizeof(V), c.stream));
    }
    hipFree(y);
    double rr0 = host_sum(c, red + (size_t)P_RR * NPART, NPART, err);
    double rel = std::sqrt(rr0) / bnorm;
    int it = 0, since_best = 0;
    const int chunk = 10;
    double best = rel;
    while (rel > tol && it < max_iter) {
        for (int k = 0; k < chunk && it < max_iter; k++, it++) bicg_iter<D, T, V>(c, K, w, red, sc, it);
        PDX_HIP(hipGetLastError());
        double rr = host_sum(c, red + (size_t)P_RR * NPART, NPART, err);
        if (!std::isfinite(rr)) {
            std::vector<double> hs(8);
            hipMemcpy(hs.data(), sc, 8 * sizeof(double), hipMemcpyDeviceToHost);
            err = "Krylov breakdown (non-finite residual) at iteration " + std::to_string(it) + " after relres " +
                  std::to_string(rel) + " (best " + std::to_string(best) + "), rho_old " + std::to_string(hs[0]) +
                  " alpha " + std::to_string(hs[1]) + " omega " + std::to_string(hs[2]) + ", value bytes " +
                  std::to_string(sizeof(V));
            *iters = it;
            return PDX_ERR_NOCONV;
        }
        rel = std::sqrt(rr) / bnorm;
        // stagnation (rounding floor of the working precision): hand back to the caller
        if (rel < 0.95 * best) { best = rel; since_best = 0; }
        else if (++since_best >= 40) break;
    }
    *iters = it;
    *relres = rel;
    return PDX_OK;
}

template <int D, typename V>
static int build_precond(Ctx &c, const double *K64, int kind, V *Minv, std::string &err) {
    int *status;
    PDX_HIP(hipMalloc((void **)&status, sizeof(int)));
    PDX_HIP(hipMemsetAsync(status, 0, sizeof(int), c.stream));
    if (kind == PDX_PREC_NONE) {
        std::vector<V> h((size_t)c.nn * c.dd, V(0));
        std::vector<int> eqh((size_t)c.nn * 3);
        PDX_HIP(hipMemcpy(eqh.data(), c.eq, eqh.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int a = 0; a < c.nn; a++)
            for (int i = 0; i < D; i++)
                if (eqh[(size_t)a * 3 + i] >= 0) h[(size_t)a * c.dd + i * D + i] = V(1);
        PDX_HIP(hipMemcpyAsync(Minv, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice, c.stream));
        PDX_HIP(hipStreamSynchronize(c.stream));
    } else {
        k_blockjacobi<D, V><<<(c.nn + 255) / 256, 256, 0, c.stream>>>(c.nn, c.adj_off, c.adj, K64, c.nblk, c.eq,
                                                                      kind == PDX_PREC_JACOBI, Minv, status);
    }
    int hs = 0;
    PDX_HIP(hipMemcpyAsync(&hs, status, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    hipFree(status);
    if (hs) { err = "singular diagonal block in the preconditioner"; return PDX_ERR_SINGULAR; }
    return PDX_OK;
}

// Solve K_ff x = b in FP64 (plain) or FP32 inner + FP64 refinement.
// b_dev: full DOF vector (double, prescribed zero).  x_dev: in/out (double).
template <int D>
static int solve_masked(Ctx &c, const double *K64, const float *K32, const pdx_solve_opts &o, const double *b_dev,
                        double *x_dev, pdx_solve_stats &st, std::string &err) {
    auto t0 = std::chrono::steady_clock::now();
    const size_t n = (size_t)c.nn * c.dim;
    if (!c.red) {
        PDX_HIP(hipMalloc((void **)&c.red, sizeof(double) * ((size_t)RED_SLOTS * NPART + 16)));
        PDX_HIP(hipMemsetAsync(c.red, 0, sizeof(double) * ((size_t)RED_SLOTS * NPART + 16), c.stream));
    }
    double *part = c.red + (size_t)P_BB * NPART;
    k_dot<double><<<RED_PARTS, 256, 0, c.stream>>>(b_dev, b_dev, n, part);
    double bnorm = std::sqrt(host_sum(c, part, NPART, err));
    if (bnorm == 0.0) {
        PDX_HIP(hipMemsetAsync(x_dev, 0, n * 8, c.stream));
        st.true_relres = 0; st.converged = 1; st.iterations = 0; st.refinements = 0;
        return PDX_OK;
    }
    if (!o.warm_start) PDX_HIP(hipMemsetAsync(x_dev, 0, n * 8, c.stream));
    int total = 0, refine = 0, rc = PDX_OK;
    double rel = 1.0;
    double *y;
    PDX_HIP(hipMalloc((void **)&y, n * 8));
    if (o.precision == PDX_FP64) {
        KrylovWs<double> w;
        if ((rc = ws_alloc(c, w, err)) || (rc = build_precond<D, double>(c, K64, o.precond, w.Minv, err))) { ws_free(w); hipFree(y); return rc; }
        PDX_HIP(hipMemcpyAsync(w.x, x_dev, n * 8, hipMemcpyDeviceToDevice, c.stream));
        for (int restart = 0; restart < 5; restart++) {
            int it = 0;
            double rr;
            rc = bicgstab<D, double, double>(c, K64, w, b_dev, bnorm, o.tol * 0.5, o.max_iter - total, &it, &rr, err);
            total += it;
            if (rc) break;
            // true residual
            spmv_dot<D, double, double>(c, K64, w.x, y, (const double *)nullptr, nullptr, (const double *)nullptr, nullptr);
            k_resid<double, double><<<RED_PARTS, 256, 0, c.stream>>>(n, b_dev, y, w.t, nullptr, part);
            rel = std::sqrt(host_sum(c, part, NPART, err)) / bnorm;
            if (rel <= o.tol || total >= o.max_iter) break;
        }
        PDX_HIP(hipMemcpyAsync(x_dev, w.x, n * 8, hipMemcpyDeviceToDevice, c.stream));
        ws_free(w);
    } else {
        if (!K32) { err = "FP32 matrix missing"; hipFree(y); return PDX_ERR_ARG; }
        KrylovWs<float> w;
        float *b32;
        PDX_HIP(hipMalloc((void **)&b32, n * 4));
        if ((rc = ws_alloc(c, w, err)) || (rc = build_precond<D, float>(c, K64, o.precond, w.Minv, err))) { ws_free(w); hipFree(y); hipFree(b32); return rc; }
        const double inner = o.inner_tol > 0 ? o.inner_tol : 1e-4;
        for (refine = 0; refine < 40; refine++) {
            // FP64 residual r = b - K x
            spmv_dot<D, double, double>(c, K64, x_dev, y, (const double *)nullptr, nullptr, (const double *)nullptr, nullptr);
            k_resid<float, double><<<RED_PARTS, 256, 0, c.stream>>>(n, b_dev, y, b32, nullptr, part);
            rel = std::sqrt(host_sum(c, part, NPART, err)) / bnorm;
            if (rel <= o.tol || total >= o.max_iter) break;
            PDX_HIP(hipMemsetAsync(w.x, 0, n * 4, c.stream));
            int it = 0;
            double rr;
            rc = bicgstab<D, float, float>(c, K32, w, b32, rel * bnorm, inner, o.max_iter - total, &it, &rr, err);
            total += it;
            if (rc) break;
            k_add_f2d<<<2048, 256, 0, c.stream>>>(n, w.x, x_dev);
        }
        ws_free(w);
        hipFree(b32);
    }
    hipFree(y);
    PDX_HIP(hipStreamSynchronize(c.stream));
    st.iterations = total;
    st.refinements = refine;
    st.true_relres = rel;
    st.converged = rel <= o.tol;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) return rc;
    if (!st.converged) { err = "Krylov solver did not reach the tolerance"; return PDX_ERR_NOCONV; }
    return PDX_OK;
}

int to_f32(Ctx &c, const double *K64, float *&K32, std::string &err) {
    size_t n = (size_t)c.dd * c.nblk;
    if (!K32) PDX_HIP(hipMalloc((void **)&K32, std::max<size_t>(n, 1) * sizeof(float)));
    k_to_f32<<<4096, 256, 0, c.stream>>>(n, K64, K32);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

// solve with an explicit node-block matrix (K, K_eff or a 1x1-block CSR)
int dev_solve_mat(Ctx &c, const double *K64, const float *K32, const pdx_solve_opts &o, const double *b_dev,
                  double *x_dev, pdx_solve_stats &st, std::string &err) {
    if (o.krylov != PDX_KRYLOV_BICGSTAB) { err = "only BiCGStab is available on the node-block operator"; return PDX_ERR_ARG; }
    if (c.dim == 3) return solve_masked<3>(c, K64, K32, o, b_dev, x_dev, st, err);
    if (c.dim == 2) return solve_masked<2>(c, K64, K32, o, b_dev, x_dev, st, err);
    return solve_masked<1>(c, K64, K32, o, b_dev, x_dev, st, err);
}

int dev_solve(Ctx &c, const pdx_solve_opts &o, const double *b_dev, double *x_dev, pdx_solve_stats &st,
              std::string &err) {
    if (!c.K_valid) { err = "K not assembled"; return PDX_ERR_ARG; }
    int rc;
    if (o.precision == PDX_FP32 && (rc = ensure_K32(c, err))) return rc;
    return dev_solve_mat(c, c.K, o.precision == PDX_FP32 ? c.K32 : nullptr, o, b_dev, x_dev, st, err);
}

}  // namespace pdx
