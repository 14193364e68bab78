// Explicit central-difference steps and the power iteration for the
// critical time step, on the node-block matrix, for CDNA4 (gfx950).
//
// Explicit step (pdsolve.cpp:5842-6090, timeIntegration_LM :5113-5132): one
// fused kernel per row class computes y = K u_n for a node's rows and updates
//   u_{n+1} = 2 u_n + dt^2 (F_ext - K u)/m - u_{n-1}   (free DOFs)
//   v_{n+2} = v_{n+1} + dt * velocity                  (prescribed DOFs)
// with three rotating state vectors, so a step is a pure function of device
// state and a run of steps can be replayed from a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

#include "device.hpp"
#include "reduce.hpp"
#include "rowblock.hpp"

namespace pdx {

// ------------------------------------------------------------------------
// explicit step
// ------------------------------------------------------------------------
// FP32 values use the difference form (see rowblock.hpp), FP64 the plain form.
// GS = lanes per short row, NT = non-temporal loads of K, XM = XCD-aware block
// order (chosen by the tuner).  The hardware deals workgroups to the 8 XCDs
// round-robin; with XM the launch has 8 x (an eighth of the short-row blocks + an
// eighth of the long-row blocks) workgroups and the ones XCD x runs take the x-th
// contiguous eighth of each, so every XCD (and its L2) owns one contiguous range of
// rows of each kind, the gathers of u_n stay local to it, and the wavefront-wide
// long rows (the PD nodes) do not all land on one XCD.  With rev
// the blocks of each range are visited last to first (alternating forms reverse
// every other step), which changes which rows run when, not what a row computes.
// column pattern classes (REL forms): class of each row, offsets of each class's relative columns
struct ColClasses {
    const int *cls, *cls_off, *cls_rel;
    const int *ksrc;  // mirrored forms: own index of every block (~partner's for a mirrored one)
};

// FO (force only, consistent mass): the row's F = F_ext - K u_n is stored to Unp1 (0 at prescribed
// DOFs) instead of the update; the acceleration then comes from M a = F.
template <int D, typename T, int GS, bool NT, bool XM, bool SOA, bool REL, bool SYM, bool FO = false>
__global__ void __launch_bounds__(256) k_explicit(const int *rows_s, int ns, int nbs, const int *rows_l, int nl,
                                                  int nbt, const int *adj_off, const int *adj, ColClasses cc, const T *K,
                                                  const double *Un, const double *Unm1, double *Unp1,
                                                  const double *m, const double *f_unit, int n_nbc, int nn,
                                                  const double *nbc_cur, double *nbc_next, const double *nbc_inc,
                                                  const int *eq, const int *presc_set, const double *ebc_inc,
                                                  double dt, int startup, int rev) {
    constexpr bool DIFF = sizeof(T) == 4;
    int lb = blockIdx.x;
    if (XM) {
        // each XCD takes one contiguous eighth of the short-row blocks and one of the long-row
        // blocks (short ones first), so the long rows' bytes are spread over all eight L2s
        const int per = gridDim.x / 8, q0 = blockIdx.x / 8, x = blockIdx.x % 8;
        const int q = rev ? per - 1 - q0 : q0;
        const int ps = (nbs + 7) / 8;
        if (q < ps) {
            lb = x * ps + q;
            if (lb >= nbs) return;
        } else {
            lb = nbs + x * (per - ps) + (q - ps);
            if (lb >= nbt) return;
        }
    } else if (rev) {
        lb = gridDim.x - 1 - blockIdx.x;
    }
    if (lb == 0)
        for (int s = threadIdx.x; s < n_nbc; s += blockDim.x) nbc_next[s] = nbc_cur[s] + nbc_inc[s];
    T acc[D];
    int r, lane, grp;
    if (lb < nbs) {
        const int gid = (lb * blockDim.x + threadIdx.x) / GS;
        lane = threadIdx.x % GS;
        if (gid >= ns) return;
        r = rows_s[gid];
        grp = GS;
    } else {
        const int gid = ((lb - nbs) * blockDim.x + threadIdx.x) / 64;
        lane = threadIdx.x % 64;
        if (gid >= nl) return;
        r = rows_l[gid];
        grp = 64;
    }
    // the row's DOF state is loaded before the row sum, so its latency overlaps the block gathers
    const size_t o = (size_t)r * D + (lane < D ? lane : 0);
    const int code = r * 3 + (lane < D ? lane : 0);
    double un = 0.0, unm1 = 0.0, mo = 1.0, fext = 0.0, inc = 0.0;
    const bool act = lane < D;
    const bool fr = act && eq[code] >= 0;
    if (act) {
        un = Un[o];
        if (fr) {
            unm1 = Unm1[o];
            mo = m[o];
            for (int s = 0; s < n_nbc; s++) fext = fext + nbc_cur[s] * f_unit[(size_t)s * nn * D + o];
        } else {
            inc = ebc_inc[presc_set[code]];
        }
    }
    const int *rel = REL ? cc.cls_rel + cc.cls_off[cc.cls[r]] : nullptr;
    if (grp == GS) row_matvec<D, T, double, GS, DIFF, NT, SOA, REL, SYM>(r, lane, adj_off, adj, K, Un, acc, rel, cc.ksrc);
    else row_matvec<D, T, double, 64, DIFF, NT, SOA, REL, SYM>(r, lane, adj_off, adj, K, Un, acc, rel, cc.ksrc);
    if (act) {
        const T v = pick<D, T>(acc, lane);
        if (FO) {
            Unp1[o] = fr ? fext - double(v) : 0.0;
        } else if (fr) {
            const double F = fext - double(v);
            if (startup) {
                const double a = F / mo;
                Unp1[o] = a * (0.5 * dt * dt) + un;
            } else {
                const double tA = dt * dt * F / mo;
                Unp1[o] = un * 2.0 + tA - unm1;
            }
        } else {
            Unp1[o] = un + inc;
        }
    }
}

// Consistent-mass update (timeIntegration_NLM, pdsolve.cpp:5134-5152, and the start-up step
// :6008-6016): u_{n+1} = 2 u_n + dt^2 a - u_{n-1}, or u_1 = a dt^2/2 + u_0; prescribed DOFs
// advance by their increment, as in the lumped kernel.
__global__ void k_nlm_update(int n, int D, const int *__restrict__ eq, const int *__restrict__ presc_set,
                             const double *__restrict__ Un, const double *__restrict__ Unm1,
                             const double *__restrict__ acc, const double *__restrict__ ebc_inc, double dt,
                             int startup, double *__restrict__ Unp1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int code = (t / D) * 3 + t % D;
    if (eq[code] >= 0) {
        if (startup) {
            Unp1[t] = acc[t] * (0.5 * dt * dt) + Un[t];
        } else {
            const double tA = dt * dt * acc[t];
            Unp1[t] = Un[t] * 2.0 + tA - Unm1[t];
        }
    } else {
        Unp1[t] = Un[t] + ebc_inc[presc_set[code]];
    }
}

// DOF values of the last stored step: free entries of u_{n+1}, prescribed entries of the
// previous buffer (the values that step's force used; the current buffer already holds the
// next step's prescribed values)
__global__ void k_merge_dofvals(int n, int D, const int *__restrict__ eq, const double *__restrict__ cur,
                                const double *__restrict__ prev, double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    out[t] = eq[(t / D) * 3 + t % D] >= 0 ? cur[t] : prev[t];
}

__global__ void k_axpy_d(size_t n, double a, const double *x, double *y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = y[i] + a * x[i];
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

int ensure_K32(Ctx &c, std::string &err) {
    if (!c.K_valid) { err = "K not assembled"; return PDX_ERR_ARG; }
    if (c.K32_valid) return PDX_OK;
    if (!c.K32) PDX_HIP(hipMalloc((void **)&c.K32, sizeof(float) * c.dd * c.nblk));
    size_t n = (size_t)c.dd * c.nblk;
    k_to_f32<<<4096, 256, 0, c.stream>>>(n, c.K, c.K32);
    PDX_HIP(hipGetLastError());
    c.K32_valid = true;
    return PDX_OK;
}

// K in row-local value order (one wave per node row), from the FP64 K
template <typename T>
__global__ void k_row_soa(int nn, int dd, const int *__restrict__ adj_off, const double *__restrict__ K, T *Ks) {
    const int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, lane = threadIdx.x & 63;
    if (r >= nn) return;
    const int b0 = adj_off[r], nb = adj_off[r + 1] - b0;
    for (int v = lane; v < nb * dd; v += 64) {
        const int b = v / dd, e = v % dd;
        Ks[(size_t)b0 * dd + (size_t)e * nb + b] = T(K[((size_t)b0 + b) * dd + e]);
    }
}

// own blocks of K (mirrored storage) in precision T
template <typename T>
__global__ void k_gather_own(long long n_own, int dd, const int *__restrict__ slot, const double *__restrict__ K,
                             T *Kown) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n_own * dd;
         v += (long long)gridDim.x * blockDim.x) {
        const long long q = v / dd;
        Kown[v] = T(K[(size_t)slot[q] * dd + (v - q * dd)]);
    }
}

// *bad = 1 when a mirrored lower block (ksrc < 0) differs from its partner's transpose in any bit
__global__ void k_check_mirror(long long nblk, int D, const int *__restrict__ ksrc, const int *__restrict__ slot,
                               const double *__restrict__ A, int *bad) {
    const int dd = D * D;
    for (long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x; b < nblk;
         b += (long long)gridDim.x * blockDim.x) {
        const int s = ksrc[b];
        if (s >= 0) continue;
        const double *lo = A + (size_t)b * dd, *up = A + (size_t)slot[~s] * dd;
        bool same = true;
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++)
                same = same && __double_as_longlong(lo[i * D + j]) == __double_as_longlong(up[j * D + i]);
        if (!same) *bad = 1;
    }
}

int check_mirrored(Ctx &c, const double *A, bool &exact, std::string &err) {
    exact = false;
    if (!c.ksrc || c.dim < 2) return PDX_OK;
    int *bad, h = 0;
    PDX_HIP(hipMalloc((void **)&bad, sizeof(int)));
    PDX_HIP(hipMemsetAsync(bad, 0, sizeof(int), c.stream));
    k_check_mirror<<<4096, 256, 0, c.stream>>>(c.nblk, c.dim, c.ksrc, c.kown_slot, A, bad);
    PDX_HIP(hipGetLastError());
    PDX_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    hipFree(bad);
    exact = h == 0;
    (void)err;
    return PDX_OK;
}

int gather_own(Ctx &c, const double *A, double *own64, float *own32, std::string &err) {
    if (!c.ksrc) { err = "no mirrored storage map"; return PDX_ERR_ARG; }
    if (own64) k_gather_own<double><<<4096, 256, 0, c.stream>>>(c.n_own, c.dd, c.kown_slot, A, own64);
    if (own32) k_gather_own<float><<<4096, 256, 0, c.stream>>>(c.n_own, c.dd, c.kown_slot, A, own32);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

// the mirrored-storage copy of K for the explicit step, rebuilt when K has been assembled again
static int ensure_Kown(Ctx &c, int prec, std::string &err) {
    const int pi = prec == PDX_FP32 ? 1 : 0;
    if (!c.K_valid) { err = "K not assembled"; return PDX_ERR_ARG; }
    if (c.kown_gen[pi] == c.K_gen) return PDX_OK;
    const size_t n = (size_t)c.dd * std::max<long long>(c.n_own, 1);
    if (pi) {
        if (!c.Kown32) PDX_HIP(hipMalloc((void **)&c.Kown32, sizeof(float) * n));
        k_gather_own<float><<<4096, 256, 0, c.stream>>>(c.n_own, c.dd, c.kown_slot, c.K, c.Kown32);
    } else {
        if (!c.Kown64) PDX_HIP(hipMalloc((void **)&c.Kown64, sizeof(double) * n));
        k_gather_own<double><<<4096, 256, 0, c.stream>>>(c.n_own, c.dd, c.kown_slot, c.K, c.Kown64);
    }
    PDX_HIP(hipGetLastError());
    c.kown_gen[pi] = c.K_gen;
    return PDX_OK;
}

int kown_for_solve(Ctx &c, int prec, Mirrored &m, std::string &err) {
    m = Mirrored{};
    if (!c.K_valid) { err = "K not assembled"; return PDX_ERR_ARG; }
    int rc;
    if (c.kown_exact_gen != c.K_gen) {
        if ((rc = check_mirrored(c, c.K, c.kown_exact, err))) return rc;
        c.kown_exact_gen = c.K_gen;
    }
    if (!c.kown_exact) return PDX_OK;
    // the FP64 copy serves the true residuals of both precisions
    if ((rc = ensure_Kown(c, PDX_FP64, err))) return rc;
    m.K64 = c.Kown64;
    if (prec == PDX_FP32) {
        if ((rc = ensure_Kown(c, PDX_FP32, err))) return rc;
        m.K32 = c.Kown32;
    }
    return PDX_OK;
}

// the row-local copy of K for the explicit step, rebuilt when K has been assembled again
static int ensure_Ks(Ctx &c, int prec, std::string &err) {
    const int pi = prec == PDX_FP32 ? 1 : 0;
    if (!c.K_valid) { err = "K not assembled"; return PDX_ERR_ARG; }
    if (c.ks_gen[pi] == c.K_gen) return PDX_OK;
    const size_t n = (size_t)c.dd * c.nblk;
    const int grid = (c.nn + 3) / 4;
    if (pi) {
        if (!c.K32s) PDX_HIP(hipMalloc((void **)&c.K32s, sizeof(float) * std::max<size_t>(n, 1)));
        k_row_soa<float><<<grid, 256, 0, c.stream>>>(c.nn, c.dd, c.adj_off, c.K, c.K32s);
    } else {
        if (!c.K64s) PDX_HIP(hipMalloc((void **)&c.K64s, sizeof(double) * std::max<size_t>(n, 1)));
        k_row_soa<double><<<grid, 256, 0, c.stream>>>(c.nn, c.dd, c.adj_off, c.K, c.K64s);
    }
    PDX_HIP(hipGetLastError());
    c.ks_gen[pi] = c.K_gen;
    return PDX_OK;
}

template <int D, typename T, int GS, bool NT, bool XM, bool SOA, bool REL, bool SYM, bool FO = false>
static void launch_explicit_v(Ctx &c, const T *K, const double *Un, const double *Unm1, double *Unp1,
                              const double *nbc_cur, double *nbc_next, int startup, int rev = 0) {
    const int nbs = (c.n_short * GS + 255) / 256, nbl = (c.n_long * 64 + 255) / 256;
    const int nbt = std::max(nbs + nbl, 1);
    // XM: 8 XCDs x (an eighth of the short blocks + an eighth of the long ones)
    const int grid = XM ? 8 * ((nbs + 7) / 8 + (nbt - nbs + 7) / 8) : nbt;
    const ColClasses cc{c.cls, c.cls_off, c.cls_rel, c.ksrc};
    k_explicit<D, T, GS, NT, XM, SOA, REL, SYM, FO><<<grid, 256, 0, c.stream>>>(
        c.rows_short, c.n_short, nbs, c.rows_long, c.n_long, nbt, c.adj_off, c.adj, cc, K, Un, Unm1, Unp1, c.mlump,
        c.f_unit, c.n_nbc, c.nn, nbc_cur, nbc_next, c.d_nbc_vel, c.eq, c.presc_set, c.d_ebc_vel, c.dt, startup, rev);
}

// one launch of form `variant` of kExplicitForms (K: the plain or the row-local copy, as the form reads)
// for step `step` (its parity sets the sweep direction of an alternating form)
template <int D, typename T>
static void launch_explicit(Ctx &c, int variant, const T *K, const double *Un, const double *Unm1, double *Unp1,
                            const double *nbc_cur, double *nbc_next, int startup, long long step) {
    const int rev = kExplicitForms[variant].alt ? int(step & 1) : 0;
    variant = kExplicitForms[variant].base;
#define PDX_EXPLICIT_FORM(V, GS, NT, XM, SOA, REL, SYM)                                                         \
    case V:                                                                                                     \
        static_assert(kExplicitForms[V].lanes == GS && kExplicitForms[V].nt == NT && kExplicitForms[V].xm == XM && \
                          kExplicitForms[V].soa == SOA && kExplicitForms[V].rel == REL &&                       \
                          kExplicitForms[V].sym == SYM,                                                         \
                      "form table");                                                                            \
        launch_explicit_v<D, T, GS, NT, XM, SOA, REL, SYM>(c, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup, rev); \
        break;
    switch (variant) {
        PDX_EXPLICIT_FORM(1, 4, false, false, false, false, false)
        PDX_EXPLICIT_FORM(2, 16, false, false, false, false, false)
        PDX_EXPLICIT_FORM(3, 8, true, false, false, false, false)
        PDX_EXPLICIT_FORM(4, 4, true, false, false, false, false)
        PDX_EXPLICIT_FORM(5, 16, true, false, false, false, false)
        PDX_EXPLICIT_FORM(6, 32, false, false, false, false, false)
        PDX_EXPLICIT_FORM(7, 32, true, false, false, false, false)
        PDX_EXPLICIT_FORM(8, 16, false, true, false, false, false)
        PDX_EXPLICIT_FORM(9, 8, false, true, false, false, false)
        PDX_EXPLICIT_FORM(10, 8, false, true, true, false, false)
        PDX_EXPLICIT_FORM(11, 16, false, true, true, false, false)
        PDX_EXPLICIT_FORM(12, 4, false, true, true, false, false)
        PDX_EXPLICIT_FORM(13, 32, false, true, true, false, false)
        PDX_EXPLICIT_FORM(14, 8, false, true, false, true, false)
        PDX_EXPLICIT_FORM(15, 16, false, true, false, true, false)
        PDX_EXPLICIT_FORM(16, 8, false, true, true, true, false)
        PDX_EXPLICIT_FORM(17, 16, false, true, true, true, false)
        PDX_EXPLICIT_FORM(18, 8, false, true, false, true, true)
        PDX_EXPLICIT_FORM(19, 16, false, true, false, true, true)
    default:
        PDX_EXPLICIT_FORM(0, 8, false, false, false, false, false)
    }
#undef PDX_EXPLICIT_FORM
}

static void launch_explicit_any(Ctx &c, int prec, int variant, const double *Un, const double *Unm1, double *Unp1,
                                const double *nbc_cur, double *nbc_next, int startup, long long step) {
    const bool soa = kExplicitForms[variant].soa, sym = kExplicitForms[variant].sym;
    if (prec == PDX_FP32) {
        const float *K = sym ? c.Kown32 : soa ? c.K32s : c.K32;
        if (c.dim == 3) launch_explicit<3, float>(c, variant, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup, step);
        else launch_explicit<2, float>(c, variant, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup, step);
    } else {
        const double *K = sym ? c.Kown64 : soa ? c.K64s : c.K;
        if (c.dim == 3) launch_explicit<3, double>(c, variant, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup, step);
        else launch_explicit<2, double>(c, variant, K, Un, Unm1, Unp1, nbc_cur, nbc_next, startup, step);
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
    for (int v = 0; v < PDX_EXPLICIT_FORMS; v++) {
        launch_explicit_any(c, prec, v, c.U[p], c.U[(p + 2) % 3], scratch, c.nbcv, nscratch, 0, 1);  // warm
        PDX_HIP(hipEventRecord(e0, c.stream));
        for (int r = 0; r < 6; r++)
            launch_explicit_any(c, prec, v, c.U[p], c.U[(p + 2) % 3], scratch, c.nbcv, nscratch, 0, r);
        PDX_HIP(hipEventRecord(e1, c.stream));
        PDX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        PDX_HIP(hipEventElapsedTime(&ms, e0, e1));
        c.tune_ms[prec][v] = ms / 6;
        if (ms < best) { best = ms; bestv = v; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(scratch);
    hipFree(nscratch);
    c.tune_explicit[prec] = bestv;
    return PDX_OK;
}

constexpr int kExplicitPeriod = 6;  // state rotation (3) x NBC buffer parity (2)

// one step on the device state: phase p = index of u_n buffer, s = step index
static void explicit_step(Ctx &c, int prec, int startup) {
    const int p = c.step_phase % 3;
    double *Un = c.U[p], *Unm1 = c.U[(p + 2) % 3], *Unp1 = c.U[(p + 1) % 3];
    const long long s = startup ? 0 : c.steps_done + 1;
    double *nbc_cur = c.nbcv + (s % 2) * std::max(c.n_nbc, 1);
    double *nbc_next = c.nbcv + ((s + 1) % 2) * std::max(c.n_nbc, 1);
    launch_explicit_any(c, prec, std::max(c.tune_explicit[prec == PDX_FP32 ? 1 : 0], 0), Un, Unm1, Unp1, nbc_cur,
                        nbc_next, startup, s);
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

// relative residual of the consistent-mass solves M a = F (the reference factorises M with
// PARDISO every step, calMatrixOperations.cpp:378-498); M is SPD and well conditioned
constexpr double kMassTol = 1e-13;

// Consistent-mass step (pdfemDynamicSolver_CSRformat with cb_lumpedMass off, pdsolve.cpp:6003-6057):
// F = F_ext - K u_n (force-only kernel of form 0), M a = F by PCG with node-block Jacobi
// warm-started from the previous acceleration, then the update.
static int explicit_step_nlm(Ctx &c, int prec, int startup, std::string &err) {
    const int p = c.step_phase % 3;
    double *Un = c.U[p], *Unm1 = c.U[(p + 2) % 3], *Unp1 = c.U[(p + 1) % 3];
    const long long s = startup ? 0 : c.steps_done + 1;
    double *nbc_cur = c.nbcv + (s % 2) * std::max(c.n_nbc, 1);
    double *nbc_next = c.nbcv + ((s + 1) % 2) * std::max(c.n_nbc, 1);
    if (prec == PDX_FP32) {
        if (c.dim == 3)
            launch_explicit_v<3, float, 8, false, false, false, false, false, true>(c, c.K32, Un, Unm1, c.nlm_F,
                                                                                    nbc_cur, nbc_next, startup);
        else
            launch_explicit_v<2, float, 8, false, false, false, false, false, true>(c, c.K32, Un, Unm1, c.nlm_F,
                                                                                    nbc_cur, nbc_next, startup);
    } else {
        if (c.dim == 3)
            launch_explicit_v<3, double, 8, false, false, false, false, false, true>(c, c.K, Un, Unm1, c.nlm_F,
                                                                                     nbc_cur, nbc_next, startup);
        else
            launch_explicit_v<2, double, 8, false, false, false, false, false, true>(c, c.K, Un, Unm1, c.nlm_F,
                                                                                     nbc_cur, nbc_next, startup);
    }
    PDX_HIP(hipGetLastError());
    pdx_solve_opts o{};
    o.krylov = PDX_KRYLOV_CG;
    o.precond = PDX_PREC_BLOCK_JACOBI;
    o.precision = PDX_FP64;
    o.tol = kMassTol;
    o.max_iter = 5000;
    o.warm_start = 1;
    pdx_solve_stats st{};
    int rc = dev_solve_mat(c, c.Mk, nullptr, o, c.nlm_F, c.nlm_a, st, err);
    c.nlm_iters += st.iterations;
    if (rc) return rc;
    const int n = c.nn * c.dim;
    k_nlm_update<<<(n + 255) / 256, 256, 0, c.stream>>>(n, c.dim, c.eq, c.presc_set, Un, Unm1, c.nlm_a,
                                                         c.d_ebc_vel, c.dt, startup, Unp1);
    PDX_HIP(hipGetLastError());
    c.step_phase = (p + 1) % 3;
    c.steps_done = s;
    return PDX_OK;
}

// K (and the FP32 / row-local / mirrored copies the step reads) for the current bond states;
// the consistent mass once
static int explicit_operators(Ctx &c, int prec, bool lumped, std::string &err) {
    if (!c.K_valid) {
        int e = dev_assemble_K(c, err);
        if (e) return e;
    }
    if (prec == PDX_FP32) {
        int e = ensure_K32(c, err);
        if (e) return e;
    }
    if (int e = ensure_Ks(c, prec, err)) return e;
    if (int e = ensure_Kown(c, prec, err)) return e;
    if (lumped && !c.mlump_valid) {
        int e = dev_lumped_mass(c, err);
        if (e) return e;
    }
    if (!lumped) {
        if (!c.Mk_valid) {
            int e = dev_mass_on_K(c, err);
            if (e) return e;
        }
        const size_t n = (size_t)c.nn * c.dim;
        if (!c.nlm_F) {
            PDX_HIP(hipMalloc((void **)&c.nlm_F, n * 8));
            PDX_HIP(hipMalloc((void **)&c.nlm_a, n * 8));
            PDX_HIP(hipMemsetAsync(c.nlm_a, 0, n * 8, c.stream));
        }
    }
    return PDX_OK;
}

int explicit_prepare(Ctx &c, int prec, bool lumped, std::string &err) {
    if (int e = explicit_operators(c, prec, lumped, err)) return e;
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
    if (lumped) {
        explicit_step(c, prec, 1);
    } else if (int e = explicit_step_nlm(c, prec, 1, err)) {
        return e;
    }
    host_advance_bc(c);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

int dev_merge_dofvals(Ctx &c, const double *free_src, const double *presc_src, double *out) {
    const int n = c.nn * c.dim;
    if (n == 0) return PDX_OK;
    k_merge_dofvals<<<(n + 255) / 256, 256, 0, c.stream>>>(n, c.dim, c.eq, free_src, presc_src, out);
    return hipGetLastError() == hipSuccess ? PDX_OK : PDX_ERR_HIP;
}

// DOF values of the last stored step into c.dofs_tmp (device), the input of the nodal stresses
int dev_explicit_dofvals(Ctx &c, const double **out, std::string &err) {
    if (c.steps_done < 0) { err = "no explicit state yet"; return PDX_ERR_ARG; }
    const int n = c.nn * c.dim;
    if (!c.dofs_tmp) PDX_HIP(hipMalloc((void **)&c.dofs_tmp, (size_t)std::max(n, 1) * 8));
    const int p = c.step_phase % 3;
    k_merge_dofvals<<<(n + 255) / 256, 256, 0, c.stream>>>(n, c.dim, c.eq, c.U[p], c.U[(p + 2) % 3], c.dofs_tmp);
    PDX_HIP(hipGetLastError());
    *out = c.dofs_tmp;
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

// Capture the 6-step graph of precision prec (three rotating state vectors, two NBC buffers:
// period 6) for replays that start at phase 0 with an even step index.  Runs before any timed
// region; the capture only records launches, so the state is not advanced.  Every exit after
// hipStreamBeginCapture ends the capture, so the stream never stays in capture mode.
static int capture_explicit_graph(Ctx &c, int prec, std::string &err) {
    const int gi = prec == PDX_FP32 ? 1 : 0;
    if (c.graph[gi]) return PDX_OK;
    const int sp = c.step_phase;
    const long long sd = c.steps_done;
    c.step_phase = 0;
    c.steps_done = 1;  // the first captured step has index 2 (even)
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int i = 0; i < kExplicitPeriod; i++) explicit_step(c, prec, 0);
        const hipError_t launch = hipGetLastError();
        e = hipStreamEndCapture(c.stream, &g);
        if (e == hipSuccess) e = launch;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&c.graph[gi], g, nullptr, nullptr, 0);
    if (g) hipGraphDestroy(g);
    c.step_phase = sp;
    c.steps_done = sd;
    if (e != hipSuccess) {
        c.graph[gi] = nullptr;
        err = std::string("explicit step graph capture: ") + hipGetErrorString(e);
        return PDX_ERR_HIP;
    }
    return PDX_OK;
}

// nsteps explicit steps (after the start-up step on the first call).  Lumped mass without a
// per-step callback: graph replays of the captured 6-step period; otherwise one step at a time,
// after_step (may be null) running after every stored step, the start-up step included, and K
// re-assembled before a step when a callback has invalidated it (broken bonds).
int dev_explicit_steps(Ctx &c, int nsteps, int prec, int use_graph, bool lumped,
                       const std::function<int(std::string &)> &after_step, std::string &err) {
    const bool startup = c.steps_done < 0;
    if (nsteps == 0 && !startup) return PDX_OK;  // nothing to run: the state is read as it is
    int e = explicit_prepare(c, prec, lumped, err);
    if (e) return e;
    if (startup && after_step && (e = after_step(err))) return e;
    if (nsteps == 0) return PDX_OK;  // the start-up step only
    if ((e = explicit_tune_if_needed(c, prec, err))) return e;
    const int period = kExplicitPeriod;
    const bool replay = lumped && !after_step && use_graph && nsteps >= period;
    // the graph is captured and instantiated before the timed region
    if (replay && (e = capture_explicit_graph(c, prec, err))) return e;
    hipEvent_t e0, e1;
    PDX_HIP(hipEventCreate(&e0));
    PDX_HIP(hipEventCreate(&e1));
    int done = 0;
    PDX_HIP(hipEventRecord(e0, c.stream));
    if (replay) {
        // align to the buffer / NBC rotation period, then replay the captured 6-step graph
        while ((c.step_phase % 3 != 0 || (c.steps_done + 1) % 2 != 0) && done < nsteps) {
            explicit_step(c, prec, 0);
            done++;
        }
        const int gi = prec == PDX_FP32 ? 1 : 0;
        while (nsteps - done >= period) {
            PDX_HIP(hipGraphLaunch(c.graph[gi], c.stream));
            c.steps_done += period;
            done += period;
        }
    }
    int rc = PDX_OK;
    while (done < nsteps) {
        if (after_step && (rc = explicit_operators(c, prec, lumped, err))) break;
        if (lumped) explicit_step(c, prec, 0);
        else if ((rc = explicit_step_nlm(c, prec, 0, err))) break;
        done++;
        if (after_step) {
            host_advance_bc(c);
            if ((rc = after_step(err))) break;
        }
    }
    PDX_HIP(hipEventRecord(e1, c.stream));
    PDX_HIP(hipEventSynchronize(e1));
    float ms = 0;
    PDX_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c.last_total_ms = ms;
    c.last_step_ms = done ? ms / done : 0;
    if (!after_step)
        for (int i = 0; i < done; i++) host_advance_bc(c);
    if (rc) return rc;
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


int to_f32(Ctx &c, const double *K64, float *&K32, std::string &err) {
    size_t n = (size_t)c.dd * c.nblk;
    if (!K32) PDX_HIP(hipMalloc((void **)&K32, std::max<size_t>(n, 1) * sizeof(float)));
    k_to_f32<<<4096, 256, 0, c.stream>>>(n, K64, K32);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}


}  // namespace pdx
