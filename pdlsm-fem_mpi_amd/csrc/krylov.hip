// Preconditioned BiCGStab on the node-block stiffness (the iterative stand-in
// for the reference's PARDISO solves, calMatrixOperations.cpp:378-498), with
//   * device-resident scalars: reducing kernels write per-block partials that
//     the consumer kernels re-reduce in a fixed order (bitwise deterministic,
//     no host synchronisation inside an iteration);
//   * FP64 or FP32 inner iterations, the latter wrapped in FP64 iterative
//     refinement on the true residual;
//   * preconditioners: none / Jacobi / node-block Jacobi (fused into the
//     vector updates), and a two-level hybrid preconditioner
//         z = Q v + S (v - A Q v),   Q = P (P^T A P)^-1 P^T
//     whose coarse space P holds the rigid-body modes of geometric node
//     aggregates (cells of the mesh split into K-connected components, so a
//     crack separates them) and whose smoother S is node-block Jacobi or the
//     node-block smoother predicted by the graph network (learned.hip);
//   * node-block ILU(0) in multicolour order (ilu0.hip);
//   * an optional fallback chain (learned / two-level -> ILU(0) -> node-block
//     Jacobi -> FP64 node-block Jacobi) when a solve stops short or a
//     preconditioner is singular.
//     The coarse matrix is inverted once per operator on the device (blocked
//     Gauss-Jordan, FP64); each application restricts inside the vector
//     update kernels, multiplies by the dense inverse, and prolongates with
//     the precomputed A P.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "device.hpp"
#include "reduce.hpp"
#include "rowblock.hpp"

namespace pdx {

int ensure_K32(Ctx &c, std::string &err);

// ------------------------------------------------------------------------
// SpMV with fused dot partials
// ------------------------------------------------------------------------
// y = mask(A x); part0 = (w0, y), part1 = (w1, y).  With RS (the d x d block
// row sums of the FP64 operator, stored in T) the product is formed as
//   sum_b A_ab (x_b - x_a) + RS_a x_a,
// algebraically identical; the block rows of a stiffness sum to ~0 (rigid
// translations), so in FP32 this avoids the cancellation of the diagonal
// against the neighbours for smooth vectors.
// With REL the block columns come from the row's column pattern class (cls, cls_off, cls_rel)
// instead of adj.  With SYM, K holds only the own blocks of the mirrored storage and src[b] is
// block b's own index, or ~(its partner's) for a lower block read as its partner's transpose (the
// same values, so the same products in the same order).
// One row per G-lane group: lane < D of the group writes y and adds its share of the dots.
template <int D, typename T, typename V, int G, bool DIFF, bool REL, bool SYM>
__device__ inline void spmv_dot_row(int r, const int *adj_off, const int *adj, const int *cls, const int *cls_off,
                                    const int *cls_rel, const int *src, const T *K, const T *RS, const V *x, V *y,
                                    const int *eq, const V *w0, const V *w1, double &d0, double &d1) {
    const int lane = threadIdx.x % G;
    T acc[D];
    const int *rel = REL ? cls_rel + cls_off[cls[r]] : nullptr;
    row_matvec<D, T, V, G, DIFF, false, false, REL, SYM>(r, lane, adj_off, adj, K, x, acc, rel, src);
    if (lane < D) {
        T v = pick<D, T>(acc, lane);
        if (DIFF) {
#pragma unroll
            for (int j = 0; j < D; j++) v += RS[(size_t)r * D * D + lane * D + j] * T(x[(size_t)r * D + j]);
        }
        const size_t o = (size_t)r * D + lane;
        const V yv = eq[r * 3 + lane] >= 0 ? V(v) : V(0);
        y[o] = yv;
        if (w0) d0 += double(w0[o]) * double(yv);
        if (w1) d1 += double(w1[o]) * double(yv);
    }
}

// Launch shape of k_spmv_dot: 256-thread blocks, the short rows 32 to a block (8 lanes each), then
// the long rows 4 to a block (a wavefront each); nbs short-row blocks of nbt, the grid a multiple
// of 8 (one partial per block)
struct SpmvShape {
    int nbs = 0, nbt = 1, grid = 8;
};
inline SpmvShape spmv_shape(const Ctx &c) {
    SpmvShape s;
    s.nbs = (c.n_short * 8 + 255) / 256;
    const int nbl = (c.n_long * 64 + 255) / 256;
    s.nbt = std::max(s.nbs + nbl, 1);
    s.grid = 8 * ((s.nbs + 7) / 8 + (s.nbt - s.nbs + 7) / 8);
    return s;
}

// The hardware deals blocks to the 8 XCDs round-robin.  Block b works on logical block
// x * ceil(nbs / 8) + q (short rows) or nbs + x * ceil((nbt - nbs) / 8) + q - ceil(nbs / 8) (long
// rows), x = b % 8, q = b / 8: each XCD (and its L2) gets one contiguous eighth of the short-row
// blocks and one of the long-row blocks, so its gathers of x stay local and the long rows are
// spread over all XCDs.  With rev, q counts down: the product visits the rows last to first, so
// it starts on the blocks the previous product read last, which may still be in the memory-side
// cache.  Each block writes its dot partials to bp0 / bp1 at slot q * 8 + x (0 for a block with
// no rows), the same slot in either direction; k_fold_parts sums them in slot order into the
// partials the vector kernels read, so both directions give the same bits.
template <int D, typename T, typename V, bool REL, bool SYM>
__global__ void __launch_bounds__(256) k_spmv_dot(const int *rows_s, int ns, int nbs, const int *rows_l, int nl,
                                                  int nbt, const int *adj_off, const int *adj, const int *cls,
                                                  const int *cls_off, const int *cls_rel, const int *src,
                                                  const T *K, const T *RS, const V *x, V *y, const int *eq,
                                                  const V *w0, double *bp0, const V *w1, double *bp1, int rev) {
    __shared__ double sm[17];
    double d0 = 0.0, d1 = 0.0;
    const int per = gridDim.x / 8, q0 = blockIdx.x / 8, xcd = blockIdx.x % 8, ps = (nbs + 7) / 8;
    const int q = rev ? per - 1 - q0 : q0;
    const int slot = q * 8 + xcd;  // the partials' slot follows the logical block, not the dispatch
    int lb = -1;
    if (q < ps) {
        if (xcd * ps + q < nbs) lb = xcd * ps + q;
    } else {
        const int l = nbs + xcd * (per - ps) + (q - ps);
        if (l < nbt) lb = l;
    }
    if (lb >= 0 && lb < nbs) {
        const int gid = lb * 32 + (int)threadIdx.x / 8;
        if (gid < ns) {
            if (RS)
                spmv_dot_row<D, T, V, 8, true, REL, SYM>(rows_s[gid], adj_off, adj, cls, cls_off, cls_rel, src, K, RS,
                                                         x, y, eq, w0, w1, d0, d1);
            else
                spmv_dot_row<D, T, V, 8, false, REL, SYM>(rows_s[gid], adj_off, adj, cls, cls_off, cls_rel, src, K, RS,
                                                          x, y, eq, w0, w1, d0, d1);
        }
    } else if (lb >= nbs) {
        const int gid = (lb - nbs) * 4 + (int)threadIdx.x / 64;
        if (gid < nl) {
            if (RS)
                spmv_dot_row<D, T, V, 64, true, REL, SYM>(rows_l[gid], adj_off, adj, cls, cls_off, cls_rel, src, K,
                                                          RS, x, y, eq, w0, w1, d0, d1);
            else
                spmv_dot_row<D, T, V, 64, false, REL, SYM>(rows_l[gid], adj_off, adj, cls, cls_off, cls_rel, src, K,
                                                           RS, x, y, eq, w0, w1, d0, d1);
        }
    }
    if (bp0) {
        d0 = block_sum(d0, sm);
        if (threadIdx.x == 0) bp0[slot] = d0;
    }
    if (bp1) {
        d1 = block_sum(d1, sm);
        if (threadIdx.x == 0) bp1[slot] = d1;
    }
}

// part[i] = the sum of bp[j] over the i-th of NPART contiguous ranges of [0, nbp): one wavefront per
// range, the lanes' loads in flight together, then a fixed shuffle tree (the same order every call)
__global__ void __launch_bounds__(64) k_fold_parts(int nbp, const double *bp0, double *p0, const double *bp1,
                                                   double *p1) {
    const int i = blockIdx.x;
    const int per = (nbp + NPART - 1) / NPART, j0 = min(nbp, i * per), j1 = min(nbp, j0 + per);
    double s0 = 0.0, s1 = 0.0;
    for (int j = j0 + (int)threadIdx.x; j < j1; j += 64) {
        if (bp0) s0 += bp0[j];
        if (bp1) s1 += bp1[j];
    }
    for (int s = 32; s > 0; s >>= 1) {
        s0 += __shfl_xor(s0, s, 64);
        s1 += __shfl_xor(s1, s, 64);
    }
    if (threadIdx.x == 0) {
        if (p0) p0[i] = s0;
        if (p1) p1[i] = s1;
    }
}

// Persistent form: rows [lb * per, (lb + 1) * per) of the list go to logical block lb, and the
// G-lane groups of the block stride through them (REL, SYM and RS as in spmv_dot_row).
template <int D, typename T, typename V, int G, bool DIFF, bool REL, bool SYM>
__device__ inline void spmv_dot_rows_p(const int *rows, int nrows, int lb, int nblocks, const int *adj_off,
                                     const int *adj, const int *cls, const int *cls_off, const int *cls_rel,
                                     const int *src, const T *K, const T *RS, const V *x, V *y, const int *eq,
                                     const V *w0, const V *w1, double &d0, double &d1) {
    const int lane = threadIdx.x % G;
    const int per = (nrows + nblocks - 1) / nblocks, g0 = lb * per, g1 = min(nrows, g0 + per);
    for (int gid = g0 + (int)threadIdx.x / G; gid < g1; gid += (int)blockDim.x / G) {
        const int r = rows[gid];
        T acc[D];
        const int *rel = REL ? cls_rel + cls_off[cls[r]] : nullptr;
        row_matvec<D, T, V, G, DIFF, false, false, REL, SYM>(r, lane, adj_off, adj, K, x, acc, rel, src);
        if (lane < D) {
            T v = pick<D, T>(acc, lane);
            if (DIFF) {
#pragma unroll
                for (int j = 0; j < D; j++) v += RS[(size_t)r * D * D + lane * D + j] * T(x[(size_t)r * D + j]);
            }
            const size_t o = (size_t)r * D + lane;
            const V yv = eq[r * 3 + lane] >= 0 ? V(v) : V(0);
            y[o] = yv;
            if (w0) d0 += double(w0[o]) * double(yv);
            if (w1) d1 += double(w1[o]) * double(yv);
        }
    }
}

// Persistent form for the 1- and 2-DOF operators: RED_PARTS blocks of SPMV_THREADS threads, logical
// block (b % 8) * (grid / 8) + b / 8 (each XCD, and its L2, gets one contiguous eighth of the rows)
// walking its share of the short, then the long rows, one partial per block straight into the NPART
// slots (no fold: their iterations are short enough that a second launch per product shows).
template <int D, typename T, typename V, bool REL, bool SYM>
__global__ void __launch_bounds__(SPMV_THREADS) k_spmv_dot_p(const int *rows_s, int ns, const int *rows_l, int nl,
                                                  const int *adj_off, const int *adj, const int *cls,
                                                  const int *cls_off, const int *cls_rel, const int *src,
                                                  const T *K, const T *RS, const V *x, V *y, const int *eq,
                                                  const V *w0, double *part0, const V *w1, double *part1) {
    __shared__ double sm[17];
    double d0 = 0.0, d1 = 0.0;
    const int nb = gridDim.x;
    const int lb = nb % 8 == 0 ? (int)(blockIdx.x % 8) * (nb / 8) + (int)blockIdx.x / 8 : (int)blockIdx.x;
    if (RS) {
        spmv_dot_rows_p<D, T, V, 8, true, REL, SYM>(rows_s, ns, lb, nb, adj_off, adj, cls, cls_off, cls_rel, src, K,
                                                  RS, x, y, eq, w0, w1, d0, d1);
        spmv_dot_rows_p<D, T, V, 64, true, REL, SYM>(rows_l, nl, lb, nb, adj_off, adj, cls, cls_off, cls_rel, src, K,
                                                   RS, x, y, eq, w0, w1, d0, d1);
    } else {
        spmv_dot_rows_p<D, T, V, 8, false, REL, SYM>(rows_s, ns, lb, nb, adj_off, adj, cls, cls_off, cls_rel, src, K,
                                                   RS, x, y, eq, w0, w1, d0, d1);
        spmv_dot_rows_p<D, T, V, 64, false, REL, SYM>(rows_l, nl, lb, nb, adj_off, adj, cls, cls_off, cls_rel, src, K,
                                                    RS, x, y, eq, w0, w1, d0, d1);
    }
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


// block row sums RS_a = sum_b A_ab (FP64), stored as T
template <int D, typename T>
__global__ void k_block_rowsum(int nn, const int *adj_off, const double *K, T *RS) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double s[D * D];
#pragma unroll
    for (int e = 0; e < D * D; e++) s[e] = 0.0;
    for (int b = adj_off[a]; b < adj_off[a + 1]; b++)
#pragma unroll
        for (int e = 0; e < D * D; e++) s[e] += K[(size_t)b * D * D + e];
#pragma unroll
    for (int e = 0; e < D * D; e++) RS[(size_t)a * D * D + e] = T(s[e]);
}

// ------------------------------------------------------------------------
// BiCGStab vector kernels (V = double or float)
// ------------------------------------------------------------------------
// z = Minv r for node a (Minv: [nn][D*D])
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

// beta of the p update; a vanishing rho_old or omega (breakdown, or exact
// convergence in low precision) restarts the search direction with p = r
__device__ inline void bicg_beta(const double *rho_part, const double *sc, int first, double *sm, double &beta,
                                 double &omg) {
    const double rho = sum_partials(rho_part, NPART, sm);
    beta = 0.0;
    omg = 0.0;
    if (!first && sc[SC_RHO_OLD] != 0.0 && sc[SC_OMEGA] != 0.0) {
        beta = (rho / sc[SC_RHO_OLD]) * (sc[SC_ALPHA] / sc[SC_OMEGA]);
        omg = sc[SC_OMEGA];
    }
    if (!isfinite(beta)) beta = 0.0;
}

__device__ inline double bicg_alpha(const double *rho_part, const double *rv_part, double *sc, double *sm) {
    const double rho = sum_partials(rho_part, NPART, sm);
    const double rv = sum_partials(rv_part, NPART, sm);
    double alpha = rv != 0.0 ? rho / rv : 0.0;
    if (!isfinite(alpha)) alpha = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_ALPHA] = alpha;
    return alpha;
}

template <int D, typename V>
__global__ void __launch_bounds__(256) k_bicg_p(int nn, const V *r, V *p, const V *v, V *ph, const V *Minv,
                                                const double *rho_part, double *sc, int first) {
    __shared__ double sm[17];
    double beta, omg;
    bicg_beta(rho_part, sc, first, sm, beta, omg);
    const V bv = V(beta), om = V(omg);
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < nn; a += gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            const size_t o = (size_t)a * D + i;
            p[o] = first ? r[o] : V(r[o] + bv * (p[o] - om * v[o]));
        }
        if (Minv) apply_blk<D, V>(Minv, p, ph, a);
    }
}

template <int D, typename V>
__global__ void __launch_bounds__(256) k_bicg_s(int nn, const V *r, const V *v, V *s, V *sh, const V *Minv,
                                                const double *rho_part, const double *rv_part, double *sc) {
    __shared__ double sm[17];
    const V al = V(bicg_alpha(rho_part, rv_part, sc, sm));
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < nn; a += gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            const size_t o = (size_t)a * D + i;
            s[o] = r[o] - al * v[o];
        }
        if (Minv) apply_blk<D, V>(Minv, s, sh, a);
    }
}

template <int D, typename V>
__global__ void __launch_bounds__(256) k_bicg_x(int nn, V *x, V *r, const V *rh, const V *ph, const V *sh,
                                                const V *s, const V *t, const double *rho_part,
                                                const double *ts_part, const double *tt_part, double *sc,
                                                double *rho_next, double *rr_part) {
    __shared__ double sm[17];
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

// ---- preconditioned conjugate gradients (symmetric operators) ----
// x += alpha p, r -= alpha q, z = Minv r; partials rho_next = (r, z), rr = (r, r)
template <int D, typename V>
__global__ void __launch_bounds__(256) k_cg_xr(int nn, V *x, V *r, V *z, const V *p, const V *q, const V *Minv,
                                               const double *rho_part, const double *pq_part, double *rho_next,
                                               double *rr_part) {
    __shared__ double sm[17];
    const double rho = sum_partials(rho_part, NPART, sm);
    const double pq = sum_partials(pq_part, NPART, sm);
    double alpha = pq != 0.0 ? rho / pq : 0.0;
    if (!isfinite(alpha)) alpha = 0.0;
    const V al = V(alpha);
    double d0 = 0.0, d1 = 0.0;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < nn; a += gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < D; i++) {
            const size_t o = (size_t)a * D + i;
            x[o] = x[o] + al * p[o];
            r[o] = r[o] - al * q[o];
        }
        apply_blk<D, V>(Minv, r, z, a);
#pragma unroll
        for (int i = 0; i < D; i++) {
            const size_t o = (size_t)a * D + i;
            d0 += double(r[o]) * double(z[o]);
            d1 += double(r[o]) * double(r[o]);
        }
    }
    d0 = block_sum(d0, sm);
    d1 = block_sum(d1, sm);
    if (threadIdx.x == 0) {
        rho_next[blockIdx.x] = d0;
        rr_part[blockIdx.x] = d1;
    }
}

// p = z + beta p, beta = rho_new / rho_old (p = z on the first iteration)
template <typename V>
__global__ void __launch_bounds__(256) k_cg_p(size_t n, V *p, const V *z, const double *rho_new_part,
                                              const double *rho_old_part, int first) {
    __shared__ double sm[17];
    const double rn = sum_partials(rho_new_part, NPART, sm);
    const double ro = sum_partials(rho_old_part, NPART, sm);
    double beta = (!first && ro != 0.0) ? rn / ro : 0.0;
    if (!isfinite(beta)) beta = 0.0;
    const V bv = V(beta);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = first ? z[i] : V(z[i] + bv * p[i]);
}

// ---- two-level variants: vector update in aggregate order + restriction ----
// One wave per aggregate: updates the vector on the aggregate's nodes and
// writes rc = P_I^T vec (NM modes) with a fixed-order wave reduction.
template <int D, int NM>
__device__ inline void wave_restrict(double acc[NM], int I, int lane, double *rc) {
#pragma unroll
    for (int m = 0; m < NM; m++) {
        double v = acc[m];
        for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
        acc[m] = v;
    }
    if (lane == 0)
#pragma unroll
        for (int m = 0; m < NM; m++) rc[(size_t)I * NM + m] = acc[m];
}

template <int D, int NM, typename V>
__global__ void __launch_bounds__(256) k_bicg_p_agg(int na, const int *agg_off, const int *agg_node, const V *r,
                                                    V *p, const V *v, const V *Pn, const double *rho_part,
                                                    double *sc, int first, double *rc) {
    __shared__ double sm[17];
    double beta, omg;
    bicg_beta(rho_part, sc, first, sm, beta, omg);
    const V bv = V(beta), om = V(omg);
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int I = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; I < na; I += nw) {
        double acc[NM];
#pragma unroll
        for (int m = 0; m < NM; m++) acc[m] = 0.0;
        for (int q = agg_off[I] + lane; q < agg_off[I + 1]; q += 64) {
            const int a = agg_node[q];
#pragma unroll
            for (int i = 0; i < D; i++) {
                const size_t o = (size_t)a * D + i;
                const V pv = first ? r[o] : V(r[o] + bv * (p[o] - om * v[o]));
                p[o] = pv;
#pragma unroll
                for (int m = 0; m < NM; m++) acc[m] += double(Pn[o * NM + m]) * double(pv);
            }
        }
        wave_restrict<D, NM>(acc, I, lane, rc);
    }
}

template <int D, int NM, typename V>
__global__ void __launch_bounds__(256) k_bicg_s_agg(int na, const int *agg_off, const int *agg_node, const V *r,
                                                    const V *v, V *s, const V *Pn, const double *rho_part,
                                                    const double *rv_part, double *sc, double *rc) {
    __shared__ double sm[17];
    const V al = V(bicg_alpha(rho_part, rv_part, sc, sm));
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int I = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; I < na; I += nw) {
        double acc[NM];
#pragma unroll
        for (int m = 0; m < NM; m++) acc[m] = 0.0;
        for (int q = agg_off[I] + lane; q < agg_off[I + 1]; q += 64) {
            const int a = agg_node[q];
#pragma unroll
            for (int i = 0; i < D; i++) {
                const size_t o = (size_t)a * D + i;
                const V sv = r[o] - al * v[o];
                s[o] = sv;
#pragma unroll
                for (int m = 0; m < NM; m++) acc[m] += double(Pn[o * NM + m]) * double(sv);
            }
        }
        wave_restrict<D, NM>(acc, I, lane, rc);
    }
}

// yc = Ainv rc (dense, nc x nc, row stride ld, a multiple of 64); one wave per row, 16-byte
// loads of the row (ld keeps them inside the matrix), FP64 sums
template <typename V>
__global__ void __launch_bounds__(256) k_coarse_mv(int nc, int ld, const V *Ainv, const double *rc, double *yc) {
    constexpr int VEC = 16 / sizeof(V);
    typedef V vec_t __attribute__((ext_vector_type(VEC)));
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < nc; row += nw) {
        const V *a = Ainv + (size_t)row * ld;
        double s = 0.0;
        for (int j0 = lane * VEC; j0 < nc; j0 += 64 * VEC) {
            const vec_t v = *(const vec_t *)(a + j0);
#pragma unroll
            for (int u = 0; u < VEC; u++)
                if (j0 + u < nc) s += double(v[u]) * rc[j0 + u];
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) yc[row] = s;
    }
}

// z = P yc + S (vin - (A P) yc), node by node (S: node-block smoother)
template <int D, int NM, typename V>
__global__ void __launch_bounds__(256) k_prolong(int nn, const V *Pn, const int *node_agg, const int *kp_off,
                                                 const int *kp_agg, const V *kp, const V *Minv, const V *vin,
                                                 const double *yc, V *z) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    const int I = node_agg[a];
    double y[D], w[D];
#pragma unroll
    for (int i = 0; i < D; i++) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < NM; m++) s += double(Pn[((size_t)a * D + i) * NM + m]) * yc[(size_t)I * NM + m];
        y[i] = s;
        w[i] = double(vin[(size_t)a * D + i]);
    }
    for (int q = kp_off[a]; q < kp_off[a + 1]; q++) {
        const int J = kp_agg[q];
#pragma unroll
        for (int i = 0; i < D; i++) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < NM; m++) s += double(kp[((size_t)q * D + i) * NM + m]) * yc[(size_t)J * NM + m];
            w[i] -= s;
        }
    }
#pragma unroll
    for (int i = 0; i < D; i++) {
        double s = y[i];
#pragma unroll
        for (int j = 0; j < D; j++) s += double(Minv[(size_t)a * D * D + i * D + j]) * w[j];
        z[(size_t)a * D + i] = V(s);
    }
}

// r = b - y (y masked already), partial (r, r); optional copy to rh
template <typename V, typename B>
__global__ void __launch_bounds__(256) k_resid(size_t n, const B *b, const double *y, V *r, V *rh, double *part) {
    __shared__ double sm[17];
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

__global__ void k_add_f2d(size_t n, const float *d, double *x) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = x[i] + double(d[i]);
}

__global__ void k_fill_d(size_t n, double *p, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// node-block Jacobi / Jacobi inverse from the diagonal blocks; prescribed rows/cols removed
template <int D, typename V>
__global__ void k_blockjacobi(int nn, const int *adj_off, const int *adj, const double *K, const int *eq,
                              int diag_only, V *Minv, int *status) {
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

// ------------------------------------------------------------------------
// coarse space: A P, the Galerkin matrix P^T A P and its dense inverse
// ------------------------------------------------------------------------
// A P on every node row: kp[q] (D x NM) for the q-th aggregate touched by row a
// (rows of prescribed DOFs zero, since the operator is masked)
template <int D, int NM>
__global__ void k_coarse_ap(int nn, const int *adj_off, const int *adj, const double *K, const double *Pn,
                            const int *blk_slot, const int *kp_off, const int *eq, double *kp) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    for (int q = kp_off[a]; q < kp_off[a + 1]; q++) {
        const int slot = q - kp_off[a];
        double s[D * NM];
#pragma unroll
        for (int e = 0; e < D * NM; e++) s[e] = 0.0;
        for (int b = adj_off[a]; b < adj_off[a + 1]; b++) {
            if (blk_slot[b] != slot) continue;
            const int c = adj[b];
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) {
                    const double k = K[(size_t)b * D * D + i * D + j];
#pragma unroll
                    for (int m = 0; m < NM; m++) s[i * NM + m] += k * Pn[((size_t)c * D + j) * NM + m];
                }
        }
#pragma unroll
        for (int i = 0; i < D; i++) {
            const bool fr = eq[a * 3 + i] >= 0;
#pragma unroll
            for (int m = 0; m < NM; m++) kp[((size_t)q * D + i) * NM + m] = fr ? s[i * NM + m] : 0.0;
        }
    }
}

// dense Galerkin matrix: one thread per coarse row (I, m); Ac[row][J*NM+m'] += P_a^T (A P)_aJ
template <int D, int NM>
__global__ void k_coarse_galerkin(int na, const int *agg_off, const int *agg_node, const double *Pn,
                                  const int *kp_off, const int *kp_agg, const double *kp, int ld, double *Ac) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= na * NM) return;
    const int I = row / NM, m = row % NM;
    double *out = Ac + (size_t)row * ld;
    for (int q0 = agg_off[I]; q0 < agg_off[I + 1]; q0++) {
        const int a = agg_node[q0];
        for (int q = kp_off[a]; q < kp_off[a + 1]; q++) {
            const int J = kp_agg[q];
#pragma unroll
            for (int m2 = 0; m2 < NM; m2++) {
                double s = 0.0;
#pragma unroll
                for (int i = 0; i < D; i++) s += Pn[((size_t)a * D + i) * NM + m] * kp[((size_t)q * D + i) * NM + m2];
                out[(size_t)J * NM + m2] += s;
            }
        }
    }
}

__global__ void k_diag_copy(int n, int ld, const double *A, double *d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = A[(size_t)i * ld + i];
}

// ---- blocked Gauss-Jordan inversion (in place, FP64, no pivoting) --------
// The Galerkin matrix of a stiffness is positive (semi)definite in its
// symmetric part; a pivot below 1e-12 of the DOF's original diagonal marks a
// coarse mode without support (e.g. a rotation of a one-node aggregate): its
// row and column are removed (zero in the inverse).
constexpr int GJB = 64;

__global__ void __launch_bounds__(1024) k_gj_diag(int ld, int k, const double *A, const double *diag0, double *Dk) {
    __shared__ double S[GJB][GJB + 1];
    __shared__ double colq[GJB];
    const int t = threadIdx.x, base = k * GJB, nt = blockDim.x;
    for (int e = t; e < GJB * GJB; e += nt) S[e / GJB][e % GJB] = A[(size_t)(base + e / GJB) * ld + base + e % GJB];
    __syncthreads();
    for (int q = 0; q < GJB; q++) {
        const double piv = S[q][q];
        const double d0 = diag0[base + q];
        const bool dropq = !(fabs(piv) > 1e-12 * fabs(d0)) || d0 == 0.0;
        if (t < GJB) colq[t] = S[t][q];
        __syncthreads();
        if (dropq) {
            if (t < GJB) { S[q][t] = 0.0; S[t][q] = 0.0; }
            __syncthreads();
            continue;
        }
        const double ip = 1.0 / piv;
        if (t < GJB && t != q) S[q][t] *= ip;
        __syncthreads();
        for (int e = t; e < GJB * GJB; e += nt) {
            const int i = e / GJB, j = e % GJB;
            if (i == q) continue;
            if (j == q) S[i][j] = -colq[i] * ip;
            else S[i][j] -= colq[i] * S[q][j];
        }
        __syncthreads();
        if (t == 0) S[q][q] = ip;
        __syncthreads();
    }
    for (int e = t; e < GJB * GJB; e += nt) Dk[e] = S[e / GJB][e % GJB];
}

typedef double gj_d4 __attribute__((ext_vector_type(4)));

// C = alpha * A_blk * B_blk (+ C when acc): 64 x 64 tiles through LDS; each of the 4
// waves computes a 32 x 32 quadrant as 2 x 2 tiles of v_mfma_f64_16x16x4f64 (operand lane l:
// A[i = l%16][k = l/16], B[k = l/16][j = l%16]; result lane l item r: row l/16 + 4 r, col l%16)
__device__ inline void tile_mm(const double *A, int lda, const double *B, int ldb, double *C, int ldc, double alpha,
                               bool acc, double (*As)[GJB + 1], double (*Bs)[GJB + 1]) {
    const int t = threadIdx.x;
    for (int e = t; e < GJB * GJB; e += 256) {
        As[e / GJB][e % GJB] = A[(size_t)(e / GJB) * lda + e % GJB];
        Bs[e / GJB][e % GJB] = B[(size_t)(e / GJB) * ldb + e % GJB];
    }
    __syncthreads();
    const int w = t >> 6, l = t & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, li = l & 15, lk = l >> 4;
    gj_d4 c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    for (int k0 = 0; k0 < GJB; k0 += 4) {
        const int kk = k0 + lk;
        const double a0 = As[wr + li][kk], a1 = As[wr + 16 + li][kk];
        const double b0 = Bs[kk][wc + li], b1 = Bs[kk][wc + 16 + li];
        c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, c00, 0, 0, 0);
        c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, c01, 0, 0, 0);
        c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, c10, 0, 0, 0);
        c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, c11, 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int row0 = wr + lk + 4 * r, row1 = row0 + 16, col0 = wc + li, col1 = col0 + 16;
        double *o00 = C + (size_t)row0 * ldc + col0, *o01 = C + (size_t)row0 * ldc + col1;
        double *o10 = C + (size_t)row1 * ldc + col0, *o11 = C + (size_t)row1 * ldc + col1;
        *o00 = acc ? *o00 + alpha * c00[r] : alpha * c00[r];
        *o01 = acc ? *o01 + alpha * c01[r] : alpha * c01[r];
        *o10 = acc ? *o10 + alpha * c10[r] : alpha * c10[r];
        *o11 = acc ? *o11 + alpha * c11[r] : alpha * c11[r];
    }
}

// block row k: A_kj <- Dk A_kj (j != k)
__global__ void __launch_bounds__(256) k_gj_row(int ld, int nb, int k, double *A, const double *Dk, double *tmp) {
    __shared__ double As[GJB][GJB + 1], Bs[GJB][GJB + 1];
    const int j = blockIdx.x;
    if (j == k) return;
    double *Akj = A + (size_t)k * GJB * ld + (size_t)j * GJB;
    double *T = tmp + (size_t)j * GJB * GJB;
    tile_mm(Dk, GJB, Akj, ld, T, GJB, 1.0, false, As, Bs);
    __syncthreads();
    for (int e = threadIdx.x; e < GJB * GJB; e += 256) Akj[(size_t)(e / GJB) * ld + e % GJB] = T[e];
}

// A_ij -= A_ik A_kj for i, j != k
__global__ void __launch_bounds__(256) k_gj_update(int ld, int nb, int k, double *A) {
    __shared__ double As[GJB][GJB + 1], Bs[GJB][GJB + 1];
    const int i = blockIdx.x / nb, j = blockIdx.x % nb;
    if (i == k || j == k) return;
    tile_mm(A + (size_t)i * GJB * ld + (size_t)k * GJB, ld, A + (size_t)k * GJB * ld + (size_t)j * GJB, ld,
            A + (size_t)i * GJB * ld + (size_t)j * GJB, ld, -1.0, true, As, Bs);
}

// block column k: A_ik <- -A_ik Dk (i != k), A_kk <- Dk
__global__ void __launch_bounds__(256) k_gj_col(int ld, int nb, int k, double *A, const double *Dk, double *tmp) {
    __shared__ double As[GJB][GJB + 1], Bs[GJB][GJB + 1];
    const int i = blockIdx.x;
    double *Aik = A + (size_t)i * GJB * ld + (size_t)k * GJB;
    if (i == k) {
        for (int e = threadIdx.x; e < GJB * GJB; e += 256) Aik[(size_t)(e / GJB) * ld + e % GJB] = Dk[e];
        return;
    }
    double *T = tmp + (size_t)i * GJB * GJB;
    tile_mm(Aik, ld, Dk, GJB, T, GJB, -1.0, false, As, Bs);
    __syncthreads();
    for (int e = threadIdx.x; e < GJB * GJB; e += 256) Aik[(size_t)(e / GJB) * ld + e % GJB] = T[e];
}

// ------------------------------------------------------------------------
// host: coarse space set-up
// ------------------------------------------------------------------------
struct TwoLevel {
    int dim = 2, nm = 3, na = 0, nc = 0, ld = 0;
    const double *K = nullptr;  // operator this was built for
    long long gen = -1;
    int agg_size = 0;
    int *agg_off = nullptr, *agg_node = nullptr, *node_agg = nullptr, *kp_off = nullptr, *kp_agg = nullptr;
    double *Pn = nullptr, *kp = nullptr, *Ainv = nullptr;  // FP64
    float *Pn32 = nullptr, *kp32 = nullptr, *Ainv32 = nullptr;
    double setup_ms = 0;
    long long serial = next_build_serial();  // distinguishes rebuilds that reuse freed addresses
    std::vector<void *> owned;
    ~TwoLevel() {
        for (void *p : owned) hipFree(p);
    }
};

template <typename P>
static int tl_alloc(TwoLevel &t, P *&p, size_t count, std::string &err) {
    PDX_HIP(hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(P)));
    t.owned.push_back((void *)p);
    return PDX_OK;
}

template <int D>
__global__ void k_block_nonzero(long long nblk, const double *K, unsigned char *nz) {
    for (long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x; b < nblk; b += (long long)gridDim.x * blockDim.x) {
        bool any = false;
#pragma unroll
        for (int e = 0; e < D * D; e++) any |= K[b * D * D + e] != 0.0;
        nz[b] = any;
    }
}

// geometric aggregates split into connected components of the nonzero block graph
static void make_aggregates(const Model &m, int D, int agg, const std::vector<unsigned char> &nz,
                            std::vector<int> &node_agg, int &na) {
    const int nn = m.nn;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int a = 0; a < nn; a++)
        for (int i = 0; i < D; i++) {
            lo[i] = std::min(lo[i], m.x[(size_t)a * 3 + i]);
            hi[i] = std::max(hi[i], m.x[(size_t)a * 3 + i]);
        }
    double vol = 1.0;
    for (int i = 0; i < D; i++) vol *= std::max(hi[i] - lo[i], 1e-300);
    const double h0 = std::pow(vol / std::max(nn, 1), 1.0 / D);
    const double cell = agg * h0;
    std::vector<long long> key(nn);
    for (int a = 0; a < nn; a++) {
        long long k = 0;
        for (int i = 0; i < D; i++) {
            long long ci = (long long)std::floor((m.x[(size_t)a * 3 + i] - lo[i]) / cell * (1.0 - 1e-12));
            k = k * 1000003LL + ci;
        }
        key[a] = k;
    }
    node_agg.assign(nn, -1);
    na = 0;
    std::vector<int> stack;
    for (int s = 0; s < nn; s++) {
        if (node_agg[s] >= 0) continue;
        node_agg[s] = na;
        stack.assign(1, s);
        while (!stack.empty()) {
            const int a = stack.back();
            stack.pop_back();
            for (int b = m.adj_off[a]; b < m.adj_off[a + 1]; b++) {
                const int c = m.adj[b];
                if (node_agg[c] >= 0 || !nz[b] || key[c] != key[a]) continue;
                node_agg[c] = na;
                stack.push_back(c);
            }
        }
        na++;
    }
}

template <int D, int NM>
static int build_two_level(Ctx &c, const double *K64, int agg, TwoLevel &t, std::string &err) {
    auto t0 = std::chrono::steady_clock::now();
    const Model &m = *c.host;
    const int nn = c.nn;
    t.dim = D;
    t.nm = NM;
    t.agg_size = agg;
    // nonzero blocks (a crack leaves zero blocks between its faces)
    unsigned char *nz_d;
    PDX_HIP(hipMalloc((void **)&nz_d, std::max<long long>(c.nblk, 1)));
    k_block_nonzero<D><<<2048, 256, 0, c.stream>>>(c.nblk, K64, nz_d);
    std::vector<unsigned char> nz(c.nblk);
    PDX_HIP(hipMemcpyAsync(nz.data(), nz_d, c.nblk, hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    hipFree(nz_d);
    std::vector<int> node_agg;
    int na = 0;
    make_aggregates(m, D, agg, nz, node_agg, na);
    t.na = na;
    t.nc = na * NM;
    t.ld = (t.nc + GJB - 1) / GJB * GJB;
    // aggregate -> nodes, centroids, mode values (scaled by the cell size)
    std::vector<int> agg_off(na + 1, 0), agg_node(nn);
    for (int a = 0; a < nn; a++) agg_off[node_agg[a] + 1]++;
    for (int I = 0; I < na; I++) agg_off[I + 1] += agg_off[I];
    {
        std::vector<int> fill(agg_off.begin(), agg_off.end() - 1);
        for (int a = 0; a < nn; a++) agg_node[fill[node_agg[a]]++] = a;
    }
    std::vector<double> cen((size_t)na * 3, 0.0);
    for (int a = 0; a < nn; a++)
        for (int i = 0; i < 3; i++) cen[(size_t)node_agg[a] * 3 + i] += m.x[(size_t)a * 3 + i];
    for (int I = 0; I < na; I++)
        for (int i = 0; i < 3; i++) cen[(size_t)I * 3 + i] /= std::max(agg_off[I + 1] - agg_off[I], 1);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int a = 0; a < nn; a++)
        for (int i = 0; i < D; i++) {
            lo[i] = std::min(lo[i], m.x[(size_t)a * 3 + i]);
            hi[i] = std::max(hi[i], m.x[(size_t)a * 3 + i]);
        }
    double vol = 1.0;
    for (int i = 0; i < D; i++) vol *= std::max(hi[i] - lo[i], 1e-300);
    const double hs = agg * std::pow(vol / std::max(nn, 1), 1.0 / D);
    std::vector<double> Pn((size_t)nn * D * NM, 0.0);
    for (int a = 0; a < nn; a++) {
        const int I = node_agg[a];
        double r[3];
        for (int i = 0; i < 3; i++) r[i] = (m.x[(size_t)a * 3 + i] - cen[(size_t)I * 3 + i]) / hs;
        double blk[D][NM] = {};
        for (int i = 0; i < D; i++) blk[i][i] = 1.0;  // translations
        if constexpr (D == 2) {
            blk[0][2] = -r[1];
            blk[1][2] = r[0];
        } else {
            // rotations about x, y, z: omega x r
            blk[1][3] = -r[2]; blk[2][3] = r[1];
            blk[0][4] = r[2];  blk[2][4] = -r[0];
            blk[0][5] = -r[1]; blk[1][5] = r[0];
        }
        for (int i = 0; i < D; i++) {
            const bool fr = m.eq[(size_t)a * 3 + i] >= 0;
            for (int k = 0; k < NM; k++) Pn[((size_t)a * D + i) * NM + k] = fr ? blk[i][k] : 0.0;
        }
    }
    // aggregates touched by each node row, slot of every block
    std::vector<int> kp_off(nn + 1, 0), kp_agg, blk_slot(c.nblk);
    std::vector<int> tmp;
    for (int a = 0; a < nn; a++) {
        tmp.clear();
        for (int b = m.adj_off[a]; b < m.adj_off[a + 1]; b++)
            if (nz[b]) tmp.push_back(node_agg[m.adj[b]]);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        for (int b = m.adj_off[a]; b < m.adj_off[a + 1]; b++) {
            if (!nz[b]) { blk_slot[b] = -1; continue; }
            blk_slot[b] = int(std::lower_bound(tmp.begin(), tmp.end(), node_agg[m.adj[b]]) - tmp.begin());
        }
        kp_agg.insert(kp_agg.end(), tmp.begin(), tmp.end());
        kp_off[a + 1] = (int)kp_agg.size();
    }
    int rc;
    int *blk_slot_d;
    if ((rc = tl_alloc(t, t.agg_off, na + 1, err)) || (rc = tl_alloc(t, t.agg_node, nn, err)) ||
        (rc = tl_alloc(t, t.node_agg, nn, err)) || (rc = tl_alloc(t, t.kp_off, nn + 1, err)) ||
        (rc = tl_alloc(t, t.kp_agg, kp_agg.size(), err)) || (rc = tl_alloc(t, t.Pn, Pn.size(), err)) ||
        (rc = tl_alloc(t, t.kp, kp_agg.size() * D * NM, err)) || (rc = tl_alloc(t, blk_slot_d, c.nblk, err)) ||
        (rc = tl_alloc(t, t.Ainv, (size_t)t.ld * t.ld, err)))
        return rc;
    PDX_HIP(hipMemcpyAsync(t.agg_off, agg_off.data(), agg_off.size() * 4, hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemcpyAsync(t.agg_node, agg_node.data(), agg_node.size() * 4, hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemcpyAsync(t.node_agg, node_agg.data(), node_agg.size() * 4, hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemcpyAsync(t.kp_off, kp_off.data(), kp_off.size() * 4, hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemcpyAsync(t.kp_agg, kp_agg.data(), std::max<size_t>(kp_agg.size(), 1) * 4, hipMemcpyHostToDevice,
                           c.stream));
    PDX_HIP(hipMemcpyAsync(t.Pn, Pn.data(), Pn.size() * 8, hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemcpyAsync(blk_slot_d, blk_slot.data(), blk_slot.size() * 4, hipMemcpyHostToDevice, c.stream));
    k_coarse_ap<D, NM><<<(nn + 127) / 128, 128, 0, c.stream>>>(nn, c.adj_off, c.adj, K64, t.Pn, blk_slot_d, t.kp_off,
                                                                c.eq, t.kp);
    // dense Galerkin matrix; padding rows/columns stay zero and drop out of the inverse
    double *A = t.Ainv, *diag0, *Dk, *tmpb;
    if ((rc = tl_alloc(t, diag0, t.ld, err)) || (rc = tl_alloc(t, Dk, GJB * GJB, err)) ||
        (rc = tl_alloc(t, tmpb, (size_t)t.ld * GJB, err)))
        return rc;
    k_fill_d<<<2048, 256, 0, c.stream>>>((size_t)t.ld * t.ld, A, 0.0);
    k_coarse_galerkin<D, NM><<<(t.nc + 63) / 64, 64, 0, c.stream>>>(na, t.agg_off, t.agg_node, t.Pn, t.kp_off,
                                                                    t.kp_agg, t.kp, t.ld, A);
    k_diag_copy<<<(t.ld + 255) / 256, 256, 0, c.stream>>>(t.ld, t.ld, A, diag0);
    const int nb = t.ld / GJB;
    for (int k = 0; k < nb; k++) {
        k_gj_diag<<<1, 1024, 0, c.stream>>>(t.ld, k, A, diag0, Dk);
        k_gj_row<<<nb, 256, 0, c.stream>>>(t.ld, nb, k, A, Dk, tmpb);
        k_gj_update<<<nb * nb, 256, 0, c.stream>>>(t.ld, nb, k, A);
        k_gj_col<<<nb, 256, 0, c.stream>>>(t.ld, nb, k, A, Dk, tmpb);
    }
    PDX_HIP(hipGetLastError());
    PDX_HIP(hipStreamSynchronize(c.stream));
    t.setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PDX_OK;
}

static int ensure_two_level_f32(Ctx &c, TwoLevel &t, std::string &err) {
    if (t.Ainv32) return PDX_OK;
    const size_t np = (size_t)c.nn * t.dim * t.nm;
    int kpn = 0;
    PDX_HIP(hipMemcpy(&kpn, t.kp_off + c.nn, 4, hipMemcpyDeviceToHost));
    int rc;
    if ((rc = tl_alloc(t, t.Pn32, np, err)) || (rc = tl_alloc(t, t.kp32, (size_t)kpn * t.dim * t.nm, err)) ||
        (rc = tl_alloc(t, t.Ainv32, (size_t)t.ld * t.ld, err)))
        return rc;
    k_copy<double, float><<<2048, 256, 0, c.stream>>>(np, t.Pn, t.Pn32);
    k_copy<double, float><<<2048, 256, 0, c.stream>>>((size_t)kpn * t.dim * t.nm, t.kp, t.kp32);
    k_copy<double, float><<<2048, 256, 0, c.stream>>>((size_t)t.ld * t.ld, t.Ainv, t.Ainv32);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

static int default_agg(int dim) { return dim == 3 ? 6 : 10; }

// the two-level data of operator K64 (cached on the context)
static int get_two_level(Ctx &c, const double *K64, int agg, TwoLevel *&out, std::string &err) {
    if (!c.host || (int)c.host->x.size() != c.nn * 3 || c.host->dim != c.dim) {
        err = "two-level preconditioner needs the node coordinates of a model (use pdx_static_solve)";
        return PDX_ERR_ARG;
    }
    if (agg <= 0) agg = default_agg(c.dim);
    TwoLevel *t = (TwoLevel *)c.two_level;
    if (t && t->K == K64 && t->gen == c.K_gen && t->agg_size == agg) {
        out = t;
        return PDX_OK;
    }
    delete t;
    c.two_level = nullptr;
    t = new TwoLevel();
    int rc = c.dim == 3 ? build_two_level<3, 6>(c, K64, agg, *t, err) : build_two_level<2, 3>(c, K64, agg, *t, err);
    if (rc) {
        delete t;
        return rc;
    }
    t->K = K64;
    t->gen = c.K_gen;
    c.two_level = t;
    out = t;
    return PDX_OK;
}

void free_two_level(Ctx &c) {
    delete (TwoLevel *)c.two_level;
    c.two_level = nullptr;
}

// ------------------------------------------------------------------------
// host: BiCGStab
// ------------------------------------------------------------------------
// device pointers a captured chunk of iterations reads; a change of any of them re-captures
struct GraphSig {
    const void *K = nullptr, *RS = nullptr, *tl = nullptr, *ilu = nullptr, *Pn = nullptr, *kp = nullptr,
               *Ainv = nullptr, *rc = nullptr, *yc = nullptr, *red = nullptr, *src = nullptr;
    long long K_gen = -1, tl_serial = -1, ilu_serial = -1;  // launch shapes follow these objects
    int krylov = -1;
    bool same(const GraphSig &o) const {
        return K == o.K && RS == o.RS && tl == o.tl && ilu == o.ilu && Pn == o.Pn && kp == o.kp && Ainv == o.Ainv &&
               rc == o.rc && yc == o.yc && red == o.red && src == o.src && K_gen == o.K_gen && tl_serial == o.tl_serial &&
               ilu_serial == o.ilu_serial && krylov == o.krylov;
    }
};

// Krylov workspace of one precision, kept on the context across solves so that the
// captured iteration graphs are replayed by every later solve with the same operator
template <typename V>
struct KrylovWs {
    V *x = nullptr, *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr,
      *ph = nullptr, *sh = nullptr, *xb = nullptr, *Minv = nullptr;
    double *rc = nullptr, *yc = nullptr;
    size_t rc_cap = 0;
    V *RSbuf = nullptr;  // FP32 block row sums of the operator (FP32 workspaces)
    double *y64 = nullptr;  // FP64 scratch vector of the start-up residual
    int *status = nullptr;  // singular-block flag of the block-diagonal preconditioners
    // chunks of CHUNK iterations: the first (iteration 0 starts a direction) and any later one
    hipGraphExec_t g_first = nullptr, g_next = nullptr;
    GraphSig sig;
    // two-level operator (null for the block-diagonal preconditioners)
    const TwoLevel *tl = nullptr;
    const V *Pn = nullptr, *kp = nullptr, *Ainv = nullptr;
    const float *Ainv32 = nullptr;  // the coarse inverse applied by both precisions
    const V *RS = nullptr;  // block row sums (difference-form SpMV) or null
    const void *ilu = nullptr;  // ILU(0) factor (ilu0.hip) or null
    const int *src = nullptr;  // own index map when the iterations read mirrored storage, else null
};

template <typename V>
static int ws_alloc(Ctx &c, KrylovWs<V> &w, std::string &err) {
    const size_t n = (size_t)c.nn * c.dim;
    V **ptrs[] = {&w.x, &w.r, &w.rh, &w.p, &w.v, &w.s, &w.t, &w.ph, &w.sh, &w.xb};
    for (auto pp : ptrs) {
        PDX_HIP(hipMalloc((void **)pp, n * sizeof(V)));
        PDX_HIP(hipMemsetAsync(*pp, 0, n * sizeof(V), c.stream));
    }
    PDX_HIP(hipMalloc((void **)&w.Minv, (size_t)c.nn * c.dd * sizeof(V)));
    PDX_HIP(hipMalloc((void **)&w.y64, n * sizeof(double)));
    PDX_HIP(hipMalloc((void **)&w.status, sizeof(int)));
    return PDX_OK;
}

template <typename V>
static void ws_drop_graphs(KrylovWs<V> &w) {
    if (w.g_first) hipGraphExecDestroy(w.g_first);
    if (w.g_next) hipGraphExecDestroy(w.g_next);
    w.g_first = w.g_next = nullptr;
}

template <typename V>
static void ws_free(KrylovWs<V> &w) {
    ws_drop_graphs(w);
    void *ptrs[] = {w.x, w.r, w.rh, w.p, w.v, w.s, w.t, w.ph, w.sh, w.xb, w.Minv, w.rc, w.yc, (void *)w.RSbuf,
                    w.y64, w.status};
    for (auto p : ptrs)
        if (p) hipFree(p);
}

// the context's workspace of precision V (allocated on first use); preconditioner fields cleared
template <typename V>
static int ws_get(Ctx &c, KrylovWs<V> *&out, std::string &err) {
    void *&slot = c.krylov_ws[sizeof(V) == 4 ? 1 : 0];
    if (!slot) {
        auto *w = new KrylovWs<V>();
        int rc = ws_alloc(c, *w, err);
        if (rc) {
            ws_free(*w);
            delete w;
            return rc;
        }
        slot = w;
    }
    out = (KrylovWs<V> *)slot;
    out->tl = nullptr;
    out->ilu = nullptr;
    out->Pn = out->kp = out->Ainv = nullptr;
    out->Ainv32 = nullptr;
    out->RS = nullptr;
    return PDX_OK;
}

void free_krylov_ws(Ctx &c) {
    if (c.solve_y) hipFree(c.solve_y);
    if (c.solve_b32) hipFree(c.solve_b32);
    c.solve_y = nullptr;
    c.solve_b32 = nullptr;
    if (c.krylov_ws[0]) {
        ws_free(*(KrylovWs<double> *)c.krylov_ws[0]);
        delete (KrylovWs<double> *)c.krylov_ws[0];
    }
    if (c.krylov_ws[1]) {
        ws_free(*(KrylovWs<float> *)c.krylov_ws[1]);
        delete (KrylovWs<float> *)c.krylov_ws[1];
    }
    c.krylov_ws[0] = c.krylov_ws[1] = nullptr;
}

constexpr int CHUNK = 10;  // iterations between host residual checks

// Run n iterations body(it), body(it + 1), ...  A full chunk (n == CHUNK) from an even
// iteration replays a captured hipGraph: the first chunk differs from the later ones only by
// iteration 0, and every later chunk launches the same kernels with the same arguments
// (the iteration index enters through its parity only).  A change of operator or
// preconditioner pointers re-captures.
template <typename V, typename Body>
static int run_chunk(Ctx &c, KrylovWs<V> &w, const GraphSig &sig, bool use_graph, int it, int n, Body &&body,
                     std::string &err) {
    // the graphs hold CHUNK iterations from an even start (iteration parity selects the rho
    // slots): any other chunk runs as plain launches
    if (!use_graph || n != CHUNK || (it & 1)) {
        for (int k = 0; k < n; k++) body(it + k);
        return PDX_OK;
    }
    if (!w.sig.same(sig)) {
        ws_drop_graphs(w);
        w.sig = sig;
    }
    hipGraphExec_t &g = it == 0 ? w.g_first : w.g_next;
    if (!g) {
        // every exit after hipStreamBeginCapture ends the capture, so the stream never stays
        // in capture mode; a launch error inside the capture fails the chunk
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (int k = 0; k < n; k++) body(it + k);
            const hipError_t launch = hipGetLastError();
            e = hipStreamEndCapture(c.stream, &gr);
            if (e == hipSuccess) e = launch;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&g, gr, nullptr, nullptr, 0);
        if (gr) hipGraphDestroy(gr);
        if (e != hipSuccess) {
            g = nullptr;
            err = std::string("Krylov chunk graph capture: ") + hipGetErrorString(e);
            return PDX_ERR_HIP;
        }
    }
    PDX_HIP(hipGraphLaunch(g, c.stream));
    return PDX_OK;
}

// Iterations of the next chunk: a full chunk, or, once the last chunk has lowered the residual,
// the count its mean rate predicts to reach tol, plus one (a chunk shorter than CHUNK runs as
// plain launches); a solve then stops within a few iterations of its tolerance instead of up
// to CHUNK - 1 past it.  Depends only on the residuals, so graph and plain launches agree.
static int next_chunk(double rel, double rel_prev, int prev_n, double tol, int left) {
    int n = std::min(CHUNK, left);
    if (prev_n > 0 && rel_prev > 0.0 && rel < rel_prev && rel > tol) {
        const double rate = std::log(rel / rel_prev) / prev_n;  // < 0
        const double need = std::log(tol / rel) / rate;  // > 0
        if (need + 1.0 < n) n = std::max(1, (int)std::ceil(need) + 1);
    }
    return n;
}

template <typename V>
static GraphSig make_sig(Ctx &c, const void *K, const KrylovWs<V> &w, int krylov) {
    GraphSig g;
    g.K = K;
    g.RS = w.RS;
    g.tl = w.tl;
    g.ilu = w.ilu;
    g.Pn = w.Pn;
    g.kp = w.kp;
    g.Ainv = w.Ainv32;
    g.rc = w.rc;
    g.yc = w.yc;
    g.red = c.red;
    g.src = w.src;
    g.K_gen = c.K_gen;
    g.tl_serial = w.tl ? w.tl->serial : -1;
    g.ilu_serial = w.ilu ? ilu0_serial(w.ilu) : -1;
    g.krylov = krylov;
    return g;
}

// y = mask(A x) with fused dot partials; A is K (all blocks) or, with src, the own blocks of
// its mirrored storage (node-block operators only)
// 1- and 2-DOF operators use the persistent form; the 3-DOF operators the blocked form, whose
// per-block dot partials live in c.red after the RED_SLOTS partial slots (solve_masked allocates
// them before any solve that asks for dots) and are folded by k_fold_parts.  Either way mirrored
// and full storage run the same kernel, so their dots and results are the same bits.
template <int D, typename T, typename V>
static void spmv_dot(Ctx &c, const T *K, const int *src, const T *RS, const V *x, V *y, const V *w0, double *p0,
                     const V *w1, double *p1) {
    if constexpr (D < 3) {
#define PDX_SPMV_DOT_P(REL, SYM)                                                                               \
    k_spmv_dot_p<D, T, V, REL, SYM><<<RED_PARTS, SPMV_THREADS, 0, c.stream>>>(                                \
        c.rows_short, c.n_short, c.rows_long, c.n_long, c.adj_off, c.adj, REL ? c.cls : nullptr,              \
        REL ? c.cls_off : nullptr, REL ? c.cls_rel : nullptr, src, K, RS, x, y, c.eq, w0, p0, w1, p1)
        if constexpr (D > 1) {
            if (src) {
                if (c.cls) PDX_SPMV_DOT_P(true, true);
                else PDX_SPMV_DOT_P(false, true);
                return;
            }
        }
        if (c.cls) PDX_SPMV_DOT_P(true, false);
        else PDX_SPMV_DOT_P(false, false);
#undef PDX_SPMV_DOT_P
    } else {
        const SpmvShape sh = spmv_shape(c);
        double *bp = c.red ? c.red + (size_t)RED_SLOTS * NPART + 16 : nullptr;
        double *bp0 = p0 ? bp : nullptr, *bp1 = p1 ? bp + sh.grid : nullptr;
        const int rev = int(c.spmv_calls++ & 1);  // every other product sweeps the rows in reverse
#define PDX_SPMV_DOT(REL, SYM)                                                                                  \
    k_spmv_dot<D, T, V, REL, SYM><<<sh.grid, 256, 0, c.stream>>>(                                             \
        c.rows_short, c.n_short, sh.nbs, c.rows_long, c.n_long, sh.nbt, c.adj_off, c.adj, REL ? c.cls : nullptr, \
        REL ? c.cls_off : nullptr, REL ? c.cls_rel : nullptr, src, K, RS, x, y, c.eq, w0, bp0, w1, bp1, rev)
        if (src) {
            if (c.cls) PDX_SPMV_DOT(true, true);
            else PDX_SPMV_DOT(false, true);
        } else if (c.cls) {
            PDX_SPMV_DOT(true, false);
        } else {
            PDX_SPMV_DOT(false, false);
        }
#undef PDX_SPMV_DOT
        if (p0 || p1) k_fold_parts<<<NPART, 64, 0, c.stream>>>(sh.grid, bp0, p0, bp1, p1);
    }
}

// one BiCGStab iteration (it = iteration index, parity selects the rho slot)
template <int D, typename T, typename V>
static void bicg_iter(Ctx &c, const T *K, KrylovWs<V> &w, double *red, double *sc, int it) {
    double *rho = red + (size_t)((it & 1) ? P_RHO1 : P_RHO0) * NPART;
    double *rho_next = red + (size_t)((it & 1) ? P_RHO0 : P_RHO1) * NPART;
    double *rv = red + (size_t)P_RV * NPART, *ts = red + (size_t)P_TS * NPART, *tt = red + (size_t)P_TT * NPART,
           *rr = red + (size_t)P_RR * NPART;
    if (w.tl) {
        constexpr int NM = D == 3 ? 6 : 3;
        const TwoLevel &t = *w.tl;
        const int gc = std::min(2048, (t.nc + 3) / 4), gn = (c.nn + 255) / 256;
        k_bicg_p_agg<D, NM, V><<<RED_PARTS, 256, 0, c.stream>>>(t.na, t.agg_off, t.agg_node, w.r, w.p, w.v, w.Pn,
                                                                rho, sc, it == 0, w.rc);
        k_coarse_mv<float><<<gc, 256, 0, c.stream>>>(t.nc, t.ld, w.Ainv32, w.rc, w.yc);
        k_prolong<D, NM, V><<<gn, 256, 0, c.stream>>>(c.nn, w.Pn, t.node_agg, t.kp_off, t.kp_agg, w.kp, w.Minv, w.p,
                                                      w.yc, w.ph);
        spmv_dot<D, T, V>(c, K, w.src, (const T *)w.RS, w.ph, w.v, w.rh, rv, (const V *)nullptr, nullptr);
        k_bicg_s_agg<D, NM, V><<<RED_PARTS, 256, 0, c.stream>>>(t.na, t.agg_off, t.agg_node, w.r, w.v, w.s, w.Pn,
                                                                rho, rv, sc, w.rc);
        k_coarse_mv<float><<<gc, 256, 0, c.stream>>>(t.nc, t.ld, w.Ainv32, w.rc, w.yc);
        k_prolong<D, NM, V><<<gn, 256, 0, c.stream>>>(c.nn, w.Pn, t.node_agg, t.kp_off, t.kp_agg, w.kp, w.Minv, w.s,
                                                      w.yc, w.sh);
    } else if (w.ilu) {
        k_bicg_p<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.r, w.p, w.v, w.ph, (const V *)nullptr, rho, sc,
                                                        it == 0);
        ilu0_apply(c, w.ilu, w.p, w.ph);
        spmv_dot<D, T, V>(c, K, w.src, (const T *)w.RS, w.ph, w.v, w.rh, rv, (const V *)nullptr, nullptr);
        k_bicg_s<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.r, w.v, w.s, w.sh, (const V *)nullptr, rho, rv, sc);
        ilu0_apply(c, w.ilu, w.s, w.sh);
    } else {
        k_bicg_p<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.r, w.p, w.v, w.ph, w.Minv, rho, sc, it == 0);
        spmv_dot<D, T, V>(c, K, w.src, (const T *)w.RS, w.ph, w.v, w.rh, rv, (const V *)nullptr, nullptr);
        k_bicg_s<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.r, w.v, w.s, w.sh, w.Minv, rho, rv, sc);
    }
    spmv_dot<D, T, V>(c, K, w.src, (const T *)w.RS, w.sh, w.t, w.s, ts, w.t, tt);
    k_bicg_x<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.x, w.r, w.rh, w.ph, w.sh, w.s, w.t, rho, ts, tt, sc,
                                                    rho_next, rr);
}

// Solve A x = b (full DOF vectors, prescribed entries zero) with preconditioned
// BiCGStab in precision V/T to relative residual tol (recurrence residual),
// starting from x.  Keeps the iterate of the best residual seen; a residual
// that grows 1e3-fold past it (loss of accuracy in low precision) ends the
// run with that iterate restored.
template <int D, typename T, typename V>
static int bicgstab(Ctx &c, const T *K, KrylovWs<V> &w, const V *b, double bnorm, double tol, int max_iter,
                    bool use_graph, int *iters, double *relres, std::string &err) {
    const size_t n = (size_t)c.nn * c.dim;
    double *red = c.red, *sc = c.red + (size_t)RED_SLOTS * NPART;
    double *y = w.y64;
    {
        V *tmp = w.t;
        spmv_dot<D, T, V>(c, K, w.src, (const T *)w.RS, w.x, tmp, (const V *)nullptr, nullptr, (const V *)nullptr, nullptr);
        k_copy<V, double><<<2048, 256, 0, c.stream>>>(n, tmp, y);
        k_resid<V, V><<<RED_PARTS, 256, 0, c.stream>>>(n, b, y, w.r, w.rh, red + (size_t)P_RR * NPART);
        k_dot<V><<<RED_PARTS, 256, 0, c.stream>>>(w.rh, w.r, n, red + (size_t)P_RHO0 * NPART);
        PDX_HIP(hipMemsetAsync(w.p, 0, n * sizeof(V), c.stream));
        PDX_HIP(hipMemsetAsync(w.v, 0, n * sizeof(V), c.stream));
    }
    double rr0 = host_sum(c, red + (size_t)P_RR * NPART, NPART, err);
    double rel = std::sqrt(rr0) / bnorm;
    int it = 0, since_best = 0;
    double best = rel;
    PDX_HIP(hipMemcpyAsync(w.xb, w.x, n * sizeof(V), hipMemcpyDeviceToDevice, c.stream));
    const GraphSig sig = make_sig(c, K, w, PDX_KRYLOV_BICGSTAB);
    double rel_prev = 0.0;
    int prev_n = 0;
    while (rel > tol && it < max_iter) {
        const int nrun = next_chunk(rel, rel_prev, prev_n, tol, max_iter - it);
        int rc = run_chunk(c, w, sig, use_graph, it, nrun,
                           [&](int i) { bicg_iter<D, T, V>(c, K, w, red, sc, i); }, err);
        if (rc) return rc;
        it += nrun;
        PDX_HIP(hipGetLastError());
        const double rr = host_sum(c, red + (size_t)P_RR * NPART, NPART, err);
        rel_prev = rel;
        prev_n = nrun;
        rel = std::sqrt(rr) / bnorm;
        if (!std::isfinite(rr) || rel > 1e3 * best) {
            // lost accuracy: fall back to the best iterate
            PDX_HIP(hipMemcpyAsync(w.x, w.xb, n * sizeof(V), hipMemcpyDeviceToDevice, c.stream));
            rel = best;
            break;
        }
        if (rel < 0.95 * best) {
            best = rel;
            since_best = 0;
            PDX_HIP(hipMemcpyAsync(w.xb, w.x, n * sizeof(V), hipMemcpyDeviceToDevice, c.stream));
        } else if ((since_best += nrun) >= 40 * CHUNK) {  // iterations without a 5% gain
            // stagnation at the rounding floor of the working precision
            PDX_HIP(hipMemcpyAsync(w.x, w.xb, n * sizeof(V), hipMemcpyDeviceToDevice, c.stream));
            rel = best;
            break;
        }
    }
    *iters = it;
    *relres = rel;
    return PDX_OK;
}

// Preconditioned CG on a symmetric operator (block-diagonal preconditioners):
// same stopping rules and best-iterate restore as bicgstab().  Workspace: p = w.p,
// q = w.v, z = w.s.
template <int D, typename T, typename V>
static int pcg(Ctx &c, const T *K, KrylovWs<V> &w, const V *b, double bnorm, double tol, int max_iter,
               bool use_graph, int *iters, double *relres, std::string &err) {
    const size_t n = (size_t)c.nn * c.dim;
    double *red = c.red;
    double *y = w.y64;
    spmv_dot<D, T, V>(c, K, w.src, (const T *)w.RS, w.x, w.t, (const V *)nullptr, nullptr, (const V *)nullptr, nullptr);
    k_copy<V, double><<<2048, 256, 0, c.stream>>>(n, w.t, y);
    k_resid<V, V><<<RED_PARTS, 256, 0, c.stream>>>(n, b, y, w.r, nullptr, red + (size_t)P_RR * NPART);
    // z = Minv r and rho0 = (r, z) through a zero step
    PDX_HIP(hipMemsetAsync(w.p, 0, n * sizeof(V), c.stream));
    PDX_HIP(hipMemsetAsync(w.v, 0, n * sizeof(V), c.stream));
    k_fill_d<<<1, 256, 0, c.stream>>>(NPART, red + (size_t)P_RV * NPART, 0.0);
    k_fill_d<<<1, 256, 0, c.stream>>>(NPART, red + (size_t)P_TMP * NPART, 1.0);  // pq != 0, alpha = 0
    k_cg_xr<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.x, w.r, w.s, w.p, w.v, w.Minv, red + (size_t)P_RV * NPART,
                                                  red + (size_t)P_TMP * NPART, red + (size_t)P_RHO0 * NPART,
                                                  red + (size_t)P_RR * NPART);
    double rel = std::sqrt(host_sum(c, red + (size_t)P_RR * NPART, NPART, err)) / bnorm;
    int it = 0, since_best = 0;
    double best = rel;
    PDX_HIP(hipMemcpyAsync(w.xb, w.x, n * sizeof(V), hipMemcpyDeviceToDevice, c.stream));
    const GraphSig sig = make_sig(c, K, w, PDX_KRYLOV_CG);
    auto body = [&](int i) {
        double *rho = red + (size_t)((i & 1) ? P_RHO1 : P_RHO0) * NPART;
        // the other slot holds the previous rho until this step's k_cg_xr overwrites it
        double *rho_new = red + (size_t)((i & 1) ? P_RHO0 : P_RHO1) * NPART;
        double *rho_prev = rho_new;
        k_cg_p<V><<<RED_PARTS, 256, 0, c.stream>>>(n, w.p, w.s, rho, rho_prev, i == 0);
        spmv_dot<D, T, V>(c, K, w.src, (const T *)w.RS, w.p, w.v, w.p, red + (size_t)P_RV * NPART, (const V *)nullptr,
                          nullptr);
        k_cg_xr<D, V><<<RED_PARTS, 256, 0, c.stream>>>(c.nn, w.x, w.r, w.s, w.p, w.v, w.Minv, rho,
                                                      red + (size_t)P_RV * NPART, rho_new,
                                                      red + (size_t)P_RR * NPART);
    };
    double rel_prev = 0.0;
    int prev_n = 0;
    while (rel > tol && it < max_iter) {
        const int nrun = next_chunk(rel, rel_prev, prev_n, tol, max_iter - it);
        int rc = run_chunk(c, w, sig, use_graph, it, nrun, body, err);
        if (rc) return rc;
        it += nrun;
        PDX_HIP(hipGetLastError());
        const double rr = host_sum(c, red + (size_t)P_RR * NPART, NPART, err);
        rel_prev = rel;
        prev_n = nrun;
        rel = std::sqrt(rr) / bnorm;
        if (!std::isfinite(rr) || rel > 1e3 * best) {
            PDX_HIP(hipMemcpyAsync(w.x, w.xb, n * sizeof(V), hipMemcpyDeviceToDevice, c.stream));
            rel = best;
            break;
        }
        if (rel < 0.95 * best) {
            best = rel;
            since_best = 0;
            PDX_HIP(hipMemcpyAsync(w.xb, w.x, n * sizeof(V), hipMemcpyDeviceToDevice, c.stream));
        } else if ((since_best += nrun) >= 40 * CHUNK) {  // iterations without a 5% gain
            PDX_HIP(hipMemcpyAsync(w.x, w.xb, n * sizeof(V), hipMemcpyDeviceToDevice, c.stream));
            rel = best;
            break;
        }
    }
    *iters = it;
    *relres = rel;
    return PDX_OK;
}

template <int D, typename V>
static int build_precond(Ctx &c, const double *K64, int kind, V *Minv, int *status, std::string &err) {
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
    } else if (kind == PDX_PREC_LEARNED) {
        if (!c.learned_Minv || c.learned_gen != c.K_gen || c.learned_for != K64) {
            err = "no learned smoother for the current operator (run pdx_gnn_apply first)";
            return PDX_ERR_ARG;
        }
        k_copy<double, V><<<2048, 256, 0, c.stream>>>((size_t)c.nn * c.dd, c.learned_Minv, Minv);
    } else {
        k_blockjacobi<D, V><<<(c.nn + 255) / 256, 256, 0, c.stream>>>(c.nn, c.adj_off, c.adj, K64, c.eq,
                                                                      kind == PDX_PREC_JACOBI, Minv, status);
    }
    int hs = 0;
    PDX_HIP(hipMemcpyAsync(&hs, status, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    if (hs) { err = "singular diagonal block in the preconditioner"; return PDX_ERR_SINGULAR; }
    return PDX_OK;
}

// preconditioner of a workspace: node-block part and, for the two-level kinds, the coarse operator
template <int D, typename V>
static int setup_precond(Ctx &c, const double *K64, const pdx_solve_opts &o, KrylovWs<V> &w, std::string &err) {
    if (o.precond == PDX_PREC_ILU0) {
        void *h;
        int rc = ilu0_prepare(c, K64, sizeof(V) == 4, h, err);
        w.ilu = rc ? nullptr : h;
        return rc;
    }
    const bool two = o.precond == PDX_PREC_TWO_LEVEL || o.precond == PDX_PREC_LEARNED;
    int rc = build_precond<D, V>(c, K64, two && o.precond != PDX_PREC_LEARNED ? PDX_PREC_BLOCK_JACOBI : o.precond,
                                 w.Minv, w.status, err);
    if (rc || !two) return rc;
    TwoLevel *t;
    if ((rc = get_two_level(c, K64, o.coarse_agg, t, err))) return rc;
    w.tl = t;
    if ((size_t)t->ld > w.rc_cap) {
        if (w.rc) hipFree(w.rc);
        if (w.yc) hipFree(w.yc);
        w.rc = w.yc = nullptr;
        w.rc_cap = 0;
        PDX_HIP(hipMalloc((void **)&w.rc, (size_t)t->ld * 8));
        PDX_HIP(hipMalloc((void **)&w.yc, (size_t)t->ld * 8));
        w.rc_cap = (size_t)t->ld;
    }
    // the coarse inverse is applied in FP32 in both precisions: it only shapes the preconditioner
    // (the Krylov iteration and its residual keep the working precision), at half the bytes
    if ((rc = ensure_two_level_f32(c, *t, err))) return rc;
    w.Ainv32 = t->Ainv32;
    if (sizeof(V) == 4) {
        w.Pn = (const V *)t->Pn32;
        w.kp = (const V *)t->kp32;
        w.Ainv = (const V *)t->Ainv32;
    } else {
        w.Pn = (const V *)t->Pn;
        w.kp = (const V *)t->kp;
        w.Ainv = (const V *)t->Ainv;
    }
    return PDX_OK;
}

// Solve K_ff x = b in FP64 (plain) or FP32 inner + FP64 refinement.
// b_dev: full DOF vector (double, prescribed zero).  x_dev: in/out (double).
// mir: own-block copies of the operator the SpMVs read instead of K64 / K32 (the
// preconditioner set-up and the row sums still read K64).
template <int D>
static int solve_masked(Ctx &c, const double *K64, const float *K32, const pdx_solve_opts &o, const double *b_dev,
                        double *x_dev, pdx_solve_stats &st, std::string &err, const Mirrored &mir) {
    auto t0 = std::chrono::steady_clock::now();
    const size_t n = (size_t)c.nn * c.dim;
    if (!c.red) {
        // the partial slots, 16 scalars, then the SpMV's per-block partials of two dots
        const size_t nred = (size_t)RED_SLOTS * NPART + 16 + 2 * (size_t)spmv_shape(c).grid;
        PDX_HIP(hipMalloc((void **)&c.red, sizeof(double) * nred));
        PDX_HIP(hipMemsetAsync(c.red, 0, sizeof(double) * nred, c.stream));
    }
    double *part = c.red + (size_t)P_BB * NPART;
    k_dot<double><<<RED_PARTS, 256, 0, c.stream>>>(b_dev, b_dev, n, part);
    double bnorm = std::sqrt(host_sum(c, part, NPART, err));
    st.iterations = 0;
    st.refinements = 0;
    st.seconds = 0;
    st.true_relres = 1.0;
    st.converged = 0;
    st.mirrored_used = D > 1 && mir.K64 && (o.precision == PDX_FP64 || mir.K32) ? 1 : 0;
    if (bnorm == 0.0) {
        PDX_HIP(hipMemsetAsync(x_dev, 0, n * 8, c.stream));
        PDX_HIP(hipStreamSynchronize(c.stream));
        st.true_relres = 0; st.converged = 1;
        st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return PDX_OK;
    }
    if (!o.warm_start) PDX_HIP(hipMemsetAsync(x_dev, 0, n * 8, c.stream));
    int total = 0, refine = 0, rc = PDX_OK;
    double rel = 1.0;
    // FP64 residual vector and FP32 right-hand side, kept on the context across solves
    if (!c.solve_y) PDX_HIP(hipMalloc((void **)&c.solve_y, n * 8));
    double *y = c.solve_y;
    const bool use_graph = !o.no_graph;
    // operator of the SpMVs: FP64 (residuals, FP64 iterations) and FP32 (FP32 iterations)
    const int *src64 = D > 1 && mir.K64 ? c.ksrc : nullptr, *src32 = D > 1 && mir.K32 ? c.ksrc : nullptr;
    const double *A64 = src64 ? mir.K64 : K64;
    const float *A32 = src32 ? mir.K32 : K32;
    if (o.precision == PDX_FP64) {
        KrylovWs<double> *wp;
        if ((rc = ws_get(c, wp, err)) || (rc = setup_precond<D, double>(c, K64, o, *wp, err))) return rc;
        KrylovWs<double> &w = *wp;
        w.src = src64;
        PDX_HIP(hipMemcpyAsync(w.x, x_dev, n * 8, hipMemcpyDeviceToDevice, c.stream));
        double prev = INFINITY;
        for (int restart = 0; restart < 8; restart++) {
            int it = 0;
            double rr;
            rc = o.krylov == PDX_KRYLOV_CG
                     ? pcg<D, double, double>(c, A64, w, b_dev, bnorm, o.tol * 0.5, o.max_iter - total, use_graph,
                                              &it, &rr, err)
                     : bicgstab<D, double, double>(c, A64, w, b_dev, bnorm, o.tol * 0.5, o.max_iter - total,
                                                   use_graph, &it, &rr, err);
            total += it;
            if (rc) break;
            spmv_dot<D, double, double>(c, A64, src64, nullptr, w.x, y, (const double *)nullptr, nullptr,
                                        (const double *)nullptr, nullptr);
            k_resid<double, double><<<RED_PARTS, 256, 0, c.stream>>>(n, b_dev, y, w.t, nullptr, part);
            rel = std::sqrt(host_sum(c, part, NPART, err)) / bnorm;
            if (rel <= o.tol || total >= o.max_iter || !(rel < 0.9 * prev)) break;
            prev = rel;
        }
        PDX_HIP(hipMemcpyAsync(x_dev, w.x, n * 8, hipMemcpyDeviceToDevice, c.stream));
    } else {
        if (!K32) { err = "FP32 matrix missing"; return PDX_ERR_ARG; }
        KrylovWs<float> *wp;
        if (!c.solve_b32) PDX_HIP(hipMalloc((void **)&c.solve_b32, n * 4));
        float *b32 = c.solve_b32;
        if ((rc = ws_get(c, wp, err)) || (rc = setup_precond<D, float>(c, K64, o, *wp, err))) return rc;
        KrylovWs<float> &w = *wp;
        w.src = src32;
        if (!w.RSbuf) PDX_HIP(hipMalloc((void **)&w.RSbuf, (size_t)c.nn * c.dd * 4));
        k_block_rowsum<D, float><<<(c.nn + 255) / 256, 256, 0, c.stream>>>(c.nn, c.adj_off, K64, w.RSbuf);
        w.RS = w.RSbuf;
        const double inner = o.inner_tol > 0 ? o.inner_tol : 1e-4;
        double prev = INFINITY;
        int stalls = 0;
        for (refine = 0; refine < 60; refine++) {
            // FP64 residual r = b - K x
            spmv_dot<D, double, double>(c, A64, src64, nullptr, x_dev, y, (const double *)nullptr, nullptr,
                                        (const double *)nullptr, nullptr);
            k_resid<float, double><<<RED_PARTS, 256, 0, c.stream>>>(n, b_dev, y, b32, nullptr, part);
            rel = std::sqrt(host_sum(c, part, NPART, err)) / bnorm;
            if (rel <= o.tol || total >= o.max_iter) break;
            if (!(rel < 0.5 * prev) && ++stalls >= 3) break;
            prev = std::min(prev, rel);
            PDX_HIP(hipMemsetAsync(w.x, 0, n * 4, c.stream));
            int it = 0;
            double rr;
            rc = o.krylov == PDX_KRYLOV_CG
                     ? pcg<D, float, float>(c, A32, w, b32, rel * bnorm, inner, o.max_iter - total, use_graph, &it,
                                            &rr, err)
                     : bicgstab<D, float, float>(c, A32, w, b32, rel * bnorm, inner, o.max_iter - total, use_graph,
                                                 &it, &rr, err);
            total += it;
            if (rc) break;
            k_add_f2d<<<2048, 256, 0, c.stream>>>(n, w.x, x_dev);
        }
    }
    PDX_HIP(hipStreamSynchronize(c.stream));
    st.iterations = total;
    st.refinements = refine;
    st.true_relres = rel;
    st.converged = rel <= o.tol;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) return rc;
    if (!st.converged) {
        err = "Krylov solver did not reach the tolerance (true relative residual " + std::to_string(rel) + " after " +
              std::to_string(total) + " iterations)";
        return PDX_ERR_NOCONV;
    }
    return PDX_OK;
}

static int solve_once(Ctx &c, const double *K64, const float *K32, const pdx_solve_opts &o, const double *b_dev,
                      double *x_dev, pdx_solve_stats &st, std::string &err, const Mirrored &mir) {
    if (c.dim == 3) return solve_masked<3>(c, K64, K32, o, b_dev, x_dev, st, err, mir);
    if (c.dim == 2) return solve_masked<2>(c, K64, K32, o, b_dev, x_dev, st, err, mir);
    return solve_masked<1>(c, K64, K32, o, b_dev, x_dev, st, err, mir);
}

// the attempt after o in the fallback chain (learned / two-level -> ILU(0) -> node-block Jacobi ->
// FP64 node-block Jacobi, all with BiCGStab); false at its end
static bool next_fallback(pdx_solve_opts &o) {
    o.krylov = PDX_KRYLOV_BICGSTAB;
    if (o.precond == PDX_PREC_LEARNED || o.precond == PDX_PREC_TWO_LEVEL) {
        o.precond = PDX_PREC_ILU0;
        return true;
    }
    if (o.precond != PDX_PREC_BLOCK_JACOBI) {
        o.precond = PDX_PREC_BLOCK_JACOBI;
        return true;
    }
    if (o.precision == PDX_FP32) {
        o.precision = PDX_FP64;
        return true;
    }
    return false;
}

// solve with an explicit node-block matrix (K, K_eff or a 1x1-block CSR)
int dev_solve_mat(Ctx &c, const double *K64, const float *K32, const pdx_solve_opts &o, const double *b_dev,
                  double *x_dev, pdx_solve_stats &st, std::string &err, const Mirrored *mir) {
    if (o.krylov != PDX_KRYLOV_BICGSTAB && o.krylov != PDX_KRYLOV_CG) { err = "unknown Krylov method"; return PDX_ERR_ARG; }
    if (o.krylov == PDX_KRYLOV_CG &&
        (o.precond == PDX_PREC_TWO_LEVEL || o.precond == PDX_PREC_LEARNED || o.precond == PDX_PREC_ILU0)) {
        err = "CG needs a symmetric preconditioner (none, Jacobi, block Jacobi); two-level and ILU(0) are not";
        return PDX_ERR_ARG;
    }
    if (o.precond < PDX_PREC_NONE || o.precond > PDX_PREC_ILU0) { err = "unknown preconditioner"; return PDX_ERR_ARG; }
    if (c.dim == 1 && (o.precond == PDX_PREC_TWO_LEVEL || o.precond == PDX_PREC_LEARNED)) {
        err = "two-level preconditioners need the node-block operator of a model (use pdx_static_solve)";
        return PDX_ERR_ARG;
    }
    pdx_solve_opts a = o;
    int iters = 0, refs = 0;
    double secs = 0;
    for (int attempt = 1;; attempt++) {
        int rc = solve_once(c, K64, K32, a, b_dev, x_dev, st, err, mir ? *mir : Mirrored{});
        iters += st.iterations;
        refs += st.refinements;
        secs += st.seconds;
        st.iterations = iters;
        st.refinements = refs;
        st.seconds = secs;
        st.precond_used = a.precond;
        st.precision_used = a.precision;
        st.attempts = attempt;
        // a learned preconditioner with no smoother predicted for this operator also falls back
        const bool learned_missing = rc == PDX_ERR_ARG && a.precond == PDX_PREC_LEARNED &&
                                     (!c.learned_Minv || c.learned_gen != c.K_gen || c.learned_for != K64);
        if (rc == PDX_OK || !o.fallback || (rc != PDX_ERR_NOCONV && rc != PDX_ERR_SINGULAR && !learned_missing))
            return rc;
        // restart from the iterate the failed attempt left, unless it is no better than zero
        const bool keep = rc == PDX_ERR_NOCONV && std::isfinite(st.true_relres) && st.true_relres < 1.0;
        if (!next_fallback(a)) return rc;
        a.warm_start = keep ? 1 : 0;
    }
}

int dev_solve(Ctx &c, const pdx_solve_opts &o, const double *b_dev, double *x_dev, pdx_solve_stats &st,
              std::string &err) {
    if (!c.K_valid) { err = "K not assembled"; return PDX_ERR_ARG; }
    int rc;
    if (o.precision == PDX_FP32 && (rc = ensure_K32(c, err))) return rc;
    Mirrored mir;
    if (o.mirrored && c.dim > 1 && (rc = kown_for_solve(c, o.precision, mir, err))) return rc;
    return dev_solve_mat(c, c.K, o.precision == PDX_FP32 ? c.K32 : nullptr, o, b_dev, x_dev, st, err, &mir);
}

// ------------------------------------------------------------------------
// coarse-corrected probes for training the learned smoother
// ------------------------------------------------------------------------
template <int D, int NM>
__global__ void __launch_bounds__(256) k_restrict_agg(int na, const int *agg_off, const int *agg_node,
                                                      const double *v, const double *Pn, double *rc) {
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int I = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; I < na; I += nw) {
        double acc[NM];
#pragma unroll
        for (int m = 0; m < NM; m++) acc[m] = 0.0;
        for (int q = agg_off[I] + lane; q < agg_off[I + 1]; q += 64) {
            const int a = agg_node[q];
#pragma unroll
            for (int i = 0; i < D; i++) {
                const size_t o = (size_t)a * D + i;
#pragma unroll
                for (int m = 0; m < NM; m++) acc[m] += Pn[o * NM + m] * v[o];
            }
        }
        wave_restrict<D, NM>(acc, I, lane, rc);
    }
}

// f = z - P yc
template <int D, int NM>
__global__ void k_sub_prolong(int nn, const double *Pn, const int *node_agg, const double *yc, const double *z,
                              double *f) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    const int I = node_agg[a];
#pragma unroll
    for (int i = 0; i < D; i++) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < NM; m++) s += Pn[((size_t)a * D + i) * NM + m] * yc[(size_t)I * NM + m];
        f[(size_t)a * D + i] = z[(size_t)a * D + i] - s;
    }
}

template <int D, int NM>
static int probe_impl(Ctx &c, const double *K64, TwoLevel &t, const double *z, double *f, double *y, int nprobe,
                      std::string &err) {
    const size_t n = (size_t)c.nn * D;
    double *kz, *rc, *yc;
    PDX_HIP(hipMalloc((void **)&kz, n * 8));
    PDX_HIP(hipMalloc((void **)&rc, (size_t)t.ld * 8));
    PDX_HIP(hipMalloc((void **)&yc, (size_t)t.ld * 8));
    const int gc = std::min(2048, (t.nc + 3) / 4);
    for (int p = 0; p < nprobe; p++) {
        const double *zp = z + p * n;
        double *fp = f + p * n, *yp = y + p * n;
        spmv_dot<D, double, double>(c, K64, nullptr, nullptr, zp, kz, (const double *)nullptr, nullptr, (const double *)nullptr,
                                    nullptr);
        k_restrict_agg<D, NM><<<RED_PARTS, 256, 0, c.stream>>>(t.na, t.agg_off, t.agg_node, kz, t.Pn, rc);
        k_coarse_mv<double><<<gc, 256, 0, c.stream>>>(t.nc, t.ld, t.Ainv, rc, yc);
        k_sub_prolong<D, NM><<<(c.nn + 255) / 256, 256, 0, c.stream>>>(c.nn, t.Pn, t.node_agg, yc, zp, fp);
        spmv_dot<D, double, double>(c, K64, nullptr, nullptr, fp, yp, (const double *)nullptr, nullptr, (const double *)nullptr,
                                    nullptr);
    }
    PDX_HIP(hipGetLastError());
    PDX_HIP(hipStreamSynchronize(c.stream));
    hipFree(kz);
    hipFree(rc);
    hipFree(yc);
    return PDX_OK;
}

// For nprobe free-DOF vectors z (device, [nprobe][nn*d], prescribed zero):
// f = (I - Q K) z and y = mask(K f), Q the two-level coarse correction of K.
int two_level_probes(Ctx &c, const double *K64, int agg, const double *z, double *f, double *y, int nprobe,
                     std::string &err) {
    TwoLevel *t;
    int rc = get_two_level(c, K64, agg, t, err);
    if (rc) return rc;
    return c.dim == 3 ? probe_impl<3, 6>(c, K64, *t, z, f, y, nprobe, err)
                      : probe_impl<2, 3>(c, K64, *t, z, f, y, nprobe, err);
}

// node-block Jacobi inverse of K (FP64, [nn][d*d], prescribed rows/columns zero)
int block_jacobi_inverse(Ctx &c, const double *K64, double *Minv, std::string &err) {
    int *status;
    PDX_HIP(hipMalloc((void **)&status, sizeof(int)));
    const int rc = c.dim == 3 ? build_precond<3, double>(c, K64, PDX_PREC_BLOCK_JACOBI, Minv, status, err)
                              : build_precond<2, double>(c, K64, PDX_PREC_BLOCK_JACOBI, Minv, status, err);
    hipFree(status);
    return rc;
}

double two_level_setup_ms(Ctx &c) {
    return c.two_level ? ((TwoLevel *)c.two_level)->setup_ms : 0.0;
}

}  // namespace pdx
