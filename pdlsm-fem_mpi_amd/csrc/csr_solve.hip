// The sparse-solve seam of the reference, calMatrixOperations::
// cluster_PARDISO_64Solver (calMatrixOperations.cpp:378-498): solve A x = b
// for a zero-based int64 CSR matrix.  The CSR is uploaded as a node-block
// matrix with 1 x 1 blocks, so the same Krylov kernels (BiCGStab, Jacobi,
// FP32 inner iterations + FP64 refinement) serve both operators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>

#include "../../include/pdx.h"
#include "device.hpp"

namespace pdx {

int to_f32(Ctx &c, const double *K64, float *&K32, std::string &err);
int set_error(int rc, const std::string &msg);

static int csr_solve(int device, long long n, const long long *ia, const long long *ja, const double *a,
                     const double *b, double *x, const pdx_solve_opts &o, pdx_solve_stats &st, std::string &err) {
    if (n <= 0 || !ia || !ja || !a || !b || !x) { err = "null or empty system"; return PDX_ERR_ARG; }
    const long long nnz = ia[n];
    if (n > 0x7fffffffLL || nnz > 0x7fffffffLL) { err = "system exceeds 32-bit indexing"; return PDX_ERR_LIMIT; }
    for (long long r = 0; r < n; r++)
        if (ia[r + 1] < ia[r]) { err = "ia not monotone"; return PDX_ERR_ARG; }
    for (long long k = 0; k < nnz; k++)
        if (ja[k] < 0 || ja[k] >= n) { err = "column index out of range"; return PDX_ERR_ARG; }
    PDX_HIP(hipSetDevice(device));
    Ctx c;
    c.device = device;
    PDX_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.dim = 1;
    c.dd = 1;
    c.nn = (int)n;
    c.nblk = nnz;
    std::vector<int> off(n + 1), col(std::max<long long>(nnz, 1)), eq(3 * n, -1), rs, rl;
    for (long long r = 0; r <= n; r++) off[r] = (int)ia[r];
    for (long long k = 0; k < nnz; k++) col[k] = (int)ja[k];
    for (long long r = 0; r < n; r++) {
        eq[3 * r] = (int)r;
        (off[r + 1] - off[r] > 48 ? rl : rs).push_back((int)r);
    }
    // rows must be sorted for the diagonal lookup of the Jacobi preconditioner
    for (long long r = 0; r < n; r++) std::sort(col.begin() + off[r], col.begin() + off[r + 1]);
    std::vector<double> av(std::max<long long>(nnz, 1));
    {
        // permute values with the same sort (stable pairs)
        std::vector<std::pair<int, double>> tmp;
        for (long long r = 0; r < n; r++) {
            tmp.clear();
            for (long long k = ia[r]; k < ia[r + 1]; k++) tmp.push_back({(int)ja[k], a[k]});
            std::stable_sort(tmp.begin(), tmp.end(), [](auto &p, auto &q) { return p.first < q.first; });
            for (size_t t = 0; t < tmp.size(); t++) av[ia[r] + t] = tmp[t].second;
        }
    }
    double *K = nullptr, *db = nullptr, *dx = nullptr;
    float *K32 = nullptr;
    auto up_i = [&](int *&d, const std::vector<int> &h) -> int {
        PDX_HIP(hipMalloc((void **)&d, std::max<size_t>(h.size(), 1) * sizeof(int)));
        if (!h.empty()) PDX_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
        c.owned.push_back(d);
        return PDX_OK;
    };
    int rc;
    if ((rc = up_i(c.adj_off, off)) || (rc = up_i(c.adj, col)) || (rc = up_i(c.eq, eq)) ||
        (rc = up_i(c.rows_short, rs)) || (rc = up_i(c.rows_long, rl)))
        return rc;
    c.n_short = (int)rs.size();
    c.n_long = (int)rl.size();
    PDX_HIP(hipMalloc((void **)&K, av.size() * 8));
    PDX_HIP(hipMalloc((void **)&db, n * 8));
    PDX_HIP(hipMalloc((void **)&dx, n * 8));
    PDX_HIP(hipMemcpy(K, av.data(), av.size() * 8, hipMemcpyHostToDevice));
    PDX_HIP(hipMemcpy(db, b, n * 8, hipMemcpyHostToDevice));
    PDX_HIP(hipMemcpy(dx, x, n * 8, hipMemcpyHostToDevice));
    if (o.precision == PDX_FP32 && (rc = to_f32(c, K, K32, err))) return rc;
    rc = dev_solve_mat(c, K, K32, o, db, dx, st, err);
    PDX_HIP(hipMemcpy(x, dx, n * 8, hipMemcpyDeviceToHost));
    hipFree(K);
    hipFree(db);
    hipFree(dx);
    if (K32) hipFree(K32);
    free_ilu0(c);
    free_krylov_ws(c);
    if (c.red) hipFree(c.red);
    c.red = nullptr;
    for (void *p : c.owned) hipFree(p);
    c.owned.clear();
    hipStreamDestroy(c.stream);
    return rc;
}

}  // namespace pdx

extern "C" int pdx_solve_csr(int device, long long n, const long long *ia, const long long *ja, const double *a,
                             const double *b, double *x, const pdx_solve_opts *op, pdx_solve_stats *st) {
    pdx_solve_opts o{};
    if (op) o = *op;
    else {
        o.krylov = PDX_KRYLOV_BICGSTAB; o.precond = PDX_PREC_JACOBI; o.precision = PDX_FP64;
        o.tol = 1e-8; o.max_iter = 20000; o.inner_tol = 1e-4; o.warm_start = 0;
    }
    std::string err;
    pdx_solve_stats s{};
    int rc = pdx::csr_solve(device, n, ia, ja, a, b, x, o, s, err);
    if (st) *st = s;
    return rc ? pdx::set_error(rc, err) : PDX_OK;
}
