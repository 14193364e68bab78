// Consistent mass on the stiffness blocks and the acceleration-form Newmark
// driver (pdsolve.cpp:6092-6290, pdfemAssembleKNFR_CSRformat :6327-6397,
// updateDispVelo :6399-6422) on the device.
//
// K is constant for intact bonds, so K_eff = beta dt^2 K + M is formed once
// (the reference re-assembles it every step with the same values) and every
// step is: predictor U_N, F = F_ext - K U_N, K_eff a_{n+1} = F (BiCGStab warm
// started from a_n), then the displacement / velocity update.
#include <hip/hip_runtime.h>

#include <functional>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "../../include/pdx.h"
#include "device.hpp"

namespace pdx {

int to_f32(Ctx &c, const double *K64, float *&K32, std::string &err);

struct MassArgs {
    const double *x;
    const int *econ, *nen, *fe_elem, *fe_blk, *nfe_off, *nfe_item;
    const int *fam_node, *fam_blk, *fam_off, *eq;
    const double *fam_v;
    double rho;
    double *M;
    long long nblk;
    int nn, nfam;
};

__constant__ double c_mp[8] = {-1., 1., 1., -1., -1., 1., 1., -1.};
__constant__ double c_mq[8] = {-1., -1., 1., 1., -1., -1., 1., 1.};
__constant__ double c_mr[8] = {-1., -1., -1., -1., 1., 1., 1., 1.};

__device__ inline double det3m(const double J[9]) {
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

// element mass rows (pdfemEleQuad4N.cpp:361-387, pdfemEleBrick8N.cpp:554-582,
// pdfemEleTri3N.cpp:49-70, pdfemEleTetra4N.cpp:55-76), node-centric: one thread per node
// forms its row of every incident element once and adds it to the d diagonal slots of the
// node's blocks (the mass is the same for every DOF direction)
template <int D>
__global__ void k_mass_rows(MassArgs A) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.nn) return;
    const double g = 0.57735026918962573, gp[2] = {-g, g};
    const size_t nb = (size_t)A.nblk;
    for (int q = A.nfe_off[a]; q < A.nfe_off[a + 1]; q++) {
        const int item = A.nfe_item[q];
        const int fp = item >> 3, la = item & 7;
        const int e = A.fe_elem[fp];
        const int n = A.nen[e];
        double xN[8][3];
        for (int b = 0; b < n; b++)
            for (int c = 0; c < 3; c++) xN[b][c] = A.x[(size_t)A.econ[(size_t)e * 8 + b] * 3 + c];
        double row[8];
        for (int b = 0; b < 8; b++) row[b] = 0.0;
        if (D == 2 && n == 4) {
            for (int ip = 0; ip < 2; ip++)
                for (int iq = 0; iq < 2; iq++) {
                    const double p = gp[ip], qq = gp[iq];
                    double N[4] = {0.25 * (1 - p) * (1 - qq), 0.25 * (1 + p) * (1 - qq), 0.25 * (1 + p) * (1 + qq),
                                   0.25 * (1 - p) * (1 + qq)};
                    double dNdp[4] = {0.25 * (-1 + qq), 0.25 * (1 - qq), 0.25 * (1 + qq), 0.25 * (-1 - qq)};
                    double dNdq[4] = {0.25 * (-1 + p), 0.25 * (-1 - p), 0.25 * (1 + p), 0.25 * (1 - p)};
                    double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
                    for (int i = 0; i < 4; i++) {
                        J00 = J00 + dNdp[i] * xN[i][0];
                        J01 = J01 + dNdp[i] * xN[i][1];
                        J10 = J10 + dNdq[i] * xN[i][0];
                        J11 = J11 + dNdq[i] * xN[i][1];
                    }
                    const double detJ = J00 * J11 - J01 * J10;
                    for (int b = 0; b < 4; b++) row[b] = row[b] + ((N[la] * A.rho) * N[b]) * (detJ * 1.0 * 1.0);
                }
        } else if (D == 2) {
            double m3[9] = {1, xN[0][0], xN[0][1], 1, xN[1][0], xN[1][1], 1, xN[2][0], xN[2][1]};
            const double fac = A.rho * (0.5 * fabs(det3m(m3))) / 12.0;
            for (int b = 0; b < 3; b++) row[b] = b == la ? fac * 2 : fac;
        } else if (n == 8) {
            for (int ip = 0; ip < 2; ip++)
                for (int iq = 0; iq < 2; iq++)
                    for (int ir = 0; ir < 2; ir++) {
                        const double p = gp[ip], qq = gp[iq], r = gp[ir];
                        double N[8], dN[8][3], J[9];
                        for (int i = 0; i < 8; i++) {
                            N[i] = 1.0 / 8 * (1 + p * c_mp[i]) * (1 + qq * c_mq[i]) * (1 + r * c_mr[i]);
                            dN[i][0] = 1.0 / 8 * c_mp[i] * (1 + qq * c_mq[i]) * (1 + r * c_mr[i]);
                            dN[i][1] = 1.0 / 8 * c_mq[i] * (1 + p * c_mp[i]) * (1 + r * c_mr[i]);
                            dN[i][2] = 1.0 / 8 * c_mr[i] * (1 + p * c_mp[i]) * (1 + qq * c_mq[i]);
                        }
                        for (int i = 0; i < 9; i++) J[i] = 0.0;
                        for (int cc = 0; cc < 3; cc++)
                            for (int i = 0; i < 8; i++) {
                                J[cc] = dN[i][0] * xN[i][cc] + J[cc];
                                J[3 + cc] = dN[i][1] * xN[i][cc] + J[3 + cc];
                                J[6 + cc] = dN[i][2] * xN[i][cc] + J[6 + cc];
                            }
                        const double s = A.rho * det3m(J) * 1.0 * 1.0 * 1.0;
                        for (int b = 0; b < 8; b++) row[b] = row[b] + (N[la] * N[b]) * s;
                    }
        } else {
            double Jv[9];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) Jv[r * 3 + c] = xN[r + 1][c] - xN[0][c];
            const double fac = A.rho * (fabs(det3m(Jv)) / 6.0) / 20.0;
            for (int b = 0; b < 4; b++) row[b] = b == la ? fac * 2 : fac;
        }
        const int *blk = &A.fe_blk[(size_t)fp * 64 + la * 8];
        for (int b = 0; b < n; b++)
            for (int ii = 0; ii < D; ii++) {
                const size_t idx = (size_t)blk[b] * D * D + ii * D + ii;
                A.M[idx] = A.M[idx] + row[b];
            }
    }
}

// PD node mass rho * cd_v on the diagonal (pdsolve.cpp:4568-4595), after the FE terms
__global__ void k_mass_pd(MassArgs A, int D) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= A.nfam) return;
    const size_t b = (size_t)A.fam_blk[A.fam_off[f]];
    const size_t nb = (size_t)A.nblk;
    for (int i = 0; i < D; i++) {
        const size_t idx = b * D * D + i * D + i;
        A.M[idx] = A.M[idx] + A.fam_v[f] * A.rho;
    }
}

__global__ void k_keff(size_t n, const double *K, const double *M, double beta, double dt, double *Ke) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        Ke[i] = K[i] * beta * dt * dt + M[i];
}

// K_eff = beta dt^2 K + M of the context's K and consistent mass into Keff [dd * nblk]
int dev_form_keff(Ctx &c, double *Keff, std::string &err) {
    if (!c.K_valid)
        if (int rc = dev_assemble_K(c, err)) return rc;
    if (!c.Mk_valid)
        if (int rc = dev_mass_on_K(c, err)) return rc;
    k_keff<<<4096, 256, 0, c.stream>>>((size_t)c.dd * c.nblk, c.K, c.Mk, c.beta, c.dt, Keff);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

__global__ void k_fill_n(double *p, double v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// predictor (pdsolve.cpp:6333-6366): full DOF vector
__global__ void k_predictor(int nn, int D, const int *eq, const int *presc_set, const double *u, const double *v,
                            const double *a, const double *val, const double *ebc_dU, double dt, double beta,
                            double *UN) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * D) return;
    const int code = (t / D) * 3 + t % D;
    if (eq[code] >= 0) UN[t] = u[t] + dt * v[t] + dt * dt * (0.5 - beta) * a[t];
    else UN[t] = val[t] + ebc_dU[presc_set[code]];
}

// prescribed entries of the ramped essential BCs take their sets' current values
__global__ void k_ebc_values(int n, int D, const int *eq, const int *presc_set, const int *vary, const double *ebc_val,
                             double *val) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int code = (t / D) * 3 + t % D;
    if (eq[code] >= 0) return;
    const int s = presc_set[code];
    if (s >= 0 && vary[s]) val[t] = ebc_val[s];
}

// F = F_ext - K U_N on free rows (y = masked K U_N)
__global__ void k_newmark_rhs(int n, const double *y, const double *f_unit, int n_nbc, const double *nbc_val,
                              const int *eq, int D, double *F) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int code = (t / D) * 3 + t % D;
    if (eq[code] < 0) { F[t] = 0.0; return; }
    double fe = 0.0;
    for (int s = 0; s < n_nbc; s++) fe = fe + nbc_val[s] * f_unit[(size_t)s * n + t];
    F[t] = fe - y[t];
}

// updateDispVelo (pdsolve.cpp:6399-6422) on free DOFs
__global__ void k_newmark_update(int n, const int *eq, int D, double *u, double *v, const double *a,
                                 const double *a1, double dt, double beta, double gamma) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int code = (t / D) * 3 + t % D;
    if (eq[code] < 0) return;
    const double t1 = u[t] + dt * v[t] + dt * dt * ((0.5 - beta) * a[t] + beta * a1[t]);
    const double t2 = v[t] + dt * ((1 - gamma) * a[t] + gamma * a1[t]);
    u[t] = t1;
    v[t] = t2;
}

int dev_mass_on_K(Ctx &c, std::string &err) {
    if (!c.Mk) PDX_HIP(hipMalloc((void **)&c.Mk, sizeof(double) * std::max<long long>(c.dd * c.nblk, 1)));
    k_fill_n<<<4096, 256, 0, c.stream>>>(c.Mk, 0.0, (size_t)c.dd * c.nblk);
    MassArgs A;
    A.x = c.x; A.econ = c.econ; A.nen = c.nen; A.fe_elem = c.fe_elem; A.fe_blk = c.fe_blk; A.nfe_off = c.nfe_off;
    A.nfe_item = c.nfe_item; A.fam_node = c.fam_node; A.fam_blk = c.fam_blk; A.fam_off = c.fam_off; A.eq = c.eq;
    A.fam_v = c.fam_v; A.rho = c.rho; A.M = c.Mk; A.nblk = c.nblk; A.nn = c.nn; A.nfam = c.nfam;
    const int th = c.nn;
    if (c.nfe) {
        if (c.dim == 3) k_mass_rows<3><<<(th + 127) / 128, 128, 0, c.stream>>>(A);
        else k_mass_rows<2><<<(th + 127) / 128, 128, 0, c.stream>>>(A);
    }
    if (c.nfam) k_mass_pd<<<(c.nfam + 127) / 128, 128, 0, c.stream>>>(A, c.dim);
    PDX_HIP(hipGetLastError());
    c.Mk_valid = true;
    return PDX_OK;
}

// x = 2 a - ap (linear extrapolation of the acceleration over the last two steps)
__global__ void k_extrapolate(int n, const double *a, const double *ap, double *x) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = 2.0 * a[t] - ap[t];
}

// pdfemDynamicNewmarkSolver_CSRformat (pdsolve.cpp:6092-6290).  after_step (may be null) runs
// after every step with the DOF values of that step (free: u_{n+1}; prescribed: the values
// advanced after the step, as calGlobalNodeStresses sees them, :6248-6261); when it invalidates
// K (broken bonds), K and K_eff are assembled again before the next step.
int dev_newmark(Ctx &c, int nsteps, const pdx_solve_opts &o, std::vector<double> &dofval, double *u_h, double *v_h,
                double *a_h, pdx_solve_stats &stats, const std::function<int(const double *, std::string &)> &after_step,
                std::string &err) {
    int rc;
    if (!c.K_valid && (rc = dev_assemble_K(c, err))) return rc;
    if (!c.Mk_valid && (rc = dev_mass_on_K(c, err))) return rc;
    const Model &m = *c.host;
    const int D = c.dim;
    const size_t n = (size_t)c.nn * D, nk = (size_t)c.dd * c.nblk;
    double *Keff, *u, *v, *a, *a1, *ap, *UN, *y, *F, *val, *dU;
    float *Keff32 = nullptr;
    // every device buffer of the run, freed on every exit
    struct Owned {
        std::vector<void *> p;
        ~Owned() {
            for (void *q : p) hipFree(q);
        }
    } owned;
    auto dalloc = [&owned](void **pp, size_t bytes) {
        const hipError_t e = hipMalloc(pp, bytes);
        if (e == hipSuccess) owned.p.push_back(*pp);
        return e;
    };
    PDX_HIP(dalloc((void **)&Keff, nk * 8));
    PDX_HIP(dalloc((void **)&u, n * 8));
    PDX_HIP(dalloc((void **)&v, n * 8));
    PDX_HIP(dalloc((void **)&a, n * 8));
    PDX_HIP(dalloc((void **)&a1, n * 8));
    PDX_HIP(dalloc((void **)&ap, n * 8));  // a_{n-1}, for the extrapolated initial guess
    PDX_HIP(dalloc((void **)&UN, n * 8));
    PDX_HIP(dalloc((void **)&y, n * 8));
    PDX_HIP(dalloc((void **)&F, n * 8));
    PDX_HIP(dalloc((void **)&val, n * 8));
    PDX_HIP(dalloc((void **)&dU, std::max<size_t>(m.ebc.size(), 1) * 8));
    k_keff<<<4096, 256, 0, c.stream>>>(nk, c.K, c.Mk, c.beta, c.dt, Keff);
    // a learned solve predicts its smoother for this K_eff with the attached network (with
    // fallback on, a missing network leaves the solve to fall back); the smoother and the
    // network's inputs of this K_eff are forgotten when it is freed
    const bool learned = o.precond == PDX_PREC_LEARNED;
    struct ForgetKeff {
        Ctx &c;
        const double *A;
        ~ForgetKeff() { forget_learned_operator(c, A); }
    } forget{c, Keff};
    if (learned) {
        rc = dev_learned_smoother(c, Keff, err);
        if (rc && !o.fallback) return rc;
        rc = PDX_OK;
    }
    if (o.precision == PDX_FP32) {
        rc = to_f32(c, Keff, Keff32, err);
        if (Keff32) owned.p.push_back(Keff32);
        if (rc) return rc;
    }
    // mirrored storage of K_eff for the SpMVs, when its mirrored blocks are exact transposes
    // (K's are by construction; the consistent mass's are checked with them)
    Mirrored mir;
    double *own64 = nullptr;
    float *own32 = nullptr;
    if (o.mirrored && c.dim > 1) {
        bool exact = false;
        if ((rc = check_mirrored(c, Keff, exact, err))) return rc;
        if (exact) {
            const size_t no = (size_t)c.dd * std::max<long long>(c.n_own, 1);
            PDX_HIP(dalloc((void **)&own64, no * 8));
            if (o.precision == PDX_FP32) PDX_HIP(dalloc((void **)&own32, no * 4));
            if ((rc = gather_own(c, Keff, own64, own32, err))) return rc;
            mir.K64 = own64;
            mir.K32 = own32;
        }
    }
    // u from DOF values (free), v = a = 0 (pdsolve.cpp:6207-6220)
    std::vector<double> uh(n, 0.0);
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < D; i++)
            if (m.eq[(size_t)k * 3 + i] >= 0) uh[(size_t)k * D + i] = dofval[(size_t)k * D + i];
    PDX_HIP(hipMemcpyAsync(u, uh.data(), n * 8, hipMemcpyHostToDevice, c.stream));
    PDX_HIP(hipMemsetAsync(v, 0, n * 8, c.stream));
    PDX_HIP(hipMemsetAsync(a, 0, n * 8, c.stream));
    PDX_HIP(hipMemsetAsync(a1, 0, n * 8, c.stream));
    PDX_HIP(hipMemsetAsync(ap, 0, n * 8, c.stream));
    std::vector<double> dUh(std::max<size_t>(m.ebc.size(), 1), 0.0);
    for (size_t s = 0; s < m.ebc.size(); s++) dUh[s] = c.ebc_vary[s] ? c.dt * c.ebc_vel[s] : 0.0;
    PDX_HIP(hipMemcpyAsync(dU, dUh.data(), dUh.size() * 8, hipMemcpyHostToDevice, c.stream));
    double *nbcv_d, *ebcv_d;
    int *ebc_vary_d;
    const size_t n_ebc = m.ebc.size();
    PDX_HIP(dalloc((void **)&nbcv_d, std::max(c.n_nbc, 1) * 8));
    PDX_HIP(dalloc((void **)&ebcv_d, std::max<size_t>(n_ebc, 1) * 8));
    PDX_HIP(dalloc((void **)&ebc_vary_d, std::max<size_t>(n_ebc, 1) * 4));
    bool any_vary = false;
    if (n_ebc) {
        PDX_HIP(hipMemcpyAsync(ebc_vary_d, c.ebc_vary.data(), n_ebc * 4, hipMemcpyHostToDevice, c.stream));
        for (size_t s = 0; s < n_ebc; s++) any_vary = any_vary || c.ebc_vary[s];
    }
    // DOF values (prescribed entries are what the predictor reads) stay on the device; the
    // ramped prescribed entries are advanced there after every step
    PDX_HIP(hipMemcpyAsync(val, dofval.data(), n * 8, hipMemcpyHostToDevice, c.stream));
    stats = pdx_solve_stats{};
    auto t0 = std::chrono::steady_clock::now();
    pdx_solve_opts oo = o;
    oo.warm_start = 1;
    const int th = (int)n;
    for (int step = 0; step < nsteps; step++) {
        if (c.n_nbc) PDX_HIP(hipMemcpyAsync(nbcv_d, c.nbc_val.data(), c.n_nbc * 8, hipMemcpyHostToDevice, c.stream));
        k_predictor<<<(th + 255) / 256, 256, 0, c.stream>>>(c.nn, D, c.eq, c.presc_set, u, v, a, val, dU, c.dt,
                                                            c.beta, UN);
        if ((rc = dev_spmv_K(c, c.K, UN, y))) { err = "spmv failed"; return rc; }
        k_newmark_rhs<<<(th + 255) / 256, 256, 0, c.stream>>>(th, y, c.f_unit, c.n_nbc, nbcv_d, c.eq, D, F);
        // initial guess a_{n+1} <- a_n (warm start), or 2 a_n - a_{n-1} (warm_start 2, from the
        // second step on)
        if (o.warm_start >= 2 && step > 0)
            k_extrapolate<<<(th + 255) / 256, 256, 0, c.stream>>>(th, a, ap, a1);
        else
            PDX_HIP(hipMemcpyAsync(a1, a, n * 8, hipMemcpyDeviceToDevice, c.stream));
        pdx_solve_stats st{};
        rc = dev_solve_mat(c, Keff, Keff32, oo, F, a1, st, err, &mir);
        stats.iterations += st.iterations;
        stats.refinements += st.refinements;
        stats.true_relres = std::max(stats.true_relres, st.true_relres);
        stats.mirrored_used = st.mirrored_used;
        // the preconditioner and precision of the last step's solve; attempts: the most any step took
        stats.precond_used = st.precond_used;
        stats.precision_used = st.precision_used;
        stats.attempts = std::max(stats.attempts, st.attempts);
        if (rc) break;
        k_newmark_update<<<(th + 255) / 256, 256, 0, c.stream>>>(th, c.eq, D, u, v, a, a1, c.dt, c.beta, c.gamma);
        std::swap(ap, a);  // ap <- a_n
        std::swap(a, a1);  // a <- a_{n+1}
        // advance BCs (pdsolve.cpp:6248-6253); the displacement is stored once, after the last
        // step (every host array below was last read by copies that the solve's
        // synchronisations have completed)
        for (size_t s = 0; s < m.ebc.size(); s++)
            if (c.ebc_vary[s]) {
                c.ebc_val[s] = c.ebc_val[s] + c.dt * c.ebc_vel[s];
                for (int k : m.ebc[s].nid) dofval[(size_t)k * D + m.ebc[s].dof] = c.ebc_val[s];
            }
        if (any_vary) {
            PDX_HIP(hipMemcpyAsync(ebcv_d, c.ebc_val.data(), n_ebc * 8, hipMemcpyHostToDevice, c.stream));
            k_ebc_values<<<(th + 255) / 256, 256, 0, c.stream>>>(th, D, c.eq, c.presc_set, ebc_vary_d, ebcv_d, val);
        }
        for (int s = 0; s < c.n_nbc; s++)
            if (c.nbc_vary[s]) c.nbc_val[s] = c.nbc_val[s] + c.dt * c.nbc_vel[s];
        if (after_step) {
            if ((rc = dev_merge_dofvals(c, u, val, UN))) { err = "merge of the DOF values failed"; break; }
            if ((rc = after_step(UN, err))) break;
            if (!c.K_valid) {
                // bonds broke: K, K_eff and its copies for the next step
                if ((rc = dev_assemble_K(c, err))) break;
                k_keff<<<4096, 256, 0, c.stream>>>(nk, c.K, c.Mk, c.beta, c.dt, Keff);
                if (learned) {
                    rc = dev_learned_smoother(c, Keff, err);
                    if (rc && !o.fallback) break;
                    rc = PDX_OK;
                }
                if (o.precision == PDX_FP32 && (rc = to_f32(c, Keff, Keff32, err))) break;
                if (mir.K64 && (rc = gather_own(c, Keff, own64, own32, err))) break;
            }
        }
    }
    // store the displacement of the last completed step in the DOF values
    // (storeDisplacementResult, pdsolve.cpp:5656-5672)
    PDX_HIP(hipMemcpyAsync(uh.data(), u, n * 8, hipMemcpyDeviceToHost, c.stream));
    PDX_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < D; i++)
            if (m.presc_set[(size_t)k * 3 + i] < 0) dofval[(size_t)k * D + i] = uh[(size_t)k * D + i];
    stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats.converged = rc == PDX_OK;
    std::vector<double> vh(n), ah(n);
    hipMemcpyAsync(vh.data(), v, n * 8, hipMemcpyDeviceToHost, c.stream);
    hipMemcpyAsync(ah.data(), a, n * 8, hipMemcpyDeviceToHost, c.stream);
    hipStreamSynchronize(c.stream);
    for (int k = 0; k < c.nn; k++)
        for (int i = 0; i < D; i++) {
            int r = m.eq[(size_t)k * 3 + i];
            if (r < 0) continue;
            if (u_h) u_h[r] = uh[(size_t)k * D + i];
            if (v_h) v_h[r] = vh[(size_t)k * D + i];
            if (a_h) a_h[r] = ah[(size_t)k * D + i];
        }
    return rc;
}

}  // namespace pdx

using namespace pdx;

extern "C" {

int pdx_mass_on_K(pdx_ctx *pc, double *Mv) {
    if (!pc || !Mv) return set_error(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    std::string err;
    int rc = dev_mass_on_K(c, err);
    if (rc) return set_error(rc, err);
    double *dk;
    if (hipMalloc((void **)&dk, std::max<long long>(c.nnz, 1) * 8) != hipSuccess) return set_error(PDX_ERR_HIP, "alloc");
    dev_extract_csr(c, c.Mk, dk);
    hipMemcpyAsync(Mv, dk, c.nnz * 8, hipMemcpyDeviceToHost, c.stream);
    hipStreamSynchronize(c.stream);
    hipFree(dk);
    return hipGetLastError() == hipSuccess ? PDX_OK : set_error(PDX_ERR_HIP, "HIP error in mass assembly");
}

}  // extern "C"
