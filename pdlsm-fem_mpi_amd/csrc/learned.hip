// Learned smoother of the two-level preconditioner (PDX_PREC_LEARNED).
//
// A message-passing network on the node graph of K predicts, for every node a,
// a bounded correction of the node-block Jacobi smoother:
//   M_a = (I + 0.5 tanh(o_a)) B_a,   B_a = (K_aa)^-1 on the free DOFs,
// used as S in  z = Q v + S (v - K Q v)  (krylov.hip).  With zero output
// weights M = B, i.e. plain two-level block Jacobi.
//
// Network (width F = 32, L = 4 layers), inputs from the Jacobi-scaled matrix
// A = S K S (S = diag(K_ii^-1/2) on free DOFs):
//   x_a  (8)   node type one-hot (FE / interface / PD), free DOFs / d,
//              log(1 + degree) / 4, ||A_aa - I||_F, mean off-diagonal ||A_ab||_F, 1
//   c_ab       ||A_ab||_F / sum_{b' != a} ||A_ab'||_F  (0 on the diagonal)
//   h0 = relu(W_in x + b_in)
//   h_{l+1} = relu(W_l [h_l ; sum_b c_ab h_l,b] + b_l) + h_l
//   o_a = W_o [h_L,a ; vec(A_aa)] + b_o           (d*d outputs)
// Loss over probes z_p (Gaussian, free DOFs):
//   f_p = (I - Q K) z_p,  y_p = K f_p,  L = sum_p ||f_p - M y_p||^2 / sum_p ||f_p||^2
// i.e. the energy left by one smoothing step after the coarse correction.
// Training runs forward, backward and Adam with the kernels below (FP32
// activations and weights, FP64 partial sums in fixed order: deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "../../include/pdx.h"
#include "device.hpp"
#include "reduce.hpp"

namespace pdx {

int two_level_probes(Ctx &c, const double *K64, int agg, const double *z, double *f, double *y, int nprobe,
                     std::string &err);
int block_jacobi_inverse(Ctx &c, const double *K64, double *Minv, std::string &err);

constexpr int GF = 32;       // hidden width
constexpr int GIN = 8;       // node input features
constexpr int GL = 4;        // message-passing layers
constexpr int NPARTW = 256;  // blocks of the weight-gradient reductions
constexpr int WACC = 9;      // weights per thread in those reductions (256 * 9 >= 32 * 65)

struct GnnLayout {
    int D = 2, DD = 4, nhead = 0;
    size_t w_in = 0, b_in = 0, w_l[GL] = {}, b_l[GL] = {}, w_o = 0, b_o = 0, total = 0;
    explicit GnnLayout(int d = 2) {
        D = d;
        DD = d * d;
        nhead = GF + DD;
        size_t o = 0;
        w_in = o; o += (size_t)GF * GIN;
        b_in = o; o += GF;
        for (int l = 0; l < GL; l++) {
            w_l[l] = o; o += (size_t)GF * 2 * GF;
            b_l[l] = o; o += GF;
        }
        w_o = o; o += (size_t)DD * nhead;
        b_o = o; o += DD;
        total = o;
    }
};

}  // namespace pdx

struct pdx_gnn {
    int dim = 2, device = 0;
    pdx::GnnLayout lay;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // weights, gradient, Adam moments (device)
    long long adam_t = 0;
};

namespace pdx {

// ------------------------------------------------------------------------
// inputs
// ------------------------------------------------------------------------
template <int D>
__global__ void k_gnn_scale(int nn, const int *adj_off, const int *adj, const double *K, const int *eq, double *S) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    int self = -1;
    for (int b = adj_off[a]; b < adj_off[a + 1]; b++)
        if (adj[b] == a) { self = b; break; }
#pragma unroll
    for (int i = 0; i < D; i++) {
        const double kii = self >= 0 ? K[(size_t)self * D * D + i * D + i] : 0.0;
        S[(size_t)a * D + i] = (eq[a * 3 + i] >= 0 && kii > 0.0) ? 1.0 / sqrt(kii) : 0.0;
    }
}

template <int D>
__global__ void k_gnn_inputs(int nn, const int *adj_off, const int *adj, const double *K, const int *eq,
                             const int *ntype, const double *S, float *X, float *cw, float *Aaa) {
    constexpr int DD = D * D;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double rowsum = 0.0, selfdev = 0.0;
    int nfree = 0;
#pragma unroll
    for (int i = 0; i < D; i++) nfree += eq[a * 3 + i] >= 0;
    const int deg = adj_off[a + 1] - adj_off[a];
#pragma unroll
    for (int e = 0; e < DD; e++) Aaa[(size_t)a * DD + e] = 0.f;
    for (int b = adj_off[a]; b < adj_off[a + 1]; b++) {
        const int c = adj[b];
        double fn = 0.0;
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) {
                const double v = K[(size_t)b * DD + i * D + j] * S[(size_t)a * D + i] * S[(size_t)c * D + j];
                fn += v * v;
                if (c == a) {
                    Aaa[(size_t)a * DD + i * D + j] = float(v);
                    const double dv = v - ((i == j && S[(size_t)a * D + i] != 0.0) ? 1.0 : 0.0);
                    selfdev += dv * dv;
                }
            }
        fn = sqrt(fn);
        cw[b] = c == a ? 0.f : float(fn);
        if (c != a) rowsum += fn;
    }
    for (int b = adj_off[a]; b < adj_off[a + 1]; b++) cw[b] = rowsum > 0.0 ? float(cw[b] / rowsum) : 0.f;
    float *x = X + (size_t)a * GIN;
    const int t = ntype[a];
    x[0] = t == 0; x[1] = t == 1; x[2] = t == 2;
    x[3] = float(nfree) / D;
    x[4] = logf(1.f + deg) / 4.f;
    x[5] = float(sqrt(selfdev));
    x[6] = deg > 1 ? float(rowsum / (deg - 1)) : 0.f;
    x[7] = 1.f;
}

// ------------------------------------------------------------------------
// forward
// ------------------------------------------------------------------------
// out = relu(W in + b) (+ res); pre = W in + b.  One thread per (node, output)
__global__ void k_dense(int nn, int nin, const float *in, const float *W, const float *b, float *pre, float *out,
                        const float *res) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * GF) return;
    const int a = t / GF, f = t % GF;
    const float *x = in + (size_t)a * nin;
    const float *w = W + (size_t)f * nin;
    float s = b[f];
    for (int k = 0; k < nin; k++) s += w[k] * x[k];
    pre[t] = s;
    float o = s > 0.f ? s : 0.f;
    if (res) o += res[t];
    out[t] = o;
}

// HM = [H ; sum_b c_ab H_b]
__global__ void k_aggregate(int nn, const int *adj_off, const int *adj, const float *cw, const float *H, float *HM) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * GF) return;
    const int a = t / GF, f = t % GF;
    float s = 0.f;
    for (int b = adj_off[a]; b < adj_off[a + 1]; b++) s += cw[b] * H[(size_t)adj[b] * GF + f];
    HM[(size_t)a * 2 * GF + f] = H[t];
    HM[(size_t)a * 2 * GF + GF + f] = s;
}

// O = W_o [H ; A_aa] + b_o; one thread per (node, output)
template <int D>
__global__ void k_head(int nn, const float *H, const float *Aaa, const float *Wo, const float *bo, float *O) {
    constexpr int DD = D * D, NH = GF + DD;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * DD) return;
    const int a = t / DD, k = t % DD;
    const float *w = Wo + (size_t)k * NH;
    float s = bo[k];
    for (int q = 0; q < GF; q++) s += w[q] * H[(size_t)a * GF + q];
#pragma unroll
    for (int q = 0; q < DD; q++) s += w[GF + q] * Aaa[(size_t)a * DD + q];
    O[t] = s;
}

// M_a = (I + 0.5 tanh(O_a)) B_a, masked to free rows/columns
template <int D, typename V>
__device__ inline void smoother_block(int a, const float *O, const float *B, const int *eq, V M[D * D]) {
    constexpr int DD = D * D;
    float T[DD];
#pragma unroll
    for (int e = 0; e < DD; e++) T[e] = 0.5f * tanhf(O[(size_t)a * DD + e]);
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            float s = B[(size_t)a * DD + i * D + j];
#pragma unroll
            for (int k = 0; k < D; k++) s += T[i * D + k] * B[(size_t)a * DD + k * D + j];
            const bool live = eq[a * 3 + i] >= 0 && eq[a * 3 + j] >= 0;
            M[i * D + j] = live ? V(s) : V(0);
        }
}

template <int D>
__global__ void k_smoother(int nn, const float *O, const float *B, const int *eq, double *M) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    double m[D * D];
    smoother_block<D, double>(a, O, B, eq, m);
#pragma unroll
    for (int e = 0; e < D * D; e++) M[(size_t)a * D * D + e] = m[e];
}

// R_p = f_p - M y_p per node; partial sums of ||R||^2 (RED_PARTS blocks, grid-stride over nodes)
template <int D>
__global__ void __launch_bounds__(256) k_gnn_resid(int nn, int P, const float *O, const float *B, const int *eq,
                                                   const float *f, const float *y, float *R, double *part) {
    __shared__ double sm[17];
    const size_t n = (size_t)nn * D;
    double acc = 0.0;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < nn; a += gridDim.x * blockDim.x) {
        float M[D * D];
        smoother_block<D, float>(a, O, B, eq, M);
        for (int p = 0; p < P; p++) {
            float yv[D];
#pragma unroll
            for (int j = 0; j < D; j++) yv[j] = y[p * n + (size_t)a * D + j];
#pragma unroll
            for (int i = 0; i < D; i++) {
                float s = f[p * n + (size_t)a * D + i];
#pragma unroll
                for (int j = 0; j < D; j++) s -= M[i * D + j] * yv[j];
                R[p * n + (size_t)a * D + i] = s;
                acc += double(s) * double(s);
            }
        }
    }
    acc = block_sum(acc, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// ------------------------------------------------------------------------
// backward
// ------------------------------------------------------------------------
// dO_a = (dM_a B_a^T) * 0.5 (1 - tanh^2 O_a),  dM_a = -scale sum_p R_p,a y_p,a^T
template <int D>
__global__ void k_gnn_dO(int nn, int P, const float *O, const float *B, const float *R, const float *y, float scale,
                         float *dO) {
    constexpr int DD = D * D;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nn) return;
    const size_t n = (size_t)nn * D;
    float dM[DD];
#pragma unroll
    for (int e = 0; e < DD; e++) dM[e] = 0.f;
    for (int p = 0; p < P; p++)
#pragma unroll
        for (int i = 0; i < D; i++) {
            const float r = R[p * n + (size_t)a * D + i];
#pragma unroll
            for (int j = 0; j < D; j++) dM[i * D + j] -= scale * r * y[p * n + (size_t)a * D + j];
        }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int k = 0; k < D; k++) {
            float s = 0.f;
#pragma unroll
            for (int j = 0; j < D; j++) s += dM[i * D + j] * B[(size_t)a * DD + k * D + j];
            const float th = tanhf(O[(size_t)a * DD + i * D + k]);
            dO[(size_t)a * DD + i * D + k] = s * 0.5f * (1.f - th * th);
        }
}

// dH_a = W_o[:, 0:F]^T dO_a
template <int D>
__global__ void k_head_back(int nn, const float *dO, const float *Wo, float *dH) {
    constexpr int DD = D * D, NH = GF + DD;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * GF) return;
    const int a = t / GF, q = t % GF;
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < DD; k++) s += Wo[(size_t)k * NH + q] * dO[(size_t)a * DD + k];
    dH[t] = s;
}

// layer backward: dU = dHout * 1[U > 0]; [dh ; dm] = W^T dU; dHin = dHout + dh; dM = dm
__global__ void k_layer_back(int nn, const float *dHout, const float *U, const float *W, float *dHin, float *dM,
                             float *dU) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * GF) return;
    const int a = t / GF, q = t % GF;
    float sh = 0.f, sm = 0.f;
    for (int o = 0; o < GF; o++) {
        const size_t ao = (size_t)a * GF + o;
        const float du = U[ao] > 0.f ? dHout[ao] : 0.f;
        sh += W[(size_t)o * 2 * GF + q] * du;
        sm += W[(size_t)o * 2 * GF + GF + q] * du;
    }
    dU[t] = U[t] > 0.f ? dHout[t] : 0.f;
    dHin[t] = dHout[t] + sh;
    dM[t] = sm;
}

// dH_b += sum over edges e = (a -> b) of c_e dM_a (transpose of the aggregation)
__global__ void k_aggr_back(int nn, const int *csc_off, const int *csc_edge, const int *csc_row, const float *cw,
                            const float *dM, float *dH) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * GF) return;
    const int b = t / GF, q = t % GF;
    float s = 0.f;
    for (int k = csc_off[b]; k < csc_off[b + 1]; k++) s += cw[csc_edge[k]] * dM[(size_t)csc_row[k] * GF + q];
    dH[t] += s;
}

__global__ void k_relu_back(int n, const float *dH, const float *U, float *dU) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dU[t] = U[t] > 0.f ? dH[t] : 0.f;
}

// partial sums of dW = sum_a dU_a [in1_a ; in2_a]^T and db = sum_a dU_a (nout outputs)
__global__ void __launch_bounds__(256) k_wgrad(int nn, int nout, const float *dU, const float *in1, int n1,
                                               const float *in2, int n2, double *part) {
    const int nin = n1 + n2, NW = nout * (nin + 1);
    __shared__ float sin_[2 * GF + 16];
    __shared__ float sdu[GF];
    double acc[WACC];
#pragma unroll
    for (int q = 0; q < WACC; q++) acc[q] = 0.0;
    const int chunk = (nn + gridDim.x - 1) / gridDim.x;
    const int a0 = blockIdx.x * chunk, a1 = min(nn, a0 + chunk);
    for (int a = a0; a < a1; a++) {
        __syncthreads();
        if (threadIdx.x < n1) sin_[threadIdx.x] = in1[(size_t)a * n1 + threadIdx.x];
        if (threadIdx.x < n2) sin_[n1 + threadIdx.x] = in2[(size_t)a * n2 + threadIdx.x];
        if (threadIdx.x < nout) sdu[threadIdx.x] = dU[(size_t)a * nout + threadIdx.x];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < WACC; q++) {
            const int w = threadIdx.x + q * 256;
            if (w < NW) {
                const int r = w / (nin + 1), j = w % (nin + 1);
                acc[q] += double(sdu[r]) * (j < nin ? double(sin_[j]) : 1.0);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < WACC; q++) {
        const int w = threadIdx.x + q * 256;
        if (w < NW) part[(size_t)blockIdx.x * NW + w] = acc[q];
    }
}

// sum the partials in block order into the flat gradient (weights row-major, then biases)
__global__ void k_wgrad_reduce(int nparts, int nout, int nin, const double *part, float *gW, float *gb) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    const int NW = nout * (nin + 1);
    if (w >= NW) return;
    double s = 0.0;
    for (int p = 0; p < nparts; p++) s += part[(size_t)p * NW + w];
    const int r = w / (nin + 1), j = w % (nin + 1);
    if (j < nin) gW[(size_t)r * nin + j] = float(s);
    else gb[r] = float(s);
}

__global__ void k_adam(size_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void k_d2f(size_t n, const double *a, float *b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = float(a[i]);
}

// ------------------------------------------------------------------------
// host: per-context work space
// ------------------------------------------------------------------------
struct GnnWork {
    long long gen = -1;
    const double *op = nullptr;  // the operator the inputs were formed from
    int nn = 0, D = 2, P = 0, agg = 0;
    long long nblk = 0;
    float *X = nullptr, *cw = nullptr, *Aaa = nullptr, *B = nullptr;
    int *csc_off = nullptr, *csc_edge = nullptr, *csc_row = nullptr;
    float *H[GL + 1] = {}, *HM[GL] = {}, *U0 = nullptr, *U[GL] = {}, *O = nullptr;
    __bf16 *Hb[GL + 1] = {};  // BF16 activations of the MFMA inference (allocated on first use)
    float *f = nullptr, *y = nullptr, *R = nullptr;
    float *dO = nullptr, *dH = nullptr, *dH2 = nullptr, *dM = nullptr, *dU = nullptr;
    double *part = nullptr, *lpart = nullptr;
    double fnorm2 = 0.0;
    std::vector<void *> owned, probe_owned;
    ~GnnWork() {
        for (void *p : owned) hipFree(p);
        for (void *p : probe_owned) hipFree(p);
    }
};

template <typename P>
static int walloc(std::vector<void *> &own, P *&p, size_t count, std::string &err) {
    PDX_HIP(hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(P)));
    own.push_back((void *)p);
    return PDX_OK;
}

void free_gnn_work(Ctx &c) {
    delete (GnnWork *)c.gnn_work;
    c.gnn_work = nullptr;
}

// the operator of the context's network work (c.gnn_op): K, or the Newmark operator formed for
// the current K
static int gnn_operator(Ctx &c, const double *&A, std::string &err) {
    if (!c.K_valid)
        if (int rc = dev_assemble_K(c, err)) return rc;
    if (c.gnn_op == 0) {
        A = c.K;
        return PDX_OK;
    }
    if (!c.gnn_keff) PDX_HIP(hipMalloc((void **)&c.gnn_keff, (size_t)c.dd * c.nblk * sizeof(double)));
    if (c.gnn_keff_gen != c.K_gen) {
        if (int rc = dev_form_keff(c, c.gnn_keff, err)) return rc;
        c.gnn_keff_gen = c.K_gen;
    }
    A = c.gnn_keff;
    return PDX_OK;
}

// features, block Jacobi inverse and activations for operator A (K or a K_eff of the current K)
static int gnn_work_for(Ctx &c, const double *A, GnnWork *&out, std::string &err) {
    if (c.dim != 2 && c.dim != 3) { err = "the learned smoother needs a 2D or 3D model"; return PDX_ERR_ARG; }
    GnnWork *g = (GnnWork *)c.gnn_work;
    if (g && g->gen == c.K_gen && g->op == A) {
        out = g;
        return PDX_OK;
    }
    delete g;
    c.gnn_work = nullptr;
    g = new GnnWork();
    const int nn = c.nn, D = c.dim, DD = D * D;
    g->nn = nn;
    g->D = D;
    g->nblk = c.nblk;
    int rc;
    auto &own = g->owned;
#define GA(p, count)                                     \
    do {                                                 \
        if ((rc = walloc(own, p, (count), err))) {       \
            delete g;                                    \
            return rc;                                   \
        }                                                \
    } while (0)
    GA(g->X, (size_t)nn * GIN);
    GA(g->cw, (size_t)c.nblk);
    GA(g->Aaa, (size_t)nn * DD);
    GA(g->B, (size_t)nn * DD);
    for (int l = 0; l <= GL; l++) GA(g->H[l], (size_t)nn * GF);
    for (int l = 0; l < GL; l++) {
        GA(g->HM[l], (size_t)nn * 2 * GF);
        GA(g->U[l], (size_t)nn * GF);
    }
    GA(g->U0, (size_t)nn * GF);
    GA(g->O, (size_t)nn * DD);
    GA(g->dO, (size_t)nn * DD);
    GA(g->dH, (size_t)nn * GF);
    GA(g->dH2, (size_t)nn * GF);
    GA(g->dM, (size_t)nn * GF);
    GA(g->dU, (size_t)nn * GF);
    GA(g->part, (size_t)NPARTW * GF * (2 * GF + 1));
    GA(g->lpart, (size_t)NPART);
    // transpose (column -> edges) index of the node graph
    const Model &m = *c.host;
    {
        std::vector<int> off(nn + 1, 0), edge(c.nblk), row(c.nblk);
        for (int a = 0; a < nn; a++)
            for (int e = m.adj_off[a]; e < m.adj_off[a + 1]; e++) off[m.adj[e] + 1]++;
        for (int a = 0; a < nn; a++) off[a + 1] += off[a];
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int a = 0; a < nn; a++)
            for (int e = m.adj_off[a]; e < m.adj_off[a + 1]; e++) {
                const int q = fill[m.adj[e]]++;
                edge[q] = e;
                row[q] = a;
            }
        GA(g->csc_off, (size_t)nn + 1);
        GA(g->csc_edge, (size_t)c.nblk);
        GA(g->csc_row, (size_t)c.nblk);
        PDX_HIP(hipMemcpy(g->csc_off, off.data(), off.size() * 4, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(g->csc_edge, edge.data(), edge.size() * 4, hipMemcpyHostToDevice));
        PDX_HIP(hipMemcpy(g->csc_row, row.data(), row.size() * 4, hipMemcpyHostToDevice));
    }
    int *ntype;
    double *S, *Bd;
    GA(ntype, (size_t)nn);
    GA(S, (size_t)nn * D);
    GA(Bd, (size_t)nn * DD);
#undef GA
    PDX_HIP(hipMemcpy(ntype, m.ntype.data(), (size_t)nn * 4, hipMemcpyHostToDevice));
    const int nb = (nn + 255) / 256;
    if (D == 3) {
        k_gnn_scale<3><<<nb, 256, 0, c.stream>>>(nn, c.adj_off, c.adj, A, c.eq, S);
        k_gnn_inputs<3><<<nb, 256, 0, c.stream>>>(nn, c.adj_off, c.adj, A, c.eq, ntype, S, g->X, g->cw, g->Aaa);
    } else {
        k_gnn_scale<2><<<nb, 256, 0, c.stream>>>(nn, c.adj_off, c.adj, A, c.eq, S);
        k_gnn_inputs<2><<<nb, 256, 0, c.stream>>>(nn, c.adj_off, c.adj, A, c.eq, ntype, S, g->X, g->cw, g->Aaa);
    }
    if ((rc = block_jacobi_inverse(c, A, Bd, err))) { delete g; return rc; }
    k_d2f<<<2048, 256, 0, c.stream>>>((size_t)nn * DD, Bd, g->B);
    PDX_HIP(hipGetLastError());
    PDX_HIP(hipStreamSynchronize(c.stream));
    g->gen = c.K_gen;
    g->op = A;
    c.gnn_work = g;
    out = g;
    return PDX_OK;
}

// the same for the context's network operator
static int gnn_work(Ctx &c, GnnWork *&out, std::string &err) {
    const double *A = nullptr;
    if (int rc = gnn_operator(c, A, err)) return rc;
    return gnn_work_for(c, A, out, err);
}

// Gaussian probes on the free DOFs (seeded, reproducible) -> f, y (FP32)
static int gnn_probes(Ctx &c, GnnWork &g, int P, unsigned long long seed, int agg, std::string &err) {
    const size_t n = (size_t)g.nn * g.D;
    if (P != g.P) {
        for (void *p : g.probe_owned) hipFree(p);
        g.probe_owned.clear();
        int rc;
        if ((rc = walloc(g.probe_owned, g.f, P * n, err)) || (rc = walloc(g.probe_owned, g.y, P * n, err)) ||
            (rc = walloc(g.probe_owned, g.R, P * n, err)))
            return rc;
        g.P = P;
    }
    const Model &m = *c.host;
    std::vector<double> z(P * n, 0.0);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int p = 0; p < P; p++)
        for (int a = 0; a < g.nn; a++)
            for (int i = 0; i < g.D; i++) {
                const double v = nd(rng);
                if (m.eq[(size_t)a * 3 + i] >= 0) z[p * n + (size_t)a * g.D + i] = v;
            }
    double *zd, *fd, *yd;
    PDX_HIP(hipMalloc((void **)&zd, P * n * 8));
    PDX_HIP(hipMalloc((void **)&fd, P * n * 8));
    PDX_HIP(hipMalloc((void **)&yd, P * n * 8));
    PDX_HIP(hipMemcpy(zd, z.data(), P * n * 8, hipMemcpyHostToDevice));
    int rc = two_level_probes(c, g.op, agg, zd, fd, yd, P, err);
    if (!rc) {
        k_dot<double><<<RED_PARTS, 256, 0, c.stream>>>(fd, fd, P * n, g.lpart);
        g.fnorm2 = host_sum(c, g.lpart, RED_PARTS, err);
        k_d2f<<<2048, 256, 0, c.stream>>>(P * n, fd, g.f);
        k_d2f<<<2048, 256, 0, c.stream>>>(P * n, yd, g.y);
        PDX_HIP(hipStreamSynchronize(c.stream));
    }
    hipFree(zd);
    hipFree(fd);
    hipFree(yd);
    g.agg = agg;
    return rc;
}

// ------------------------------------------------------------------------
// BF16 MFMA inference
// ------------------------------------------------------------------------
// Layers on v_mfma_f32_16x16x32_bf16 (A 16x32, B 32x16, D 16x16 FP32; lane l holds
// A[l%16][(l/16)*8 + v], B[(l/16)*8 + v][l%16], v < 8, and D[(l/16)*4 + v][l%16], v < 4, the
// layout checked on gfx950).  One wavefront takes 16 nodes: A = the nodes' [h ; sum_b c_ab h_b]
// (two 32-wide K chunks), B = W^T (two 16-wide output tiles): 4 MFMAs per 16 nodes.  The
// activations are stored in BF16 (the gathers of the aggregation read half the bytes); sums,
// bias, ReLU and the residual are FP32.
typedef __bf16 gnn_bf16x8 __attribute__((ext_vector_type(8)));
typedef float gnn_f4 __attribute__((ext_vector_type(4)));

// h0 = relu(W_in x + b_in) in BF16, one thread per (node, feature)
__global__ void k_dense_in_bf16(int nn, const float *X, const float *W, const float *b, __bf16 *H) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * GF) return;
    const int a = t / GF, f = t % GF;
    float s = b[f];
#pragma unroll
    for (int k = 0; k < GIN; k++) s += W[f * GIN + k] * X[(size_t)a * GIN + k];
    H[t] = (__bf16)(s > 0.f ? s : 0.f);
}

__global__ void __launch_bounds__(256) k_gnn_layer_mfma(int nn, const int *__restrict__ adj_off,
                                                        const int *__restrict__ adj, const float *__restrict__ cw,
                                                        const __bf16 *__restrict__ H, const float *__restrict__ W,
                                                        const float *__restrict__ bias, __bf16 *__restrict__ Hn) {
    const int l = threadIdx.x & 63;
    const int n0 = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 16;
    if (n0 >= nn) return;  // whole wavefronts leave together
    const int q = l >> 4, node = n0 + (l & 15);
    gnn_bf16x8 bw[2][2];  // [K chunk][output tile]: B[k][j] = W[j][k]
#pragma unroll
    for (int kc = 0; kc < 2; kc++)
#pragma unroll
        for (int jt = 0; jt < 2; jt++)
#pragma unroll
            for (int v = 0; v < 8; v++) bw[kc][jt][v] = (__bf16)W[(jt * 16 + (l & 15)) * (2 * GF) + kc * GF + q * 8 + v];
    gnn_bf16x8 ah, am;
    float mv[8];
#pragma unroll
    for (int v = 0; v < 8; v++) mv[v] = 0.f;
    if (node < nn) {
#pragma unroll
        for (int v = 0; v < 8; v++) ah[v] = H[(size_t)node * GF + q * 8 + v];
        for (int b = adj_off[node]; b < adj_off[node + 1]; b++) {
            const float w = cw[b];
            const __bf16 *hb = H + (size_t)adj[b] * GF + q * 8;
#pragma unroll
            for (int v = 0; v < 8; v++) mv[v] += w * float(hb[v]);
        }
    } else {
#pragma unroll
        for (int v = 0; v < 8; v++) ah[v] = (__bf16)0.f;
    }
#pragma unroll
    for (int v = 0; v < 8; v++) am[v] = (__bf16)mv[v];
    gnn_f4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bw[0][0], d0, 0, 0, 0);
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bw[1][0], d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bw[0][1], d1, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bw[1][1], d1, 0, 0, 0);
    const int j0 = l & 15, j1 = 16 + (l & 15);
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int nd = n0 + q * 4 + v;
        if (nd < nn) {
            const float o0 = d0[v] + bias[j0], o1 = d1[v] + bias[j1];
            Hn[(size_t)nd * GF + j0] = (__bf16)((o0 > 0.f ? o0 : 0.f) + float(H[(size_t)nd * GF + j0]));
            Hn[(size_t)nd * GF + j1] = (__bf16)((o1 > 0.f ? o1 : 0.f) + float(H[(size_t)nd * GF + j1]));
        }
    }
}

// O = W_o [H ; A_aa] + b_o from BF16 H; one thread per (node, output)
template <int D>
__global__ void k_head_bf16(int nn, const __bf16 *H, const float *Aaa, const float *Wo, const float *bo, float *O) {
    constexpr int DD = D * D, NH = GF + DD;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn * DD) return;
    const int a = t / DD, k = t % DD;
    const float *w = Wo + (size_t)k * NH;
    float s = bo[k];
    for (int q = 0; q < GF; q++) s += w[q] * float(H[(size_t)a * GF + q]);
#pragma unroll
    for (int q = 0; q < DD; q++) s += w[GF + q] * Aaa[(size_t)a * DD + q];
    O[t] = s;
}

static int gnn_forward_bf16(Ctx &c, GnnWork &g, const pdx_gnn &net, std::string &err);

static void gnn_forward(Ctx &c, GnnWork &g, const pdx_gnn &net) {
    const int nn = g.nn;
    const GnnLayout &L = net.lay;
    const float *w = net.w;
    const int nt = (nn * GF + 255) / 256;
    k_dense<<<nt, 256, 0, c.stream>>>(nn, GIN, g.X, w + L.w_in, w + L.b_in, g.U0, g.H[0], nullptr);
    for (int l = 0; l < GL; l++) {
        k_aggregate<<<nt, 256, 0, c.stream>>>(nn, c.adj_off, c.adj, g.cw, g.H[l], g.HM[l]);
        k_dense<<<nt, 256, 0, c.stream>>>(nn, 2 * GF, g.HM[l], w + L.w_l[l], w + L.b_l[l], g.U[l], g.H[l + 1], g.H[l]);
    }
    const int nh = (nn * L.DD + 255) / 256;
    if (g.D == 3) k_head<3><<<nh, 256, 0, c.stream>>>(nn, g.H[GL], g.Aaa, w + L.w_o, w + L.b_o, g.O);
    else k_head<2><<<nh, 256, 0, c.stream>>>(nn, g.H[GL], g.Aaa, w + L.w_o, w + L.b_o, g.O);
}

static int gnn_forward_bf16(Ctx &c, GnnWork &g, const pdx_gnn &net, std::string &err) {
    const int nn = g.nn;
    const GnnLayout &L = net.lay;
    const float *w = net.w;
    int rc;
    for (int l = 0; l <= GL; l++)
        if (!g.Hb[l] && (rc = walloc(g.owned, g.Hb[l], (size_t)nn * GF, err))) return rc;
    k_dense_in_bf16<<<(nn * GF + 255) / 256, 256, 0, c.stream>>>(nn, g.X, w + L.w_in, w + L.b_in, g.Hb[0]);
    const int waves = (nn + 15) / 16, blocks = (waves + 3) / 4;
    for (int l = 0; l < GL; l++)
        k_gnn_layer_mfma<<<blocks, 256, 0, c.stream>>>(nn, c.adj_off, c.adj, g.cw, g.Hb[l], w + L.w_l[l], w + L.b_l[l],
                                                       g.Hb[l + 1]);
    const int nh = (nn * L.DD + 255) / 256;
    if (g.D == 3) k_head_bf16<3><<<nh, 256, 0, c.stream>>>(nn, g.Hb[GL], g.Aaa, w + L.w_o, w + L.b_o, g.O);
    else k_head_bf16<2><<<nh, 256, 0, c.stream>>>(nn, g.Hb[GL], g.Aaa, w + L.w_o, w + L.b_o, g.O);
    PDX_HIP(hipGetLastError());
    return PDX_OK;
}

// loss of the current weights and its gradient into net.g
static int gnn_loss_backward(Ctx &c, GnnWork &g, pdx_gnn &net, double &loss, std::string &err) {
    const int nn = g.nn, D = g.D, DD = D * D;
    const GnnLayout &L = net.lay;
    gnn_forward(c, g, net);
    if (D == 3) k_gnn_resid<3><<<RED_PARTS, 256, 0, c.stream>>>(nn, g.P, g.O, g.B, c.eq, g.f, g.y, g.R, g.lpart);
    else k_gnn_resid<2><<<RED_PARTS, 256, 0, c.stream>>>(nn, g.P, g.O, g.B, c.eq, g.f, g.y, g.R, g.lpart);
    loss = host_sum(c, g.lpart, RED_PARTS, err) / g.fnorm2;
    const float scale = float(2.0 / g.fnorm2);
    const int nb = (nn + 255) / 256, nt = (nn * GF + 255) / 256;
    if (D == 3) {
        k_gnn_dO<3><<<nb, 256, 0, c.stream>>>(nn, g.P, g.O, g.B, g.R, g.y, scale, g.dO);
        k_head_back<3><<<nt, 256, 0, c.stream>>>(nn, g.dO, net.w + L.w_o, g.dH);
    } else {
        k_gnn_dO<2><<<nb, 256, 0, c.stream>>>(nn, g.P, g.O, g.B, g.R, g.y, scale, g.dO);
        k_head_back<2><<<nt, 256, 0, c.stream>>>(nn, g.dO, net.w + L.w_o, g.dH);
    }
    float *gw = net.g;
    k_wgrad<<<NPARTW, 256, 0, c.stream>>>(nn, DD, g.dO, g.H[GL], GF, g.Aaa, DD, g.part);
    k_wgrad_reduce<<<(DD * (L.nhead + 1) + 255) / 256, 256, 0, c.stream>>>(NPARTW, DD, L.nhead, g.part, gw + L.w_o,
                                                                           gw + L.b_o);
    float *dHout = g.dH, *dHin = g.dH2;
    for (int l = GL - 1; l >= 0; l--) {
        k_layer_back<<<nt, 256, 0, c.stream>>>(nn, dHout, g.U[l], net.w + L.w_l[l], dHin, g.dM, g.dU);
        k_wgrad<<<NPARTW, 256, 0, c.stream>>>(nn, GF, g.dU, g.HM[l], 2 * GF, nullptr, 0, g.part);
        k_wgrad_reduce<<<(GF * (2 * GF + 1) + 255) / 256, 256, 0, c.stream>>>(NPARTW, GF, 2 * GF, g.part,
                                                                             gw + L.w_l[l], gw + L.b_l[l]);
        k_aggr_back<<<nt, 256, 0, c.stream>>>(nn, g.csc_off, g.csc_edge, g.csc_row, g.cw, g.dM, dHin);
        std::swap(dHout, dHin);
    }
    k_relu_back<<<nt, 256, 0, c.stream>>>(nn * GF, dHout, g.U0, g.dU);
    k_wgrad<<<NPARTW, 256, 0, c.stream>>>(nn, GF, g.dU, g.X, GIN, nullptr, 0, g.part);
    k_wgrad_reduce<<<(GF * (GIN + 1) + 255) / 256, 256, 0, c.stream>>>(NPARTW, GF, GIN, g.part, gw + L.w_in,
                                                                      gw + L.b_in);
    PDX_HIP(hipGetLastError());
    return std::isfinite(loss) ? PDX_OK : (err = "non-finite loss", PDX_ERR_NOCONV);
}

static int check_pair(const pdx_ctx *pc, const pdx_gnn *g, std::string &err) {
    if (!pc || !g) { err = "null argument"; return PDX_ERR_ARG; }
    if (pc->c->dim != g->dim) { err = "network and model dimensions differ"; return PDX_ERR_ARG; }
    if (pc->c->device != g->device) { err = "network and context live on different devices"; return PDX_ERR_ARG; }
    return PDX_OK;
}

// drop what was computed for operator A (about to be freed): the smoother and the network's inputs
void forget_learned_operator(Ctx &c, const double *A) {
    if (c.learned_for == A) c.learned_for = nullptr;
    if (c.gnn_work && ((GnnWork *)c.gnn_work)->op == A) free_gnn_work(c);
}

// The smoother of the attached network for operator A (a Newmark K_eff of the current K): the
// same inference as pdx_gnn_apply, on A's inputs.
int dev_learned_smoother(Ctx &c, const double *A, std::string &err) {
    if (!c.gnn_net) {
        err = "a learned Newmark solve needs a network attached to the context (pdx_gnn_attach)";
        return PDX_ERR_ARG;
    }
    const pdx_gnn &net = *(const pdx_gnn *)c.gnn_net;
    GnnWork *w;
    if (int rc = gnn_work_for(c, A, w, err)) return rc;
    gnn_forward(c, *w, net);
    if (!c.learned_Minv) PDX_HIP(hipMalloc((void **)&c.learned_Minv, (size_t)c.nn * c.dd * 8));
    const int nb = (c.nn + 255) / 256;
    if (c.dim == 3) k_smoother<3><<<nb, 256, 0, c.stream>>>(c.nn, w->O, w->B, c.eq, c.learned_Minv);
    else k_smoother<2><<<nb, 256, 0, c.stream>>>(c.nn, w->O, w->B, c.eq, c.learned_Minv);
    PDX_HIP(hipGetLastError());
    c.learned_gen = c.K_gen;
    c.learned_for = A;
    return PDX_OK;
}

}  // namespace pdx

using namespace pdx;

static int gfail(int rc, const std::string &msg) { return set_error(rc, msg); }

extern "C" {

int pdx_gnn_create(int dim, int device, unsigned long long seed, pdx_gnn **out) {
    if (!out || (dim != 2 && dim != 3)) return gfail(PDX_ERR_ARG, "pdx_gnn_create: dim must be 2 or 3");
    std::string err;
    auto run = [&]() -> int {
        PDX_HIP(hipSetDevice(device));
        pdx_gnn *g = new pdx_gnn();
        g->dim = dim;
        g->device = device;
        g->lay = GnnLayout(dim);
        const size_t n = g->lay.total;
        float **ptrs[] = {&g->w, &g->g, &g->m, &g->v};
        for (auto pp : ptrs) {
            if (hipMalloc((void **)pp, n * 4) != hipSuccess || hipMemset(*pp, 0, n * 4) != hipSuccess) {
                err = "pdx_gnn_create: device allocation failed";
                pdx_gnn_destroy(g);
                return PDX_ERR_HIP;
            }
        }
        // He-style init of the hidden layers, zero output head (M = block Jacobi)
        std::vector<float> w(n, 0.f);
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        const GnnLayout &L = g->lay;
        for (size_t i = 0; i < (size_t)GF * GIN; i++) w[L.w_in + i] = nd(rng) * std::sqrt(2.f / GIN);
        for (int l = 0; l < GL; l++)
            for (size_t i = 0; i < (size_t)GF * 2 * GF; i++) w[L.w_l[l] + i] = nd(rng) * 0.5f * std::sqrt(1.f / GF);
        PDX_HIP(hipMemcpy(g->w, w.data(), n * 4, hipMemcpyHostToDevice));
        *out = g;
        return PDX_OK;
    };
    int rc = run();
    return rc ? gfail(rc, err) : PDX_OK;
}

void pdx_gnn_destroy(pdx_gnn *g) {
    if (!g) return;
    hipSetDevice(g->device);
    float *ptrs[] = {g->w, g->g, g->m, g->v};
    for (auto p : ptrs)
        if (p) hipFree(p);
    delete g;
}

long long pdx_gnn_num_params(const pdx_gnn *g) { return g ? (long long)g->lay.total : 0; }

int pdx_gnn_get_params(const pdx_gnn *g, float *w) {
    if (!g || !w) return gfail(PDX_ERR_ARG, "null argument");
    hipSetDevice(g->device);
    if (hipMemcpy(w, g->w, g->lay.total * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return gfail(PDX_ERR_HIP, "pdx_gnn_get_params: copy failed");
    return PDX_OK;
}

int pdx_gnn_set_params(pdx_gnn *g, const float *w) {
    if (!g || !w) return gfail(PDX_ERR_ARG, "null argument");
    hipSetDevice(g->device);
    const size_t n = g->lay.total * 4;
    if (hipMemcpy(g->w, w, n, hipMemcpyHostToDevice) != hipSuccess || hipMemset(g->m, 0, n) != hipSuccess ||
        hipMemset(g->v, 0, n) != hipSuccess)
        return gfail(PDX_ERR_HIP, "pdx_gnn_set_params: copy failed");
    g->adam_t = 0;
    return PDX_OK;
}

int pdx_gnn_train(pdx_ctx *pc, pdx_gnn *g, int n_probes, unsigned long long probe_seed, int coarse_agg, int steps,
                  float lr, double *loss) {
    std::string err;
    int rc = check_pair(pc, g, err);
    if (rc) return gfail(rc, err);
    if (n_probes < 1 || steps < 0) return gfail(PDX_ERR_ARG, "pdx_gnn_train: n_probes >= 1, steps >= 0");
    Ctx &c = *pc->c;
    hipSetDevice(c.device);
    GnnWork *w;
    if ((rc = gnn_work(c, w, err)) || (rc = gnn_probes(c, *w, n_probes, probe_seed, coarse_agg, err)))
        return gfail(rc, err);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int s = 0; s <= steps; s++) {
        double l;
        if ((rc = gnn_loss_backward(c, *w, *g, l, err))) return gfail(rc, err);
        if (loss) loss[s] = l;
        if (s == steps) break;
        g->adam_t++;
        const float c1 = 1.f - std::pow(b1, (float)g->adam_t), c2 = 1.f - std::pow(b2, (float)g->adam_t);
        k_adam<<<64, 256, 0, c.stream>>>(g->lay.total, g->w, g->g, g->m, g->v, lr, b1, b2, eps, c1, c2);
    }
    if (hipStreamSynchronize(c.stream) != hipSuccess) return gfail(PDX_ERR_HIP, "pdx_gnn_train: stream error");
    return PDX_OK;
}

int pdx_gnn_loss_grad(pdx_ctx *pc, pdx_gnn *g, int n_probes, unsigned long long probe_seed, int coarse_agg,
                      double *loss, float *grad) {
    std::string err;
    int rc = check_pair(pc, g, err);
    if (rc) return gfail(rc, err);
    Ctx &c = *pc->c;
    hipSetDevice(c.device);
    GnnWork *w;
    if ((rc = gnn_work(c, w, err)) || (rc = gnn_probes(c, *w, n_probes, probe_seed, coarse_agg, err)))
        return gfail(rc, err);
    double l;
    if ((rc = gnn_loss_backward(c, *w, *g, l, err))) return gfail(rc, err);
    if (loss) *loss = l;
    if (hipStreamSynchronize(c.stream) != hipSuccess) return gfail(PDX_ERR_HIP, "stream error");
    if (grad && hipMemcpy(grad, g->g, g->lay.total * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return gfail(PDX_ERR_HIP, "copy failed");
    return PDX_OK;
}

int pdx_gnn_inputs(pdx_ctx *pc, int n_probes, unsigned long long probe_seed, int coarse_agg, float *X, float *cw,
                   float *Aaa, float *B, float *f, float *y) {
    if (!pc) return gfail(PDX_ERR_ARG, "null argument");
    Ctx &c = *pc->c;
    hipSetDevice(c.device);
    std::string err;
    GnnWork *w;
    int rc;
    if ((rc = gnn_work(c, w, err))) return gfail(rc, err);
    if (n_probes > 0 && (rc = gnn_probes(c, *w, n_probes, probe_seed, coarse_agg, err))) return gfail(rc, err);
    const size_t nn = c.nn, DD = (size_t)c.dim * c.dim, n = nn * c.dim;
    struct {
        void *dst;
        const void *src;
        size_t bytes;
    } cp[] = {{X, w->X, nn * GIN * 4}, {cw, w->cw, (size_t)c.nblk * 4}, {Aaa, w->Aaa, nn * DD * 4},
              {B, w->B, nn * DD * 4},  {f, w->f, n_probes * n * 4},     {y, w->y, n_probes * n * 4}};
    for (auto &e : cp)
        if (e.dst && e.bytes && hipMemcpy(e.dst, e.src, e.bytes, hipMemcpyDeviceToHost) != hipSuccess)
            return gfail(PDX_ERR_HIP, "copy failed");
    return PDX_OK;
}

int pdx_gnn_apply(pdx_ctx *pc, const pdx_gnn *g) {
    std::string err;
    int rc = check_pair(pc, g, err);
    if (rc) return gfail(rc, err);
    Ctx &c = *pc->c;
    hipSetDevice(c.device);
    GnnWork *w;
    if ((rc = gnn_work(c, w, err))) return gfail(rc, err);
    gnn_forward(c, *w, *g);
    if (!c.learned_Minv && hipMalloc((void **)&c.learned_Minv, (size_t)c.nn * c.dd * 8) != hipSuccess)
        return gfail(PDX_ERR_HIP, "alloc failed");
    const int nb = (c.nn + 255) / 256;
    if (c.dim == 3) k_smoother<3><<<nb, 256, 0, c.stream>>>(c.nn, w->O, w->B, c.eq, c.learned_Minv);
    else k_smoother<2><<<nb, 256, 0, c.stream>>>(c.nn, w->O, w->B, c.eq, c.learned_Minv);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c.stream) != hipSuccess)
        return gfail(PDX_ERR_HIP, "pdx_gnn_apply: kernel failed");
    c.learned_gen = c.K_gen;
    c.learned_for = w->op;
    return PDX_OK;
}

int pdx_gnn_operator(pdx_ctx *pc, int mode) {
    if (!pc || (mode != 0 && mode != 1)) return gfail(PDX_ERR_ARG, "pdx_gnn_operator: mode is 0 (K) or 1 (Newmark)");
    Ctx &c = *pc->c;
    hipSetDevice(c.device);
    c.gnn_op = mode;
    return PDX_OK;
}

int pdx_gnn_attach(pdx_ctx *pc, const pdx_gnn *g) {
    if (!pc) return gfail(PDX_ERR_ARG, "null argument");
    if (g) {
        std::string err;
        if (int rc = check_pair(pc, g, err)) return gfail(rc, err);
    }
    pc->c->gnn_net = g;
    return PDX_OK;
}

// Inference into the learned smoother of the context's current K, FP32 (bf16 == 0, the kernels
// of pdx_gnn_apply) or BF16 activations on MFMA (bf16 != 0); M [n_node*d*d] (or NULL) gets the
// smoother, ms (or NULL) the inference time (layers, head and smoother; HIP events).
int pdx_gnn_infer(pdx_ctx *pc, const pdx_gnn *g, int bf16, double *M, double *ms) {
    std::string err;
    int rc = check_pair(pc, g, err);
    if (rc) return gfail(rc, err);
    Ctx &c = *pc->c;
    hipSetDevice(c.device);
    GnnWork *w;
    if ((rc = gnn_work(c, w, err))) return gfail(rc, err);
    if (!c.learned_Minv && hipMalloc((void **)&c.learned_Minv, (size_t)c.nn * c.dd * 8) != hipSuccess)
        return gfail(PDX_ERR_HIP, "alloc failed");
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return gfail(PDX_ERR_HIP, "event");
    hipEventRecord(e0, c.stream);
    if (bf16) {
        if ((rc = gnn_forward_bf16(c, *w, *g, err))) return gfail(rc, err);
    } else {
        gnn_forward(c, *w, *g);
    }
    const int nb = (c.nn + 255) / 256;
    if (c.dim == 3) k_smoother<3><<<nb, 256, 0, c.stream>>>(c.nn, w->O, w->B, c.eq, c.learned_Minv);
    else k_smoother<2><<<nb, 256, 0, c.stream>>>(c.nn, w->O, w->B, c.eq, c.learned_Minv);
    hipEventRecord(e1, c.stream);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
        return gfail(PDX_ERR_HIP, "pdx_gnn_infer: kernel failed");
    c.learned_for = w->op;
    float t = 0.f;
    hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (ms) *ms = t;
    if (M && hipMemcpy(M, c.learned_Minv, (size_t)c.nn * c.dd * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return gfail(PDX_ERR_HIP, "copy failed");
    c.learned_gen = c.K_gen;
    return PDX_OK;
}

}  // extern "C"
