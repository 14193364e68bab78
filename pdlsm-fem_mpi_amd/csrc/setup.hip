// Model set-up on the GPU: the cell-list family search and the node graph.
//
// Family search (initialPDN_Fami, pdsolve.cpp:1442-1568; the same rule serves the node-cloud
// batch): the members of a query node are the node itself, then every node of the 3 x 3 x 3
// cells around the query's cell -- x outer, y, z inner, each cell's nodes in node order -- at a
// squared distance d2 with tiny < d2 and (d2 < delta_q^2 + tiny or d2 < delta_j^2 + tiny).  One
// wavefront per query: the lanes test 64 candidates of a cell at a time and the ballot's prefix
// counts place the hits, so the members come out in the host scan's order.  A count pass sizes
// the lists, a fill pass writes them.
//
// Node graph (setCSRIndexes_gloStiffMat, pdsolve.cpp:959-1121): per node the sorted, unique
// union of its family's members and the nodes of its FE elements.  One wavefront per node
// gathers the candidates into LDS, sorts them (bitonic) and keeps the first of each run.
//
// Distances are formed without FMA contraction and the cell coordinates by IEEE division, as on
// the host, so the lists are the host's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <string>
#include <vector>

#include "device.hpp"
#include "model.hpp"
#include "setup.hpp"

namespace pdx {

namespace {

constexpr int kWave = 64;

__device__ inline int query_cell(double p, double lo, double cs, int nb, bool clamp) {
    if (!clamp) return int((p - lo) / cs);
    return min(nb - 1, max(0, int(floor((p - lo) / cs))));
}

// one wavefront per query; FILL = false counts into cnt, FILL = true writes mem[off[q] ...]
template <bool FILL>
__global__ void __launch_bounds__(256) k_cell_search(long long nq, const int *__restrict__ qnode,
                                                     const double *__restrict__ x, const double *__restrict__ delta,
                                                     const int *__restrict__ cell_off,
                                                     const int *__restrict__ cell_node, CellGrid g, double tiny,
                                                     int *__restrict__ cnt, const long long *__restrict__ off,
                                                     int *__restrict__ mem) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (q >= nq) return;  // the whole wavefront leaves together
    const int lane = threadIdx.x % kWave;
    const int k = qnode ? qnode[q] : int(q);
    const double px = x[(size_t)k * 3], py = x[(size_t)k * 3 + 1], pz = x[(size_t)k * 3 + 2];
    const double dk = delta[k];
    const double rk = dk * dk + tiny;
    int ci[3];
    ci[0] = query_cell(px, g.lo[0], g.cs, g.nb[0], g.clamp && 0 < g.dim);
    ci[1] = query_cell(py, g.lo[1], g.cs, g.nb[1], g.clamp && 1 < g.dim);
    ci[2] = g.clamp && g.dim < 3 ? 0 : query_cell(pz, g.lo[2], g.cs, g.nb[2], g.clamp);
    long long base = 0;
    if (FILL) {
        base = off[q];
        if (lane == 0) mem[base] = k;
    }
    int c = 1;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int xi = ci[0] - 1; xi <= ci[0] + 1; xi++)
        for (int yi = ci[1] - 1; yi <= ci[1] + 1; yi++)
            for (int zi = ci[2] - 1; zi <= ci[2] + 1; zi++) {
                if (xi < 0 || yi < 0 || zi < 0 || xi >= g.nb[0] || yi >= g.nb[1] || zi >= g.nb[2]) continue;
                const long long cc = xi + (long long)g.nb[0] * (yi + (long long)g.nb[1] * zi);
                const int t1 = cell_off[cc + 1];
                for (int t0 = cell_off[cc]; t0 < t1; t0 += kWave) {
                    const int t = t0 + lane;
                    bool hit = false;
                    int j = 0;
                    if (t < t1) {
                        j = cell_node[t];
                        const double ax = x[(size_t)j * 3] - px, ay = x[(size_t)j * 3 + 1] - py,
                                     az = x[(size_t)j * 3 + 2] - pz;
                        const double d2 = ax * ax + ay * ay + az * az;
                        const double dj = delta[j];
                        hit = d2 > tiny && (d2 < rk || d2 < dj * dj + tiny);
                    }
                    const unsigned long long m = __ballot(hit);
                    if (FILL && hit) mem[base + c + __popcll(m & below)] = j;
                    c += __popcll(m);
                }
            }
    if (!FILL && lane == 0) cnt[q] = c;
}

// candidates of node k's row: its family's members, then the nodes of its FE elements
__device__ inline int graph_candidates(int k, const int *famid, const int *fam_off, const int *fam_mem,
                                       const int *ne_off, const int *ne_item, const int *econ, const int *nen,
                                       int lane, int *buf) {
    int n = 0;
    const int f = famid[k];
    if (f >= 0) {
        const int a = fam_off[f], b = fam_off[f + 1];
        for (int t = a + lane; t < b; t += kWave) buf[t - a] = fam_mem[t];
        n = b - a;
    }
    for (int t = ne_off[k]; t < ne_off[k + 1]; t++) {
        const int e = ne_item[t], ne = nen[e];
        if (lane < ne) buf[n + lane] = econ[(size_t)e * 8 + lane];
        n += ne;
    }
    return n;
}

// one wavefront (one workgroup) per node; cap = the power of two the LDS buffer holds
template <bool FILL>
__global__ void __launch_bounds__(64) k_node_graph(int nn, int cap, const int *__restrict__ famid,
                                                   const int *__restrict__ fam_off, const int *__restrict__ fam_mem,
                                                   const int *__restrict__ ne_off, const int *__restrict__ ne_item,
                                                   const int *__restrict__ econ, const int *__restrict__ nen,
                                                   int *__restrict__ cnt, const long long *__restrict__ off,
                                                   int *__restrict__ adj) {
    extern __shared__ int buf[];
    const int k = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = graph_candidates(k, famid, fam_off, fam_mem, ne_off, ne_item, econ, nen, lane, buf);
    int p2 = 1;
    while (p2 < n) p2 <<= 1;
    for (int t = n + lane; t < p2; t += kWave) buf[t] = INT_MAX;
    __syncthreads();
    // bitonic sort of buf[0, p2)
    for (int size = 2; size <= p2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < p2 / 2; t += kWave) {
                const int lo = 2 * t - (t & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const int a = buf[lo], b = buf[hi];
                if ((a > b) == up) {
                    buf[lo] = b;
                    buf[hi] = a;
                }
            }
            __syncthreads();
        }
    // the first of each run of equal values, in order
    const unsigned long long below = (1ull << lane) - 1ull;
    int c = 0;
    const long long base = FILL ? off[k] : 0;
    for (int t0 = 0; t0 < n; t0 += kWave) {
        const int t = t0 + lane;
        const bool keep = t < n && (t == 0 || buf[t] != buf[t - 1]);
        const unsigned long long m = __ballot(keep);
        if (FILL && keep) adj[base + c + __popcll(m & below)] = buf[t];
        c += __popcll(m);
    }
    if (!FILL && lane == 0) cnt[k] = c;
    (void)cap;
}

// device buffers of one set-up call, freed on every path
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs() {
        for (void *q : p) hipFree(q);
    }
    template <class T>
    hipError_t alloc(T *&dst, size_t n) {
        dst = nullptr;
        const hipError_t e = hipMalloc((void **)&dst, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(dst);
        return e;
    }
    template <class T>
    hipError_t upload(T *&dst, const T *src, size_t n, hipStream_t s) {
        hipError_t e = alloc(dst, n);
        if (e == hipSuccess && n) e = hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s);
        return e;
    }
};

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() {
        if (s) {
            hipStreamSynchronize(s);
            hipStreamDestroy(s);
        }
    }
};

// exclusive scan of counts into offsets (host; the offsets are also the caller's result)
static long long scan_counts(const std::vector<int> &cnt, std::vector<long long> &off, int &mx) {
    off.assign(cnt.size() + 1, 0);
    mx = 0;
    for (size_t i = 0; i < cnt.size(); i++) {
        off[i + 1] = off[i] + cnt[i];
        mx = std::max(mx, cnt[i]);
    }
    return off.back();
}

std::atomic<int> g_setup_device{-1};

}  // namespace

int setup_device() { return g_setup_device.load(); }

void set_setup_device(int device) { g_setup_device.store(device); }

int dev_cell_search(int device, int nn, const double *x, const double *delta, const std::vector<int> &cell_off,
                    const std::vector<int> &cell_node, const CellGrid &g, long long nq, const int *qnode,
                    double tiny, std::vector<long long> &off, std::vector<int> *mem, int &max_count,
                    int **dmem_out, std::string &err) {
    if (dmem_out) *dmem_out = nullptr;
    const long long ncell = (long long)g.nb[0] * g.nb[1] * g.nb[2];
    if ((long long)cell_off.size() != ncell + 1 || (long long)cell_node.size() != (long long)nn) {
        err = "cell search: cell lists do not match the grid";
        return PDX_ERR_ARG;
    }
    PDX_HIP(hipSetDevice(device));
    Stream st;
    PDX_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    DevBufs B;
    double *dx, *dd;
    int *dco, *dcn, *dq = nullptr, *dcnt;
    PDX_HIP(B.upload(dx, x, (size_t)nn * 3, st.s));
    PDX_HIP(B.upload(dd, delta, (size_t)nn, st.s));
    PDX_HIP(B.upload(dco, cell_off.data(), cell_off.size(), st.s));
    PDX_HIP(B.upload(dcn, cell_node.data(), cell_node.size(), st.s));
    if (qnode) PDX_HIP(B.upload(dq, qnode, (size_t)nq, st.s));
    PDX_HIP(B.alloc(dcnt, (size_t)nq));
    const int per = 256 / kWave;
    const long long blocks = (nq + per - 1) / per;
    if (blocks > INT_MAX) {
        err = "cell search: too many queries";
        return PDX_ERR_LIMIT;
    }
    if (nq > 0) {
        k_cell_search<false><<<int(blocks), 256, 0, st.s>>>(nq, dq, dx, dd, dco, dcn, g, tiny, dcnt, nullptr, nullptr);
        PDX_HIP(hipGetLastError());
    }
    std::vector<int> cnt((size_t)nq);
    if (nq) PDX_HIP(hipMemcpyAsync(cnt.data(), dcnt, (size_t)nq * 4, hipMemcpyDeviceToHost, st.s));
    PDX_HIP(hipStreamSynchronize(st.s));
    const long long total = scan_counts(cnt, off, max_count);
    long long *doff;
    PDX_HIP(B.upload(doff, off.data(), off.size(), st.s));
    int *dmem = nullptr;
    PDX_HIP(hipMalloc((void **)&dmem, std::max<long long>(total, 1) * 4));
    if (nq > 0) {
        k_cell_search<true><<<int(blocks), 256, 0, st.s>>>(nq, dq, dx, dd, dco, dcn, g, tiny, nullptr, doff, dmem);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            hipFree(dmem);
            err = std::string("k_cell_search: ") + hipGetErrorString(e);
            return PDX_ERR_HIP;
        }
    }
    if (mem) {
        mem->resize((size_t)total);
        const hipError_t e = total ? hipMemcpyAsync(mem->data(), dmem, (size_t)total * 4, hipMemcpyDeviceToHost, st.s)
                                   : hipSuccess;
        if (e != hipSuccess) {
            hipFree(dmem);
            err = std::string("cell search copy: ") + hipGetErrorString(e);
            return PDX_ERR_HIP;
        }
    }
    const hipError_t e = hipStreamSynchronize(st.s);
    if (e != hipSuccess || !dmem_out) hipFree(dmem);
    if (e != hipSuccess) {
        err = std::string("cell search: ") + hipGetErrorString(e);
        return PDX_ERR_HIP;
    }
    if (dmem_out) *dmem_out = dmem;
    return PDX_OK;
}

int dev_node_graph(int device, const Model &m, const std::vector<int> &ne_off, const std::vector<int> &ne_item,
                   std::vector<int> &adj_off, std::vector<int> &adj, std::string &err) {
    const int nn = m.nn;
    // the largest candidate list sets the LDS buffer
    long long most = 1;
    for (int k = 0; k < nn; k++) {
        long long n = m.famid[k] >= 0 ? m.fam_off[m.famid[k] + 1] - m.fam_off[m.famid[k]] : 0;
        for (int t = ne_off[k]; t < ne_off[k + 1]; t++) n += m.nen[ne_item[t]];
        most = std::max(most, n);
    }
    int cap = 1;
    while (cap < most) cap <<= 1;
    if (cap > 16384) {
        err = "node graph on the GPU: a node has " + std::to_string(most) +
              " candidate neighbours, more than the 16384 one workgroup sorts in LDS";
        return PDX_ERR_LIMIT;
    }
    PDX_HIP(hipSetDevice(device));
    Stream st;
    PDX_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    DevBufs B;
    int *dfamid, *dfoff, *dfmem, *dneo, *dnei, *decon, *dnen, *dcnt;
    PDX_HIP(B.upload(dfamid, m.famid.data(), m.famid.size(), st.s));
    PDX_HIP(B.upload(dfoff, m.fam_off.data(), m.fam_off.size(), st.s));
    PDX_HIP(B.upload(dfmem, m.fam_mem.data(), m.fam_mem.size(), st.s));
    PDX_HIP(B.upload(dneo, ne_off.data(), ne_off.size(), st.s));
    PDX_HIP(B.upload(dnei, ne_item.data(), ne_item.size(), st.s));
    PDX_HIP(B.upload(decon, m.econ.data(), m.econ.size(), st.s));
    PDX_HIP(B.upload(dnen, m.nen.data(), m.nen.size(), st.s));
    PDX_HIP(B.alloc(dcnt, (size_t)nn));
    const size_t lds = (size_t)cap * 4;
    if (nn > 0) {
        k_node_graph<false><<<nn, kWave, lds, st.s>>>(nn, cap, dfamid, dfoff, dfmem, dneo, dnei, decon, dnen, dcnt,
                                                       nullptr, nullptr);
        PDX_HIP(hipGetLastError());
    }
    std::vector<int> cnt(nn);
    if (nn) PDX_HIP(hipMemcpyAsync(cnt.data(), dcnt, (size_t)nn * 4, hipMemcpyDeviceToHost, st.s));
    PDX_HIP(hipStreamSynchronize(st.s));
    std::vector<long long> off;
    int mx = 0;
    const long long total = scan_counts(cnt, off, mx);
    if (total > INT_MAX) {
        err = "node graph: more than 2^31 blocks";
        return PDX_ERR_LIMIT;
    }
    long long *doff;
    int *dadj;
    PDX_HIP(B.upload(doff, off.data(), off.size(), st.s));
    PDX_HIP(B.alloc(dadj, (size_t)total));
    if (nn > 0) {
        k_node_graph<true><<<nn, kWave, lds, st.s>>>(nn, cap, dfamid, dfoff, dfmem, dneo, dnei, decon, dnen, nullptr,
                                                      doff, dadj);
        PDX_HIP(hipGetLastError());
    }
    adj.resize((size_t)total);
    if (total) PDX_HIP(hipMemcpyAsync(adj.data(), dadj, (size_t)total * 4, hipMemcpyDeviceToHost, st.s));
    PDX_HIP(hipStreamSynchronize(st.s));
    adj_off.resize((size_t)nn + 1);
    for (int k = 0; k <= nn; k++) adj_off[k] = int(off[k]);
    return PDX_OK;
}

}  // namespace pdx
