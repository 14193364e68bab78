// Deterministic reductions shared by the explicit, eigenvalue, Krylov and
// graph-network code: every reducing kernel writes NPART partial sums (one per
// block of RED_PARTS blocks, or, for the 3-DOF Krylov SpMV, per-block partials
// folded into NPART in block order); consumers re-reduce the partials in a fixed
// order, so results are bitwise reproducible.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "device.hpp"

namespace pdx {

constexpr int NPART = RED_PARTS;  // partial sums per reduction (one per block)
constexpr int SPMV_THREADS = 1024;  // block size of the persistent Krylov SpMV (16 waves)

// block reduction of one value per thread (blockDim.x a multiple of 64, <= 1024)
__device__ inline double block_sum(double v, double *sm) {
    const int t = threadIdx.x;
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    if ((t & 63) == 0) sm[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sm[w];
        sm[16] = r;
    }
    __syncthreads();
    r = sm[16];
    __syncthreads();
    return r;
}

// every block re-reduces n partial sums in the same fixed order
__device__ inline double sum_partials(const double *p, int n, double *sm) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += p[i];
    return block_sum(v, sm);
}

template <typename V>
__global__ void __launch_bounds__(256) k_dot(const V *a, const V *b, size_t n, double *part) {
    __shared__ double sm[17];
    double v = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v += double(a[i]) * double(b[i]);
    v = block_sum(v, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
    if (threadIdx.x == 0 && blockIdx.x == 0)
        for (int i = gridDim.x; i < NPART; i++) part[i] = 0.0;
}

// host-side sum of n device partials (synchronises the stream)
inline double host_sum(Ctx &c, const double *part, int n, std::string &err) {
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

}  // namespace pdx
