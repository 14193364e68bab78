// Row segment of the node-block matrix: G lanes walk the blocks of one node
// row, each lane loading whole d x d blocks (contiguous, wide loads) and the
// d-vector of the column node, then the G partial sums are combined with
// cross-lane shuffles.  Shared by the explicit step, the SpMV and the Krylov
// SpMV kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace pdx {

// acc = sum_b K_b * x_c  (plain form), or sum_b K_b * (x_c - x_r) (difference
// form: every full block row of K sums to zero because each family, element
// and boundary operator annihilates rigid translations, so both forms are
// equal in exact arithmetic; the difference form keeps FP32 products away
// from the cancellation of large displacement terms).
// SOA: K is in row-local value order (value e of block b of a row whose blocks are [b0, b0 + nb)
// at b0*DD + e*nb + (b - b0)): lanes on consecutive blocks then read consecutive words.
// REL: the block columns are read as r + rel[k] from the row's column pattern class (rel points at
// the class's offsets) instead of from adj.
// SYM: K holds the own blocks only; src[b] is block b's own index, or ~(own index of its
// partner) for a lower block that equals its partner's transpose.
template <int D, typename T, typename X, int G, bool DIFF, bool NT = false, bool SOA = false, bool REL = false,
          bool SYM = false>
__device__ inline void row_matvec(int r, int lane, const int *__restrict__ adj_off, const int *__restrict__ adj,
                                  const T *__restrict__ K, const X *__restrict__ x, T acc[D],
                                  const int *__restrict__ rel = nullptr, const int *__restrict__ src = nullptr) {
    constexpr int DD = D * D;
#pragma unroll
    for (int i = 0; i < D; i++) acc[i] = T(0);
    X xr[D];
    if (DIFF) {
#pragma unroll
        for (int j = 0; j < D; j++) xr[j] = x[(size_t)r * D + j];
    }
    const int b0 = adj_off[r], b1 = adj_off[r + 1], nb = b1 - b0;
    for (int b = b0 + lane; b < b1; b += G) {
        const int c = REL ? r + rel[b - b0] : adj[b];
        T kb[DD];
        if constexpr (SYM) {
            const int s = src[b];
            const bool tr = s < 0;
            const T *kp = K + (size_t)(tr ? ~s : s) * DD;
            T raw[DD];
#pragma unroll
            for (int e = 0; e < DD; e++) raw[e] = kp[e];
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) kb[i * D + j] = tr ? raw[j * D + i] : raw[i * D + j];
        } else {
#pragma unroll
            for (int e = 0; e < DD; e++) {
                const size_t k = SOA ? (size_t)b0 * DD + (size_t)e * nb + (b - b0) : (size_t)b * DD + e;
                kb[e] = NT ? __builtin_nontemporal_load(&K[k]) : K[k];
            }
        }
        T xv[D];
#pragma unroll
        for (int j = 0; j < D; j++) xv[j] = DIFF ? T(x[(size_t)c * D + j] - xr[j]) : T(x[(size_t)c * D + j]);
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) acc[i] += kb[i * D + j] * xv[j];
    }
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1)
#pragma unroll
        for (int i = 0; i < D; i++) acc[i] += __shfl_xor(acc[i], s, G);
}

// acc[lane] for lane < D (0 for the other lanes), as an OR of the selected values' bit patterns:
// a chain of `if (lane == i) v = acc[i]` is folded by the compiler into one read at a lane-dependent
// offset, which keeps acc in memory (LDS or scratch) instead of registers
template <int D, typename T>
__device__ inline T pick(const T acc[D], int lane) {
    using U = std::conditional_t<sizeof(T) == 8, unsigned long long, unsigned int>;
    static_assert(sizeof(T) == sizeof(U), "pick: 4- or 8-byte values");
    U bits = 0;
#pragma unroll
    for (int i = 0; i < D; i++) bits |= lane == i ? __builtin_bit_cast(U, acc[i]) : U(0);
    return __builtin_bit_cast(T, bits);
}

}  // namespace pdx
