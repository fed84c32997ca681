// bemx_lu_kernels.hpp -- device kernels shared by the FP64 and the mixed LU.
#pragma once
#include "bemx_common.hpp"

namespace bemx {

struct cf {
    float x, y;
};
__host__ __device__ inline cf mkf(float x, float y) { return cf{x, y}; }
__host__ __device__ inline cf operator+(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cf operator-(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cf operator*(cf a, cf b) {
    return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline cf cinvf(cf a) {
    float d = 1.0f / (a.x * a.x + a.y * a.y);
    return cf{a.x * d, -a.y * d};
}
__device__ inline cd up(cf a) { return cd{(double)a.x, (double)a.y}; }
__device__ inline cd up(cd a) { return a; }

template <class T>
__device__ inline T zero_of();
template <>
__device__ inline cd zero_of<cd>() { return mk(0.0, 0.0); }
template <>
__device__ inline cf zero_of<cf>() { return mkf(0.f, 0.f); }

__device__ inline cd shfl_c(cd v, int src) { return mk(__shfl(v.x, src, 64), __shfl(v.y, src, 64)); }
__device__ inline cf shfl_c(cf v, int src) { return mkf(__shfl(v.x, src, 64), __shfl(v.y, src, 64)); }

constexpr int TB = 64;            // triangular block size (rows of L11 / U11)
constexpr int TRSM_COLS_PER_WG = 16;

// X = L11^-1 (P X) for the TB rows k0.. of columns [c0, c0+nc), in place.
// L11 = unit lower triangle of A[k0:k0+nb, k0:k0+nb].  One wavefront per
// column, lane r = block row r; with PERM the column is first gathered by
// the block permutation lperm (row r <- row lperm[r]).
template <class T, bool PERM>
__global__ __launch_bounds__(256) void k_trsm_lower_cols(T* __restrict__ A, size_t lda, int k0,
                                                         int nb, int c0, int nc,
                                                         const int* __restrict__ lperm) {
    __shared__ T Ls[TB * TB];
    __shared__ int lp[TB];
    for (int t = threadIdx.x; t < nb * nb; t += 256) {
        int c = t / nb, r = t - c * nb;
        Ls[c * TB + r] = A[(size_t)(k0 + c) * lda + k0 + r];
    }
    if (PERM && threadIdx.x < nb) lp[threadIdx.x] = lperm[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cbeg = c0 + blockIdx.x * TRSM_COLS_PER_WG;
    const int cend = min(c0 + nc, cbeg + TRSM_COLS_PER_WG);
    for (int c = cbeg + wave; c < cend; c += 4) {
        T* col = A + (size_t)c * lda + k0;
        T x = zero_of<T>();
        if (lane < nb) x = col[PERM ? lp[lane] : lane];
        for (int i = 0; i < nb - 1; ++i) {
            T xi = shfl_c(x, i);
            if (lane > i && lane < nb) x = x - Ls[i * TB + lane] * xi;
        }
        if (lane < nb) col[lane] = x;
    }
}

}  // namespace bemx
