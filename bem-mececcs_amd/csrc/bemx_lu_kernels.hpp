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
// a - x y in four fused multiply-adds (the operator form costs six
// instructions: a multiply, a fused multiply-add and a subtraction per part)
__device__ inline cf cfms(cf a, cf x, cf y) {
    return cf{fmaf(-x.x, y.x, fmaf(x.y, y.y, a.x)), fmaf(-x.x, y.y, fmaf(-x.y, y.x, a.y))};
}
__device__ inline cd cfms(cd a, cd x, cd y) {
    return cd{fma(-x.x, y.x, fma(x.y, y.y, a.x)), fma(-x.x, y.y, fma(-x.y, y.x, a.y))};
}

// 1 / a for the FP32 factors, which only precondition: v_rcp_f32 (1 ulp)
// of |a|^2 instead of a full-precision division.
__device__ inline cf cinvf(cf a) {
    const float d = __builtin_amdgcn_rcpf(a.x * a.x + a.y * a.y);
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

// Wave-uniform copy of lane src's complex double (v_readlane, no LDS).
__device__ inline cd readlane_c(cd v, int src) {
    long long bx = __double_as_longlong(v.x), by = __double_as_longlong(v.y);
    int xl = __builtin_amdgcn_readlane((int)bx, src), xh = __builtin_amdgcn_readlane((int)(bx >> 32), src);
    int yl = __builtin_amdgcn_readlane((int)by, src), yh = __builtin_amdgcn_readlane((int)(by >> 32), src);
    return mk(__longlong_as_double(((long long)xh << 32) | (unsigned)xl),
              __longlong_as_double(((long long)yh << 32) | (unsigned)yl));
}

__device__ inline cd shfl_c(cd v, int src) { return mk(__shfl(v.x, src, 64), __shfl(v.y, src, 64)); }
__device__ inline cf shfl_c(cf v, int src) { return mkf(__shfl(v.x, src, 64), __shfl(v.y, src, 64)); }

// Split-BF16 planes of a factor panel, the operands of the mixed LU's
// trailing update: plane 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo, with
// hi = bf16(x), lo = bf16(x - hi); element (index, k) of plane pl at
// pl*pstride + index*64 + k, index = global row (L21) or column (U12).
__device__ inline void split_bf16(float v, __bf16& hi, __bf16& lo) {
    hi = (__bf16)v;
    lo = (__bf16)(v - (float)hi);
}
__device__ inline void store_planes(__bf16* __restrict__ pl, size_t pstride, size_t idx, cf v) {
    __bf16 h, l;
    split_bf16(v.x, h, l);
    pl[idx] = h;
    pl[pstride + idx] = l;
    split_bf16(v.y, h, l);
    pl[2 * pstride + idx] = h;
    pl[3 * pstride + idx] = l;
}

constexpr int TB = 64;            // triangular block size (rows of L11 / U11)
constexpr int TRSM_COLS_PER_WG = 64;

// X = L11^-1 (P X) for the TB rows k0.. of columns [c0, c0+nc), in place
// (and, for complex float with planes != NULL, split into the bf16 planes).
// L11 = unit lower triangle of A[k0:k0+nb, k0:k0+nb].  A workgroup of 256
// threads loads L11 and its 64 columns (gathered by the block permutation
// lperm, row r <- row lperm[r], with PERM) with all 32 loads of a thread in
// flight, then solves one lane per column, by chunks of 16 rows: wave 0
// solves chunk g in registers, then each of the four waves updates a
// quarter of every later chunk with it.  L11 reads are wave-uniform LDS broadcasts, the
// column reads conflict-free.
// Ls (TB x TB) and X (TB x (TB + 1)) are LDS buffers of the calling kernel.
template <class T, bool PERM>
__device__ __forceinline__ void trsm_lower_cols_body(int bid, T* __restrict__ A, size_t lda,
                                                     int k0, int nb, int c0, int nc,
                                                     const int* __restrict__ lperm,
                                                     __bf16* __restrict__ planes,
                                                     size_t pstride, T* __restrict__ Ls,
                                                     T* __restrict__ X) {
    constexpr int XS = TB + 1;  // padded row stride: the transposing staging is not bank-bound
    // Ls[i*TB + r] = L11[r][i];  X[r*XS + lane] = row r of the lane's column
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int cb = c0 + bid * TRSM_COLS_PER_WG;
    const int ncols = min(TRSM_COLS_PER_WG, c0 + nc - cb);
    {
        T lv[16], xv[16];
        // lane = row of column j (coalesced), wave w takes columns w, w+4, ...
        const int src = (lane < nb) ? (PERM ? lperm[lane] : lane) : 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int t = tid + 256 * q, i = t >> 6, r = t & 63;
            lv[q] = (i < nb && r < nb) ? A[(size_t)(k0 + i) * lda + k0 + r] : zero_of<T>();
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = w + 4 * q;
            xv[q] = (lane < nb && j < ncols) ? A[(size_t)(cb + j) * lda + k0 + src] : zero_of<T>();
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) Ls[tid + 256 * q] = lv[q];
#pragma unroll
        for (int q = 0; q < 16; ++q) X[lane * XS + w + 4 * q] = xv[q];
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < TB; g += 16) {
        if (w == 0) {
            T a[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) a[r] = X[(g + r) * XS + lane];
#pragma unroll
            for (int i = 0; i < 15; ++i) {
#pragma unroll
                for (int r = i + 1; r < 16; ++r) a[r] = cfms(a[r], Ls[(g + i) * TB + g + r], a[i]);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) X[(g + r) * XS + lane] = a[r];
        }
        __syncthreads();
        // every wave updates rows 4w..4w+3 of each later chunk, so the next
        // chunk is complete after a quarter of one chunk's update
        if (g + 16 < TB) {
            T xg[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) xg[i] = X[(g + i) * XS + lane];
#pragma unroll
            for (int g2 = g + 16; g2 < TB; g2 += 16) {
                const int rw = g2 + 4 * w;
                T a[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = X[(rw + r) * XS + lane];
#pragma unroll
                for (int i = 0; i < 16; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r) a[r] = cfms(a[r], Ls[(g + i) * TB + rw + r], xg[i]);
#pragma unroll
                for (int r = 0; r < 4; ++r) X[(rw + r) * XS + lane] = a[r];
            }
        }
        __syncthreads();
    }
#pragma unroll 4
    for (int j = w; j < TRSM_COLS_PER_WG; j += 4) {
        if (lane < nb && j < ncols) A[(size_t)(cb + j) * lda + k0 + lane] = X[lane * XS + j];
        if constexpr (sizeof(T) == sizeof(cf)) {
            // U12 columns also go out split (planes; k = lane), when asked
            if (planes && j < ncols)
                store_planes(planes, pstride, (size_t)(cb + j) * TB + lane,
                             lane < nb ? X[lane * XS + j] : zero_of<T>());
        }
    }
}

template <class T, bool PERM>
__global__ __launch_bounds__(256) void k_trsm_lower_cols(T* __restrict__ A, size_t lda, int k0,
                                                         int nb, int c0, int nc,
                                                         const int* __restrict__ lperm,
                                                         __bf16* __restrict__ planes,
                                                         size_t pstride) {
    __shared__ T Ls[TB * TB];
    __shared__ T X[TB * (TB + 1)];
    trsm_lower_cols_body<T, PERM>(blockIdx.x, A, lda, k0, nb, c0, nc, lperm, planes, pstride, Ls,
                                  X);
}

}  // namespace bemx
