// lu_mixed.hip -- mixed-precision solve: FP32 complex LU whose trailing update
// runs as split-BF16x3 products on MFMA, used as the preconditioner of an
// FP64 GMRES iterative refinement (GMRES-IR) against the FP64 matrix.
//
// Factorisation (blocks of NBM = 64 columns, A32 = (complex float) A64):
//   k_diag_getrf32     partial pivoting inside the 64 x 64 diagonal block (LDS)
//   k_trsm_lower_cols  U12 = L11^-1 P A12   (one wavefront per column)
//   k_trsm_upper_rows  L21 = A21 U11^-1     (one wavefront per row, tile in LDS)
//   k_cgemm_bf16x3     A22 -= L21 U12; every FP32 operand x is split into
//                      hi = bf16(x), lo = bf16(x - hi) and each real product is
//                      hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with
//                      FP32 accumulation; complex = 4 real products.
// Preconditioner M^-1 v (FP64 vectors, FP32 factors): k_trsv_lower_coop and
// k_trsv_upper_coop, one cooperative launch each; workgroups own 64-row
// blocks, accumulate the contributions of finished blocks as their flags
// come up and publish their own block, so the dependency chain costs one
// small GEMV and one in-register solve per block instead of one launch.
// Pivoting is restricted to the diagonal block: the factors only
// precondition.  The accuracy of the answer comes from the refinement, whose
// residual b - A64 x is formed in FP64 with the FP64 matrix; if it stalls the
// caller falls back to the FP64 partial-pivot LU.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "../../include/bemx.h"
#include "bemx_lu.hpp"
#include "bemx_lu_kernels.hpp"

namespace bemx {

constexpr int NBM = 64;
static_assert(NBM == TB, "the mixed LU block is the triangular-kernel block");
constexpr int FLAG_SPIN_LIMIT = 1 << 22;

__global__ void k_cvt_z2c(const cd* __restrict__ a, size_t lda, cf* __restrict__ o, size_t ldo,
                          int n) {
    size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t tot = (size_t)n * n;
    if (t >= tot) return;
    size_t c = t / n, r = t - c * n;
    cd v = a[c * lda + r];
    o[c * ldo + r] = cf{(float)v.x, (float)v.y};
}

// Diagonal block: GEPP within its nb rows (largest |re|+|im|, lowest row on
// ties).  lperm[i] = block row whose content ends at block row i.
__global__ __launch_bounds__(256) void k_diag_getrf32(cf* __restrict__ A, size_t lda, int k0,
                                                      int nb, int* __restrict__ lperm,
                                                      int* __restrict__ status) {
    __shared__ cf S[NBM * NBM];
    __shared__ int pos[NBM];
    __shared__ int s_p;
    const int tid = threadIdx.x;
    for (int t = tid; t < nb * nb; t += 256) {
        int c = t / nb, r = t - c * nb;
        S[c * NBM + r] = A[(size_t)(k0 + c) * lda + k0 + r];
    }
    if (tid < NBM) pos[tid] = tid;
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (tid < 64) {
            float bv = -1.f;
            int br = 1 << 30;
            if (tid >= j && tid < nb) {
                cf v = S[j * NBM + tid];
                bv = fabsf(v.x) + fabsf(v.y);
                br = tid;
            }
            for (int off = 32; off >= 1; off >>= 1) {
                float ov = __shfl_xor(bv, off, 64);
                int orr = __shfl_xor(br, off, 64);
                if (ov > bv || (ov == bv && orr < br)) {
                    bv = ov;
                    br = orr;
                }
            }
            if (tid == 0) {
                s_p = br;
                if (bv == 0.f) atomicOr(status, ST_PIVOT_ZERO);
            }
        }
        __syncthreads();
        const int p = s_p;
        if (p != j) {
            for (int c = tid; c < nb; c += 256) {
                cf t = S[c * NBM + j];
                S[c * NBM + j] = S[c * NBM + p];
                S[c * NBM + p] = t;
            }
            if (tid == 0) {
                int t = pos[j];
                pos[j] = pos[p];
                pos[p] = t;
            }
        }
        __syncthreads();
        cf u = S[j * NBM + j];
        cf inv = (u.x != 0.f || u.y != 0.f) ? cinvf(u) : mkf(0.f, 0.f);
        if (tid > j && tid < nb) S[j * NBM + tid] = S[j * NBM + tid] * inv;
        __syncthreads();
        const int rest = nb - j - 1;
        for (int t = tid; t < rest * rest; t += 256) {
            int c = j + 1 + t / rest, r = j + 1 + t % rest;
            S[c * NBM + r] = S[c * NBM + r] - S[j * NBM + r] * S[c * NBM + j];
        }
        __syncthreads();
    }
    for (int t = tid; t < nb * nb; t += 256) {
        int c = t / nb, r = t - c * nb;
        A[(size_t)(k0 + c) * lda + k0 + r] = S[c * NBM + r];
    }
    if (tid < nb) lperm[tid] = pos[tid];
}

// L21 = A21 U11^-1 for rows [r0, r0+nr): a workgroup stages 32 rows x nb
// columns in LDS; one wavefront per row, lane c = column c:
// x_c = (a_c - sum_{i<c} x_i U_ic) / U_cc.
constexpr int TRR_ROWS = 32;
__global__ __launch_bounds__(256) void k_trsm_upper_rows(cf* __restrict__ A, size_t lda, int k0,
                                                         int nb, int r0, int nr) {
    __shared__ cf Ut[NBM * NBM];  // Ut[i*NBM + c] = U[i][c]
    __shared__ cf T[NBM * (TRR_ROWS + 1)];
    __shared__ cf dinv[NBM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = tid; t < nb * nb; t += 256) {
        int c = t / nb, i = t - c * nb;
        Ut[i * NBM + c] = A[(size_t)(k0 + c) * lda + k0 + i];
    }
    const int rb = r0 + blockIdx.x * TRR_ROWS;
    const int rn = min(TRR_ROWS, r0 + nr - rb);
    for (int t = tid; t < nb * TRR_ROWS; t += 256) {
        int c = t / TRR_ROWS, rho = t - c * TRR_ROWS;
        T[c * (TRR_ROWS + 1) + rho] =
            rho < rn ? A[(size_t)(k0 + c) * lda + rb + rho] : mkf(0.f, 0.f);
    }
    __syncthreads();
    if (tid < nb) dinv[tid] = cinvf(Ut[tid * NBM + tid]);
    __syncthreads();
    for (int rho = wave; rho < rn; rho += 4) {
        cf a = lane < nb ? T[lane * (TRR_ROWS + 1) + rho] : mkf(0.f, 0.f);
        cf acc = mkf(0.f, 0.f), xf = mkf(0.f, 0.f);
        for (int c = 0; c < nb; ++c) {
            cf xv = (a - acc) * dinv[c];
            cf xb = shfl_c(xv, c);
            if (lane == c) xf = xb;
            if (lane > c && lane < nb) acc = acc + xb * Ut[c * NBM + lane];
        }
        if (lane < nb) T[lane * (TRR_ROWS + 1) + rho] = xf;
    }
    __syncthreads();
    for (int t = tid; t < nb * TRR_ROWS; t += 256) {
        int c = t / TRR_ROWS, rho = t - c * TRR_ROWS;
        if (rho < rn) A[(size_t)(k0 + c) * lda + rb + rho] = T[c * (TRR_ROWS + 1) + rho];
    }
}

// Bounded wait until *flag >= target (see wait_count in lu.hip).
__device__ inline void wait_ge(int* flag, int target, int* status) {
    int it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++it > FLAG_SPIN_LIMIT ||
            (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
             ST_SYNC_TIMEOUT)) {
            atomicOr(status, ST_SYNC_TIMEOUT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ inline void publish(int* flag, int value) {
    __hip_atomic_store(flag, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// y = L^-1 P b, block forward substitution.  Workgroup w owns blocks
// w, w+P, ...; for its block it accumulates b - sum_k L[blk,k] y_k over the
// finished blocks k < blk (waiting on flags[k] == epoch), applies the block
// permutation, solves with the unit lower L_bb (wave 0, lane = row) and
// publishes y_blk.  Blocks only wait on lower-numbered blocks, and every
// workgroup is resident (cooperative launch), so the chain always drains.
__global__ __launch_bounds__(256) void k_trsv_lower_coop(const cf* __restrict__ A, size_t lda,
                                                         int n, int nblk,
                                                         const int* __restrict__ lperm,
                                                         const cd* __restrict__ b,
                                                         cd* __restrict__ y, int* __restrict__ flags,
                                                         int epoch, int* __restrict__ status) {
    __shared__ cd acc[NBM];
    __shared__ cd part[4][NBM];
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int r0 = blk * NBM, nb = min(NBM, n - r0);
        if (tid < NBM) acc[tid] = tid < nb ? b[r0 + tid] : mk(0, 0);
        __syncthreads();
        for (int k = 0; k < blk; ++k) {
            if (tid == 0) wait_ge(&flags[k], epoch, status);
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            cd s = mk(0, 0);
            if (lane < nb) {
                const cf* Lc = A + (size_t)(k * NBM + 16 * q) * lda + r0 + lane;
                const cd* yk = y + k * NBM + 16 * q;
#pragma unroll 4
                for (int i = 0; i < 16; ++i) s = s + up(Lc[(size_t)i * lda]) * yk[i];
            }
            part[q][lane] = s;
            __syncthreads();
            if (tid < nb) acc[tid] = acc[tid] - ((part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]));
            __syncthreads();
        }
        if (q == 0) {
            cd v = lane < nb ? acc[lperm[(size_t)blk * NBM + lane]] : mk(0, 0);
            for (int i = 0; i < nb - 1; ++i) {
                cd vi = shfl_c(v, i);
                if (lane > i && lane < nb) v = v - up(A[(size_t)(r0 + i) * lda + r0 + lane]) * vi;
            }
            if (lane < nb) y[r0 + lane] = v;
        }
        __threadfence();
        __syncthreads();
        if (tid == 0) publish(&flags[blk], epoch);
    }
}

// x = U^-1 y, block backward substitution (blocks in reverse order).
__global__ __launch_bounds__(256) void k_trsv_upper_coop(const cf* __restrict__ A, size_t lda,
                                                         int n, int nblk,
                                                         const cd* __restrict__ yv,
                                                         cd* __restrict__ x, int* __restrict__ flags,
                                                         int epoch, int* __restrict__ status) {
    __shared__ cd acc[NBM];
    __shared__ cd part[4][NBM];
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    for (int bi = blockIdx.x; bi < nblk; bi += gridDim.x) {
        const int blk = nblk - 1 - bi;
        const int r0 = blk * NBM, nb = min(NBM, n - r0);
        if (tid < NBM) acc[tid] = tid < nb ? yv[r0 + tid] : mk(0, 0);
        __syncthreads();
        for (int k = nblk - 1; k > blk; --k) {
            if (tid == 0) wait_ge(&flags[k], epoch, status);
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const int kb = min(NBM, n - k * NBM);
            cd s = mk(0, 0);
            if (lane < nb) {
                const cf* Uc = A + (size_t)(k * NBM + 16 * q) * lda + r0 + lane;
                const cd* xk = x + k * NBM + 16 * q;
                for (int i = 0; i < 16 && 16 * q + i < kb; ++i) s = s + up(Uc[(size_t)i * lda]) * xk[i];
            }
            part[q][lane] = s;
            __syncthreads();
            if (tid < nb) acc[tid] = acc[tid] - ((part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]));
            __syncthreads();
        }
        if (q == 0) {
            cd v = lane < nb ? acc[lane] : mk(0, 0);
            for (int i = nb - 1; i >= 0; --i) {
                if (lane == i) v = v * cinv(up(A[(size_t)(r0 + i) * lda + r0 + i]));
                cd vi = shfl_c(v, i);
                if (lane < i) v = v - up(A[(size_t)(r0 + i) * lda + r0 + lane]) * vi;
            }
            if (lane < nb) x[r0 + lane] = v;
        }
        __threadfence();
        __syncthreads();
        if (tid == 0) publish(&flags[blk], epoch);
    }
}

// ---------------------------------------------------------------------------
// C[m x nc] -= A[m x K] B[K x nc], complex FP32, K <= 64, split-BF16x3 MFMA.
// Workgroup tile 64 x 64, 4 waves of 32 x 32 (2 x 2 MFMA 16x16 blocks).
// LDS holds the bf16 hi/lo planes of A ([row][k]) and B ([col][k]) so a lane
// reads its 8 consecutive k with one 16-byte load.
// v_mfma_f32_16x16x32_bf16 layout: A[i][k]: lane i + 16 (k/8), element k%8;
// B[k][j]: lane j + 16 (k/8), element k%8; D[i][j]: lane j + 16 (i/4), reg i%4.
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int CT = 64;
constexpr int CK = 64;
constexpr int KP = CK + 8;  // padded k stride (bf16 elements)

__device__ inline void split_bf16(float v, __bf16& hi, __bf16& lo) {
    hi = (__bf16)v;
    lo = (__bf16)(v - (float)hi);
}

__global__ __launch_bounds__(256) void k_cgemm_bf16x3(int m, int nc, int K,
                                                      const cf* __restrict__ A, size_t lda,
                                                      const cf* __restrict__ B, size_t ldb,
                                                      cf* __restrict__ C, size_t ldc) {
    // planes: 0 re_hi, 1 re_lo, 2 im_hi, 3 im_lo
    __shared__ __attribute__((aligned(16))) __bf16 sA[4][CT * KP];
    __shared__ __attribute__((aligned(16))) __bf16 sB[4][CT * KP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bm = blockIdx.x * CT, bn = blockIdx.y * CT;
    for (int t = tid; t < CK * CT; t += 256) {
        int k = t / CT, i = t - k * CT;
        cf v = (k < K && bm + i < m) ? A[(size_t)k * lda + bm + i] : cf{0.f, 0.f};
        __bf16 h, l;
        split_bf16(v.x, h, l);
        sA[0][i * KP + k] = h;
        sA[1][i * KP + k] = l;
        split_bf16(v.y, h, l);
        sA[2][i * KP + k] = h;
        sA[3][i * KP + k] = l;
    }
    for (int t = tid; t < CK * CT; t += 256) {
        int j = t / CK, k = t - j * CK;
        cf v = (k < K && bn + j < nc) ? B[(size_t)(bn + j) * ldb + k] : cf{0.f, 0.f};
        __bf16 h, l;
        split_bf16(v.x, h, l);
        sB[0][j * KP + k] = h;
        sB[1][j * KP + k] = l;
        split_bf16(v.y, h, l);
        sB[2][j * KP + k] = h;
        sB[3][j * KP + k] = l;
    }
    __syncthreads();
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    f4 accR[2][2], accI[2][2];
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
            int col = bn + wn + 16 * y + li;
            for (int r = 0; r < 4; ++r) {
                int row = bm + wm + 16 * x + (lane >> 4) * 4 + r;
                cf v = (row < m && col < nc) ? C[(size_t)col * ldc + row] : cf{0.f, 0.f};
                accR[x][y][r] = v.x;
                accI[x][y][r] = v.y;
            }
        }
    for (int k = 0; k < CK; k += 32) {
        bf8 arh[2], arl[2], aih[2], ail[2], narh[2], narl[2], naih[2], nail[2];
        for (int x = 0; x < 2; ++x) {
            int off = (wm + 16 * x + li) * KP + k + lk;
            arh[x] = *(const bf8*)&sA[0][off];
            arl[x] = *(const bf8*)&sA[1][off];
            aih[x] = *(const bf8*)&sA[2][off];
            ail[x] = *(const bf8*)&sA[3][off];
            narh[x] = -arh[x];
            narl[x] = -arl[x];
            naih[x] = -aih[x];
            nail[x] = -ail[x];
        }
        for (int y = 0; y < 2; ++y) {
            int off = (wn + 16 * y + li) * KP + k + lk;
            bf8 brh = *(const bf8*)&sB[0][off];
            bf8 brl = *(const bf8*)&sB[1][off];
            bf8 bih = *(const bf8*)&sB[2][off];
            bf8 bil = *(const bf8*)&sB[3][off];
            for (int x = 0; x < 2; ++x) {
                f4 r = accR[x][y], i = accI[x][y];
                // Cr -= Ar Br - Ai Bi
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narh[x], brh, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narh[x], brl, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narl[x], brh, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih[x], bih, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih[x], bil, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ail[x], bih, r, 0, 0, 0);
                // Ci -= Ar Bi + Ai Br
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narh[x], bih, i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narh[x], bil, i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narl[x], bih, i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih[x], brh, i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih[x], brl, i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nail[x], brh, i, 0, 0, 0);
                accR[x][y] = r;
                accI[x][y] = i;
            }
        }
    }
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
            int col = bn + wn + 16 * y + li;
            for (int r = 0; r < 4; ++r) {
                int row = bm + wm + 16 * x + (lane >> 4) * 4 + r;
                if (row < m && col < nc) C[(size_t)col * ldc + row] = cf{accR[x][y][r], accI[x][y][r]};
            }
        }
}

// ---------------------------------------------------------------------------
// FP64 vector kernels of the refinement.
// ---------------------------------------------------------------------------
constexpr int MV_COLS = 128;  // columns per GEMV chunk

// part[chunk][row] = sum over the chunk's columns of A[row, col] x[col]
__global__ __launch_bounds__(256) void k_zgemv_part(const cd* __restrict__ A, size_t lda, int n,
                                                    const cd* __restrict__ x,
                                                    cd* __restrict__ part) {
    const int row = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * MV_COLS;
    const int c1 = min(n, c0 + MV_COLS);
    __shared__ cd xs[MV_COLS];
    if (threadIdx.x < c1 - c0) xs[threadIdx.x] = x[c0 + threadIdx.x];
    __syncthreads();
    if (row >= n) return;
    cd s = mk(0, 0);
    const cd* a = A + (size_t)c0 * lda + row;
    for (int c = 0; c < c1 - c0; ++c) s = s + a[(size_t)c * lda] * xs[c];
    part[(size_t)blockIdx.y * n + row] = s;
}

// r = b - sum(part)   (sub = 1)   or   r = sum(part)   (sub = 0)
__global__ void k_gemv_reduce(int n, int nchunk, const cd* __restrict__ part,
                              const cd* __restrict__ b, cd* __restrict__ r, int sub) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cd s = mk(0, 0);
    for (int k = 0; k < nchunk; ++k) s = s + part[(size_t)k * n + i];
    r[i] = sub ? b[i] - s : s;
}

// partial[blk][i] = sum over the block's rows of conj(V_i) w   (i < nv), and
// partial[blk][nv] = sum |w|^2 (real part).
__global__ __launch_bounds__(256) void k_dots(int n, const cd* __restrict__ V, size_t ldv, int nv,
                                              const cd* __restrict__ w, cd* __restrict__ partial) {
    __shared__ double red[2][4];
    const int row = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    cd wv = row < n ? w[row] : mk(0, 0);
    for (int i = 0; i <= nv; ++i) {
        cd d;
        if (i < nv) {
            cd v = row < n ? V[(size_t)i * ldv + row] : mk(0, 0);
            d = mk(v.x * wv.x + v.y * wv.y, v.x * wv.y - v.y * wv.x);
        } else {
            d = mk(wv.x * wv.x + wv.y * wv.y, 0.0);
        }
        double a = wave_sum(d.x), b = wave_sum(d.y);
        if (lane == 0) {
            red[0][wave] = a;
            red[1][wave] = b;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            partial[(size_t)blockIdx.x * (nv + 1) + i] =
                mk(red[0][0] + red[0][1] + red[0][2] + red[0][3],
                   red[1][0] + red[1][1] + red[1][2] + red[1][3]);
        __syncthreads();
    }
}

// w = (w - sum_i h_i V_i) * s
__global__ void k_orth_update(int n, const cd* __restrict__ V, size_t ldv, int nv,
                              const cd* __restrict__ h, cd* __restrict__ w, cd* __restrict__ out,
                              double s) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    cd v = w[row];
    for (int i = 0; i < nv; ++i) v = v - h[i] * V[(size_t)i * ldv + row];
    out[row] = v * s;
}

// x += sum_i y_i V_i
__global__ void k_combine(int n, const cd* __restrict__ V, size_t ldv, int nv,
                          const cd* __restrict__ y, cd* __restrict__ x) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    cd v = x[row];
    for (int i = 0; i < nv; ++i) v = v + y[i] * V[(size_t)i * ldv + row];
    x[row] = v;
}

__global__ void k_copy_z(const cd* __restrict__ s, cd* __restrict__ d, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

__global__ void k_fill_i(int* p, int v, int cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
#define MX_CHK(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            err = std::string("mixed solve: ") + hipGetErrorString(e_);      \
            return e_ == hipErrorOutOfMemory ? BEMX_ERR_NOMEM : BEMX_ERR_HIP; \
        }                                                                     \
    } while (0)

// Workspace slots 9.. of LuWork (0..8 belong to the FP64 path).
enum { M_A32 = 9, M_PERM, M_FLAGS, M_STATUS, M_VEC, M_V, M_PART, M_HOST, M_PTMP };

struct MixedFactor {
    cf* A = nullptr;
    int* lperm = nullptr;
    int* flags = nullptr;
    int* status = nullptr;
    int* epoch = nullptr;   // host-side counter living in LuWork
    int n = 0, nblk = 0, coop = 1;
};

static int factor_mixed(LuWork& w, hipStream_t st, int n, const cd* A64, size_t lda,
                        MixedFactor& F, std::string& err) {
    hipError_t e;
    const int nblk = (n + NBM - 1) / NBM;
    F.n = n;
    F.nblk = nblk;
    F.A = (cf*)w.get(M_A32, sizeof(cf) * (size_t)n * n, &e);
    MX_CHK(e);
    F.lperm = (int*)w.get(M_PERM, sizeof(int) * (size_t)nblk * NBM, &e);
    MX_CHK(e);
    const size_t cap_before = w.cap[M_FLAGS];
    F.flags = (int*)w.get(M_FLAGS, sizeof(int) * (size_t)nblk, &e);
    MX_CHK(e);
    if (w.flags_init != (void*)F.flags || w.cap[M_FLAGS] != cap_before) {
        // fresh flag array: clear it once
        hipLaunchKernelGGL(k_fill_i, dim3((nblk + 255) / 256), dim3(256), 0, st, F.flags, 0, nblk);
        w.flags_init = F.flags;
    }
    F.epoch = &w.epoch;
    F.status = (int*)w.get(M_STATUS, sizeof(int), &e);
    MX_CHK(e);
    hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, st, F.status, 0, 1);
    int dev = 0, ncu = 0, per = 0;
    MX_CHK(hipGetDevice(&dev));
    MX_CHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    MX_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_trsv_lower_coop, 256, 0));
    int per2 = 0;
    MX_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, k_trsv_upper_coop, 256, 0));
    F.coop = std::max(1, std::min(nblk, ncu * std::min(per, per2)));
    size_t tot = (size_t)n * n;
    hipLaunchKernelGGL(k_cvt_z2c, dim3((tot + 255) / 256), dim3(256), 0, st, A64, lda, F.A,
                       (size_t)n, n);
    for (int k = 0; k < nblk; ++k) {
        const int k0 = k * NBM, nb = std::min(NBM, n - k0);
        int* lp = F.lperm + (size_t)k * NBM;
        hipLaunchKernelGGL(k_diag_getrf32, dim3(1), dim3(256), 0, st, F.A, (size_t)n, k0, nb, lp,
                           F.status);
        const int c0 = k0 + nb, rest = n - c0;
        if (rest <= 0) continue;
        hipLaunchKernelGGL((k_trsm_lower_cols<cf, true>),
                           dim3((rest + TRSM_COLS_PER_WG - 1) / TRSM_COLS_PER_WG), dim3(256), 0,
                           st, F.A, (size_t)n, k0, nb, c0, rest, (const int*)lp);
        hipLaunchKernelGGL(k_trsm_upper_rows, dim3((rest + TRR_ROWS - 1) / TRR_ROWS), dim3(256),
                           0, st, F.A, (size_t)n, k0, nb, c0, rest);
        hipLaunchKernelGGL(k_cgemm_bf16x3, dim3((rest + CT - 1) / CT, (rest + CT - 1) / CT),
                           dim3(256), 0, st, rest, rest, nb, F.A + (size_t)k0 * n + c0, (size_t)n,
                           F.A + (size_t)c0 * n + k0, (size_t)n, F.A + (size_t)c0 * n + c0,
                           (size_t)n);
    }
    MX_CHK(hipGetLastError());
    int hs = 0;
    MX_CHK(hipMemcpyAsync(&hs, F.status, sizeof(int), hipMemcpyDeviceToHost, st));
    MX_CHK(hipStreamSynchronize(st));
    if (hs & ST_PIVOT_ZERO) {
        err = "mixed solve: zero pivot in a diagonal block";
        return BEMX_ERR_SINGULAR;
    }
    return BEMX_OK;
}

// z = M^-1 v = U^-1 L^-1 P v   (tmp: work vector of n; z may alias v)
static hipError_t precond(hipStream_t st, const MixedFactor& F, const cd* v, cd* z, cd* tmp) {
    int n = F.n, nblk = F.nblk;
    size_t lda = (size_t)n;
    const cf* A = F.A;
    const int* lp = F.lperm;
    int* flags = F.flags;
    int* status = F.status;
    int e1 = ++*F.epoch;
    void* a1[] = {(void*)&A, &lda, &n, &nblk, (void*)&lp, (void*)&v, &tmp, &flags, &e1, &status};
    hipError_t e = hipLaunchCooperativeKernel((const void*)k_trsv_lower_coop, dim3(F.coop),
                                              dim3(256), a1, 0, st);
    if (e != hipSuccess) return e;
    int e2 = ++*F.epoch;
    const cd* tv = tmp;
    void* a2[] = {(void*)&A, &lda, &n, &nblk, (void*)&tv, &z, &flags, &e2, &status};
    return hipLaunchCooperativeKernel((const void*)k_trsv_upper_coop, dim3(F.coop), dim3(256), a2,
                                      0, st);
}

static void gemv(hipStream_t st, int n, const cd* A, size_t lda, const cd* x, const cd* b, cd* r,
                 cd* part, int sub) {
    const int nchunk = (n + MV_COLS - 1) / MV_COLS;
    hipLaunchKernelGGL(k_zgemv_part, dim3((n + 255) / 256, nchunk), dim3(256), 0, st, A, lda, n, x,
                       part);
    hipLaunchKernelGGL(k_gemv_reduce, dim3((n + 255) / 256), dim3(256), 0, st, n, nchunk, part, b, r,
                       sub);
}

typedef std::complex<double> zx;

// GMRES-IR of one right-hand side against the FP32 factors F.
static int refine_one(LuWork& w, hipStream_t st, const MixedFactor& F, int n, const cd* A64,
                      size_t lda, cd* b, int* iters, double* relres_out, std::string& err) {
    hipError_t e;
    const int MR = 12;        // GMRES restart length
    const int MAXOUT = 12;    // refinement steps
    const double TOL = 2e-15; // normwise backward error target
    const int nb256 = (n + 255) / 256;
    cd* vec = (cd*)w.get(M_VEC, sizeof(cd) * (size_t)n * 5, &e);
    MX_CHK(e);
    cd *x = vec, *r = vec + n, *wv = vec + 2 * (size_t)n, *tmp = vec + 3 * (size_t)n,
       *bsave = vec + 4 * (size_t)n;
    cd* V = (cd*)w.get(M_V, sizeof(cd) * (size_t)n * (MR + 1), &e);
    MX_CHK(e);
    const int nchunk = (n + MV_COLS - 1) / MV_COLS;
    cd* part = (cd*)w.get(M_PART, sizeof(cd) * std::max((size_t)nchunk * n,
                                                         (size_t)nb256 * (MR + 2)), &e);
    MX_CHK(e);
    cd* hdev = (cd*)w.get(M_HOST, sizeof(cd) * (MR + 2), &e);
    MX_CHK(e);
    std::vector<cd> hpart((size_t)nb256 * (MR + 2));
    auto dots = [&](const cd* Vb, int nv, const cd* wvec, std::vector<zx>& out, double& nrm2) -> int {
        hipLaunchKernelGGL(k_dots, dim3(nb256), dim3(256), 0, st, n, Vb, (size_t)n, nv, wvec, part);
        MX_CHK(hipMemcpyAsync(hpart.data(), part, sizeof(cd) * nb256 * (nv + 1),
                              hipMemcpyDeviceToHost, st));
        MX_CHK(hipStreamSynchronize(st));
        out.assign(nv, zx(0, 0));
        nrm2 = 0;
        for (int bk = 0; bk < nb256; ++bk) {
            for (int i = 0; i < nv; ++i)
                out[i] += zx(hpart[(size_t)bk * (nv + 1) + i].x, hpart[(size_t)bk * (nv + 1) + i].y);
            nrm2 += hpart[(size_t)bk * (nv + 1) + nv].x;
        }
        return 0;
    };
    // norms for the backward error: ||A||_inf estimate via |A| row sums is
    // costly; use ||b|| and the residual relative to ||A x|| instead.
    hipLaunchKernelGGL(k_copy_z, dim3(nb256), dim3(256), 0, st, b, bsave, n);
    MX_CHK(precond(st, F, bsave, x, tmp));  // x0 = M^-1 b
    std::vector<zx> h;
    double nrm2 = 0, bnorm = 0, relres = 1.0;
    int it_total = 0;
    dots(bsave, 0, bsave, h, nrm2);
    bnorm = std::sqrt(nrm2);
    bool converged = false;
    for (int outer = 0; outer < MAXOUT; ++outer) {
        gemv(st, n, A64, lda, x, bsave, r, part, 1);  // r = b - A x (FP64)
        double rn2 = 0, xn2 = 0;
        dots(r, 0, r, h, rn2);
        dots(x, 0, x, h, xn2);
        (void)xn2;
        double prev = relres;
        relres = std::sqrt(rn2) / std::max(bnorm, 1e-300);
        // ||A x|| ~ ||b||, so ||r|| <= TOL ||b|| bounds the normwise backward
        // error ||r|| / (||A|| ||x|| + ||b||) by TOL; a residual that stops
        // shrinking at a level near the FP64 rounding of A x is accepted too.
        if (relres <= TOL || (outer >= 2 && relres > 0.5 * prev && relres <= 1e-13)) {
            converged = true;
            break;
        }
        // GMRES on M^-1 A d = M^-1 r
        MX_CHK(precond(st, F, r, V, tmp));
        double beta2 = 0;
        dots(V, 0, V, h, beta2);
        double beta = std::sqrt(beta2);
        if (beta == 0.0) {
            converged = true;
            break;
        }
        hipLaunchKernelGGL(k_orth_update, dim3(nb256), dim3(256), 0, st, n, V, (size_t)n, 0, hdev,
                           V, V, 1.0 / beta);
        std::vector<std::vector<zx>> H(MR + 1, std::vector<zx>(MR, zx(0, 0)));
        std::vector<zx> cs(MR), sn(MR), g(MR + 1, zx(0, 0));
        g[0] = beta;
        int k = 0;
        for (; k < MR; ++k) {
            ++it_total;
            gemv(st, n, A64, lda, V + (size_t)k * n, nullptr, wv, part, 0);
            MX_CHK(precond(st, F, wv, wv, tmp));
            // classical Gram-Schmidt, two passes
            std::vector<zx> hcol(k + 1, zx(0, 0));
            double wn2 = 0;
            for (int pass = 0; pass < 2; ++pass) {
                std::vector<zx> hh;
                dots(V, k + 1, wv, hh, wn2);
                std::vector<cd> hc(k + 1);
                for (int i = 0; i <= k; ++i) {
                    hcol[i] += hh[i];
                    hc[i] = mk(hh[i].real(), hh[i].imag());
                }
                MX_CHK(hipMemcpyAsync(hdev, hc.data(), sizeof(cd) * (k + 1), hipMemcpyHostToDevice,
                                      st));
                hipLaunchKernelGGL(k_orth_update, dim3(nb256), dim3(256), 0, st, n, V, (size_t)n,
                                   k + 1, hdev, wv, wv, 1.0);
            }
            dots(V, 0, wv, h, wn2);
            double hn = std::sqrt(wn2);
            for (int i = 0; i <= k; ++i) H[i][k] = hcol[i];
            H[k + 1][k] = hn;
            if (hn > 0)
                hipLaunchKernelGGL(k_orth_update, dim3(nb256), dim3(256), 0, st, n, V, (size_t)n, 0,
                                   hdev, wv, V + (size_t)(k + 1) * n, 1.0 / hn);
            // Givens
            for (int i = 0; i < k; ++i) {
                zx t = std::conj(cs[i]) * H[i][k] + std::conj(sn[i]) * H[i + 1][k];
                H[i + 1][k] = -sn[i] * H[i][k] + cs[i] * H[i + 1][k];
                H[i][k] = t;
            }
            double a = std::abs(H[k][k]), bb = std::abs(H[k + 1][k]);
            double den = std::sqrt(a * a + bb * bb);
            if (den == 0) {
                cs[k] = 1;
                sn[k] = 0;
            } else {
                cs[k] = H[k][k] / den;
                sn[k] = H[k + 1][k] / den;
            }
            H[k][k] = std::conj(cs[k]) * H[k][k] + std::conj(sn[k]) * H[k + 1][k];
            H[k + 1][k] = 0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = std::conj(cs[k]) * g[k];
            if (std::abs(g[k + 1]) <= 1e-8 * beta || hn == 0.0) {
                ++k;
                break;
            }
        }
        // back substitution H y = g, then x += V y
        std::vector<zx> y(k);
        for (int i = k - 1; i >= 0; --i) {
            zx t = g[i];
            for (int j = i + 1; j < k; ++j) t -= H[i][j] * y[j];
            y[i] = t / H[i][i];
        }
        std::vector<cd> yc(k);
        for (int i = 0; i < k; ++i) yc[i] = mk(y[i].real(), y[i].imag());
        MX_CHK(hipMemcpyAsync(hdev, yc.data(), sizeof(cd) * k, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_combine, dim3(nb256), dim3(256), 0, st, n, V, (size_t)n, k, hdev, x);
    }
    MX_CHK(hipGetLastError());
    int hst = 0;
    MX_CHK(hipMemcpyAsync(&hst, F.status, sizeof(int), hipMemcpyDeviceToHost, st));
    MX_CHK(hipStreamSynchronize(st));
    if (hst & ST_SYNC_TIMEOUT) {
        err = "mixed solve: a device-side wait of the triangular solves timed out";
        return BEMX_ERR_TIMEOUT;
    }
    *iters = it_total;
    *relres_out = relres;
    if (!converged) {
        MX_CHK(hipStreamSynchronize(st));
        err = "mixed solve: GMRES-IR did not reach the FP64 residual level";
        return BEMX_ERR_NOCONV;
    }
    hipLaunchKernelGGL(k_copy_z, dim3(nb256), dim3(256), 0, st, x, b, n);
    MX_CHK(hipStreamSynchronize(st));
    return BEMX_OK;
}

int gmres_ir(LuWork& w, hipStream_t st, int n, const cd* A64, size_t lda, cd* b, int nrhs,
             double stats[4], std::string& err, int* nsolved) {
    *nsolved = 0;
    MixedFactor F;
    int s = factor_mixed(w, st, n, A64, lda, F, err);
    if (s) return s;
    int iters = 0;
    double worst = 0.0;
    for (int j = 0; j < nrhs; ++j) {
        int it = 0;
        double rr = 0.0;
        s = refine_one(w, st, F, n, A64, lda, b + (size_t)j * n, &it, &rr, err);
        iters += it;
        worst = std::max(worst, rr);
        if (stats) {
            stats[0] = iters;
            stats[1] = worst;
        }
        if (s) return s;
        *nsolved = j + 1;
    }
    return BEMX_OK;
}

}  // namespace bemx
