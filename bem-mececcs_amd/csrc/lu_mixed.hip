// lu_mixed.hip -- mixed-precision solve: FP32 complex LU whose trailing update
// runs as split-BF16x3 products on MFMA, used as the preconditioner of an
// FP64 GMRES iterative refinement (GMRES-IR) against the FP64 matrix.
//
// Factorisation (blocks of NBM = 64 columns, A32 = (complex float) A64):
//   k_diag_getrf32     partial pivoting inside the 64 x 64 diagonal block
//                      (registers: lane = row, wave = 16 columns)
//   k_trsm_lower_cols  U12 = L11^-1 P A12   (one lane per column)
//   k_trsm_upper_rows  L21 = A21 U11^-1     (one lane per row)
//   k_cgemm_planes     A22 -= L21 U12; every FP32 operand x is split into
//                      hi = bf16(x), lo = bf16(x - hi) (by the two triangular
//                      kernels, into bf16 planes) and each real product is
//                      hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16 with
//                      FP32 accumulation; complex = 4 real products.
// The factorisation looks one block ahead over two streams: the next
// diagonal block is factored while the bulk update of the current step runs.
// Preconditioner M^-1 v (FP64 vectors, FP32 factors): k_trsv_lower_coop and
// k_trsv_upper_coop, one cooperative launch each; workgroups own 64-row
// blocks, accumulate the contributions of finished blocks as their flags
// come up and publish their own block, so the dependency chain costs one
// small GEMV and one product with the block's explicit inverse
// (k_diag_inverses, computed on the side stream during the factorisation)
// per block instead of one launch.
// Pivoting is restricted to the diagonal block: the factors only
// precondition.  The accuracy of the answer comes from the refinement, whose
// residual b - A64 x is formed in FP64 with the FP64 matrix; if it stalls the
// caller falls back to the FP64 partial-pivot LU.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
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

// Maximum over the wavefront with DPP (no LDS): prefix maxima inside each
// row of 16 lanes (row_shr 1, 2, 4, 8), then across rows (row_bcast 15 into
// rows 1 and 3, row_bcast 31 into rows 2 and 3); lane 63 ends with the
// maximum.  Lanes whose DPP source lies outside the row take 0, the identity.
__device__ inline unsigned wave_max_u32(unsigned v) {
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Wave-uniform copy of lane src's complex float.
__device__ inline cf readlane_f(cf v, int src) {
    return mkf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), src)),
               __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), src)));
}

// Diagonal block: GEPP within its nb rows (largest |re|+|im| to 17 mantissa
// bits, lowest block row on ties), held in registers: wave q owns columns
// [16q, 16q+16), lane r
// owns block row r.  Rows are not moved while factoring: a row is marked
// used when it becomes the pivot of step j (its rank), and the factors are
// written back in rank order.  lperm[i] = block row whose content ends at
// block row i.  One barrier per step: the wave owning column j+1 applies
// step j to that column first and publishes the pivot and multipliers of
// step j+1 (double-buffered in LDS) before finishing its step-j update.
__global__ __launch_bounds__(256) void k_diag_getrf32(cf* __restrict__ A, size_t lda, int k0,
                                                      int nb, int* __restrict__ lperm,
                                                      int* __restrict__ status) {
    __shared__ cf lcol[2][NBM];
    __shared__ int s_piv[2];
    const int tid = threadIdx.x, lane = tid & 63;
    const int q = __builtin_amdgcn_readfirstlane(tid >> 6);
    cf v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int c = 16 * q + i;
        v[i] = (lane < nb && c < nb) ? A[(size_t)(k0 + c) * lda + k0 + lane] : mkf(0.f, 0.f);
    }
    bool used = lane >= nb;
    int rank = lane;  // rows >= nb keep their place
    // pivot of column (wave-local index) ii; multipliers stored in v[ii]
    auto search = [&](cf& col, int buf) {
        // key = |re|+|im| with its 6 lowest mantissa bits replaced by 63 - lane:
        // the wave maximum picks the largest candidate (to 17 mantissa bits,
        // ample for pivoting a preconditioner), lowest row on ties
        const unsigned key =
            used ? 0u
                 : ((__float_as_uint(fabsf(col.x) + fabsf(col.y)) & ~63u) | (unsigned)(63 - lane));
        const int p = 63 - (int)(wave_max_u32(key) & 63u);
        const cf pv = readlane_f(col, p);
        const bool zero = pv.x == 0.f && pv.y == 0.f;
        const cf inv = zero ? mkf(0.f, 0.f) : cinvf(pv);
        cf l = mkf(0.f, 0.f);
        if (!used && lane != p) {
            l = col * inv;
            col = l;
        }
        lcol[buf][lane] = l;
        if (lane == 0) {
            s_piv[buf] = p;
            if (zero) atomicOr(status, ST_PIVOT_ZERO);
        }
    };
    if (q == 0 && nb > 0) search(v[0], 0);
    __syncthreads();
    for (int jc = 0; jc < 4; ++jc) {
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const int j = 16 * jc + jj;
            if (j < nb) {
                const int p = __builtin_amdgcn_readfirstlane(s_piv[j & 1]);
                const cf l = lcol[j & 1][lane];
                if (lane == p) {
                    used = true;
                    rank = j;
                }
                const int qn = jj == 15 ? jc + 1 : jc;  // owner of column j+1
                const int in = (jj + 1) & 15;
                const bool owner = q == qn && j + 1 < nb;
                if (owner) {
                    v[in] = v[in] - l * readlane_f(v[in], p);
                    search(v[in], (j + 1) & 1);
                }
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    if (16 * q + i > j && !(owner && i == in)) v[i] = v[i] - l * readlane_f(v[i], p);
                }
                __syncthreads();
            }
        }
    }
    if (lane < nb) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int c = 16 * q + i;
            if (c < nb) A[(size_t)(k0 + c) * lda + k0 + rank] = v[i];
        }
        if (q == 0) lperm[rank] = lane;
    }
}

// L21 = A21 U11^-1 for rows [r0, r0+nr).  256 threads load U11 and 64
// rows with all 32 loads of a thread in flight; the rows are then solved one
// lane per row, left to right by chunks of 16 columns: wave 0 solves chunk g
// in registers, then the four waves update the later chunks with it in
// parallel.  U11 reads are wave-uniform broadcasts.  The result goes back to
// A and, split, to the planes la.
constexpr int TRR_ROWS = 64;
__device__ __forceinline__ void trsm_upper_rows_body(int bid, cf* __restrict__ A, size_t lda,
                                                     int k0, int nb, int r0, int nr,
                                                     __bf16* __restrict__ la, size_t pstride) {
    constexpr int XS = NBM + 1;
    __shared__ cf Us[NBM * NBM];  // Us[c2*NBM + c] = U11[c][c2]
    __shared__ cf X[NBM * XS];    // X[c*XS + lane] = column c of the lane's row
    __shared__ cf dinv[NBM];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int rb = r0 + bid * TRR_ROWS;
    const bool ok = rb + lane < r0 + nr;
    {
        cf uv[16], xv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int t = tid + 256 * q, c2 = t >> 6, c = t & 63;
            uv[q] = (c2 < nb && c < nb) ? A[(size_t)(k0 + c2) * lda + k0 + c] : mkf(0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = w + 4 * q;
            xv[q] = (ok && c < nb) ? A[(size_t)(k0 + c) * lda + rb + lane] : mkf(0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) Us[tid + 256 * q] = uv[q];
#pragma unroll
        for (int q = 0; q < 16; ++q) X[(w + 4 * q) * XS + lane] = xv[q];
    }
    __syncthreads();
    if (tid < NBM) {
        cf u = Us[tid * NBM + tid];
        dinv[tid] = (u.x != 0.f || u.y != 0.f) ? cinvf(u) : mkf(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < NBM; g += 16) {
        if (w == 0) {
            cf a[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) a[c] = X[(g + c) * XS + lane];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                a[c] = a[c] * dinv[g + c];
#pragma unroll
                for (int c2 = c + 1; c2 < 16; ++c2) a[c2] = a[c2] - a[c] * Us[(g + c2) * NBM + g + c];
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) X[(g + c) * XS + lane] = a[c];
        }
        __syncthreads();
        const int g2 = g + 16 * (1 + w);  // the chunk this wave updates
        if (g2 < NBM) {
            cf a[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) a[c] = X[(g2 + c) * XS + lane];
#pragma unroll 4
            for (int i = 0; i < 16; ++i) {
                const cf xi = X[(g + i) * XS + lane];
#pragma unroll
                for (int c = 0; c < 16; ++c) a[c] = a[c] - xi * Us[(g2 + c) * NBM + g + i];
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) X[(g2 + c) * XS + lane] = a[c];
        }
        __syncthreads();
    }
#pragma unroll 4
    for (int c = w; c < NBM; c += 4)
        if (ok && c < nb) A[(size_t)(k0 + c) * lda + rb + lane] = X[c * XS + lane];
    // planes: t -> (row rr, k), k fastest, so the 2-byte stores are coalesced
#pragma unroll 4
    for (int t = tid; t < TRR_ROWS * NBM; t += 256) {
        const int rr = t >> 6, k = t & 63;
        if (rb + rr < r0 + nr)
            store_planes(la, pstride, (size_t)(rb + rr) * NBM + k,
                         k < nb ? X[k * XS + rr] : mkf(0.f, 0.f));
    }
}

__global__ __launch_bounds__(256) void k_trsm_upper_rows(cf* __restrict__ A, size_t lda, int k0,
                                                         int nb, int r0, int nr,
                                                         __bf16* __restrict__ la, size_t pstride) {
    trsm_upper_rows_body(blockIdx.x, A, lda, k0, nb, r0, nr, la, pstride);
}

// Bounded wait until *flag >= target (see wait_count in lu.hip); relaxed
// polls, the caller issues one acquire fence after the wait.
__device__ inline void wait_ge(int* flag, int target, int* status) {
    int it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
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

// Explicit inverses of diagonal block k for the preconditioner's block
// solves: Li = L11^-1 P (workgroup 0, one lane per column of P) and
// Ui = U11^-1 (workgroup 1, one lane per row of I), column-major 64 x 64.
// Solved like the two triangular kernels (16-wide chunks: wave 0 solves a
// chunk in registers, the four waves update the later chunks).  A partial
// last block is padded with the identity.
__global__ __launch_bounds__(256) void k_diag_inverses(const cf* __restrict__ A, size_t lda,
                                                       int k0, int nb,
                                                       const int* __restrict__ lperm,
                                                       cf* __restrict__ Li, cf* __restrict__ Ui) {
    constexpr int XS = NBM + 1;
    __shared__ cf M[NBM * NBM];  // M[c*NBM + r] = block[r][c]
    __shared__ cf X[NBM * XS];
    __shared__ cf dinv[NBM];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    {
        cf mv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int t = tid + 256 * q, c = t >> 6, r = t & 63;
            mv[q] = (c < nb && r < nb) ? A[(size_t)(k0 + c) * lda + k0 + r]
                                       : mkf(r == c ? 1.f : 0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) M[tid + 256 * q] = mv[q];
    }
    if (blockIdx.x == 0) {
        // X[r][col] = (P I)[r][col]: row r of P I is row lperm[r] of I
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = w + 4 * q;
            const int src = r < nb ? lperm[r] : r;
            X[r * XS + lane] = mkf(src == lane ? 1.f : 0.f, 0.f);
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < NBM; g += 16) {
            if (w == 0) {
                cf a[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) a[r] = X[(g + r) * XS + lane];
#pragma unroll
                for (int i = 0; i < 15; ++i) {
#pragma unroll
                    for (int r = i + 1; r < 16; ++r) a[r] = a[r] - M[(g + i) * NBM + g + r] * a[i];
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) X[(g + r) * XS + lane] = a[r];
            }
            __syncthreads();
            const int g2 = g + 16 * (1 + w);
            if (g2 < NBM) {
                cf a[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) a[r] = X[(g2 + r) * XS + lane];
#pragma unroll 4
                for (int i = 0; i < 16; ++i) {
                    const cf xi = X[(g + i) * XS + lane];
#pragma unroll
                    for (int r = 0; r < 16; ++r) a[r] = a[r] - M[(g + i) * NBM + g2 + r] * xi;
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) X[(g2 + r) * XS + lane] = a[r];
            }
            __syncthreads();
        }
        // Li column col = lane's column: Li[col*64 + r] = X[r][col]
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int col = w + 4 * q;
            Li[col * NBM + lane] = X[lane * XS + col];
        }
    } else {
        // rows of I, solved x U = e_r left to right (lane = row r)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = w + 4 * q;
            X[c * XS + lane] = mkf(c == lane ? 1.f : 0.f, 0.f);
        }
        __syncthreads();
        if (tid < NBM) {
            const cf u = M[tid * NBM + tid];
            dinv[tid] = (u.x != 0.f || u.y != 0.f) ? cinvf(u) : mkf(0.f, 0.f);
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < NBM; g += 16) {
            if (w == 0) {
                cf a[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) a[c] = X[(g + c) * XS + lane];
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    a[c] = a[c] * dinv[g + c];
#pragma unroll
                    for (int c2 = c + 1; c2 < 16; ++c2) a[c2] = a[c2] - a[c] * M[(g + c2) * NBM + g + c];
                }
#pragma unroll
                for (int c = 0; c < 16; ++c) X[(g + c) * XS + lane] = a[c];
            }
            __syncthreads();
            const int g2 = g + 16 * (1 + w);
            if (g2 < NBM) {
                cf a[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) a[c] = X[(g2 + c) * XS + lane];
#pragma unroll 4
                for (int i = 0; i < 16; ++i) {
                    const cf xi = X[(g + i) * XS + lane];
#pragma unroll
                    for (int c = 0; c < 16; ++c) a[c] = a[c] - xi * M[(g2 + c) * NBM + g + i];
                }
#pragma unroll
                for (int c = 0; c < 16; ++c) X[(g2 + c) * XS + lane] = a[c];
            }
            __syncthreads();
        }
        // Ui[c*64 + r] = X[c][r]: row r = lane
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = w + 4 * q;
            Ui[c * NBM + lane] = X[c * XS + lane];
        }
    }
}

// Wave-uniform copy of lane src's complex double (v_readlane, no LDS).
__device__ inline cd readlane_c(cd v, int src) {
    long long bx = __double_as_longlong(v.x), by = __double_as_longlong(v.y);
    int xl = __builtin_amdgcn_readlane((int)bx, src), xh = __builtin_amdgcn_readlane((int)(bx >> 32), src);
    int yl = __builtin_amdgcn_readlane((int)by, src), yh = __builtin_amdgcn_readlane((int)(by >> 32), src);
    return mk(__longlong_as_double(((long long)xh << 32) | (unsigned)xl),
              __longlong_as_double(((long long)yh << 32) | (unsigned)yl));
}

// Wave-level wait for a block flag: every lane polls the same address (one
// transaction), then one agent-scope acquire.
__device__ inline void wave_wait_ge(int* flag, int target, int* status) {
    wait_ge(flag, target, status);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// ---------------------------------------------------------------------------
// Preconditioner block solves over 128-row blocks (two factor blocks each),
// which halves the number of block hand-offs of the dependency chain.
// k_pair_inverses builds, for pair p = (blocks a = 2p, b = 2p+1), the
// inverses of the 128 x 128 diagonal blocks of both triangular factors from
// the 64 x 64 inverses and the coupling blocks:
//   lower  [Li_a, 0; X, Li_b],  X = -Li_b L_ba Li_a
//   upper  [Ui_a, Y; 0, Ui_b],  Y = -Ui_a U_ab Ui_b
// (Li = L^-1 P of a block, Ui = U^-1; L_ba / U_ab as stored in A, rows of
// block b in their unpermuted order, which is how the block solves use
// them).  Column-major 128 x 128; a missing or partial block b contributes
// zero coupling.
// ---------------------------------------------------------------------------
constexpr int NB2 = 2 * NBM;

// Grid (2, pairs): x = 0 lower, 1 upper; y = pair.
__global__ __launch_bounds__(256) void k_pair_inverses(const cf* __restrict__ A, size_t lda, int n,
                                                       const cf* __restrict__ inv, int nblk,
                                                       cf* __restrict__ Ml, cf* __restrict__ Mu) {
    __shared__ cf S0[NBM * NBM], S1[NBM * NBM], T[NBM * NBM];
    const int tid = threadIdx.x;
    const int p = blockIdx.y;
    const bool lower = blockIdx.x == 0;
    const int a = 2 * p, b = 2 * p + 1;
    const int a0 = a * NBM, b0 = b * NBM;
    const int nbb = b < nblk ? min(NBM, n - b0) : 0;  // rows/columns of block b
    const cf* Ia = inv + (size_t)(lower ? a : nblk + a) * NBM * NBM;
    const cf* Ib = inv + (size_t)(lower ? b : nblk + b) * NBM * NBM;
    cf* M = (lower ? Ml : Mu) + (size_t)p * NB2 * NB2;
    // the two diagonal quadrants, and zero for the one that stays zero
    for (int t = tid; t < NBM * NBM; t += 256) {
        const int c = t >> 6, r = t & 63;
        M[(size_t)c * NB2 + r] = Ia[t];
        M[(size_t)(NBM + c) * NB2 + NBM + r] =
            nbb > 0 ? Ib[t] : mkf(r == c ? 1.f : 0.f, 0.f);
        if (lower) M[(size_t)(NBM + c) * NB2 + r] = mkf(0.f, 0.f);
        else M[(size_t)c * NB2 + NBM + r] = mkf(0.f, 0.f);
    }
    if (nbb == 0) {
        for (int t = tid; t < NBM * NBM; t += 256) {
            const int c = t >> 6, r = t & 63;
            if (lower) M[(size_t)c * NB2 + NBM + r] = mkf(0.f, 0.f);
            else M[(size_t)(NBM + c) * NB2 + r] = mkf(0.f, 0.f);
        }
        return;
    }
    // lower: S0 = L_ba (rows b, cols a), S1 = Li_a;  T = S0 S1;  then X = -Li_b T
    // upper: S0 = Ui_a, S1 = U_ab (rows a, cols b);  T = S0 S1;  then Y = -T Ui_b
    for (int t = tid; t < NBM * NBM; t += 256) {
        const int c = t >> 6, r = t & 63;
        if (lower) {
            S0[t] = r < nbb ? A[(size_t)(a0 + c) * lda + b0 + r] : mkf(0.f, 0.f);
            S1[t] = Ia[t];
        } else {
            S0[t] = Ia[t];
            S1[t] = c < nbb ? A[(size_t)(b0 + c) * lda + a0 + r] : mkf(0.f, 0.f);
        }
    }
    __syncthreads();
    // thread: column c = tid >> 2, rows r = (tid & 3) + 4 i
    const int c = tid >> 2, rq = tid & 3;
    {
        cf acc[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = mkf(0.f, 0.f);
        for (int k = 0; k < NBM; ++k) {
            const cf bk = S1[c * NBM + k];
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[i] = acc[i] + S0[k * NBM + rq + 4 * i] * bk;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) T[c * NBM + rq + 4 * i] = acc[i];
    }
    __syncthreads();
    {
        // lower: X = -Li_b T (Li_b from global, column-major); upper: Y = -T Ui_b
        cf acc[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = mkf(0.f, 0.f);
        for (int k = 0; k < NBM; ++k) {
            if (lower) {
                const cf bk = T[c * NBM + k];
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[i] = acc[i] + Ib[k * NBM + rq + 4 * i] * bk;
            } else {
                const cf bk = Ib[c * NBM + k];
#pragma unroll
                for (int i = 0; i < 16; ++i) acc[i] = acc[i] + T[k * NBM + rq + 4 * i] * bk;
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = rq + 4 * i;
            const cf v = mkf(-acc[i].x, -acc[i].y);
            if (lower) M[(size_t)c * NB2 + NBM + r] = v;
            else M[(size_t)(NBM + c) * NB2 + r] = v;
        }
    }
}

// Block diagonal step of the 128-row block solves: out = Minv acc, the
// eight waves each taking 32 columns of one row half; joined through LDS,
// the two row halves stored by waves 0 and 4, one flag published.
__device__ inline void diag128_apply_publish(const cf mv[32], const cd* acc, cd (*part)[NBM],
                                             int w, int lane, int nb, cd* out, int* flag,
                                             int epoch) {
    const int rh = w >> 2, q = w & 3;
    cd s = mk(0, 0);
#pragma unroll
    for (int i = 0; i < 32; ++i) s = s + up(mv[i]) * acc[32 * q + i];
    part[w][lane] = s;
    __syncthreads();
    if (q == 0) {
        const int row = rh * NBM + lane;
        const cd v = (part[4 * rh][lane] + part[4 * rh + 1][lane]) +
                     (part[4 * rh + 2][lane] + part[4 * rh + 3][lane]);
        if (row < nb) out[row] = v;
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0) publish(flag, epoch);
}

// y = L^-1 P b over 128-row blocks.  Workgroup g owns blocks g, g+G, ...;
// wave w = (row half rh, column quarter q) accumulates its 64 rows x 32
// columns of every finished block as that block's flag comes up (factor
// tile loaded before the wait), then y_blk = Minv acc.  Blocks only wait on
// lower-numbered blocks and every workgroup is resident (cooperative
// launch), so the chain always drains.
__global__ __launch_bounds__(512) void k_trsv_lower_coop(const cf* __restrict__ A, size_t lda,
                                                         int n, int nblk2,
                                                         const cf* __restrict__ Minv,
                                                         const cd* __restrict__ b,
                                                         cd* __restrict__ y, int* __restrict__ flags,
                                                         int epoch, int* __restrict__ status) {
    __shared__ cd part[8][NBM];
    __shared__ cd acc[NB2];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, rh = w >> 2, q = w & 3;
    for (int blk = blockIdx.x; blk < nblk2; blk += gridDim.x) {
        const int r0 = blk * NB2, nb = min(NB2, n - r0);
        const int row = rh * NBM + lane;
        cf mv[32];
        const cf* Mb = Minv + (size_t)blk * NB2 * NB2 + (size_t)(32 * q) * NB2 + row;
#pragma unroll
        for (int i = 0; i < 32; ++i) mv[i] = Mb[(size_t)i * NB2];
        cd s = mk(0, 0);
        const int rowc = r0 + min(row, nb - 1);
        for (int k = 0; k < blk; ++k) {
            const int c0 = k * NB2 + 32 * q;
            const cf* Lc = A + (size_t)c0 * lda + rowc;
            cf l[32];  // the factor tile does not depend on the flag: load it first
#pragma unroll
            for (int i = 0; i < 32; ++i) l[i] = Lc[(size_t)i * lda];
            wave_wait_ge(&flags[k], epoch, status);
            const cd yl = y[c0 + (lane & 31)];
#pragma unroll
            for (int i = 0; i < 32; ++i) s = s + up(l[i]) * readlane_c(yl, i);
        }
        part[w][lane] = s;
        __syncthreads();
        if (q == 0)
            acc[row] = row < nb ? b[r0 + row] - ((part[4 * rh][lane] + part[4 * rh + 1][lane]) +
                                                 (part[4 * rh + 2][lane] + part[4 * rh + 3][lane]))
                                : mk(0, 0);
        __syncthreads();
        diag128_apply_publish(mv, acc, part, w, lane, nb, y + r0, &flags[blk], epoch);
        __syncthreads();
    }
}

// x = U^-1 y over 128-row blocks (in reverse order), same scheme.
__global__ __launch_bounds__(512) void k_trsv_upper_coop(const cf* __restrict__ A, size_t lda,
                                                         int n, int nblk2,
                                                         const cf* __restrict__ Minv,
                                                         const cd* __restrict__ yv,
                                                         cd* __restrict__ x, int* __restrict__ flags,
                                                         int epoch, int* __restrict__ status) {
    __shared__ cd part[8][NBM];
    __shared__ cd acc[NB2];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, rh = w >> 2, q = w & 3;
    for (int bi = blockIdx.x; bi < nblk2; bi += gridDim.x) {
        const int blk = nblk2 - 1 - bi;
        const int r0 = blk * NB2, nb = min(NB2, n - r0);
        const int row = rh * NBM + lane;
        cf mv[32];
        const cf* Mb = Minv + (size_t)blk * NB2 * NB2 + (size_t)(32 * q) * NB2 + row;
#pragma unroll
        for (int i = 0; i < 32; ++i) mv[i] = Mb[(size_t)i * NB2];
        cd s = mk(0, 0);
        const int rowc = r0 + min(row, nb - 1);
        for (int k = nblk2 - 1; k > blk; --k) {
            const int kb = min(NB2, n - k * NB2);
            const int c0 = k * NB2 + 32 * q;
            const int cnt = min(32, kb - 32 * q);  // columns of this wave in block k
            const cf* Uc = A + (size_t)c0 * lda + rowc;
            if (cnt == 32) {
                cf u[32];  // the factor tile does not depend on the flag: load it first
#pragma unroll
                for (int i = 0; i < 32; ++i) u[i] = Uc[(size_t)i * lda];
                wave_wait_ge(&flags[k], epoch, status);
                const cd yl = x[c0 + (lane & 31)];
#pragma unroll
                for (int i = 0; i < 32; ++i) s = s + up(u[i]) * readlane_c(yl, i);
            } else {
                wave_wait_ge(&flags[k], epoch, status);
                if (cnt <= 0) continue;
                const cd yl = x[c0 + min(lane & 31, cnt - 1)];
                for (int i = 0; i < cnt; ++i) s = s + up(Uc[(size_t)i * lda]) * readlane_c(yl, i);
            }
        }
        part[w][lane] = s;
        __syncthreads();
        if (q == 0)
            acc[row] = row < nb ? yv[r0 + row] - ((part[4 * rh][lane] + part[4 * rh + 1][lane]) +
                                                  (part[4 * rh + 2][lane] + part[4 * rh + 3][lane]))
                                : mk(0, 0);
        __syncthreads();
        diag128_apply_publish(mv, acc, part, w, lane, nb, x + r0, &flags[blk], epoch);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// C[m x nc] -= L21 U12, complex FP32 with K = 64, split-BF16x3 on MFMA.  The
// operands come pre-split as bf16 planes (written by the two triangular
// kernels), so a lane loads its 8 consecutive k of one plane with a single
// 16-byte global load, straight into the MFMA operand layout; no LDS.
// Workgroup tile 64 x 64, 4 waves of 32 x 32 (2 x 2 MFMA 16x16 blocks).
// Each real product x*y is xh*yh + xh*yl + xl*yh (FP32 accumulation);
// complex = 4 real products, 12 MFMA per 16x16x32 complex block.
// v_mfma_f32_16x16x32_bf16 layout: A[i][k]: lane i + 16 (k/8), element k%8;
// B[k][j]: lane j + 16 (k/8), element k%8; D[i][j]: lane j + 16 (i/4), reg i%4.
// Tiles are assigned so that consecutive workgroups of one XCD (ids that
// differ by the XCD count) sweep neighbouring tiles and share operands in
// that XCD's L2.
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
constexpr int CT = 64;
constexpr int NXCD = 8;

__device__ inline bf8 neg_bf8(bf8 x) {
    u4 u = __builtin_bit_cast(u4, x);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf8, u);
}

// One 64 x 64 tile b of the update (tile order: row tiles fastest).
__device__ __forceinline__ void cgemm_tile(int b, int gx, int m, int nc,
                                           const __bf16* __restrict__ la, int ar0,
                                           const __bf16* __restrict__ ub, int bc0,
                                           size_t pstride, cf* __restrict__ C, size_t ldc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bm = (b % gx) * CT, bn = (b / gx) * CT;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    // operand fragments of both k-steps first, so that they and the C tile
    // below are in flight together: pa/pb[k-step][block][plane]
    bf8 pa[2][2][4], pb[2][2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const __bf16* p = la + (size_t)(ar0 + min(bm + wm + 16 * x + li, m - 1)) * CT + lk;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int pl = 0; pl < 4; ++pl) pa[ks][x][pl] = *(const bf8*)(p + 32 * ks + pl * pstride);
    }
#pragma unroll
    for (int y = 0; y < 2; ++y) {
        const __bf16* p = ub + (size_t)(bc0 + min(bn + wn + 16 * y + li, nc - 1)) * CT + lk;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int pl = 0; pl < 4; ++pl) pb[ks][y][pl] = *(const bf8*)(p + 32 * ks + pl * pstride);
    }
    f4 accR[2][2], accI[2][2];
    // C tile into the accumulators: a lane's 4 D registers are 4 consecutive
    // rows of one column, 32 contiguous bytes.  Interior tiles of an aligned
    // C take a branch-free path (all eight 16-byte loads in flight at once);
    // edge tiles go element by element.
    const bool fast = bm + CT <= m && bn + CT <= nc && (ldc & 3) == 0 &&
                      ((uintptr_t)C & 31) == 0;
    if (fast) {
        float4 v0[2][2], v1[2][2];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const cf* cp = C + (size_t)(bn + wn + 16 * y + li) * ldc + bm + wm + 16 * x +
                               (lane >> 4) * 4;
                v0[x][y] = *(const float4*)cp;
                v1[x][y] = *(const float4*)(cp + 2);
            }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                accR[x][y] = f4{v0[x][y].x, v0[x][y].z, v1[x][y].x, v1[x][y].z};
                accI[x][y] = f4{v0[x][y].y, v0[x][y].w, v1[x][y].y, v1[x][y].w};
            }
    } else {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int col = bn + wn + 16 * y + li;
                const int row0 = bm + wm + 16 * x + (lane >> 4) * 4;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    cf v = (row0 + r < m && col < nc) ? C[(size_t)col * ldc + row0 + r]
                                                      : cf{0.f, 0.f};
                    accR[x][y][r] = v.x;
                    accI[x][y][r] = v.y;
                }
            }
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        bf8 narh[2], narl[2], naih[2], nail[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            narh[x] = neg_bf8(pa[ks][x][0]);
            narl[x] = neg_bf8(pa[ks][x][1]);
            naih[x] = neg_bf8(pa[ks][x][2]);
            nail[x] = neg_bf8(pa[ks][x][3]);
        }
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const bf8 brh = pb[ks][y][0], brl = pb[ks][y][1], bih = pb[ks][y][2],
                      bil = pb[ks][y][3];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const bf8 aih = pa[ks][x][2], ail = pa[ks][x][3];
                f4 r = accR[x][y], i = accI[x][y];
                // Cr -= Ar Br - Ai Bi
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narh[x], brh, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narh[x], brl, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(narl[x], brh, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih, bih, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih, bil, r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ail, bih, r, 0, 0, 0);
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
    if (fast) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                cf* cp = C + (size_t)(bn + wn + 16 * y + li) * ldc + bm + wm + 16 * x +
                         (lane >> 4) * 4;
                const f4 r = accR[x][y], i = accI[x][y];
                *(float4*)cp = float4{r[0], i[0], r[1], i[1]};
                *(float4*)(cp + 2) = float4{r[2], i[2], r[3], i[3]};
            }
    } else {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int col = bn + wn + 16 * y + li;
                const int row0 = bm + wm + 16 * x + (lane >> 4) * 4;
                if (col >= nc) continue;
                const f4 r = accR[x][y], i = accI[x][y];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (row0 + q < m) C[(size_t)col * ldc + row0 + q] = cf{r[q], i[q]};
            }
    }
}

__global__ __launch_bounds__(256) void k_cgemm_planes(int m, int nc, const __bf16* __restrict__ la,
                                                      int ar0, const __bf16* __restrict__ ub,
                                                      int bc0, size_t pstride,
                                                      cf* __restrict__ C, size_t ldc) {
    const int gx = (m + CT - 1) / CT, gy = (nc + CT - 1) / CT;
    const int tiles = gx * gy;
    // XCD-aware tile order: workgroup id runs on XCD id % NXCD; give each XCD
    // a contiguous range of tiles.  A grid smaller than the tile count loops
    // (its size a multiple of NXCD keeps every id on its XCD), which leaves
    // CUs to the kernels of the other stream.
    const int per = (tiles + NXCD - 1) / NXCD;
    for (int id = blockIdx.x; id < per * NXCD; id += gridDim.x) {
        const int b = (id % NXCD) * per + id / NXCD;
        if (b < tiles) cgemm_tile(b, gx, m, nc, la, ar0, ub, bc0, pstride, C, ldc);
    }
}

// Number of workgroups of k_cgemm_planes for an m x nc update, at most
// `cap` (a multiple of NXCD; 0 = one workgroup per tile).
static int cgemm_groups(int m, int nc, int cap = 0) {
    const int tiles = ((m + CT - 1) / CT) * ((nc + CT - 1) / CT);
    const int g = ((tiles + NXCD - 1) / NXCD) * NXCD;
    return cap > 0 ? std::min(g, cap) : g;
}

// Workgroups of the bulk update on the side stream (BEMX_BULK_WGS; default
// 0 = one per tile).  A cap keeps the bulk update from holding every CU, so
// the panel kernels of the main stream find room sooner; measured on the
// 6480-DOF solve, caps of 256..512 cost more in the bulk update than they
// gained on the panel chain.
static int bulk_wg_cap() {
    static const int cap = [] {
        const char* v = getenv("BEMX_BULK_WGS");
        const int c = v ? atoi(v) : 0;
        return c > 0 ? ((c + NXCD - 1) / NXCD) * NXCD : 0;
    }();
    return cap;
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
enum { M_A32 = 9, M_PERM, M_FLAGS, M_STATUS, M_VEC, M_V, M_PART, M_HOST, M_PTMP, M_PLANES, M_INV,
       M_MINV };

struct MixedFactor {
    cf* A = nullptr;
    int* lperm = nullptr;
    int* flags = nullptr;
    int* status = nullptr;
    int* epoch = nullptr;   // host-side counter living in LuWork
    cf* inv = nullptr;      // per block: L_bb^-1 P_bb, then (after nblk blocks) U_bb^-1
    cf* minv = nullptr;     // per 128-row pair: lower inverse, then (after npair) upper
    int n = 0, nblk = 0, npair = 0, coop = 1;
    size_t ld = 0;  // leading dimension of A
};

// The launch sequence of the factorisation (no host synchronisation, no
// allocation), so it can be captured into a graph.
static hipError_t enqueue_factor(LuWork& w, hipStream_t st, int n, const cd* A64, size_t lda,
                                 const MixedFactor& F, __bf16* planes, size_t pstride) {
    hipError_t first = hipSuccess;
    auto chk = [&](hipError_t e) {
        if (first == hipSuccess && e != hipSuccess) first = e;
    };
    const int nblk = F.nblk;
    hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, st, F.status, 0, 1);
    size_t tot = (size_t)n * n;
    const size_t ld = F.ld;
    hipLaunchKernelGGL(k_cvt_z2c, dim3((tot + 255) / 256), dim3(256), 0, st, A64, lda, F.A,
                       ld, n);
    // Look-ahead over two streams.  Step k (block columns: k, N = k+1, R =
    // the rest):
    //   st:   getrf(k), L21(k); [wait bulk(k-1)] U12(k) on N, update of N
    //   side: [wait L21(k)] U12(k) on R, update of R (the bulk)
    // so getrf(k+1) and L21(k+1) run while the bulk update of step k does.
    const hipStream_t sd = w.side;
    bool bulk_pending = false;
    for (int k = 0; k < nblk; ++k) {
        const int k0 = k * NBM, nb = std::min(NBM, n - k0);
        int* lp = F.lperm + (size_t)k * NBM;
        __bf16* la = planes + (size_t)(k & 1) * 8 * pstride;
        __bf16* ub = la + 4 * pstride;
        hipLaunchKernelGGL(k_diag_getrf32, dim3(1), dim3(256), 0, st, F.A, ld, k0, nb, lp,
                           F.status);
        cf* li = F.inv + (size_t)k * NBM * NBM;
        cf* ui = F.inv + (size_t)(nblk + k) * NBM * NBM;
        const int c0 = k0 + nb, rest = n - c0;
        if (rest <= NBM) {
            // no bulk update left to hide it behind: inverses on the main stream
            hipLaunchKernelGGL(k_diag_inverses, dim3(2), dim3(256), 0, st, (const cf*)F.A, ld, k0,
                               nb, (const int*)lp, li, ui);
        }
        if (rest <= 0) break;
        hipLaunchKernelGGL(k_trsm_upper_rows, dim3((rest + TRR_ROWS - 1) / TRR_ROWS), dim3(256),
                           0, st, F.A, ld, k0, nb, c0, rest, la, pstride);
        const int nn1 = std::min(NBM, rest), nr = rest - nn1;
        if (nr > 0) chk(hipEventRecord(w.ev_panel, st));
        if (bulk_pending) chk(hipStreamWaitEvent(st, w.ev_bulk, 0));
        hipLaunchKernelGGL((k_trsm_lower_cols<cf, true>),
                           dim3((nn1 + TRSM_COLS_PER_WG - 1) / TRSM_COLS_PER_WG), dim3(256), 0,
                           st, F.A, ld, k0, nb, c0, nn1, (const int*)lp, ub, pstride);
        hipLaunchKernelGGL(k_cgemm_planes, dim3(cgemm_groups(rest, nn1)), dim3(256), 0, st, rest,
                           nn1, (const __bf16*)la, c0, (const __bf16*)ub, c0, pstride,
                           F.A + c0 * ld + c0, ld);
        bulk_pending = false;
        if (nr > 0) {
            const int cr = c0 + nn1;
            chk(hipStreamWaitEvent(sd, w.ev_panel, 0));
            hipLaunchKernelGGL(k_diag_inverses, dim3(2), dim3(256), 0, sd, (const cf*)F.A, ld, k0,
                               nb, (const int*)lp, li, ui);
            hipLaunchKernelGGL((k_trsm_lower_cols<cf, true>),
                               dim3((nr + TRSM_COLS_PER_WG - 1) / TRSM_COLS_PER_WG), dim3(256), 0,
                               sd, F.A, ld, k0, nb, cr, nr, (const int*)lp, ub, pstride);
            hipLaunchKernelGGL(k_cgemm_planes, dim3(cgemm_groups(rest, nr, bulk_wg_cap())), dim3(256), 0, sd,
                               rest, nr, (const __bf16*)la, c0, (const __bf16*)ub, cr, pstride,
                               F.A + cr * ld + c0, ld);
            chk(hipEventRecord(w.ev_bulk, sd));
            bulk_pending = true;
        }
    }
    if (bulk_pending) chk(hipStreamWaitEvent(st, w.ev_bulk, 0));
    // 128-row diagonal-block inverses of both factors for the block solves
    hipLaunchKernelGGL(k_pair_inverses, dim3(2, F.npair), dim3(256), 0, st, (const cf*)F.A, ld,
                       n, (const cf*)F.inv, nblk, F.minv, F.minv + (size_t)F.npair * NB2 * NB2);
    chk(hipGetLastError());
    return first;
}

static int factor_mixed(LuWork& w, hipStream_t st, int n, const cd* A64, size_t lda,
                        MixedFactor& F, std::string& err) {
    hipError_t e;
    const int nblk = (n + NBM - 1) / NBM;
    F.n = n;
    F.nblk = nblk;
    F.ld = ((size_t)n + 31) & ~(size_t)31;  // 16-byte aligned 4-row groups
    F.A = (cf*)w.get(M_A32, sizeof(cf) * F.ld * n, &e);
    MX_CHK(e);
    F.lperm = (int*)w.get(M_PERM, sizeof(int) * (size_t)nblk * NBM, &e);
    MX_CHK(e);
    F.inv = (cf*)w.get(M_INV, sizeof(cf) * (size_t)nblk * NBM * NBM * 2, &e);
    MX_CHK(e);
    F.npair = (nblk + 1) / 2;
    F.minv = (cf*)w.get(M_MINV, sizeof(cf) * (size_t)F.npair * NB2 * NB2 * 2, &e);
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
    int dev = 0, ncu = 0, per = 0;
    MX_CHK(hipGetDevice(&dev));
    MX_CHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    MX_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_trsv_lower_coop, 512, 0));
    int per2 = 0;
    MX_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, k_trsv_upper_coop, 512, 0));
    F.coop = std::max(1, std::min(F.npair, ncu * std::min(per, per2)));
    MX_CHK(w.ensure_side());
    // bf16 planes of L21 and U12, double-buffered by step parity: step k+1
    // writes one buffer while the bulk update of step k still reads the other
    const size_t pstride = (size_t)n * NBM;
    __bf16* planes = (__bf16*)w.get(M_PLANES, sizeof(__bf16) * pstride * 4 * 4, &e);
    MX_CHK(e);
    // The launch sequence depends only on n and the buffer addresses: it is
    // captured once into a graph and replayed while those stay the same.
    const void* key[8] = {(const void*)(size_t)n, A64, (const void*)lda, F.A, planes, F.lperm,
                          F.status, w.side};
    if (w.use_graphs) {
        if (!w.gexec || memcmp(key, w.gkey, sizeof key) != 0) {
            if (w.gexec) MX_CHK(hipGraphExecDestroy(w.gexec));
            w.gexec = nullptr;
            hipGraph_t g = nullptr;
            MX_CHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const hipError_t e1 = enqueue_factor(w, st, n, A64, lda, F, planes, pstride);
            const hipError_t e2 = hipStreamEndCapture(st, &g);
            if (e1 != hipSuccess || e2 != hipSuccess) {
                if (g) hipGraphDestroy(g);
                MX_CHK(e1 != hipSuccess ? e1 : e2);
            }
            const hipError_t e3 = hipGraphInstantiate(&w.gexec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            MX_CHK(e3);
            memcpy(w.gkey, key, sizeof key);
        }
        MX_CHK(hipGraphLaunch(w.gexec, st));
    } else {
        MX_CHK(enqueue_factor(w, st, n, A64, lda, F, planes, pstride));
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
    int n = F.n, nblk = F.npair;
    size_t lda = F.ld;
    const cf* A = F.A;
    const cf* li = F.minv;
    const cf* ui = F.minv + (size_t)F.npair * NB2 * NB2;
    int* flags = F.flags;
    int* status = F.status;
    int e1 = ++*F.epoch;
    void* a1[] = {(void*)&A, &lda, &n, &nblk, (void*)&li, (void*)&v, &tmp, &flags, &e1, &status};
    hipError_t e = hipLaunchCooperativeKernel((const void*)k_trsv_lower_coop, dim3(F.coop),
                                              dim3(512), a1, 0, st);
    if (e != hipSuccess) return e;
    int e2 = ++*F.epoch;
    const cd* tv = tmp;
    void* a2[] = {(void*)&A, &lda, &n, &nblk, (void*)&ui, (void*)&tv, &z, &flags, &e2, &status};
    return hipLaunchCooperativeKernel((const void*)k_trsv_upper_coop, dim3(F.coop), dim3(512), a2,
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
