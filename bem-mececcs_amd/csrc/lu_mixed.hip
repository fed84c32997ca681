// lu_mixed.hip -- mixed-precision solve: FP32 complex LU whose trailing update
// runs as split-BF16x3 products on MFMA, used as the preconditioner of an
// FP64 GMRES iterative refinement (GMRES-IR) against the FP64 matrix.
//
// Factorisation (blocks of NBM = 64 columns, A32 = (complex float) A64), in
// the block form A = L U with unit block-lower L and block-upper U whose
// diagonal blocks are kept as explicit inverses:
//   k_diag_gj / gj_factor_body  Gauss-Jordan elimination of the 64 x 64
//                      diagonal block with partial pivoting inside the block
//                      (registers: lane = row, wave = a panel of columns),
//                      recorded (multipliers, pivots)
//   k_cgemm_update_inverse  the next block's elimination (workgroup 0) and
//                      its inverse Xinv from the record (8 more workgroups)
//                      beside the update of the current step
//   k_panel_inv        L21 = A21 Xinv (MFMA) for all rows below and the
//                      split planes of U12 = A12 for all columns right
//   k_cgemm_update2    A22 -= L21 U12 over K = 64 .. 512; every FP32 operand x
//                      is split into hi = bf16(x), lo = bf16(x - hi) (by
//                      k_panel_inv, into bf16 planes) and each real
//                      product is hi*hi + hi*lo + lo*hi on
//                      v_mfma_f32_16x16x32_bf16 with FP32 accumulation;
//                      complex = 4 real products.
// Blocks are factored in groups (enqueue_factor): inside a group the update
// is applied to the group's own columns and rows only, in the same launch as
// the elimination of the next diagonal block (k_cgemm_update_inverse); the rest
// of the matrix gets one update per group with K = 64 x group size, on a
// second stream, while the next group is factored.
// Preconditioner M^-1 v (FP64 vectors, FP32 factors): k_trsv_coop, one
// cooperative launch for the lower and then the upper block solve; two
// workgroups own each 128-row block (pair of diagonal blocks), 64 rows each,
// accumulate the contributions of finished blocks as their flags come up and
// publish their half, so the dependency chain costs one 64 x 128 product
// with a precomputed coupling block (k_pair_coupling) per 128 rows; the
// 128-row inverses come from k_pair_inverses after the factorisation.
// Pivoting is restricted to the diagonal block: the factors only
// precondition.  The accuracy of the answer comes from the refinement, whose
// residual b - A64 x is formed in FP64 with the FP64 matrix; if it stalls the
// caller falls back to the FP64 partial-pivot LU.
#include <algorithm>
#include <cmath>
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

// Batches.  B problems of one order share every launch of the factorisation
// and the refinement; problem z is blockIdx.z (k_pair_coupling: z / 4).  Each
// problem's workspace is one arena, the arenas bs bytes apart, so a kernel
// moves every workspace pointer it is given by z * bs; the caller's FP64
// matrices and right-hand sides have strides of their own.  A problem's
// arithmetic does not depend on B or on its place in the batch.
template <class T>
__device__ __forceinline__ T* zp(T* p, size_t bs, int z) {
    return p ? (T*)((const char*)p + (size_t)z * bs) : p;
}

__global__ void k_cvt_z2c(const cd* __restrict__ a, size_t lda, cf* __restrict__ o, size_t ldo,
                          int n, size_t sa, size_t bs) {
    a = zp(a, sa, blockIdx.z);
    o = zp(o, bs, blockIdx.z);
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

// Gauss-Jordan elimination of the 64 x 64 diagonal block with partial
// pivoting (largest |re|+|im| to 17 mantissa bits among the rows not yet
// pivots, lowest row on ties), held in registers with lane r owning row r.
// Rows are not moved and not normalised: step j picks pivot row p_j (rank
// j) and subtracts l_r = a_rj / a_pj times row p_j from every other row r
// (rows above the pivots too, l_p = 0), so the A part ends as the pivots d_j
// at (p_j, j).  The record of the elimination -- lmul[j NBM + r] = l_r of
// step j, piv[j] = p_j, inv[j] = 1 / d_j -- determines A^-1: applied to the
// identity it gives E, and row j of A^-1 is row p_j of E times inv[j]
// (gj_identity_cols).  A partial block (nb < 64) is padded with the
// identity, so its inverse is padded likewise.
// Wave q owns the A columns [PW q, PW q + PW) and factors them as a panel
// alone (the searches and the panel updates stay in the wave), publishing
// each step as soon as its panel is updated: the record in LDS (and in rec,
// global memory, for the kernels that build the inverse), then the step
// count s_done with a workgroup-scope release.  The waves to its right apply
// published steps to their columns as they come (polling s_done, sleeping
// between polls), so when wave q publishes its last step, wave q+1 has that
// one step left to apply before it factors the next panel: no barrier on
// the chain.  The factoring wave raises its issue priority over the
// replaying waves that share its SIMD.
// The body runs on 64 * NBM / PW threads; lmul (NBM x NBM), s_inv, s_piv
// (NBM each) and s_done are LDS.
// Bounded wait until *flag >= target (see wait_count in lu.hip); relaxed
// polls, the caller issues one acquire fence after the wait.  Another
// workgroup's time-out (status) is looked at every 64 polls only, so a poll
// is one memory round trip.
__device__ inline void wait_ge(int* flag, int target, int* status) {
    int it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if ((++it & 63) == 0 &&
            (it > FLAG_SPIN_LIMIT ||
             (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
              ST_SYNC_TIMEOUT))) {
            atomicOr(status, ST_SYNC_TIMEOUT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ inline void publish(int* flag, int value) {
    __hip_atomic_store(flag, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-level wait for a block flag: every lane polls the same address (one
// transaction), then one agent-scope acquire.
__device__ inline void wave_wait_ge(int* flag, int target, int* status) {
    wait_ge(flag, target, status);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

constexpr size_t REC_BYTES = sizeof(cf) * (NBM * NBM + NBM) + sizeof(int) * NBM;
__device__ inline cf* rec_lmul(char* rec) { return (cf*)rec; }
__device__ inline cf* rec_inv(char* rec) { return (cf*)rec + NBM * NBM; }
__device__ inline int* rec_piv(char* rec) { return (int*)((cf*)rec + NBM * NBM + NBM); }

template <int PW>
__device__ __forceinline__ void gj_factor_body(const cf* __restrict__ A, size_t lda, int nb,
                                               char* __restrict__ rec, int* __restrict__ status,
                                               cf* lmul, cf* s_inv, int* s_piv, int* s_done,
                                               int* __restrict__ gpan = nullptr) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int q = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int first = PW * q;  // this wave's first column
    cf v[PW];
#pragma unroll
    for (int i = 0; i < PW; ++i) {
        const int c = first + i;
        v[i] = (lane < nb && c < nb) ? A[(size_t)c * lda + lane] : mkf(lane == c ? 1.f : 0.f, 0.f);
    }
    if (tid == 0) *s_done = 0;
    __syncthreads();  // the count is zero before any wave polls it
    cf* __restrict__ g_lmul = rec_lmul(rec);
    cf* __restrict__ g_inv = rec_inv(rec);
    int* __restrict__ g_piv = rec_piv(rec);
    // the panels to the left, as they are published
    bool used = false;
    int avail = 0;
    for (int j = 0; j < first; ++j) {
        if (j >= avail) {
            int spins = 0;
            while ((avail = __hip_atomic_load(s_done, __ATOMIC_ACQUIRE,
                                              __HIP_MEMORY_SCOPE_WORKGROUP)) <= j) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > FLAG_SPIN_LIMIT) {
                    if (lane == 0) atomicOr(status, ST_SYNC_TIMEOUT);
                    avail = NBM;
                    break;
                }
            }
        }
        const int p = __builtin_amdgcn_readfirstlane(s_piv[j]);
        const cf l = lmul[j * NBM + lane];
        used = used || lane == p;
#pragma unroll
        for (int c = 0; c < PW; ++c) v[c] = cfms(v[c], l, readlane_f(v[c], p));
    }
    // this wave's panel
    __builtin_amdgcn_s_setprio(2);
    bool zero = false;
#pragma unroll
    for (int i = 0; i < PW; ++i) {
        const int j = first + i;
        // key = |re|+|im| with its 6 lowest mantissa bits replaced by 63 -
        // lane: the wave maximum picks the largest candidate, lowest row on
        // ties
        const unsigned key = used ? 0u
                                  : ((__float_as_uint(fabsf(v[i].x) + fabsf(v[i].y)) & ~63u) |
                                     (unsigned)(63 - lane));
        const int p = 63 - (int)(wave_max_u32(key) & 63u);
        // branch-free from here to the next search (an exactly zero pivot
        // gives zero multipliers and is reported)
        const cf pv = readlane_f(v[i], p);
        const float d2 = pv.x * pv.x + pv.y * pv.y;
        zero |= d2 == 0.f;
        const float rd = d2 > 0.f ? __builtin_amdgcn_rcpf(d2) : 0.f;
        const cf inv = mkf(pv.x * rd, -pv.y * rd);
        const bool piv = lane == p;
        const cf l = piv ? mkf(0.f, 0.f) : v[i] * inv;
        used = used || piv;
        lmul[j * NBM + lane] = l;
        g_lmul[j * NBM + lane] = l;
        if (lane == 0) {
            s_piv[j] = p;
            s_inv[j] = inv;
            g_piv[j] = p;
            g_inv[j] = inv;
        }
#pragma unroll
        for (int c = i + 1; c < PW; ++c) v[c] = cfms(v[c], l, readlane_f(v[c], p));
        // publish after the panel update, when the LDS writes have long
        // completed (the release waits for them); the empty asm keeps the
        // updated columns ahead of that wait
#pragma unroll
        for (int c = i + 1; c < PW; ++c) asm volatile("" ::"v"(v[c].x), "v"(v[c].y) : "memory");
        if (lane == 0)
            __hip_atomic_store(s_done, j + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __builtin_amdgcn_s_setprio(0);
    if (zero && lane == 0 && first < nb) atomicOr(status, ST_PIVOT_ZERO);
    // the panel's record is in memory for other workgroups (an agent-scope
    // release of this wave's own record stores, after the next wave has
    // already been handed the panel through LDS)
    if (gpan && lane == 0) publish(&gpan[q], 1);
}

// Columns [c0e, c0e + EC) of E (the identity under the 64 recorded steps;
// record in LDS), lane = row of E; rank = the step whose pivot is this
// lane's row.  Column m stays e_m (steps with another pivot row leave it
// alone) until row m is the pivot, so a column is updated only from that
// step on (act: the wave-uniform set of such columns).  Row rank of A^-1 is
// then e[] * inv[rank].
template <int EC>
__device__ __forceinline__ void gj_identity_cols(const cf* lmul, const int* piv, int c0e,
                                                 cf (&e)[EC], int& rank) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < EC; ++i) e[i] = mkf(lane == c0e + i ? 1.f : 0.f, 0.f);
    rank = lane;
    unsigned act = 0;
    for (int j = 0; j < NBM; ++j) {
        const int p = __builtin_amdgcn_readfirstlane(piv[j]);
        const cf l = lmul[j * NBM + lane];
        rank = lane == p ? j : rank;
        const int m = p - c0e;
        if (m >= 0 && m < EC) act |= 1u << m;
#pragma unroll
        for (int c = 0; c < EC; ++c)
            if ((act >> c) & 1u) e[c] = cfms(e[c], l, readlane_f(e[c], p));
    }
}

// The elimination of the diagonal block and, when the inverse is not built
// from the record by the next step's k_panel_inv (the last block), the
// inverse itself into Xinv (column-major, leading dimension NBM) after one
// barrier, every wave taking 64 / waves columns.
template <int PW>
__device__ __forceinline__ void gj_block(const cf* __restrict__ A, size_t lda, int nb,
                                         char* __restrict__ rec, cf* __restrict__ Xinv,
                                         bool with_inverse, int* __restrict__ status, cf* lmul,
                                         cf* s_inv, int* s_piv, int* s_done) {
    gj_factor_body<PW>(A, lda, nb, rec, status, lmul, s_inv, s_piv, s_done);
    if (!with_inverse) return;
    __syncthreads();
    constexpr int EC = PW;  // 64 / (64 / PW waves)
    const int q = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    cf e[EC];
    int rank;
    gj_identity_cols<EC>(lmul, s_piv, EC * q, e, rank);
    const cf sc = s_inv[rank];
#pragma unroll
    for (int i = 0; i < EC; ++i) Xinv[(size_t)(EC * q + i) * NBM + rank] = e[i] * sc;
}

template <int PW>
__global__ __launch_bounds__(64 * (NBM / PW)) void k_diag_gj(const cf* __restrict__ A, size_t lda,
                                                             int nb, char* __restrict__ rec,
                                                             cf* __restrict__ Xinv,
                                                             int with_inverse,
                                                             int* __restrict__ status, size_t bs) {
    __shared__ cf lmul[NBM * NBM];
    __shared__ cf s_inv[NBM];
    __shared__ int s_piv[NBM];
    __shared__ int s_done;
    const int z = blockIdx.z;
    gj_block<PW>(zp(A, bs, z), lda, nb, zp(rec, bs, z), zp(Xinv, bs, z), with_inverse != 0,
                 zp(status, bs, z), lmul, s_inv, s_piv, &s_done);
}

// ---------------------------------------------------------------------------
// Preconditioner block solves over 128-row blocks (two factor blocks each),
// which halves the number of block hand-offs of the dependency chain.
// k_pair_inverses builds, for pair p = (blocks a = 2p, b = 2p+1), the
// inverses of the 128 x 128 diagonal blocks of both block factors from the
// 64 x 64 diagonal-block inverses and the coupling blocks:
//   lower  [I, 0; -L_ba, I]     (unit diagonal blocks)
//   upper  [Ui_a, Y; 0, Ui_b],  Y = -Ui_a U_ab Ui_b
// (Ui = inverse of the block's diagonal block; L_ba / U_ab as stored in A).
// Column-major 128 x 128; a missing or partial block b contributes zero
// coupling.
// ---------------------------------------------------------------------------
constexpr int NB2 = 2 * NBM;

// Grid (2, pairs): x = 0 lower, 1 upper; y = pair.
__global__ __launch_bounds__(256) void k_pair_inverses(const cf* __restrict__ A, size_t lda, int n,
                                                       const cf* __restrict__ inv, int nblk,
                                                       cf* __restrict__ Ml, cf* __restrict__ Mu,
                                                       size_t bs) {
    __shared__ cf S0[NBM * NBM], S1[NBM * NBM], T[NBM * NBM];
    A = zp(A, bs, blockIdx.z);
    inv = zp(inv, bs, blockIdx.z);
    Ml = zp(Ml, bs, blockIdx.z);
    Mu = zp(Mu, bs, blockIdx.z);
    const int tid = threadIdx.x;
    const int p = blockIdx.y;
    const bool lower = blockIdx.x == 0;
    const int a = 2 * p, b = 2 * p + 1;
    const int a0 = a * NBM, b0 = b * NBM;
    const int nbb = b < nblk ? min(NBM, n - b0) : 0;  // rows/columns of block b
    cf* M = (lower ? Ml : Mu) + (size_t)p * NB2 * NB2;
    if (lower) {
        for (int t = tid; t < NBM * NBM; t += 256) {
            const int c = t >> 6, r = t & 63;
            const cf one = mkf(r == c ? 1.f : 0.f, 0.f);
            M[(size_t)c * NB2 + r] = one;
            M[(size_t)(NBM + c) * NB2 + NBM + r] = one;
            M[(size_t)(NBM + c) * NB2 + r] = mkf(0.f, 0.f);
            cf x = mkf(0.f, 0.f);
            if (r < nbb) {
                x = A[(size_t)(a0 + c) * lda + b0 + r];
                x = mkf(-x.x, -x.y);
            }
            M[(size_t)c * NB2 + NBM + r] = x;
        }
        return;
    }
    const cf* Ia = inv + (size_t)a * NBM * NBM;
    const cf* Ib = inv + (size_t)(b < nblk ? b : a) * NBM * NBM;
    // upper: the two diagonal quadrants, zero below them
    for (int t = tid; t < NBM * NBM; t += 256) {
        const int c = t >> 6, r = t & 63;
        M[(size_t)c * NB2 + r] = Ia[t];
        M[(size_t)(NBM + c) * NB2 + NBM + r] = nbb > 0 ? Ib[t] : mkf(r == c ? 1.f : 0.f, 0.f);
        M[(size_t)c * NB2 + NBM + r] = mkf(0.f, 0.f);
        if (nbb == 0) M[(size_t)(NBM + c) * NB2 + r] = mkf(0.f, 0.f);
    }
    if (nbb == 0) return;
    // S0 = Ui_a, S1 = U_ab (rows a, cols b);  T = S0 S1;  then Y = -T Ui_b
    for (int t = tid; t < NBM * NBM; t += 256) {
        const int c = t >> 6, r = t & 63;
        S0[t] = Ia[t];
        S1[t] = c < nbb ? A[(size_t)(b0 + c) * lda + a0 + r] : mkf(0.f, 0.f);
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
        cf acc[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = mkf(0.f, 0.f);
        for (int k = 0; k < NBM; ++k) {
            const cf bk = Ib[c * NBM + k];
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[i] = acc[i] + T[k * NBM + rq + 4 * i] * bk;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
            M[(size_t)(NBM + c) * NB2 + rq + 4 * i] = mkf(-acc[i].x, -acc[i].y);
    }
}



// Coupling blocks for the block solves: lower, pair p >= 1,
// Wl_p = Ml_p L_{p,p-1}; upper, pair p < npair - 1, Wu_p = Mu_p U_{p,p+1}
// (Ml/Mu the 128 x 128 inverses from k_pair_inverses, L/U blocks of A as the
// block solves read them; a partial pair's missing rows (lower) or columns
// (upper) are zero).  With them a block's solution is
//   y_p = Ml_p (b_p - sum_{k < p-1} L_{p,k} y_k) - Wl_p y_{p-1}
// so only the last product waits for the previous block.  Grid (2, npair,
// 4 B): z = 4 problem + 32-column quarter of W; 256 threads of 4 rows x 4
// columns.
__global__ __launch_bounds__(256) void k_pair_coupling(const cf* __restrict__ A, size_t lda,
                                                       int n, int npair,
                                                       const cf* __restrict__ Ml,
                                                       const cf* __restrict__ Mu,
                                                       cf* __restrict__ Wl, cf* __restrict__ Wu,
                                                       size_t bs) {
    __shared__ cf Ms[16][NB2];
    __shared__ cf Bs[16][32];
    const int z = blockIdx.z >> 2;
    A = zp(A, bs, z);
    Ml = zp(Ml, bs, z);
    Mu = zp(Mu, bs, z);
    Wl = zp(Wl, bs, z);
    Wu = zp(Wu, bs, z);
    const bool lower = blockIdx.x == 0;
    const int p = blockIdx.y, cq = blockIdx.z & 3;
    if (lower ? p == 0 : p == npair - 1) return;
    const int tid = threadIdx.x, tr = tid & 31, tc = tid >> 5;
    const cf* M = (lower ? Ml : Mu) + (size_t)p * NB2 * NB2;
    cf* W = (lower ? Wl : Wu) + (size_t)p * NB2 * NB2;
    const int rb = p * NB2;                          // rows of B: pair p
    const int cb = (lower ? p - 1 : p + 1) * NB2;    // columns of B
    const int nbr = min(NB2, n - rb), nbc = min(NB2, n - cb);
    cf acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = mkf(0.f, 0.f);
    for (int k0 = 0; k0 < NB2; k0 += 16) {
        for (int t = tid; t < 16 * NB2; t += 256) {
            const int kk = t >> 7, r = t & (NB2 - 1);
            Ms[kk][r] = M[(size_t)(k0 + kk) * NB2 + r];
        }
        for (int t = tid; t < 16 * 32; t += 256) {
            const int kk = t & 15, c = t >> 4;
            const int k = k0 + kk, col = cq * 32 + c;
            Bs[kk][c] = (k < nbr && col < nbc) ? A[(size_t)(cb + col) * lda + rb + k]
                                              : mkf(0.f, 0.f);
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 16; ++kk) {
            cf m[4], bv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) m[r] = Ms[kk][4 * tr + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][4 * tc + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = acc[r][c] + m[r] * bv[c];
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            W[(size_t)(cq * 32 + 4 * tc + c) * NB2 + 4 * tr + r] = acc[r][c];
}

// Hand-off of a block's solution to the next blocks of the chain without a
// separate flag: entry i also goes to ll[4 i .. 4 i + 3] as four 64-bit words
// {32-bit half of re / im, tag}, each stored as one agent-scope atomic, and
// the reader polls the words themselves until all four carry its tag (the
// launch epoch), so seeing the data and reading it is one memory round trip
// instead of flag, fence, data.
typedef unsigned long long u64;
__device__ inline void ll_put(u64* __restrict__ ll, int i, cd v, int tag) {
    const u64 t = (u64)(unsigned)tag << 32;
    const u64 bx = (u64)__double_as_longlong(v.x), by = (u64)__double_as_longlong(v.y);
    u64* p = ll + 4 * (size_t)i;
    __hip_atomic_store(p + 0, t | (bx & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, t | (bx >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 2, t | (by & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 3, t | (by >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Entry i of this lane once every lane of the wave has its entry with the
// tag (bounded like wait_ge).
__device__ inline cd ll_wait(u64* __restrict__ ll, int i, int tag, int* status) {
    u64* p = ll + 4 * (size_t)i;
    u64 w0, w1, w2, w3;
    int it = 0;
    for (;;) {
        w0 = __hip_atomic_load(p + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w1 = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w2 = __hip_atomic_load(p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w3 = __hip_atomic_load(p + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned t = (unsigned)tag;
        const bool ok = (unsigned)(w0 >> 32) == t && (unsigned)(w1 >> 32) == t &&
                        (unsigned)(w2 >> 32) == t && (unsigned)(w3 >> 32) == t;
        if (__all(ok)) break;
        if ((++it & 63) == 0 &&
            (it > FLAG_SPIN_LIMIT ||
             (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
              ST_SYNC_TIMEOUT))) {
            atomicOr(status, ST_SYNC_TIMEOUT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return mk(__longlong_as_double((long long)(((w1 & 0xffffffffull) << 32) | (w0 & 0xffffffffull))),
              __longlong_as_double((long long)(((w3 & 0xffffffffull) << 32) | (w2 & 0xffffffffull))));
}

// The first poll of entry i's hand-off words, issued ahead of other loads
// (a wave's loads complete in issue order, so a poll issued behind a factor
// tile would wait for the tile); ll_finish takes it if every lane's words
// carry the tag and polls on (ll_wait) otherwise.
__device__ inline void ll_issue(u64* __restrict__ ll, int i, u64 (&q)[4]) {
    u64* p = ll + 4 * (size_t)i;
#pragma unroll
    for (int j = 0; j < 4; ++j) q[j] = __hip_atomic_load(p + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline cd ll_finish(u64* __restrict__ ll, int i, int tag, int* status, const u64 (&q)[4]) {
    const unsigned t = (unsigned)tag;
    const bool ok = (unsigned)(q[0] >> 32) == t && (unsigned)(q[1] >> 32) == t &&
                    (unsigned)(q[2] >> 32) == t && (unsigned)(q[3] >> 32) == t;
    if (!__all(ok)) return ll_wait(ll, i, tag, status);
    return mk(__longlong_as_double((long long)(((q[1] & 0xffffffffull) << 32) | (q[0] & 0xffffffffull))),
              __longlong_as_double((long long)(((q[3] & 0xffffffffull) << 32) | (q[2] & 0xffffffffull))));
}

// sum_i f[i] * y_i over 16 columns (FP32 factor entries, FP64 vector), y_i
// held by lane i (i < 16) and broadcast with v_readlane.
__device__ inline cd dot16(const cf (&f)[16], cd yl) {
    cd s0 = mk(0, 0), s1 = mk(0, 0), s2 = mk(0, 0), s3 = mk(0, 0);
#pragma unroll
    for (int i = 0; i < 16; i += 4) {
        s0 = s0 + up(f[i]) * readlane_c(yl, i);
        s1 = s1 + up(f[i + 1]) * readlane_c(yl, i + 1);
        s2 = s2 + up(f[i + 2]) * readlane_c(yl, i + 2);
        s3 = s3 + up(f[i + 3]) * readlane_c(yl, i + 3);
    }
    return (s0 + s1) + (s2 + s3);
}

// Sum over the 8 waves' partials of this lane's row, in a fixed order.
__device__ inline cd sum8(const cd (*part)[NBM], int lane) {
    return ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane])) +
           ((part[4][lane] + part[5][lane]) + (part[6][lane] + part[7][lane]));
}

// Block solves over 128-row blocks, each block split between two workgroups
// of 64 rows (one lane per row; wave w takes the 16 columns [16 w, 16 w + 16)
// of every block it multiplies), so the product that waits on the previous
// block is 64 x 128 per workgroup.  Workgroup 2 g + h owns row half h of
// blocks g, g + G, ...  Per block: the block's inverse rows, its coupling
// rows and its right-hand side are loaded first (none depends on the
// chain); then the contributions of the finished blocks, each solution
// block read from its tagged hand-off words (one memory round trip, no
// flag, fence and second load), polled before the next factor tile is
// requested and with that tile loaded one block ahead; the halves of the
// accumulated right-hand side exchanged with the partner workgroup through
// hand-off words (acc_ll); the product with the block's 128 x 128 inverse,
// and the coupling to the previous block as its hand-off words arrive; one
// reduction, then this half's solution goes out as hand-off words and plain
// stores, and flags[2 blk + h] is raised (the upper solve waits on it before
// reading this block of y).  Blocks only wait on lower-numbered blocks and on
// the partner half, which publishes before it waits, and every workgroup is
// resident (cooperative launch), so the chain always drains.
__device__ __forceinline__ void trsv_lower_body(const cf* __restrict__ A, size_t lda, int n,
                                                int nblk2, const cf* __restrict__ Minv,
                                                const cf* __restrict__ Wc,
                                                const cd* __restrict__ b, cd* __restrict__ y,
                                                u64* __restrict__ yll, u64* __restrict__ acc_ll,
                                                int* __restrict__ flags, int epoch,
                                                int* __restrict__ status, cd (*part)[NBM],
                                                cd* accs) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int h = blockIdx.x & 1, G = gridDim.x >> 1;
    for (int blk = blockIdx.x >> 1; blk < nblk2; blk += G) {
        const int r0 = blk * NB2, nb = min(NB2, n - r0);
        const int row = NBM * h + lane;  // row within the block
        const bool rok = row < nb;
        const int rowc = r0 + min(row, nb - 1);
        cf mv[16], wv[16];
        const cf* Mb = Minv + (size_t)blk * NB2 * NB2 + (size_t)(16 * w) * NB2 + row;
#pragma unroll
        for (int i = 0; i < 16; ++i) mv[i] = Mb[(size_t)i * NB2];
        if (blk > 0) {
            const cf* Wb = Wc + (size_t)blk * NB2 * NB2 + (size_t)(16 * w) * NB2 + row;
#pragma unroll
            for (int i = 0; i < 16; ++i) wv[i] = Wb[(size_t)i * NB2];
        }
        const cd bv = (w == 0 && rok) ? b[r0 + row] : mk(0, 0);
        cd s = mk(0, 0);
        const int nk = blk - 1;  // blocks 0 .. blk - 2
        auto tile = [&](int k, cf (&l)[16]) {
            const cf* Lc = A + (size_t)(k * NB2 + 16 * w) * lda + rowc;
#pragma unroll
            for (int i = 0; i < 16; ++i) l[i] = Lc[(size_t)i * lda];
        };
        auto step = [&](int k, const cf (&cur)[16], cf (&nxt)[16]) {
            const int i = k * NB2 + 16 * w + (lane & 15);
            u64 q[4];
            ll_issue(yll, i, q);
            if (k + 1 < nk) tile(k + 1, nxt);
            s = s + dot16(cur, ll_finish(yll, i, epoch, status, q));
        };
        cf la[16], lb[16];
        if (nk > 0) tile(0, la);
        int k = 0;
        for (; k + 1 < nk; k += 2) {
            step(k, la, lb);
            step(k + 1, lb, la);
        }
        if (k < nk) step(k, la, lb);
        part[w][lane] = s;
        __syncthreads();
        if (w == 0) {
            const cd a = rok ? bv - sum8(part, lane) : mk(0, 0);
            accs[row] = a;
            ll_put(acc_ll, r0 + row, a, epoch);
        } else if (w == 1) {
            accs[NBM * (1 - h) + lane] =
                ll_wait(acc_ll, r0 + NBM * (1 - h) + lane, epoch, status);
        }
        __syncthreads();
        cd t = mk(0, 0);
#pragma unroll
        for (int i = 0; i < 16; ++i) t = t + up(mv[i]) * accs[16 * w + i];
        if (blk > 0)  // coupling to the previous block
            t = t - dot16(wv, ll_wait(yll, (blk - 1) * NB2 + 16 * w + (lane & 15), epoch, status));
        part[w][lane] = t;
        __syncthreads();
        if (w == 0 && rok) {
            const cd z = sum8(part, lane);
            ll_put(yll, r0 + row, z, epoch);  // the chain's hand-off first
            y[r0 + row] = z;
        }
        // the barrier completes these stores; the release store of the flag
        // then makes them visible at device scope
        __syncthreads();
        if (tid == 0) publish(&flags[2 * blk + h], epoch);
    }
}

// x = U^-1 y over 128-row blocks (in reverse order), same scheme; the
// solution blocks to the right come from their hand-off words (a partial
// last block, n not a multiple of 128, is taken first, outside the
// pipelined loop).  ylow: epoch of the lower solve that produced yv; this
// half's entries of yv are read after its lower flag reached ylow (x may
// alias the lower solve's right-hand side: those entries were read before
// that flag was set).  The hand-off words of block blk hold y (tag ylow)
// until this block writes x into them: every lower block that reads them
// has finished by then (each upper block starts after its own lower block,
// and block blk needs every upper block to its right).
__device__ __forceinline__ void trsv_upper_body(const cf* __restrict__ A, size_t lda, int n,
                                                int nblk2, const cf* __restrict__ Minv,
                                                const cf* __restrict__ Wc,
                                                const cd* __restrict__ yv, cd* __restrict__ x,
                                                u64* __restrict__ yll, u64* __restrict__ acc_ll,
                                                int* __restrict__ flags, int epoch, int ylow,
                                                int* __restrict__ status, cd (*part)[NBM],
                                                cd* accs) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int h = blockIdx.x & 1, G = gridDim.x >> 1;
    const int klast = nblk2 - 1;
    const int kblast = n - klast * NB2;  // columns of the last block
    for (int bi = blockIdx.x >> 1; bi < nblk2; bi += G) {
        const int blk = nblk2 - 1 - bi;
        const int r0 = blk * NB2, nb = min(NB2, n - r0);
        const int row = NBM * h + lane;
        const bool rok = row < nb;
        const int rowc = r0 + min(row, nb - 1);
        cf mv[16], wv[16];
        const cf* Mb = Minv + (size_t)blk * NB2 * NB2 + (size_t)(16 * w) * NB2 + row;
#pragma unroll
        for (int i = 0; i < 16; ++i) mv[i] = Mb[(size_t)i * NB2];
        if (blk + 1 < nblk2) {
            const cf* Wb = Wc + (size_t)blk * NB2 * NB2 + (size_t)(16 * w) * NB2 + row;
#pragma unroll
            for (int i = 0; i < 16; ++i) wv[i] = Wb[(size_t)i * NB2];
        }
        cd s = mk(0, 0);
        int khi = klast;  // first block of the pipelined loop (full blocks)
        if (kblast < NB2) {
            khi = klast - 1;
            const int c0 = klast * NB2 + 16 * w;
            const int cnt = min(16, kblast - 16 * w);  // columns of this wave
            if (klast > blk + 1 && cnt > 0) {
                const cd xk = ll_wait(yll, c0 + min(lane & 15, cnt - 1), epoch, status);
                const cf* Uc = A + (size_t)c0 * lda + rowc;
                for (int i = 0; i < cnt; ++i) s = s + up(Uc[(size_t)i * lda]) * readlane_c(xk, i);
            }
        }
        const int nk = khi - (blk + 1);  // full blocks khi .. blk + 2
        auto tile = [&](int j, cf (&u)[16]) {
            const cf* Uc = A + (size_t)((khi - j) * NB2 + 16 * w) * lda + rowc;
#pragma unroll
            for (int i = 0; i < 16; ++i) u[i] = Uc[(size_t)i * lda];
        };
        auto step = [&](int j, const cf (&cur)[16], cf (&nxt)[16]) {
            const int i = (khi - j) * NB2 + 16 * w + (lane & 15);
            u64 q[4];
            ll_issue(yll, i, q);
            if (j + 1 < nk) tile(j + 1, nxt);
            s = s + dot16(cur, ll_finish(yll, i, epoch, status, q));
        };
        cf ua[16], ub[16];
        if (nk > 0) tile(0, ua);
        int j = 0;
        for (; j + 1 < nk; j += 2) {
            step(j, ua, ub);
            step(j + 1, ub, ua);
        }
        if (j < nk) step(j, ua, ub);
        wave_wait_ge(&flags[2 * blk + h], ylow, status);
        part[w][lane] = s;
        __syncthreads();
        if (w == 0) {
            const cd a = rok ? yv[r0 + row] - sum8(part, lane) : mk(0, 0);
            accs[row] = a;
            ll_put(acc_ll, r0 + row, a, epoch);
        } else if (w == 1) {
            accs[NBM * (1 - h) + lane] =
                ll_wait(acc_ll, r0 + NBM * (1 - h) + lane, epoch, status);
        }
        __syncthreads();
        cd t = mk(0, 0);
#pragma unroll
        for (int i = 0; i < 16; ++i) t = t + up(mv[i]) * accs[16 * w + i];
        if (blk + 1 < nblk2) {
            // columns of a partial next block have zero coupling; their
            // index is clamped into x
            t = t - dot16(wv, ll_wait(yll, min((blk + 1) * NB2 + 16 * w + (lane & 15), n - 1),
                                      epoch, status));
        }
        part[w][lane] = t;
        __syncthreads();
        if (w == 0 && rok) {
            const cd z = sum8(part, lane);
            ll_put(yll, r0 + row, z, epoch);
            x[r0 + row] = z;
        }
        __syncthreads();
        if (tid == 0) publish(&flags[2 * blk + h], epoch);
    }
}

// z = U^-1 L^-1 P v: the lower block solve, then the upper one, in one
// cooperative launch (every workgroup resident; each waits only on blocks
// that other resident workgroups finish).  tmp receives y = L^-1 P v.
__global__ __launch_bounds__(512) void k_trsv_coop(const cf* __restrict__ A, size_t lda, int n,
                                                   int nblk2, const cf* __restrict__ Ml,
                                                   const cf* __restrict__ Wl,
                                                   const cf* __restrict__ Mu,
                                                   const cf* __restrict__ Wu,
                                                   const cd* __restrict__ v, cd* __restrict__ tmp,
                                                   cd* __restrict__ z, u64* __restrict__ yll,
                                                   int* __restrict__ flags, int e1, int e2,
                                                   int* __restrict__ status, size_t bs,
                                                   unsigned act) {
    __shared__ cd part[8][NBM];
    __shared__ cd accs[NB2];
    if (!((act >> blockIdx.z) & 1u)) return;  // a converged problem of the batch
    {
        const int q = blockIdx.z;
        A = zp(A, bs, q);
        Ml = zp(Ml, bs, q);
        Wl = zp(Wl, bs, q);
        Mu = zp(Mu, bs, q);
        Wu = zp(Wu, bs, q);
        v = zp(v, bs, q);
        tmp = zp(tmp, bs, q);
        z = zp(z, bs, q);
        yll = zp(yll, bs, q);
        flags = zp(flags, bs, q);
        status = zp(status, bs, q);
    }
    u64* acc_ll = yll + 4 * (size_t)nblk2 * NB2;  // the right-hand-side halves' hand-off
    trsv_lower_body(A, lda, n, nblk2, Ml, Wl, v, tmp, yll, acc_ll, flags, e1, status, part, accs);
    trsv_upper_body(A, lda, n, nblk2, Mu, Wu, tmp, z, yll, acc_ll, flags, e2, e1, status, part,
                    accs);
}

// ---------------------------------------------------------------------------
// Split-BF16x3 complex products on MFMA.  Each real product x*y is
// xh*yh + xh*yl + xl*yh (FP32 accumulation); complex = 4 real products,
// 12 MFMA per 16x16x32 complex block.  The operands come pre-split as bf16
// planes (written by the triangular kernels), so a lane loads its 8
// consecutive k of one plane with a single 16-byte global load, straight
// into the MFMA operand layout.  Workgroup tile 64 x 64, 4 waves of 32 x 32
// (2 x 2 MFMA 16x16 blocks).
// v_mfma_f32_16x16x32_bf16 layout: A[i][k]: lane i + 16 (k/8), element k%8;
// B[k][j]: lane j + 16 (k/8), element k%8; D[i][j]: lane j + 16 (i/4), reg i%4.
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

// The block column and row of step k once the diagonal block's inverse Xinv
// is known: L21 = A21 Xinv for the rows [c0, c0+nr) and U12 = A12 (the rows
// of the block as they stand) for the columns [c0, c0+nc), in one launch:
// the first ceil(nr/64) workgroups take 64-row blocks, the others 64-column
// blocks.  Row blocks: Xinv is split into bf16 hi/lo planes in LDS (the MFMA
// B operand); each wave takes 16 rows, its A operand (8 consecutive k of one
// row per lane) loaded straight from A and split in registers, and computes
// 16 x 64 with 2 x 4 complex 16x16x32 blocks of 12 MFMA (the split-BF16x3
// product of the trailing update, FP32 accumulation).  The 64 x 64 result is
// staged in LDS and stored to A and, split, to the planes la, both
// coalesced.  Column blocks only split A12 into the planes ub.
constexpr int TRR_ROWS = 64;
constexpr int PI_BK = NBM + 8;  // padded k stride of the LDS planes

// Split one complex float into the four bf16 planes of an MFMA operand
// (element e of lane's 8 consecutive k).
__device__ __forceinline__ void split_into(bf8 (&f)[4], int e, cf x) {
    __bf16 h, l;
    split_bf16(x.x, h, l);
    f[0][e] = h;
    f[1][e] = l;
    split_bf16(x.y, h, l);
    f[2][e] = h;
    f[3][e] = l;
}

// 16 rows x 64 columns, K = 64, of a complex product on MFMA: A operand fa
// (rows 16 w.., this lane's 8 consecutive k per 32-wide chunk kc, split),
// B operand from the LDS planes Bp (plane stride NBM PI_BK, element
// [col PI_BK + k]); 12 MFMA per complex 16x16x32 block (split-BF16x3).
// D layout: lane -> column 16 y + (lane & 15), rows 4 (lane >> 4) + (0..3).
__device__ __forceinline__ void cmma_16x64(const bf8 (&fa)[2][4], const __bf16* __restrict__ Bp,
                                           f4 (&accR)[4], f4 (&accI)[4]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        accR[y] = f4{0.f, 0.f, 0.f, 0.f};
        accI[y] = f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
        const bf8 naih = neg_bf8(fa[kc][2]), nail = neg_bf8(fa[kc][3]);
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            bf8 fb[4];
#pragma unroll
            for (int pl = 0; pl < 4; ++pl)
                fb[pl] = *(const bf8*)&Bp[pl * NBM * PI_BK + (16 * y + li) * PI_BK + 32 * kc + 8 * lq];
            f4 r = accR[y], i = accI[y];
            // + Ar Br - Ai Bi
            r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][0], fb[0], r, 0, 0, 0);
            r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][0], fb[1], r, 0, 0, 0);
            r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][1], fb[0], r, 0, 0, 0);
            r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih, fb[2], r, 0, 0, 0);
            r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih, fb[3], r, 0, 0, 0);
            r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nail, fb[2], r, 0, 0, 0);
            // + Ar Bi + Ai Br
            i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][0], fb[2], i, 0, 0, 0);
            i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][0], fb[3], i, 0, 0, 0);
            i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][1], fb[2], i, 0, 0, 0);
            i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][2], fb[0], i, 0, 0, 0);
            i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][2], fb[1], i, 0, 0, 0);
            i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[kc][3], fb[0], i, 0, 0, 0);
            accR[y] = r;
            accI[y] = i;
        }
    }
}

// Workgroup 0 of a step inside a group also applies the step to the next
// diagonal block (corner = its size, 0 otherwise): that block has only this
// step's update pending, its rows are the workgroup's rows, and U12 there is
// A12 as it stands, so corner -= L21 A12 is one more 64 x 64 x 64 product
// (the split-BF16x3 product of the trailing update, the same operands).
__global__ __launch_bounds__(256) void k_panel_inv(cf* __restrict__ A, size_t lda, int k0, int c0,
                                                   int nr, int nc, const cf* __restrict__ Xinv,
                                                   __bf16* __restrict__ la,
                                                   __bf16* __restrict__ ub, size_t pstride,
                                                   int corner, size_t bs) {
    constexpr int XS = NBM + 1;
    __shared__ __attribute__((aligned(16))) __bf16 Bp[4 * NBM * PI_BK];  // [pl][col BK + k]
    __shared__ cf X[NBM * XS];                                           // X[col XS + row]
    A = zp(A, bs, blockIdx.z);
    Xinv = zp(Xinv, bs, blockIdx.z);
    la = zp(la, bs, blockIdx.z);
    ub = zp(ub, bs, blockIdx.z);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nrb = (nr + TRR_ROWS - 1) / TRR_ROWS;
    if ((int)blockIdx.x >= nrb) {
        const int cb = c0 + ((int)blockIdx.x - nrb) * NBM;
#pragma unroll 4
        for (int t = tid; t < NBM * NBM; t += 256) {
            const int j = t >> 6, k = t & 63;
            if (cb + j < c0 + nc)
                store_planes(ub, pstride, (size_t)(cb + j) * NBM + k,
                             A[(size_t)(cb + j) * lda + k0 + k]);
        }
        return;
    }
    const int rb = c0 + (int)blockIdx.x * TRR_ROWS;
    const int li = lane & 15, lq = lane >> 4;
    const int row = rb + 16 * w + li;
    const bool rok = row < c0 + nr;
    const int cn = blockIdx.x == 0 ? corner : 0;  // uniform
    // A operand: k = 32 kc + 8 lq + e of this lane's row, loads in flight
    // while Xinv is split into LDS
    cf ax[2][8];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc)
#pragma unroll
        for (int e = 0; e < 8; ++e)
            ax[kc][e] = rok ? A[(size_t)(k0 + 32 * kc + 8 * lq + e) * lda + row] : mkf(0.f, 0.f);
    // the corner's A12 block (uq[q]: column t >> 6, k = t & 63 of t = tid +
    // 256 q) and its current values in the D layout, in flight as well
    cf uq[16], cc[4][4];
    if (cn > 0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int t = tid + 256 * q, j = t >> 6, k = t & 63;
            uq[q] = j < cn ? A[(size_t)(c0 + j) * lda + k0 + k] : mkf(0.f, 0.f);
        }
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = 16 * w + 4 * lq + e, c = 16 * y + li;
                cc[y][e] = (r < cn && c < cn) ? A[(size_t)(c0 + c) * lda + c0 + r] : mkf(0.f, 0.f);
            }
    }
#pragma unroll 4
    for (int t = tid; t < NBM * NBM; t += 256) {
        const int col = t >> 6, k = t & 63;  // Xinv[col 64 + k] = Xinv[k][col]
        const cf x = Xinv[t];
        __bf16 h, l;
        split_bf16(x.x, h, l);
        Bp[0 * NBM * PI_BK + col * PI_BK + k] = h;
        Bp[1 * NBM * PI_BK + col * PI_BK + k] = l;
        split_bf16(x.y, h, l);
        Bp[2 * NBM * PI_BK + col * PI_BK + k] = h;
        Bp[3 * NBM * PI_BK + col * PI_BK + k] = l;
    }
    bf8 fa[2][4];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc)
#pragma unroll
        for (int e = 0; e < 8; ++e) split_into(fa[kc], e, ax[kc][e]);
    __syncthreads();
    f4 accR[4], accI[4];
    cmma_16x64(fa, Bp, accR, accI);
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            X[(16 * y + li) * XS + 16 * w + 4 * lq + e] = mkf(accR[y][e], accI[y][e]);
    __syncthreads();
    // this block's L21 first: the stores drain while workgroup 0 applies
    // the step to the next diagonal block
    const bool ok = rb + lane < c0 + nr;
#pragma unroll 4
    for (int c = w; c < NBM; c += 4)
        if (ok) A[(size_t)(k0 + c) * lda + rb + lane] = X[c * XS + lane];
    // planes: t -> (row rr, k), k fastest, so the 2-byte stores are coalesced
#pragma unroll 4
    for (int t = tid; t < TRR_ROWS * NBM; t += 256) {
        const int rr = t >> 6, k = t & 63;
        if (rb + rr < c0 + nr) store_planes(la, pstride, (size_t)(rb + rr) * NBM + k, X[k * XS + rr]);
    }
    if (cn > 0) {
        // corner -= L21 A12: A12 into the planes (Bp is free), L21 rows of
        // this workgroup from X
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int t = tid + 256 * q, j = t >> 6, k = t & 63;
            __bf16 h, l;
            split_bf16(uq[q].x, h, l);
            Bp[0 * NBM * PI_BK + j * PI_BK + k] = h;
            Bp[1 * NBM * PI_BK + j * PI_BK + k] = l;
            split_bf16(uq[q].y, h, l);
            Bp[2 * NBM * PI_BK + j * PI_BK + k] = h;
            Bp[3 * NBM * PI_BK + j * PI_BK + k] = l;
        }
#pragma unroll
        for (int kc = 0; kc < 2; ++kc)
#pragma unroll
            for (int e = 0; e < 8; ++e)
                split_into(fa[kc], e, X[(32 * kc + 8 * lq + e) * XS + 16 * w + li]);
        __syncthreads();
        cmma_16x64(fa, Bp, accR, accI);
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = 16 * w + 4 * lq + e, c = 16 * y + li;
                if (r < cn && c < cn)
                    A[(size_t)(c0 + c) * lda + c0 + r] =
                        mkf(cc[y][e].x - accR[y][e], cc[y][e].y - accI[y][e]);
            }
    }
}

// ---------------------------------------------------------------------------
// Trailing update over K = 64 * nks: C -= sum over chunks q of L21_q U12_q,
// chunk q's planes at base + q * sstride (L21 planes first, U12 planes
// 4 * pstride after them).  Up to two output regions per launch.
// ---------------------------------------------------------------------------
struct UpdRegion {
    int m, nc;   // rows and columns of the region
    int r0, c0;  // first row / column (absolute: the plane index)
    cf* C;       // A(r0, c0)
};

// k_cgemm_update2: 128 x 128 workgroup tiles staged through LDS.
// 512 threads = 8 waves of 64 x 32 (4 x 2 MFMA blocks).  Per stage of 32 k
// the workgroup copies the tile's L21 and U12 planes (2 x 4 planes x 128 x
// 32 bf16 = 64 KB) into LDS once, so every operand byte is read from L2 once
// per workgroup instead of once per wave; the next stage's planes are loaded
// into registers while the MFMAs of the current one run.  In LDS a row of a
// plane is 32 k = 64 bytes; its four 16-byte parts are stored XOR-swizzled
// (swz) so the rows one ds_read_b128 reads together hit distinct banks.  The products accumulate from zero and C
// is loaded during the first stage and updated once at the end.  Tile order
// inside a region: panels of `panel` column tiles, row tiles outer and the
// panel's column tiles inner, so the workgroups resident on one XCD at a
// time share a few L21 and U12 tiles in that XCD's L2.
constexpr int BT = 128;            // workgroup tile
constexpr int KS = 32;             // k per stage
constexpr int PLN = BT * KS;       // bf16 per plane per operand per stage

// 16-byte part `part` of row r of a plane in LDS, XOR-swizzled by
// ((r >> 1) ^ (r >> 2)) & 3: the rows that one ds_read_b128 serves together
// (8 consecutive rows per 128 bytes, or 16 per 256 bytes) land in distinct
// bank groups under either width.
__device__ __forceinline__ int swz(int r, int part) {
    return r * KS + ((part ^ (((r >> 1) ^ (r >> 2)) & 3)) << 3);
}

template <int NBUF>
__device__ __forceinline__ void upd2_tile(const UpdRegion& R, int t, const __bf16* __restrict__ base,
                                          size_t pstride, size_t sstride, int nks, size_t ldc,
                                          __bf16* __restrict__ la, __bf16* __restrict__ lb,
                                          int panel, bool skip_corner = false) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int gx = (R.m + BT - 1) / BT, gy = (R.nc + BT - 1) / BT;
    const int p = t / (gx * panel);
    const int within = t - p * gx * panel;
    const int pw = min(panel, gy - p * panel);
    const int bm = (within / pw) * BT, bn = (p * panel + within % pw) * BT;
    // copy assignment: thread -> (row, 16-byte part) of every plane
    const int crow = tid >> 2, cpart = tid & 3;
    const __bf16* ga = base + (size_t)(R.r0 + min(bm + crow, R.m - 1)) * CT + cpart * 8;
    const __bf16* gb = base + 4 * pstride + (size_t)(R.c0 + min(bn + crow, R.nc - 1)) * CT + cpart * 8;
    const int coff = swz(crow, cpart);
    // compute assignment: wave tile rows 64 wr.., columns 32 wc..
    const int wr = w & 1, wc = w >> 1;
    const int li = lane & 15, lq = lane >> 4;
    // planes of the next stage, loaded into registers while a stage runs
    u4 ra[4], rb[4];
    auto gload = [&](int s) {
        const size_t off = (size_t)(s >> 1) * sstride + (s & 1) * KS;
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) {
            ra[pl] = *(const u4*)(ga + off + pl * pstride);
            rb[pl] = *(const u4*)(gb + off + pl * pstride);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) {
            *(u4*)(la + buf * 4 * PLN + pl * PLN + coff) = ra[pl];
            *(u4*)(lb + buf * 4 * PLN + pl * PLN + coff) = rb[pl];
        }
    };
    const int nst = 2 * nks;
    gload(0);
    // C of this wave's 64 x 32 block, in flight during the first stage
    cf* __restrict__ C = R.C;
    const bool fast = bm + BT <= R.m && bn + BT <= R.nc && (ldc & 3) == 0 &&
                      ((uintptr_t)C & 31) == 0;
    float4 c0[4][2], c1[4][2];
    auto cload = [&]() {
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const cf* cp = C + (size_t)(bn + 32 * wc + 16 * y + li) * ldc + bm + 64 * wr +
                               16 * x + lq * 4;
                c0[x][y] = *(const float4*)cp;
                c1[x][y] = *(const float4*)(cp + 2);
            }
    };
    if (fast) {
        cload();
    } else {
        // partial tile: the same registers, element by element with bounds,
        // all in flight together (loaded after the stores they would
        // otherwise wait behind, they serialise)
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int col = bn + 32 * wc + 16 * y + li;
                const int row0 = bm + 64 * wr + 16 * x + lq * 4;
                cf t[4];
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
                    t[qq] = (col < R.nc && row0 + qq < R.m) ? C[(size_t)col * ldc + row0 + qq]
                                                            : mkf(0.f, 0.f);
                c0[x][y] = float4{t[0].x, t[0].y, t[1].x, t[1].y};
                c1[x][y] = float4{t[2].x, t[2].y, t[3].x, t[3].y};
            }
    }
    f4 accR[4][2], accI[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            accR[x][y] = f4{0.f, 0.f, 0.f, 0.f};
            accI[x][y] = f4{0.f, 0.f, 0.f, 0.f};
        }
    auto compute = [&](int buf) {
        const __bf16* __restrict__ xa = la + buf * 4 * PLN;
        const __bf16* __restrict__ xb = lb + buf * 4 * PLN;
        bf8 fb[2][4];
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int pl = 0; pl < 4; ++pl)
                fb[y][pl] = *(const bf8*)(xb + pl * PLN + swz(32 * wc + 16 * y + li, lq));
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            bf8 fa[4];
#pragma unroll
            for (int pl = 0; pl < 4; ++pl)
                fa[pl] = *(const bf8*)(xa + pl * PLN + swz(64 * wr + 16 * x + li, lq));
            const bf8 naih = neg_bf8(fa[2]), nail = neg_bf8(fa[3]);
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                f4 r = accR[x][y], i = accI[x][y];
                // + Ar Br - Ai Bi
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[y][0], r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[y][1], r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[1], fb[y][0], r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih, fb[y][2], r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(naih, fb[y][3], r, 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nail, fb[y][2], r, 0, 0, 0);
                // + Ar Bi + Ai Br
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[y][2], i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[y][3], i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[1], fb[y][2], i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[2], fb[y][0], i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[2], fb[y][1], i, 0, 0, 0);
                i = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[3], fb[y][0], i, 0, 0, 0);
                accR[x][y] = r;
                accI[x][y] = i;
            }
        }
    };
    if (NBUF == 1) {
        // one LDS buffer: two barriers per stage
        for (int s = 0; s < nst; ++s) {
            __syncthreads();  // every wave is done with the previous stage
            lstore(0);
            __syncthreads();
            if (s + 1 < nst) gload(s + 1);
            compute(0);
        }
    } else {
        // two LDS buffers: stage s+1 is stored while other waves still
        // compute stage s; one barrier per stage
        __syncthreads();  // the previous tile is done with both buffers
        lstore(0);
        if (nst > 1) gload(1);
        __syncthreads();
        for (int s = 0; s < nst; ++s) {
            compute(s & 1);
            if (s + 1 < nst) {
                lstore((s + 1) & 1);
                if (s + 2 < nst) gload(s + 2);
            }
            __syncthreads();
        }
    }
    // C -= L21 U12 (the tile's top-left 64 x 64 corner, waves 0 and 2, is
    // left to another workgroup when skip_corner is set)
    if (skip_corner && wr == 0 && wc < 2) {
    } else if (fast) {
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                cf* cp = C + (size_t)(bn + 32 * wc + 16 * y + li) * ldc + bm + 64 * wr + 16 * x +
                         lq * 4;
                const f4 r = accR[x][y], i = accI[x][y];
                *(float4*)cp = float4{c0[x][y].x - r[0], c0[x][y].y - i[0], c0[x][y].z - r[1],
                                      c0[x][y].w - i[1]};
                *(float4*)(cp + 2) = float4{c1[x][y].x - r[2], c1[x][y].y - i[2],
                                            c1[x][y].z - r[3], c1[x][y].w - i[3]};
            }
    } else {
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int col = bn + 32 * wc + 16 * y + li;
                const int row0 = bm + 64 * wr + 16 * x + lq * 4;
                if (col >= R.nc) continue;
                const f4 r = accR[x][y], i = accI[x][y];
                const cf v[4] = {mkf(c0[x][y].x, c0[x][y].y), mkf(c0[x][y].z, c0[x][y].w),
                                 mkf(c1[x][y].x, c1[x][y].y), mkf(c1[x][y].z, c1[x][y].w)};
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
                    if (row0 + qq < R.m)
                        C[(size_t)col * ldc + row0 + qq] = cf{v[qq].x - r[qq], v[qq].y - i[qq]};
            }
    }
}

__device__ inline int upd2_tiles(const UpdRegion& R) {
    return R.m > 0 && R.nc > 0 ? ((R.m + BT - 1) / BT) * ((R.nc + BT - 1) / BT) : 0;
}

// Grid: any multiple of NXCD (workgroup id runs on XCD id % NXCD; each XCD
// takes a contiguous range of the tiles of both regions).
template <int NBUF>
__global__ __launch_bounds__(512) void k_cgemm_update2(UpdRegion r0, UpdRegion r1,
                                                       const __bf16* __restrict__ base,
                                                       size_t pstride, size_t sstride, int nks,
                                                       size_t ldc, int panel, size_t bs) {
    __shared__ __attribute__((aligned(16))) __bf16 lds_a[NBUF * 4 * PLN];
    __shared__ __attribute__((aligned(16))) __bf16 lds_b[NBUF * 4 * PLN];
    r0.C = zp(r0.C, bs, blockIdx.z);
    r1.C = zp(r1.C, bs, blockIdx.z);
    base = zp(base, bs, blockIdx.z);
    const int t0 = upd2_tiles(r0), tiles = t0 + upd2_tiles(r1);
    const int per = (tiles + NXCD - 1) / NXCD;
    for (int id = blockIdx.x; id < per * NXCD; id += gridDim.x) {
        const int t = (id % NXCD) * per + id / NXCD;
        if (t < t0)
            upd2_tile<NBUF>(r0, t, base, pstride, sstride, nks, ldc, lds_a, lds_b, panel);
        else if (t < tiles)
            upd2_tile<NBUF>(r1, t - t0, base, pstride, sstride, nks, ldc, lds_a, lds_b, panel);
    }
}

constexpr int XINV_WGS = 8;  // inverse-builder workgroups of k_cgemm_update_inverse
constexpr int GJ_PANELS = NBM / 8;  // panels of gj_factor_body<8>

// Update of step k fused with the elimination and inversion of diagonal
// block k+1, which is the top-left nb1 x nb1 corner of r0.  Workgroup 0
// applies the update to that corner itself (one partial tile; with
// do_corner = 0 k_panel_inv has done it) and eliminates it with
// gj_factor_body<8> (LDS of the tile reused), the record into rec1,
// raising gpan[q] once panel q of it is in memory; workgroups 1 .. ntile
// update r0 without that corner, and r1; the last XINV_WGS workgroups build
// the inverse Xinv1 from the record, one identity column per wave, panel by
// panel as the flags come up (the column is updated only from its pivot
// step on), so the inverse is done shortly after the elimination.  The waits
// cannot block workgroup 0: at most XINV_WGS workgroups wait, far fewer
// than the CUs.  The next block's elimination and inversion thus run beside
// the update instead of after it.  Over a batch of nz problems the grid is
// one-dimensional in role order -- the nz eliminations, then the nz
// XINV_WGS inverse builders, then each problem's ntile tile workgroups --
// since workgroups are dispatched in order: every problem's elimination
// starts at once instead of behind the tiles of the problems before it.
__global__ __launch_bounds__(512) void k_cgemm_update_inverse(UpdRegion r0, UpdRegion r1,
                                                              const __bf16* __restrict__ base,
                                                              size_t pstride, size_t sstride,
                                                              int nks, size_t ldc, int panel,
                                                              int ntile, int nb1,
                                                              int do_corner,
                                                              char* __restrict__ rec1,
                                                              int* __restrict__ gpan,
                                                              cf* __restrict__ Xinv1,
                                                              int* __restrict__ status,
                                                              size_t bs, int nz) {
    __shared__ __attribute__((aligned(16))) __bf16 lds_a[2 * 4 * PLN];
    __shared__ __attribute__((aligned(16))) __bf16 lds_b[2 * 4 * PLN];
    static_assert(sizeof(lds_a) >= sizeof(cf) * (NBM + 1) * NBM + sizeof(int) * (NBM + 1),
                  "elimination scratch fits the tile buffer");
    // role order: problem z and its workgroup number bid (0: elimination,
    // 1 .. ntile: tiles, ntile + 1 ..: inverse builders)
    int z, bid;
    {
        const int L = blockIdx.x;
        if (L < nz) {
            z = L;
            bid = 0;
        } else if (L < nz * (1 + XINV_WGS)) {
            z = (L - nz) / XINV_WGS;
            bid = ntile + 1 + (L - nz) % XINV_WGS;
        } else {
            const int q = L - nz * (1 + XINV_WGS);
            z = q / ntile;
            bid = 1 + q % ntile;
        }
        r0.C = zp(r0.C, bs, z);
        r1.C = zp(r1.C, bs, z);
        base = zp(base, bs, z);
        rec1 = zp(rec1, bs, z);
        gpan = zp(gpan, bs, z);
        Xinv1 = zp(Xinv1, bs, z);
        status = zp(status, bs, z);
    }
    if (bid == 0) {
        if (do_corner) {
            const UpdRegion corner{nb1, nb1, r0.r0, r0.c0, r0.C};
            upd2_tile<2>(corner, 0, base, pstride, sstride, nks, ldc, lds_a, lds_b, panel);
            __syncthreads();  // the corner's stores before the elimination reads it
        }
        cf* lmul = (cf*)lds_a;
        cf* s_inv = lmul + NBM * NBM;
        int* s_piv = (int*)(s_inv + NBM);
        gj_factor_body<8>(r0.C, ldc, nb1, rec1, status, lmul, s_inv, s_piv, s_piv + NBM, gpan);
        return;
    }
    if (bid > ntile) {
        // column m of the inverse from the record
        const int lane = threadIdx.x & 63;
        const int m = (bid - ntile - 1) * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
        const cf* __restrict__ g_lmul = rec_lmul(rec1);
        const int* __restrict__ g_piv = rec_piv(rec1);
        cf e = mkf(lane == m ? 1.f : 0.f, 0.f);
        int rank = lane;
        bool active = false;
        for (int pn = 0; pn < GJ_PANELS; ++pn) {
            wave_wait_ge(&gpan[pn], 1, status);
            cf lv[8];
            int pv[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                lv[i] = g_lmul[(8 * pn + i) * NBM + lane];
                pv[i] = __builtin_amdgcn_readfirstlane(g_piv[8 * pn + i]);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int p = pv[i];
                rank = lane == p ? 8 * pn + i : rank;
                active = active || p == m;
                if (active) e = cfms(e, lv[i], readlane_f(e, p));
            }
        }
        // row rank of A^-1: this lane's row of E times 1 / d_rank
        const cf sc = rec_inv(rec1)[rank];
        Xinv1[(size_t)m * NBM + rank] = e * sc;
        return;
    }
    const int t0 = upd2_tiles(r0), tiles = t0 + upd2_tiles(r1);
    const int per = (tiles + NXCD - 1) / NXCD;
    for (int id = bid - 1; id < per * NXCD; id += ntile) {
        const int t = (id % NXCD) * per + id / NXCD;
        if (t < t0)
            upd2_tile<2>(r0, t, base, pstride, sstride, nks, ldc, lds_a, lds_b, panel, t == 0);
        else if (t < tiles)
            upd2_tile<2>(r1, t - t0, base, pstride, sstride, nks, ldc, lds_a, lds_b, panel);
    }
}

static int upd2_tiles_host(const UpdRegion& R) {
    return R.m > 0 && R.nc > 0 ? ((R.m + BT - 1) / BT) * ((R.nc + BT - 1) / BT) : 0;
}


// Launch of k_cgemm_update2 for regions a and b (b may be empty) of each of
// the nz problems of a batch; nothing when both are empty.  cap: workgroups
// over the whole batch (0: one per tile).
static void launch_update(hipStream_t s, const UpdRegion& a, const UpdRegion& b,
                          const __bf16* base, size_t pstride, size_t sstride, int nks,
                          size_t ldc, int nz, size_t bs, int cap = 0) {
    const int tiles = upd2_tiles_host(a) + upd2_tiles_host(b);
    if (tiles == 0) return;
    int g = ((tiles + NXCD - 1) / NXCD) * NXCD;
    // a cap only for updates of up to 16 tiles per capped workgroup (of a
    // problem; a batch's problems share the cap's workgroups): a larger one
    // is the bottleneck itself and gets the whole device.  In a batch of 4
    // on the 6480-DOF patch the early bulk updates are over that bound;
    // capping them anyway (192, 128, 64 workgroups) was measured 3%, 8% and
    // 15% slower.
    const int capz = cap > 0 ? std::max(NXCD, ((cap / nz + NXCD - 1) / NXCD) * NXCD) : 0;
    if (capz > 0 && tiles <= 16 * capz) g = std::min(g, capz);
    // 4 column tiles per panel of the tile order (2, 8 and 16 measured the
    // same on the 25920-DOF solve); two LDS stage buffers (one buffer with
    // two barriers per stage measured the same or 1% slower)
    hipLaunchKernelGGL(k_cgemm_update2<2>, dim3(g, 1, nz), dim3(512), 0, s, a, b, base, pstride,
                       sstride, nks, ldc, 4, bs);
}

// k_cgemm_update_inverse for regions a (non-empty; block k+1 of nb1 columns
// at its top-left corner) and b, for each of the nz problems of a batch.
static void launch_update_inverse(hipStream_t s, const UpdRegion& a, const UpdRegion& b,
                                  const __bf16* base, size_t pstride, size_t sstride, int nks,
                                  size_t ldc, int nb1, bool do_corner, char* rec1, int* gpan,
                                  cf* Xinv1, int* status, int nz, size_t bs) {
    const int tiles = upd2_tiles_host(a) + upd2_tiles_host(b);
    const int g = ((tiles + NXCD - 1) / NXCD) * NXCD;
    hipLaunchKernelGGL(k_cgemm_update_inverse, dim3(nz * (1 + g + XINV_WGS)), dim3(512), 0, s, a,
                       b, base, pstride, sstride, nks, ldc, 4, g, nb1, (int)do_corner, rec1, gpan,
                       Xinv1, status, bs, nz);
}

// Workgroups of the bulk update on the side stream (BEMX_BULK_WGS; default
// 128, 0 = one per tile).  The bulk update then loops over its tiles on
// half the CUs and leaves the rest to the critical chain of the next group
// (panel kernels, eliminations, in-group updates), which otherwise queues
// behind thousands of bulk workgroups.  Updates of more than 16 tiles per
// capped workgroup (the early bulk updates of large systems, which bound
// the factorisation) are not capped.  Measured on the 6480-DOF solve:
// 16.9 ms with one workgroup per tile, 16.1 ms with 128 (96: 16.4, 192:
// 16.6, 64: 17.3), and 73.8 vs 70.1 solves/s through two request slots.
static int bulk_wg_cap() {
    static const int cap = [] {
        const char* v = getenv("BEMX_BULK_WGS");
        const int c = v ? atoi(v) : 128;
        return c > 0 ? ((c + NXCD - 1) / NXCD) * NXCD : 0;
    }();
    return cap;
}

// ---------------------------------------------------------------------------
// FP64 vector kernels of the refinement.
// ---------------------------------------------------------------------------
constexpr int MV_COLS = 128;  // columns per GEMV chunk

// Whether problem blockIdx.z takes part in a launch of the refinement (bit
// z of act; converged problems drop out and keep their solution).
__device__ __forceinline__ bool z_active(unsigned act) { return (act >> blockIdx.z) & 1u; }

// part[chunk][row] = sum over the chunk's columns of A[row, col] x[col]
__global__ __launch_bounds__(256) void k_zgemv_part(const cd* __restrict__ A, size_t lda, int n,
                                                    const cd* __restrict__ x,
                                                    cd* __restrict__ part, size_t sa, size_t bs,
                                                    unsigned act) {
    if (!z_active(act)) return;
    A = zp(A, sa, blockIdx.z);
    x = zp(x, bs, blockIdx.z);
    part = zp(part, bs, blockIdx.z);
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
                              const cd* __restrict__ b, cd* __restrict__ r, int sub, size_t bs,
                              unsigned act) {
    if (!z_active(act)) return;
    part = zp(part, bs, blockIdx.z);
    b = zp(b, bs, blockIdx.z);
    r = zp(r, bs, blockIdx.z);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cd s = mk(0, 0);
    for (int k = 0; k < nchunk; ++k) s = s + part[(size_t)k * n + i];
    r[i] = sub ? b[i] - s : s;
}

// partial[blk][i] = sum over the block's rows of conj(V_i) w   (i < nv), and
// partial[blk][nv] = sum |w|^2 (real part).
__global__ __launch_bounds__(256) void k_dots(int n, const cd* __restrict__ V, size_t ldv, int nv,
                                              const cd* __restrict__ w, cd* __restrict__ partial,
                                              size_t bs, unsigned act) {
    __shared__ double red[2][4];
    if (!z_active(act)) return;
    V = zp(V, bs, blockIdx.z);
    w = zp(w, bs, blockIdx.z);
    partial = zp(partial, bs, blockIdx.z);
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
                              double s, size_t bs, unsigned act) {
    if (!z_active(act)) return;
    V = zp(V, bs, blockIdx.z);
    h = zp(h, bs, blockIdx.z);
    w = zp(w, bs, blockIdx.z);
    out = zp(out, bs, blockIdx.z);
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    cd v = w[row];
    for (int i = 0; i < nv; ++i) v = v - h[i] * V[(size_t)i * ldv + row];
    out[row] = v * s;
}

// out[i] = sum over the nparts partial rows of entry i (i < nv1), in a fixed
// order: the reduction of k_dots on the device, same result on every run.
__global__ void k_dots_reduce(int nparts, int nv1, const cd* __restrict__ part,
                              cd* __restrict__ out, size_t bs, unsigned act) {
    if (!z_active(act)) return;
    part = zp(part, bs, blockIdx.z);
    out = zp(out, bs, blockIdx.z);
    const int i = threadIdx.x;
    if (i >= nv1) return;
    cd s = mk(0, 0);
    for (int bk = 0; bk < nparts; ++bk) s = s + part[(size_t)bk * nv1 + i];
    out[i] = s;
}

// out = w / sqrt(nrm2->x) (zero when the norm is zero)
__global__ void k_scale_by_norm(int n, const cd* __restrict__ w, cd* __restrict__ out,
                                const cd* __restrict__ nrm2, size_t bs, unsigned act) {
    if (!z_active(act)) return;
    w = zp(w, bs, blockIdx.z);
    out = zp(out, bs, blockIdx.z);
    nrm2 = zp(nrm2, bs, blockIdx.z);
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const double hn = sqrt(nrm2->x);
    out[row] = hn > 0.0 ? w[row] * (1.0 / hn) : mk(0, 0);
}

constexpr int GM_RESTART = 12;  // longest GMRES cycle of the refinement

// One GMRES cycle's state on the device: the norms the host reads back
// (head), the Hessenberg matrix (column-major, GM_RESTART + 1 rows), the
// Givens rotations, the rotated right-hand side g and the least-squares
// solution y.
struct GmHead {
    cd bn2, rn2, beta2;  // ||b||^2, ||r||^2, ||M^-1 r||^2 (real parts)
    double beta, gres;   // beta and |g[kk]|, the cycle's residual estimate
    int kk, stop;        // Arnoldi steps in the solution; converged / broken down
};
struct GmState {
    GmHead head;
    cd H[(GM_RESTART + 1) * GM_RESTART];
    cd cs[GM_RESTART], sn[GM_RESTART], g[GM_RESTART + 1], y[GM_RESTART];
};

__device__ inline cd conjz(cd a) { return mk(a.x, -a.y); }
__device__ inline double absz(cd a) { return hypot(a.x, a.y); }
// a / b by Smith's method (no overflow of |b|^2)
__device__ inline cd divz(cd a, cd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        const double t = b.y / b.x, d = b.x + b.y * t;
        return mk((a.x + a.y * t) / d, (a.y - a.x * t) / d);
    }
    const double t = b.x / b.y, d = b.x * t + b.y;
    return mk((a.x * t + a.y) / d, (a.y * t - a.x) / d);
}

// Start of a cycle: beta = ||M^-1 r||, g = beta e1.
__global__ void k_gm_start(GmState* gs, size_t bs, unsigned act) {
    if (threadIdx.x != 0 || !z_active(act)) return;
    gs = zp(gs, bs, blockIdx.z);
    GmHead& h = gs->head;
    h.beta = sqrt(h.beta2.x);
    for (int i = 0; i <= GM_RESTART; ++i) gs->g[i] = mk(0.0, 0.0);
    gs->g[0] = mk(h.beta, 0.0);
    h.gres = h.beta;
    h.kk = 0;
    h.stop = h.beta == 0.0;
}

// Arnoldi step k into the cycle: column k of H from the two Gram-Schmidt
// passes' projections (hp, hq) and the new vector's norm (hn2), the earlier
// rotations applied to it, a new rotation that zeroes H[k+1][k], and g
// rotated with it.  After convergence (|g[k+1]| <= tol beta) or a breakdown
// (zero norm) later steps leave the state alone.
__global__ void k_gm_step(GmState* gs, int k, const cd* __restrict__ hp,
                          const cd* __restrict__ hq, const cd* __restrict__ hn2, double tol,
                          size_t bs, unsigned act) {
    if (threadIdx.x != 0 || !z_active(act)) return;
    gs = zp(gs, bs, blockIdx.z);
    hp = zp(hp, bs, blockIdx.z);
    hq = zp(hq, bs, blockIdx.z);
    hn2 = zp(hn2, bs, blockIdx.z);
    GmHead& h = gs->head;
    if (h.stop) return;
    constexpr int LD = GM_RESTART + 1;
    cd* col = gs->H + (size_t)k * LD;
    for (int i = 0; i <= k; ++i) col[i] = hp[i] + hq[i];
    const double hn = sqrt(hn2[0].x);
    col[k + 1] = mk(hn, 0.0);
    cd *cs = gs->cs, *sn = gs->sn, *g = gs->g;
    for (int i = 0; i < k; ++i) {
        const cd t = conjz(cs[i]) * col[i] + conjz(sn[i]) * col[i + 1];
        col[i + 1] = -(sn[i] * col[i]) + cs[i] * col[i + 1];
        col[i] = t;
    }
    const double a = absz(col[k]), bb = absz(col[k + 1]);
    const double den = sqrt(a * a + bb * bb);
    if (den == 0.0) {
        cs[k] = mk(1.0, 0.0);
        sn[k] = mk(0.0, 0.0);
    } else {
        cs[k] = col[k] * (1.0 / den);
        sn[k] = col[k + 1] * (1.0 / den);
    }
    col[k] = conjz(cs[k]) * col[k] + conjz(sn[k]) * col[k + 1];
    col[k + 1] = mk(0.0, 0.0);
    g[k + 1] = -(sn[k] * g[k]);
    g[k] = conjz(cs[k]) * g[k];
    h.kk = k + 1;
    h.gres = absz(g[k + 1]);
    if (h.gres <= tol * h.beta || hn == 0.0) h.stop = 1;
}

// y = H^-1 g over the cycle's kk steps (back substitution).
__global__ void k_gm_solve(GmState* gs, size_t bs, unsigned act) {
    if (threadIdx.x != 0 || !z_active(act)) return;
    gs = zp(gs, bs, blockIdx.z);
    constexpr int LD = GM_RESTART + 1;
    const int kk = gs->head.kk;
    for (int i = kk - 1; i >= 0; --i) {
        cd t = gs->g[i];
        for (int j = i + 1; j < kk; ++j) t = t - gs->H[(size_t)j * LD + i] * gs->y[j];
        gs->y[i] = divz(t, gs->H[(size_t)i * LD + i]);
    }
}

// x += sum_{i < kk} y_i V_i
__global__ void k_combine_gm(int n, const cd* __restrict__ V, size_t ldv,
                             const GmState* __restrict__ gs, cd* __restrict__ x, size_t bs,
                             unsigned act) {
    if (!z_active(act)) return;
    V = zp(V, bs, blockIdx.z);
    gs = zp(gs, bs, blockIdx.z);
    x = zp(x, bs, blockIdx.z);
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int kk = gs->head.kk;
    cd v = x[row];
    for (int i = 0; i < kk; ++i) v = v + gs->y[i] * V[(size_t)i * ldv + row];
    x[row] = v;
}

// d = s for each active problem (s and d with strides of their own)
__global__ void k_copy_z(const cd* __restrict__ s, cd* __restrict__ d, int n, size_t ss,
                         size_t ds, unsigned act) {
    if (!z_active(act)) return;
    s = zp(s, ss, blockIdx.z);
    d = zp(d, ds, blockIdx.z);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

__global__ void k_fill_i(int* p, int v, int cnt, size_t bs) {
    p = zp(p, bs, blockIdx.z);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = v;
}

// What the host reads back per problem after a factorisation or a GMRES
// cycle: the cycle's head and the status word, gathered from the arenas into
// one contiguous array so that one copy serves the whole batch.
struct ReadRec {
    GmHead head;
    int status;
};

__global__ void k_gather_heads(const GmState* gs, const int* status, size_t bs,
                               ReadRec* __restrict__ out) {
    if (threadIdx.x != 0) return;
    const int z = blockIdx.z;
    if (gs) out[z].head = zp(gs, bs, z)->head;
    out[z].status = *zp(status, bs, z);
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

// Workspace of the mixed path: LuWork slot M_ARENA holds one arena per
// problem of a batch (every buffer of the factorisation and the refinement,
// at the offsets of arena_layout), slot M_RREC the host-read records.
// Slots 0..8 belong to the FP64 path.
enum { M_ARENA = 9, M_RREC = 10 };
enum { A_A32, A_REC, A_FLAGS, A_STATUS, A_VEC, A_V, A_PART, A_HOST, A_PLANES, A_INV, A_MINV,
       A_WC, A_GJF, A_LL, A_PARTS };

struct MixedFactor {
    cf* A = nullptr;
    int* flags = nullptr;
    int* status = nullptr;
    int* epoch = nullptr;   // host-side counter living in LuWork
    cf* inv = nullptr;      // per block: inverse of its diagonal block (column-major 64 x 64)
    char* rec = nullptr;    // two elimination records (steps k even / odd)
    int* gjflag = nullptr;  // per block and panel: that part of its record is in memory
    unsigned long long* yll = nullptr;  // tagged hand-off words of the block solves
    cf* minv = nullptr;     // per 128-row pair: lower inverse, then (after npair) upper
    cf* wc = nullptr;       // per pair: lower coupling Wl, then (after npair) upper Wu
    int n = 0, nblk = 0, npair = 0, coop = 1;
    size_t ld = 0;  // leading dimension of A
    // the batch: nz problems, problem z's arena at arena + z bs (the
    // pointers above are problem 0's)
    int nz = 1;
    size_t bs = 0;
    char* arena = nullptr;
    ReadRec* rrec = nullptr;  // nz host-read records
};

// Blocks per group of the factorisation (BEMX_GROUP, 1..8; default 8): the
// trailing update outside a group runs once per group with K = 64 * group
// (measured on MI355X: 4 and 8 are equal at 6480 DOF, 8 is 11% faster at
// 25920 DOF than 4).
static int group_blocks() {
    static const int g = [] {
        const char* v = getenv("BEMX_GROUP");
        const int x = v ? atoi(v) : 8;
        return std::min(8, std::max(1, x));
    }();
    return g;
}


// Offsets of the buffers in one problem's arena for order n (each 256-byte
// aligned) and the arena's size (a multiple of 64 KB).
struct ArenaLayout {
    size_t off[A_PARTS];
    size_t bytes;
};

static ArenaLayout arena_layout(int n) {
    const size_t ld = ((size_t)n + 31) & ~(size_t)31, nblk = (size_t)(n + NBM - 1) / NBM;
    const size_t npair = (nblk + 1) / 2, nchunk = (size_t)(n + MV_COLS - 1) / MV_COLS;
    const size_t nb256 = (size_t)(n + 255) / 256, MR = GM_RESTART;
    size_t sz[A_PARTS];
    sz[A_A32] = sizeof(cf) * ld * n;                    // FP32 copy of the matrix
    sz[A_INV] = sizeof(cf) * nblk * NBM * NBM;          // diagonal-block inverses
    sz[A_REC] = 2 * REC_BYTES;                          // two elimination records
    sz[A_GJF] = sizeof(int) * nblk * GJ_PANELS;         // record panel flags
    sz[A_MINV] = sizeof(cf) * npair * NB2 * NB2 * 2;    // 128-row inverses of L and U
    sz[A_WC] = sizeof(cf) * npair * NB2 * NB2 * 2;      // coupling blocks
    sz[A_FLAGS] = sizeof(int) * 2 * npair;              // block-solve flags
    sz[A_LL] = sizeof(u64) * 4 * 2 * npair * NB2;       // tagged hand-off words
    sz[A_STATUS] = sizeof(int);
    sz[A_PLANES] = sizeof(__bf16) * (size_t)n * NBM * 8 * 2 * group_blocks();
    sz[A_VEC] = sizeof(cd) * n * 5;                     // x, r, w, tmp, b
    sz[A_V] = sizeof(cd) * n * (MR + 1);                // GMRES basis
    sz[A_PART] = sizeof(cd) * std::max(nchunk * n, nb256 * (MR + 2));
    sz[A_HOST] = sizeof(cd) * (MR + 2) * 3 + sizeof(GmState);  // Arnoldi scratch, cycle state
    ArenaLayout L;
    size_t o = 0;
    for (int i = 0; i < A_PARTS; ++i) {
        L.off[i] = o;
        o += (sz[i] + 255) & ~(size_t)255;
    }
    L.bytes = (o + 65535) & ~(size_t)65535;
    return L;
}

// The host-read records are sized for the largest batch whatever the batch,
// so that a larger batch later never reallocates them (a reallocation drops
// every captured graph of the workspace).
static size_t rrec_bytes(int) {
    return std::max<size_t>(sizeof(ReadRec) * MIXED_MAX_BATCH, 256);
}

size_t mixed_workspace_bytes(int n, int nz) {
    return (size_t)nz * arena_layout(n).bytes + rrec_bytes(nz);
}

hipError_t preload_mixed_code() {
    hipFuncAttributes a;
    return hipFuncGetAttributes(&a, (const void*)k_trsv_coop);
}

// The FP32 copies of the batch's matrices (matrix z at A64 + z sa bytes),
// launched before the factorisation's graph, which then depends on the
// workspace only.
static void enqueue_convert(hipStream_t st, int n, const cd* A64, size_t lda, size_t sa,
                            const MixedFactor& F) {
    const size_t tot = (size_t)n * n;
    hipLaunchKernelGGL(k_cvt_z2c, dim3((tot + 255) / 256, 1, F.nz), dim3(256), 0, st, A64, lda,
                       F.A, F.ld, n, sa, F.bs);
}

// The launch sequence of the factorisation (no host synchronisation, no
// allocation), so it can be captured into a graph; every launch covers the
// F.nz problems of the batch, whose FP32 matrices are in place.
//
// Block LU in the form A = L U with unit block-lower L (L21 = A21 A11^-1 per
// step) and block-upper U whose diagonal blocks are the Schur-complement
// blocks A11 themselves, kept as their inverses Xinv (Gauss-Jordan with
// partial pivoting inside the block), and whose off-diagonal blocks are the
// rows A12 as they stand at their step: the Schur complements are those of
// the partial-pivot-in-block LU, with no triangular solves on the chain.
// Blocks of 64 columns are factored in groups of s = group_blocks() (W = 64 s
// columns).  Step k of group g (its columns end at gend):
//   st:   L21(k) = A21 Xinv(k) on all rows below and the planes of U12(k) on
//         all columns to the right (k_panel_inv, bf16 planes into the
//         group's slot of k); then the update of the group's own columns
//         (rows below k) and of the group's own rows (columns beyond gend),
//         K = 64, in one launch with the elimination of block k+1
//         (k_cgemm_update_inverse; block 0 by k_diag_gj alone).  (A
//         left-looking form, one K = 64 j update of block column and row k
//         per step, was measured 0.7 ms slower per 6480-DOF solve: its
//         64-wide regions half-fill the 128-wide tiles.)
// After the last step of group g, with K = W from the group's planes:
//   st:   [wait Z(g-1)] X(g): the next group's columns, all rows below gend;
//         Y(g): the next group's rows, the columns beyond it; with the
//         elimination of the next group's first block in the same launch;
//   side: Z(g): everything beyond the next group (the bulk),
// so group g+1 is factored while the bulk update of group g runs.  Planes
// are double-buffered by group parity: group g+1 writes the slots Z(g-1) read,
// after the wait for Z(g-1).
static hipError_t enqueue_factor(LuWork& w, hipStream_t st, int n, const MixedFactor& F,
                                 __bf16* planes, size_t pstride) {
    hipError_t first = hipSuccess;
    auto chk = [&](hipError_t e) {
        if (first == hipSuccess && e != hipSuccess) first = e;
    };
    const int nblk = F.nblk, nz = F.nz;
    const size_t bs = F.bs;
    const int sg = group_blocks();
    const size_t sstride = 8 * pstride;  // one slot: 4 L21 planes, 4 U12 planes
    hipLaunchKernelGGL(k_fill_i, dim3(1, 1, nz), dim3(64), 0, st, F.status, 0, 1, bs);
    // elimination-record flags (per block and panel), raised by
    // k_cgemm_update_inverse
    hipLaunchKernelGGL(k_fill_i, dim3((nblk * GJ_PANELS + 255) / 256, 1, nz), dim3(256), 0, st,
                       F.gjflag, 0, nblk * GJ_PANELS, bs);
    const size_t ld = F.ld;
    auto at = [&](int r, int c) { return F.A + (size_t)c * ld + r; };
    const UpdRegion none{0, 0, 0, 0, nullptr};
    const hipStream_t sd = w.side;
    bool bulk_pending = false;
    for (int k = 0; k < nblk; ++k) {
        const int g = k / sg, j = k - g * sg;
        const int k0 = k * NBM, nb = std::min(NBM, n - k0);
        const int c0 = k0 + nb, rest = n - c0;
        const int gend = std::min(n, (g + 1) * sg * NBM);
        cf* xinv = F.inv + (size_t)k * NBM * NBM;
        __bf16* gbase = planes + (size_t)(g & 1) * sg * sstride;
        __bf16* la = gbase + (size_t)j * sstride;
        __bf16* ub = la + 4 * pstride;
        char* rec1 = F.rec + (size_t)((k + 1) & 1) * REC_BYTES;
        // block 0 alone; block k+1 is eliminated and inverted inside step
        // k's update
        if (k == 0)
            hipLaunchKernelGGL(k_diag_gj<4>, dim3(1, 1, nz), dim3(1024), 0, st, (const cf*)F.A, ld,
                               nb, F.rec, xinv, 1, F.status, bs);
        if (rest <= 0) break;
        const int nrb = (rest + TRR_ROWS - 1) / TRR_ROWS;
        const int ncb = (rest + NBM - 1) / NBM;
        const int nb1 = std::min(NBM, rest);
        // inside a group the next diagonal block is updated by k_panel_inv
        const bool in_group = c0 < gend;
        hipLaunchKernelGGL(k_panel_inv, dim3(nrb + ncb, 1, nz), dim3(256), 0, st, F.A, ld, k0, c0,
                           rest, rest, (const cf*)xinv, la, ub, pstride, in_group ? nb1 : 0, bs);
        cf* xinv1 = xinv + NBM * NBM;
        int* gjf1 = F.gjflag + (size_t)(k + 1) * GJ_PANELS;
        if (in_group) {
            // inside the group: its columns right of k, and its rows below k
            const UpdRegion cs{rest, gend - c0, c0, c0, at(c0, c0)};
            const UpdRegion rs{gend - c0, n - gend, c0, gend, at(c0, gend)};
            launch_update_inverse(st, cs, n > gend ? rs : none, la, pstride, sstride, 1, ld, nb1,
                                  false, rec1, gjf1, xinv1, F.status, nz, bs);
            continue;
        }
        // last block of group g: apply the group to the rest, K = 64 * sg
        const int nend = std::min(n, gend + sg * NBM);
        if (bulk_pending) chk(hipStreamWaitEvent(st, w.ev_bulk, 0));
        bulk_pending = false;
        const UpdRegion X{n - gend, nend - gend, gend, gend, at(gend, gend)};
        const UpdRegion Y{nend - gend, n - nend, gend, nend, at(gend, nend)};
        launch_update_inverse(st, X, n > nend ? Y : none, gbase, pstride, sstride, sg, ld, nb1,
                              true, rec1, gjf1, xinv1, F.status, nz, bs);
        if (n > nend) {
            const UpdRegion Z{n - nend, n - nend, nend, nend, at(nend, nend)};
            chk(hipEventRecord(w.ev_panel, st));
            chk(hipStreamWaitEvent(sd, w.ev_panel, 0));
            launch_update(sd, Z, none, gbase, pstride, sstride, sg, ld, nz, bs, bulk_wg_cap());
            chk(hipEventRecord(w.ev_bulk, sd));
            bulk_pending = true;
        }
    }
    if (bulk_pending) chk(hipStreamWaitEvent(st, w.ev_bulk, 0));
    // 128-row diagonal-block inverses of both factors for the block solves
    hipLaunchKernelGGL(k_pair_inverses, dim3(2, F.npair, nz), dim3(256), 0, st, (const cf*)F.A,
                       ld, n, (const cf*)F.inv, nblk, F.minv,
                       F.minv + (size_t)F.npair * NB2 * NB2, bs);
    // coupling blocks of consecutive pairs for the block solves
    hipLaunchKernelGGL(k_pair_coupling, dim3(2, F.npair, 4 * nz), dim3(256), 0, st,
                       (const cf*)F.A, ld, n, F.npair, (const cf*)F.minv,
                       (const cf*)(F.minv + (size_t)F.npair * NB2 * NB2), F.wc,
                       F.wc + (size_t)F.npair * NB2 * NB2, bs);
    chk(hipGetLastError());
    return first;
}

// The factorisation of the nz problems of a batch.  fcode[z]: BEMX_OK or
// BEMX_ERR_SINGULAR (a zero pivot in one of problem z's diagonal blocks).
static int factor_mixed(LuWork& w, hipStream_t st, int n, const cd* A64, size_t lda, size_t sa,
                        int nz, MixedFactor& F, int* fcode, std::string& err,
                        const std::function<int()>* after) {
    hipError_t e;
    const int nblk = (n + NBM - 1) / NBM;
    F.n = n;
    F.nblk = nblk;
    F.ld = ((size_t)n + 31) & ~(size_t)31;  // 16-byte aligned 4-row groups
    F.nz = nz;
    // Every buffer of the mixed path, the refinement's too, before anything
    // is captured: a captured graph holds the addresses of the buffers it
    // touches, and any later reallocation would drop it.
    const ArenaLayout L = arena_layout(n);
    F.bs = L.bytes;
    char* ar = (char*)w.get(M_ARENA, (size_t)nz * L.bytes, &e);
    MX_CHK(e);
    F.rrec = (ReadRec*)w.get(M_RREC, rrec_bytes(nz), &e);
    MX_CHK(e);
    F.arena = ar;
    F.A = (cf*)(ar + L.off[A_A32]);
    F.inv = (cf*)(ar + L.off[A_INV]);
    F.rec = ar + L.off[A_REC];
    F.gjflag = (int*)(ar + L.off[A_GJF]);
    F.npair = (nblk + 1) / 2;
    F.minv = (cf*)(ar + L.off[A_MINV]);
    F.wc = (cf*)(ar + L.off[A_WC]);
    F.flags = (int*)(ar + L.off[A_FLAGS]);  // block-solve flags: one per 64-row half of a 128-row block
    F.yll = (u64*)(ar + L.off[A_LL]);
    F.status = (int*)(ar + L.off[A_STATUS]);
    F.epoch = &w.epoch;
    if (w.arena_gen != w.gen || w.arena_n != n) {
        w.arena_gen = w.gen;
        w.arena_n = n;
        w.arena_clear = 0;
    }
    if (w.arena_clear < nz) {
        // Flags and hand-off tags hold launch epochs, which only grow, so
        // an arena's are cleared once for its layout, before its first use,
        // and never inside a solve.
        const int z0 = w.arena_clear, cnt = nz - z0;
        const int nflags = 2 * F.npair, nll = F.npair * NB2;
        hipLaunchKernelGGL(k_fill_i, dim3((nflags + 255) / 256, 1, cnt), dim3(256), 0, st,
                           (int*)((char*)F.flags + (size_t)z0 * F.bs), 0, nflags, F.bs);
        hipLaunchKernelGGL(k_fill_i, dim3((16 * nll + 255) / 256, 1, cnt), dim3(256), 0, st,
                           (int*)((char*)F.yll + (size_t)z0 * F.bs), 0, 16 * nll, F.bs);
        w.arena_clear = nz;
    }
    int dev = 0, ncu = 0, per = 0;
    MX_CHK(hipGetDevice(&dev));
    MX_CHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    MX_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_trsv_coop, 512, 0));
    // two workgroups per 128-row block, at most one workgroup per CU over
    // the batch (two sharing a CU slow each other's products down more than
    // looping over blocks costs); every workgroup of the batch resident
    F.coop = 2 * std::max(1, std::min(F.npair, std::min(ncu, ncu * per) / (2 * nz)));
    MX_CHK(w.ensure_side());
    // bf16 planes of L21 and U12: one slot (8 planes) per block of a group,
    // double-buffered by group parity (group g+1 writes one set while the bulk
    // update of group g still reads the other)
    const size_t pstride = (size_t)n * NBM;
    __bf16* planes = (__bf16*)(ar + L.off[A_PLANES]);
    // The launch sequence after the FP32 copies depends only on n, the batch
    // size and the workspace buffers: it is captured once per size class and
    // batch size into a graph and replayed while the workspace keeps its
    // buffers (every reallocation bumps w.gen and drops the cached graphs).
    MX_CHK(hipEventRecord(w.ev_f0, st));
    enqueue_convert(st, n, A64, lda, sa, F);
    if (w.use_graphs) {
        auto key_of = [&](int z, const void** key) {
            key[0] = (const void*)(size_t)n;
            key[1] = (const void*)(size_t)w.gen;
            key[2] = (const void*)(size_t)z;
            for (int i = 3; i < LuWork::KEYS; ++i) key[i] = nullptr;
        };
        // captures the factorisation of batch size Fc.nz (nothing runs)
        auto capture = [&](const MixedFactor& Fc, hipGraphExec_t* out) -> hipError_t {
            hipGraph_t g = nullptr;
            hipError_t e0 = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e0 != hipSuccess) return e0;
            const hipError_t e1 = enqueue_factor(w, st, n, Fc, planes, pstride);
            const hipError_t e2 = hipStreamEndCapture(st, &g);
            if (e1 != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return e1 != hipSuccess ? e1 : e2;
            }
            const hipError_t e3 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            return e3;
        };
        const void* key[LuWork::KEYS];
        key_of(nz, key);
        hipGraphExec_t ex = w.find_graph(key);
        const bool miss = !ex;
        if (miss) {
            MX_CHK(capture(F, &ex));
            w.add_graph(key, ex);
            ++w.captures;
        } else {
            ++w.replays;
        }
        MX_CHK(hipGraphLaunch(ex, st));
        if (miss && w.ahead > 1) {
            // The first capture of this order also captures the other batch
            // sizes up to w.ahead that the arena holds, while this batch runs:
            // a later batch of any of them replays (a request loop then
            // captures nothing after its first request of each mesh size).
            // Every buffer enqueue_factor touches is problem 0's or an offset
            // of it by the arena stride, so one arena serves every batch size.
            const int room = std::min(std::min(w.ahead, MIXED_MAX_BATCH),
                                      (int)(w.cap[M_ARENA] / L.bytes));
            for (int z = 1; z <= room; ++z) {
                if (z == nz) continue;
                key_of(z, key);
                if (w.find_graph(key)) continue;
                MixedFactor Fz = F;
                Fz.nz = z;
                hipGraphExec_t ez = nullptr;
                MX_CHK(capture(Fz, &ez));
                w.add_graph(key, ez);
                ++w.precaptures;
            }
        }
    } else {
        MX_CHK(enqueue_factor(w, st, n, F, planes, pstride));
    }
    MX_CHK(hipEventRecord(w.ev_f1, st));
    w.factor_timed = true;
    hipLaunchKernelGGL(k_gather_heads, dim3(1, 1, nz), dim3(64), 0, st, (const GmState*)nullptr,
                       (const int*)F.status, F.bs, F.rrec);
    MX_CHK(hipGetLastError());
    if (after) {
        const int s = (*after)();
        if (s) return s;
    }
    ReadRec hr[MIXED_MAX_BATCH];
    MX_CHK(hipMemcpyAsync(hr, F.rrec, sizeof(ReadRec) * nz, hipMemcpyDeviceToHost, st));
    MX_CHK(hipStreamSynchronize(st));
    for (int z = 0; z < nz; ++z) {
        if (hr[z].status & ST_SYNC_TIMEOUT) {
            err = "mixed solve: a device-side wait of the factorisation timed out";
            return BEMX_ERR_TIMEOUT;
        }
        fcode[z] = (hr[z].status & ST_PIVOT_ZERO) ? BEMX_ERR_SINGULAR : BEMX_OK;
    }
    return BEMX_OK;
}

// z = M^-1 v = U^-1 L^-1 P v for the problems of the batch in act (tmp:
// work vector of n; z may alias v; every pointer is problem 0's)
static hipError_t precond(hipStream_t st, const MixedFactor& F, const cd* v, cd* z, cd* tmp,
                          unsigned act) {
    int n = F.n, nblk = F.npair;
    size_t lda = F.ld, bs = F.bs;
    const cf* A = F.A;
    const cf* li = F.minv;
    const cf* ui = F.minv + (size_t)F.npair * NB2 * NB2;
    const cf* wl = F.wc;
    const cf* wu = F.wc + (size_t)F.npair * NB2 * NB2;
    int* flags = F.flags;
    int* status = F.status;
    int e1 = ++*F.epoch;
    int e2 = ++*F.epoch;
    u64* yll = F.yll;
    void* a[] = {(void*)&A, &lda, &n, &nblk, (void*)&li, (void*)&wl, (void*)&ui, (void*)&wu,
                 (void*)&v, &tmp, &z, &yll, &flags, &e1, &e2, &status, &bs, &act};
    return hipLaunchCooperativeKernel((const void*)k_trsv_coop, dim3(F.coop, 1, F.nz), dim3(512),
                                      a, 0, st);
}

// r = b - A x (sub = 1) or r = A x (sub = 0) for the problems in act
// (matrix z at A + z sa bytes; the vectors in the arenas).
static void gemv(hipStream_t st, const MixedFactor& F, int n, const cd* A, size_t lda, size_t sa,
                 const cd* x, const cd* b, cd* r, cd* part, int sub, unsigned act) {
    const int nchunk = (n + MV_COLS - 1) / MV_COLS;
    hipLaunchKernelGGL(k_zgemv_part, dim3((n + 255) / 256, nchunk, F.nz), dim3(256), 0, st, A, lda,
                       n, x, part, sa, F.bs, act);
    hipLaunchKernelGGL(k_gemv_reduce, dim3((n + 255) / 256, 1, F.nz), dim3(256), 0, st, n, nchunk,
                       (const cd*)part, b, r, sub, F.bs, act);
}

// GMRES-IR of one right-hand side per problem against the FP32 factors F:
// problem z's right-hand side at b + z sb bytes, for the problems in act0.
// The outer iteration (residual b - A x in FP64, then a GMRES cycle on
// M^-1 A d = M^-1 r) runs on the device from start to end: the Arnoldi
// steps, the Givens rotations (k_gm_step), the least-squares solution and
// x += V y (k_gm_solve, k_combine_gm), then the next residual and its norm.
// The host reads the norms and the cycle's state of every problem once per
// outer iteration to decide which problems go on.  A problem's cycle takes
// `ch` Arnoldi steps (2 at first; doubled up to GM_RESTART while a cycle ends
// short of its inner tolerance) and stops early on the device once
// |g| <= 1e-8 beta or on breakdown; a launch of step k covers the problems
// whose cycle has more than k steps, so each problem follows its own
// sequence, that of a batch of one.  Per problem: code[z] (BEMX_OK, b
// overwritten with x; BEMX_ERR_NOCONV, b intact), iters[z], relres[z].
static int refine_batch(LuWork& w, hipStream_t st, const MixedFactor& F, int n, const cd* A64,
                        size_t lda, size_t sa, cd* b, size_t sb, unsigned act0, int* code,
                        int* iters, double* relres_out, std::string& err) {
    (void)w;
    const int MR = GM_RESTART;
    const int MAXOUT = 12;    // refinement steps
    const double TOL = 2e-15; // normwise backward error target
    const double TOL_IN = 1e-8;  // inner (GMRES cycle) reduction
    const int nb256 = (n + 255) / 256, nz = F.nz;
    const size_t bs = F.bs;
    const ArenaLayout L = arena_layout(n);
    cd* vec = (cd*)(F.arena + L.off[A_VEC]);
    cd *x = vec, *r = vec + n, *wv = vec + 2 * (size_t)n, *tmp = vec + 3 * (size_t)n,
       *bsave = vec + 4 * (size_t)n;
    cd* V = (cd*)(F.arena + L.off[A_V]);
    cd* part = (cd*)(F.arena + L.off[A_PART]);
    // device scratch of the Arnoldi step (projections of the two
    // Gram-Schmidt passes and the norm, slots of MR + 2 entries each), then
    // the cycle's state
    const int HS = MR + 2;
    cd* hdev = (cd*)(F.arena + L.off[A_HOST]);
    cd *h1 = hdev + HS, *h2 = hdev + 2 * HS;
    GmState* gs = (GmState*)(hdev + 3 * HS);
    const dim3 gv(nb256, 1, nz);
    // k_dots over nv vectors + reduction into out (nv + 1 entries, the last
    // one ||wvec||^2), all on the device
    auto dots_dev = [&](const cd* Vb, int nv, const cd* wvec, cd* out, unsigned m) {
        hipLaunchKernelGGL(k_dots, gv, dim3(256), 0, st, n, Vb, (size_t)n, nv, wvec, part, bs, m);
        hipLaunchKernelGGL(k_dots_reduce, dim3(1, 1, nz), dim3(64), 0, st, nb256, nv + 1,
                           (const cd*)part, out, bs, m);
    };
    auto pre = [&](const cd* v, cd* z, unsigned m) -> int {
        const hipError_t ep = precond(st, F, v, z, tmp, m);
        if (ep == hipErrorIllegalState) {
            err = coop_refused_message("mixed solve, block solves");
            return BEMX_ERR_HIP;
        }
        MX_CHK(ep);
        return BEMX_OK;
    };
    // every problem's norms, cycle state and status word in one round trip
    ReadRec hr[MIXED_MAX_BATCH];
    auto readback = [&]() -> int {
        hipLaunchKernelGGL(k_gather_heads, dim3(1, 1, nz), dim3(64), 0, st, (const GmState*)gs,
                           (const int*)F.status, bs, F.rrec);
        MX_CHK(hipMemcpyAsync(hr, F.rrec, sizeof(ReadRec) * nz, hipMemcpyDeviceToHost, st));
        MX_CHK(hipStreamSynchronize(st));
        for (int z = 0; z < nz; ++z)
            if (hr[z].status & ST_SYNC_TIMEOUT) {
                err = "mixed solve: a device-side wait of the triangular solves timed out";
                return BEMX_ERR_TIMEOUT;
            }
        return BEMX_OK;
    };
    hipLaunchKernelGGL(k_copy_z, gv, dim3(256), 0, st, (const cd*)b, bsave, n, sb, bs, act0);
    int s = pre(bsave, x, act0);  // x0 = M^-1 b
    if (s) return s;
    dots_dev(bsave, 0, bsave, &gs->head.bn2, act0);
    gemv(st, F, n, A64, lda, sa, x, bsave, r, part, 1, act0);  // r = b - A x (FP64)
    dots_dev(r, 0, r, &gs->head.rn2, act0);
    if ((s = readback())) return s;
    double bnorm[MIXED_MAX_BATCH], relres[MIXED_MAX_BATCH], prev[MIXED_MAX_BATCH];
    int ch[MIXED_MAX_BATCH], it_total[MIXED_MAX_BATCH];
    for (int z = 0; z < nz; ++z) {
        bnorm[z] = std::sqrt(hr[z].head.bn2.x);
        relres[z] = prev[z] = std::sqrt(hr[z].head.rn2.x) / std::max(bnorm[z], 1e-300);
        ch[z] = 2;
        it_total[z] = 0;
    }
    unsigned act = act0, conv = 0;
    for (int outer = 0;; ++outer) {
        // ||A x|| ~ ||b||, so ||r|| <= TOL ||b|| bounds the normwise backward
        // error ||r|| / (||A|| ||x|| + ||b||) by TOL; a residual that stops
        // shrinking at a level near the FP64 rounding of A x is accepted too.
        for (int z = 0; z < nz; ++z) {
            if (!((act >> z) & 1u)) continue;
            if (relres[z] <= TOL || (outer >= 2 && relres[z] > 0.5 * prev[z] && relres[z] <= 1e-13)) {
                conv |= 1u << z;
                act &= ~(1u << z);
            } else if (outer == MAXOUT) {
                act &= ~(1u << z);
            }
        }
        if (!act) break;
        int chmax = 0;
        for (int z = 0; z < nz; ++z)
            if ((act >> z) & 1u) chmax = std::max(chmax, ch[z]);
        // GMRES cycle on M^-1 A d = M^-1 r: v0 = M^-1 r / beta
        if ((s = pre(r, V, act))) return s;
        dots_dev(V, 0, V, &gs->head.beta2, act);
        hipLaunchKernelGGL(k_scale_by_norm, gv, dim3(256), 0, st, n, (const cd*)V, V,
                           (const cd*)&gs->head.beta2, bs, act);
        hipLaunchKernelGGL(k_gm_start, dim3(1, 1, nz), dim3(64), 0, st, gs, bs, act);
        for (int k = 0; k < chmax; ++k) {
            unsigned mk = 0;  // the problems whose cycle has a step k
            for (int z = 0; z < nz; ++z)
                if (((act >> z) & 1u) && ch[z] > k) mk |= 1u << z;
            gemv(st, F, n, A64, lda, sa, V + (size_t)k * n, nullptr, wv, part, 0, mk);
            if ((s = pre(wv, wv, mk))) return s;
            // classical Gram-Schmidt, two passes, and the normalisation
            dots_dev(V, k + 1, wv, hdev, mk);
            hipLaunchKernelGGL(k_orth_update, gv, dim3(256), 0, st, n, (const cd*)V, (size_t)n,
                               k + 1, (const cd*)hdev, wv, wv, 1.0, bs, mk);
            dots_dev(V, k + 1, wv, h1, mk);
            hipLaunchKernelGGL(k_orth_update, gv, dim3(256), 0, st, n, (const cd*)V, (size_t)n,
                               k + 1, (const cd*)h1, wv, wv, 1.0, bs, mk);
            dots_dev(V, 0, wv, h2, mk);
            hipLaunchKernelGGL(k_scale_by_norm, gv, dim3(256), 0, st, n, (const cd*)wv,
                               V + (size_t)(k + 1) * n, (const cd*)h2, bs, mk);
            hipLaunchKernelGGL(k_gm_step, dim3(1, 1, nz), dim3(64), 0, st, gs, k, (const cd*)hdev,
                               (const cd*)h1, (const cd*)h2, TOL_IN, bs, mk);
        }
        // y = H^-1 g, x += V y, then the next residual
        hipLaunchKernelGGL(k_gm_solve, dim3(1, 1, nz), dim3(64), 0, st, gs, bs, act);
        hipLaunchKernelGGL(k_combine_gm, gv, dim3(256), 0, st, n, (const cd*)V, (size_t)n,
                           (const GmState*)gs, x, bs, act);
        gemv(st, F, n, A64, lda, sa, x, bsave, r, part, 1, act);
        dots_dev(r, 0, r, &gs->head.rn2, act);
        if ((s = readback())) return s;
        for (int z = 0; z < nz; ++z) {
            if (!((act >> z) & 1u)) continue;
            const GmHead& hd = hr[z].head;
            it_total[z] += hd.kk;
            prev[z] = relres[z];
            relres[z] = std::sqrt(hd.rn2.x) / std::max(bnorm[z], 1e-300);
            // a cycle that used all its steps short of the inner tolerance: a
            // longer one next time
            if (!hd.stop && hd.kk == ch[z]) ch[z] = std::min(2 * ch[z], MR);
        }
    }
    MX_CHK(hipGetLastError());
    for (int z = 0; z < nz; ++z) {
        if (!((act0 >> z) & 1u)) continue;
        iters[z] = it_total[z];
        relres_out[z] = relres[z];
        code[z] = ((conv >> z) & 1u) ? BEMX_OK : BEMX_ERR_NOCONV;
    }
    if (conv)
        hipLaunchKernelGGL(k_copy_z, gv, dim3(256), 0, st, (const cd*)x, b, n, bs, sb, conv);
    MX_CHK(hipGetLastError());
    return BEMX_OK;
}

int gmres_ir_batch(LuWork& w, hipStream_t st, int n, int nz, const cd* A64, size_t lda, size_t sa,
                   cd* b, size_t sb, int nrhs, int* code, int* nsolved, double* stats,
                   std::string& err, const std::function<int()>* after) {
    if (nz < 1 || nz > MIXED_MAX_BATCH || n < 1 || nrhs < 1) {
        err = "mixed solve: batch of 1..16 problems, order and right-hand sides at least 1";
        return BEMX_ERR_ARG;
    }
    MixedFactor F;
    int fc[MIXED_MAX_BATCH];
    int s = factor_mixed(w, st, n, A64, lda, sa, nz, F, fc, err, after);
    if (s) return s;
    unsigned act = 0;
    for (int z = 0; z < nz; ++z) {
        nsolved[z] = 0;
        for (int i = 0; i < 4; ++i) stats[4 * z + i] = 0.0;
        code[z] = fc[z];
        if (fc[z] == BEMX_OK) act |= 1u << z;
    }
    for (int j = 0; j < nrhs && act; ++j) {
        int c[MIXED_MAX_BATCH] = {0}, it[MIXED_MAX_BATCH] = {0};
        double rr[MIXED_MAX_BATCH] = {0};
        s = refine_batch(w, st, F, n, A64, lda, sa, b + (size_t)j * n, sb, act, c, it, rr, err);
        if (s) return s;
        for (int z = 0; z < nz; ++z) {
            if (!((act >> z) & 1u)) continue;
            stats[4 * z] += it[z];
            stats[4 * z + 1] = std::max(stats[4 * z + 1], rr[z]);
            if (c[z] != BEMX_OK) {
                code[z] = c[z];
                act &= ~(1u << z);
            } else {
                nsolved[z] = j + 1;
            }
        }
    }
    return BEMX_OK;
}

int gmres_ir(LuWork& w, hipStream_t st, int n, const cd* A64, size_t lda, cd* b, int nrhs,
             double stats[4], std::string& err, int* nsolved) {
    int code = BEMX_OK;
    double st4[4] = {0, 0, 0, 0};
    *nsolved = 0;
    const int s = gmres_ir_batch(w, st, n, 1, A64, lda, 0, b, 0, nrhs, &code, nsolved, st4, err);
    if (stats) {
        stats[0] = st4[0];
        stats[1] = st4[1];
    }
    if (s) return s;
    if (code == BEMX_ERR_SINGULAR) err = "mixed solve: zero pivot in a diagonal block";
    if (code == BEMX_ERR_NOCONV) err = "mixed solve: GMRES-IR did not reach the FP64 residual level";
    return code;
}

}  // namespace bemx
