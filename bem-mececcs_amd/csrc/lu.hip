// lu.hip -- dense complex solve of ZH x = ZFI on MI355X (LINSOLVE).
//
// Plain FP64 path (ZGESV semantics, Linsolve.for:68-109): right-looking
// blocked LU with partial pivoting, panel width NB = 64, panels in groups of
// FP64_GROUP.
//   k_panel_fp64   the m x NB panel held in LDS across P co-resident
//                  workgroups; per column one inter-workgroup exchange of the
//                  local pivot candidates (largest |re|+|im|, lowest row on
//                  ties, as LAPACK's IZAMAX) through tagged words, then the
//                  row swap and the rank-1 update in LDS.  Also emits the
//                  panel's row permutation.  Orders whose panel needs more
//                  workgroups of 128 rows than the device holds at once
//                  (above 32768 on 256 CUs) use the tall form, whose
//                  workgroups own 256 or more rows: the first 128 in LDS,
//                  the rest updated in place in the matrix.
//   k_perm_trsm    applies that permutation to other columns (inside the
//                  group to its earlier L columns and its later columns,
//                  after the group to the columns right of it; L columns of
//                  earlier groups are left in place: the forward solve
//                  applies a group's permutations before its L blocks) and
//                  solves U12 = L11^-1 A12 for the columns right of the
//                  panel in the same launch (k_trsm_lower_cols alone for the
//                  group's later blocks), one lane per column.
//   k_zgemm_mfma, k_zgemm_mfma2  A22 -= L21 U12 as four real FP64 MFMA
//                  products (v_mfma_f64_16x16x4f64): Cr -= Ar Br - Ai Bi,
//                  Ci -= Ar Bi + Ai Br; K = 64 inside a group, K = 256 for
//                  the rest of the matrix after it.
// After a group, its update of the next group's columns runs on the
// solve's stream and the update of everything beyond them on a second
// stream (look-ahead), so the next group's panels overlap the bulk update.
// The solve walks the groups forward (permute, then per panel L11 solve and
// L21 update) and the panels backward (U11 solve, U update) with one kernel
// per step.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/bemx.h"
#include "bemx_lu.hpp"
#include "bemx_lu_kernels.hpp"

namespace bemx {

void* LuWork::get(int slot, size_t bytes, hipError_t* e) {
    *e = hipSuccess;
    if (bytes <= cap[slot] && buf[slot]) return buf[slot];
    // a captured graph may hold the old address of this buffer
    drop_graphs();
    ++gen;
    if (buf[slot]) hipFree(buf[slot]);
    buf[slot] = nullptr;
    cap[slot] = 0;
    *e = hipMalloc(&buf[slot], bytes ? bytes : 256);
    if (*e == hipSuccess) cap[slot] = bytes;
    return buf[slot];
}

size_t LuWork::bytes() const {
    size_t b = 0;
    for (int i = 0; i < SLOTS; ++i) b += buf[i] ? (cap[i] ? cap[i] : 256) : 0;
    return b;
}

hipGraphExec_t LuWork::find_graph(const void* const key[KEYS]) {
    for (GraphEntry& g : graphs) {
        if (!g.exec) continue;
        bool same = true;
        for (int i = 0; i < KEYS; ++i) same = same && g.key[i] == key[i];
        if (same) {
            g.used = ++tick;
            return g.exec;
        }
    }
    return nullptr;
}

void LuWork::add_graph(const void* const key[KEYS], hipGraphExec_t exec) {
    GraphEntry* slot = &graphs[0];
    for (GraphEntry& g : graphs) {
        if (!g.exec) {
            slot = &g;
            break;
        }
        if (g.used < slot->used) slot = &g;
    }
    if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
    for (int i = 0; i < KEYS; ++i) slot->key[i] = key[i];
    slot->exec = exec;
    slot->used = ++tick;
}

void LuWork::drop_graphs() {
    for (GraphEntry& g : graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = GraphEntry();
    }
}

void LuWork::release() {
    if (side) (void)hipStreamSynchronize(side);
    drop_graphs();
    for (int i = 0; i < SLOTS; ++i) {
        if (buf[i]) hipFree(buf[i]);
        buf[i] = nullptr;
        cap[i] = 0;
    }
    ++gen;
    if (side) hipStreamDestroy(side);
    if (ev_panel) hipEventDestroy(ev_panel);
    if (ev_bulk) hipEventDestroy(ev_bulk);
    if (ev_f0) hipEventDestroy(ev_f0);
    if (ev_f1) hipEventDestroy(ev_f1);
    side = nullptr;
    ev_panel = ev_bulk = ev_f0 = ev_f1 = nullptr;
    factor_timed = false;
}

LuWork::LuWork() {
    // BEMX_GRAPHS=0 replaces the graph replay of the mixed factorisation by
    // eager launches (for A/B timing)
    const char* g = getenv("BEMX_GRAPHS");
    use_graphs = !(g && g[0] == '0');
}

hipError_t LuWork::ensure_side() {
    hipError_t e = hipSuccess;
    if (!side) e = hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
    if (e == hipSuccess && !ev_panel) e = hipEventCreateWithFlags(&ev_panel, hipEventDisableTiming);
    if (e == hipSuccess && !ev_bulk) e = hipEventCreateWithFlags(&ev_bulk, hipEventDisableTiming);
    if (e == hipSuccess && !ev_f0) e = hipEventCreate(&ev_f0);
    if (e == hipSuccess && !ev_f1) e = hipEventCreate(&ev_f1);
    return e;
}

constexpr int NB = 64;          // panel width (FP64 path)
constexpr int PANEL_ROWS = 128;  // rows of the panel per workgroup (128 x 64 x 16 B = 128 KB LDS)
constexpr int SPIN_LIMIT = 1 << 22;

__device__ inline double cabs1(cd a) { return fabs(a.x) + fabs(a.y); }

// Wave maximum of a 64-bit unsigned key, uniform: DPP row shifts by 1, 2,
// 4, 8 within each row of 16 lanes, then row_bcast 15 into rows 1 and 3 and
// row_bcast 31 into rows 2 and 3, so lane 63 ends with the maximum; lanes
// whose DPP source lies outside the row take 0, the identity.  A few VALU
// instructions per step instead of an LDS permute per 32 bits.
__device__ inline unsigned long long wave_max_u64(unsigned long long k) {
    unsigned hi = (unsigned)(k >> 32), lo = (unsigned)k;
#define BEMX_MAX_STEP(ctrl, rmask)                                                       \
    {                                                                                    \
        const unsigned nh = (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi, ctrl, rmask, \
                                                                  0xf, false);           \
        const unsigned nl = (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo, ctrl, rmask, \
                                                                  0xf, false);           \
        const bool gt = nh > hi || (nh == hi && nl > lo);                                \
        hi = gt ? nh : hi;                                                               \
        lo = gt ? nl : lo;                                                               \
    }
    BEMX_MAX_STEP(0x111, 0xf)
    BEMX_MAX_STEP(0x112, 0xf)
    BEMX_MAX_STEP(0x114, 0xf)
    BEMX_MAX_STEP(0x118, 0xf)
    BEMX_MAX_STEP(0x142, 0xa)
    BEMX_MAX_STEP(0x143, 0xc)
#undef BEMX_MAX_STEP
    return ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)hi, 63) << 32) |
           (unsigned)__builtin_amdgcn_readlane((int)lo, 63);
}

// Wave minimum of a 32-bit unsigned value (the same DPP pattern, identity
// 0xffffffff), uniform.
__device__ inline unsigned wave_min_u32(unsigned v) {
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x111, 0xf, 0xf, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x112, 0xf, 0xf, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x114, 0xf, 0xf, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x118, 0xf, 0xf, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x142, 0xa, 0xf, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x143, 0xc, 0xf, false));
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Order-preserving key of a candidate |re|+|im| (>= 0; the bits of a
// non-negative double order like its value), 0 for no candidate (v < 0).
__device__ inline unsigned long long cand_key(double v) {
    return v < 0.0 ? 0ull : (unsigned long long)__double_as_longlong(v) + 1ull;
}
__device__ inline double key_val(unsigned long long k) {
    return k ? __longlong_as_double((long long)(k - 1ull)) : -1.0;
}

// a / u for the panel's multipliers with every operation written out (the
// fused multiply-adds explicit), so both forms of the panel round it alike
// whatever the compiler contracts around them.
__device__ inline cd panel_scale(cd a, cd u) {
    const double d = 1.0 / fma(u.x, u.x, u.y * u.y);
    const double ix = u.x * d, iy = -u.y * d;
    return cd{fma(a.x, ix, -(a.y * iy)), fma(a.x, iy, a.y * ix)};
}

// (val, row) ordering: larger val wins; equal val -> smaller row.
__device__ inline bool better(double v, int r, double bv, int br) {
    return v > bv || (v == bv && r < br);
}

// Tagged exchange words of the panel: every 64-bit word carries the panel
// launch's tag in its high half and 32 bits of data in its low half, and is
// written and read as one relaxed agent-scope atomic, so a reader that sees
// the tag in every word it needs has the data: no counter, fence or second
// read.  A column's candidate slot is three words (the two halves of
// |re|+|im|, the row); a row entry is four (the halves of re and im).  Tags
// are unique in the process (tag_block), so words left from an earlier
// panel or in reused memory never carry the tag being waited for.
typedef unsigned long long u64;
__device__ inline void tput(u64* p, unsigned tag, unsigned v) {
    __hip_atomic_store(p, ((u64)tag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void tput_d(u64* p, unsigned tag, double v) {
    const u64 b = (u64)__double_as_longlong(v);
    tput(p, tag, (unsigned)b);
    tput(p + 1, tag, (unsigned)(b >> 32));
}
__device__ inline u64 tget(const u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool tok(u64 w, unsigned tag) { return (unsigned)(w >> 32) == tag; }
__device__ inline double tval_d(u64 lo, u64 hi) {
    return __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
}

// Bounded poll of n consecutive tagged words at p (n <= 4) until all carry
// the tag; gives up (and flags ST_SYNC_TIMEOUT) after SPIN_LIMIT polls, or
// at once if an earlier wait already gave up, so every wave of the grid
// always runs to its end.  Lanes poll their own words; the loop ends when
// every lane of the wave has its words.
template <int N>
__device__ inline void tpoll(const u64* p, unsigned tag, u64 (&w)[N], bool active, int* status) {
    int it = 0;
    for (;;) {
        bool ok = true;
        if (active) {
#pragma unroll
            for (int i = 0; i < N; ++i) w[i] = tget(p + i);
#pragma unroll
            for (int i = 0; i < N; ++i) ok = ok && tok(w[i], tag);
        }
        if (__all(ok)) return;
        if ((++it & 63) == 0 &&
            (it > SPIN_LIMIT ||
             (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
              ST_SYNC_TIMEOUT))) {
            atomicOr(status, ST_SYNC_TIMEOUT);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// tpoll for two groups of consecutive words at once (their loads in flight
// together): NA words at pa into wa (if active_a), NB_ at pb into wb (if
// active_b).
template <int NA, int NB_>
__device__ inline void tpoll_pair(const u64* pa, const u64* pb, unsigned tag, u64 (&wa)[NA],
                                  u64 (&wb)[NB_], bool active_a, bool active_b, int* status) {
    int it = 0;
    for (;;) {
        bool ok = true;
        if (active_a) {
#pragma unroll
            for (int i = 0; i < NA; ++i) wa[i] = tget(pa + i);
        }
        if (active_b) {
#pragma unroll
            for (int i = 0; i < NB_; ++i) wb[i] = tget(pb + i);
        }
        if (active_a) {
#pragma unroll
            for (int i = 0; i < NA; ++i) ok = ok && tok(wa[i], tag);
        }
        if (active_b) {
#pragma unroll
            for (int i = 0; i < NB_; ++i) ok = ok && tok(wb[i], tag);
        }
        if (__all(ok)) return;
        if ((++it & 63) == 0 &&
            (it > SPIN_LIMIT ||
             (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
              ST_SYNC_TIMEOUT))) {
            atomicOr(status, ST_SYNC_TIMEOUT);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// ---------------------------------------------------------------------------
// Panel factorisation.  Workgroup p owns panel rows [r0, r1) (global row
// indices, r0 = k0 + p*rpw) held column-major in LDS (TALL: the first
// PANEL_ROWS of them; rows from PANEL_ROWS on stay in the matrix, where the
// same thread updates them, one thread per row).  Column j
// (diagonal row d = k0 + j) costs one inter-workgroup exchange through
// tagged words:
//   pick   wave 0 polls the P candidate slots of column j (best |re|+|im|
//          among each workgroup's owned rows >= d, lowest row on ties),
//          picks the winner, the same in every workgroup, and polls the
//          winner's row; wave 1 meanwhile polls row d, which its owner
//          publishes (two memory round trips);
//   apply  swap rows d and piv where owned, scale column j, update column
//          j+1 (only), find the local candidate for column j+1, complete
//          that row and row d+1 for the columns after j+1 and publish them;
//   then   the rest of the rank-1 update of columns j+2.. runs while the
//          other workgroups reach their own publish (look-ahead by one
//          column: the exchange of column j+1 overlaps this update).
// At the end workgroup 0 turns the pivot sequence into the list of rows
// whose content changed (dst) and where it came from (src).
// ---------------------------------------------------------------------------
template <bool TALL>
__global__ __launch_bounds__(256) void k_panel_fp64(cd* __restrict__ A, size_t lda, int n, int k0,
                                                    int nb, int rpw, int* __restrict__ ipiv,
                                                    u64* __restrict__ slots,
                                                    u64* __restrict__ cand, u64* __restrict__ drow,
                                                    int* __restrict__ perm,
                                                    long long* __restrict__ info,
                                                    int* __restrict__ status, unsigned tag) {
    __shared__ cd sA[NB * PANEL_ROWS];
    __shared__ cd urow[NB], dbuf[NB];
    __shared__ cd urow2[2][NB];  // the plain form's pivot rows, by step parity
    __shared__ int s_piv2[2];
    __shared__ double s_pval2[2];
    __shared__ double rv[4];
    __shared__ int rr[4];
    __shared__ int s_piv, s_cand;
    __shared__ int s_pivs[NB];  // pivot row of each column
    __shared__ double s_pval, s_cval;
    const int P = gridDim.x, p = blockIdx.x, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r0 = k0 + p * rpw;
    const int r1 = min(n, r0 + rpw);
    const int nr = r1 - r0;
    const int nl = min(nr, PANEL_ROWS);  // rows held in LDS
    // entry (column c, owned row r) of the panel
    auto el = [&](int c, int r) -> cd& {
        if (!TALL || r < PANEL_ROWS) return sA[c * PANEL_ROWS + r];
        return A[(size_t)(k0 + c) * lda + r0 + r];
    };
    for (int t = tid; t < nl * nb; t += 256) {
        int c = t / nl, r = t - c * nl;
        sA[c * PANEL_ROWS + r] = A[(size_t)(k0 + c) * lda + r0 + r];
    }
    __syncthreads();

    // local candidate of column jj among owned rows >= dmin -> s_cand, s_cval
    auto local_cand = [&](int jj, int dmin) {
        double bv = -1.0;
        int br = 0x7fffffff;
        if (TALL) {
            for (int r = tid; r < nr; r += 256)
                if (r0 + r >= dmin) {
                    const double v = cabs1(el(jj, r));
                    if (better(v, r0 + r, bv, br)) {
                        bv = v;
                        br = r0 + r;
                    }
                }
        } else if (tid < nr && r0 + tid >= dmin) {
            bv = cabs1(sA[jj * PANEL_ROWS + tid]);
            br = r0 + tid;
        }
        for (int off = 32; off >= 1; off >>= 1) {
            double ov = __shfl_xor(bv, off, 64);
            int orr = __shfl_xor(br, off, 64);
            if (better(ov, orr, bv, br)) {
                bv = ov;
                br = orr;
            }
        }
        if (lane == 0) {
            rv[wave] = bv;
            rr[wave] = br;
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < 4; ++w)
                if (better(rv[w], rr[w], bv, br)) {
                    bv = rv[w];
                    br = rr[w];
                }
            s_cand = br;
            s_cval = bv;
        }
        __syncthreads();
    };
    // publish the candidate row and row k0+jj (complete for column jj on) of
    // step jj, and the candidate slot; each word stands alone, so no barrier
    // or fence orders them
    auto publish = [&](int jj) {
        const int myrow = s_cand;
        if (myrow != 0x7fffffff && tid < nb) {
            const cd v = el(tid, myrow - r0);
            u64* q = cand + (((size_t)jj * P + p) * NB + tid) * 4;
            tput_d(q, tag, v.x);
            tput_d(q + 2, tag, v.y);
        }
        const int d = k0 + jj;
        if (d >= r0 && d < r1 && tid >= 64 && tid < 64 + nb) {
            const cd v = el(tid - 64, d - r0);
            u64* q = drow + ((size_t)jj * NB + tid - 64) * 4;
            tput_d(q, tag, v.x);
            tput_d(q + 2, tag, v.y);
        }
        if (tid == 0) {
            u64* q = slots + ((size_t)jj * P + p) * 3;
            tput_d(q, tag, s_cval);
            tput(q + 2, tag, (unsigned)myrow);
        }
    };

    local_cand(0, k0);
    publish(0);
    const int r = tid & (PANEL_ROWS - 1), half = tid >> 7;  // update mapping: row, column parity
    if constexpr (!TALL) {
        // Plain form (at most 128 rows, all in LDS), two barriers per column:
        //   poll   wave 0: the slots and row d together, then the winner's row
        //          unless it is row d, into the buffers of the step's parity
        //          (waves 1-3 may still run the previous step's rank-1 update,
        //          which reads the other one);
        //   [A]    swap: row d <- pivot row; row piv <- old row d except its
        //          columns j, j+1, which the thread of that row takes from dbuf;
        //          thread r (< 128) scales column j of its row and updates
        //          column j+1, and its row's |re|+|im| there is its candidate
        //          key; a wave maximum per wave;
        //   [E]    every thread combines the two waves' candidates; thread c of
        //          wave 0 (wave 1) completes entry c of the candidate row (row
        //          d+1) for columns j+2.. and publishes it with no further
        //          barrier; then waves 1-3 do the rest of the rank-1 update, in
        //          batches of eight independent LDS reads, ordered before the
        //          next column by the next step's [A].
        for (int j = 0; j < nb; ++j) {
            const int d = k0 + j, buf = j & 1;
            cd* __restrict__ U = urow2[buf];
            if (wave == 0) {
                double wv = -1.0;
                int wr = 0x7fffffff;
                u64 wd[4];  // row d, published a column earlier, polled with the first slots
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = lane + 64 * u;
                    if (64 * u >= P) break;  // wave-uniform
                    u64 w3[3];
                    if (u == 0)
                        tpoll_pair<3, 4>(slots + ((size_t)j * P + min(q, P - 1)) * 3,
                                         drow + ((size_t)j * NB + min(lane, nb - 1)) * 4, tag,
                                         w3, wd, q < P, lane < nb, status);
                    else
                        tpoll<3>(slots + ((size_t)j * P + min(q, P - 1)) * 3, tag, w3, q < P,
                                 status);
                    const double cv = tval_d(w3[0], w3[1]);
                    const int cr = (int)(unsigned)w3[2];
                    if (q < P && better(cv, cr, wv, wr)) {
                        wv = cv;
                        wr = cr;
                    }
                }
                // the largest value, then the lowest row holding it; workgroup
                // q owns the rows from k0 + q rpw on
                const unsigned long long kmax = wave_max_u64(cand_key(wv));
                wr = (int)wave_min_u32(cand_key(wv) == kmax ? (unsigned)wr : 0xffffffffu);
                wv = key_val(kmax);
                const int wp = (wr - k0) / rpw;
                // the winner's row: row d itself when the diagonal wins (no
                // second round trip; its owner published it both as its
                // candidate and as row d, from the same values), else polled
                const cd rd = mk(tval_d(wd[0], wd[1]), tval_d(wd[2], wd[3]));
                cd ru = rd;
                if (wr != d) {
                    u64 w4[4];
                    tpoll<4>(cand + (((size_t)j * P + wp) * NB + min(lane, nb - 1)) * 4, tag, w4,
                             lane < nb, status);
                    ru = mk(tval_d(w4[0], w4[1]), tval_d(w4[2], w4[3]));
                }
                if (lane < nb) {
                    U[lane] = ru;
                    dbuf[lane] = rd;
                }
                if (lane == 0) {
                    s_piv2[buf] = wr;
                    s_pval2[buf] = wv;
                }
            }
            __syncthreads();  // [A]
            const int piv = s_piv2[buf];
            const double pval = s_pval2[buf];
            if (tid == 0) s_pivs[j] = piv;
            if (p == 0 && tid == 0) {
                ipiv[d] = piv;
                if (pval == 0.0 && *info == 0) *info = d + 1;
            }
            if (piv != d && tid < nb) {
                if (d >= r0 && d < r1) sA[tid * PANEL_ROWS + d - r0] = U[tid];
                if (piv >= r0 && piv < r1 && tid != j && tid != j + 1)
                    sA[tid * PANEL_ROWS + piv - r0] = dbuf[tid];
            }
            const bool more = j + 1 < nb;
            double bv = -1.0;
            int br = 0x7fffffff;
            if (tid < nr && r0 + tid > d) {
                const bool moved = r0 + tid == piv && piv != d;
                cd aj = moved ? dbuf[j] : sA[j * PANEL_ROWS + tid];
                cd aj1 = more ? (moved ? dbuf[j + 1] : sA[(j + 1) * PANEL_ROWS + tid])
                              : mk(0.0, 0.0);
                if (pval != 0.0) {
                    aj = panel_scale(aj, U[j]);
                    if (more) aj1 = cfms(aj1, aj, U[j + 1]);
                }
                sA[j * PANEL_ROWS + tid] = aj;
                if (more) {
                    sA[(j + 1) * PANEL_ROWS + tid] = aj1;
                    bv = cabs1(aj1);
                    br = r0 + tid;
                }
            }
            if (wave < 2) {
                // lane order is row order: the lowest lane holding the largest
                // value is the lowest row
                const unsigned long long k = cand_key(bv), kmax = wave_max_u64(k);
                const unsigned long long tie = __ballot(k == kmax);
                if (lane == 0) {
                    rv[wave] = key_val(kmax);
                    rr[wave] = kmax ? r0 + 64 * wave + __ffsll((long long)tie) - 1 : 0x7fffffff;
                }
            }
            __syncthreads();  // [E]
            int sp1 = -1, sp2 = -1;  // local rows already completed for columns j+2..
            if (more) {
                double cv = rv[0];
                int cr = rr[0];
                if (better(rv[1], rr[1], cv, cr)) {
                    cv = rv[1];
                    cr = rr[1];
                }
                if (cr != 0x7fffffff) sp1 = cr - r0;
                const bool own_d1 = d + 1 >= r0 && d + 1 < r1;
                if (own_d1 && d + 1 != cr) sp2 = d + 1 - r0;
                // wave 0: the candidate row (and row d+1 when it is the
                // candidate); wave 1: row d+1 otherwise; thread c of the wave
                // owns entry c
                const int sr = wave == 0 ? sp1 : (wave == 1 ? sp2 : -1);
                const int c = lane;
                if (sr >= 0 && c < nb) {
                    cd v = sA[c * PANEL_ROWS + sr];
                    if (c >= j + 2 && pval != 0.0) {
                        v = cfms(v, sA[j * PANEL_ROWS + sr], U[c]);
                        sA[c * PANEL_ROWS + sr] = v;
                    }
                    if (wave == 0) {
                        u64* q = cand + (((size_t)(j + 1) * P + p) * NB + c) * 4;
                        tput_d(q, tag, v.x);
                        tput_d(q + 2, tag, v.y);
                    }
                    if (wave == 1 || (own_d1 && d + 1 == cr)) {
                        u64* q = drow + ((size_t)(j + 1) * NB + c) * 4;
                        tput_d(q, tag, v.x);
                        tput_d(q + 2, tag, v.y);
                    }
                }
                if (tid == 0) {
                    u64* q = slots + ((size_t)(j + 1) * P + p) * 3;
                    tput_d(q, tag, cv);
                    tput(q + 2, tag, (unsigned)cr);
                }
            }
            // the rest of the rank-1 update (columns j+2..) by waves 1-3, while
            // wave 0 goes on to the next column's polls: item (row, third) of
            // the 128 x 3 items, the third q taking the columns j+2+q, step 3
            if (wave > 0 && pval != 0.0) {
#pragma unroll
                for (int it = 0; it < 2; ++it) {
                    const int item = tid - 64 + 192 * it, rw = item & (PANEL_ROWS - 1);
                    if (rw < nr && r0 + rw > d && rw != sp1 && rw != sp2) {
                        const cd l = sA[j * PANEL_ROWS + rw];
                        int c = j + 2 + (item >> 7);
                        for (; c + 21 < nb; c += 24) {
                            cd v[8], u[8];
#pragma unroll
                            for (int k = 0; k < 8; ++k) {
                                v[k] = sA[(c + 3 * k) * PANEL_ROWS + rw];
                                u[k] = U[c + 3 * k];
                            }
#pragma unroll
                            for (int k = 0; k < 8; ++k)
                                sA[(c + 3 * k) * PANEL_ROWS + rw] = cfms(v[k], l, u[k]);
                        }
                        for (; c < nb; c += 3)
                            sA[c * PANEL_ROWS + rw] = cfms(sA[c * PANEL_ROWS + rw], l, U[c]);
                    }
                }
            }
        }
        __syncthreads();
    } else {
        for (int j = 0; j < nb; ++j) {
            const int d = k0 + j;
            if (wave == 0) {
                // the P slots (lane q: slots q, q + 64, ...; P <= 4 * 64), then
                // the winner's row
                double wv = -1.0;
                int wr = 0x7fffffff, wp = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = lane + 64 * u;
                    if (64 * u >= P) break;  // wave-uniform
                    u64 w3[3];
                    tpoll<3>(slots + ((size_t)j * P + min(q, P - 1)) * 3, tag, w3, q < P, status);
                    const double cv = tval_d(w3[0], w3[1]);
                    const int cr = (int)(unsigned)w3[2];
                    if (q < P && better(cv, cr, wv, wr)) {
                        wv = cv;
                        wr = cr;
                        wp = q;
                    }
                }
                for (int off = 32; off >= 1; off >>= 1) {
                    double ov = __shfl_xor(wv, off, 64);
                    int orr = __shfl_xor(wr, off, 64);
                    int op = __shfl_xor(wp, off, 64);
                    if (better(ov, orr, wv, wr)) {
                        wv = ov;
                        wr = orr;
                        wp = op;
                    }
                }
                u64 w4[4];
                tpoll<4>(cand + (((size_t)j * P + wp) * NB + min(lane, nb - 1)) * 4, tag, w4,
                         lane < nb, status);
                if (lane < nb) urow[lane] = mk(tval_d(w4[0], w4[1]), tval_d(w4[2], w4[3]));
                if (lane == 0) {
                    s_piv = wr;
                    s_pval = wv;
                }
            } else if (wave == 1) {
                u64 w4[4];
                tpoll<4>(drow + ((size_t)j * NB + min(lane, nb - 1)) * 4, tag, w4, lane < nb,
                         status);
                if (lane < nb) dbuf[lane] = mk(tval_d(w4[0], w4[1]), tval_d(w4[2], w4[3]));
            }
            __syncthreads();
            const int piv = s_piv;
            const double pval = s_pval;
            if (tid == 0) s_pivs[j] = piv;
            if (p == 0 && tid == 0) {
                ipiv[d] = piv;
                if (pval == 0.0 && *info == 0) *info = d + 1;
            }
            __syncthreads();
            // apply: swap, scale column j, update column j+1
            if (piv != d) {
                if (d >= r0 && d < r1 && tid < nb) el(tid, d - r0) = urow[tid];
                if (piv >= r0 && piv < r1 && tid < nb) el(tid, piv - r0) = dbuf[tid];
            }
            __syncthreads();

            for (int ri = tid; ri < (TALL ? nr : min(nr, 256)); ri += 256) {
                if (r0 + ri > d && pval != 0.0) {
                    const cd l = panel_scale(el(j, ri), urow[j]);
                    el(j, ri) = l;
                    if (j + 1 < nb) el(j + 1, ri) = cfms(el(j + 1, ri), l, urow[j + 1]);
                }
            }
            __syncthreads();
            int sp1 = -1, sp2 = -1;  // local rows already completed for columns j+2..
            if (j + 1 < nb) {
                local_cand(j + 1, d + 1);
                const int cr = s_cand;
                if (cr != 0x7fffffff) sp1 = cr - r0;
                if (d + 1 >= r0 && d + 1 < r1 && d + 1 != cr) sp2 = d + 1 - r0;
                // complete the candidate row (threads 0..63) and row d+1 (64..127)
                const int sr = tid < 64 ? sp1 : (tid < 128 ? sp2 : -1);
                const int c = j + 2 + lane;
                if (sr >= 0 && c < nb && pval != 0.0)
                    el(c, sr) = cfms(el(c, sr), el(j, sr), urow[c]);
                __syncthreads();
                publish(j + 1);
            }
            // the rest of the rank-1 update (columns j+2..), overlapping the exchange
            if (TALL) {
                // one thread per row; a row's entries in the matrix are loaded
                // eight at a time so their latencies overlap
                for (int rr = tid; rr < nr; rr += 256) {
                    if (r0 + rr > d && rr != sp1 && rr != sp2 && pval != 0.0) {
                        const cd l = el(j, rr);
                        int c = j + 2;
                        if (rr >= PANEL_ROWS) {
                            cd* col = &A[(size_t)k0 * lda + r0 + rr];
                            for (; c + 8 <= nb; c += 8) {
                                cd v[8];
#pragma unroll
                                for (int u = 0; u < 8; ++u) v[u] = col[(size_t)(c + u) * lda];
#pragma unroll
                                for (int u = 0; u < 8; ++u)
                                    col[(size_t)(c + u) * lda] = cfms(v[u], l, urow[c + u]);
                            }
                        }
                        for (; c < nb; ++c) el(c, rr) = cfms(el(c, rr), l, urow[c]);
                    }
                }
            } else if (r < nr && r0 + r > d && r != sp1 && r != sp2 && pval != 0.0) {
                const cd l = sA[j * PANEL_ROWS + r];
                for (int c = j + 2 + half; c < nb; c += 2)
                    sA[c * PANEL_ROWS + r] = cfms(sA[c * PANEL_ROWS + r], l, urow[c]);
            }
            __syncthreads();
        }
    }
    for (int t = tid; t < nl * nb; t += 256) {
        int c = t / nl, r2 = t - c * nl;
        A[(size_t)(k0 + c) * lda + r0 + r2] = sA[c * PANEL_ROWS + r2];
    }
    if (p == 0 && wave == 0) {
        // Replay the swaps on the rows they touch (at most 2 NB): slot i of
        // the list lives in lane i % 64, register i / 64; key = position,
        // content = original row now there.  Lookups are wave ballots, so
        // the whole replay stays in registers.
        int key[2] = {-1, -1}, content[2] = {-1, -1};
        int cnt = 0;
        auto find = [&](int row) -> int {
            const unsigned long long b0 = __ballot(lane < cnt && key[0] == row);
            const unsigned long long b1 = __ballot(lane + 64 < cnt && key[1] == row);
            if (b0) return __ffsll((long long)b0) - 1;
            if (b1) return 63 + __ffsll((long long)b1);
            const int i = cnt++;
            if (lane == (i & 63)) {
                if (i < 64) {
                    key[0] = row;
                    content[0] = row;
                } else {
                    key[1] = row;
                    content[1] = row;
                }
            }
            return i;
        };
        auto get = [&](int i) {
            return i < 64 ? __shfl(content[0], i, 64) : __shfl(content[1], i - 64, 64);
        };
        auto set = [&](int i, int v) {
            if (lane == (i & 63)) {
                if (i < 64)
                    content[0] = v;
                else
                    content[1] = v;
            }
        };
        for (int j = 0; j < nb; ++j) {
            const int a = find(k0 + j), b = find(s_pivs[j]);
            const int ca = get(a), cb = get(b);
            set(a, cb);
            set(b, ca);
        }
        const bool ch0 = lane < cnt && content[0] != key[0];
        const bool ch1 = lane + 64 < cnt && content[1] != key[1];
        const unsigned long long m0 = __ballot(ch0), m1 = __ballot(ch1);
        const unsigned long long below = (1ull << lane) - 1;
        const int n0 = __popcll(m0);
        const int i0 = __popcll(m0 & below), i1 = n0 + __popcll(m1 & below);
        if (ch0) {
            perm[1 + i0] = key[0];
            perm[1 + 2 * NB + i0] = content[0];
        }
        if (ch1) {
            perm[1 + i1] = key[1];
            perm[1 + 2 * NB + i1] = content[1];
        }
        if (lane == 0) perm[0] = n0 + __popcll(m1);
    }
}

// The row changes of panels pk .. pk + cnt - 1, then U12 = L11^-1 A12 for the
// columns [c0, c0 + nc) (k_trsm_lower_cols' body, 64 columns per workgroup),
// in one launch; workgroups beyond those of the solve apply the changes alone
// to the columns [l0, l0 + nl), four per workgroup.  In a solve workgroup
// wave w permutes the block's columns w, w + 4, ..., all sixteen columns'
// sources read before any destination is written, and the solve reads them
// after a barrier.
__global__ __launch_bounds__(256) void k_perm_trsm(cd* __restrict__ A, size_t lda, int k0, int nb,
                                                   int c0, int nc, int l0, int nl,
                                                   const int* __restrict__ perm, int cnt) {
    __shared__ cd Ls[TB * TB];
    __shared__ cd X[TB * (TB + 1)];
    const int ntr = (nc + TRSM_COLS_PER_WG - 1) / TRSM_COLS_PER_WG;
    const int w = threadIdx.x >> 6, s = threadIdx.x & 63;
    if ((int)blockIdx.x >= ntr) {
        const int c = l0 + ((int)blockIdx.x - ntr) * 4 + w;
        const bool colok = c < l0 + nl;
        for (int p = 0; p < cnt; ++p) {
            const int* __restrict__ pm = perm + (size_t)p * (1 + 4 * NB);
            const int m = pm[0];
            cd v[2];
            for (int h = 0; h < 2; ++h) {
                const int i = s + 64 * h;
                if (colok && i < m) v[h] = A[(size_t)c * lda + pm[1 + 2 * NB + i]];
            }
            __syncthreads();
            for (int h = 0; h < 2; ++h) {
                const int i = s + 64 * h;
                if (colok && i < m) A[(size_t)c * lda + pm[1 + i]] = v[h];
            }
            __syncthreads();
        }
        return;
    }
    const int cb = c0 + blockIdx.x * TRSM_COLS_PER_WG;
    for (int p = 0; p < cnt; ++p) {
        const int* __restrict__ pm = perm + (size_t)p * (1 + 4 * NB);
        const int m = pm[0];
        cd v[16][2];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int c = cb + w + 4 * u;
            for (int h = 0; h < 2; ++h) {
                const int i = s + 64 * h;
                if (c < c0 + nc && i < m) v[u][h] = A[(size_t)c * lda + pm[1 + 2 * NB + i]];
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int c = cb + w + 4 * u;
            for (int h = 0; h < 2; ++h) {
                const int i = s + 64 * h;
                if (c < c0 + nc && i < m) A[(size_t)c * lda + pm[1 + i]] = v[u][h];
            }
        }
        __syncthreads();
    }
    trsm_lower_cols_body<cd, false>(blockIdx.x, A, lda, k0, nb, c0, nc, nullptr, nullptr, 0, Ls,
                                    X);
}

// ---------------------------------------------------------------------------
// C[m x nc] -= A[m x K] B[K x nc], complex FP64, column-major, K <= NB.
// 256 threads = 4 waves; a workgroup owns a 64 x 64 tile of C, each wave a
// 32 x 32 quarter made of 2 x 2 MFMA 16x16 blocks with separate real and
// imaginary accumulators.  The K dimension is staged through LDS in chunks
// of 32 as separate real/imaginary planes.
// MFMA 16x16x4 f64 operand layout: A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, D[i][j] in lane j + 16(i%4), register i/4 (unlike the f32-output
// 16x16 forms, whose D rows run 4 per lane group).
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int GB = 64;   // tile rows/cols
constexpr int GK = 32;   // K chunk

__global__ __launch_bounds__(256) void k_zgemm_mfma(int m, int nc, int K, const cd* __restrict__ A,
                                                    size_t lda, const cd* __restrict__ B,
                                                    size_t ldb, cd* __restrict__ C, size_t ldc) {
    __shared__ double sAr[GK][GB + 1], sAi[GK][GB + 1];
    __shared__ double sBr[GK][GB + 1], sBi[GK][GB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bm = blockIdx.x * GB, bn = blockIdx.y * GB;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    v4d accR[2][2], accI[2][2];
    // load C into the accumulators
    for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj) {
            int col = bn + wn + 16 * bj + li;
            for (int r = 0; r < 4; ++r) {
                int row = bm + wm + 16 * bi + 4 * r + lk;
                cd v = (row < m && col < nc) ? C[(size_t)col * ldc + row] : mk(0, 0);
                accR[bi][bj][r] = v.x;
                accI[bi][bj][r] = v.y;
            }
        }
    for (int kc = 0; kc < K; kc += GK) {
        const int kn = min(GK, K - kc);
        __syncthreads();
        for (int t = tid; t < GK * GB; t += 256) {
            int k = t / GB, i = t - k * GB;
            cd v = (k < kn && bm + i < m) ? A[(size_t)(kc + k) * lda + bm + i] : mk(0, 0);
            sAr[k][i] = v.x;
            sAi[k][i] = v.y;
        }
        for (int t = tid; t < GK * GB; t += 256) {
            int j = t / GK, k = t - j * GK;
            cd v = (k < kn && bn + j < nc) ? B[(size_t)(bn + j) * ldb + kc + k] : mk(0, 0);
            sBr[k][j] = v.x;
            sBi[k][j] = v.y;
        }
        __syncthreads();
        for (int k = 0; k < GK; k += 4) {
            const int kk = k + lk;
            double ar[2], ai[2], br[2], bi_[2];
            for (int b = 0; b < 2; ++b) {
                ar[b] = sAr[kk][wm + 16 * b + li];
                ai[b] = sAi[kk][wm + 16 * b + li];
                br[b] = sBr[kk][wn + 16 * b + li];
                bi_[b] = sBi[kk][wn + 16 * b + li];
            }
            for (int x = 0; x < 2; ++x)
                for (int y = 0; y < 2; ++y) {
                    accR[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[x], br[y], accR[x][y], 0, 0, 0);
                    accR[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], bi_[y], accR[x][y], 0, 0, 0);
                    accI[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[x], bi_[y], accI[x][y], 0, 0, 0);
                    accI[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[x], br[y], accI[x][y], 0, 0, 0);
                }
        }
    }
    for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj) {
            int col = bn + wn + 16 * bj + li;
            for (int r = 0; r < 4; ++r) {
                int row = bm + wm + 16 * bi + 4 * r + lk;
                if (row < m && col < nc) C[(size_t)col * ldc + row] = mk(accR[bi][bj][r], accI[bi][bj][r]);
            }
        }
}

// The same update for K <= 64 with every load issued at once: the 64 x K
// block of A, the K x 64 block of B and the C tile (48 loads per thread in
// flight), A and B into LDS (the whole K), one barrier, then the MFMAs in
// the order k_zgemm_mfma issues them (the same sums, rounded alike).  The
// small updates on the factorisation's chain (inside a group, K = 64) are
// load-latency bound in the chunked form.
constexpr int G64K = 64;
__global__ __launch_bounds__(256) void k_zgemm64(int m, int nc, int K, const cd* __restrict__ A,
                                                 size_t lda, const cd* __restrict__ B, size_t ldb,
                                                 cd* __restrict__ C, size_t ldc) {
    __shared__ double sAr[G64K][GB + 1], sAi[G64K][GB + 1];
    __shared__ double sBr[G64K][GB + 1], sBi[G64K][GB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bm = blockIdx.x * GB, bn = blockIdx.y * GB;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    cd av[16], bv[16], cv[16];
    // staging: entry t = tid + 256 u is (k = t / 64, i = t % 64) of A and
    // (j = t / 64, k = t % 64) of B
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int t = tid + 256 * u, k = t >> 6, i = t & 63;
        av[u] = (k < K && bm + i < m) ? A[(size_t)k * lda + bm + i] : mk(0.0, 0.0);
        bv[u] = (i < K && bn + k < nc) ? B[(size_t)(bn + k) * ldb + i] : mk(0.0, 0.0);
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int col = bn + wn + 16 * bj + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = bm + wm + 16 * bi + 4 * r + lk;
                cv[(bi * 2 + bj) * 4 + r] =
                    (row < m && col < nc) ? C[(size_t)col * ldc + row] : mk(0.0, 0.0);
            }
        }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int t = tid + 256 * u, k = t >> 6, i = t & 63;
        sAr[k][i] = av[u].x;
        sAi[k][i] = av[u].y;
        sBr[i][k] = bv[u].x;
        sBi[i][k] = bv[u].y;
    }
    v4d accR[2][2], accI[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                accR[bi][bj][r] = cv[(bi * 2 + bj) * 4 + r].x;
                accI[bi][bj][r] = cv[(bi * 2 + bj) * 4 + r].y;
            }
    __syncthreads();
    for (int k = 0; k < K; k += 4) {
        const int kk = k + lk;
        double ar[2], ai[2], br[2], bi_[2];
        for (int b = 0; b < 2; ++b) {
            ar[b] = sAr[kk][wm + 16 * b + li];
            ai[b] = sAi[kk][wm + 16 * b + li];
            br[b] = sBr[kk][wn + 16 * b + li];
            bi_[b] = sBi[kk][wn + 16 * b + li];
        }
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y) {
                accR[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[x], br[y], accR[x][y], 0, 0, 0);
                accR[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[x], bi_[y], accR[x][y], 0, 0, 0);
                accI[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar[x], bi_[y], accI[x][y], 0, 0, 0);
                accI[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[x], br[y], accI[x][y], 0, 0, 0);
            }
    }
    for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj) {
            const int col = bn + wn + 16 * bj + li;
            for (int r = 0; r < 4; ++r) {
                const int row = bm + wm + 16 * bi + 4 * r + lk;
                if (row < m && col < nc)
                    C[(size_t)col * ldc + row] = mk(accR[bi][bj][r], accI[bi][bj][r]);
            }
        }
}

// Trailing update, second form: C[m x nc] -= A[m x K] B[K x nc] (K <= NB)
// with 128 x 128 workgroup tiles.  512 threads = 8 waves of 64 x 32 (4 x 2
// MFMA blocks of 16 x 16), the product accumulated from zero and C read and
// updated once at the end.  K runs in chunks of 16 through two LDS buffers
// of real and imaginary planes (rows padded by 2 doubles); the next chunk's
// operands are loaded into registers while the MFMAs of the current one run.
// Tiles are numbered so that each XCD takes a contiguous row-major range
// (workgroup id runs on XCD id % 8), sharing A tiles in its L2.  The grid is
// a multiple of 8 workgroups; with fewer workgroups than tiles each one
// takes every (grid / 8)-th tile of its XCD's range in turn.
constexpr int ZB = 128, ZK = 16, ZLD = ZB + 2, ZXCD = 8;
__global__ __launch_bounds__(512) void k_zgemm_mfma2(int m, int nc, int K,
                                                     const cd* __restrict__ A, size_t lda,
                                                     const cd* __restrict__ B, size_t ldb,
                                                     cd* __restrict__ C, size_t ldc) {
    __shared__ double sA[2][2][ZK][ZLD];
    __shared__ double sB[2][2][ZK][ZLD];
    const int gx = (m + ZB - 1) / ZB, gy = (nc + ZB - 1) / ZB, tiles = gx * gy;
    const int per = (tiles + ZXCD - 1) / ZXCD;
    const int xcd = blockIdx.x % ZXCD, step = gridDim.x / ZXCD;
    for (int i = blockIdx.x / ZXCD; i < per; i += step) {
    const int t = xcd * per + i;
    if (t >= tiles) break;
    const int bm = (t / gy) * ZB, bn = (t % gy) * ZB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = 64 * (wave & 1), wn = 32 * (wave >> 1);
    const int li = lane & 15, lk = lane >> 4;
    // staging assignment: A row ai, k = 4 aq .. 4 aq + 3; B column bj, same k
    const int ai = tid & 127, aq = tid >> 7;
    const int bj = tid >> 2, bq = tid & 3;
    const bool arow = bm + ai < m, bcol = bn + bj < nc;
    cd ra[4], rb[4];
    auto gload = [&](int kc) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = kc + 4 * aq + q;
            ra[q] = (arow && k < K) ? A[(size_t)k * lda + bm + ai] : mk(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = kc + 4 * bq + q;
            rb[q] = (bcol && k < K) ? B[(size_t)(bn + bj) * ldb + k] : mk(0.0, 0.0);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sA[buf][0][4 * aq + q][ai] = ra[q].x;
            sA[buf][1][4 * aq + q][ai] = ra[q].y;
            sB[buf][0][4 * bq + q][bj] = rb[q].x;
            sB[buf][1][4 * bq + q][bj] = rb[q].y;
        }
    };
    v4d accR[4][2], accI[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            accR[x][y] = v4d{0.0, 0.0, 0.0, 0.0};
            accI[x][y] = v4d{0.0, 0.0, 0.0, 0.0};
        }
    auto compute = [&](int buf) {
#pragma unroll
        for (int s = 0; s < ZK / 4; ++s) {
            const int k = 4 * s + lk;
            double ar[4], an[4], br[2], bi[2];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                ar[x] = sA[buf][0][k][wm + 16 * x + li];
                an[x] = -sA[buf][1][k][wm + 16 * x + li];
            }
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                br[y] = sB[buf][0][k][wn + 16 * y + li];
                bi[y] = sB[buf][1][k][wn + 16 * y + li];
            }
            // re += ar br - ai bi, im += ar bi + ai br: the 16 first products,
            // then the 16 second ones (16 independent MFMAs between two that
            // update the same accumulator)
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    accR[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], br[y], accR[x][y], 0, 0, 0);
                    accI[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[x], bi[y], accI[x][y], 0, 0, 0);
                }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    accR[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(an[x], bi[y], accR[x][y], 0, 0, 0);
                    accI[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-an[x], br[y], accI[x][y], 0, 0, 0);
                }
        }
    };
    const int nch = (K + ZK - 1) / ZK;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        if (c + 1 < nch) gload(ZK * (c + 1));
        compute(c & 1);
        if (c + 1 < nch) lstore((c + 1) & 1);
        __syncthreads();
    }
    // D[i][j] of a 16 x 16 block: lane j + 16 (i % 4), register i / 4
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int col = bn + wn + 16 * y + li;
            if (col >= nc) continue;
            cd* cp = C + (size_t)col * ldc;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = bm + wm + 16 * x + 4 * r + lk;
                if (row < m) {
                    const cd v = cp[row];
                    cp[row] = mk(v.x - accR[x][y][r], v.y - accI[x][y][r]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Solve with the factors, one right-hand side, with one work vector y.
// Forward, group by group: the group's row changes on b (k_permute_group),
// then per panel k_fwd_blk: y_k = L11^-1 b_k into y, and b_r -= L21[r,:] y_k
// in place for the rows below the block.  Backward, last panel first,
// k_bwd_blk: x_k = U11^-1 y_k into b, and y_r -= U[r, k] x_k in place for
// the rows above the block.  A step reads its block's entries of the vector
// it updates and writes its block's result to the other vector, so no
// workgroup writes what another one of the same step reads.
// Each workgroup takes 64 rows; thread (r, q) = (tid & 63, tid >> 6) loads
// row r's entries of the block columns 16q .. 16q + 15 before the block
// solve, so those loads overlap it, and the four partial sums are added in
// a fixed order.  Every workgroup solves the 64 x 64 triangular block itself
// (one wave, lane = row of the block, v_readlane broadcasts).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void blk_row_loads(const cd* __restrict__ A, size_t lda, int k0, int nb,
                                              int row, bool rok, int q, cd (&a)[16]) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int c = 16 * q + i;
        a[i] = (rok && c < nb) ? A[(size_t)(k0 + c) * lda + row] : mk(0.0, 0.0);
    }
}

// The nb x nb diagonal block at (k0, k0), column-major with stride NB, the
// rest of the 64 x 64 tile the identity.
__device__ __forceinline__ void blk_stage(const cd* __restrict__ A, size_t lda, int k0, int nb,
                                          cd* __restrict__ T) {
    const int tid = threadIdx.x;
    cd v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int t = tid + 256 * u, c = t >> 6, rr = t & 63;
        v[u] = (c < nb && rr < nb) ? A[(size_t)(k0 + c) * lda + k0 + rr]
                                   : mk(rr == c ? 1.0 : 0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) T[tid + 256 * u] = v[u];
}

// out[row] = base - (s0 + s1 + s2 + s3) with s_q thread (r, q)'s partial sum.
__device__ __forceinline__ cd blk_reduce(const cd (&a)[16], const cd* __restrict__ xs, int q, int r,
                                         cd (*part)[64], cd base) {
    cd s = mk(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < 16; ++i) s = s + a[i] * xs[16 * q + i];
    if (q > 0) part[q - 1][r] = s;
    __syncthreads();
    return base - (((s + part[0][r]) + part[1][r]) + part[2][r]);
}

__global__ __launch_bounds__(256) void k_fwd_blk(const cd* __restrict__ A, size_t lda, int n,
                                                 int k0, int nb, cd* __restrict__ b,
                                                 cd* __restrict__ y) {
    __shared__ cd T[NB * NB];
    __shared__ cd ys[NB];
    __shared__ cd part[3][64];
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int row = k0 + nb + blockIdx.x * 64 + r;
    const bool rok = row < n;
    cd a[16];
    blk_row_loads(A, lda, k0, nb, row, rok, q, a);
    const cd br = (q == 0 && rok) ? b[row] : mk(0.0, 0.0);
    blk_stage(A, lda, k0, nb, T);
    __syncthreads();
    if (q == 0) {
        // unit lower: lane r subtracts L[r][i] y_i for every i < r
        cd yr = r < nb ? b[k0 + r] : mk(0.0, 0.0);
        for (int i = 0; i + 1 < nb; ++i) {
            const cd yi = readlane_c(yr, i);
            if (r > i) yr = cfms(yr, T[i * NB + r], yi);
        }
        ys[r] = yr;
        if (blockIdx.x == 0 && r < nb) y[k0 + r] = yr;
    }
    __syncthreads();
    const cd v = blk_reduce(a, ys, q, r, part, br);
    if (q == 0 && rok) b[row] = v;
}

__global__ __launch_bounds__(256) void k_bwd_blk(const cd* __restrict__ A, size_t lda, int k0,
                                                 int nb, cd* __restrict__ y, cd* __restrict__ x) {
    __shared__ cd T[NB * NB];
    __shared__ cd xs[NB];
    __shared__ cd part[3][64];
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int row = blockIdx.x * 64 + r;
    const bool rok = row < k0;
    cd a[16];
    blk_row_loads(A, lda, k0, nb, row, rok, q, a);
    const cd yr0 = (q == 0 && rok) ? y[row] : mk(0.0, 0.0);
    blk_stage(A, lda, k0, nb, T);
    __syncthreads();
    if (q == 0) {
        // upper: lane r divides by U[r][r] at step r, then subtracts
        // U[r][i] x_i for every i > r
        const cd dinv = cinv(T[r * NB + r]);
        cd xr = r < nb ? y[k0 + r] : mk(0.0, 0.0);
        for (int i = nb - 1; i >= 0; --i) {
            if (r == i) xr = xr * dinv;
            const cd xi = readlane_c(xr, i);
            if (r < i) xr = cfms(xr, T[i * NB + r], xi);
        }
        xs[r] = xr;
        if (blockIdx.x == 0 && r < nb) x[k0 + r] = xr;
    }
    __syncthreads();
    const cd v = blk_reduce(a, xs, q, r, part, yr0);
    if (q == 0 && rok) y[row] = v;
}

__global__ void k_fill_int(int* p, int v, size_t cnt) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = v;
}

// b <- the row changes of cnt consecutive panels, in panel order (perm of
// panel i at perm + i (1 + 4 NB) = {m, dst[0..2NB), src[0..2NB)}), in
// place: every source of a panel is read before any destination is written.
__global__ __launch_bounds__(2 * NB) void k_permute_group(cd* __restrict__ b,
                                                          const int* __restrict__ perm, int cnt) {
    const int t = threadIdx.x;
    for (int i = 0; i < cnt; ++i) {
        const int* pm = perm + (size_t)i * (1 + 4 * NB);
        const int m = pm[0];
        cd v = mk(0.0, 0.0);
        if (t < m) v = b[pm[1 + 2 * NB + t]];
        __syncthreads();
        if (t < m) b[pm[1 + t]] = v;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Host orchestration.
// ---------------------------------------------------------------------------
#define LU_CHK(expr)                                                        \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            err = std::string("LU: ") + hipGetErrorString(e_);              \
            return e_ == hipErrorOutOfMemory ? BEMX_ERR_NOMEM : BEMX_ERR_HIP; \
        }                                                                   \
    } while (0)

enum { W_IPIV, W_SLOTS, W_CAND, W_DROW, W_PERM, W_INFO, W_STATUS, W_BTMP, W_N };

// Bytes of buffer i of the FP64 factorisation and solve for order n:
// pivots, the panel's tagged exchange words (per column and workgroup a
// 3-word candidate slot and a 64-entry candidate row of 4 words per entry;
// per column the 64-entry row d), the panels' row permutations, INFO, the
// status word and the solve's work vector.
constexpr size_t fp64_size(int n, int i) {
    return i == W_IPIV     ? sizeof(int) * (size_t)n
           : i == W_SLOTS  ? 3 * sizeof(u64) * NB * (size_t)((n + PANEL_ROWS - 1) / PANEL_ROWS)
           : i == W_CAND   ? 4 * sizeof(u64) * NB * (size_t)((n + PANEL_ROWS - 1) / PANEL_ROWS) * NB
           : i == W_DROW   ? 4 * sizeof(u64) * NB * NB
           : i == W_PERM   ? sizeof(int) * (size_t)((n + NB - 1) / NB) * (1 + 4 * NB)
           : i == W_INFO   ? sizeof(long long)
           : i == W_STATUS ? sizeof(int)
                           : sizeof(cd) * (size_t)n;
}

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The buffers of fp64_size laid out back to back, 256-byte aligned.
constexpr size_t fp64_carved_bytes(int n) {
    size_t t = 0;
    for (int i = 0; i < W_N; ++i) t += align256(fp64_size(n, i));
    return t;
}

size_t fp64_factor_bytes(int n) {
    size_t t = 0;
    for (int i = 0; i < W_N; ++i) t += std::max<size_t>(fp64_size(n, i), 256);
    return t;
}

// Device pointers of one FP64 workspace.
struct Fp64Ws {
    int* ipiv = nullptr;
    u64* slots = nullptr;
    u64* cand = nullptr;
    u64* drow = nullptr;
    int* perm = nullptr;
    long long* info = nullptr;
    int* status = nullptr;
    cd* btmp = nullptr;
    void set(void* const p[W_N]) {
        ipiv = (int*)p[W_IPIV];
        slots = (u64*)p[W_SLOTS];
        cand = (u64*)p[W_CAND];
        drow = (u64*)p[W_DROW];
        perm = (int*)p[W_PERM];
        info = (long long*)p[W_INFO];
        status = (int*)p[W_STATUS];
        btmp = (cd*)p[W_BTMP];
    }
};

// The workspace in LuWork's grow-only buffers (allocates on growth).
static int fp64_ws_alloc(LuWork& w, int n, Fp64Ws& ws, std::string& err) {
    void* p[W_N];
    hipError_t e;
    for (int i = 0; i < W_N; ++i) {
        p[i] = w.get(i, fp64_size(n, i), &e);
        LU_CHK(e);
    }
    ws.set(p);
    return BEMX_OK;
}

// The workspace carved out of one block of fp64_carved_bytes(n) bytes.
static void fp64_ws_carve(char* base, int n, Fp64Ws& ws) {
    void* p[W_N];
    size_t off = 0;
    for (int i = 0; i < W_N; ++i) {
        p[i] = base + off;
        off += align256(fp64_size(n, i));
    }
    ws.set(p);
}

// Static device workspace of bemx_linsolve_device_noalloc: the FP64
// factorisation's buffers for orders up to NOALLOC_MAX_N (35.3 MB).  Each
// GPU the library runs on holds its own copy until the process ends; every
// call on that GPU shares it.  Host code takes its address from
// hipGetSymbolAddress (scratch_setup, once per device when a context is
// made on it) and never uses the array's name, which on the host is only a
// stand-in for the symbol.
constexpr int NOALLOC_MAX_N = 32768;
constexpr size_t NOALLOC_BYTES = fp64_carved_bytes(NOALLOC_MAX_N);
static __device__ __align__(256) char scratch[NOALLOC_BYTES];
constexpr int MAX_DEVICES = 64;
static std::atomic<void*> scratch_addr[MAX_DEVICES];
static std::mutex scratch_mu;

hipError_t scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_addr[device].load()) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_addr[device].store(p);
    return hipSuccess;
}

int noalloc_max_n() { return NOALLOC_MAX_N; }

std::string coop_refused_message(const char* what) {
    return std::string(what) +
           ": cooperative launch refused (hipErrorIllegalState): another process on this GPU "
           "holds the device's cooperative-launch resource; run one solver process per GPU";
}

hipError_t preload_lu_code() {
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, (const void*)k_panel_fp64<false>);
    if (e == hipSuccess) e = hipFuncGetAttributes(&a, (const void*)k_panel_fp64<true>);
    return e;
}

// Panels per group of the FP64 factorisation: inside a group each panel
// updates only the group's columns (K = 64); the rest of the matrix gets one
// update per group with K = 64 x FP64_GROUP, which reads and writes it a
// quarter as often as an update per panel.  Row interchanges follow LAPACK's
// convention inside a group (a panel's changes also reach the group's earlier
// L columns, so the group's L rows share one order) and the per-panel one
// across groups; the forward solve applies a group's changes before its
// blocks.
constexpr int FP64_GROUP = 4;

// C -= A B for the trailing update: 128 x 128 tiles from 1024 x 1024 on
// (on the 6480-DOF solve 32 against 27 TFLOP/s at m = n = 3216 .. 6416), the
// 64 x 64 tiles below, where 128 x 128 tiles leave most CUs idle (m = n =
// 416: 16 tiles, 49 against 23 us).
// cap > 0: at most cap workgroups of the 128 x 128 form (a multiple of 8),
// each taking several tiles, so the rest of the device stays free for other
// streams' work.
static void launch_zgemm(hipStream_t st, int m, int nc, int K, const cd* A, size_t lda,
                         const cd* B, size_t ldb, cd* C, size_t ldc, int cap = 0) {
    if (m >= 1024 && nc >= 1024) {
        const int tiles = ((m + ZB - 1) / ZB) * ((nc + ZB - 1) / ZB);
        const int per = (tiles + ZXCD - 1) / ZXCD;
        const int wgs = cap > 0 ? std::min(per * ZXCD, cap) : per * ZXCD;
        hipLaunchKernelGGL(k_zgemm_mfma2, dim3(wgs), dim3(512), 0, st, m, nc, K, A, lda, B, ldb,
                           C, ldc);
    } else if (K <= G64K) {
        hipLaunchKernelGGL(k_zgemm64, dim3((m + GB - 1) / GB, (nc + GB - 1) / GB), dim3(256), 0,
                           st, m, nc, K, A, lda, B, ldb, C, ldc);
    } else {
        hipLaunchKernelGGL(k_zgemm_mfma, dim3((m + GB - 1) / GB, (nc + GB - 1) / GB), dim3(256), 0,
                           st, m, nc, K, A, lda, B, ldb, C, ldc);
    }
}

// Tags of the panel exchange words: n consecutive tags, never 0, unique in
// the process until the 32-bit counter wraps (4e9 panel launches).
static std::atomic<unsigned> g_panel_tag{0};
static unsigned tag_block(int n) {
    for (;;) {
        const unsigned t = g_panel_tag.fetch_add((unsigned)n) + 1;
        if (t != 0 && t + (unsigned)n > t) return t;  // skip a block that wraps through 0
    }
}

// Rows per panel workgroup for a panel of m rows when at most cap
// workgroups can be co-resident: PANEL_ROWS while that suffices, else the
// smallest multiple of PANEL_ROWS that fits (the tall form).  The environment
// variable BEMX_PANEL_ROWS (a multiple of PANEL_ROWS, read at every
// factorisation) raises it, so the tall form can be compared with the plain
// one at orders both handle.
static int forced_panel_rows() {
    const char* v = std::getenv("BEMX_PANEL_ROWS");
    const int r = v ? std::atoi(v) : 0;
    return r > PANEL_ROWS && r % PANEL_ROWS == 0 ? r : PANEL_ROWS;
}
static int panel_rows(int m, int cap, int forced) {
    const int need = (m + PANEL_ROWS - 1) / PANEL_ROWS;
    const int rpw = cap > 0 && need > cap ? PANEL_ROWS * ((need + cap - 1) / cap) : PANEL_ROWS;
    return std::max(rpw, forced);
}

// tw (may be NULL): the workspace whose events time the factorisation.
static int factor_fp64(const Fp64Ws& ws, hipStream_t st, int n, cd* A, size_t lda, LuWork* tw,
                       std::string& err) {
    const int npanel = (n + NB - 1) / NB;
    int* ipiv = ws.ipiv;
    u64* slots = ws.slots;
    u64* cand = ws.cand;
    u64* drow = ws.drow;
    int* perm = ws.perm;
    long long* info = ws.info;
    int* status = ws.status;
    const unsigned tag0 = tag_block(npanel);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, (int*)info, 0, 2);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, status, 0, 1);
    int dev = 0, ncu = 0, perCU = 0, perCUt = 0;
    LU_CHK(hipGetDevice(&dev));
    LU_CHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    LU_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_panel_fp64<false>, 256, 0));
    LU_CHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCUt, k_panel_fp64<true>, 256, 0));
    // co-resident panel workgroups: the device's limit, and at most 256
    // (the candidate slots are polled by one wave, four per lane)
    const int cap = std::min(256, ncu * std::min(perCU, perCUt));
    const int forced = forced_panel_rows();
    if (tw) {
        LU_CHK(tw->ensure_side());
        LU_CHK(hipEventRecord(tw->ev_f0, st));
    }
    auto permk_of = [&](int pk) { return perm + (size_t)pk * (1 + 4 * NB); };
    // panels pk .. pk + cnt - 1's row changes on the columns [c0, c0 + nc),
    // which are then solved for with panel k0's L11, and on [l0, l0 + nl)
    auto perm_trsm = [&](hipStream_t s, int k0, int nb, int c0, int nc, int l0, int nl, int pk,
                         int cnt) {
        const int wgs = (nc + TRSM_COLS_PER_WG - 1) / TRSM_COLS_PER_WG + (nl + 3) / 4;
        if (wgs > 0)
            hipLaunchKernelGGL(k_perm_trsm, dim3(wgs), dim3(256), 0, s, A, lda, k0, nb, c0, nc,
                               l0, nl, (const int*)permk_of(pk), cnt);
    };
    auto trsm = [&](hipStream_t s, int k0, int nb, int c0, int nc) {
        hipLaunchKernelGGL((k_trsm_lower_cols<cd, false>),
                           dim3((nc + TRSM_COLS_PER_WG - 1) / TRSM_COLS_PER_WG), dim3(256), 0, s,
                           A, lda, k0, nb, c0, nc, (const int*)nullptr, (__bf16*)nullptr,
                           (size_t)0);
    };
    // Group [g0, pend) applied to the columns [cr, cr + nr) right of it: the
    // group's row changes in panel order, its 64-row blocks of U solved for
    // in turn (each updating the group's later rows), then one update with
    // K = the group's width.
    auto apply_group = [&](hipStream_t s, int g0, int pend, int cr, int nr, int cap) {
        if (nr <= 0) return;
        const int G0 = g0 * NB, Ge = std::min(n, pend * NB);
        for (int pk = g0; pk < pend; ++pk) {
            const int k0 = pk * NB, nb = std::min(NB, n - k0);
            if (pk == g0)  // the group's row changes, then the first block's solve
                perm_trsm(s, k0, nb, cr, nr, 0, 0, g0, pend - g0);
            else
                trsm(s, k0, nb, cr, nr);
            const int mrow = Ge - (k0 + nb);
            if (mrow > 0)
                launch_zgemm(s, mrow, nr, nb, A + (size_t)k0 * lda + k0 + nb, lda,
                             A + (size_t)cr * lda + k0, lda, A + (size_t)cr * lda + k0 + nb, lda,
                             cap);
        }
        launch_zgemm(s, n - Ge, nr, Ge - G0, A + (size_t)G0 * lda + Ge, lda,
                     A + (size_t)cr * lda + G0, lda, A + (size_t)cr * lda + Ge, lda, cap);
    };
    // Look-ahead (with a workspace that has a second stream): after group g,
    // the next group's columns get group g on st, and every column beyond
    // them gets it on the side stream, which runs while st factors the next
    // group's panels.  st waits for the side stream before it applies the
    // next group to anything beyond its own columns (those columns still
    // lack group g there), and once more at the end.
    // BEMX_FP64_LOOKAHEAD=0 keeps every update on st (for A/B timing)
    static const bool lookahead = [] {
        const char* v = std::getenv("BEMX_FP64_LOOKAHEAD");
        return !(v && v[0] == '0');
    }();
    // BEMX_FP64_SIDE_WGS: workgroups of the side stream's updates (a
    // multiple of 8; 0 = as many as the next panels leave free).  192 of the
    // 256 CUs leave room for the panels and the next group's updates on st
    // (6480 unknowns: 47.8 ms against 51.1 with one workgroup per tile and
    // 49.3 with 160).
    static const int side_wgs = [] {
        const char* v = std::getenv("BEMX_FP64_SIDE_WGS");
        const int c = v ? std::atoi(v) : 192;
        return c > 0 ? (c + ZXCD - 1) / ZXCD * ZXCD : 0;
    }();
    LuWork* const la = lookahead ? tw : nullptr;
    const hipStream_t sd = la ? la->side : st;
    bool side_pending = false;
    for (int g0 = 0; g0 < npanel; g0 += FP64_GROUP) {
        const int pend = std::min(npanel, g0 + FP64_GROUP);
        const int G0 = g0 * NB, Ge = std::min(n, pend * NB);
        for (int pk = g0; pk < pend; ++pk) {
            const int k0 = pk * NB;
            const int nb = std::min(NB, n - k0);
            const int m = n - k0;
            int rpw = panel_rows(m, cap, forced);
            const int P = (m + rpw - 1) / rpw;
            if (cap < 1 || P > cap) {
                err = "LU panel needs more co-resident workgroups than the device holds";
                return BEMX_ERR_ARG;
            }
            int* permk = permk_of(pk);
            unsigned tag = tag0 + (unsigned)pk;
            void* args[] = {&A, &lda, (void*)&n, (void*)&k0, (void*)&nb, &rpw, &ipiv, &slots,
                            &cand, &drow, &permk, &info, &status, &tag};
            const void* kfn = rpw > PANEL_ROWS ? (const void*)k_panel_fp64<true>
                                               : (const void*)k_panel_fp64<false>;
            const hipError_t ec =
                hipLaunchCooperativeKernel(kfn, dim3(P), dim3(256), args, 0, st);
            if (ec == hipErrorIllegalState) {
                err = coop_refused_message("LU panel");
                return BEMX_ERR_HIP;
            }
            LU_CHK(ec);
            // this panel's row changes on the group's earlier L columns (the
            // group's L rows stay in one order) and on its later columns,
            // which are then solved for and updated with K = 64
            const int c0 = k0 + nb, ncg = Ge - c0;
            perm_trsm(st, k0, nb, c0, ncg, G0, k0 - G0, pk, 1);
            if (ncg > 0)
                launch_zgemm(st, n - c0, ncg, nb, A + (size_t)k0 * lda + c0, lda,
                             A + (size_t)c0 * lda + k0, lda, A + (size_t)c0 * lda + c0, lda);
            LU_CHK(hipGetLastError());
        }
        const int nr = n - Ge;
        if (nr <= 0) continue;
        if (side_pending) LU_CHK(hipStreamWaitEvent(st, la->ev_bulk, 0));
        side_pending = false;
        // The second stream may hold only the CUs that the next group's
        // panels leave free (one workgroup per CU each).  On C of the ncu
        // CUs the bulk update takes ncu / C times as long as on all of them,
        // so it goes beside the next group only when that stretch costs less
        // than the group's panels take (about 1.2 ms for four panels, nearly
        // independent of n); the bulk's own time is estimated at 47 TFLOP/s.
        const int pnext = (nr + panel_rows(nr, cap, forced) - 1) / panel_rows(nr, cap, forced);
        const int room = (ncu - pnext) / ZXCD * ZXCD;
        const int nn = std::min(nr, FP64_GROUP * NB);  // the next group's columns
        const int cside = side_wgs > 0 ? std::min(side_wgs, room) : room;
        const double bulk_ms = 8.0 * nr * (double)(nr - nn) * (Ge - G0) / 47e9;
        const bool ahead =
            la && nr > nn && cside >= 64 && bulk_ms * ((double)ncu / cside - 1.0) < 1.2;
        if (!ahead) {
            apply_group(st, g0, pend, Ge, nr, 0);
        } else {
            apply_group(st, g0, pend, Ge, nn, 0);
            LU_CHK(hipEventRecord(la->ev_panel, st));
            LU_CHK(hipStreamWaitEvent(sd, la->ev_panel, 0));
            apply_group(sd, g0, pend, Ge + nn, nr - nn, cside);
            LU_CHK(hipEventRecord(la->ev_bulk, sd));
            side_pending = true;
        }
        LU_CHK(hipGetLastError());
    }
    if (side_pending) LU_CHK(hipStreamWaitEvent(st, la->ev_bulk, 0));
    if (tw) {
        LU_CHK(hipEventRecord(tw->ev_f1, st));
        tw->factor_timed = true;
    }
    return BEMX_OK;
}

static int solve_fp64(const Fp64Ws& ws, hipStream_t st, int n, const cd* A, size_t lda, cd* b,
                      std::string& err) {
    const int npanel = (n + NB - 1) / NB;
    cd* y = ws.btmp;
    for (int g0 = 0; g0 < npanel; g0 += FP64_GROUP) {
        const int pend = std::min(npanel, g0 + FP64_GROUP);
        hipLaunchKernelGGL(k_permute_group, dim3(1), dim3(2 * NB), 0, st, b,
                           ws.perm + (size_t)g0 * (1 + 4 * NB), pend - g0);
        for (int pk = g0; pk < pend; ++pk) {
            const int k0 = pk * NB, nb = std::min(NB, n - k0);
            const int rows = n - k0 - nb;
            hipLaunchKernelGGL(k_fwd_blk, dim3(std::max(1, (rows + 63) / 64)), dim3(256), 0, st,
                               A, lda, n, k0, nb, b, y);
        }
    }
    for (int pk = npanel - 1; pk >= 0; --pk) {
        const int k0 = pk * NB, nb = std::min(NB, n - k0);
        hipLaunchKernelGGL(k_bwd_blk, dim3(std::max(1, (k0 + 63) / 64)), dim3(256), 0, st, A,
                           lda, k0, nb, y, b);
    }
    LU_CHK(hipGetLastError());
    return BEMX_OK;
}

// INFO and the status word of the factorisation just enqueued (waits for it).
static int factor_status(const Fp64Ws& ws, hipStream_t st, long long* info, std::string& err) {
    long long hinfo = 0;
    int hstatus = 0;
    LU_CHK(hipMemcpyAsync(&hinfo, ws.info, sizeof(long long), hipMemcpyDeviceToHost, st));
    LU_CHK(hipMemcpyAsync(&hstatus, ws.status, sizeof(int), hipMemcpyDeviceToHost, st));
    LU_CHK(hipStreamSynchronize(st));
    if (hstatus & ST_SYNC_TIMEOUT) {
        err = "LU panel: inter-workgroup wait timed out";
        return BEMX_ERR_TIMEOUT;
    }
    if (info) *info = hinfo;
    if (hinfo > 0) {
        err = "LU: exactly singular pivot (INFO > 0)";
        return BEMX_ERR_SINGULAR;
    }
    return BEMX_OK;
}

int lu_solve(LuWork& w, hipStream_t st, int n, cd* A, size_t lda, cd* b, int nrhs, int mode,
             long long* info, double stats[4], std::string& err) {
    if (nrhs < 1) {
        err = "LU: no right-hand side";
        return BEMX_ERR_ARG;
    }
    if (stats) stats[0] = stats[1] = stats[2] = stats[3] = 0.0;
    w.factor_timed = false;
    int first = 0;  // first column the FP64 path still has to solve
    if (mode == BEMX_MODE_MIXED) {
        int sm = gmres_ir(w, st, n, A, lda, b, nrhs, stats, err, &first);
        if (sm != BEMX_ERR_NOCONV && sm != BEMX_ERR_SINGULAR) return sm;
        // the plain FP64 path is the fallback; A and the unsolved columns of
        // b are still intact
        if (stats) stats[2] = 1.0;
    }
    Fp64Ws ws;
    int s = fp64_ws_alloc(w, n, ws, err);
    if (!s) s = factor_fp64(ws, st, n, A, lda, &w, err);
    if (!s) s = factor_status(ws, st, info, err);
    if (s) return s;
    for (int j = first; j < nrhs; ++j) {
        s = solve_fp64(ws, st, n, A, lda, b + (size_t)j * n, err);
        if (s) return s;
    }
    return BEMX_OK;
}

int lu_solve_batch(LuWork& w, hipStream_t st, int n, int nz, cd* A, size_t lda, size_t sa, cd* b,
                   size_t sb, int nrhs, int mode, double* stats, std::string& err,
                   const std::function<int()>* after) {
    if (nz < 1 || nz > MIXED_MAX_BATCH || nrhs < 1) {
        err = "LU: batch of 1..16 problems with at least one right-hand side";
        return BEMX_ERR_ARG;
    }
    for (int i = 0; i < 4 * nz; ++i) stats[i] = 0.0;
    auto mat = [&](int z) { return (cd*)((char*)A + (size_t)z * sa); };
    auto rhs = [&](int z) { return (cd*)((char*)b + (size_t)z * sb); };
    if (mode != BEMX_MODE_MIXED) {
        if (after) {
            const int s = (*after)();
            if (s) return s;
        }
        for (int z = 0; z < nz; ++z) {
            long long info = 0;
            const int s = lu_solve(w, st, n, mat(z), lda, rhs(z), nrhs, mode, &info, stats + 4 * z,
                                   err);
            if (s) return s;
        }
        return BEMX_OK;
    }
    w.factor_timed = false;
    int code[MIXED_MAX_BATCH], nsolved[MIXED_MAX_BATCH];
    int s = gmres_ir_batch(w, st, n, nz, A, lda, sa, b, sb, nrhs, code, nsolved, stats, err,
                           after);
    if (s) return s;
    for (int z = 0; z < nz; ++z) {
        if (code[z] == BEMX_OK) continue;
        if (code[z] != BEMX_ERR_NOCONV && code[z] != BEMX_ERR_SINGULAR) return code[z];
        // the plain FP64 path is the fallback; the matrix and the unsolved
        // columns of this problem are still intact
        stats[4 * z + 2] = 1.0;
        Fp64Ws ws;
        long long info = 0;
        s = fp64_ws_alloc(w, n, ws, err);
        if (!s) s = factor_fp64(ws, st, n, mat(z), lda, &w, err);
        if (!s) s = factor_status(ws, st, &info, err);
        for (int j = nsolved[z]; !s && j < nrhs; ++j)
            s = solve_fp64(ws, st, n, mat(z), lda, rhs(z) + (size_t)j * n, err);
        if (s) return s;
    }
    return BEMX_OK;
}

int lu_solve_fp64_noalloc(int device, hipStream_t st, int n, cd* A, size_t lda, cd* b,
                          long long* info, std::string& err) {
    if (n < 1 || n > NOALLOC_MAX_N) {
        err = "LU: order outside 1..32768, the static workspace's range";
        return BEMX_ERR_ARG;
    }
    void* base = (device >= 0 && device < MAX_DEVICES) ? scratch_addr[device].load() : nullptr;
    if (!base) {
        err = "LU: the static workspace of this device is not set up (no context made on it)";
        return BEMX_ERR_STATE;
    }
    Fp64Ws ws;
    fp64_ws_carve((char*)base, n, ws);
    int s = factor_fp64(ws, st, n, A, lda, nullptr, err);
    if (!s) s = factor_status(ws, st, info, err);
    if (!s) s = solve_fp64(ws, st, n, A, lda, b, err);
    if (s) return s;
    LU_CHK(hipStreamSynchronize(st));
    return BEMX_OK;
}

}  // namespace bemx
