// bemx_lu.hpp -- dense complex solve on the device (LINSOLVE).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <string>

#include "bemx_common.hpp"

namespace bemx {

// Device workspace of the factorisation (grow-only).
struct LuWork {
    static constexpr int SLOTS = 32;  // slots 0..8: FP64 path, 9..: mixed path
    void* buf[SLOTS] = {nullptr};
    size_t cap[SLOTS] = {0};
    // Bumped whenever a buffer is (re)allocated: a captured graph holds the
    // addresses of every buffer it touches, so one captured before a
    // reallocation must never be replayed after it.
    unsigned long long gen = 0;
    // Flags of the cooperative triangular solves hold the epoch of the
    // launch that set them; epochs only grow, so they are cleared only when
    // the mixed path's arenas change (a new allocation or another order n,
    // which moves every buffer inside an arena), and never inside a solve:
    // arenas 0 .. arena_clear - 1 hold cleared flags of the layout of order
    // arena_n in workspace generation arena_gen.
    int epoch = 0;
    unsigned long long arena_gen = ~0ull;
    int arena_n = -1, arena_clear = 0;
    // second stream and events of the look-ahead factorisation (lazy)
    hipStream_t side = nullptr;
    hipEvent_t ev_panel = nullptr, ev_bulk = nullptr;
    // events around the last factorisation on the solve's stream (its device
    // time; factor_timed says whether they bracket one)
    hipEvent_t ev_f0 = nullptr, ev_f1 = nullptr;
    bool factor_timed = false;
    hipError_t ensure_side();
    // Captured factorisation graphs, one per size class and batch size:
    // keyed by the order, the workspace generation and the batch size (and
    // unused zeros); least recently used entries are evicted beyond GRAPHS.
    static constexpr int GRAPHS = 64, KEYS = 6;
    struct GraphEntry {
        const void* key[KEYS] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        hipGraphExec_t exec = nullptr;
        unsigned long long used = 0;
    };
    GraphEntry graphs[GRAPHS];
    unsigned long long tick = 0;
    long long captures = 0, replays = 0;  // graph statistics since creation
    // Batch sizes 1 .. ahead of an order are captured together at the
    // order's first capture (those the arena holds; 1: only the size asked
    // for); precaptures counts the graphs captured that way.  A pre-captured
    // graph's first launch counts as a replay.
    int ahead = 1;
    long long precaptures = 0;
    bool use_graphs = true;
    hipGraphExec_t find_graph(const void* const key[KEYS]);
    // takes ownership of exec; evicts the least recently used entry if full
    void add_graph(const void* const key[KEYS], hipGraphExec_t exec);
    void drop_graphs();
    LuWork();
    void* get(int slot, size_t bytes, hipError_t* e);
    size_t bytes() const;  // device bytes held by the workspace
    void release();
};

// Solves A X = B in place for nrhs columns of B (A overwritten by its
// factors, B by X).
// mode: 0 = FP64 partial-pivot LU; 1 = FP32 LU + FP64 GMRES-IR.
// stats: [GMRES iterations, final ||r||/||b||, 1 if the mixed path fell back
// to FP64, 0].
// FP32 LU (split-BF16x3 MFMA update) preconditioning FP64 GMRES-IR; A is
// read only, b is overwritten with x on success and left intact otherwise.
// b holds nrhs columns of n; *nsolved receives how many leading columns
// were solved when an error is returned.
int gmres_ir(LuWork& w, hipStream_t st, int n, const cd* A, size_t lda, cd* b, int nrhs,
             double stats[4], std::string& err, int* nsolved);

// Most problems in one batch of the mixed path.
constexpr int MIXED_MAX_BATCH = 16;

// The mixed solve of nz (1 .. MIXED_MAX_BATCH) problems of order n in one
// batch: problem z's matrix at A + z sa bytes (read only) and its nrhs
// right-hand sides at b + z sb bytes (leading dimension n; overwritten with
// the solutions that converge).  Each problem's arithmetic is that of
// gmres_ir alone.  Per problem: code[z] (BEMX_OK, BEMX_ERR_SINGULAR for a
// zero pivot in a diagonal block, BEMX_ERR_NOCONV when the refinement
// stalls), nsolved[z] leading columns solved, stats[4 z ..] = [GMRES
// iterations, worst final ||r||/||b||, 0, 0].  The return value reports a
// failure of the whole batch (HIP error, timed-out device wait).
// after (may be null) runs on the host once the factorisation is enqueued,
// before anything waits for it (a caller's work for other streams, such as
// the next batch's assembly, goes in behind the factorisation's launches).
int gmres_ir_batch(LuWork& w, hipStream_t st, int n, int nz, const cd* A, size_t lda, size_t sa,
                   cd* b, size_t sb, int nrhs, int* code, int* nsolved, double* stats,
                   std::string& err, const std::function<int()>* after = nullptr);

// lu_solve for a batch: the mixed solve of every problem (gmres_ir_batch),
// then the plain FP64 LU for each problem whose mixed solve failed
// (stats[4 z + 2] = 1).  mode BEMX_MODE_FP64 solves the problems one by one.
int lu_solve_batch(LuWork& w, hipStream_t st, int n, int nz, cd* A, size_t lda, size_t sa, cd* b,
                   size_t sb, int nrhs, int mode, double* stats, std::string& err,
                   const std::function<int()>* after = nullptr);

int lu_solve(LuWork& w, hipStream_t st, int n, cd* A, size_t lda, cd* b, int nrhs, int mode,
             long long* info, double stats[4], std::string& err);

// Device bytes of every workspace buffer of the mixed path (factorisation
// and refinement) for order n.
size_t mixed_workspace_bytes(int n, int nz = 1);

// Device bytes of the FP64 factorisation's workspace for order n (pivots,
// panel candidates, permutations, solve vector).
size_t fp64_factor_bytes(int n);

// FP64 solve of one right-hand side with its workspace in the static
// device array of the device (no allocation); n <= noalloc_max_n().
int lu_solve_fp64_noalloc(int device, hipStream_t st, int n, cd* A, size_t lda, cd* b,
                          long long* info, std::string& err);
int noalloc_max_n();
// Looks up the static array's device address on `device` (current).
hipError_t scratch_setup(int device);

// Error text for a cooperative launch the runtime refuses because another
// process on the GPU holds its cooperative-launch resource.
std::string coop_refused_message(const char* what);

// Query one kernel of each kernel file so that its device code is loaded by
// the calling thread (a context's creator) rather than by whichever thread
// makes the first launch.
hipError_t preload_lu_code();
hipError_t preload_mixed_code();
hipError_t preload_assembly_code();

}  // namespace bemx
