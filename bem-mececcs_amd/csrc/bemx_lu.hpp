// bemx_lu.hpp -- dense complex solve on the device (LINSOLVE).
#pragma once
#include <hip/hip_runtime.h>

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
    // launch that set them; epochs only grow, so the array is cleared once
    // per workspace generation, and never inside a solve.
    int epoch = 0;
    unsigned long long cleared_gen = ~0ull;
    // second stream and events of the look-ahead factorisation (lazy)
    hipStream_t side = nullptr;
    hipEvent_t ev_panel = nullptr, ev_bulk = nullptr;
    // events around the last factorisation on the solve's stream (its device
    // time; factor_timed says whether they bracket one)
    hipEvent_t ev_f0 = nullptr, ev_f1 = nullptr;
    bool factor_timed = false;
    hipError_t ensure_side();
    // Captured factorisation graphs, one per size class: keyed by the order,
    // the matrix address, its leading dimension and the workspace generation;
    // least recently used entries are evicted beyond GRAPHS.
    static constexpr int GRAPHS = 8;
    struct GraphEntry {
        const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
        hipGraphExec_t exec = nullptr;
        unsigned long long used = 0;
    };
    GraphEntry graphs[GRAPHS];
    unsigned long long tick = 0;
    long long captures = 0, replays = 0;  // graph statistics since creation
    bool use_graphs = true;
    hipGraphExec_t find_graph(const void* const key[4]);
    // takes ownership of exec; evicts the least recently used entry if full
    void add_graph(const void* const key[4], hipGraphExec_t exec);
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

int lu_solve(LuWork& w, hipStream_t st, int n, cd* A, size_t lda, cd* b, int nrhs, int mode,
             long long* info, double stats[4], std::string& err);

// Device bytes of every workspace buffer of the mixed path (factorisation
// and refinement) for order n.
size_t mixed_workspace_bytes(int n);

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
