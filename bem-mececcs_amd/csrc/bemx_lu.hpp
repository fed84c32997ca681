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
    // Flags of the cooperative triangular solves hold the epoch of the
    // launch that set them; epochs only grow, so the array is cleared once,
    // when it is (re)allocated, and never inside a solve.
    int epoch = 0;
    void* flags_init = nullptr;
    // second stream and events of the look-ahead factorisation (lazy)
    hipStream_t side = nullptr;
    hipEvent_t ev_panel = nullptr, ev_bulk = nullptr;
    hipError_t ensure_side();
    // graph of the mixed factorisation and the key it was captured for
    bool use_graphs = true;
    hipGraphExec_t gexec = nullptr;
    const void* gkey[8] = {nullptr};
    LuWork();
    void* get(int slot, size_t bytes, hipError_t* e);
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

// Device bytes of the mixed factorisation's workspace for order n beyond the
// FP32 copy of the matrix (bf16 planes, diagonal-block inverses).
size_t mixed_factor_extra_bytes(int n);

}  // namespace bemx
