// bemx_common.hpp -- shared device helpers for the MI355X (gfx950) BEM kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bemx {

// Complex number as two doubles (interleaved re, im -- the Fortran COMPLEX*16
// layout).  All arithmetic is written out so the compiler fuses it into FMAs.
struct cd {
    double x, y;
};

__host__ __device__ inline cd mk(double x, double y) { return cd{x, y}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cd operator-(cd a) { return cd{-a.x, -a.y}; }
__host__ __device__ inline cd operator*(cd a, cd b) {
    return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline cd operator*(cd a, double s) { return cd{a.x * s, a.y * s}; }
__host__ __device__ inline cd operator*(double s, cd a) { return cd{a.x * s, a.y * s}; }
__host__ __device__ inline cd cinv(cd a) {
    double d = 1.0 / (a.x * a.x + a.y * a.y);
    return cd{a.x * d, -a.y * d};
}
__host__ __device__ inline cd operator/(cd a, cd b) { return a * cinv(b); }

// Small-|z| form of the dynamic kernel scalars, z = ks R.  Each of psi R,
// chi R, psi' R^2, chi' R^2 (Solfund.for:42-49) is P(z) e^z - r^2 Q(r z)
// e^(r z) with r = cs/cp (real: the damping is the same for both waves) and
// P, Q polynomials in 1/z of degree 2, whose 1/z^2 and 1/z terms cancel
// between the two waves; evaluated as written, FP64 rounding is amplified
// by ~1/|z|^2.  Their power series sum_m f_m z^m (m >= 0) has coefficients
// f_m = a_m - b_m r^(m+2), a_m = sum_j p_j / (m-j)!, b_m = sum_j q_j / (m-j)!
// (p_j, q_j the coefficients of z^j in P and Q, j = -2..1): a_m and b_m are
// pure numbers, so the series need only r at run time.  Below |z| = tau the
// kernels sum the series: off the diagonal tau = 0.35 with 16 terms (the
// closed form loses at most ~1.2e-13 above it, the series truncates below
// 1e-17), in the diagonal blocks tau = 0.5 with 18 terms (3e-20).
constexpr int NSER_OFF = 16, NSER_DIAG = 18;
constexpr double TAU_OFF = 0.35, TAU_DIAG = 0.5;
enum { SER_PSI = 0, SER_CHI, SER_PSID, SER_CHID };
constexpr int SER_P[4][4] = {{1, -1, 1, 0}, {3, -3, 1, 0}, {-3, 3, -2, 1}, {-9, 9, -4, 1}};
constexpr int SER_Q[4][4] = {{1, -1, 0, 0}, {3, -3, 1, 0}, {-3, 3, -1, 0}, {-9, 9, -4, 1}};

constexpr double cfact(int k) {
    double f = 1.0;
    for (int i = 2; i <= k; ++i) f *= i;
    return f;
}

// a_m (q = false) or b_m (q = true) of function f
constexpr double ser_ab(int f, int m, bool q) {
    double s = 0.0;
    for (int j = -2; j <= 1 && j <= m; ++j) s += (q ? SER_Q[f][j + 2] : SER_P[f][j + 2]) / cfact(m - j);
    return s;
}

// Per-frequency constants of the dynamic Kelvin/Stokes kernel
// (Solfund.for:29-40): k = -i w / c for the S and P waves, their inverses
// (1/z = (1/k)(1/R) replaces a complex division per point), c2 = (cs/cp)^2,
// c1 - 2 = (cp/cs)^2 - 2, c0 = 1/(4 pi mu*).
struct DynConst {
    cd ks, kp;     // wave numbers
    cd as, ap;     // 1/ks, 1/kp
    cd as2, ap2;   // 1/ks^2, 1/kp^2
    cd c2;         // (cs/cp)^2
    cd c1m2;       // (cp/cs)^2 - 2
    cd c0;         // 1/(4 pi mu*)
    cd zge;        // mu*
    cd lm;         // lambda*/mu*  (stress kernels)
    double q4pi;   // 1/(4 pi)
    double c1s, c2s, c3s, c4s;  // static constants C1..C4 (Ghmatece.for:60-63)
    double ks_abs2;             // |ks|^2: the series is used where |ks|^2 R^2 < tau^2
    double ser_rinv;            // 1/r, r = cs/cp
    double ser_rtop_off;        // r^(NSER_OFF + 1)
    double ser_rtop_diag;       // r^(NSER_DIAG + 1)
    double ser_c1m2;            // 1/r^2 - 2
    // Diagonal blocks (dynamic minus static kernels, Solfundif.for:56-69)
    // with the series: the m = 0 terms are the static limits, which cancel
    // the static kernels up to these residuals (zero in exact arithmetic for
    // T; O(damping) for U, whose static kernel has the real G):
    cd u_d0, u_r0;              // c0 f_psi,0 - C1 C2, c0 f_chi,0 + C1
    double t_a0, t_b0, t_c0;    // f_A,0 / (4 pi) - C3 C4, f_B,0 / (4 pi) - 3 C3, f_C,0 / (4 pi) + C3 C4
};

// One frequency of a batched sweep assembly: its constants, its exp table
// (2 nt entries, spacing h) and its outputs.
struct DynFreq {
    DynConst K;
    cd* etab;
    int nt;
    double h;
    cd *zh, *rhs_part, *rhs_diag, *zfi;
};

// Gauss point of an element: position and weight OME(ig) OME(jg) DET.
struct GaussPt {
    double x, y, z, w;
};

// Per-element data: outward unit normal and collocation point (centroid).
struct ElemData {
    double nx, ny, nz, cx, cy, cz, pad0, pad1;
};

// One problem of a batched solve on a context (ctx_solve_items): its
// material (bemx_material's zge, zcs, zcp) and frequency, its load cases
// (host, 3 NBE x nrhs interleaved complex; null: the context's), where its
// solution goes (host, may be null) and, out, its statistics [GMRES
// iterations, final ||r||/||b||, 1 if it fell back to FP64, 0].
struct SolveItem {
    double zge[2], zcs[2], zcp[2];
    double fr;
    const double* zdfi;
    double* x;
    double stats[4];
};

// Status bits written by kernels (OR-ed into a device word).
enum : int {
    ST_DEGENERATE = 1,  // DET < 1e-5 at some Gauss point (Nonsingd.for:89-95)
    ST_PIVOT_ZERO = 2,  // exactly singular pivot in the LU
    ST_SYNC_TIMEOUT = 4 // a bounded inter-workgroup wait gave up
};

// exp(z) / R for z = k R, with k = (kx, ky): exp(kx R) (cos(ky R), sin(ky R)) / R
__device__ inline cd exp_over_r(cd k, double R, double invR) {
    double e = exp(k.x * R) * invR;
    double s, c;
    sincos(k.y * R, &s, &c);
    return cd{e * c, e * s};
}

// Wave-wide sum of a double over the 64 lanes in a fixed butterfly order,
// so every lane ends with the same, run-to-run reproducible value.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace bemx
