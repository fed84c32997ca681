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
};

// Gauss point of an element: position and weight OME(ig) OME(jg) DET.
struct GaussPt {
    double x, y, z, w;
};

// Per-element data: outward unit normal and collocation point (centroid).
struct ElemData {
    double nx, ny, nz, cx, cy, cz, pad0, pad1;
};

// Status bits written by kernels (OR-ed into a device word).
enum : int {
    ST_DEGENERATE = 1,  // DET < 1e-5 at some Gauss point (Nonsingd.for:152)
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
