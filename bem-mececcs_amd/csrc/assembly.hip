// assembly.hip -- H/G influence-matrix assembly kernels for MI355X (gfx950).
//
// FP64 throughout.  Kernels and the reference routines whose results they
// reproduce:
//   k_element_setup  NORMVEC (Normvec.for:4-39), centroids (Inputece.for:100-102)
//                    and the per-element Gauss geometry of Nonsingd.for:45-106
//   k_static_hdiag   GHMATECE off-diagonal loop + RIGID_BODY
//                    (Ghmatece.for:92-150, 258-282): the static H row sums,
//                    without ever storing the static H matrix
//   k_gest_diag      GEST_SINGULAR / SINGGE / NONSINGE_G (Ghmatece.for:288-323,
//                    Singge.for:11-202)
//   k_dyn_diag       SING_DE / SOLFUNDIF + HESTdiag/GESTdiag (Ghmatecd.for:139-153,
//                    Sing_de.for, Solfundif.for)
//   k_dyn_offdiag_h  NONSINGD / SOLFUND (Ghmatecd.for:155-162, Nonsingd.for,
//                    Solfund.for) for the H part of every column
//   k_dyn_offdiag_g  the G part of the KODE-swapped columns (Ghmatecd.for:168,
//                    316-333) and the fused RHS ZFI = ZG ZDFI (Ghmatecd.for:172)
//   k_interior       INTEREC displacements and SIGMAEC stresses
//                    (Interec.for:86-135, 188-283, Sigmaec.for)
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "bemx_common.hpp"

namespace bemx {

// ---------------------------------------------------------------------------
// Element setup: normal, centroid, and the NPG x NPG Gauss points and weights
// of every element, stored twice: element-major gp[j*npg2+q] (wave-uniform
// reads in the assembly) and point-major gpT[q*n+j] (coalesced per-lane reads
// in the interior kernel).  use_given != 0 keeps the caller's normals and
// centroids (the reference driver computes them once, Main.for:120).
// ---------------------------------------------------------------------------
__global__ void k_element_setup(int n, const int* __restrict__ cone, const double* __restrict__ cx,
                                const double* __restrict__ cy, const double* __restrict__ cz,
                                const double* __restrict__ gi, const double* __restrict__ ome,
                                int npg, int use_given, ElemData* __restrict__ elem,
                                GaussPt* __restrict__ gp, GaussPt* __restrict__ gpT,
                                int* __restrict__ status) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double co[4][3];
    for (int k = 0; k < 4; ++k) {
        int p = cone[k * n + j] - 1;
        co[k][0] = cx[p];
        co[k][1] = cy[p];
        co[k][2] = cz[p];
    }
    if (!use_given) {
        double a = (co[1][1] - co[0][1]) * (co[2][2] - co[0][2]) -
                   (co[1][2] - co[0][2]) * (co[2][1] - co[0][1]);
        double b = (co[1][2] - co[0][2]) * (co[2][0] - co[0][0]) -
                   (co[1][0] - co[0][0]) * (co[2][2] - co[0][2]);
        double c = (co[1][0] - co[0][0]) * (co[2][1] - co[0][1]) -
                   (co[1][1] - co[0][1]) * (co[2][0] - co[0][0]);
        double r = sqrt(a * a + b * b + c * c);
        ElemData e;
        e.nx = a / r;
        e.ny = b / r;
        e.nz = c / r;
        e.cx = (co[0][0] + co[1][0] + co[2][0] + co[3][0]) * 0.25;
        e.cy = (co[0][1] + co[1][1] + co[2][1] + co[3][1]) * 0.25;
        e.cz = (co[0][2] + co[1][2] + co[2][2] + co[3][2]) * 0.25;
        e.pad0 = e.pad1 = 0.0;
        elem[j] = e;
    }
    int npg2 = npg * npg;
    int bad = 0;
    const double ecx = elem[j].cx, ecy = elem[j].cy, ecz = elem[j].cz;
    double rad = 0.0;
    for (int jg = 0; jg < npg; ++jg) {
        double eta = gi[jg], sp = 1.0 + eta, sm = 1.0 - eta;
        for (int ig = 0; ig < npg; ++ig) {
            double xi = gi[ig], rp = 1.0 + xi, rm = 1.0 - xi;
            double dxi[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
            double det_[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
            double f[4] = {0.25 * rm * sm, 0.25 * rp * sm, 0.25 * rp * sp, 0.25 * rm * sp};
            double j1[3], j2[3], xg[3];
            for (int d = 0; d < 3; ++d) {
                double s1 = 0.0, s2 = 0.0, s3 = 0.0;
                for (int k = 0; k < 4; ++k) {
                    s1 += dxi[k] * co[k][d];
                    s2 += det_[k] * co[k][d];
                    s3 += co[k][d] * f[k];
                }
                j1[d] = s1;
                j2[d] = s2;
                xg[d] = s3;
            }
            double a = j1[1] * j2[2] - j2[1] * j1[2];
            double b = j2[0] * j1[2] - j1[0] * j2[2];
            double c = j1[0] * j2[1] - j2[0] * j1[1];
            double det = sqrt(a * a + b * b + c * c);
            if (det < 1.0e-5) bad = 1;
            GaussPt g{xg[0], xg[1], xg[2], ome[ig] * ome[jg] * det};
            rad = fmax(rad, sqrt((xg[0] - ecx) * (xg[0] - ecx) + (xg[1] - ecy) * (xg[1] - ecy) +
                                 (xg[2] - ecz) * (xg[2] - ecz)));
            int q = jg * npg + ig;
            gp[(size_t)j * npg2 + q] = g;
            gpT[(size_t)q * n + j] = g;
        }
    }
    elem[j].pad0 = rad;  // largest distance of a Gauss point from the collocation point
    if (bad) atomicOr(status, ST_DEGENERATE);
}

// ---------------------------------------------------------------------------
// Static H rigid-body diagonal.  For collocation row I (one per lane) and the
// elements J of this split: T^st summed over J != I and its Gauss points, in
// the factored form  T_ij = C3 [C4 d_ij S0 + 3 S2_ij + C4 (A_ij - A_ji)]  with
// S0 = sum w drn/R^2, S2 = sum w drn r_i r_j/R^2 and A_ij = sum_J n_i(J) S1_j(J),
// S1 = sum w r_j/R^2  (Solfune.for:27-31).  part[s][I][9] gets this split's
// sum; k_static_hdiag_reduce negates and sums the splits in a fixed order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_static_hdiag(int n, int nbe, int npg2, int jper,
                                                     const ElemData* __restrict__ elem,
                                                     const GaussPt* __restrict__ gp, double c3,
                                                     double c4, double* __restrict__ part) {
    int I = blockIdx.x * 64 + threadIdx.x;
    int Ic = I < nbe ? I : nbe - 1;
    int j0 = blockIdx.y * jper;
    int j1 = min(n, j0 + jper);
    double xc = elem[Ic].cx, yc = elem[Ic].cy, zc = elem[Ic].cz;
    double s0 = 0, s2[6] = {0, 0, 0, 0, 0, 0}, A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int J = j0; J < j1; ++J) {
        double nx = elem[J].nx, ny = elem[J].ny, nz = elem[J].nz;
        const GaussPt* g = gp + (size_t)J * npg2;
        double t0 = 0, t2[6] = {0, 0, 0, 0, 0, 0}, t1[3] = {0, 0, 0};
        for (int q = 0; q < npg2; ++q) {
            GaussPt p = g[q];
            double r1 = p.x - xc, r2 = p.y - yc, r3 = p.z - zc;
            double R = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
            double iR = 1.0 / R;
            double d1 = r1 * iR, d2 = r2 * iR, d3 = r3 * iR;
            double wr = p.w * iR * iR;
            double drn = d1 * nx + d2 * ny + d3 * nz;
            double wd = wr * drn;
            t0 += wd;
            t2[0] += wd * d1 * d1;
            t2[1] += wd * d2 * d1;
            t2[2] += wd * d3 * d1;
            t2[3] += wd * d2 * d2;
            t2[4] += wd * d3 * d2;
            t2[5] += wd * d3 * d3;
            t1[0] += wr * d1;
            t1[1] += wr * d2;
            t1[2] += wr * d3;
        }
        if (J != I) {
            s0 += t0;
            for (int k = 0; k < 6; ++k) s2[k] += t2[k];
            double nn[3] = {nx, ny, nz};
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a) A[a + 3 * b] += nn[a] * t1[b];
        }
    }
    if (I >= nbe) return;
    const int sidx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};  // (i,j) column-major -> packed sym
    double* out = part + ((size_t)blockIdx.y * nbe + I) * 9;
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) {
            double v = 3.0 * s2[sidx[a + 3 * b]] + c4 * (A[a + 3 * b] - A[b + 3 * a]);
            if (a == b) v += c4 * s0;
            out[a + 3 * b] = c3 * v;
        }
}

__global__ void k_static_hdiag_reduce(int nbe, int nsplit, const double* __restrict__ part,
                                      double* __restrict__ hdiag) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbe * 9) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * nbe * 9 + t];
    hdiag[t] = -s;
}

// ---------------------------------------------------------------------------
// Static G diagonal: four quads (P_a, P_b, xm, xm) collapsed onto the
// centroid, NPG x NPG Gauss each, Kelvin U^st = (C1/R)(C2 d_ij + r_i r_j).
// One wavefront per element; lanes stride over the 4 NPG^2 points.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_gest_diag(int n, int nbe, const int* __restrict__ cone,
                                                  const double* __restrict__ cx,
                                                  const double* __restrict__ cy,
                                                  const double* __restrict__ cz,
                                                  const ElemData* __restrict__ elem,
                                                  const double* __restrict__ gi,
                                                  const double* __restrict__ ome, int npg,
                                                  double c1, double c2,
                                                  double* __restrict__ gdiag,
                                                  int* __restrict__ status) {
    int I = blockIdx.x;
    int lane = threadIdx.x;
    double P[4][3];
    for (int k = 0; k < 4; ++k) {
        int p = cone[k * n + I] - 1;
        P[k][0] = cx[p];
        P[k][1] = cy[p];
        P[k][2] = cz[p];
    }
    double xm[3] = {elem[I].cx, elem[I].cy, elem[I].cz};
    int npg2 = npg * npg;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int bad = 0;
    for (int t = lane; t < 4 * npg2; t += 64) {
        int it = t / npg2, q = t - it * npg2;
        int jg = q / npg, ig = q - jg * npg;
        const double* pa = P[it];
        const double* pb = P[(it + 1) & 3];
        double eta = gi[jg], sp = 1.0 + eta, sm = 1.0 - eta;
        double xi = gi[ig], rp = 1.0 + xi, rm = 1.0 - xi;
        double dxi[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
        double det_[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
        double f[4] = {0.25 * rm * sm, 0.25 * rp * sm, 0.25 * rp * sp, 0.25 * rm * sp};
        double j1[3], j2[3], xg[3];
        for (int d = 0; d < 3; ++d) {
            double co[4] = {pa[d], pb[d], xm[d], xm[d]};
            double s1 = 0, s2 = 0, s3 = 0;
            for (int k = 0; k < 4; ++k) {
                s1 += dxi[k] * co[k];
                s2 += det_[k] * co[k];
                s3 += co[k] * f[k];
            }
            j1[d] = s1;
            j2[d] = s2;
            xg[d] = s3;
        }
        double a = j1[1] * j2[2] - j2[1] * j1[2];
        double b = j2[0] * j1[2] - j1[0] * j2[2];
        double c = j1[0] * j2[1] - j2[0] * j1[1];
        double det = sqrt(a * a + b * b + c * c);
        if (det < 1.0e-5) bad = 1;
        double w = ome[ig] * ome[jg] * det;
        double r1 = xg[0] - xm[0], r2 = xg[1] - xm[1], r3 = xg[2] - xm[2];
        double R = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
        double iR = 1.0 / R;
        double d1 = r1 * iR, d2 = r2 * iR, d3 = r3 * iR;
        double wc = w * c1 * iR;
        acc[0] += wc * (c2 + d1 * d1);
        acc[1] += wc * (d2 * d1);
        acc[2] += wc * (d3 * d1);
        acc[3] += wc * (c2 + d2 * d2);
        acc[4] += wc * (d3 * d2);
        acc[5] += wc * (c2 + d3 * d3);
    }
    for (int k = 0; k < 6; ++k) acc[k] = wave_sum(acc[k]);
    if (bad) atomicOr(status, ST_DEGENERATE);
    if (lane == 0) {
        double* o = gdiag + (size_t)I * 9;
        o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2];
        o[3] = acc[1]; o[4] = acc[3]; o[5] = acc[4];
        o[6] = acc[2]; o[7] = acc[4]; o[8] = acc[5];
    }
}

// ---------------------------------------------------------------------------
// Dynamic scalar functions at distance R (Solfund.for:29-53):
//   chi, and with WithPsi also psi; and A = psi' - chi/R, B = 4 chi/R - 2 chi',
//   C = (c1-2)(A + B/2 - 3 chi/R) - 2 chi/R.
// 1/z is formed as (1/k)(1/R).
// ---------------------------------------------------------------------------
// Es = exp(ks R) / R and Epc = c2 exp(kp R) / R are given.
template <bool WithPsi>
__device__ inline void dyn_scalars_e(const DynConst& K, double R, double iR, cd Es, cd Epc,
                                     cd& psi, cd& chi, cd& A, cd& B, cd& C) {
    double iR2 = iR * iR;
    cd as = K.as * iR, as2 = K.as2 * iR2;
    cd ap = K.ap * iR, ap2 = K.ap2 * iR2;
    cd zs = K.ks * R, zp = K.kp * R;
    // chi = (1 + 3/zs^2 - 3/zs) Es - c2 (1 + 3/zp^2 - 3/zp) Ep
    cd fs = mk(1.0 + 3.0 * (as2.x - as.x), 3.0 * (as2.y - as.y));
    cd fp = mk(1.0 + 3.0 * (ap2.x - ap.x), 3.0 * (ap2.y - ap.y));
    chi = fs * Es - fp * Epc;
    if (WithPsi) {
        cd gs = mk(1.0 + as2.x - as.x, as2.y - as.y);
        cd gp = ap2 - ap;
        psi = gs * Es - gp * Epc;
    }
    // psi' R = (-2 + zs + 3/zs - 3/zs^2) Es - c2 (-1 + 3/zp - 3/zp^2) Ep
    cd hs = mk(-2.0 + zs.x + 3.0 * (as.x - as2.x), zs.y + 3.0 * (as.y - as2.y));
    cd hp = mk(-1.0 + 3.0 * (ap.x - ap2.x), 3.0 * (ap.y - ap2.y));
    cd psid = (hs * Es - hp * Epc) * iR;
    // chi' R = (zs - 4 + 9/zs - 9/zs^2) Es - c2 (zp - 4 + 9/zp - 9/zp^2) Ep
    cd ks_ = mk(zs.x - 4.0 + 9.0 * (as.x - as2.x), zs.y + 9.0 * (as.y - as2.y));
    cd kp_ = mk(zp.x - 4.0 + 9.0 * (ap.x - ap2.x), zp.y + 9.0 * (ap.y - ap2.y));
    cd chid = (ks_ * Es - kp_ * Epc) * iR;
    cd chr = chi * iR;
    A = psid - chr;
    B = 4.0 * chr - 2.0 * chid;
    C = K.c1m2 * (A + 0.5 * B - 3.0 * chr) - 2.0 * chr;
}

// t z + c for complex t, z and real c.
__device__ inline cd mad_re(cd t, cd z, double c) {
    return cd{fma(t.x, z.x, fma(-t.y, z.y, c)), fma(t.x, z.y, t.y * z.x)};
}

// Horner steps M, M-1, .., M0 of the series of chi R, psi' R^2, chi' R^2
// (and psi R with Psi): t = t z + f_m, f_m = a_m - b_m rho, rho = r^(m+2)
// on entry and divided by r between steps.  a_m, b_m are compile-time
// constants.
template <int M, int M0, bool Psi>
__device__ inline void ser_steps(cd z, double rho, double rinv, cd& tc, cd& tpd, cd& tcd,
                                 cd& tp) {
    constexpr double ac = ser_ab(SER_CHI, M, false), bc = ser_ab(SER_CHI, M, true);
    constexpr double apd = ser_ab(SER_PSID, M, false), bpd = ser_ab(SER_PSID, M, true);
    constexpr double acd = ser_ab(SER_CHID, M, false), bcd = ser_ab(SER_CHID, M, true);
    tc = mad_re(tc, z, fma(-bc, rho, ac));
    tpd = mad_re(tpd, z, fma(-bpd, rho, apd));
    tcd = mad_re(tcd, z, fma(-bcd, rho, acd));
    if constexpr (Psi) {
        constexpr double ap = ser_ab(SER_PSI, M, false), bp = ser_ab(SER_PSI, M, true);
        tp = mad_re(tp, z, fma(-bp, rho, ap));
    }
    if constexpr (M > M0) ser_steps<M - 1, M0, Psi>(z, rho * rinv, rinv, tc, tpd, tcd, tp);
}

// chi R, psi' R^2, chi' R^2 (and psi R) summed over m = M0..N-1 of their
// series in z = ks R; rtop = r^(N+1).  With M0 = 1 the static limits (the
// m = 0 terms) are left out.
template <int N, int M0, bool Psi>
__device__ inline void ser_sums(const DynConst& K, cd z, double rtop, cd& chiR, cd& psidR2,
                                cd& chidR2, cd& psiR) {
    cd tc = mk(0.0, 0.0), tpd = tc, tcd = tc, tp = tc;
    ser_steps<N - 1, M0, Psi>(z, rtop, K.ser_rinv, tc, tpd, tcd, tp);
    if constexpr (M0 == 1) {
        tc = tc * z;
        tpd = tpd * z;
        tcd = tcd * z;
        if constexpr (Psi) tp = tp * z;
    }
    chiR = tc;
    psidR2 = tpd;
    chidR2 = tcd;
    psiR = tp;
}

// A R^2, B R^2, C R^2 from chi R, psi' R^2, chi' R^2 (Solfund.for:51-53
// multiplied through by R^2; A + B/2 - 3 chi = psi' - chi' - 2 chi).
__device__ inline void abc_from(const DynConst& K, cd chiR, cd psidR2, cd chidR2, cd& AR2,
                               cd& BR2, cd& CR2) {
    AR2 = psidR2 - chiR;
    BR2 = 4.0 * chiR - 2.0 * chidR2;
    CR2 = K.ser_c1m2 * (psidR2 - chidR2 - 2.0 * chiR) - 2.0 * chiR;
}

template <bool WithPsi>
__device__ inline void dyn_scalars(const DynConst& K, double R, double iR, cd& psi, cd& chi,
                                   cd& A, cd& B, cd& C) {
    if (K.ks_abs2 * (R * R) < TAU_OFF * TAU_OFF) {
        const cd z = K.ks * R;
        const double iR2 = iR * iR;
        cd psiR, chiR, psidR2, chidR2, AR2, BR2, CR2;
        ser_sums<NSER_OFF, 0, WithPsi>(K, z, K.ser_rtop_off, chiR, psidR2, chidR2, psiR);
        abc_from(K, chiR, psidR2, chidR2, AR2, BR2, CR2);
        psi = psiR * iR;
        chi = chiR * iR;
        A = AR2 * iR2;
        B = BR2 * iR2;
        C = CR2 * iR2;
        return;
    }
    dyn_scalars_e<WithPsi>(K, R, iR, exp_over_r(K.ks, R, iR), K.c2 * exp_over_r(K.kp, R, iR),
                           psi, chi, A, B, C);
}

// Squared distance with explicit FMAs: the near-field split of the H
// assembly (series below |ks| R = TAU_OFF, closed form above) must classify
// every Gauss point the same way in both kernels.
__device__ inline double dist2(double r1, double r2, double r3) {
    return fma(r3, r3, fma(r2, r2, r1 * r1));
}

__device__ inline bool near_field(const DynConst& K, double RR) {
    return K.ks_abs2 * RR < TAU_OFF * TAU_OFF;
}

// ---------------------------------------------------------------------------
// exp(k R) for the two wave numbers without an exp and a sincos per Gauss
// point: exp(k R) = exp(k x_i) exp(k d), x_i = i h the nearest point of a
// grid whose spacing h is a power of two (so i = rint(R / h), x_i and
// d = R - x_i are exact), exp(k x_i) from a table built once per frequency,
// and exp(k d) = exp(u) (cos v + i sin v), u = Re(k) d, v = Im(k) d, from
// Taylor polynomials with the constant coefficients 1/n! (exp to degree 7,
// cos to 6, sin to 7).  h is at most 1 / (16 max|k|), so |u|, |v| <= 1/32
// and each truncation error is below (1/32)^8 / 8! ~ 2e-17.
// ---------------------------------------------------------------------------
constexpr int EXP_TAB_MAX = 1 << 18;  // entries per wave number

struct ExpTab {
    const cd* ts;  // exp(ks x_i)
    const cd* tp;  // c2 exp(kp x_i)
    double h, ih;
    int nt;
};

// exp(k x) with the rounding error of each product k x carried to first
// order, so a table entry is as accurate as exp and sincos themselves.
__device__ inline cd exp_tab_entry(cd k, double x) {
    const double ax = k.x * x, ex = fma(k.x, x, -ax);
    const double ay = k.y * x, ey = fma(k.y, x, -ay);
    const double e = exp(ax);
    const double m = fma(e, ex, e);
    double s, c;
    sincos(ay, &s, &c);
    return cd{m * fma(-ey, s, c), m * fma(ey, c, s)};
}

__global__ void k_exp_table(int nt, double h, cd ks, cd kp, cd c2, cd* __restrict__ ts,
                            cd* __restrict__ tp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const double x = (double)i * h;
    ts[i] = exp_tab_entry(ks, x);
    tp[i] = c2 * exp_tab_entry(kp, x);
}

// exp(k d) for |Re(k) d|, |Im(k) d| <= 1/32.
__device__ inline cd exp_small(cd k, double d) {
    const double u = k.x * d, v = k.y * d, v2 = v * v;
    double e = fma(u, 1.0 / 5040, 1.0 / 720);
    e = fma(e, u, 1.0 / 120);
    e = fma(e, u, 1.0 / 24);
    e = fma(e, u, 1.0 / 6);
    e = fma(e, u, 0.5);
    e = fma(e, u, 1.0);
    e = fma(e, u, 1.0);
    double c = fma(v2, -1.0 / 720, 1.0 / 24);
    c = fma(c, v2, -0.5);
    c = fma(c, v2, 1.0);
    double sn = fma(v2, -1.0 / 5040, 1.0 / 120);
    sn = fma(sn, v2, -1.0 / 6);
    sn = fma(sn * v2, v, v);
    return cd{e * c, e * sn};
}

// Es = exp(ks R) and Epc = c2 exp(kp R) from the tables.
__device__ inline void exp_pair_tab(const DynConst& K, const ExpTab& T, double R, cd& Es,
                                    cd& Epc) {
    const double fi = __builtin_rint(R * T.ih);
    const double d = fma(-fi, T.h, R);
    const int i = min(max((int)fi, 0), T.nt - 1);
    const cd ts = T.ts[i], tp = T.tp[i];
    const cd es = exp_small(K.ks, d), ep = exp_small(K.kp, d);
    Es = cd{fma(ts.x, es.x, -ts.y * es.y), fma(ts.x, es.y, ts.y * es.x)};
    Epc = cd{fma(tp.x, ep.x, -tp.y * ep.y), fma(tp.x, ep.y, tp.y * ep.x)};
}

// KODE column swap of one 3x3 block (Ghmatecd.for:316-333 as called at
// :168): for a KODE=0 column c, H'(:,c) = -G(:,c) and G'(:,c) = -mu* H(:,c).
__device__ inline void swap_cols(const cd H[9], const cd G[9], int flag, cd zge, cd Hout[9],
                                 cd Gout[9]) {
    for (int k = 0; k < 3; ++k) {
        bool sw = (flag >> (1 + k)) & 1;
        for (int i = 0; i < 3; ++i) {
            cd h = H[i + 3 * k], g = G[i + 3 * k];
            Hout[i + 3 * k] = sw ? -g : h;
            Gout[i + 3 * k] = sw ? -(h * zge) : g;
        }
    }
}

constexpr int MAXRHS = 4;  // load cases per assembly

// ---------------------------------------------------------------------------
// Dynamic diagonal blocks: plain Gauss of (dynamic - static) kernels over the
// element itself, collocation at its centroid, plus HESTdiag/GESTdiag.
// One wavefront per element, one lane per Gauss point.  Writes the swapped
// block and, per load case c, G'_II ZDFI_c(I) into rhs_diag[c].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_dyn_diag(int nbe, int npg2, const ElemData* __restrict__ elem,
                                                 const GaussPt* __restrict__ gp,
                                                 const double* __restrict__ hdiag,
                                                 const double* __restrict__ gdiag,
                                                 const int* __restrict__ colflag,
                                                 const cd* __restrict__ zdfi, int nrhs,
                                                 size_t ldr, DynConst K, cd* __restrict__ zh,
                                                 size_t ldzh, cd* __restrict__ rhs_diag) {
    int I = blockIdx.x;
    int lane = threadIdx.x;
    ElemData e = elem[I];
    double nv[3] = {e.nx, e.ny, e.nz};
    double accH[18], accG[18];
    for (int k = 0; k < 18; ++k) accH[k] = accG[k] = 0.0;
    for (int q = lane; q < npg2; q += 64) {
        GaussPt p = gp[(size_t)I * npg2 + q];
        double r[3] = {p.x - e.cx, p.y - e.cy, p.z - e.cz};
        double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        double iR = 1.0 / R;
        double rd[3] = {r[0] * iR, r[1] * iR, r[2] * iR};
        double drn = rd[0] * nv[0] + rd[1] * nv[1] + rd[2] * nv[2];
        // dynamic minus static kernels (Solfundif.for:56-69) as
        //   U = ud d_ij - ur r_i r_j,
        //   T = ta (drn d_ij + r_j n_i) + tb drn r_i r_j + tc r_i n_j:
        // below SER_TAU from the series without its static limits plus the
        // residuals of those limits, above it from the closed form minus
        // the static kernels
        cd ud, ur, ta, tb, tc;
        const double iR2 = iR * iR;
        if (K.ks_abs2 * (R * R) < TAU_DIAG * TAU_DIAG) {
            const cd z = K.ks * R;
            cd psiR, chiR, psidR2, chidR2, AR2, BR2, CR2;
            ser_sums<NSER_DIAG, 1, true>(K, z, K.ser_rtop_diag, chiR, psidR2, chidR2, psiR);
            abc_from(K, chiR, psidR2, chidR2, AR2, BR2, CR2);
            ud = (K.u_d0 + K.c0 * psiR) * iR;
            ur = (K.u_r0 + K.c0 * chiR) * iR;
            ta = (mk(K.t_a0, 0.0) + AR2 * K.q4pi) * iR2;
            tb = (mk(K.t_b0, 0.0) + BR2 * K.q4pi) * iR2;
            tc = (mk(K.t_c0, 0.0) + CR2 * K.q4pi) * iR2;
        } else {
            cd psi, chi, A, B, C;
            dyn_scalars_e<true>(K, R, iR, exp_over_r(K.ks, R, iR),
                                K.c2 * exp_over_r(K.kp, R, iR), psi, chi, A, B, C);
            const double ust = K.c1s * iR, tst = K.c3s * iR2;
            ud = K.c0 * psi - mk(ust * K.c2s, 0.0);
            ur = K.c0 * chi + mk(ust, 0.0);
            ta = A * K.q4pi - mk(tst * K.c4s, 0.0);
            tb = B * K.q4pi - mk(3.0 * tst, 0.0);
            tc = C * K.q4pi + mk(tst * K.c4s, 0.0);
        }
        double w = p.w;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                double dij = (i == j) ? 1.0 : 0.0;
                cd u = ud * dij - ur * (rd[i] * rd[j]);
                cd t = ta * (drn * dij + rd[j] * nv[i]) + tb * (rd[i] * rd[j] * drn) +
                       tc * (rd[i] * nv[j]);
                int k = i + 3 * j;
                accH[2 * k] += t.x * w;
                accH[2 * k + 1] += t.y * w;
                accG[2 * k] += u.x * w;
                accG[2 * k + 1] += u.y * w;
            }
    }
    for (int k = 0; k < 18; ++k) {
        accH[k] = wave_sum(accH[k]);
        accG[k] = wave_sum(accG[k]);
    }
    if (lane != 0) return;
    cd H[9], G[9], Ho[9], Gn[9];
    for (int k = 0; k < 9; ++k) {
        H[k] = mk(accH[2 * k] + hdiag[(size_t)I * 9 + k], accH[2 * k + 1]);
        G[k] = mk(accG[2 * k] + gdiag[(size_t)I * 9 + k], accG[2 * k + 1]);
    }
    swap_cols(H, G, colflag[I], K.zge, Ho, Gn);
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i) zh[(size_t)(3 * I + k) * ldzh + 3 * I + i] = Ho[i + 3 * k];
    for (int c = 0; c < nrhs; ++c) {
        const cd* t = zdfi + c * ldr + 3 * I;
        for (int i = 0; i < 3; ++i)
            rhs_diag[c * ldr + 3 * I + i] = Gn[i] * t[0] + Gn[i + 3] * t[1] + Gn[i + 6] * t[2];
    }
}

// 1/sqrt(x) to FP64 accuracy: v_rsq_f64 refined by two Newton steps.
__device__ inline double rsqrt_refined(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = x * y;
    y = fma(y * 0.5, fma(-h, y, 1.0), y);
    h = x * y;
    y = fma(y * 0.5, fma(-h, y, 1.0), y);
    return y;
}

// ---------------------------------------------------------------------------
// Off-diagonal dynamic blocks.  A workgroup of WAVES wavefronts covers
// 64*WAVES collocation rows (one per lane) and a tile of JT column elements
// whose Gauss points are staged in LDS and read by every lane at the same
// address (broadcast).  Per element J the lane accumulates H in the factored
// form  H_ij = d_ij SD + n_i SP_j + SRQ_ij  (SD = sum A' drn, SP = sum A' r_j,
// SRQ = sum r_i (B' drn r_j + C' n_j), primes = times w/(4 pi)), which is
// the Gauss sum of Solfund.for:58-59 with n fixed per element.  With NeedG
// it also integrates U (psi, chi) into G.
// ---------------------------------------------------------------------------
template <bool NeedG>
__device__ inline void offdiag_element(const DynConst& K, const GaussPt* __restrict__ sg, int npg2,
                                       double xc, double yc, double zc, const double nv[3],
                                       cd H[9], cd G[9]) {
    cd SD = mk(0, 0), SP[3], SRQ[9], Gpsi = mk(0, 0), Gchi[6];
    for (int k = 0; k < 3; ++k) SP[k] = mk(0, 0);
    for (int k = 0; k < 9; ++k) SRQ[k] = mk(0, 0);
    for (int k = 0; k < 6; ++k) Gchi[k] = mk(0, 0);
    for (int q = 0; q < npg2; ++q) {
        GaussPt p = sg[q];
        double r1 = p.x - xc, r2 = p.y - yc, r3 = p.z - zc;
        double iR = rsqrt_refined(r1 * r1 + r2 * r2 + r3 * r3);
        double R = (r1 * r1 + r2 * r2 + r3 * r3) * iR;
        double rd[3] = {r1 * iR, r2 * iR, r3 * iR};
        double drn = rd[0] * nv[0] + rd[1] * nv[1] + rd[2] * nv[2];
        cd psi, chi, A, B, C;
        dyn_scalars<NeedG>(K, R, iR, psi, chi, A, B, C);
        double wq = p.w * K.q4pi;
        cd Ap = A * wq, Bp = B * (wq * drn), Cp = C * wq;
        SD = SD + Ap * drn;
        cd Q[3];
        for (int j = 0; j < 3; ++j) {
            SP[j] = SP[j] + Ap * rd[j];
            Q[j] = Bp * rd[j] + Cp * nv[j];
        }
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) SRQ[i + 3 * j] = SRQ[i + 3 * j] + Q[j] * rd[i];
        if (NeedG) {
            Gpsi = Gpsi + psi * p.w;
            cd cw = chi * p.w;
            Gchi[0] = Gchi[0] + cw * (rd[0] * rd[0]);
            Gchi[1] = Gchi[1] + cw * (rd[1] * rd[0]);
            Gchi[2] = Gchi[2] + cw * (rd[2] * rd[0]);
            Gchi[3] = Gchi[3] + cw * (rd[1] * rd[1]);
            Gchi[4] = Gchi[4] + cw * (rd[2] * rd[1]);
            Gchi[5] = Gchi[5] + cw * (rd[2] * rd[2]);
        }
    }
    const int sidx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            cd h = SRQ[i + 3 * j] + SP[j] * nv[i];
            if (i == j) h = h + SD;
            H[i + 3 * j] = h;
            if (NeedG) {
                cd g = -Gchi[sidx[i + 3 * j]];
                if (i == j) g = g + Gpsi;
                G[i + 3 * j] = K.c0 * g;
            }
        }
}

// H of one element with the exp tables.  The exponentials come without the
// factor 1/R, so the scalars below are chi R, psi' R^2, chi' R^2 and
// A R^2, B R^2, C R^2 (the relations of dyn_scalars_e multiplied through),
// and 1/R^2 joins the quadrature weight.  SRQ_ij = sum B' drn r_i r_j +
// (sum C' r_i) n_j: the first sum is symmetric (6 entries), the second a
// vector.  The diagonal term sum A' drn is SP . n (n is fixed per element).
// The staged weights already hold the factor 1/(4 pi).
// Factored H sums of one element: SP = sum A' r, SC = sum C' r, SB = the
// symmetric sum B' drn r r^T (6 entries).
struct HSums {
    cd SP[3], SC[3], SB[6];
    __device__ void zero() {
        for (int k = 0; k < 3; ++k) SP[k] = SC[k] = mk(0, 0);
        for (int k = 0; k < 6; ++k) SB[k] = mk(0, 0);
    }
    // H_ij = delta_ij sum A' drn + n_i sum A' r_j + sum r_i (B' drn r_j + C' n_j),
    // with sum A' drn = SP . n
    __device__ void to_block(const double nv[3], cd H[9]) const {
        const cd SD = SP[0] * nv[0] + SP[1] * nv[1] + SP[2] * nv[2];
        const int sidx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                cd h = SB[sidx[i + 3 * j]] + SC[i] * nv[j] + SP[j] * nv[i];
                if (i == j) h = h + SD;
                H[i + 3 * j] = h;
            }
    }
};

// One Gauss point's contribution to the H sums of one frequency (table
// route), from the point's frequency-independent geometry: R, 1/R, 1/R^2,
// the unit vector rd, drn and the weight w (with 1/(4 pi)).  The one- and
// two-frequency kernels both call it, so each frequency's arithmetic is the
// same in both.
__device__ __forceinline__ void h_point_tab(const DynConst& K, const ExpTab& T, double R,
                                            double iR, double iR2, const double rd[3], double drn,
                                            double w, HSums& S) {
    cd A, B, C;
    {
        cd Es, Epc;
        exp_pair_tab(K, T, R, Es, Epc);
        cd as = K.as * iR, as2 = K.as2 * iR2;
        cd ap = K.ap * iR, ap2 = K.ap2 * iR2;
        cd zs = K.ks * R;
        cd fs = mk(1.0 + 3.0 * (as2.x - as.x), 3.0 * (as2.y - as.y));
        cd fp = mk(1.0 + 3.0 * (ap2.x - ap.x), 3.0 * (ap2.y - ap.y));
        cd chiR = fs * Es - fp * Epc;
        cd hs = mk(-2.0 + zs.x + 3.0 * (as.x - as2.x), zs.y + 3.0 * (as.y - as2.y));
        cd hp = mk(-1.0 + 3.0 * (ap.x - ap2.x), 3.0 * (ap.y - ap2.y));
        cd psidR2 = hs * Es - hp * Epc;
        cd zp = K.kp * R;
        cd ks_ = mk(zs.x - 4.0 + 9.0 * (as.x - as2.x), zs.y + 9.0 * (as.y - as2.y));
        cd kp_ = mk(zp.x - 4.0 + 9.0 * (ap.x - ap2.x), zp.y + 9.0 * (ap.y - ap2.y));
        cd chidR2 = ks_ * Es - kp_ * Epc;
        A = psidR2 - chiR;
        B = 4.0 * chiR - 2.0 * chidR2;
        C = K.c1m2 * (A + 0.5 * B - 3.0 * chiR) - 2.0 * chiR;
    }
    double wq = w * iR2;
    cd Ap = A * wq, Bp = B * (wq * drn), Cp = C * wq;
    for (int j = 0; j < 3; ++j) {
        S.SP[j] = S.SP[j] + Ap * rd[j];
        S.SC[j] = S.SC[j] + Cp * rd[j];
    }
    S.SB[0] = S.SB[0] + Bp * (rd[0] * rd[0]);
    S.SB[1] = S.SB[1] + Bp * (rd[1] * rd[0]);
    S.SB[2] = S.SB[2] + Bp * (rd[2] * rd[0]);
    S.SB[3] = S.SB[3] + Bp * (rd[1] * rd[1]);
    S.SB[4] = S.SB[4] + Bp * (rd[2] * rd[1]);
    S.SB[5] = S.SB[5] + Bp * (rd[2] * rd[2]);
}

// H blocks of NF frequencies for one element (table route): the geometry of
// each Gauss point once, then each frequency's kernel scalars.  Points with
// |ks| R < TAU_OFF are left to k_dyn_offdiag_near (decided per frequency).
template <int NF>
__device__ inline void offdiag_element_h_tab(const DynConst* K, const ExpTab* T,
                                             const GaussPt* __restrict__ sg, int npg2, double xc,
                                             double yc, double zc, const double nv[3],
                                             cd (*H)[9]) {
    HSums S[NF];
    for (int f = 0; f < NF; ++f) S[f].zero();
    for (int q = 0; q < npg2; ++q) {
        GaussPt p = sg[q];
        double r1 = p.x - xc, r2 = p.y - yc, r3 = p.z - zc;
        const double RR = dist2(r1, r2, r3);
        bool any = false;
        for (int f = 0; f < NF; ++f) any = any || !near_field(K[f], RR);
        if (!any) continue;
        double iR = rsqrt_refined(RR);
        double R = RR * iR;
        double rd[3] = {r1 * iR, r2 * iR, r3 * iR};
        double drn = rd[0] * nv[0] + rd[1] * nv[1] + rd[2] * nv[2];
        double iR2 = iR * iR;
        for (int f = 0; f < NF; ++f)
            if (!near_field(K[f], RR)) h_point_tab(K[f], T[f], R, iR, iR2, rd, drn, p.w, S[f]);
    }
    for (int f = 0; f < NF; ++f) S[f].to_block(nv, H[f]);
}

// H of every off-diagonal block; columns with KODE = 0 are left to
// k_dyn_offdiag_g, which writes their swapped content.
template <int WAVES, int JT, bool TAB>
__global__ __launch_bounds__(64 * WAVES) void k_dyn_offdiag_h(int nbe, int npg2,
                                                              const ElemData* __restrict__ elem,
                                                              const GaussPt* __restrict__ gp,
                                                              const int* __restrict__ colflag,
                                                              DynConst K, ExpTab T,
                                                              cd* __restrict__ zh, size_t ldzh) {
    extern __shared__ GaussPt sgp[];
    const int j0 = blockIdx.y * JT;
    const int jn = min(JT, nbe - j0);
    for (int t = threadIdx.x; t < jn * npg2; t += 64 * WAVES) {
        GaussPt g = gp[(size_t)j0 * npg2 + t];
        if (TAB) g.w *= K.q4pi;
        sgp[t] = g;
    }
    __syncthreads();
    const int I = blockIdx.x * (64 * WAVES) + threadIdx.x;
    const bool rowok = I < nbe;
    const int Ic = rowok ? I : nbe - 1;
    const double xc = elem[Ic].cx, yc = elem[Ic].cy, zc = elem[Ic].cz;
    for (int jj = 0; jj < jn; ++jj) {
        const int J = j0 + jj;
        const int flag = colflag[J];
        const double nv[3] = {elem[J].nx, elem[J].ny, elem[J].nz};
        cd H[1][9], G[9];
        if constexpr (TAB)
            offdiag_element_h_tab<1>(&K, &T, sgp + jj * npg2, npg2, xc, yc, zc, nv, H);
        else
            offdiag_element<false>(K, sgp + jj * npg2, npg2, xc, yc, zc, nv, H[0], G);
        if (rowok && I != J) {
            for (int k = 0; k < 3; ++k) {
                if ((flag >> (1 + k)) & 1) continue;
                cd* col = zh + (size_t)(3 * J + k) * ldzh + 3 * I;
                col[0] = H[0][3 * k];
                col[1] = H[0][3 * k + 1];
                col[2] = H[0][3 * k + 2];
            }
        }
    }
}

// Two frequencies of a sweep in one launch (table route): the same tiles,
// the Gauss-point geometry shared, each frequency's H into its own ZH.
struct DynPair {
    DynConst K[2];
    ExpTab T[2];
    cd* zh[2];
};
template <int WAVES, int JT>
__global__ __launch_bounds__(64 * WAVES) void k_dyn_offdiag_h2(int nbe, int npg2,
                                                               const ElemData* __restrict__ elem,
                                                               const GaussPt* __restrict__ gp,
                                                               const int* __restrict__ colflag,
                                                               DynPair P, size_t ldzh) {
    extern __shared__ GaussPt sgp[];
    const int j0 = blockIdx.y * JT;
    const int jn = min(JT, nbe - j0);
    for (int t = threadIdx.x; t < jn * npg2; t += 64 * WAVES) {
        GaussPt g = gp[(size_t)j0 * npg2 + t];
        g.w *= P.K[0].q4pi;
        sgp[t] = g;
    }
    __syncthreads();
    const int I = blockIdx.x * (64 * WAVES) + threadIdx.x;
    const bool rowok = I < nbe;
    const int Ic = rowok ? I : nbe - 1;
    const double xc = elem[Ic].cx, yc = elem[Ic].cy, zc = elem[Ic].cz;
    for (int jj = 0; jj < jn; ++jj) {
        const int J = j0 + jj;
        const int flag = colflag[J];
        const double nv[3] = {elem[J].nx, elem[J].ny, elem[J].nz};
        cd H[2][9];
        offdiag_element_h_tab<2>(P.K, P.T, sgp + jj * npg2, npg2, xc, yc, zc, nv, H);
        if (rowok && I != J) {
            for (int f = 0; f < 2; ++f)
                for (int k = 0; k < 3; ++k) {
                    if ((flag >> (1 + k)) & 1) continue;
                    cd* col = P.zh[f] + (size_t)(3 * J + k) * ldzh + 3 * I;
                    col[0] = H[f][3 * k];
                    col[1] = H[f][3 * k + 1];
                    col[2] = H[f][3 * k + 2];
                }
        }
    }
}

// Near-field part of the H blocks that k_dyn_offdiag_h leaves out: the
// Gauss points with |ks| R < TAU_OFF, where the closed form of the kernel
// scalars loses digits, summed from their series and added to ZH (same
// factored accumulation, same columns).  A workgroup takes a tile of
// NF_T rows x NF_T column elements, keeps the pairs (I, J) whose element J
// has a Gauss point within TAU_OFF / |ks| of row I's collocation point
// (centroid distance < reach + the element's Gauss-point radius) in an LDS
// list, and gives each listed pair to one lane, which loops over J's Gauss
// points.  Tiles far from the diagonal band keep no pair and end after the
// distance tests; the pair order in the list does not affect results (each
// pair's sum is one lane's, in Gauss-point order).
constexpr int NF_T = 32;
__global__ __launch_bounds__(256) void k_dyn_offdiag_near(int nbe, int npg2,
                                                          const ElemData* __restrict__ elem,
                                                          const GaussPt* __restrict__ gp,
                                                          const int* __restrict__ colflag,
                                                          DynConst K, cd* __restrict__ zh,
                                                          size_t ldzh) {
    __shared__ double rx[NF_T], ry[NF_T], rz[NF_T];
    __shared__ double jx[NF_T], jy[NF_T], jz[NF_T], jr[NF_T];
    __shared__ unsigned short list[NF_T * NF_T];
    __shared__ int count;
    const int i0 = blockIdx.x * NF_T, j0 = blockIdx.y * NF_T;
    const int t = threadIdx.x;
    if (t < NF_T) {
        const ElemData e = elem[min(i0 + t, nbe - 1)];
        rx[t] = e.cx;
        ry[t] = e.cy;
        rz[t] = e.cz;
    } else if (t < 2 * NF_T) {
        const ElemData e = elem[min(j0 + t - NF_T, nbe - 1)];
        jx[t - NF_T] = e.cx;
        jy[t - NF_T] = e.cy;
        jz[t - NF_T] = e.cz;
        jr[t - NF_T] = e.pad0;
    }
    if (t == 0) count = 0;
    __syncthreads();
    const double reach = TAU_OFF / sqrt(K.ks_abs2);
    // NF_T^2 / 256 rounds with every lane active (the ballot needs them all)
    for (int c = t; c < NF_T * NF_T; c += 256) {
        const int il = c % NF_T, jl = c / NF_T;
        const int I = i0 + il, J = j0 + jl;
        bool nf = false;
        if (I < nbe && J < nbe && I != J && (colflag[J] & 14) != 14) {
            const double d = sqrt(dist2(jx[jl] - rx[il], jy[jl] - ry[il], jz[jl] - rz[il]));
            nf = d < reach + jr[jl] + 1e-9 * (reach + d);
        }
        const unsigned long long m = __ballot(nf);
        if (m) {
            int base = 0;
            if ((t & 63) == 0) base = atomicAdd(&count, __popcll(m));
            base = __shfl(base, 0);
            const int below = __builtin_amdgcn_mbcnt_hi(
                (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            if (nf) list[base + below] = (unsigned short)c;
        }
    }
    __syncthreads();
    const int npairs = count;
    for (int k = t; k < npairs; k += 256) {
        const int c = list[k];
        const int il = c % NF_T, jl = c / NF_T;
        const int I = i0 + il, J = j0 + jl;
        const double xc = rx[il], yc = ry[il], zc = rz[il];
        const int flag = colflag[J];
        const double nv[3] = {elem[J].nx, elem[J].ny, elem[J].nz};
        const GaussPt* g = gp + (size_t)J * npg2;
        cd SP[3], SC[3], SB[6];
        for (int q = 0; q < 3; ++q) SP[q] = SC[q] = mk(0, 0);
        for (int q = 0; q < 6; ++q) SB[q] = mk(0, 0);
        for (int q = 0; q < npg2; ++q) {
            const GaussPt p = g[q];
            const double r1 = p.x - xc, r2 = p.y - yc, r3 = p.z - zc;
            const double RR = dist2(r1, r2, r3);
            if (!near_field(K, RR)) continue;
            const double iR = rsqrt_refined(RR);
            const double R = RR * iR;
            const double rd[3] = {r1 * iR, r2 * iR, r3 * iR};
            const double drn = rd[0] * nv[0] + rd[1] * nv[1] + rd[2] * nv[2];
            cd chiR, psidR2, chidR2, unused, A, B, C;
            ser_sums<NSER_OFF, 0, false>(K, K.ks * R, K.ser_rtop_off, chiR, psidR2, chidR2,
                                         unused);
            abc_from(K, chiR, psidR2, chidR2, A, B, C);
            const double wq = (p.w * K.q4pi) * (iR * iR);
            const cd Ap = A * wq, Bp = B * (wq * drn), Cp = C * wq;
            for (int j = 0; j < 3; ++j) {
                SP[j] = SP[j] + Ap * rd[j];
                SC[j] = SC[j] + Cp * rd[j];
            }
            SB[0] = SB[0] + Bp * (rd[0] * rd[0]);
            SB[1] = SB[1] + Bp * (rd[1] * rd[0]);
            SB[2] = SB[2] + Bp * (rd[2] * rd[0]);
            SB[3] = SB[3] + Bp * (rd[1] * rd[1]);
            SB[4] = SB[4] + Bp * (rd[2] * rd[1]);
            SB[5] = SB[5] + Bp * (rd[2] * rd[2]);
        }
        const cd SD = SP[0] * nv[0] + SP[1] * nv[1] + SP[2] * nv[2];
        const int sidx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        for (int kk = 0; kk < 3; ++kk) {
            if ((flag >> (1 + kk)) & 1) continue;
            cd* col = zh + (size_t)(3 * J + kk) * ldzh + 3 * I;
            for (int i = 0; i < 3; ++i) {
                cd h = SB[sidx[i + 3 * kk]] + SC[i] * nv[kk] + SP[kk] * nv[i];
                if (i == kk) h = h + SD;
                col[i] = col[i] + h;
            }
        }
    }
}

// Flagged elements (nonzero traction in some load case, or a KODE = 0
// column): H and G, the swapped KODE = 0 columns, and per load case c the
// partial RHS  sum_J G'_IJ ZDFI_c(J)  of this chunk of the flagged list into
// rhs_part[chunk][c][row].  A workgroup covers 64 rows (one per lane) and
// one chunk of JTG flagged elements, wave w taking element w of the chunk,
// so the few flagged columns are spread over many workgroups; the waves'
// RHS terms are summed through LDS in element order (a fixed order: the
// same result on every run and in every batch).
template <int JTG>
__global__ __launch_bounds__(64 * JTG) void k_dyn_offdiag_g(
    int nbe, int npg2, const ElemData* __restrict__ elem, const GaussPt* __restrict__ gp,
    const int* __restrict__ colflag, const int* __restrict__ flist, const int* __restrict__ fcount,
    const cd* __restrict__ zdfi, int nrhs, size_t ldr, DynConst K, cd* __restrict__ zh,
    size_t ldzh, cd* __restrict__ rhs_part) {
    extern __shared__ GaussPt sgp[];
    __shared__ cd rb[JTG][MAXRHS][3][64];
    const int cnt = *fcount;
    const int lane = threadIdx.x & 63, jj = threadIdx.x >> 6;
    const int I = blockIdx.x * 64 + lane;
    const bool rowok = I < nbe;
    const int Ic = rowok ? I : nbe - 1;
    const double xc = elem[Ic].cx, yc = elem[Ic].cy, zc = elem[Ic].cz;
    // chunks blockIdx.y, + gridDim.y, ... of the flagged list (the grid is
    // not sized by the flagged count, which only the device knows)
    for (int chunk = blockIdx.y; chunk * JTG < cnt; chunk += gridDim.y) {
        const int f0 = chunk * JTG;
        const int fn = min(JTG, cnt - f0);
        __syncthreads();  // the previous chunk is done with sgp and rb
        for (int t = threadIdx.x; t < fn * npg2; t += 64 * JTG) {
            int jf = t / npg2, q = t - jf * npg2;
            sgp[t] = gp[(size_t)flist[f0 + jf] * npg2 + q];
        }
        __syncthreads();
        cd b[MAXRHS][3];
        for (int c = 0; c < MAXRHS; ++c) b[c][0] = b[c][1] = b[c][2] = mk(0, 0);
        if (jj < fn) {
            const int J = flist[f0 + jj];
            const int flag = colflag[J];
            const double nv[3] = {elem[J].nx, elem[J].ny, elem[J].nz};
            cd H[9], G[9], Ho[9], Gn[9];
            offdiag_element<true>(K, sgp + jj * npg2, npg2, xc, yc, zc, nv, H, G);
            swap_cols(H, G, flag, K.zge, Ho, Gn);
            if (rowok && I != J) {
                for (int k = 0; k < 3; ++k) {
                    if (!((flag >> (1 + k)) & 1)) continue;
                    cd* col = zh + (size_t)(3 * J + k) * ldzh + 3 * I;
                    col[0] = Ho[3 * k];
                    col[1] = Ho[3 * k + 1];
                    col[2] = Ho[3 * k + 2];
                }
                for (int c = 0; c < nrhs; ++c) {
                    const cd* t = zdfi + c * ldr + 3 * J;
                    cd t0 = t[0], t1 = t[1], t2 = t[2];
                    for (int i = 0; i < 3; ++i) b[c][i] = Gn[i] * t0 + Gn[i + 3] * t1 + Gn[i + 6] * t2;
                }
            }
        }
        for (int c = 0; c < nrhs; ++c)
            for (int i = 0; i < 3; ++i) rb[jj][c][i][lane] = b[c][i];
        __syncthreads();
        if (jj == 0 && rowok)
            for (int c = 0; c < nrhs; ++c) {
                cd* o = rhs_part + ((size_t)chunk * nrhs + c) * ldr + 3 * I;
                for (int i = 0; i < 3; ++i) {
                    cd acc = rb[0][c][i][lane];
                    for (int w = 1; w < JTG; ++w) acc = acc + rb[w][c][i][lane];
                    o[i] = acc;
                }
            }
    }
}

// ZFI_c = diagonal part + the chunk partials of the flagged list, chunks
// summed in order (deterministic).
__global__ void k_rhs_reduce(int nn, int nrhs, size_t ldr, int jtg, const int* __restrict__ fcount,
                             const cd* __restrict__ part, const cd* __restrict__ diag,
                             cd* __restrict__ zfi) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (r >= nn || c >= nrhs) return;
    const int chunks = (*fcount + jtg - 1) / jtg;
    cd s = diag[c * ldr + r];
    for (int t = 0; t < chunks; ++t) s = s + part[((size_t)t * nrhs + c) * ldr + r];
    zfi[c * ldr + r] = s;
}

// Column flags from KODE and the load cases (bit 0: the element needs G --
// nonzero traction in some case or a KODE = 0 column; bit 1+k: column 3J+k
// has KODE = 0) and the ascending list of flagged elements.  One workgroup.
__global__ __launch_bounds__(256) void k_colflags(int nbe, const int* __restrict__ kode,
                                                  const cd* __restrict__ zdfi, int nrhs, size_t ldr,
                                                  int* __restrict__ flag, int* __restrict__ flist,
                                                  int* __restrict__ fcount) {
    __shared__ int wcnt[4];
    __shared__ int base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) base = 0;
    __syncthreads();
    for (int j0 = 0; j0 < nbe; j0 += 256) {
        const int J = j0 + tid;
        int f = 0;
        if (J < nbe) {
            for (int k = 0; k < 3; ++k) {
                if (kode[3 * J + k] == 0) f |= (1 << (1 + k)) | 1;
                for (int c = 0; c < nrhs; ++c) {
                    cd t = zdfi[c * ldr + 3 * J + k];
                    if (t.x != 0.0 || t.y != 0.0) f |= 1;
                }
            }
            flag[J] = f;
        }
        const bool on = f & 1;
        unsigned long long m = __ballot(on);
        int below = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += wcnt[w];
        if (on) flist[off + below] = J;
        __syncthreads();
        if (tid == 0) base += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        __syncthreads();
    }
    if (tid == 0) *fcount = base;
}

// INTEREC's un-swap (Interec.for:59-67): for KODE = 0, tmp = ZFI mu*,
// ZFI = ZDFI, ZDFI = tmp.
__global__ void k_unswap(int nn, const int* __restrict__ kode, cd* __restrict__ zfi,
                         cd* __restrict__ zdfi, cd zge) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn || kode[i] != 0) return;
    cd t = zfi[i] * zge;
    zfi[i] = zdfi[i];
    zdfi[i] = t;
}

// ---------------------------------------------------------------------------
// Interior points.  One wavefront per (point k, block of 64 elements), one
// lane per element J.  Per Gauss point the displacement kernels (Solfund) and
// the stress kernels (Sigmaec.for:139-198) are contracted at once with the
// element's traction t = ZDFI_J and displacement u = ZFI_J:
//   du_i     += w [ (G t)_i - (T u)_i ]
//   sigma_ab += w [ sum_m t_m ZD(m,a,b) - u_m ZS(m,a,b) ]   (symmetric in a,b)
// and the wave reduces its 64 elements in a fixed order into part[k][blk].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_interior(int n, int nbe, int npg2, const ElemData* __restrict__ elem,
                                                 const GaussPt* __restrict__ gpT,
                                                 const double* __restrict__ cxi,
                                                 const double* __restrict__ cyi,
                                                 const double* __restrict__ czi,
                                                 const cd* __restrict__ zfi,
                                                 const cd* __restrict__ zdfi, DynConst K,
                                                 cd* __restrict__ part) {
    const int k = blockIdx.y;
    const int J = blockIdx.x * 64 + threadIdx.x;
    const bool ok = J < nbe;
    const int Jc = ok ? J : nbe - 1;
    const double xk = cxi[k], yk = cyi[k], zk = czi[k];
    const double nv[3] = {elem[Jc].nx, elem[Jc].ny, elem[Jc].nz};
    cd t[3], u[3];
    for (int m = 0; m < 3; ++m) {
        t[m] = ok ? zdfi[3 * J + m] : mk(0, 0);
        u[m] = ok ? zfi[3 * J + m] : mk(0, 0);
    }
    const cd nu = u[0] * nv[0] + u[1] * nv[1] + u[2] * nv[2];
    cd du[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
    cd sg[6];
    for (int e = 0; e < 6; ++e) sg[e] = mk(0, 0);
    const cd mu4 = K.zge * K.q4pi;
    for (int q = 0; q < npg2; ++q) {
        GaussPt p = gpT[(size_t)q * n + Jc];
        double r1 = p.x - xk, r2 = p.y - yk, r3 = p.z - zk;
        double R = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
        double iR = 1.0 / R, iR2 = iR * iR;
        double rd[3] = {r1 * iR, r2 * iR, r3 * iR};
        double drn = rd[0] * nv[0] + rd[1] * nv[1] + rd[2] * nv[2];
        // psi, chi and their first derivatives, as in dyn_scalars
        cd Es = exp_over_r(K.ks, R, iR);
        cd Ep = exp_over_r(K.kp, R, iR);
        cd as = K.as * iR, as2 = K.as2 * iR2, ap = K.ap * iR, ap2 = K.ap2 * iR2;
        cd zs = K.ks * R, zp = K.kp * R;
        cd Epc = K.c2 * Ep;
        cd psi = mk(1.0 + as2.x - as.x, as2.y - as.y) * Es - (ap2 - ap) * Epc;
        cd chi = mk(1.0 + 3.0 * (as2.x - as.x), 3.0 * (as2.y - as.y)) * Es -
                 mk(1.0 + 3.0 * (ap2.x - ap.x), 3.0 * (ap2.y - ap.y)) * Epc;
        cd psid = (mk(-2.0 + zs.x + 3.0 * (as.x - as2.x), zs.y + 3.0 * (as.y - as2.y)) * Es -
                   mk(-1.0 + 3.0 * (ap.x - ap2.x), 3.0 * (ap.y - ap2.y)) * Epc) * iR;
        cd chid = (mk(zs.x - 4.0 + 9.0 * (as.x - as2.x), zs.y + 9.0 * (as.y - as2.y)) * Es -
                   mk(zp.x - 4.0 + 9.0 * (ap.x - ap2.x), zp.y + 9.0 * (ap.y - ap2.y)) * Epc) * iR;
        cd chr = chi * iR;
        cd A = psid - chr;
        cd B = 4.0 * chr - 2.0 * chid;
        cd C = K.c1m2 * (A + 0.5 * B - 3.0 * chr) - 2.0 * chr;
        // second derivatives (Sigmaec.for:146-159); q1 = -zp, q2 = -zs,
        // exp(-q) / (q^2 R^3) = (exp(z)/R) (1/z^2) / R^2
        cd q1 = -zp, q2 = -zs;
        cd q1s = q1 * q1, q2s = q2 * q2;
        double iR4 = iR2 * iR2;
        cd e1 = Epc * K.ap2 * iR4;  // c2 exp(zp) / (zp^2 R^3) . R^0
        cd e2 = Es * K.as2 * iR4;
        cd p1a = mk(12.0 + 12.0 * q1.x, 12.0 * q1.y) + 5.0 * q1s + q1s * q1;
        cd p2a = mk(12.0 + 12.0 * q2.x, 12.0 * q2.y) + 7.0 * q2s + 3.0 * (q2s * q2) + q2s * q2s;
        cd p1b = mk(36.0 + 36.0 * q1.x, 36.0 * q1.y) + 17.0 * q1s + 5.0 * (q1s * q1) + q1s * q1s;
        cd p2b = mk(36.0 + 36.0 * q2.x, 36.0 * q2.y) + 17.0 * q2s + 5.0 * (q2s * q2) + q2s * q2s;
        cd fhidrr = e2 * p2a - e1 * p1a;
        cd cappadrr = e2 * p2b - e1 * p1b;
        // Sigmaec.for:161-176
        cd cr = chi * iR, cr2 = chi * iR2, cdr = chid * iR, fdr = psid * iR;
        cd cte1 = 2.0 * cr;
        cd cte2 = K.lm * (psid - chid - cte1);
        cd cte3 = psid - cr;
        cd cte4 = 2.0 * (chid - cte1);
        cd cte5 = 4.0 * (cappadrr - 5.0 * cdr + 8.0 * cr2);
        cd cte6 = fhidrr - fdr - 3.0 * cdr + 6.0 * cr2;
        cd cte7 = 2.0 * cdr - 4.0 * cr2;
        cd cte8 = K.lm * (cappadrr + cdr - 4.0 * cr2 - fhidrr + fdr);
        cd cte9 = 2.0 * (cte7 + cte8);
        cd cte10 = 4.0 * cr2 + 4.0 * (K.lm * (cdr + 2.0 * cr2 - fdr));
        cd cte11 = (K.lm * K.lm) * (cappadrr + 4.0 * cdr + 2.0 * cr2 - fhidrr - 2.0 * fdr);
        cd cte12 = 2.0 * (fdr - cr2);
        // contractions
        cd rt = t[0] * rd[0] + t[1] * rd[1] + t[2] * rd[2];
        cd ru = u[0] * rd[0] + u[1] * rd[1] + u[2] * rd[2];
        double w = p.w;
        cd c0w = K.c0 * w;
        cd Tw = A * (K.q4pi * w);
        cd Bw = B * (K.q4pi * w * drn);
        cd Cw = C * (K.q4pi * w);
        for (int i = 0; i < 3; ++i) {
            cd gt = c0w * (psi * t[i] - chi * (rd[i] * rt));
            cd tu = Tw * (u[i] * drn + ru * nv[i]) + Bw * (ru * rd[i]) + Cw * (nu * rd[i]);
            du[i] = du[i] + gt - tu;
        }
        cd dA = (cte1 - cte2) * rt;  // coefficient of d_ab
        cd sA = (cte9 * ru) * drn + (cte10 + cte11) * nu;
        int e = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b, ++e) {
                double rab = rd[a] * rd[b];
                cd zd = -cte3 * (t[b] * rd[a] + t[a] * rd[b]) + cte4 * (rt * rab);
                cd zsv = (cte5 * (ru * rab) - cte6 * (u[b] * rd[a] + u[a] * rd[b])) * drn +
                         cte9 * (nu * rab) - cte6 * (ru * (rd[a] * nv[b] + rd[b] * nv[a])) -
                         cte12 * (u[a] * nv[b] + u[b] * nv[a]);
                if (a == b) {
                    zd = zd + dA;
                    zsv = zsv + sA;
                }
                sg[e] = sg[e] + (zd * K.q4pi - zsv * mu4) * w;
            }
    }
    double v[18];
    for (int i = 0; i < 3; ++i) {
        v[2 * i] = ok ? du[i].x : 0.0;
        v[2 * i + 1] = ok ? du[i].y : 0.0;
    }
    for (int e = 0; e < 6; ++e) {
        v[6 + 2 * e] = ok ? sg[e].x : 0.0;
        v[7 + 2 * e] = ok ? sg[e].y : 0.0;
    }
    for (int i = 0; i < 18; ++i) v[i] = wave_sum(v[i]);
    if (threadIdx.x == 0) {
        cd* o = part + ((size_t)k * gridDim.x + blockIdx.x) * 9;
        for (int i = 0; i < 9; ++i) o[i] = mk(v[2 * i], v[2 * i + 1]);
    }
}

// Sum the element blocks of each interior point in order; expand the
// symmetric stress to the 9 entries in the order (11,21,31,12,22,32,13,23,33)
// of Interec.for:227-280.
__global__ void k_interior_reduce(int l, int nblk, const cd* __restrict__ part,
                                  cd* __restrict__ zdsol, cd* __restrict__ zssol) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= l) return;
    cd s[9];
    for (int i = 0; i < 9; ++i) s[i] = mk(0, 0);
    for (int b = 0; b < nblk; ++b)
        for (int i = 0; i < 9; ++i) s[i] = s[i] + part[((size_t)k * nblk + b) * 9 + i];
    for (int i = 0; i < 3; ++i) zdsol[3 * k + i] = s[i];
    // packed (00,01,02,11,12,22) at s[3..8]
    const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int e = 0; e < 9; ++e) zssol[9 * k + e] = s[3 + map[e]];
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from bemx_api.hip).
// ---------------------------------------------------------------------------
hipError_t launch_element_setup(hipStream_t st, int n, const int* cone, const double* cx,
                                const double* cy, const double* cz, const double* gi,
                                const double* ome, int npg, int use_given, ElemData* elem,
                                GaussPt* gp, GaussPt* gpT, int* status) {
    hipLaunchKernelGGL(k_element_setup, dim3((n + 127) / 128), dim3(128), 0, st, n, cone, cx, cy,
                       cz, gi, ome, npg, use_given, elem, gp, gpT, status);
    return hipGetLastError();
}

hipError_t launch_static_hdiag(hipStream_t st, int n, int nbe, int npg2, const ElemData* elem,
                               const GaussPt* gp, double c3, double c4, double* part, int nsplit,
                               double* hdiag) {
    int jper = (n + nsplit - 1) / nsplit;
    hipLaunchKernelGGL(k_static_hdiag, dim3((nbe + 63) / 64, nsplit), dim3(64), 0, st, n, nbe,
                       npg2, jper, elem, gp, c3, c4, part);
    hipLaunchKernelGGL(k_static_hdiag_reduce, dim3((nbe * 9 + 255) / 256), dim3(256), 0, st, nbe,
                       nsplit, part, hdiag);
    return hipGetLastError();
}

hipError_t launch_gest_diag(hipStream_t st, int n, int nbe, const int* cone, const double* cx,
                            const double* cy, const double* cz, const ElemData* elem,
                            const double* gi, const double* ome, int npg, double c1, double c2,
                            double* gdiag, int* status) {
    hipLaunchKernelGGL(k_gest_diag, dim3(nbe), dim3(64), 0, st, n, nbe, cone, cx, cy, cz, elem,
                       gi, ome, npg, c1, c2, gdiag, status);
    return hipGetLastError();
}

constexpr int OD_WAVES = 4;
// 4 column elements per workgroup: 4860 workgroups of equal work, so the
// last partial round of resident workgroups is a small part of the launch
// (8 elements, 2430 workgroups in 2.4 rounds: 3.69 ms at 2160 elements
// against 3.48 ms; 2 elements measured the same as 4)
constexpr int OD_JT = 4;
constexpr int OG_JT = 4;

int offdiag_tiles(int nbe) { return (nbe + OG_JT - 1) / OG_JT; }

// Entries of each exp table of the H kernel for the wave numbers of K and
// distances up to rmax, with the grid spacing in *h; 0 when more than
// EXP_TAB_MAX entries would be needed or BEMX_EXP_TABLE=0 is set (the H
// kernel then evaluates exp and sincos at every Gauss point).
int exp_table_entries(const DynConst& K, double rmax, double* h) {
    const char* env = getenv("BEMX_EXP_TABLE");
    if (env && env[0] == '0') return 0;
    const double kmax = std::max(std::hypot(K.ks.x, K.ks.y), std::hypot(K.kp.x, K.kp.y));
    if (!(kmax > 0.0) || !(rmax > 0.0) || !std::isfinite(kmax * rmax)) return 0;
    const double hh = std::exp2(std::floor(std::log2(1.0 / (16.0 * kmax))));
    const double nt = std::floor(rmax / hh) + 2.0;
    if (!(nt <= (double)EXP_TAB_MAX)) return 0;
    *h = hh;
    return (int)nt;
}

// Dynamic assembly of ZH and of ZFI for nrhs load cases (zdfi, rhs_diag,
// zfi: nrhs columns of leading dimension ldr = 3 NBE).  flist/fcount: NBE + 1
// ints of scratch; rhs_part: offdiag_tiles(nbe) * nrhs * ldr complex.
hipError_t launch_dyn_assembly(hipStream_t st, int nbe, int npg2, const ElemData* elem,
                               const GaussPt* gp, const double* hdiag, const double* gdiag,
                               const int* kode, const cd* zdfi, int nrhs, int* colflag,
                               int* flist, const DynConst& K, cd* etab, int nt, double h,
                               cd* zh, size_t ldzh, cd* rhs_part, cd* rhs_diag, cd* zfi,
                               hipEvent_t ev_od0, hipEvent_t ev_od1) {
    const size_t ldr = 3 * (size_t)nbe;
    // exp tables of the H kernel (etab: 2 nt entries)
    ExpTab T{};
    const bool tab = etab && nt > 0;
    if (tab) {
        T.ts = etab;
        T.tp = etab + nt;
        T.h = h;
        T.ih = 1.0 / h;
        T.nt = nt;
        hipLaunchKernelGGL(k_exp_table, dim3((nt + 255) / 256), dim3(256), 0, st, nt, h, K.ks,
                           K.kp, K.c2, etab, etab + nt);
    }
    int* fcount = flist + nbe;
    hipLaunchKernelGGL(k_colflags, dim3(1), dim3(256), 0, st, nbe, kode, zdfi, nrhs, ldr, colflag,
                       flist, fcount);
    hipLaunchKernelGGL(k_dyn_diag, dim3(nbe), dim3(64), 0, st, nbe, npg2, elem, gp, hdiag, gdiag,
                       colflag, zdfi, nrhs, ldr, K, zh, ldzh, rhs_diag);
    const int rowblk = (nbe + 64 * OD_WAVES - 1) / (64 * OD_WAVES);
    if (ev_od0) (void)hipEventRecord(ev_od0, st);
    const dim3 odgrid(rowblk, (nbe + OD_JT - 1) / OD_JT), odblock(64 * OD_WAVES);
    const size_t odlds = (size_t)OD_JT * npg2 * sizeof(GaussPt);
    if (tab) {
        hipLaunchKernelGGL((k_dyn_offdiag_h<OD_WAVES, OD_JT, true>), odgrid, odblock, odlds, st,
                           nbe, npg2, elem, gp, colflag, K, T, zh, ldzh);
        if (ev_od1) (void)hipEventRecord(ev_od1, st);
        const int nft = (nbe + NF_T - 1) / NF_T;
        hipLaunchKernelGGL(k_dyn_offdiag_near, dim3(nft, nft), dim3(256), 0, st, nbe, npg2, elem,
                           gp, colflag, K, zh, ldzh);
    } else {
        hipLaunchKernelGGL((k_dyn_offdiag_h<OD_WAVES, OD_JT, false>), odgrid, odblock, odlds, st,
                           nbe, npg2, elem, gp, colflag, K, T, zh, ldzh);
        if (ev_od1) (void)hipEventRecord(ev_od1, st);
    }
    hipLaunchKernelGGL((k_dyn_offdiag_g<OG_JT>),
                       dim3((nbe + 63) / 64, std::min(offdiag_tiles(nbe), 64)),
                       dim3(64 * OG_JT), (size_t)OG_JT * npg2 * sizeof(GaussPt), st, nbe, npg2,
                       elem, gp, colflag, flist, fcount, zdfi, nrhs, ldr, K, zh, ldzh, rhs_part);
    hipLaunchKernelGGL(k_rhs_reduce, dim3((3 * nbe + 255) / 256, nrhs), dim3(256), 0, st, 3 * nbe,
                       nrhs, ldr, OG_JT, fcount, rhs_part, rhs_diag, zfi);
    return hipGetLastError();
}

// Dynamic assembly of two frequencies of a sweep (same mesh, boundary data
// and load cases): per frequency the exp table, the diagonal blocks, the
// near field, the flagged columns and the right-hand sides, as
// launch_dyn_assembly; the column flags once; the H blocks of both
// frequencies in one k_dyn_offdiag_h2 launch.  Each frequency's ZH and ZFI
// are bit-identical to launch_dyn_assembly's.  Both frequencies use the
// table route (nt > 0).
hipError_t launch_dyn_assembly_pair(hipStream_t st, int nbe, int npg2, const ElemData* elem,
                                    const GaussPt* gp, const double* hdiag, const double* gdiag,
                                    const int* kode, const cd* zdfi, int nrhs, int* colflag,
                                    int* flist, const DynFreq (&F)[2], size_t ldzh,
                                    hipEvent_t ev_od0, hipEvent_t ev_od1) {
    const size_t ldr = 3 * (size_t)nbe;
    DynPair P{};
    for (int f = 0; f < 2; ++f) {
        if (!F[f].etab || F[f].nt <= 0) return hipErrorInvalidValue;
        P.K[f] = F[f].K;
        P.T[f] = ExpTab{F[f].etab, F[f].etab + F[f].nt, F[f].h, 1.0 / F[f].h, F[f].nt};
        P.zh[f] = F[f].zh;
        hipLaunchKernelGGL(k_exp_table, dim3((F[f].nt + 255) / 256), dim3(256), 0, st, F[f].nt,
                           F[f].h, F[f].K.ks, F[f].K.kp, F[f].K.c2, F[f].etab,
                           F[f].etab + F[f].nt);
    }
    int* fcount = flist + nbe;
    hipLaunchKernelGGL(k_colflags, dim3(1), dim3(256), 0, st, nbe, kode, zdfi, nrhs, ldr, colflag,
                       flist, fcount);
    for (int f = 0; f < 2; ++f)
        hipLaunchKernelGGL(k_dyn_diag, dim3(nbe), dim3(64), 0, st, nbe, npg2, elem, gp, hdiag,
                           gdiag, colflag, zdfi, nrhs, ldr, F[f].K, F[f].zh, ldzh,
                           F[f].rhs_diag);
    const int rowblk = (nbe + 64 * OD_WAVES - 1) / (64 * OD_WAVES);
    if (ev_od0) (void)hipEventRecord(ev_od0, st);
    hipLaunchKernelGGL((k_dyn_offdiag_h2<OD_WAVES, OD_JT>),
                       dim3(rowblk, (nbe + OD_JT - 1) / OD_JT), dim3(64 * OD_WAVES),
                       (size_t)OD_JT * npg2 * sizeof(GaussPt), st, nbe, npg2, elem, gp, colflag, P,
                       ldzh);
    if (ev_od1) (void)hipEventRecord(ev_od1, st);
    const int nft = (nbe + NF_T - 1) / NF_T;
    for (int f = 0; f < 2; ++f) {
        hipLaunchKernelGGL(k_dyn_offdiag_near, dim3(nft, nft), dim3(256), 0, st, nbe, npg2, elem,
                           gp, colflag, F[f].K, F[f].zh, ldzh);
        hipLaunchKernelGGL((k_dyn_offdiag_g<OG_JT>),
                           dim3((nbe + 63) / 64, std::min(offdiag_tiles(nbe), 64)),
                           dim3(64 * OG_JT), (size_t)OG_JT * npg2 * sizeof(GaussPt), st, nbe,
                           npg2, elem, gp, colflag, flist, fcount, zdfi, nrhs, ldr, F[f].K,
                           F[f].zh, ldzh, F[f].rhs_part);
        hipLaunchKernelGGL(k_rhs_reduce, dim3((3 * nbe + 255) / 256, nrhs), dim3(256), 0, st,
                           3 * nbe, nrhs, ldr, OG_JT, fcount, F[f].rhs_part, F[f].rhs_diag,
                           F[f].zfi);
    }
    return hipGetLastError();
}

hipError_t launch_unswap(hipStream_t st, int nn, const int* kode, cd* zfi, cd* zdfi, cd zge) {
    hipLaunchKernelGGL(k_unswap, dim3((nn + 255) / 256), dim3(256), 0, st, nn, kode, zfi, zdfi,
                       zge);
    return hipGetLastError();
}

int interior_blocks(int nbe) { return (nbe + 63) / 64; }

hipError_t launch_interior(hipStream_t st, int n, int nbe, int npg2, int l, const ElemData* elem,
                           const GaussPt* gpT, const double* cxi, const double* cyi,
                           const double* czi, const cd* zfi, const cd* zdfi, const DynConst& K,
                           cd* part, cd* zdsol, cd* zssol) {
    int nblk = interior_blocks(nbe);
    hipLaunchKernelGGL(k_interior, dim3(nblk, l), dim3(64), 0, st, n, nbe, npg2, elem, gpT, cxi,
                       cyi, czi, zfi, zdfi, K, part);
    hipLaunchKernelGGL(k_interior_reduce, dim3((l + 63) / 64), dim3(64), 0, st, l, nblk, part,
                       zdsol, zssol);
    return hipGetLastError();
}

size_t exp_table_max_bytes() { return sizeof(cd) * 2 * (size_t)EXP_TAB_MAX; }

hipError_t preload_assembly_code() {
    hipFuncAttributes a;
    return hipFuncGetAttributes(&a, (const void*)k_exp_table);
}

}  // namespace bemx
