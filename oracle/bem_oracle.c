/*
 * bem_oracle.c -- CPU FP64 restatement of the reference BEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against; it is linked by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg, never by the product library.
 *
 * It restates, in plain C99 real / real complex arithmetic (real = double
 * in the default build), the
 * algorithm of the reference (giulianobelinassi/BEM-MECECCS, FPREC=double
 * CPU build).  Each routine cites the reference file:line it follows.
 * Layout conventions are the reference's Fortran ones:
 *   - connectivity CONE(N,4), column-major, 1-based node ids:  cone[k*n + j]
 *   - normals ETAS(3,N):  etas[3*j + c]
 *   - 3x3 blocks (i,j) column-major:  b[i + 3*j]
 *   - HESTdiag/GESTdiag(3,3,NBE):  d[9*e + i + 3*j]
 *   - ZH(3NBE,3NBE) column-major with leading dimension 3NBE
 *   - stress kernels ZD(3,3,3)/ZS(3,3,3):  z[k + 3*i + 9*j]
 * Error status: 0 ok, 1 degenerate Jacobian (DET < 1e-5), 2 bad argument.
 *
 * Parity pins (see tests/test_oracle_*.py): every function here is checked
 * against the reference's own Fortran routines compiled under oracle/_ref,
 * against the reference's golden output decks (SSOLO240D/SSOLO2160D), and
 * against closed forms (Legendre rules, Kelvin static limit, Hooke's law on
 * finite-difference gradients, LAPACK zgesv).
 */
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

/* The working precision.  The default build is the FP64 restatement of the
 * reference; the build with -DORC_LONG_DOUBLE evaluates the same formulas in
 * x87 extended precision (64-bit significand) and serves as the reference
 * value where the FP64 formulas lose digits (the dynamic kernels at small
 * k R), through the orcld_* entry points at the end of this file. */
#ifdef ORC_LONG_DOUBLE
typedef long double real;
typedef long double complex zc;
#define ORC_CMPLX(a, b) CMPLXL((a), (b))
#else
typedef double real;
typedef double complex zc;
#define ORC_CMPLX(a, b) CMPLX((a), (b))
#endif

#define ORC_OK 0
#define ORC_DEGENERATE 1
#define ORC_BADARG 2

/* ------------------------------------------------------------------ */
/* Gauss-Legendre rule on [-1,1].  Gauleg.for:9-36 (Newton on P_n,    */
/* EPS = 3e-14, initial guess cos(3.141592654 (i-1/4)/(n+1/2))).       */
/* ------------------------------------------------------------------ */
void orc_gauleg(int n, real *x, real *w)
{
    const real x1 = -1.0, x2 = 1.0;
    const real xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    int m = (n + 1) / 2;
    for (int i = 1; i <= m; ++i) {
        real z = cos(3.141592654 * (i - 0.25) / (n + 0.5));
        real z1, pp, p1, p2, p3;
        do {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > 3e-14);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

/* Unit normal (P2-P1)x(P3-P1)/|.|.  Normvec.for:4-39 */
void orc_normvec(int n, int np, const int *cone, const real *cx,
                 const real *cy, const real *cz, real *etas)
{
    (void)np;
    for (int j = 0; j < n; ++j) {
        int n1 = cone[j] - 1, n2 = cone[n + j] - 1, n3 = cone[2 * n + j] - 1;
        real a = (cy[n2] - cy[n1]) * (cz[n3] - cz[n1]) -
                   (cz[n2] - cz[n1]) * (cy[n3] - cy[n1]);
        real b = (cz[n2] - cz[n1]) * (cx[n3] - cx[n1]) -
                   (cx[n2] - cx[n1]) * (cz[n3] - cz[n1]);
        real c = (cx[n2] - cx[n1]) * (cy[n3] - cy[n1]) -
                   (cy[n2] - cy[n1]) * (cx[n3] - cx[n1]);
        real r = sqrt(a * a + b * b + c * c);
        etas[3 * j + 0] = a / r;
        etas[3 * j + 1] = b / r;
        etas[3 * j + 2] = c / r;
    }
}

/* Collocation points: centroid of the 4 nodes.  Inputece.for:93-102 */
void orc_centroids(int n, const int *cone, const real *cx, const real *cy,
                   const real *cz, real *cxm, real *cym, real *czm)
{
    for (int j = 0; j < n; ++j) {
        int a = cone[j] - 1, b = cone[n + j] - 1, c = cone[2 * n + j] - 1,
            d = cone[3 * n + j] - 1;
        cxm[j] = (cx[a] + cx[b] + cx[c] + cx[d]) * 0.25;
        cym[j] = (cy[a] + cy[b] + cy[c] + cy[d]) * 0.25;
        czm[j] = (cz[a] + cz[b] + cz[c] + cz[d]) * 0.25;
    }
}

/* Element corner coordinates CO(4,3) of element j (0-based).
 * Ghmatecd.for:116-134 */
static void element_corners(int n, const int *cone, const real *cx,
                            const real *cy, const real *cz, int j,
                            real co[4][3])
{
    for (int k = 0; k < 4; ++k) {
        int p = cone[k * n + j] - 1;
        co[k][0] = cx[p];
        co[k][1] = cy[p];
        co[k][2] = cz[p];
    }
}

/* One Gauss point of the bilinear quad map: position xg and the weight
 * OME(ig)*OME(jg)*DET.  Nonsingd.for:45-106 (shape functions F, their
 * derivatives P, Jacobian XJ, DET = |XJ1 x XJ2|, DET < 1e-5 fatal). */
static int quad_point(const real co[4][3], real xi, real wxi, real eta,
                      real weta, real xg[3], real *wdet)
{
    real sp = 1.0 + eta, sm = 1.0 - eta, rp = 1.0 + xi, rm = 1.0 - xi;
    real p[2][4], f[4], xj[2][3];
    p[0][0] = -0.25 * sm;
    p[0][1] = 0.25 * sm;
    p[0][2] = 0.25 * sp;
    p[0][3] = -0.25 * sp;
    f[0] = 0.25 * rm * sm;
    f[1] = 0.25 * rp * sm;
    f[2] = 0.25 * rp * sp;
    f[3] = 0.25 * rm * sp;
    p[1][0] = -0.25 * rm;
    p[1][1] = -0.25 * rp;
    p[1][2] = 0.25 * rp;
    p[1][3] = 0.25 * rm;
    for (int i = 0; i < 2; ++i)
        for (int c = 0; c < 3; ++c) {
            real t = 0.0;
            for (int k = 0; k < 4; ++k) t = t + p[i][k] * co[k][c];
            xj[i][c] = t;
        }
    real a = xj[0][1] * xj[1][2] - xj[1][1] * xj[0][2];
    real b = xj[1][0] * xj[0][2] - xj[0][0] * xj[1][2];
    real c = xj[0][0] * xj[1][1] - xj[1][0] * xj[0][1];
    real det = sqrt(a * a + b * b + c * c);
    if (det < 1.0e-5) return ORC_DEGENERATE;
    for (int d = 0; d < 3; ++d) {
        real s = 0.0;
        for (int k = 0; k < 4; ++k) s = s + co[k][d] * f[k];
        xg[d] = s;
    }
    *wdet = wxi * weta * det;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Point kernels                                                       */
/* ------------------------------------------------------------------ */

/* Static traction kernel T(3,3).  Solfune.for:7-34 */
void orc_solfune(const real xp[3], const real xg[3], const real eta[3],
                 real c3, real c4, real *t)
{
    real r1 = xg[0] - xp[0], r2 = xg[1] - xp[1], r3 = xg[2] - xp[2];
    real r = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
    real rdn = (r1 * eta[0] + r2 * eta[1] + r3 * eta[2]) / r;
    real rd[3] = {r1 / r, r2 / r, r3 / r};
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            real dij = (i == j) ? 1.0 : 0.0;
            t[i + 3 * j] = (c3 / (r * r)) *
                           (rdn * (c4 * dij + 3.0 * rd[i] * rd[j]) +
                            c4 * (rd[j] * eta[i] - rd[i] * eta[j]));
        }
}

/* Static displacement kernel U(3,3).  Singge.for:175-202 */
void orc_solfune_g(const real xp[3], const real xg[3], real c1,
                   real c2, real *u)
{
    real r1 = xg[0] - xp[0], r2 = xg[1] - xp[1], r3 = xg[2] - xp[2];
    real r = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
    real rd[3] = {r1 / r, r2 / r, r3 / r};
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            real dij = (i == j) ? 1.0 : 0.0;
            u[i + 3 * j] = (c1 / r) * (c2 * dij + rd[i] * rd[j]);
        }
}

/* Dynamic scalar functions psi, chi, psi', chi' and A, B, C.
 * Solfund.for:29-53 (identical in Solfundif.for:29-53). */
struct dyn_scalars {
    zc c0, fhi, cappa, fhidr, cappadr, aa, bb, cc;
};

static void dyn_scalars(real r, zc zge, zc zcs, zc zcp, real pi,
                        real fr, struct dyn_scalars *s)
{
    zc zwi = ORC_CMPLX(0.0, fr);
    zc zc0 = 1.0 / (ORC_CMPLX(4.0 * pi, 0.0) * zge);
    zc zc1 = (zcp / zcs) * (zcp / zcs);
    zc zc2 = (zcs / zcp) * (zcs / zcp);
    zc zkp = -zwi / zcp;
    zc zks = -zwi / zcs;
    zc zzp = zkp * r;
    zc zzs = zks * r;
    zc zezp = exp(zzp);
    zc zezs = exp(zzs);
    zc zp2 = zzp * zzp;
    zc zs2 = zzs * zzs;
    real r2 = r * r;
    s->c0 = zc0;
    s->fhi = (1.0 + 1.0 / zs2 - 1.0 / zzs) * zezs / r -
             zc2 * (1.0 / zp2 - 1.0 / zzp) * zezp / r;
    s->cappa = (1.0 + 3.0 / zs2 - 3.0 / zzs) * zezs / r -
               zc2 * (1.0 + 3.0 / zp2 - 3.0 / zzp) * zezp / r;
    s->fhidr = (-2.0 + zzs + 3.0 / zzs - 3.0 / zs2) * zezs / r2 -
               zc2 * (-1.0 + 3.0 / zzp - 3.0 / zp2) * zezp / r2;
    s->cappadr = (zzs - 4.0 + 9.0 / zzs - 9.0 / zs2) * zezs / r2 -
                 zc2 * (zzp - 4.0 + 9.0 / zzp - 9.0 / zp2) * zezp / r2;
    s->aa = s->fhidr - s->cappa / r;
    s->bb = 4.0 * s->cappa / r - 2.0 * s->cappadr;
    s->cc = (zc1 - 2.0) * (s->aa + 0.5 * s->bb - 3.0 * s->cappa / r) -
            2.0 * s->cappa / r;
}

/* Dynamic displacement ZU(3,3) and traction ZT(3,3) kernels.
 * Solfund.for:7-63 */
void orc_solfund(const real xp[3], const real xg[3], const real rn[3],
                 zc zge, zc zcs, zc zcp, real pi, real fr, zc *zu, zc *zt)
{
    real r1 = xg[0] - xp[0], r2 = xg[1] - xp[1], r3 = xg[2] - xp[2];
    real r = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
    real drn = (r1 * rn[0] + r2 * rn[1] + r3 * rn[2]) / r;
    real rd[3] = {r1 / r, r2 / r, r3 / r};
    struct dyn_scalars s;
    dyn_scalars(r, zge, zcs, zcp, pi, fr, &s);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            real dij = (i == j) ? 1.0 : 0.0;
            zu[i + 3 * j] = s.c0 * (s.fhi * dij - s.cappa * rd[j] * rd[i]);
            zt[i + 3 * j] = (1.0 / (4.0 * pi)) *
                            ((s.aa * (drn * dij + rd[j] * rn[i])) +
                             rd[i] * rd[j] * drn * s.bb + rd[i] * rn[j] * s.cc);
        }
}

/* Regularised difference (dynamic - static) kernels.  Solfundif.for:7-72 */
void orc_solfundif(const real xp[3], const real xg[3], const real rn[3],
                   zc zge, zc zcs, zc zcp, real c1, real c2, real c3,
                   real c4, real pi, real fr, zc *zudif, zc *ztdif)
{
    real r1 = xg[0] - xp[0], r2 = xg[1] - xp[1], r3 = xg[2] - xp[2];
    real r = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
    real drn = (r1 * rn[0] + r2 * rn[1] + r3 * rn[2]) / r;
    real rd[3] = {r1 / r, r2 / r, r3 / r};
    struct dyn_scalars s;
    dyn_scalars(r, zge, zcs, zcp, pi, fr, &s);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real dij = (i == j) ? 1.0 : 0.0;
            zudif[i + 3 * j] = s.c0 * (s.fhi * dij - s.cappa * rd[j] * rd[i]) -
                               ORC_CMPLX((c1 / r) * (c2 * dij + rd[i] * rd[j]), 0.0);
            ztdif[i + 3 * j] =
                (1.0 / (4.0 * pi)) * ((s.aa * (drn * dij + rd[j] * rn[i])) +
                                      rd[i] * rd[j] * drn * s.bb +
                                      rd[i] * rn[j] * s.cc) -
                ORC_CMPLX((c3 / (r * r)) * (drn * (c4 * dij + 3.0 * rd[i] * rd[j]) +
                                        c4 * (rd[j] * rn[i] - rd[i] * rn[j])),
                      0.0);
        }
}

/* ------------------------------------------------------------------ */
/* Element quadratures                                                 */
/* ------------------------------------------------------------------ */

/* NPG x NPG Gauss of the static traction kernel.  Nonsinge.for:12-121 */
int orc_nonsinge(const real co[4][3], const real xp[3], const real eta[3],
                 real c3, real c4, const real *gi, const real *ome,
                 int npg, real *helem)
{
    real t[9], xg[3], w;
    for (int k = 0; k < 9; ++k) helem[k] = 0.0;
    for (int jg = 0; jg < npg; ++jg)
        for (int ig = 0; ig < npg; ++ig) {
            if (quad_point(co, gi[ig], ome[ig], gi[jg], ome[jg], xg, &w))
                return ORC_DEGENERATE;
            orc_solfune(xp, xg, eta, c3, c4, t);
            for (int k = 0; k < 9; ++k) helem[k] = helem[k] + t[k] * w;
        }
    return ORC_OK;
}

/* NPG x NPG Gauss of U over a quad.  Singge.for:83-173 (NONSINGE_G) */
static int nonsinge_g(const real co[4][3], const real xp[3], real c1,
                      real c2, const real *gi, const real *ome, int npg,
                      real *gelem)
{
    real u[9], xg[3], w;
    for (int k = 0; k < 9; ++k) gelem[k] = 0.0;
    for (int jg = 0; jg < npg; ++jg)
        for (int ig = 0; ig < npg; ++ig) {
            if (quad_point(co, gi[ig], ome[ig], gi[jg], ome[jg], xg, &w))
                return ORC_DEGENERATE;
            orc_solfune_g(xp, xg, c1, c2, u);
            for (int k = 0; k < 9; ++k) gelem[k] = gelem[k] + u[k] * w;
        }
    return ORC_OK;
}

/* Weakly singular static G diagonal block: four quads collapsed onto the
 * centroid, edges (N1,N2),(N2,N3),(N3,N4),(N4,N1).  Singge.for:11-81 */
int orc_singge(const real corners[4][3], const real xm[3], real c1,
               real c2, const real *gi, const real *ome, int npg,
               real *gelem)
{
    real co[4][3], gp[9];
    for (int k = 0; k < 9; ++k) gelem[k] = 0.0;
    for (int d = 0; d < 3; ++d) co[2][d] = co[3][d] = xm[d];
    for (int it = 0; it < 4; ++it) {
        int a = it, b = (it + 1) % 4;
        for (int d = 0; d < 3; ++d) {
            co[0][d] = corners[a][d];
            co[1][d] = corners[b][d];
        }
        int st = nonsinge_g(co, xm, c1, c2, gi, ome, npg, gp);
        if (st) return st;
        for (int k = 0; k < 9; ++k) gelem[k] = gelem[k] + gp[k];
    }
    return ORC_OK;
}

/* NPG x NPG Gauss of the dynamic kernels.  Nonsingd.for:45-118 */
int orc_nonsingd(const real co[4][3], const real xp[3], const real eta[3],
                 zc zge, zc zcs, zc zcp, real pi, real fr,
                 const real *gi, const real *ome, int npg, zc *zhelem,
                 zc *zgelem)
{
    zc zu[9], zt[9];
    real xg[3], w;
    for (int k = 0; k < 9; ++k) zhelem[k] = zgelem[k] = 0.0;
    for (int jg = 0; jg < npg; ++jg)
        for (int ig = 0; ig < npg; ++ig) {
            if (quad_point(co, gi[ig], ome[ig], gi[jg], ome[jg], xg, &w))
                return ORC_DEGENERATE;
            orc_solfund(xp, xg, eta, zge, zcs, zcp, pi, fr, zu, zt);
            for (int k = 0; k < 9; ++k) {
                zhelem[k] = zhelem[k] + zt[k] * w;
                zgelem[k] = zgelem[k] + zu[k] * w;
            }
        }
    return ORC_OK;
}

/* Diagonal block by plain Gauss of the difference kernel.  Sing_de.for:13-113 */
int orc_sing_de(const real co[4][3], const real xp[3], const real eta[3],
                zc zge, zc zcs, zc zcp, real c1, real c2, real c3,
                real c4, real pi, real fr, const real *gi,
                const real *ome, int npg, zc *zhdif, zc *zgdif)
{
    zc zu[9], zt[9];
    real xg[3], w;
    for (int k = 0; k < 9; ++k) zhdif[k] = zgdif[k] = 0.0;
    for (int jg = 0; jg < npg; ++jg)
        for (int ig = 0; ig < npg; ++ig) {
            if (quad_point(co, gi[ig], ome[ig], gi[jg], ome[jg], xg, &w))
                return ORC_DEGENERATE;
            orc_solfundif(xp, xg, eta, zge, zcs, zcp, c1, c2, c3, c4, pi, fr,
                          zu, zt);
            for (int k = 0; k < 9; ++k) {
                zhdif[k] = zhdif[k] + zt[k] * w;
                zgdif[k] = zgdif[k] + zu[k] * w;
            }
        }
    return ORC_OK;
}

/* Interior stress kernels ZD(3,3,3), ZS(3,3,3).  Sigmaec.for:9-212 */
int orc_sigmaec(const real co[4][3], const real xp[3], const real rn[3],
                real pi, real fr, zc zge, real rho, zc zcs, zc zcp,
                const real *gi, const real *ome, int npg, zc *zd, zc *zs)
{
    real xg[3], w;
    for (int k = 0; k < 27; ++k) zd[k] = zs[k] = 0.0;
    for (int jg = 0; jg < npg; ++jg)
        for (int ig = 0; ig < npg; ++ig) {
            if (quad_point(co, gi[ig], ome[ig], gi[jg], ome[jg], xg, &w))
                return ORC_DEGENERATE;
            real r1 = xg[0] - xp[0], r2 = xg[1] - xp[1], r3 = xg[2] - xp[2];
            real r = sqrt(r1 * r1 + r2 * r2 + r3 * r3);
            real drn = (r1 * rn[0] + r2 * rn[1] + r3 * rn[2]) / r;
            real rd[3] = {r1 / r, r2 / r, r3 / r};
            struct dyn_scalars s;
            dyn_scalars(r, zge, zcs, zcp, pi, fr, &s);
            /* Fedelinski form, Sigmaec.for:139-176 */
            zc zwi = ORC_CMPLX(0.0, fr);
            zc zk1 = zwi / zcp, zk2 = zwi / zcs;
            zc zek1 = exp(-zk1 * r), zek2 = exp(-zk2 * r);
            zc zmu = zge;
            zc zlamb = (zcp * zcp * rho) - (2.0 * zmu);
            zc q1 = zk1 * r, q2 = zk2 * r;
            zc q1_2 = q1 * q1, q2_2 = q2 * q2;
            zc csp2 = (zcs / zcp) * (zcs / zcp);
            real r3c = r * r * r;
            zc fhidrr = -csp2 * zek1 * (12.0 + 12.0 * q1 + 5.0 * q1_2 + q1_2 * q1) *
                            (1.0 / (q1_2 * r3c)) +
                        zek2 *
                            (12.0 + 12.0 * q2 + 7.0 * q2_2 + 3.0 * (q2_2 * q2) +
                             q2_2 * q2_2) *
                            (1.0 / (q2_2 * r3c));
            zc cappadrr =
                -csp2 * zek1 *
                    (36.0 + 36.0 * q1 + 17.0 * q1_2 + 5.0 * (q1_2 * q1) +
                     q1_2 * q1_2) *
                    (1.0 / (q1_2 * r3c)) +
                zek2 *
                    (36.0 + 36.0 * q2 + 17.0 * q2_2 + 5.0 * (q2_2 * q2) +
                     q2_2 * q2_2) *
                    (1.0 / (q2_2 * r3c));
            zc cappa = s.cappa, fhidr = s.fhidr, cappadr = s.cappadr;
            real rr = r * r;
            zc lm = zlamb / zmu;
            zc cte1 = 2.0 * (cappa / r);
            zc cte2 = lm * (fhidr - cappadr - cte1);
            zc cte3 = (fhidr - cappa / r);
            zc cte4 = 2.0 * (cappadr - cte1);
            zc cte5 = 4.0 * (cappadrr - 5.0 * (cappadr / r) + 8.0 * (cappa / rr));
            zc cte6 = fhidrr - (fhidr / r) - 3.0 * (cappadr / r) + 6.0 * (cappa / rr);
            zc cte7 = 2.0 * (cappadr / r) - 4.0 * (cappa / rr);
            zc cte8 = lm * (cappadrr + (cappadr / r) - 4.0 * (cappa / rr) - fhidrr +
                            (fhidr / r));
            zc cte9 = 2.0 * (cte7 + cte8);
            zc cte10 = 4.0 * (cappa / rr) +
                       4.0 * lm * ((cappadr / r) + 2.0 * (cappa / rr) - (fhidr / r));
            zc cte11 = (lm * lm) * (cappadrr + 4.0 * (cappadr / r) +
                                    2.0 * (cappa / rr) - fhidrr - 2.0 * (fhidr / r));
            zc cte12 = 2.0 * ((fhidr / r) - (cappa / rr));
            real q4pi = 1.0 / (4.0 * pi);
            for (int k = 0; k < 3; ++k)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        real dji = (j == i), djk = (j == k), dik = (i == k);
                        zc dp = q4pi * ((cte1 - cte2) * dji * rd[k] -
                                        cte3 * (djk * rd[i] + dik * rd[j]) +
                                        cte4 * rd[j] * rd[i] * rd[k]);
                        zc sp = (zmu / (4.0 * pi)) *
                                ((cte5 * rd[j] * rd[i] * rd[k] -
                                  cte6 * (djk * rd[i] + dik * rd[j]) +
                                  cte9 * dji * rd[k]) *
                                     drn +
                                 cte9 * rd[j] * rd[i] * rn[k] -
                                 cte6 * (rd[i] * rn[j] + rd[j] * rn[i]) * rd[k] +
                                 (cte10 + cte11) * dji * rn[k] -
                                 cte12 * (dik * rn[j] + djk * rn[i]));
                        int idx = k + 3 * i + 9 * j;
                        zd[idx] = zd[idx] + dp * w;
                        zs[idx] = zs[idx] + sp * w;
                    }
        }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Assembly routines                                                   */
/* ------------------------------------------------------------------ */

/* Mesh-wide geometry (centroids, normals, Gauss rule) for one call. */
struct orc_mesh {
    int n, nbe, np, npg;
    const real *cx, *cy, *cz;
    const int *cone;
    real *cxm, *cym, *czm, *etas, *gi, *ome;
};

static int mesh_setup(struct orc_mesh *m, int n, int nbe, int np, int npg,
                      const real *cx, const real *cy, const real *cz,
                      const int *cone)
{
    if (n <= 0 || nbe <= 0 || nbe > n || np <= 0 || npg <= 0) return ORC_BADARG;
    m->n = n;
    m->nbe = nbe;
    m->np = np;
    m->npg = npg;
    m->cx = cx;
    m->cy = cy;
    m->cz = cz;
    m->cone = cone;
    m->cxm = malloc(sizeof(real) * n);
    m->cym = malloc(sizeof(real) * n);
    m->czm = malloc(sizeof(real) * n);
    m->etas = malloc(sizeof(real) * 3 * n);
    m->gi = malloc(sizeof(real) * npg);
    m->ome = malloc(sizeof(real) * npg);
    orc_gauleg(npg, m->gi, m->ome);
    orc_normvec(n, np, cone, cx, cy, cz, m->etas);
    orc_centroids(n, cone, cx, cy, cz, m->cxm, m->cym, m->czm);
    return ORC_OK;
}

static void mesh_free(struct orc_mesh *m)
{
    free(m->cxm);
    free(m->cym);
    free(m->czm);
    free(m->etas);
    free(m->gi);
    free(m->ome);
}

/* Static constants.  Ghmatece.for:55-72 */
void orc_static_constants(real ge, real rnu, real *consts)
{
    real pi = 4.0 * atan(1.0);
    real rmu = ge; /* Inputece.for:49 */
    consts[0] = pi;
    consts[1] = 1.0 / (16.0 * pi * rmu * (1.0 - rnu));
    consts[2] = 3.0 - (4.0 * rnu);
    consts[3] = -1.0 / (8.0 * pi * (1.0 - rnu));
    consts[4] = 1.0 - (2.0 * rnu);
}

/* Material: zge = G(1+2i DAM), zcs = sqrt(zge/rho),
 * zcp = zcs sqrt((2-2nu)/(1-2nu)).  Inputecd.for:58-60 */
void orc_material(real ge, real rnu, real dam, real rho, zc *out3)
{
    zc zge = ORC_CMPLX(ge, ge * 2.0 * dam);
    zc zcs = sqrt(zge / rho);
    zc zcp = zcs * sqrt((2.0 - 2.0 * rnu) / (1.0 - 2.0 * rnu));
    out3[0] = zge;
    out3[1] = zcs;
    out3[2] = zcp;
}

/* GHMATECE: static rigid-body H diagonal and weakly singular G diagonal.
 * Ghmatece.for:9-256 (off-diagonal loop :92-150, RIGID_BODY :258-282,
 * GEST_SINGULAR :288-323).  consts out: pi, c1, c2, c3, c4.
 * If hest != NULL it receives the full static H (3NBE x 3N, col-major). */
int orc_ghmatece(int n, int nbe, int np, int npg, const real *cx,
                 const real *cy, const real *cz, const int *cone,
                 real ge, real rnu, real *hestdiag, real *gestdiag,
                 real *consts, real *hest_out)
{
    struct orc_mesh m;
    int st = mesh_setup(&m, n, nbe, np, npg, cx, cy, cz, cone);
    if (st) return st;
    orc_static_constants(ge, rnu, consts);
    real c1 = consts[1], c2 = consts[2], c3 = consts[3], c4 = consts[4];
    size_t ld = (size_t)3 * nbe;
    real *hest = calloc(ld * 3 * (size_t)n, sizeof(real));
    int err = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
    for (int j = 0; j < n; ++j) {
        real co[4][3], helem[9];
        element_corners(n, cone, cx, cy, cz, j, co);
        for (int i = 0; i < nbe; ++i) {
            if (i == j) continue;
            real xp[3] = {m.cxm[i], m.cym[i], m.czm[i]};
            err |= orc_nonsinge(co, xp, &m.etas[3 * j], c3, c4, m.gi, m.ome, npg,
                                helem);
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a)
                    hest[(size_t)(3 * j + b) * ld + 3 * i + a] = helem[a + 3 * b];
        }
    }
    if (!err) {
        /* RIGID_BODY: HESTdiag(:,:,I) = -sum_{J=1..N} HEST(I,J), J ascending */
#pragma omp parallel for
        for (int i = 0; i < nbe; ++i) {
            real acc[9] = {0};
            for (int j = 0; j < n; ++j)
                for (int b = 0; b < 3; ++b)
                    for (int a = 0; a < 3; ++a)
                        acc[a + 3 * b] =
                            acc[a + 3 * b] - hest[(size_t)(3 * j + b) * ld + 3 * i + a];
            memcpy(&hestdiag[9 * i], acc, sizeof acc);
        }
        /* GEST_SINGULAR */
#pragma omp parallel for reduction(| : err)
        for (int j = 0; j < nbe; ++j) {
            real co[4][3];
            element_corners(n, cone, cx, cy, cz, j, co);
            real xm[3] = {m.cxm[j], m.cym[j], m.czm[j]};
            err |= orc_singge(co, xm, c1, c2, m.gi, m.ome, npg, &gestdiag[9 * j]);
        }
    }
    if (hest_out && !err) memcpy(hest_out, hest, sizeof(real) * ld * 3 * n);
    free(hest);
    mesh_free(&m);
    return err ? ORC_DEGENERATE : ORC_OK;
}

/* GHMATECD: dynamic ZH, ZG; KODE column swap; ZFI = ZG ZDFI.
 * Ghmatecd.for:11-239 (assembly :113-166, BOUNDARY_CONDITIONS call :168
 * with the actual arguments swapped relative to :316, ZGEMV :172-181).
 * zg_out (optional) receives ZG after the swap. */
int orc_ghmatecd(int n, int nbe, int np, int npg, const real *cx,
                 const real *cy, const real *cz, const int *cone,
                 const real *hestdiag, const real *gestdiag, const int *kode,
                 const zc *zdfi, real fr, zc zge, zc zcs, zc zcp,
                 const real *consts, zc *zh, zc *zfi, zc *zg_out)
{
    struct orc_mesh m;
    int st = mesh_setup(&m, n, nbe, np, npg, cx, cy, cz, cone);
    if (st) return st;
    real pi = consts[0], c1 = consts[1], c2 = consts[2], c3 = consts[3],
           c4 = consts[4];
    size_t nn = (size_t)3 * nbe;
    zc *zg = zg_out ? zg_out : malloc(sizeof(zc) * nn * nn);
    int err = 0;
#pragma omp parallel for schedule(dynamic, 2) reduction(| : err)
    for (int j = 0; j < nbe; ++j) {
        real co[4][3];
        zc hb[9], gb[9];
        element_corners(n, cone, cx, cy, cz, j, co);
        for (int i = 0; i < nbe; ++i) {
            real xp[3] = {m.cxm[i], m.cym[i], m.czm[i]};
            if (i == j) {
                err |= orc_sing_de(co, xp, &m.etas[3 * j], zge, zcs, zcp, c1, c2,
                                   c3, c4, pi, fr, m.gi, m.ome, npg, hb, gb);
                for (int k = 0; k < 9; ++k) {
                    gb[k] = gb[k] + gestdiag[9 * i + k];
                    hb[k] = hb[k] + hestdiag[9 * i + k];
                }
            } else {
                err |= orc_nonsingd(co, xp, &m.etas[3 * j], zge, zcs, zcp, pi, fr,
                                    m.gi, m.ome, npg, hb, gb);
            }
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a) {
                    size_t idx = (size_t)(3 * j + b) * nn + 3 * i + a;
                    zh[idx] = hb[a + 3 * b];
                    zg[idx] = gb[a + 3 * b];
                }
        }
    }
    if (!err) {
        /* KODE = 0 column: tmp = ZH*mu*; ZH = -ZG; ZG = -tmp */
        for (size_t c = 0; c < nn; ++c) {
            if (kode[c] != 0) continue;
            for (size_t r = 0; r < nn; ++r) {
                zc zch = zh[c * nn + r] * zge;
                zh[c * nn + r] = -zg[c * nn + r];
                zg[c * nn + r] = -zch;
            }
        }
        /* ZFI = ZG * ZDFI, column order */
        for (size_t r = 0; r < nn; ++r) zfi[r] = 0.0;
        for (size_t c = 0; c < nn; ++c) {
            zc x = zdfi[c];
            if (x == 0.0) continue;
            for (size_t r = 0; r < nn; ++r) zfi[r] += zg[c * nn + r] * x;
        }
    }
    if (!zg_out) free(zg);
    mesh_free(&m);
    return err ? ORC_DEGENERATE : ORC_OK;
}

/* Selected 3x3 blocks (bi[k], bj[k]) of ZH and ZG before the KODE swap,
 * each formed exactly as orc_ghmatecd forms it (Ghmatecd.for:113-166):
 * NONSINGD off the diagonal, SING_DE + HESTdiag/GESTdiag on it.  Blocks come
 * back column-major (a + 3 b), 9 per block. */
int orc_ghmatecd_blocks(int n, int nbe, int np, int npg, const real *cx,
                        const real *cy, const real *cz, const int *cone,
                        const real *hestdiag, const real *gestdiag, real fr,
                        zc zge, zc zcs, zc zcp, const real *consts, int nblk,
                        const int *bi, const int *bj, zc *hblk, zc *gblk)
{
    struct orc_mesh m;
    int st = mesh_setup(&m, n, nbe, np, npg, cx, cy, cz, cone);
    if (st) return st;
    real pi = consts[0], c1 = consts[1], c2 = consts[2], c3 = consts[3],
           c4 = consts[4];
    int err = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(| : err)
    for (int k = 0; k < nblk; ++k) {
        const int i = bi[k], j = bj[k];
        if (i < 0 || i >= nbe || j < 0 || j >= nbe) {
            err |= ORC_DEGENERATE;
            continue;
        }
        real co[4][3];
        element_corners(n, cone, cx, cy, cz, j, co);
        real xp[3] = {m.cxm[i], m.cym[i], m.czm[i]};
        zc *hb = hblk + 9 * (size_t)k, *gb = gblk + 9 * (size_t)k;
        if (i == j) {
            err |= orc_sing_de(co, xp, &m.etas[3 * j], zge, zcs, zcp, c1, c2, c3, c4,
                               pi, fr, m.gi, m.ome, npg, hb, gb);
            for (int q = 0; q < 9; ++q) {
                gb[q] = gb[q] + gestdiag[9 * i + q];
                hb[q] = hb[q] + hestdiag[9 * i + q];
            }
        } else {
            err |= orc_nonsingd(co, xp, &m.etas[3 * j], zge, zcs, zcp, pi, fr, m.gi,
                                m.ome, npg, hb, gb);
        }
    }
    mesh_free(&m);
    return err ? ORC_DEGENERATE : ORC_OK;
}

/* INTEREC: KODE un-swap, interior displacements ZDSOL(3L) and stresses
 * ZSSOL(9L).  Interec.for:59-67 (un-swap), :86-135 (ZDSOL), :188-283
 * (ZSSOL, entry e = I + 3(J-1) of ZD(K,I,J)).  zfi and zdfi are updated in
 * place.  Points are summed over elements J in ascending order. */
int orc_interec(int n, int nbe, int np, int npg, const real *cx,
                const real *cy, const real *cz, const int *cone, zc *zfi,
                zc *zdfi, const int *kode, int l, const real *cxi,
                const real *cyi, const real *czi, zc zge, zc zcs, zc zcp,
                real fr, real rho, real pi, zc *zdsol, zc *zssol)
{
    struct orc_mesh m;
    int st = mesh_setup(&m, n, nbe, np, npg, cx, cy, cz, cone);
    if (st) return st;
    int nn = 3 * nbe;
    for (int i = 0; i < nn; ++i) {
        if (kode[i] == 0) {
            zc zch = zfi[i] * zge;
            zfi[i] = zdfi[i];
            zdfi[i] = zch;
        }
    }
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int k = 0; k < l; ++k) {
        real xp[3] = {cxi[k], cyi[k], czi[k]};
        zc du[3] = {0, 0, 0}, sg[9] = {0};
        zc hb[9], gb[9], zd[27], zs[27];
        for (int j = 0; j < nbe; ++j) {
            real co[4][3];
            element_corners(n, cone, cx, cy, cz, j, co);
            err |= orc_nonsingd(co, xp, &m.etas[3 * j], zge, zcs, zcp, pi, fr,
                                m.gi, m.ome, npg, hb, gb);
            const zc *t = &zdfi[3 * j], *u = &zfi[3 * j];
            for (int a = 0; a < 3; ++a) {
                zc gt = gb[a] * t[0] + gb[a + 3] * t[1] + gb[a + 6] * t[2];
                zc hu = hb[a] * u[0] + hb[a + 3] * u[1] + hb[a + 6] * u[2];
                du[a] = du[a] + (gt - hu);
            }
            err |= orc_sigmaec(co, xp, &m.etas[3 * j], pi, fr, zge, rho, zcs, zcp,
                               m.gi, m.ome, npg, zd, zs);
            for (int jj = 0; jj < 3; ++jj)
                for (int ii = 0; ii < 3; ++ii) {
                    int e = ii + 3 * jj;
                    zc v = sg[e];
                    for (int q = 0; q < 3; ++q) v = v + t[q] * zd[q + 3 * ii + 9 * jj];
                    for (int q = 0; q < 3; ++q) v = v - u[q] * zs[q + 3 * ii + 9 * jj];
                    sg[e] = v;
                }
        }
        for (int a = 0; a < 3; ++a) zdsol[3 * k + a] = du[a];
        for (int e = 0; e < 9; ++e) zssol[9 * k + e] = sg[e];
    }
    mesh_free(&m);
    return err ? ORC_DEGENERATE : ORC_OK;
}

/* Plain partial-pivot complex LU + solve (ZGESV semantics, one RHS), used
 * as an independent in-oracle check of the LAPACK solve on small systems.
 * Linsolve.for:68-109.  Returns info (0 ok, k>0: zero pivot at column k). */
int orc_zgesv(int nn, zc *a, zc *b)
{
    for (int k = 0; k < nn; ++k) {
        int p = k;
        real best = fabs(a[(size_t)k * nn + k]);
        for (int r = k + 1; r < nn; ++r) {
            real v = fabs(a[(size_t)k * nn + r]);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        if (best == 0.0) return k + 1;
        if (p != k) {
            for (int c = 0; c < nn; ++c) {
                zc t = a[(size_t)c * nn + k];
                a[(size_t)c * nn + k] = a[(size_t)c * nn + p];
                a[(size_t)c * nn + p] = t;
            }
            zc t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
        zc piv = a[(size_t)k * nn + k];
        for (int r = k + 1; r < nn; ++r) a[(size_t)k * nn + r] /= piv;
        for (int c = k + 1; c < nn; ++c) {
            zc f = a[(size_t)c * nn + k];
            if (f == 0.0) continue;
            for (int r = k + 1; r < nn; ++r)
                a[(size_t)c * nn + r] -= a[(size_t)k * nn + r] * f;
        }
    }
    for (int k = 0; k < nn; ++k)
        for (int r = k + 1; r < nn; ++r) b[r] -= a[(size_t)k * nn + r] * b[k];
    for (int k = nn - 1; k >= 0; --k) {
        b[k] /= a[(size_t)k * nn + k];
        for (int r = 0; r < k; ++r) b[r] -= a[(size_t)k * nn + r] * b[k];
    }
    return 0;
}

/* The NPG x NPG Gauss points and weights of one element, in the loop order
 * of Nonsingd.for:108-119 (JG outer, IG inner).  Exposed for tests. */
int orc_quad_rule(const real co[4][3], const real *gi, const real *ome,
                  int npg, real *xg, real *w)
{
    for (int jg = 0; jg < npg; ++jg)
        for (int ig = 0; ig < npg; ++ig) {
            int q = jg * npg + ig;
            if (quad_point(co, gi[ig], ome[ig], gi[jg], ome[jg], &xg[3 * q], &w[q]))
                return ORC_DEGENERATE;
        }
    return ORC_OK;
}

#ifdef ORC_LONG_DOUBLE
/* ------------------------------------------------------------------ */
/* Entry points of the extended-precision build: the same inputs as    */
/* the FP64 build (double arrays; complex scalars as (re, im) pairs),  */
/* every operation in long double, results rounded to double.          */
/* ------------------------------------------------------------------ */
static real *ld_copy(const double *a, size_t n)
{
    real *r = malloc(sizeof(real) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) r[i] = a[i];
    return r;
}

static zc ld_z(const double *p) { return ORC_CMPLX(p[0], p[1]); }

int orcld_ghmatece(int n, int nbe, int np, int npg, const double *cx,
                   const double *cy, const double *cz, const int *cone, double ge,
                   double rnu, double *hestdiag, double *gestdiag, double *consts)
{
    real *x = ld_copy(cx, np), *y = ld_copy(cy, np), *z = ld_copy(cz, np);
    real *hd = malloc(sizeof(real) * 9 * nbe), *gd = malloc(sizeof(real) * 9 * nbe);
    real c[5];
    int st = orc_ghmatece(n, nbe, np, npg, x, y, z, cone, ge, rnu, hd, gd, c, NULL);
    for (int i = 0; i < 9 * nbe; ++i) {
        hestdiag[i] = (double)hd[i];
        gestdiag[i] = (double)gd[i];
    }
    for (int i = 0; i < 5; ++i) consts[i] = (double)c[i];
    free(x); free(y); free(z); free(hd); free(gd);
    return st;
}

/* zmat: zge, zcs, zcp as 6 doubles; zdfi, zh, zfi interleaved complex. */
int orcld_ghmatecd(int n, int nbe, int np, int npg, const double *cx,
                   const double *cy, const double *cz, const int *cone,
                   const double *hestdiag, const double *gestdiag, const int *kode,
                   const double *zdfi, double fr, const double *zmat,
                   const double *consts, double *zh, double *zfi)
{
    size_t nn = (size_t)3 * nbe;
    real *x = ld_copy(cx, np), *y = ld_copy(cy, np), *z = ld_copy(cz, np);
    real *hd = ld_copy(hestdiag, 9 * (size_t)nbe), *gd = ld_copy(gestdiag, 9 * (size_t)nbe);
    real *c = ld_copy(consts, 5);
    zc *zd = malloc(sizeof(zc) * nn), *h = malloc(sizeof(zc) * nn * nn),
       *f = malloc(sizeof(zc) * nn);
    for (size_t i = 0; i < nn; ++i) zd[i] = ld_z(zdfi + 2 * i);
    int st = orc_ghmatecd(n, nbe, np, npg, x, y, z, cone, hd, gd, kode, zd, fr,
                          ld_z(zmat), ld_z(zmat + 2), ld_z(zmat + 4), c, h, f, NULL);
    for (size_t i = 0; i < nn * nn; ++i) {
        zh[2 * i] = (double)creal(h[i]);
        zh[2 * i + 1] = (double)cimag(h[i]);
    }
    for (size_t i = 0; i < nn; ++i) {
        zfi[2 * i] = (double)creal(f[i]);
        zfi[2 * i + 1] = (double)cimag(f[i]);
    }
    free(x); free(y); free(z); free(hd); free(gd); free(c); free(zd); free(h); free(f);
    return st;
}

int orcld_ghmatecd_blocks(int n, int nbe, int np, int npg, const double *cx,
                          const double *cy, const double *cz, const int *cone,
                          const double *hestdiag, const double *gestdiag, double fr,
                          const double *zmat, const double *consts, int nblk,
                          const int *bi, const int *bj, double *hblk, double *gblk)
{
    real *x = ld_copy(cx, np), *y = ld_copy(cy, np), *z = ld_copy(cz, np);
    real *hd = ld_copy(hestdiag, 9 * (size_t)nbe), *gd = ld_copy(gestdiag, 9 * (size_t)nbe);
    real *c = ld_copy(consts, 5);
    zc *hb = malloc(sizeof(zc) * 9 * (size_t)(nblk ? nblk : 1)),
       *gb = malloc(sizeof(zc) * 9 * (size_t)(nblk ? nblk : 1));
    int st = orc_ghmatecd_blocks(n, nbe, np, npg, x, y, z, cone, hd, gd, fr, ld_z(zmat),
                                 ld_z(zmat + 2), ld_z(zmat + 4), c, nblk, bi, bj, hb, gb);
    for (size_t i = 0; i < 9 * (size_t)nblk; ++i) {
        hblk[2 * i] = (double)creal(hb[i]);
        hblk[2 * i + 1] = (double)cimag(hb[i]);
        gblk[2 * i] = (double)creal(gb[i]);
        gblk[2 * i + 1] = (double)cimag(gb[i]);
    }
    free(x); free(y); free(z); free(hd); free(gd); free(c); free(hb); free(gb);
    return st;
}

/* Point kernels ZU, ZT (Solfund.for:7-63) in extended precision. */
void orcld_solfund(const double xp[3], const double xg[3], const double rn[3],
                   const double *zmat, double pi, double fr, double *zu, double *zt)
{
    real p[3] = {xp[0], xp[1], xp[2]}, g[3] = {xg[0], xg[1], xg[2]},
         e[3] = {rn[0], rn[1], rn[2]};
    zc u[9], t[9];
    orc_solfund(p, g, e, ld_z(zmat), ld_z(zmat + 2), ld_z(zmat + 4), pi, fr, u, t);
    for (int k = 0; k < 9; ++k) {
        zu[2 * k] = (double)creal(u[k]);
        zu[2 * k + 1] = (double)cimag(u[k]);
        zt[2 * k] = (double)creal(t[k]);
        zt[2 * k + 1] = (double)cimag(t[k]);
    }
}
#endif /* ORC_LONG_DOUBLE */
