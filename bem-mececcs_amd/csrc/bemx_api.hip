// bemx_api.hip -- host side of libbemx.so: the device context, the C ABI of
// include/bemx.h and the orchestration of the assembly, solve and interior
// kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/bemx.h"
#include "bemx_common.hpp"
#include "bemx_lu.hpp"

using namespace bemx;

namespace bemx {
hipError_t launch_element_setup(hipStream_t, int, const int*, const double*, const double*,
                                const double*, const double*, const double*, int, int,
                                ElemData*, GaussPt*, GaussPt*, int*);
hipError_t launch_static_hdiag(hipStream_t, int, int, int, const ElemData*, const GaussPt*,
                               double, double, double*, int, double*);
hipError_t launch_gest_diag(hipStream_t, int, int, const int*, const double*, const double*,
                            const double*, const ElemData*, const double*, const double*, int,
                            double, double, double*, int*);
hipError_t launch_dyn_assembly(hipStream_t, int, int, const ElemData*, const GaussPt*,
                               const double*, const double*, const int*, const cd*, int, int*,
                               int*, const DynConst&, cd*, int, double, cd*, size_t, cd*, cd*,
                               cd*, hipEvent_t, hipEvent_t);
hipError_t launch_dyn_assembly_pair(hipStream_t, int, int, const ElemData*, const GaussPt*,
                                    const double*, const double*, const int*, const cd*, int,
                                    int*, int*, const DynFreq (&)[2], size_t, hipEvent_t,
                                    hipEvent_t);
int offdiag_tiles(int nbe);
size_t exp_table_max_bytes();
int exp_table_entries(const DynConst& K, double rmax, double* h);
hipError_t launch_unswap(hipStream_t, int, const int*, cd*, cd*, cd);
int interior_blocks(int nbe);
int batch_size();
hipError_t launch_interior(hipStream_t, int, int, int, int, const ElemData*, const GaussPt*,
                           const double*, const double*, const double*, const cd*, const cd*,
                           const DynConst&, cd*, cd*, cd*);
}  // namespace bemx

constexpr int NLAST = 11;  // doubles of bemx_last_stats

// A grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct bemx_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;
    // mesh
    int n = 0, nbe = 0, np = 0, npg = 0;
    bool mesh_ok = false;
    DevBuf cone, cx, cy, cz, gi, ome, elem, gp, gpT, status;
    double rmax = 0.0;  // diagonal of the box holding every node and collocation point
    // static pass
    bool static_ok = false;
    double consts[5] = {0, 0, 0, 0, 0};  // pi, c1..c4
    DevBuf hdiag, gdiag, hpart;
    // dynamic data
    bool dyn_ok = false;
    DevBuf kode, zdfi, colflag, flist, rhs_part, rhs_diag, zfi, zh, etab;
    int nrhs = 1;              // load cases (columns of zdfi / zfi)
    // load cases with KODE columns of their own (bemx_send_load_cases_kode):
    // host copies, grouped by identical KODE column; empty when every case
    // shares one KODE (the device buffers then hold the cases directly)
    std::vector<double> case_zdfi;            // 6 NBE doubles per case
    std::vector<int> case_kode;               // 3 NBE per case
    std::vector<std::vector<int>> case_groups;  // case indices per KODE set
    bool zh_valid = false;     // zh holds the current assembled matrix
    const void* zh_host = nullptr;  // host array the last GHMATECD copied ZH to
    // interior
    DevBuf ipts, ipart, zdsol, zssol, ifi, idfi, ikode;
    // frequency sweep (bemx_solve_sweep): two halves of up to SWB frequency
    // slots each (one batch of frequencies is assembled into one half while
    // the other half's batch is solved); the slots' matrices and right-hand
    // sides are contiguous (a batch's solve takes them at one stride), the
    // assembly's stream and the halves' events
    static constexpr int SWB = bemx::MIXED_MAX_BATCH, NSW = 2 * SWB;
    DevBuf sw_zh, sw_zfi, sw_rhs_part[NSW], sw_rhs_diag[NSW], sw_etab[NSW], sw_zdfi[NSW];
    hipStream_t st_asm = nullptr;
    hipEvent_t sw_asm[2] = {nullptr, nullptr}, sw_done[2] = {nullptr, nullptr};
    double sw_factor_ms = 0.0;  // device time of the last batched solves' factorisations
    // solver
    bemx::LuWork lu;
    // ev0..ev1: whole solve; ev2..ev3: off-diagonal assembly kernel;
    // ev1 of assembly end is ev4
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr, ev4 = nullptr;
    double last[NLAST] = {0};
};

namespace {

int fail(bemx_ctx* c, int code, const char* fmt, const char* a = "", long b = 0) {
    char buf[512];
    snprintf(buf, sizeof buf, fmt, a, b);
    c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ctx, e_ == hipErrorOutOfMemory ? BEMX_ERR_NOMEM : BEMX_ERR_HIP, \
                        "%s (line %ld)", hipGetErrorString(e_), (long)__LINE__);       \
    } while (0)

int check_status(bemx_ctx* c) {
    int s = 0;
    HIPCHK(c, hipMemcpyAsync(&s, c->status.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (s & ST_DEGENERATE)
        return fail(c, BEMX_ERR_DEGENERATE, "%sdegenerate element: Jacobian below 1e-5", "", 0);
    if (s & ST_SYNC_TIMEOUT)
        return fail(c, BEMX_ERR_TIMEOUT, "%sdevice-side wait timed out", "", 0);
    return BEMX_OK;
}

// Run-time constants of the small-|z| series (see DynConst): r = cs/cp and
// its powers, and the residuals of the static limits (the m = 0 terms) that
// the diagonal blocks' difference kernels subtract, in long double on the
// host.  A R^2 = psi' R^2 - chi R, B R^2 = 4 chi R - 2 chi' R^2 and
// C R^2 = (1/r^2 - 2)(psi' R^2 - chi' R^2 - 2 chi R) - 2 chi R
// (Solfund.for:51-53 multiplied through by R^2).
void fill_series(DynConst& K) {
    const long double r2 = K.c2.x, r = std::sqrt(r2), c1m2 = 1.0L / r2 - 2.0L;
    auto f0 = [&](int f) { return (long double)ser_ab(f, 0, false) - ser_ab(f, 0, true) * r2; };
    const long double chi = f0(SER_CHI), psid = f0(SER_PSID), chid = f0(SER_CHID);
    const long double fa = psid - chi, fb = 4.0L * chi - 2.0L * chid;
    const long double fc = c1m2 * (psid - chid - 2.0L * chi) - 2.0L * chi;
    long double rt = r;
    for (int i = 1; i < NSER_OFF + 1; ++i) rt *= r;
    K.ser_rtop_off = (double)rt;
    for (int i = NSER_OFF + 1; i < NSER_DIAG + 1; ++i) rt *= r;
    K.ser_rtop_diag = (double)rt;
    K.ser_rinv = (double)(1.0L / r);
    K.ser_c1m2 = (double)c1m2;
    K.ks_abs2 = K.ks.x * K.ks.x + K.ks.y * K.ks.y;
    const long double q4pi = K.q4pi, c3 = K.c3s, c4 = K.c4s;
    K.t_a0 = (double)(q4pi * fa - c3 * c4);
    K.t_b0 = (double)(q4pi * fb - 3.0L * c3);
    K.t_c0 = (double)(q4pi * fc + c3 * c4);
    const cd fpsi = mk((double)f0(SER_PSI), 0.0), fchi = mk((double)chi, 0.0);
    K.u_d0 = K.c0 * fpsi - mk(K.c1s * K.c2s, 0.0);
    K.u_r0 = K.c0 * fchi + mk(K.c1s, 0.0);
}

DynConst make_dyn_const(const bemx_ctx* c, cd zge, cd zcs, cd zcp, double fr, double rho) {
    DynConst K;
    const double pi = c->consts[0];
    cd zwi = mk(0.0, fr);
    K.ks = -(zwi / zcs);
    K.kp = -(zwi / zcp);
    K.as = cinv(K.ks);
    K.ap = cinv(K.kp);
    K.as2 = K.as * K.as;
    K.ap2 = K.ap * K.ap;
    cd r = zcs / zcp;
    K.c2 = r * r;
    cd r1 = zcp / zcs;
    K.c1m2 = r1 * r1 - mk(2.0, 0.0);
    K.c0 = cinv(mk(4.0 * pi, 0.0) * zge);
    K.zge = zge;
    K.lm = ((zcp * zcp) * rho - 2.0 * zge) / zge;
    K.q4pi = 1.0 / (4.0 * pi);
    K.c1s = c->consts[1];
    K.c2s = c->consts[2];
    K.c3s = c->consts[3];
    K.c4s = c->consts[4];
    fill_series(K);
    return K;
}

cd ld_c(const double* p) { return mk(p[0], p[1]); }

}  // namespace

namespace bemx {

// Every device buffer of a context, the status word excepted.
static void ctx_buffers(bemx_ctx* c, std::vector<DevBuf*>& out) {
    out = {&c->cone, &c->cx, &c->cy, &c->cz, &c->gi, &c->ome, &c->elem, &c->gp, &c->gpT,
           &c->hdiag, &c->gdiag, &c->hpart, &c->kode, &c->zdfi, &c->colflag, &c->flist,
           &c->rhs_part, &c->rhs_diag, &c->zfi, &c->zh, &c->etab, &c->ipts, &c->ipart,
           &c->zdsol, &c->zssol, &c->ifi, &c->idfi, &c->ikode};
    out.push_back(&c->sw_zh);
    out.push_back(&c->sw_zfi);
    for (int i = 0; i < bemx_ctx::NSW; ++i)
        for (DevBuf* d : {&c->sw_rhs_part[i], &c->sw_rhs_diag[i], &c->sw_etab[i], &c->sw_zdfi[i]})
            out.push_back(d);
}

size_t ctx_device_bytes(bemx_ctx* c) {
    std::vector<DevBuf*> bufs;
    ctx_buffers(c, bufs);
    size_t b = c->status.p ? 256 : 0;
    for (DevBuf* d : bufs) b += d->p ? std::max<size_t>(d->cap, 256) : 0;
    return b + c->lu.bytes();
}

void ctx_release_buffers(bemx_ctx* c) {
    hipSetDevice(c->device);
    if (c->st) hipStreamSynchronize(c->st);
    if (c->st_asm) hipStreamSynchronize(c->st_asm);
    std::vector<DevBuf*> bufs;
    ctx_buffers(c, bufs);
    for (DevBuf* d : bufs) d->release();
    c->status.release();
    c->lu.release();
    c->mesh_ok = c->static_ok = c->dyn_ok = c->zh_valid = false;
    c->zh_host = nullptr;
}

void ctx_release_workspace(bemx_ctx* c) {
    ctx_release_buffers(c);
    // keep the status word: every entry point expects it
    (void)c->status.ensure(sizeof(int));
}

// Upper bound of the device bytes a context holds after serving meshes of at
// most these sizes through bemx_send_shared_data, bemx_ghmatece,
// bemx_send_load_cases, bemx_solve and batches of up to nb problems through
// ctx_solve_items: every buffer those calls grow (they only grow), at its
// largest.  The FP64 solver workspace is counted in both modes, since the
// mixed path falls back to it.  Monotone in each argument, so the bound over
// several requests is the bound at their maxima.
size_t ctx_solve_bytes_bound(int n, int nbe, int np, int npg, int nrhs, int mode, int nb) {
    auto r = [](size_t b) { return std::max<size_t>(b, 256); };
    const size_t nn = 3 * (size_t)nbe, npg2 = (size_t)npg * npg, rows = (nbe + 63) / 64;
    size_t b = r(sizeof(int));  // status word
    b += r(sizeof(int) * 4 * (size_t)n) + 3 * r(sizeof(double) * np) + 2 * r(sizeof(double) * npg) +
         r(sizeof(ElemData) * (size_t)n) + 2 * r(sizeof(GaussPt) * (size_t)n * npg2);
    // static diagonals; the split partials of bemx_ghmatece hold 9 nbe nsplit
    // doubles with nsplit <= 2048 / rows + 1 and nbe <= 64 rows
    b += 2 * r(sizeof(double) * 9 * nbe) + r(sizeof(double) * 9 * 64 * (2048 + rows));
    b += r(sizeof(cd) * nn * nrhs) + r(sizeof(int) * nn);
    b += r(sizeof(int) * nbe) + r(sizeof(int) * (nbe + 1)) +
         r(sizeof(cd) * nn * offdiag_tiles(nbe) * nrhs) + 2 * r(sizeof(cd) * nn * nrhs) +
         r(sizeof(cd) * nn * nn) + r(exp_table_max_bytes());
    b += fp64_factor_bytes((int)nn);
    nb = std::max(1, nb);
    const int nar = std::min(nb, batch_size());  // arenas of the mixed workspace
    if (nb > 1) {
        // the batch's frequency slots: matrices and right-hand sides in one
        // buffer each, partials, exponential tables and load cases per slot
        b += r(sizeof(cd) * nn * nn * nb) + r(sizeof(cd) * nn * nrhs * nb);
        b += nb * (r(sizeof(cd) * nn * offdiag_tiles(nbe) * nrhs) + r(sizeof(cd) * nn * nrhs) +
                   r(exp_table_max_bytes()) + r(sizeof(cd) * nn * nrhs));
    }
    if (mode == BEMX_MODE_MIXED) b += mixed_workspace_bytes((int)nn, nar);
    return b;
}

void ctx_graph_counts(const bemx_ctx* c, long long* captures, long long* replays,
                      long long* precaptures) {
    *captures = c->lu.captures;
    *replays = c->lu.replays;
    *precaptures = c->lu.precaptures;
}

void ctx_capture_ahead(bemx_ctx* c, int ahead) { c->lu.ahead = ahead; }

}  // namespace bemx

extern "C" {

int bemx_create(int device, bemx_ctx** out) {
    if (!out) return BEMX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return BEMX_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return BEMX_ERR_HIP;
    bemx_ctx* c = new bemx_ctx();
    c->device = device;
    // The creating thread also loads the device code of every kernel file,
    // so no other thread (a request-loop worker) makes the process's first
    // load of it.
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess ||
        hipEventCreate(&c->ev4) != hipSuccess || c->status.ensure(sizeof(int)) != hipSuccess ||
        preload_assembly_code() != hipSuccess || preload_lu_code() != hipSuccess ||
        preload_mixed_code() != hipSuccess || scratch_setup(device) != hipSuccess) {
        bemx_destroy(c);
        return BEMX_ERR_HIP;
    }
    *out = c;
    return BEMX_OK;
}

int bemx_destroy(bemx_ctx* c) {
    if (!c) return BEMX_ERR_ARG;
    hipSetDevice(c->device);
    if (c->st) hipStreamSynchronize(c->st);
    bemx::ctx_release_buffers(c);
    hipEvent_t evs[] = {c->ev0,     c->ev1,     c->ev2,      c->ev3,     c->ev4,
                        c->sw_asm[0], c->sw_asm[1], c->sw_done[0], c->sw_done[1]};
    for (hipEvent_t e : evs)
        if (e) hipEventDestroy(e);
    if (c->st_asm) hipStreamDestroy(c->st_asm);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
    return BEMX_OK;
}

const char* bemx_last_error(const bemx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int bemx_gauleg(int npg, double* x, double* w) {
    // Newton iteration on P_n from the Chebyshev-like initial guess, stopping
    // at |dz| <= 3e-14, as Gauleg.for:9-36.
    if (npg <= 0 || !x || !w) return BEMX_ERR_ARG;
    int m = (npg + 1) / 2;
    for (int i = 1; i <= m; ++i) {
        double z = std::cos(3.141592654 * (i - 0.25) / (npg + 0.5)), z1, pp, p1, p2, p3;
        do {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= npg; ++j) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = npg * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (std::fabs(z - z1) > 3e-14);
        x[i - 1] = -z;
        x[npg - i] = z;
        w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
        w[npg - i] = w[i - 1];
    }
    return BEMX_OK;
}

int bemx_send_shared_data(bemx_ctx* c, const double* cx, const double* cy, const double* cz,
                          const double* cxm, const double* cym, const double* czm,
                          const double* etas, const double* gi, const double* ome,
                          const int* cone, int np, int npg, int n, int nbe) {
    if (!c) return BEMX_ERR_ARG;
    if (!cx || !cy || !cz || !cone || np <= 0 || npg <= 0 || npg > 16 || n <= 0 || nbe <= 0 ||
        nbe > n)
        return fail(c, BEMX_ERR_ARG, "%sbad mesh sizes", "", 0);
    for (long k = 0; k < 4L * n; ++k)
        if (cone[k] < 1 || cone[k] > np)
            return fail(c, BEMX_ERR_ARG, "%sconnectivity entry %ld outside 1..NP", "", k);
    bool given = cxm && cym && czm && etas;
    HIPCHK(c, hipSetDevice(c->device));
    c->mesh_ok = c->static_ok = c->dyn_ok = c->zh_valid = false;
    c->n = n;
    c->nbe = nbe;
    c->np = np;
    c->npg = npg;
    {
        double lo[3], hi[3];
        const double* xyz[3] = {cx, cy, cz};
        for (int d = 0; d < 3; ++d) lo[d] = hi[d] = xyz[d][0];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < np; ++k) {
                lo[d] = std::min(lo[d], xyz[d][k]);
                hi[d] = std::max(hi[d], xyz[d][k]);
            }
        if (given) {
            const double* m[3] = {cxm, cym, czm};
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < n; ++k) {
                    lo[d] = std::min(lo[d], m[d][k]);
                    hi[d] = std::max(hi[d], m[d][k]);
                }
        }
        c->rmax = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                            (hi[2] - lo[2]) * (hi[2] - lo[2]));
    }
    int npg2 = npg * npg;
    std::vector<double> g(npg), w(npg);
    if (gi && ome) {
        g.assign(gi, gi + npg);
        w.assign(ome, ome + npg);
    } else {
        bemx_gauleg(npg, g.data(), w.data());
    }
    HIPCHK(c, c->cone.ensure(sizeof(int) * 4 * n));
    HIPCHK(c, c->cx.ensure(sizeof(double) * np));
    HIPCHK(c, c->cy.ensure(sizeof(double) * np));
    HIPCHK(c, c->cz.ensure(sizeof(double) * np));
    HIPCHK(c, c->gi.ensure(sizeof(double) * npg));
    HIPCHK(c, c->ome.ensure(sizeof(double) * npg));
    HIPCHK(c, c->elem.ensure(sizeof(ElemData) * n));
    HIPCHK(c, c->gp.ensure(sizeof(GaussPt) * (size_t)n * npg2));
    HIPCHK(c, c->gpT.ensure(sizeof(GaussPt) * (size_t)n * npg2));
    HIPCHK(c, hipMemcpyAsync(c->cone.p, cone, sizeof(int) * 4 * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->cx.p, cx, sizeof(double) * np, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->cy.p, cy, sizeof(double) * np, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->cz.p, cz, sizeof(double) * np, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->gi.p, g.data(), sizeof(double) * npg, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->ome.p, w.data(), sizeof(double) * npg, hipMemcpyHostToDevice, c->st));
    if (given) {
        std::vector<ElemData> e(n);
        for (int j = 0; j < n; ++j)
            e[j] = ElemData{etas[3 * j], etas[3 * j + 1], etas[3 * j + 2], cxm[j], cym[j], czm[j],
                            0.0, 0.0};
        HIPCHK(c, hipMemcpyAsync(c->elem.p, e.data(), sizeof(ElemData) * n,
                                 hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    HIPCHK(c, hipMemsetAsync(c->status.p, 0, sizeof(int), c->st));
    HIPCHK(c, launch_element_setup(c->st, n, c->cone.as<int>(), c->cx.as<double>(),
                                   c->cy.as<double>(), c->cz.as<double>(), c->gi.as<double>(),
                                   c->ome.as<double>(), npg, given ? 1 : 0, c->elem.as<ElemData>(),
                                   c->gp.as<GaussPt>(), c->gpT.as<GaussPt>(), c->status.as<int>()));
    int s = check_status(c);
    if (s) return s;
    c->mesh_ok = true;
    return BEMX_OK;
}

int bemx_deallocate_shared_data(bemx_ctx* c) {
    if (!c) return BEMX_ERR_ARG;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->st);
    DevBuf* bufs[] = {&c->cone, &c->cx, &c->cy, &c->cz, &c->gi, &c->ome, &c->elem, &c->gp,
                      &c->gpT, &c->hdiag, &c->gdiag, &c->hpart, &c->zh};
    for (DevBuf* b : bufs) b->release();
    c->mesh_ok = c->static_ok = c->zh_valid = false;
    return BEMX_OK;
}

int bemx_ghmatece(bemx_ctx* c, double ge, double rnu, double* hestdiag, double* gestdiag,
                  double* delta, double* pi, double* c1, double* c2, double* c3, double* c4) {
    if (!c) return BEMX_ERR_ARG;
    if (!c->mesh_ok) return fail(c, BEMX_ERR_STATE, "%sGHMATECE before the mesh was sent", "", 0);
    if (!(ge > 0.0) || !(rnu < 0.5))
        return fail(c, BEMX_ERR_ARG, "%sneeds G > 0 and nu < 1/2", "", 0);
    HIPCHK(c, hipSetDevice(c->device));
    // Ghmatece.for:55-63
    const double PI = 4.0 * std::atan(1.0);
    c->consts[0] = PI;
    c->consts[1] = 1.0 / (16.0 * PI * ge * (1.0 - rnu));
    c->consts[2] = 3.0 - (4.0 * rnu);
    c->consts[3] = -1.0 / (8.0 * PI * (1.0 - rnu));
    c->consts[4] = 1.0 - (2.0 * rnu);
    const int nbe = c->nbe, n = c->n, npg2 = c->npg * c->npg;
    // splits of the element range so that the grid holds >= 2048 waves
    int rows = (nbe + 63) / 64;
    int nsplit = std::max(1, std::min((2048 + rows - 1) / rows, (n + 15) / 16));
    HIPCHK(c, c->hdiag.ensure(sizeof(double) * 9 * nbe));
    HIPCHK(c, c->gdiag.ensure(sizeof(double) * 9 * nbe));
    HIPCHK(c, c->hpart.ensure(sizeof(double) * 9 * (size_t)nbe * nsplit));
    HIPCHK(c, hipMemsetAsync(c->status.p, 0, sizeof(int), c->st));
    HIPCHK(c, launch_static_hdiag(c->st, n, nbe, npg2, c->elem.as<ElemData>(),
                                  c->gp.as<GaussPt>(), c->consts[3], c->consts[4],
                                  c->hpart.as<double>(), nsplit, c->hdiag.as<double>()));
    HIPCHK(c, launch_gest_diag(c->st, n, nbe, c->cone.as<int>(), c->cx.as<double>(),
                               c->cy.as<double>(), c->cz.as<double>(), c->elem.as<ElemData>(),
                               c->gi.as<double>(), c->ome.as<double>(), c->npg, c->consts[1],
                               c->consts[2], c->gdiag.as<double>(), c->status.as<int>()));
    if (hestdiag)
        HIPCHK(c, hipMemcpyAsync(hestdiag, c->hdiag.p, sizeof(double) * 9 * nbe,
                                 hipMemcpyDeviceToHost, c->st));
    if (gestdiag)
        HIPCHK(c, hipMemcpyAsync(gestdiag, c->gdiag.p, sizeof(double) * 9 * nbe,
                                 hipMemcpyDeviceToHost, c->st));
    int s = check_status(c);
    if (s) return s;
    if (delta)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) delta[i + 3 * j] = (i == j) ? 1.0 : 0.0;
    if (pi) *pi = c->consts[0];
    if (c1) *c1 = c->consts[1];
    if (c2) *c2 = c->consts[2];
    if (c3) *c3 = c->consts[3];
    if (c4) *c4 = c->consts[4];
    c->static_ok = true;
    return BEMX_OK;
}

static int upload_cases(bemx_ctx* c, int nrhs, const double* zdfi, const int* kode) {
    size_t nn = 3 * (size_t)c->nbe;
    c->nrhs = nrhs;
    HIPCHK(c, c->zdfi.ensure(sizeof(cd) * nn * nrhs));
    HIPCHK(c, c->kode.ensure(sizeof(int) * nn));
    HIPCHK(c, hipMemcpyAsync(c->zdfi.p, zdfi, sizeof(cd) * nn * nrhs, hipMemcpyHostToDevice,
                             c->st));
    HIPCHK(c, hipMemcpyAsync(c->kode.p, kode, sizeof(int) * nn, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return BEMX_OK;
}

int bemx_send_load_cases(bemx_ctx* c, int nrhs, const double* zdfi, const int* kode) {
    if (!c) return BEMX_ERR_ARG;
    if (!c->mesh_ok) return fail(c, BEMX_ERR_STATE, "%sdynamic data before the mesh", "", 0);
    if (!zdfi || !kode) return fail(c, BEMX_ERR_ARG, "%snull ZDFI/KODE", "", 0);
    if (nrhs < 1 || nrhs > 4) return fail(c, BEMX_ERR_ARG, "%s1..4 load cases, got %ld", "", nrhs);
    HIPCHK(c, hipSetDevice(c->device));
    c->case_groups.clear();
    int s = upload_cases(c, nrhs, zdfi, kode);
    if (s) return s;
    c->dyn_ok = true;
    return BEMX_OK;
}

int bemx_send_load_cases_kode(bemx_ctx* c, int nrhs, const double* zdfi, const int* kode) {
    if (!c) return BEMX_ERR_ARG;
    if (!c->mesh_ok) return fail(c, BEMX_ERR_STATE, "%sdynamic data before the mesh", "", 0);
    if (!zdfi || !kode) return fail(c, BEMX_ERR_ARG, "%snull ZDFI/KODE", "", 0);
    if (nrhs < 1 || nrhs > 4) return fail(c, BEMX_ERR_ARG, "%s1..4 load cases, got %ld", "", nrhs);
    const size_t nn = 3 * (size_t)c->nbe;
    std::vector<std::vector<int>> groups;
    for (int j = 0; j < nrhs; ++j) {
        bool found = false;
        for (auto& g : groups)
            if (!memcmp(kode + nn * g[0], kode + nn * j, sizeof(int) * nn)) {
                g.push_back(j);
                found = true;
                break;
            }
        if (!found) groups.push_back({j});
    }
    if (groups.size() == 1) return bemx_send_load_cases(c, nrhs, zdfi, kode);
    HIPCHK(c, hipSetDevice(c->device));
    c->case_zdfi.assign(zdfi, zdfi + 2 * nn * nrhs);
    c->case_kode.assign(kode, kode + nn * nrhs);
    c->case_groups = groups;
    c->nrhs = nrhs;
    c->dyn_ok = true;
    return BEMX_OK;
}

int bemx_send_dyn_data(bemx_ctx* c, const double* zdfi, const int* kode) {
    return bemx_send_load_cases(c, 1, zdfi, kode);
}

// Assembly of ZH and ZFI on the device (no host copies).
static int assemble_dev(bemx_ctx* c, const DynConst& K) {
    const int nbe = c->nbe;
    const size_t nn = 3 * (size_t)nbe;
    const int tiles = offdiag_tiles(nbe);
    const int nr = c->nrhs;
    HIPCHK(c, c->colflag.ensure(sizeof(int) * nbe));
    HIPCHK(c, c->flist.ensure(sizeof(int) * (nbe + 1)));
    HIPCHK(c, c->rhs_part.ensure(sizeof(cd) * nn * tiles * nr));
    HIPCHK(c, c->rhs_diag.ensure(sizeof(cd) * nn * nr));
    HIPCHK(c, c->zfi.ensure(sizeof(cd) * nn * nr));
    HIPCHK(c, c->zh.ensure(sizeof(cd) * nn * nn));
    double h = 0.0;
    const int nt = exp_table_entries(K, c->rmax, &h);
    if (nt) HIPCHK(c, c->etab.ensure(sizeof(cd) * 2 * (size_t)nt));
    c->zh_valid = false;
    HIPCHK(c, launch_dyn_assembly(c->st, nbe, c->npg * c->npg, c->elem.as<ElemData>(),
                                  c->gp.as<GaussPt>(), c->hdiag.as<double>(),
                                  c->gdiag.as<double>(), c->kode.as<int>(), c->zdfi.as<cd>(), nr,
                                  c->colflag.as<int>(), c->flist.as<int>(), K,
                                  nt ? c->etab.as<cd>() : nullptr, nt, h, c->zh.as<cd>(), nn,
                                  c->rhs_part.as<cd>(), c->rhs_diag.as<cd>(), c->zfi.as<cd>(),
                                  c->ev2, c->ev3));
    HIPCHK(c, hipEventRecord(c->ev4, c->st));
    c->zh_valid = true;
    return BEMX_OK;
}

int bemx_ghmatecd(bemx_ctx* c, const double* zge, const double* zcs, const double* zcp,
                  double c1, double c2, double c3, double c4, double fr,
                  const double* hestdiag, const double* gestdiag, double* zfi, double* zh) {
    if (!c) return BEMX_ERR_ARG;
    if (!c->mesh_ok || !c->dyn_ok)
        return fail(c, BEMX_ERR_STATE, "%sGHMATECD needs the mesh and the dynamic data", "", 0);
    if (!zge || !zcs || !zcp || !(fr != 0.0))
        return fail(c, BEMX_ERR_ARG, "%sGHMATECD needs the material and a nonzero frequency", "", 0);
    if (!c->case_groups.empty())
        return fail(c, BEMX_ERR_ARG,
                    "%sGHMATECD assembles for one KODE: the load cases have %ld KODE sets", "",
                    (long)c->case_groups.size());
    HIPCHK(c, hipSetDevice(c->device));
    const int nbe = c->nbe;
    const size_t nn = 3 * (size_t)nbe;
    if (hestdiag || gestdiag) {
        if (!(hestdiag && gestdiag))
            return fail(c, BEMX_ERR_ARG, "%spass both HESTdiag and GESTdiag, or neither", "", 0);
        HIPCHK(c, c->hdiag.ensure(sizeof(double) * 9 * nbe));
        HIPCHK(c, c->gdiag.ensure(sizeof(double) * 9 * nbe));
        HIPCHK(c, hipMemcpyAsync(c->hdiag.p, hestdiag, sizeof(double) * 9 * nbe,
                                 hipMemcpyHostToDevice, c->st));
        HIPCHK(c, hipMemcpyAsync(c->gdiag.p, gestdiag, sizeof(double) * 9 * nbe,
                                 hipMemcpyHostToDevice, c->st));
        const double PI = 4.0 * std::atan(1.0);
        c->consts[0] = PI;
        c->consts[1] = c1;
        c->consts[2] = c2;
        c->consts[3] = c3;
        c->consts[4] = c4;
    } else if (!c->static_ok) {
        return fail(c, BEMX_ERR_STATE, "%sno static diagonals: call GHMATECE first", "", 0);
    } else {
        c->consts[1] = c1;
        c->consts[2] = c2;
        c->consts[3] = c3;
        c->consts[4] = c4;
    }
    DynConst K = make_dyn_const(c, ld_c(zge), ld_c(zcs), ld_c(zcp), fr, 1.0);
    int s = assemble_dev(c, K);
    if (s) return s;
    if (zfi)
        HIPCHK(c, hipMemcpyAsync(zfi, c->zfi.p, sizeof(cd) * nn * c->nrhs, hipMemcpyDeviceToHost,
                                 c->st));
    if (zh) HIPCHK(c, hipMemcpyAsync(zh, c->zh.p, sizeof(cd) * nn * nn, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->zh_host = zh;
    return BEMX_OK;
}

int bemx_linsolve(bemx_ctx* c, int nn, double* zh, double* zfi, int mode, long long* info) {
    if (!c) return BEMX_ERR_ARG;
    if (info) *info = 0;
    if (nn <= 0 || !zfi) return fail(c, BEMX_ERR_ARG, "%sbad LINSOLVE arguments", "", 0);
    if (mode != BEMX_MODE_FP64 && mode != BEMX_MODE_MIXED)
        return fail(c, BEMX_ERR_ARG, "%sunknown solver mode %ld", "", mode);
    HIPCHK(c, hipSetDevice(c->device));
    const size_t n = nn;
    bool resident = c->zh_valid && (size_t)3 * c->nbe == n && (zh == nullptr || zh == c->zh_host);
    if (!resident) {
        if (!zh) return fail(c, BEMX_ERR_STATE, "%sno matrix on the device or host", "", 0);
        HIPCHK(c, c->zh.ensure(sizeof(cd) * n * n));
        HIPCHK(c, hipMemcpyAsync(c->zh.p, zh, sizeof(cd) * n * n, hipMemcpyHostToDevice, c->st));
    }
    HIPCHK(c, c->zfi.ensure(sizeof(cd) * n));
    HIPCHK(c, hipMemcpyAsync(c->zfi.p, zfi, sizeof(cd) * n, hipMemcpyHostToDevice, c->st));
    long long inf = 0;
    double st[4] = {0, 0, 0, 0};
    int s = bemx::lu_solve(c->lu, c->st, (int)n, c->zh.as<cd>(), n, c->zfi.as<cd>(), 1, mode,
                           &inf, st, c->err);
    c->zh_valid = false;  // the factorisation overwrote the matrix
    double last[NLAST] = {st[0], st[1], st[2], 0, 0, 0, 0, (double)c->lu.captures, 0, 0,
                          (double)c->lu.replays};
    memcpy(c->last, last, sizeof last);
    if (info) *info = inf;
    if (s) return s;
    HIPCHK(c, hipMemcpyAsync(zfi, c->zfi.p, sizeof(cd) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return BEMX_OK;
}

int bemx_linsolve_device_noalloc(bemx_ctx* c, int nn, void* a, long long lda, void* b,
                                 long long* info) {
    if (!c) return BEMX_ERR_ARG;
    if (info) *info = 0;
    if (!a || !b || nn < 1 || lda < nn)
        return fail(c, BEMX_ERR_ARG, "%sbad device LINSOLVE arguments", "", 0);
    if (nn > bemx::noalloc_max_n())
        return fail(c, BEMX_ERR_ARG, "%sorder above the static workspace's maximum %ld", "",
                    (long)bemx::noalloc_max_n());
    HIPCHK(c, hipSetDevice(c->device));
    long long inf = 0;
    const int s = bemx::lu_solve_fp64_noalloc(c->device, c->st, nn, (cd*)a, (size_t)lda, (cd*)b,
                                              &inf, c->err);
    if (info) *info = inf;
    return s;
}

int bemx_interec(bemx_ctx* c, double* zfi, double* zdfi, const int* kode, int l,
                 const double* cxi, const double* cyi, const double* czi, const double* zge,
                 const double* zcs, const double* zcp, double fr, double rho, double* zdsol,
                 double* zssol) {
    if (!c) return BEMX_ERR_ARG;
    if (!c->mesh_ok) return fail(c, BEMX_ERR_STATE, "%sINTEREC before the mesh", "", 0);
    if (!zfi || !zdfi || !kode || l < 0 || !zge || !zcs || !zcp || !(fr != 0.0))
        return fail(c, BEMX_ERR_ARG, "%sbad INTEREC arguments", "", 0);
    const int nbe = c->nbe;
    const size_t nn = 3 * (size_t)nbe;
    // un-swap on the host arrays, as the reference does in place
    cd z = ld_c(zge);
    for (size_t i = 0; i < nn; ++i) {
        if (kode[i] != 0) continue;
        cd f = mk(zfi[2 * i], zfi[2 * i + 1]) * z;
        zfi[2 * i] = zdfi[2 * i];
        zfi[2 * i + 1] = zdfi[2 * i + 1];
        zdfi[2 * i] = f.x;
        zdfi[2 * i + 1] = f.y;
    }
    if (l == 0) return BEMX_OK;  // Interec.for:71
    if (!cxi || !cyi || !czi || !zdsol || !zssol)
        return fail(c, BEMX_ERR_ARG, "%snull interior arrays", "", 0);
    HIPCHK(c, hipSetDevice(c->device));
    int nblk = interior_blocks(nbe);
    HIPCHK(c, c->ipts.ensure(sizeof(double) * 3 * l));
    HIPCHK(c, c->ipart.ensure(sizeof(cd) * 9 * (size_t)l * nblk));
    HIPCHK(c, c->zdsol.ensure(sizeof(cd) * 3 * l));
    HIPCHK(c, c->zssol.ensure(sizeof(cd) * 9 * l));
    HIPCHK(c, c->ifi.ensure(sizeof(cd) * nn));
    HIPCHK(c, c->idfi.ensure(sizeof(cd) * nn));
    double* pts = c->ipts.as<double>();
    HIPCHK(c, hipMemcpyAsync(pts, cxi, sizeof(double) * l, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(pts + l, cyi, sizeof(double) * l, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(pts + 2 * l, czi, sizeof(double) * l, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->ifi.p, zfi, sizeof(cd) * nn, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->idfi.p, zdfi, sizeof(cd) * nn, hipMemcpyHostToDevice, c->st));
    DynConst K = make_dyn_const(c, ld_c(zge), ld_c(zcs), ld_c(zcp), fr, rho);
    if (c->consts[0] == 0.0) {  // INTEREC only needs pi from the static pass
        c->consts[0] = 4.0 * std::atan(1.0);
        K = make_dyn_const(c, ld_c(zge), ld_c(zcs), ld_c(zcp), fr, rho);
    }
    HIPCHK(c, launch_interior(c->st, c->n, nbe, c->npg * c->npg, l, c->elem.as<ElemData>(),
                              c->gpT.as<GaussPt>(), pts, pts + l, pts + 2 * l, c->ifi.as<cd>(),
                              c->idfi.as<cd>(), K, c->ipart.as<cd>(), c->zdsol.as<cd>(),
                              c->zssol.as<cd>()));
    HIPCHK(c, hipMemcpyAsync(zdsol, c->zdsol.p, sizeof(cd) * 3 * l, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(zssol, c->zssol.p, sizeof(cd) * 9 * l, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return BEMX_OK;
}

}  // extern "C"

// Assembly + factorisation (+ refinement) + solve of the load cases now in
// the device buffers; the solution columns go to out (host, may be NULL),
// statistics are accumulated into acc (GMRES iterations summed, residual
// and fallback as worst, device times summed).
static int solve_pass(bemx_ctx* c, const DynConst& K, int mode, double* out, double acc[NLAST]) {
    const size_t nn = 3 * (size_t)c->nbe;
    HIPCHK(c, hipEventRecord(c->ev0, c->st));
    int s = assemble_dev(c, K);
    if (s) return s;
    long long inf = 0;
    double st[4] = {0, 0, 0, 0};
    s = bemx::lu_solve(c->lu, c->st, (int)nn, c->zh.as<cd>(), nn, c->zfi.as<cd>(), c->nrhs, mode,
                       &inf, st, c->err);
    c->zh_valid = false;
    if (s) return s;
    HIPCHK(c, hipEventRecord(c->ev1, c->st));
    if (out)
        HIPCHK(c, hipMemcpyAsync(out, c->zfi.p, sizeof(cd) * nn * c->nrhs, hipMemcpyDeviceToHost,
                                 c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f, ms_od = 0.f, ms_as = 0.f, ms_f = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    hipEventElapsedTime(&ms_od, c->ev2, c->ev3);
    hipEventElapsedTime(&ms_as, c->ev0, c->ev4);
    if (c->lu.factor_timed) hipEventElapsedTime(&ms_f, c->lu.ev_f0, c->lu.ev_f1);
    acc[0] += st[0];
    acc[1] = std::max(acc[1], st[1]);
    acc[2] = std::max(acc[2], st[2]);
    acc[3] += ms;
    acc[4] += ms_od;
    acc[5] += ms_as;
    acc[6] += ms - ms_as;
    acc[8] += ms_f;
    acc[9] += ms - ms_as - ms_f;
    return BEMX_OK;
}

extern "C" {

int bemx_solve(bemx_ctx* c, const double* zge, const double* zcs, const double* zcp, double fr,
               int mode, double* zfi_out, double* stats) {
    if (!c) return BEMX_ERR_ARG;
    if (!c->mesh_ok || !c->dyn_ok || !c->static_ok)
        return fail(c, BEMX_ERR_STATE, "%ssolve needs mesh, static pass and dynamic data", "", 0);
    if (!zge || !zcs || !zcp || !(fr != 0.0))
        return fail(c, BEMX_ERR_ARG, "%sbad material or frequency", "", 0);
    if (mode != BEMX_MODE_FP64 && mode != BEMX_MODE_MIXED)
        return fail(c, BEMX_ERR_ARG, "%sunknown solver mode %ld", "", mode);
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nn = 3 * (size_t)c->nbe;
    DynConst K = make_dyn_const(c, ld_c(zge), ld_c(zcs), ld_c(zcp), fr, 1.0);
    double acc[NLAST] = {0};
    int s = BEMX_OK;
    if (c->case_groups.empty()) {
        s = solve_pass(c, K, mode, zfi_out, acc);
    } else {
        // one KODE column swap, assembly and factorisation per KODE set
        // (Ghmatecd.for:316-333 applies one KODE to the whole matrix)
        const int ncase = c->nrhs;
        std::vector<double> z, x;
        for (const std::vector<int>& g : c->case_groups) {
            const int ng = (int)g.size();
            z.resize(2 * nn * ng);
            for (int q = 0; q < ng; ++q)
                memcpy(&z[2 * nn * q], &c->case_zdfi[2 * nn * g[q]], sizeof(double) * 2 * nn);
            s = upload_cases(c, ng, z.data(), &c->case_kode[nn * g[0]]);
            if (!s) {
                x.assign(2 * nn * ng, 0.0);
                s = solve_pass(c, K, mode, zfi_out ? x.data() : nullptr, acc);
            }
            if (s) break;
            if (zfi_out)
                for (int q = 0; q < ng; ++q)
                    memcpy(zfi_out + 2 * nn * g[q], &x[2 * nn * q], sizeof(double) * 2 * nn);
        }
        c->nrhs = ncase;
    }
    if (s) return s;
    acc[7] = (double)c->lu.captures;
    acc[10] = (double)c->lu.replays;
    memcpy(c->last, acc, sizeof acc);
    if (stats) {
        stats[0] = acc[0];
        stats[1] = acc[1];
        stats[2] = acc[2];
        stats[3] = acc[3];
    }
    return BEMX_OK;
}

}  // extern "C"

// Problems per batch of the batched solves (BEMX_SWEEP_BATCH, 1..16; default
// 8).  Measured on the 6480-DOF sweep, one context: 118, 141 and 152
// solves/s with batches of 2, 4 and 8.
static int sweep_batch() {
    static const int b = [] {
        const char* v = getenv("BEMX_SWEEP_BATCH");
        const int x = v ? atoi(v) : 8;
        return std::min(bemx::MIXED_MAX_BATCH, std::max(1, x));
    }();
    return b;
}

namespace bemx {

int batch_size() { return sweep_batch(); }

// Batched solves on one context whose mesh, static pass and KODE are in
// place: items go in batches of up to batch_size() problems.  A batch is
// assembled (two problems per launch where they share their load cases) on
// the context's second stream into one half of the frequency slots while
// the previous batch is solved on the context's stream as one batch
// (lu_solve_batch); the next batch's assembly is enqueued right behind the
// current batch's factorisation.  Each problem's arithmetic is that of
// bemx_solve.
// batch_done (may be null) is called with (first item, count) once a
// batch's solutions are in their host arrays.
int ctx_solve_items(bemx_ctx* c, int nitem, SolveItem* items, int mode,
                    const std::function<void(int, int)>* batch_done) {
    if (nitem < 1) return BEMX_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const int nbe = c->nbe, npg2 = c->npg * c->npg, nr = c->nrhs;
    const size_t nn = 3 * (size_t)nbe;
    std::vector<DynConst> K(nitem);
    std::vector<int> nt(nitem);
    std::vector<double> h(nitem);
    int ntmax = 0;
    for (int f = 0; f < nitem; ++f) {
        const SolveItem& it = items[f];
        K[f] = make_dyn_const(c, ld_c(it.zge), ld_c(it.zcs), ld_c(it.zcp), it.fr, 1.0);
        nt[f] = exp_table_entries(K[f], c->rmax, &h[f]);
        ntmax = std::max(ntmax, nt[f]);
    }
    // problems per batch: at most batch_size(), and the matrices and mixed
    // workspaces of the slots within half of the device memory now free
    const size_t zh_bytes = sizeof(cd) * nn * nn, zfi_bytes = sizeof(cd) * nn * nr;
    int nbat = std::min(sweep_batch(), nitem);
    const int halves = nitem > nbat ? 2 : 1;
    {
        size_t fr = 0, tot = 0;
        HIPCHK(c, hipMemGetInfo(&fr, &tot));
        fr += c->sw_zh.cap + c->lu.bytes();
        while (nbat > 1 && halves * nbat * (zh_bytes + zfi_bytes) +
                                   mixed_workspace_bytes((int)nn, nbat) > fr / 2)
            --nbat;
    }
    // batches as even as the count allows (20 problems in batches of at most
    // 8: 7, 7 and 6 rather than 8, 8 and 4)
    nbat = (nitem + (nitem + nbat - 1) / nbat - 1) / ((nitem + nbat - 1) / nbat);
    // the first factorisation of this order also captures the graphs of the
    // other batch sizes a later call may run
    c->lu.ahead = sweep_batch();
    const int nslot = std::min(nitem, halves * nbat);
    HIPCHK(c, c->colflag.ensure(sizeof(int) * nbe));
    HIPCHK(c, c->flist.ensure(sizeof(int) * (nbe + 1)));
    HIPCHK(c, c->sw_zh.ensure(zh_bytes * nslot));
    HIPCHK(c, c->sw_zfi.ensure(zfi_bytes * nslot));
    bool own_cases = false;
    for (int f = 0; f < nitem; ++f) own_cases = own_cases || items[f].zdfi;
    for (int i = 0; i < nslot; ++i) {
        HIPCHK(c, c->sw_rhs_part[i].ensure(sizeof(cd) * nn * offdiag_tiles(nbe) * nr));
        HIPCHK(c, c->sw_rhs_diag[i].ensure(sizeof(cd) * nn * nr));
        if (ntmax) HIPCHK(c, c->sw_etab[i].ensure(sizeof(cd) * 2 * (size_t)ntmax));
        if (own_cases) HIPCHK(c, c->sw_zdfi[i].ensure(zfi_bytes));
    }
    if (!c->st_asm) HIPCHK(c, hipStreamCreateWithFlags(&c->st_asm, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        if (!c->sw_asm[i]) HIPCHK(c, hipEventCreateWithFlags(&c->sw_asm[i], hipEventDisableTiming));
        if (!c->sw_done[i])
            HIPCHK(c, hipEventCreateWithFlags(&c->sw_done[i], hipEventDisableTiming));
    }
    c->zh_valid = false;
    const ElemData* elem = c->elem.as<ElemData>();
    const GaussPt* gp = c->gp.as<GaussPt>();
    auto zh_of = [&](int sl) { return (cd*)((char*)c->sw_zh.p + zh_bytes * sl); };
    auto zfi_of = [&](int sl) { return (cd*)((char*)c->sw_zfi.p + zfi_bytes * sl); };
    // the load cases of problem f in slot sl on the device
    auto cases_of = [&](int f, int sl) {
        return items[f].zdfi ? c->sw_zdfi[sl].as<cd>() : c->zdfi.as<cd>();
    };
    // batch b: problems nbat b .. into the slots of half b & 1, two per
    // assembly launch when they share their load cases
    auto assemble = [&](int b) -> int {
        const int f0 = nbat * b, nf = std::min(nbat, nitem - f0), s0 = nbat * (b & 1);
        for (int i = 0; i < nf; ++i)
            if (items[f0 + i].zdfi)
                HIPCHK(c, hipMemcpyAsync(c->sw_zdfi[s0 + i].p, items[f0 + i].zdfi, zfi_bytes,
                                         hipMemcpyHostToDevice, c->st_asm));
        for (int i = 0; i < nf; i += 2) {
            const int np = std::min(2, nf - i);
            DynFreq F[2];
            for (int q = 0; q < np; ++q) {
                const int f = f0 + i + q, s = s0 + i + q;
                F[q] = DynFreq{K[f], nt[f] ? c->sw_etab[s].as<cd>() : nullptr, nt[f], h[f],
                               zh_of(s), c->sw_rhs_part[s].as<cd>(), c->sw_rhs_diag[s].as<cd>(),
                               zfi_of(s)};
            }
            const bool pair = np == 2 && nt[f0 + i] > 0 && nt[f0 + i + 1] > 0 &&
                              !items[f0 + i].zdfi && !items[f0 + i + 1].zdfi;
            if (pair) {
                HIPCHK(c, launch_dyn_assembly_pair(c->st_asm, nbe, npg2, elem, gp,
                                                   c->hdiag.as<double>(), c->gdiag.as<double>(),
                                                   c->kode.as<int>(), c->zdfi.as<cd>(), nr,
                                                   c->colflag.as<int>(), c->flist.as<int>(), F,
                                                   nn, nullptr, nullptr));
            } else {
                for (int q = 0; q < np; ++q)
                    HIPCHK(c, launch_dyn_assembly(c->st_asm, nbe, npg2, elem, gp,
                                                  c->hdiag.as<double>(), c->gdiag.as<double>(),
                                                  c->kode.as<int>(), cases_of(f0 + i + q, s0 + i + q),
                                                  nr, c->colflag.as<int>(), c->flist.as<int>(),
                                                  F[q].K, F[q].etab, F[q].nt, F[q].h, F[q].zh,
                                                  nn, F[q].rhs_part, F[q].rhs_diag, F[q].zfi,
                                                  nullptr, nullptr));
            }
        }
        HIPCHK(c, hipEventRecord(c->sw_asm[b & 1], c->st_asm));
        return BEMX_OK;
    };
    const int nb = (nitem + nbat - 1) / nbat;
    std::vector<double> st4(4 * nbat);
    c->sw_factor_ms = 0.0;
    HIPCHK(c, hipEventRecord(c->ev0, c->st));
    HIPCHK(c, hipStreamWaitEvent(c->st_asm, c->ev0, 0));  // after the context's earlier work
    int s = assemble(0);
    for (int b = 0; b < nb && !s; ++b) {
        HIPCHK(c, hipStreamWaitEvent(c->st, c->sw_asm[b & 1], 0));
        // the next batch's assembly beside this batch's solves, after the
        // solves that last used its slots
        const std::function<int()> next = [&]() -> int {
            if (b + 1 >= nb) return BEMX_OK;
            if (b >= 1) HIPCHK(c, hipStreamWaitEvent(c->st_asm, c->sw_done[(b + 1) & 1], 0));
            return assemble(b + 1);
        };
        const int f0 = nbat * b, nf = std::min(nbat, nitem - f0), s0 = nbat * (b & 1);
        c->lu.factor_timed = false;
        s = lu_solve_batch(c->lu, c->st, (int)nn, nf, zh_of(s0), nn, zh_bytes, zfi_of(s0),
                           zfi_bytes, nr, mode, st4.data(), c->err, &next);
        if (s) break;
        if (c->lu.factor_timed) {  // the batch's factorisation, waited for by its refinement
            float ms_f = 0.f;
            if (hipEventElapsedTime(&ms_f, c->lu.ev_f0, c->lu.ev_f1) == hipSuccess)
                c->sw_factor_ms += ms_f;
        }
        for (int i = 0; i < nf; ++i) {
            SolveItem& it = items[f0 + i];
            if (it.x)
                HIPCHK(c, hipMemcpyAsync(it.x, zfi_of(s0 + i), zfi_bytes, hipMemcpyDeviceToHost,
                                         c->st));
            for (int q = 0; q < 4; ++q) it.stats[q] = st4[4 * i + q];
            it.stats[3] = 0.0;
        }
        HIPCHK(c, hipEventRecord(c->sw_done[b & 1], c->st));
        if (batch_done) {
            HIPCHK(c, hipStreamSynchronize(c->st));
            (*batch_done)(f0, nf);
        }
    }
    hipStreamSynchronize(c->st_asm);
    if (s) {
        hipStreamSynchronize(c->st);
        return s;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return BEMX_OK;
}

}  // namespace bemx

extern "C" {

int bemx_solve_sweep(bemx_ctx* c, const double* zge, const double* zcs, const double* zcp,
                     int nfreq, const double* freqs, int mode, double* zfi_out, double* stats) {
    if (!c) return BEMX_ERR_ARG;
    if (!c->mesh_ok || !c->dyn_ok || !c->static_ok)
        return fail(c, BEMX_ERR_STATE, "%ssweep needs mesh, static pass and dynamic data", "", 0);
    if (!zge || !zcs || !zcp || !freqs || nfreq < 1)
        return fail(c, BEMX_ERR_ARG, "%sbad material or frequency list", "", 0);
    for (int f = 0; f < nfreq; ++f)
        if (!(freqs[f] != 0.0)) return fail(c, BEMX_ERR_ARG, "%szero frequency at %ld", "", f);
    if (mode != BEMX_MODE_FP64 && mode != BEMX_MODE_MIXED)
        return fail(c, BEMX_ERR_ARG, "%sunknown solver mode %ld", "", mode);
    if (!c->case_groups.empty())
        return fail(c, BEMX_ERR_ARG,
                    "%sload cases with KODE sets of their own are solved with bemx_solve", "", 0);
    const size_t nn = 3 * (size_t)c->nbe;
    std::vector<bemx::SolveItem> items(nfreq);
    for (int f = 0; f < nfreq; ++f) {
        bemx::SolveItem& it = items[f];
        for (int q = 0; q < 2; ++q) {
            it.zge[q] = zge[q];
            it.zcs[q] = zcs[q];
            it.zcp[q] = zcp[q];
        }
        it.fr = freqs[f];
        it.zdfi = nullptr;
        it.x = zfi_out ? zfi_out + 2 * nn * c->nrhs * f : nullptr;
    }
    const int s = bemx::ctx_solve_items(c, nfreq, items.data(), mode, nullptr);
    if (s) return s;
    double acc[NLAST] = {0};
    for (int f = 0; f < nfreq; ++f) {
        const double* st4 = items[f].stats;
        if (stats)
            for (int q = 0; q < 4; ++q) stats[4 * f + q] = st4[q];
        acc[0] += st4[0];
        acc[1] = std::max(acc[1], st4[1]);
        acc[2] = std::max(acc[2], st4[2]);
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    acc[3] = ms;
    acc[7] = (double)c->lu.captures;
    acc[8] = c->sw_factor_ms;
    acc[10] = (double)c->lu.replays;
    memcpy(c->last, acc, sizeof acc);
    return BEMX_OK;
}

int bemx_last_stats(const bemx_ctx* c, double* out, int nout) {
    if (!c || !out || nout < 1) return BEMX_ERR_ARG;
    for (int i = 0; i < nout && i < NLAST; ++i) out[i] = c->last[i];
    return BEMX_OK;
}

int bemx_device_buffers(bemx_ctx* c, void** zh, void** zfi, void** stream) {
    if (!c) return BEMX_ERR_ARG;
    if (zh) *zh = c->zh.p;
    if (zfi) *zfi = c->zfi.p;
    if (stream) *stream = (void*)c->st;
    return BEMX_OK;
}

}  // extern "C"
