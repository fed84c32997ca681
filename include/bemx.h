/*
 * bemx.h -- C ABI of the MI355X-native BEM hot path (libbemx.so).
 *
 * Drop-in replacement for the reference's four assembly/solve subroutines
 * and the device entry points they sat on (giulianobelinassi/BEM-MECECCS,
 * the .cu files of src/kernels).  Plain pointers and sizes only.  Arrays use the
 * reference's Fortran layouts:
 *   - real arrays are double (FPREC=double build)
 *   - complex arrays / scalars are interleaved (re, im) doubles
 *   - CONE(N,4) column-major int32 with 1-based node ids
 *   - ETAS(3,N), HESTdiag/GESTdiag(3,3,NBE) column-major
 *   - ZH(3NBE,3NBE) column-major, leading dimension 3NBE
 * Every entry point returns a bemx_status; nothing inside the library exits
 * the process (the reference exit(1)s on CUDA/MAGMA errors,
 * kernels/shared.cu:33-41, Linsolve_cu.cu:10-17).  A context may be used by
 * one host thread at a time; separate contexts are independent.
 */
#ifndef BEMX_H
#define BEMX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bemx_ctx bemx_ctx;

typedef enum {
    BEMX_OK = 0,
    BEMX_ERR_DEGENERATE = 1, /* Jacobian DET < 1e-5 (Nonsingd.for:89-95); the
                                reference's GPU status "Matriz Singular" */
    BEMX_ERR_SINGULAR = 2,   /* exactly zero pivot (LAPACK INFO > 0) */
    BEMX_ERR_ARG = 3,        /* invalid argument / size */
    BEMX_ERR_HIP = 4,        /* HIP runtime error */
    BEMX_ERR_NOMEM = 5,      /* device allocation failed / admission refused */
    BEMX_ERR_STATE = 6,      /* entry point called out of order */
    BEMX_ERR_NOCONV = 7,     /* iterative refinement did not reach tolerance */
    BEMX_ERR_TIMEOUT = 8     /* a bounded device-side wait gave up */
} bemx_status;

/* Solver modes for LINSOLVE / bemx_solve. */
#define BEMX_MODE_FP64 0  /* partial-pivot LU in FP64 (ZGESV semantics) */
#define BEMX_MODE_MIXED 1 /* FP32 LU (split-BF16x3 MFMA update) + FP64 GMRES-IR */

/* Context on one device.  No reference counterpart: replaces the module
 * globals of kernels/shared.cu:6-31. */
int bemx_create(int device, bemx_ctx **ctx);
int bemx_destroy(bemx_ctx *ctx);
const char *bemx_last_error(const bemx_ctx *ctx);

/* Gauss-Legendre rule on [-1,1].  Replaces GAULEG (Gauleg.for:9-36). */
int bemx_gauleg(int npg, double *gi, double *ome);

/* Upload the mesh and build its device geometry (normals, centroids, Gauss
 * points).  Replaces send_shared_data_to_gpu_ (kernels/shared.cu:84-165).
 * cxm/cym/czm/etas may be NULL: the device then computes them as NORMVEC
 * (Normvec.for:4-39) and INPUTECE (Inputece.for:100-102) do; gi/ome may be
 * NULL: bemx_gauleg is used. */
int bemx_send_shared_data(bemx_ctx *ctx, const double *cx, const double *cy, const double *cz,
                          const double *cxm, const double *cym, const double *czm,
                          const double *etas, const double *gi, const double *ome,
                          const int *cone, int np, int npg, int n, int nbe);

/* Release the mesh.  Replaces deallocate_shared_gpu_data_ (kernels/shared.cu:184-214). */
int bemx_deallocate_shared_data(bemx_ctx *ctx);

/* Static pass.  Replaces GHMATECE (Ghmatece.for:9-256) = cuda_ghmatece_
 * (kernels/Ghmatece_cu.cu:264-371) + GEST_SINGULAR + cuda_send_gest_data_
 * (:373-380).  Outputs (any may be NULL): HESTdiag(3,3,NBE), GESTdiag(3,3,NBE),
 * DELTA(3,3), PI, C1..C4.  The diagonals stay resident for bemx_ghmatecd. */
int bemx_ghmatece(bemx_ctx *ctx, double ge, double rnu, double *hestdiag, double *gestdiag,
                  double *delta, double *pi, double *c1, double *c2, double *c3, double *c4);

/* Boundary data of the dynamic problem: ZDFI(3NBE) complex, KODE(3NBE).
 * Replaces send_dyn_data_gpu_ (kernels/shared.cu:168-182). */
int bemx_send_dyn_data(bemx_ctx *ctx, const double *zdfi, const int *kode);

/* Boundary data of nrhs (1..4) load cases on the same mesh and KODE:
 * ZDFI(3NBE, nrhs) complex, column-major.  Each load case is a column of
 * ZDFI as send_dyn_data_gpu_ (kernels/shared.cu:168-182) takes one; the
 * reference solves one load case per run, so bemx_ghmatecd then returns
 * ZFI(3NBE, nrhs) and bemx_solve solves all columns with one factorisation. */
int bemx_send_load_cases(bemx_ctx *ctx, int nrhs, const double *zdfi, const int *kode);

/* Load cases with a KODE column each: ZDFI(3NBE, nrhs) and KODE(3NBE, nrhs),
 * column-major, nrhs 1..4.  A KODE set is one column swap of the matrix
 * (BOUNDARY_CONDITIONS, Ghmatecd.for:316-333, called at :168 with one KODE
 * per run), so bemx_solve groups the cases by identical KODE columns and
 * assembles, factors and solves once per group, returning every case in its
 * own column.  With a single KODE set this is bemx_send_load_cases;
 * bemx_ghmatecd refuses several sets (BEMX_ERR_ARG). */
int bemx_send_load_cases_kode(bemx_ctx *ctx, int nrhs, const double *zdfi, const int *kode);

/* Dynamic assembly at one frequency.  Replaces GHMATECD (Ghmatecd.for:11-239)
 * = cuda_ghmatecd_ (kernels/Ghmatecd_cu.cu:305-483): ZH (3NBE x 3NBE) after
 * the KODE column swap, and ZFI = ZG ZDFI.  zge/zcs/zcp point to complex
 * scalars.  hestdiag/gestdiag (host) may be NULL to use the ones resident
 * from bemx_ghmatece; c1..c4 are the static constants.  zh may be NULL: the
 * matrix then stays on the device only (bemx_linsolve uses it). */
int bemx_ghmatecd(bemx_ctx *ctx, const double *zge, const double *zcs, const double *zcp,
                  double c1, double c2, double c3, double c4, double fr,
                  const double *hestdiag, const double *gestdiag, double *zfi, double *zh);

/* Dense complex solve ZH x = ZFI, x returned in zfi.  Replaces LINSOLVE
 * (Linsolve.for:1-66) = LINSOLVE_CPU/ZGESV (:68-109) and cuda_linsolve_
 * (kernels/Linsolve_cu.cu:53-126).  If zh is NULL, or is the host array the
 * last bemx_ghmatecd wrote, the device-resident matrix is factored (the
 * reference hands device_zh from GHMATECD to LINSOLVE the same way);
 * otherwise zh is uploaded.  mode: BEMX_MODE_FP64 or BEMX_MODE_MIXED.
 * info (may be NULL): LAPACK-style (0, or k > 0 for a zero pivot at k). */
int bemx_linsolve(bemx_ctx *ctx, int nn, double *zh, double *zfi, int mode, long long *info);

/* Dense FP64 solve of a device-resident system that allocates no device
 * memory.  No counterpart in the reference: for callers that cannot call
 * hipMalloc (tight serving loops, harnesses that forbid allocation).  It
 * reads INFO back and synchronises the stream, so it is not for use under
 * graph capture.  a: nn x nn complex, column-major with leading
 * dimension lda >= nn, overwritten by its LU factors; b: nn complex,
 * overwritten by the solution; both device pointers on the context's
 * device.  ZGESV semantics and kernels of bemx_linsolve's FP64 path, on the
 * context's stream; returns when the solve is done.  The pivoting
 * workspace is a static device array of 35.3 MB per GPU, so nn <= 32768
 * (larger orders: BEMX_ERR_ARG); calls that overlap on one GPU share that
 * array and race.  info as bemx_linsolve. */
int bemx_linsolve_device_noalloc(bemx_ctx *ctx, int nn, void *a, long long lda, void *b,
                                 long long *info);

/* Interior points.  Replaces INTEREC (Interec.for:9-288) = the KODE
 * un-swap of ZFI/ZDFI (in place, as the reference), cuda_interec1_
 * (kernels/Interec1_cu.cu:44-237) for ZDSOL(3L) and the SIGMAEC stress loop
 * (Interec.for:186-283) for ZSSOL(9L). */
int bemx_interec(bemx_ctx *ctx, double *zfi, double *zdfi, const int *kode, int l,
                 const double *cxi, const double *cyi, const double *czi, const double *zge,
                 const double *zcs, const double *zcp, double fr, double rho, double *zdsol,
                 double *zssol);

/* Device-resident solve of one frequency: assembly + factor (+ refine) +
 * solve, with the boundary data from bemx_send_dyn_data.  No reference
 * counterpart (it is GHMATECD + LINSOLVE without the host round trip of ZH);
 * used by the sweep, the serving loop and the benchmark.  zfi_out (host,
 * may be NULL) receives the solution, 3NBE x nrhs for the load cases of
 * bemx_send_load_cases (one factorisation serves all of them); stats (may
 * be NULL, 4 doubles): GMRES iterations, worst final ||r||/||b||, 1 if the
 * mixed path fell back to FP64, assembly+solve device time in ms. */
int bemx_solve(bemx_ctx *ctx, const double *zge, const double *zcs, const double *zcp,
               double fr, int mode, double *zfi_out, double *stats);

/* Frequency sweep on the resident mesh, static pass and load cases (one
 * KODE for every case): for each of nfreq frequencies the dynamic assembly,
 * factorisation (+ refinement in BEMX_MODE_MIXED) and solve of bemx_solve.
 * Replaces the frequency loop of the reference's main program
 * (Main.for:167-200: GHMATECD + LINSOLVE per frequency) for one material.
 * Frequencies go in batches of up to BEMX_SWEEP_BATCH (environment, 1..16,
 * default 8; fewer when device memory is short; sizes as even as the count
 * allows, e.g. 7, 7, 6 for 20 frequencies): a batch is assembled two
 * frequencies per kernel launch on a second stream while the previous batch
 * is solved, and in BEMX_MODE_MIXED the batch's factorisations and
 * refinements share every launch.  Each frequency's solution is
 * bit-identical to bemx_solve's.  zfi_out (may be NULL): nfreq blocks of
 * 3 NBE x nrhs complex solutions in frequency order; stats (may be NULL): 4
 * doubles per frequency (GMRES iterations, final relative residual, 1 if
 * the FP64 fallback ran, 0).  Device memory: two batches of frequency slots
 * (ZH, ZFI and assembly partials each) and the batch's mixed workspace. */
int bemx_solve_sweep(bemx_ctx *ctx, const double *zge, const double *zcs, const double *zcp,
                     int nfreq, const double *freqs, int mode, double *zfi_out, double *stats);

/* Statistics of the last bemx_linsolve / bemx_solve (the first nout of 11
 * doubles): GMRES iterations, final ||r||/||b||, 1 if the mixed path fell
 * back to FP64, solve device time (ms), off-diagonal assembly kernel time
 * (ms), whole assembly time (ms), factor+refine+solve time (ms),
 * factorisation graphs captured by the context so far, factorisation time
 * (ms), refinement/solve time after it (ms), graph replays so far.  Times
 * are device times from HIP events, filled by bemx_solve; after
 * bemx_solve_sweep: iterations summed, worst residual and fallback, the
 * sweep's device time, graph counts and the factorisations' device time
 * summed over its batches (mixed mode).  No
 * reference counterpart (the reference prints per-routine wall times,
 * Ghmatecd.for:183-184). */
int bemx_last_stats(const bemx_ctx *ctx, double *out, int nout);

/* Device buffer addresses of the last solve (for zero-copy consumers). */
int bemx_device_buffers(bemx_ctx *ctx, void **zh, void **zfi, void **hip_stream);

/* Complex material constants: mu* = G(1 + 2i DAM), cs = sqrt(mu* / rho),
 * cp = cs sqrt((2-2nu)/(1-2nu)).  Replaces the INPUTECD lines that form
 * ZGE/ZCS/ZCP (Inputecd.for:58-60).  Outputs are complex (re, im). */
int bemx_material(double ge, double rnu, double dam, double rho, double *zge, double *zcs,
                  double *zcp);

/* ---- Request loop (continuous batching on one GPU) ----------------------
 * No reference counterpart: the reference solves one deck per process
 * (Main.for).  nslots worker slots (1..16), each a context with its own
 * streams, take queued requests oldest first as soon as the request's HBM
 * workspace fits in hbm_budget_bytes (<= 0: 80% of the free HBM at
 * creation); their work overlaps on the device.  A slot takes with the
 * oldest request up to BEMX_SWEEP_BATCH - 1 more queued requests on the same
 * mesh with the same KODE and number of load cases and solves them as one
 * batch (as bemx_solve_sweep does; each result bit-identical to the request
 * solved alone).  A slot is charged the
 * workspace bound (bemx_request_footprint) at the largest sizes it has
 * served, since its buffers only grow; idle slots give their workspace back
 * when the oldest request does not fit.  mode: BEMX_MODE_FP64 or BEMX_MODE_MIXED for
 * every request.  Mesh arrays as bemx_send_shared_data, material as
 * bemx_material, ZDFI(3NBE, nrhs) / KODE(3NBE) as bemx_send_load_cases;
 * everything is copied at submit.  A request larger than the whole budget
 * is refused with BEMX_ERR_NOMEM. */
typedef struct bemx_server bemx_server;
/* Device bytes a slot is charged once it has served a request of these
 * sizes (nbe boundary elements, n elements, np nodes, npg Gauss points per
 * direction, nrhs load cases): an upper bound of every buffer its context
 * holds after the request.  0 for invalid sizes. */
double bemx_request_footprint(int nbe, int n, int np, int npg, int nrhs, int mode);
int bemx_server_create(int device, int nslots, double hbm_budget_bytes, int mode,
                       bemx_server **srv);
int bemx_server_submit(bemx_server *srv, const double *cx, const double *cy, const double *cz,
                       const int *cone, int np, int n, int nbe, int npg, double ge, double rnu,
                       double dam, double rho, double fr, int nrhs, const double *zdfi,
                       const int *kode, int kode_cols, long long *id);
/* kode_cols: 1 when KODE(3NBE) holds for every load case, nrhs for
 * KODE(3NBE, nrhs) with a column per case (as bemx_send_load_cases_kode). */
/* Blocks until request id is done; returns its status.  x (may be NULL):
 * 3NBE x nrhs complex solution.  info (may be NULL, 11 doubles): latency
 * ms (submit to done), queue ms, service ms (of its batch), GMRES
 * iterations, worst relative residual, 1 if it fell back to FP64, slot,
 * device ms of the solve, start and done times in ms since the server was
 * created, and the number of requests in its batch.  msg
 * (may be NULL) receives the error text.  Each id can be waited once. */
int bemx_server_wait(bemx_server *srv, long long id, double *x, double *info, char *msg,
                     int msglen);
/* Statistics of the request loop (nout <= 10 doubles): requests completed,
 * factorisation graphs captured for a batch, graph replays, bytes charged
 * now, peak bytes charged, the budget, workspaces given back to make room,
 * the number of batches after which a slot held more device bytes than it
 * was charged (0 unless the bound is wrong), batches run, and graphs
 * captured ahead (a slot's first batch of a mesh size also captures the
 * other batch sizes its workspace holds, up to the batch limit; a later
 * batch of such a size replays).  Captures plus replays equal batches. */
int bemx_server_stats(bemx_server *srv, double *out, int nout);
/* Stops the workers after their current request; requests still queued
 * are dropped.  Each worker releases its slot's workspace before it
 * returns.  Must not run concurrently with bemx_server_wait. */
int bemx_server_destroy(bemx_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* BEMX_H */
