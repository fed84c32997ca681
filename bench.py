"""Benchmark: BEM solves/s and p50/p99 latency, mixed precision vs plain FP64.

Default run (`python bench.py`): the frequency sweep on the 48 x 45 flat
free-surface patch over [-5, 5]^2 (2160 boundary elements + 5 enclosing,
6480 complex DOF), mu = 1, nu = 0.25, rho = 1, damping 0.01, 8 x 8 Gauss
points, unit vertical traction on the central 2 x 2 elements, 256
frequencies log-spaced in [0.1, 10] rad/s.  A step is one frequency: the
dynamic assembly of ZH and ZFI (FP64), the factorisation and the
refinement, with the mesh, its static pass and the boundary data resident
in HBM before the timed region.  The frequencies go through the request
loop (bemx_server) on --slots slots, which takes the queued frequencies in
batches (BEMX_SWEEP_BATCH, 8 by default) whose mixed factorisations and
refinements share every launch; `value` is the mixed path's solves/s over
the sweep and `fp64` the plain FP64 path's over the same sweep (the same
loop and batched assembly, one FP64 LU per frequency), every mixed result
is compared with the FP64 result of its frequency.  Beside it:
  single_solve  the same patch at omega = 2 on one context, mixed and FP64
                interleaved: latency of one solve alone and the device times
                of assembly, factorisation and refinement (HIP events)
  serve         a slice of the serving stream (--requests Poisson arrivals at
                --rate/s, meshes 16x15 .. 64x60, omega ~ U[0.1, 10], damping
                ~ U[0.005, 0.05], 1..4 load cases): submit-to-done p50/p99,
                graphs captured after warm-up, peak HBM charge
  roofline      the batched mixed factorisation as the sweep runs it
                (split-BF16x3 MFMA trailing update): 8 n^3 BF16 flops per
                frequency over the batches' factorisation device time
  roofline_single_factor  the same for one factorisation alone
  roofline_assembly  k_dyn_offdiag_h (FP64 vector ALU): 242 flops per triple
  cpu_baseline  the C oracle (OpenMP FP64 GHMATECD + LAPACK zgesv) solving
                one frequency of the sweep on the host cores: the same unit
                of work as a GPU step
Other workloads (--workload): `solve` (the omega = 2 patch alone), `serve`
(the full serving stream, 2000 requests by default), `large` (96 x 90 patch
over [-10, 10]^2, 8640 BE, 25920 DOF, omega = 0.5, 6 x 6 Gauss).
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bem-mececcs_amd"))
# Hardware queues of this process (read by the HIP runtime at its first call):
# HIP maps every stream onto one of them, the internal stream of each
# instantiated graph's side branch included, so with several request slots
# the default of 4 queues serialises unrelated work (measured on the serving
# stream: p99 84 ms with 8 queues, 74 ms with 16).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

import numpy as np  # noqa: E402

METRIC = ("BEM solves/s + p50/p99 latency at ≤1e-10 err vs FP64, mixed vs plain FP64")
DTYPE = "f64 assembly + f32/bf16x3 factor + f64 refine"

# FP64 operations per (collocation point, element, Gauss point) triple of the
# off-diagonal assembly kernel k_dyn_offdiag_h, counted from the factored
# formula it evaluates: 234 add/mul (an fma counts 2) + 8 special functions
# (sqrt, 1/R, 2 exp, 2 sin, 2 cos) counted as one each, however they are
# evaluated (the Newton steps that refine 1/R and the table-and-polynomial
# evaluation of the exponentials are implementation details, not counted).
FLOPS_PER_TRIPLE = 242
FP64_PEAK_TFLOPS = 78.6    # MI355X dense FP64 (vector ALU and matrix cores alike)
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix cores
HBM_PEAK_GBS = 8000.0
# HBM bytes on the 48 x 45 patch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs of `bench.py --workload solve` (FETCH_SIZE x 2 and WRITE_SIZE x 1: the
# factors that a streaming copy of known size, k_cvt_z2c, reads back), per
# k_dyn_offdiag_h launch and per factorisation (every kernel of it), recorded
# under profiles/.
OFFDIAG_TRAFFIC_BYTES = 735.6e6
FACTOR_TRAFFIC_BYTES = 10.22e9
FP64_FACTOR_TRAFFIC_BYTES = 29.26e9
# Counter figures of the same runs: FP64 VALU flops counted by
# SQ_INSTS_VALU_FLOPS_FP64 (per wave instruction, x 64 lanes) per
# k_dyn_offdiag_h launch, and MFMA flops (SQ_INSTS_VALU_MFMA_MOPS_* x 512)
# and MFMA-busy SIMD cycles (SQ_VALU_MFMA_BUSY_CYCLES) per factorisation.
OFFDIAG_COUNTED_FLOPS = 92.8e9
FACTOR_COUNTED_FLOPS = 2261.3e9
FACTOR_MFMA_BUSY_CYCLES = 2208.3e6
FP64_FACTOR_COUNTED_FLOPS = 731.0e9
FP64_FACTOR_MFMA_BUSY_CYCLES = 22842.6e6
SIMDS, CLOCK_HZ = 1024, 2.4e9  # MI355X: 256 CUs x 4 SIMDs, peak engine clock
# Speed of the reference's own CPU GHMATECD (its NONSINGD/SING_DE, built with
# its Makefile's -Ofast -march=native -fopenmp) relative to the C oracle,
# measured once on identical hardware (oracle/calibrate.py; profiles/
# cpu_calibration.md): time(reference) / time(oracle) on the same mesh and
# thread count.
REF_OVER_ORACLE = 0.743


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_err(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def host_threads():
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v and v.isdigit() else (os.cpu_count() or 1)


def base_record(args, workload, value, ms_per_step, steps):
    return {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE,
        "data": "synthetic (generated flat-patch meshes)",
        "config": dict(workload, parallelism="single GPU, %d request-loop slots" % args.slots),
    }


def stream_through_server(api, mode, slots, items, label, warm_items=()):
    """Submit every item at once to a request loop of `slots` slots; items
    are (mesh, kode, zdfi, (ge, rnu, dam, rho), omega).  warm_items run
    first on the same loop, untimed, so each slot has its workspace, its
    mesh and static pass and its captured factorisation before the timed
    items.  Returns (x list, info list, wall seconds, stats after warm-up,
    stats at the end)."""
    srv = api.Server(0, nslots=slots, mode=mode)
    try:
        if warm_items:
            for i in [srv.submit(*it) for it in warm_items]:
                srv.wait(i)
        warm = srv.stats()
        t0 = time.perf_counter()
        ids = [srv.submit(*it) for it in items]
        outs = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t0
        end = srv.stats()
    finally:
        srv.close()
    svc = [o[1]["service_ms"] for o in outs]
    log("%s: %d solves in %.3f s on %d slots (%.2f solves/s), service p50 %.2f ms" %
        (label, len(items), wall, slots, len(items) / wall, percentile(svc, 50)))
    return [o[0] for o in outs], [o[1] for o in outs], wall, warm, end


def sweep_frequencies(dyn, k):
    """k frequencies of the 256-point sweep, evenly spaced over its indices
    (all 256 for k = 256)."""
    f = np.asarray(dyn.freqs, dtype=np.float64)
    idx = np.round(np.linspace(0, len(f) - 1, k)).astype(int) if k <= len(f) else \
        np.arange(k) % len(f)
    return [float(f[i]) for i in idx]


# ---------------------------------------------------------------------------
def single_solve(args, api, decks, steps):
    """The 48 x 45 patch at omega = 2 on one context: mixed and FP64
    interleaved, latency of one solve alone and the device-time breakdown."""
    mesh, dyn = decks.config_problem("c2")
    fr = float(dyn.freqs[0])
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx = api.Context(0)
    try:
        ctx.send_shared_data(mesh)
        ctx.ghmatece(mesh.ge, mesh.rnu)
        ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
        for _ in range(max(1, args.warmup)):
            ctx.solve(zmat, fr, api.MODE_FP64)
            ctx.solve(zmat, fr, api.MODE_MIXED)
        lat = {"mixed": [], "fp64": []}
        dev = {"mixed": [], "fp64": []}
        errs, iters = [], []
        for step in range(steps):
            out = {}
            for name, mode in (("fp64", api.MODE_FP64), ("mixed", api.MODE_MIXED)):
                t0 = time.perf_counter()
                x, _ = ctx.solve(zmat, fr, mode)
                lat[name].append((time.perf_counter() - t0) * 1e3)
                dev[name].append(ctx.last_stats())
                out[name] = x
            errs.append(rel_err(out["mixed"], out["fp64"]))
            iters.append(dev["mixed"][-1]["gmres_iters"])
            sm, s64 = dev["mixed"][-1], dev["fp64"][-1]
            log("omega=2 step %d: mixed %.2f ms (assembly %.2f, factor %.2f, refine %.2f, %d it), "
                "fp64 %.2f ms (factor %.2f), rel err %.2e" %
                (step, lat["mixed"][-1], sm["assembly_ms"], sm["lu_factor_ms"],
                 sm["lu_refine_ms"], sm["gmres_iters"], lat["fp64"][-1], s64["lu_factor_ms"],
                 errs[-1]))
    finally:
        ctx.close()

    def mean(name, key):
        return float(np.mean([d[key] for d in dev[name]]))

    n = 3 * mesh.nbe
    t_fac = mean("mixed", "lu_factor_ms")
    t_fac64 = mean("fp64", "lu_factor_ms")
    t_od = float(np.mean([d["offdiag_ms"] for d in dev["mixed"] + dev["fp64"]]))
    triples = mesh.nbe * mesh.nbe * mesh.npg ** 2
    bf16_flops = 8.0 * n ** 3          # 3 BF16 products x 8 n^3 / 3 complex-LU flops
    fp64_flops = 8.0 * n ** 3 / 3.0
    rec = {
        "workload": "flat patch 48x45 (2160 BE + 5 EE, 6480 complex DOF), omega=2.0, 8x8 Gauss",
        "mixed": {"p50_ms": percentile(lat["mixed"], 50), "p99_ms": percentile(lat["mixed"], 99),
                  "device_ms": mean("mixed", "solve_ms"), "assembly_ms": mean("mixed", "assembly_ms"),
                  "factor_ms": t_fac, "refine_ms": mean("mixed", "lu_refine_ms"),
                  "gmres_iters_max": int(max(iters))},
        "fp64": {"p50_ms": percentile(lat["fp64"], 50), "p99_ms": percentile(lat["fp64"], 99),
                 "device_ms": mean("fp64", "solve_ms"), "assembly_ms": mean("fp64", "assembly_ms"),
                 "factor_ms": t_fac64, "solve_ms": mean("fp64", "lu_refine_ms")},
        "max_rel_err_mixed_vs_fp64": max(errs),
    }
    roof_lu = {"kernel": "mixed factorisation (graph: k_panel_inv, k_cgemm_update_inverse, "
                         "k_cgemm_update2, ...; split-BF16x3 trailing update on MFMA)",
               "bound": "mfma", "achieved": bf16_flops / (t_fac * 1e-3) / 1e12,
               "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": bf16_flops / (t_fac * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
               "traffic": FACTOR_TRAFFIC_BYTES, "avg_ms": t_fac,
               "algorithmic_flops": bf16_flops,
               "counted_mfma_flops": FACTOR_COUNTED_FLOPS,
               "mfma_busy": FACTOR_MFMA_BUSY_CYCLES / (SIMDS * CLOCK_HZ * t_fac * 1e-3),
               "note": "device time of the whole factorisation from HIP events around its graph "
                       "launch; flops = 3 BF16 products x 8n^3/3, n = 6480; traffic, counted "
                       "flops and MFMA-busy cycles (over 1024 SIMDs at 2.4 GHz) from the "
                       "rocprofv3 counters"}
    roof_fp64 = {"kernel": "FP64 factorisation (k_panel_fp64, k_zgemm_mfma, ...)",
                 "bound": "mfma", "achieved": fp64_flops / (t_fac64 * 1e-3) / 1e12,
                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": fp64_flops / (t_fac64 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                 "avg_ms": t_fac64, "algorithmic_flops": fp64_flops,
                 "traffic": FP64_FACTOR_TRAFFIC_BYTES, "counted_mfma_flops": FP64_FACTOR_COUNTED_FLOPS,
                 "mfma_busy": FP64_FACTOR_MFMA_BUSY_CYCLES / (SIMDS * CLOCK_HZ * t_fac64 * 1e-3)}
    ach = FLOPS_PER_TRIPLE * triples / (t_od * 1e-3) / 1e12
    roof_asm = {"kernel": "k_dyn_offdiag_h (dynamic H assembly)", "bound": "valu",
                "pipe": "FP64 vector ALU", "achieved": ach, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                "traffic": OFFDIAG_TRAFFIC_BYTES, "avg_ms": t_od,
                "flops_per_triple": FLOPS_PER_TRIPLE, "triples": triples,
                "counted_valu_flops": OFFDIAG_COUNTED_FLOPS,
                "counted_frac": OFFDIAG_COUNTED_FLOPS / (t_od * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    return rec, roof_lu, roof_fp64, roof_asm


def serve_slice(args, api, decks, n_requests, rate, check):
    """Poisson arrivals of the serving mix through the request loop: every
    slot warmed on the largest class first (its workspace then never grows)
    and on every class, then the timed stream."""
    meshes, kodes, reqs = decks.serving_stream(n_requests, rate, seed=0)
    srv = api.Server(0, nslots=args.slots, mode=api.MODE_MIXED)
    try:
        big = len(meshes) - 1
        order = [big] + list(range(big))
        for c in order:
            m = meshes[c]
            z = np.zeros((3 * m.nbe, 4), np.complex128)
            z[np.arange(4), np.arange(4)] = 1.0
            for w in [srv.submit(m, kodes[c], z, (m.ge, m.rnu, 0.01, 1.0), 1.0)
                      for _ in range(args.slots)]:
                srv.wait(w)
        warm = srv.stats()
        log("serve: warm-up done (%d graphs captured, %d ahead)" %
            (warm["graph_captures"], warm["graph_precaptures"]))
        ids = [None] * len(reqs)

        def feeder():
            t0 = time.perf_counter()
            for k, r in enumerate(reqs):
                dt = r["t_arrival"] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                m = meshes[r["cls"]]
                ids[k] = srv.submit(m, kodes[r["cls"]], r["zdfi"], (m.ge, m.rnu, r["dam"], 1.0),
                                    r["omega"])

        t0 = time.perf_counter()
        th = threading.Thread(target=feeder)
        th.start()
        outs = []
        for k in range(len(reqs)):
            while ids[k] is None:
                time.sleep(0.0005)
            outs.append(srv.wait(ids[k]))
            if (k + 1) % 250 == 0:
                log("serve: %d/%d requests done" % (k + 1, len(reqs)))
        th.join()
        wall = time.perf_counter() - t0
        end = srv.stats()
    finally:
        srv.close()
    lat = [o[1]["latency_ms"] for o in outs]
    # accuracy: a sample of requests re-solved alone on the plain FP64 path
    errs = []
    ctx = api.Context(0)
    try:
        for k in range(0, len(reqs), max(1, len(reqs) // check)):
            r = reqs[k]
            m = meshes[r["cls"]]
            ctx.send_shared_data(m)
            ctx.ghmatece(m.ge, m.rnu)
            ctx.send_load_cases(r["zdfi"], kodes[r["cls"]])
            x64, _ = ctx.solve(api.material(m.ge, m.rnu, r["dam"], 1.0), r["omega"],
                               api.MODE_FP64)
            errs.append(rel_err(outs[k][0].reshape(x64.shape, order="F"), x64))
    finally:
        ctx.close()
    solves = sum(r["zdfi"].shape[1] for r in reqs)
    log("serve: %d requests in %.2f s, p50 %.1f ms, p99 %.1f ms, captures after warm-up %d" %
        (len(reqs), wall, percentile(lat, 50), percentile(lat, 99),
         end["graph_captures"] - warm["graph_captures"]))
    return {
        "workload": "serving stream: %d requests, Poisson arrivals at %.1f/s, meshes 16x15..64x60 "
                    "(240-3840 BE), omega U[0.1,10], damping U[0.005,0.05], 1-4 load cases" %
                    (len(reqs), rate),
        "requests": len(reqs), "requests_per_s": len(reqs) / wall,
        "load_cases_per_s": solves / wall,
        "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
        "latency_note": "submit-to-done latency per request (queueing included)",
        "service_p50_ms": percentile([o[1]["service_ms"] for o in outs], 50),
        "service_p99_ms": percentile([o[1]["service_ms"] for o in outs], 99),
        "graph_captures_after_warmup": end["graph_captures"] - warm["graph_captures"],
        "graph_replays": end["graph_replays"] - warm["graph_replays"],
        "graph_precaptures_after_warmup": end["graph_precaptures"] - warm["graph_precaptures"],
        "peak_charged_bytes": end["peak_charged_bytes"], "budget_bytes": end["budget_bytes"],
        "over_charge": end["over_charge"],
        "max_rel_err_mixed_vs_fp64": max(errs), "checked_requests": len(errs),
        "fallbacks_to_fp64": int(sum(o[1]["fell_back"] for o in outs)),
    }


def cpu_baseline(decks, mesh, dyn, fr):
    """The C oracle solving one frequency of the sweep (GHMATECD + LAPACK
    zgesv, the unit a GPU step does) on the host cores.  Returns the record
    and the reference solution for the accuracy check: LAPACK on the
    oracle's matrix with its diagonal blocks from the extended-precision
    build (the FP64 difference kernels of those blocks lose digits at small
    k R; not timed)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    orc = oracle.Oracle("c")
    threads = host_threads()
    hd, gd, consts = orc.ghmatece(mesh)  # resident before the timed GPU region too
    zmat = orc.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    zdfi = dyn.bc.astype(np.complex128)
    t0 = time.perf_counter()
    zh, zfi = orc.ghmatecd(mesh, hd, gd, consts, dyn.kode, zdfi, fr, zmat)
    t1 = time.perf_counter()
    x = orc.linsolve(zh, zfi)
    t2 = time.perf_counter()
    del zh
    tcpu = t2 - t0
    rec = {"value": 1.0 / tcpu, "unit": "solves/s", "cores": threads, "kind": "port",
           "sample": "one frequency (omega=%.4g) of the sweep: GHMATECD %.1f s + LAPACK zgesv "
                     "%.1f s by the C oracle (OpenMP) and numpy on %d threads" %
                     (fr, t1 - t0, t2 - t1, threads),
           "reference_over_port_time": REF_OVER_ORACLE,
           "reference_estimate_solves_per_s": 1.0 / (tcpu * REF_OVER_ORACLE)}
    log("cpu oracle: %.1f s per solve on %d threads" % (tcpu, threads))
    zh, zfi, _ = orc.ghmatecd_accurate(mesh, hd, gd, consts, dyn.kode, zdfi, fr, zmat)
    x_acc = orc.linsolve(zh, zfi)
    return rec, x_acc, x


# ---------------------------------------------------------------------------
def workload_default(args, api, decks):
    mesh, dyn = decks.config_problem("c4")
    log("frequency sweep: nbe=%d dof=%d npg=%d, %d frequencies" %
        (mesh.nbe, 3 * mesh.nbe, mesh.npg, args.steps))
    single, roof_lu, roof_fp64, roof_asm = single_solve(args, api, decks, args.single_steps)

    freqs = sweep_frequencies(dyn, args.steps)
    p = (mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    bc = dyn.bc.astype(np.complex128)
    items = [(mesh, dyn.kode, bc, p, f) for f in freqs]
    warm = [(mesh, dyn.kode, bc, p, freqs[(k * 37) % len(freqs)])
            for k in range(max(1, args.warmup) * args.slots)]
    xs, infos, wall, w0, w1 = stream_through_server(api, api.MODE_MIXED, args.slots, items,
                                                    "mixed sweep", warm_items=warm)
    x64, infos64, wall64, _, _ = stream_through_server(api, api.MODE_FP64, args.slots, items,
                                                       "fp64 sweep", warm_items=warm)
    errs = [rel_err(a, b) for a, b in zip(xs, x64)]
    value = len(items) / wall
    fp64_value = len(items) / wall64
    svc = [i["service_ms"] for i in infos]
    engine = sweep_engine(api, mesh, dyn, freqs, x64)

    serve = serve_slice(args, api, decks, args.requests, args.rate, args.check) \
        if args.requests > 0 else None

    cpu, err_oracle, err_fp64_oracle = None, None, None
    if not args.no_cpu:
        k = int(np.argmin(np.abs(np.log(np.asarray(freqs) / 2.0))))
        cpu, xo, xo64 = cpu_baseline(decks, mesh, dyn, freqs[k])
        err_oracle = rel_err(xs[k], xo)
        err_fp64_oracle = rel_err(xo64, xo)
        log("mixed vs oracle at omega=%.4g: %.2e (FP64 oracle vs accurate oracle %.2e)" %
            (freqs[k], err_oracle, err_fp64_oracle))

    out = base_record(args, {"workload": "frequency sweep: flat patch 48x45 over [-5,5]^2 "
                                         "(2160 BE + 5 EE, 6480 complex DOF), 8x8 Gauss, %d "
                                         "frequencies log-spaced in [0.1, 10], one full assembly "
                                         "+ factor + refine per frequency" % len(items)},
                      value, 1e3 / value, len(items))
    out.update({
        "p50_ms": percentile(svc, 50), "p99_ms": percentile(svc, 99),
        "latency_note": "p50/p99: per-frequency service time in the request loop (assembly + "
                        "factor + refine); the serving stream's submit-to-done p50/p99 are under "
                        "'serve'",
        "fp64": {"value": fp64_value, "unit": "solves/s",
                 "p50_ms": percentile([i["service_ms"] for i in infos64], 50),
                 "p99_ms": percentile([i["service_ms"] for i in infos64], 99)},
        "speedup_mixed_vs_fp64": value / fp64_value,
        "max_rel_err_mixed_vs_fp64": max(errs),
        "rel_err_mixed_vs_oracle": err_oracle,
        "rel_err_fp64_oracle_vs_accurate_oracle": err_fp64_oracle,
        "gmres_iters_max": int(max(i["gmres_iters"] for i in infos)),
        "fallbacks_to_fp64": int(sum(i["fell_back"] for i in infos)),
        "graph_captures_after_warmup": w1["graph_captures"] - w0["graph_captures"],
        "single_solve": single,
        "sweep_engine": engine,
        "serve": serve,
        "roofline": engine.get("roofline") or roof_lu,
        "roofline_single_factor": roof_lu,
        "roofline_assembly": roof_asm,
        "roofline_fp64_factor": roof_fp64,
        "cpu_baseline": cpu,
    })
    return out


def workload_solve(args, api, decks):
    single, roof_lu, roof_fp64, roof_asm = single_solve(args, api, decks, args.steps)
    value = 1e3 / single["mixed"]["p50_ms"]
    out = base_record(args, {"workload": single["workload"]}, value, single["mixed"]["p50_ms"],
                      args.steps)
    out["config"]["parallelism"] = "single GPU, one context"
    out.update({"p50_ms": single["mixed"]["p50_ms"], "p99_ms": single["mixed"]["p99_ms"],
                "fp64": single["fp64"], "single_solve": single, "roofline": roof_lu,
                "roofline_assembly": roof_asm, "roofline_fp64_factor": roof_fp64})
    return out


def sweep_engine(api, mesh, dyn, freqs, x64):
    """The same sweep through bemx_solve_sweep on one context: batches of
    frequencies assembled (two per launch) on a second stream beside the
    previous batch's batched solve; warmed by an 8-frequency sweep
    (workspace, captured factorisation graphs).  Every result is compared
    with the request loop's FP64 one."""
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx = api.Context(0)
    try:
        ctx.send_shared_data(mesh)
        ctx.ghmatece(mesh.ge, mesh.rnu)
        ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
        ctx.solve_sweep(zmat, freqs[:8], api.MODE_MIXED)
        t0 = time.perf_counter()
        xe, st = ctx.solve_sweep(zmat, freqs, api.MODE_MIXED)
        wall = time.perf_counter() - t0
        t_fac = ctx.last_stats()["lu_factor_ms"]  # summed over the sweep's batches
        # one batch alone (its assembly done before it): the batched
        # factorisation with nothing beside it, three times
        nb = min(8, len(freqs))
        t_one = []
        for _ in range(3):
            ctx.solve_sweep(zmat, freqs[:nb], api.MODE_MIXED)
            t_one.append(ctx.last_stats()["lu_factor_ms"])
        t_one = float(np.median(t_one))
    finally:
        ctx.close()
    err = max(rel_err(a, b) for a, b in zip(xe, x64))
    log("sweep engine: %d solves in %.3f s on one context (%.2f solves/s), max err vs fp64 %.2e, "
        "factorisations %.1f ms in all" % (len(freqs), wall, len(freqs) / wall, err, t_fac))
    n = 3 * mesh.nbe
    fl = 8.0 * n ** 3  # 3 BF16 products x 8 n^3 / 3 per factorisation
    ach = nb * fl / (t_one * 1e-3) / 1e12 if t_one > 0 else None
    ach_sweep = len(freqs) * fl / (t_fac * 1e-3) / 1e12 if t_fac > 0 else None
    roof = {"kernel": "batched mixed factorisation (%d problems in every launch of k_panel_inv, "
                      "k_cgemm_update_inverse, k_cgemm_update2, ...; split-BF16x3 MFMA trailing "
                      "update)" % nb,
            "bound": "mfma", "achieved": ach, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / BF16_PEAK_TFLOPS if ach else None, "traffic": None,
            "avg_ms": t_one, "problems_per_launch": nb, "algorithmic_flops_per_problem": fl,
            "in_sweep_achieved": ach_sweep, "in_sweep_ms_per_solve": t_fac / len(freqs),
            "note": "device time of one batch's factorisation (HIP events around it, the batch "
                    "assembled before it; median of 3); in_sweep_*: the same over the timed sweep, "
                    "where the next batch's assembly runs beside each factorisation; flops = 3 "
                    "BF16 products x 8n^3/3 per problem, n = 6480"}
    return {"value": len(freqs) / wall, "unit": "solves/s", "wall_s": wall,
            "max_rel_err_vs_fp64": err, "gmres_iters_max": int(st[:, 0].max()),
            "fallbacks_to_fp64": int(st[:, 2].sum()), "factor_ms_total": t_fac,
            "roofline": roof,
            "note": "bemx_solve_sweep on one context: batches of frequencies assembled two "
                    "per launch (k_dyn_offdiag_h2) on a second stream beside the previous "
                    "batch's factorisations and refinements, which share every launch"}


def warm_main_thread(api, decks):
    """One solve of the smallest serving class on the main thread, so the
    process's first HIP work is not made by the request loop's workers."""
    meshes, kodes, _ = decks.serving_stream(1, 1.0, seed=0)
    m = meshes[0]
    ctx = api.Context(0)
    try:
        ctx.send_shared_data(m)
        ctx.ghmatece(m.ge, m.rnu)
        z = np.zeros((3 * m.nbe, 1), np.complex128)
        z[0, 0] = 1.0
        ctx.send_load_cases(z, kodes[0])
        for mode in (api.MODE_MIXED, api.MODE_FP64):
            ctx.solve(api.material(m.ge, m.rnu, 0.01, 1.0), 1.0, mode)
    finally:
        ctx.close()


def workload_serve(args, api, decks):
    warm_main_thread(api, decks)
    serve = serve_slice(args, api, decks, args.requests, args.rate, args.check)
    out = base_record(args, {"workload": serve["workload"]}, serve["load_cases_per_s"],
                      1e3 / serve["load_cases_per_s"], serve["requests"])
    out.update({"p50_ms": serve["p50_ms"], "p99_ms": serve["p99_ms"], "serve": serve,
                "value_note": "load cases solved per second (1-4 per request)"})
    return out


def workload_large(args, api, decks):
    mesh, dyn = decks.config_problem("c3", seed=args.seed)
    fr = float(dyn.freqs[0])
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx = api.Context(0)
    try:
        ctx.send_shared_data(mesh)
        ctx.ghmatece(mesh.ge, mesh.rnu)
        ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
        log("workload large: nbe=%d dof=%d nu=%.3f mu=%.3f" %
            (mesh.nbe, 3 * mesh.nbe, mesh.rnu, mesh.ge))
        ctx.solve(zmat, fr, api.MODE_MIXED)
        lat, fac = [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            xm, _ = ctx.solve(zmat, fr, api.MODE_MIXED)
            lat.append((time.perf_counter() - t0) * 1e3)
            stm = ctx.last_stats()
            fac.append(stm["lu_factor_ms"])
            log("mixed %.1f ms (assembly %.1f, factor %.1f, refine %.1f, %d GMRES it)" %
                (lat[-1], stm["assembly_ms"], stm["lu_factor_ms"], stm["lu_refine_ms"],
                 stm["gmres_iters"]))
        t0 = time.perf_counter()
        x64, _ = ctx.solve(zmat, fr, api.MODE_FP64)
        t64 = time.perf_counter() - t0
    finally:
        ctx.close()
    n = 3 * mesh.nbe
    t_fac = float(np.mean(fac))
    value = 1e3 / float(np.mean(lat))
    out = base_record(args, {"workload": "large single solve: flat patch 96x90 over [-10,10]^2 "
                                         "(8640 BE, 25920 complex DOF), omega=0.5, 6x6 Gauss"},
                      value, float(np.mean(lat)), args.steps)
    out["config"]["parallelism"] = "single GPU, one context"
    ach = 8.0 * n ** 3 / (t_fac * 1e-3) / 1e12
    out.update({"p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                "fp64": {"value": 1.0 / t64, "unit": "solves/s"},
                "speedup_mixed_vs_fp64": value * t64,
                "max_rel_err_mixed_vs_fp64": rel_err(xm, x64),
                "roofline": {"kernel": "mixed factorisation (split-BF16x3 MFMA)", "bound": "mfma",
                             "achieved": ach, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": ach / BF16_PEAK_TFLOPS, "traffic": None, "avg_ms": t_fac}})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256,
                    help="sweep frequencies timed (default: all 256); other workloads: solves")
    ap.add_argument("--warmup", type=int, default=2, help="untimed requests per slot")
    ap.add_argument("--workload", default="sweep", choices=["sweep", "solve", "serve", "large"])
    ap.add_argument("--slots", type=int, default=2, help="request-loop slots")
    ap.add_argument("--single-steps", type=int, default=10,
                    help="single-solve latency steps beside the sweep")
    ap.add_argument("--requests", type=int, default=None,
                    help="serving stream requests (default 300 beside the sweep, 2000 alone)")
    ap.add_argument("--rate", type=float, default=20.0, help="serving arrivals per second")
    ap.add_argument("--check", type=int, default=10, help="serving requests re-solved in FP64")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the solve is not sharded across devices "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.requests is None:
        args.requests = 2000 if args.workload == "serve" else 300
    if args.workload in ("solve", "large") and args.steps == 256:
        args.steps = 10 if args.workload == "solve" else 5
    from bemx import api, decks
    fn = {"sweep": workload_default, "solve": workload_solve, "serve": workload_serve,
          "large": workload_large}[args.workload]
    print(json.dumps(fn(args, api, decks)), flush=True)


if __name__ == "__main__":
    main()
