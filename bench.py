"""Benchmark: BEM solves/s and p50/p99 latency, mixed precision vs plain FP64.

Workloads (--workload):
  solve  (default) the 48 x 45 flat free-surface patch (2160 boundary
         elements + 5 enclosing, 6480 complex DOF), mu = 1, nu = 0.25,
         rho = 1, damping 0.01, omega = 2.0, 8 x 8 Gauss points, unit
         vertical traction on the central 2 x 2 elements.  A step is one
         solve at one frequency with everything resident in HBM before the
         timed region: dynamic assembly of ZH and the RHS (FP64), factor,
         refine, solve.  Single-solve latency is measured on one context
         (mixed and FP64 interleaved, every mixed result compared with the
         FP64 result of the same step); `value` is the throughput of the
         request loop (bemx_server) running the same solves on --slots
         concurrent slots.
  sweep  the same mesh, 256 frequencies log-spaced in [0.1, 10] rad/s, each
         a full assembly + factor + refine, through the request loop.
  serve  2000 requests with Poisson arrivals: meshes from {16x15, 24x22,
         32x30, 48x45, 64x60}, omega ~ U[0.1, 10], damping ~ U[0.005, 0.05],
         1..4 load cases; p50/p99 of submit-to-done latency.
  large  96 x 90 patch over [-10, 10]^2 (8640 elements, 25920 complex DOF),
         omega = 0.5, 6 x 6 Gauss, nu and mu drawn by seed; mixed path with
         the plain FP64 path as reference.
The mixed path (FP32 LU with split-BF16x3 MFMA trailing update + FP64
GMRES-IR) is the headline; the plain FP64 partial-pivot path is measured
beside it and the accuracy contract is <= 1e-10 relative to FP64.

The CPU baseline (default workload) is the oracle (C, OpenMP, FP64, LAPACK
zgesv via numpy) on the same problem, one solve, on the host's cores.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bem-mececcs_amd"))

import numpy as np  # noqa: E402

METRIC = ("BEM solves/s + p50/p99 latency at ≤1e-10 err vs FP64, mixed vs plain FP64")

# FP64 operations per (collocation point, element, Gauss point) triple of the
# off-diagonal assembly kernel k_dyn_offdiag_h, counted from the factored
# formula it evaluates: 234 add/mul (an fma counts 2) + 8 special functions
# (sqrt, 1/R, 2 exp, 2 sin, 2 cos) counted as one each, however they are
# evaluated (the Newton steps that refine 1/R and the table-and-polynomial
# evaluation of the exponentials are implementation details, not counted).
FLOPS_PER_TRIPLE = 242
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
# HBM bytes per launch of k_dyn_offdiag_h on the default workload, from the
# rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a plain streaming copy of
# known size) recorded in profiles/pmc_traffic_c2.md: 48.4 MB read + 671.6 MB
# written (the ZH store itself is 671.8 MB).
OFFDIAG_TRAFFIC_BYTES_C2 = 720.0e6
HBM_PEAK_GBS = 8000.0
DTYPE = "f64 assembly + f32/bf16x3 factor + f64 refine"


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def base_record(args, workload, value, ms_per_step, steps):
    return {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE,
        "data": "synthetic (generated flat-patch meshes)",
        "config": dict(workload, parallelism="single GPU, %s" % (
            "%d request-loop slots" % args.slots if args.slots else "one context")),
    }


def stream_through_server(api, mode, slots, items, label, warm_items=()):
    """Submit every item at once to a request loop of `slots` slots; items
    are (mesh, kode, zdfi, (ge, rnu, dam, rho), omega).  warm_items run
    first on the same loop, untimed, so each slot has its workspace, its
    mesh and static pass and its captured factorisation before the timed
    items.  Returns (x list, info list, wall seconds) of the timed items."""
    srv = api.Server(0, nslots=slots, mode=mode)
    try:
        if warm_items:
            for i in [srv.submit(*it) for it in warm_items]:
                srv.wait(i)
        t0 = time.perf_counter()
        ids = [srv.submit(*it) for it in items]
        outs = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    svc = [o[1]["service_ms"] for o in outs]
    log("%s: %d solves in %.3f s on %d slots (%.2f solves/s), service p50 %.2f ms" %
        (label, len(items), wall, slots, len(items) / wall, percentile(svc, 50)))
    return [o[0] for o in outs], [o[1] for o in outs], wall


def one_context_sweep(api, mesh, dyn, freqs, mode, label):
    """The frequencies of a sweep solved back to back on one context (no
    request loop); returns (x list, info list, wall seconds)."""
    ctx = api.Context(0)
    try:
        ctx.send_shared_data(mesh)
        ctx.ghmatece(mesh.ge, mesh.rnu)
        ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
        zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
        ctx.solve(zmat, freqs[0], mode)  # warm
        xs, infos = [], []
        t0 = time.perf_counter()
        for f in freqs:
            t1 = time.perf_counter()
            x, _ = ctx.solve(zmat, f, mode)
            st = ctx.last_stats()
            xs.append(x)
            infos.append({"service_ms": (time.perf_counter() - t1) * 1e3,
                          "gmres_iters": st["gmres_iters"], "fell_back": float(st["fell_back"])})
        wall = time.perf_counter() - t0
    finally:
        ctx.close()
    log("%s: %d solves in %.3f s on one context (%.2f solves/s)" %
        (label, len(freqs), wall, len(freqs) / wall))
    return xs, infos, wall


# ---------------------------------------------------------------------------
def workload_solve(args, api, decks):
    mesh, dyn = decks.config_problem("c2")
    fr = float(dyn.freqs[0])
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx = api.Context(0)
    ctx.send_shared_data(mesh)
    ctx.ghmatece(mesh.ge, mesh.rnu)
    ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
    log("workload c2: nbe=%d n=%d dof=%d npg=%d omega=%g" %
        (mesh.nbe, mesh.n, 3 * mesh.nbe, mesh.npg, fr))
    for _ in range(args.warmup):
        ctx.solve(zmat, fr, api.MODE_FP64)
        ctx.solve(zmat, fr, api.MODE_MIXED)
    log("warmup done")

    # ---- single-solve latency: mixed and plain FP64, interleaved
    lat_mixed, lat_fp64, errs, iters, fallbacks, offdiag_ms = [], [], [], [], 0, []
    for step in range(args.steps):
        t0 = time.perf_counter()
        x64, _ = ctx.solve(zmat, fr, api.MODE_FP64)
        t1 = time.perf_counter()
        st64 = ctx.last_stats()
        xm, _ = ctx.solve(zmat, fr, api.MODE_MIXED)
        t2 = time.perf_counter()
        stm = ctx.last_stats()
        lat_fp64.append((t1 - t0) * 1e3)
        lat_mixed.append((t2 - t1) * 1e3)
        offdiag_ms.extend([st64["offdiag_ms"], stm["offdiag_ms"]])
        iters.append(stm["gmres_iters"])
        fallbacks += int(stm["fell_back"])
        errs.append(rel_err(xm, x64))
        log("step %d: fp64 %.2f ms (assembly %.2f, factor+solve %.2f), mixed %.2f ms "
            "(assembly %.2f, factor+refine %.2f, %d GMRES it), rel err %.2e" %
            (step, lat_fp64[-1], st64["assembly_ms"], st64["factor_ms"], lat_mixed[-1],
             stm["assembly_ms"], stm["factor_ms"], stm["gmres_iters"], errs[-1]))
    ctx.close()

    # ---- throughput: the same solves through the request loop
    item = (mesh, dyn.kode, dyn.bc.astype(np.complex128), (mesh.ge, mesh.rnu, dyn.dam, dyn.rho),
            fr)
    n_stream = max(args.steps, 2 * args.slots)
    warm = [item] * (max(1, args.warmup) * args.slots)
    xs, infos, wall = stream_through_server(api, api.MODE_MIXED, args.slots, [item] * n_stream,
                                            "mixed stream", warm_items=warm)
    n64 = max(2, n_stream // 2)
    x64s, _, wall64 = stream_through_server(api, api.MODE_FP64, args.slots, [item] * n64,
                                            "fp64 stream", warm_items=[item] * args.slots)
    errs.extend(rel_err(x, x64s[0]) for x in xs)
    value = n_stream / wall
    fp64_value = n64 / wall64

    triples = mesh.nbe * mesh.nbe * mesh.npg ** 2
    od = float(np.mean(offdiag_ms))
    achieved = FLOPS_PER_TRIPLE * triples / (od * 1e-3) / 1e12

    cpu = None
    err_oracle = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        orc = oracle.Oracle("c")
        threads = os.cpu_count() or 1
        t0 = time.perf_counter()
        ref = orc.solve_problem(mesh, dyn)
        tcpu = time.perf_counter() - t0
        err_oracle = rel_err(xm, ref["x"])
        cpu = {"value": 1.0 / tcpu, "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": "one full solve of the same workload (GHMATECE, GHMATECD, LAPACK "
                         "zgesv, INTEREC at 64 points) by the C oracle with OpenMP on %d "
                         "threads: %.1f s" % (threads, tcpu)}
        log("cpu oracle solve %.1f s; mixed vs oracle rel err %.2e" % (tcpu, err_oracle))

    out = base_record(args, {"workload": "flat patch 48x45 (2160 BE + 5 EE, 6480 complex DOF), "
                                         "omega=2.0, 8x8 Gauss, one frequency per solve"},
                      value, 1e3 / value, n_stream)
    out.update({
        "p50_ms": percentile(lat_mixed, 50),
        "p99_ms": percentile(lat_mixed, 99),
        "latency_note": "p50/p99: one solve alone on one context; value: request-loop "
                        "throughput on %d slots" % args.slots,
        "stream_service_p50_ms": percentile([i["service_ms"] for i in infos], 50),
        "stream_service_p99_ms": percentile([i["service_ms"] for i in infos], 99),
        "single_context_solves_per_s": 1e3 / float(np.mean(lat_mixed)),
        "fp64": {"value": fp64_value, "unit": "solves/s", "p50_ms": percentile(lat_fp64, 50),
                 "p99_ms": percentile(lat_fp64, 99),
                 "single_context_solves_per_s": 1e3 / float(np.mean(lat_fp64))},
        "speedup_mixed_vs_fp64": value / fp64_value,
        "max_rel_err_mixed_vs_fp64": max(errs),
        "rel_err_mixed_vs_oracle": err_oracle,
        "gmres_iters": int(np.max(iters)),
        "fallbacks_to_fp64": fallbacks,
        "roofline": {"kernel": "k_dyn_offdiag_h (dynamic H assembly)",
                     "pipe": "FP64 vector ALU", "bound": "mfma", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": OFFDIAG_TRAFFIC_BYTES_C2,
                     "traffic_unit": "bytes per launch (rocprofv3 PMC)",
                     "avg_ms": od, "flops_per_triple": FLOPS_PER_TRIPLE,
                     "triples": triples},
        "cpu_baseline": cpu,
    })
    return out


def workload_sweep(args, api, decks):
    mesh, dyn = decks.config_problem("c4")
    freqs = [float(f) for f in dyn.freqs][: args.nfreq]
    p = (mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    bc = dyn.bc.astype(np.complex128)
    items = [(mesh, dyn.kode, bc, p, f) for f in freqs]
    if args.slots == 0:
        # no request loop: one context in this thread, frequencies back to back
        xs, infos, wall = one_context_sweep(api, mesh, dyn, freqs, api.MODE_MIXED, "mixed sweep")
        x64, _, wall64 = one_context_sweep(api, mesh, dyn, freqs, api.MODE_FP64, "fp64 sweep")
    else:
        xs, infos, wall = stream_through_server(api, api.MODE_MIXED, args.slots, items,
                                                "mixed sweep", warm_items=items[: args.slots])
        x64, _, wall64 = stream_through_server(api, api.MODE_FP64, args.slots, items, "fp64 sweep",
                                               warm_items=items[: args.slots])
    errs = [rel_err(a, b) for a, b in zip(xs, x64)]
    svc = [i["service_ms"] for i in infos]
    value = len(items) / wall
    out = base_record(args, {"workload": "frequency sweep: flat patch 48x45 (6480 complex DOF), "
                                         "%d frequencies log-spaced in [0.1, 10]" % len(items)},
                      value, 1e3 / value, len(items))
    out.update({
        "p50_ms": percentile(svc, 50), "p99_ms": percentile(svc, 99),
        "latency_note": "per-frequency service time (assembly + factor + refine) in the loop",
        "fp64": {"value": len(items) / wall64, "unit": "solves/s"},
        "speedup_mixed_vs_fp64": value / (len(items) / wall64),
        "max_rel_err_mixed_vs_fp64": max(errs),
        "gmres_iters_max": int(max(i["gmres_iters"] for i in infos)),
        "fallbacks_to_fp64": int(sum(i["fell_back"] for i in infos)),
    })
    return out


def workload_serve(args, api, decks):
    meshes, kodes, reqs = decks.serving_stream(args.requests, args.rate, seed=0)
    srv = api.Server(0, nslots=args.slots, mode=api.MODE_MIXED)
    # warm every size class on every slot's workspace (4 load cases: the largest)
    warm = []
    for c, m in enumerate(meshes):
        z = np.zeros((3 * m.nbe, 4), np.complex128)
        z[np.arange(4), np.arange(4)] = 1.0
        for _ in range(args.slots):
            warm.append(srv.submit(m, kodes[c], z, (m.ge, m.rnu, 0.01, 1.0), 1.0))
    for w in warm:
        srv.wait(w)
    log("warm done")
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
            log("%d/%d requests done" % (k + 1, len(reqs)))
    th.join()
    wall = time.perf_counter() - t0
    srv.close()
    lat = [o[1]["latency_ms"] for o in outs]
    # accuracy: a sample of requests re-solved alone on the plain FP64 path
    errs = []
    ctx = api.Context(0)
    for k in range(0, len(reqs), max(1, len(reqs) // args.check)):
        r = reqs[k]
        m = meshes[r["cls"]]
        ctx.send_shared_data(m)
        ctx.ghmatece(m.ge, m.rnu)
        ctx.send_load_cases(r["zdfi"], kodes[r["cls"]])
        x64, _ = ctx.solve(api.material(m.ge, m.rnu, r["dam"], 1.0), r["omega"], api.MODE_FP64)
        xm = outs[k][0]
        errs.append(rel_err(xm.reshape(x64.shape, order="F"), x64))
    ctx.close()
    value = len(reqs) / wall
    out = base_record(args, {"workload": "serving stream: %d requests, Poisson arrivals at %.1f/s, "
                                         "meshes 16x15..64x60 (240-3840 BE), omega U[0.1,10], "
                                         "damping U[0.005,0.05], 1-4 load cases" %
                                         (len(reqs), args.rate)},
                      value, 1e3 / value, len(reqs))
    out.update({
        "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
        "latency_note": "submit-to-done latency per request (queueing included)",
        "max_rel_err_mixed_vs_fp64": max(errs), "checked_requests": len(errs),
        "fallbacks_to_fp64": int(sum(o[1]["fell_back"] for o in outs)),
        "offered_rate": args.rate,
    })
    return out


def workload_large(args, api, decks):
    mesh, dyn = decks.config_problem("c3", seed=args.seed)
    fr = float(dyn.freqs[0])
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx = api.Context(0)
    ctx.send_shared_data(mesh)
    ctx.ghmatece(mesh.ge, mesh.rnu)
    ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
    log("workload c3: nbe=%d dof=%d nu=%.3f mu=%.3f" % (mesh.nbe, 3 * mesh.nbe, mesh.rnu, mesh.ge))
    ctx.solve(zmat, fr, api.MODE_MIXED)
    lat = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        xm, _ = ctx.solve(zmat, fr, api.MODE_MIXED)
        lat.append((time.perf_counter() - t0) * 1e3)
        stm = ctx.last_stats()
        log("mixed %.1f ms (assembly %.1f, factor+refine %.1f, %d GMRES it)" %
            (lat[-1], stm["assembly_ms"], stm["factor_ms"], stm["gmres_iters"]))
    t0 = time.perf_counter()
    x64, _ = ctx.solve(zmat, fr, api.MODE_FP64)
    t64 = time.perf_counter() - t0
    ctx.close()
    value = 1e3 / float(np.mean(lat))
    out = base_record(args, {"workload": "large single solve: flat patch 96x90 over [-10,10]^2 "
                                         "(8640 BE, 25920 complex DOF), omega=0.5, 6x6 Gauss"},
                      value, float(np.mean(lat)), args.steps)
    out.update({"p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                "fp64": {"value": 1.0 / t64, "unit": "solves/s"},
                "speedup_mixed_vs_fp64": value * t64,
                "max_rel_err_mixed_vs_fp64": rel_err(xm, x64)})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="solve", choices=["solve", "sweep", "serve", "large"])
    ap.add_argument("--slots", type=int, default=2,
                    help="request-loop slots (sweep: 0 = one context, no request loop)")
    ap.add_argument("--nfreq", type=int, default=256, help="sweep: number of frequencies")
    ap.add_argument("--requests", type=int, default=2000, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: arrivals per second")
    ap.add_argument("--check", type=int, default=20, help="serve: requests re-solved in FP64")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the solve is not sharded across devices "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    from bemx import api, decks
    fn = {"solve": workload_solve, "sweep": workload_sweep, "serve": workload_serve,
          "large": workload_large}[args.workload]
    print(json.dumps(fn(args, api, decks)))


if __name__ == "__main__":
    main()
