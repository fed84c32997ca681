"""Benchmark: BEM solves/s and p50/p99 latency, mixed precision vs plain FP64.

Workload: the 48 x 45 flat free-surface patch (2160 boundary elements + 5
enclosing, 6480 complex DOF), mu = 1, nu = 0.25, rho = 1, damping 0.01,
omega = 2.0, 8 x 8 Gauss points, unit vertical traction on the central 2 x 2
elements.  One step = one solve at one frequency, everything resident in HBM
before the timed region: dynamic assembly of ZH and the RHS (FP64), factor,
refine, solve.  The mixed path (FP32 LU with split-BF16x3 MFMA trailing
update + FP64 GMRES-IR) is the headline `value`; the plain FP64 partial-pivot
path is timed beside it, and every mixed result is compared with the FP64
result of the same step (the accuracy contract is <= 1e-10 relative).

The CPU baseline is the oracle (C, OpenMP, FP64, LAPACK zgesv via numpy) on
the same problem, one solve, on the host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bem-mececcs_amd"))

import numpy as np  # noqa: E402

METRIC = ("BEM solves/s + p50/p99 latency at ≤1e-10 err vs FP64, mixed vs plain FP64")

# FP64 operations per (collocation point, element, Gauss point) triple of the
# off-diagonal assembly kernel k_dyn_offdiag_h, counted from its source:
# 234 add/mul (an fma counts 2) + 8 special functions (sqrt, 1/R, 2 exp,
# 2 sin, 2 cos) counted as one each; the Newton steps that refine 1/R are an
# implementation detail and not counted.
FLOPS_PER_TRIPLE = 242
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the solve is not sharded across devices "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from bemx import api, decks

    mesh, dyn = decks.config_problem(args.config)
    fr = float(dyn.freqs[0])
    zmat = api.material(mesh.ge, mesh.rnu, dyn.dam, dyn.rho)
    ctx = api.Context(0)
    ctx.send_shared_data(mesh)
    ctx.ghmatece(mesh.ge, mesh.rnu)
    ctx.send_dyn_data(dyn.bc.astype(np.complex128), dyn.kode)
    print("workload %s: nbe=%d n=%d dof=%d npg=%d omega=%g" %
          (args.config, mesh.nbe, mesh.n, 3 * mesh.nbe, mesh.npg, fr), file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        ctx.solve(zmat, fr, api.MODE_FP64)
        ctx.solve(zmat, fr, api.MODE_MIXED)
    print("warmup done", file=sys.stderr, flush=True)

    # ---- timed: mixed path (headline) and plain FP64 path, interleaved
    lat_mixed, lat_fp64, errs, iters, fallbacks, offdiag_ms = [], [], [], [], 0, []
    t_mixed = t_fp64 = 0.0
    for step in range(args.steps):
        t0 = time.perf_counter()
        x64, _ = ctx.solve(zmat, fr, api.MODE_FP64)
        t1 = time.perf_counter()
        st64 = ctx.last_stats()
        t2 = time.perf_counter()
        xm, _ = ctx.solve(zmat, fr, api.MODE_MIXED)
        t3 = time.perf_counter()
        stm = ctx.last_stats()
        t_fp64 += t1 - t0
        t_mixed += t3 - t2
        lat_fp64.append((t1 - t0) * 1e3)
        lat_mixed.append((t3 - t2) * 1e3)
        offdiag_ms.extend([st64["offdiag_ms"], stm["offdiag_ms"]])
        iters.append(stm["gmres_iters"])
        fallbacks += int(stm["fell_back"])
        errs.append(float(np.max(np.abs(xm - x64)) / np.max(np.abs(x64))))
        print("step %d: fp64 %.2f ms (assembly %.2f, factor+solve %.2f), mixed %.2f ms "
              "(assembly %.2f, factor+refine %.2f, %d GMRES it), rel err %.2e" %
              (step, lat_fp64[-1], st64["assembly_ms"], st64["factor_ms"], lat_mixed[-1],
               stm["assembly_ms"], stm["factor_ms"], stm["gmres_iters"], errs[-1]),
              file=sys.stderr, flush=True)

    value = args.steps / t_mixed
    fp64_value = args.steps / t_fp64
    triples = mesh.nbe * mesh.nbe * mesh.npg ** 2
    od = float(np.mean(offdiag_ms))
    achieved = FLOPS_PER_TRIPLE * triples / (od * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        orc = oracle.Oracle("c")
        threads = os.cpu_count() or 1
        t0 = time.perf_counter()
        ref = orc.solve_problem(mesh, dyn)
        tcpu = time.perf_counter() - t0
        err_oracle = float(np.max(np.abs(xm - ref["x"])) / np.max(np.abs(ref["x"])))
        cpu = {"value": 1.0 / tcpu, "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": "one full solve of the same workload (GHMATECE, GHMATECD, LAPACK "
                         "zgesv, INTEREC at 64 points) by the C oracle with OpenMP on %d "
                         "threads: %.1f s" % (threads, tcpu)}
        print("cpu oracle solve %.1f s; mixed vs oracle rel err %.2e" % (tcpu, err_oracle),
              file=sys.stderr, flush=True)
    else:
        err_oracle = None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 assembly + f32/bf16x3 factor + f64 refine",
        "data": "synthetic (generated flat-patch mesh)",
        "config": {"workload": "flat patch 48x45 (2160 BE + 5 EE, 6480 complex DOF), "
                               "omega=2.0, 8x8 Gauss, one frequency per solve",
                   "parallelism": "single GPU"},
        "p50_ms": percentile(lat_mixed, 50),
        "p99_ms": percentile(lat_mixed, 99),
        "fp64": {"value": fp64_value, "unit": "solves/s", "p50_ms": percentile(lat_fp64, 50),
                 "p99_ms": percentile(lat_fp64, 99)},
        "speedup_mixed_vs_fp64": value / fp64_value,
        "max_rel_err_mixed_vs_fp64": max(errs),
        "rel_err_mixed_vs_oracle": err_oracle,
        "gmres_iters": int(np.max(iters)),
        "fallbacks_to_fp64": fallbacks,
        "roofline": {"kernel": "k_dyn_offdiag_h (dynamic H assembly)",
                     "pipe": "FP64 vector ALU", "bound": "mfma", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "avg_ms": od, "flops_per_triple": FLOPS_PER_TRIPLE,
                     "triples": triples},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
