#!/usr/bin/env python
"""IBIM exterior Helmholtz solve throughput on one MI355X.

A step is one full solve of the hot path on a level set already resident in
HBM: WENO gradient + curvatures, tube extraction / projection / weights, the
right-hand side, the solve, and the field at 500 exterior receivers.

Two paths are timed on the same workload:
  mixed  complex BiCG on the FP32 operator -- stored as complex64 in HBM while
         N^2 * 8 B fits, else FP32 entries recomputed per sweep with an FP64
         range-reduced phase -- with FP64 residual replacement by the FP64
         operator (method IBIM_SOLVE_RELIABLE);
  fp64   the plain FP64 GPU path: dense complex128 operator assembled in HBM,
         the reference's complex BiCG.
Both run with one host check per iteration (chunk 1), so every sweep launch in
the timed region does a full pass.
`value` is the mixed path's solves/s; both paths' densities are checked against
an FP64 solve to tol 1e-11 (rel err must be <= 1e-6) and the field against the
analytic point-source field.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload torus128]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "helmholtz-3d-torrone_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

import ibim  # noqa: E402
import shapes  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
C_EVAL = 26  # flops per entry evaluation (fixed convention)

WORKLOADS = {
    # Torus (R=1, r=0.4) on 128^3 over [-2,2]^3, eps = tau = 2h, k = 10, interior
    # point source at (1, 0, 0.05): exterior Dirichlet problem, double layer.
    "torus128": dict(shape="torus", n=128, k=10.0, eps_h=2.0, tau_h=2.0, x0=(1.0, 0.0, 0.05)),
    # Unit sphere on 64^3, eps = 3h, k = 2, source at (0.1, 0.2, 0.3).
    "sphere64": dict(shape="sphere", n=64, k=2.0, eps_h=3.0, tau_h=2.0, x0=(0.1, 0.2, 0.3)),
    # small case for quick checks
    "sphere32": dict(shape="sphere", n=32, k=2.0, eps_h=3.0, tau_h=2.0, x0=(0.1, 0.2, 0.3)),
}
TOL = 1e-8          # relative residual target of both timed paths
REF_TOL = 1e-11     # the FP64 solve the densities are compared against
REPLACE = 1e-2      # residual drop between FP64 residual replacements
NRECV = 500
C64_MAX_BYTES = 120e9  # store the complex64 operator up to this size


def level_of(w):
    return shapes.torus(w["n"]) if w["shape"] == "torus" else shapes.sphere(w["n"])


def make_problem(level, w):
    pb = ibim.Problem()
    pb.set_level(level)
    return pb


def params_of(w):
    h = 4.0 / w["n"]
    return ibim.params(k=w["k"], eps=w["eps_h"] * h, tau=w["tau_h"] * h, dirichlet=1)


def mixed_stores(count):
    return float(count) * count * 8 <= C64_MAX_BYTES


def step(pb, w, recv, mixed):
    pb.geometry(5)
    count = pb.tube(params_of(w), ibim.TUBE_TRADITIONAL)
    if not mixed:
        pb.assemble(ibim.C128)
    elif mixed_stores(count):
        pb.assemble(ibim.C64)
    pb.rhs(ibim.RHS_POINT_SOURCE, x0=w["x0"])
    if mixed:
        info, _ = pb.solve(ibim.SOLVE_RELIABLE, maxit=3000, tol=TOL, inner_tol=REPLACE, chunk=1)
    else:
        info, _ = pb.solve(ibim.SOLVE_BICG, maxit=3000, tol=TOL, chunk=1)
    u = pb.evaluate(recv, ibim.EVAL_DOUBLE)
    x = pb.density()
    return info, x, u


def timed(pb, w, recv, mixed, steps, warmup):
    for _ in range(warmup):
        step(pb, w, recv, mixed)
    pb.sync()
    pb.reset_timers()
    pb.set_profiling(True)
    pb.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        info, x, u = step(pb, w, recv, mixed)
    pb.sync()
    t = time.perf_counter() - t0
    pb.set_profiling(False)
    return t / steps, info, x, u


def cpu_baseline(w, count, iters, budget_s=20.0):
    """The oracle (FP64 C + OpenMP, the reference's loop nests) on a bounded sample:
    assembly of a block of rows and reference-style sweeps over it, extrapolated
    to one full solve with the FP64 path's iteration count."""
    import oracle_ctypes as O

    level = level_of(w)
    n = w["n"]
    h = 4.0 / n
    prm = O.make_params(k=w["k"], eps=w["eps_h"] * h, tau=w["tau_h"] * h, dirichlet=1)
    geo = O.geometry(level, n)
    node, flat = O.tube(level, geo, n, prm, 0)
    nrows = min(count, 1024)
    t0 = time.perf_counter()
    are, aim = O.kernel_rows(flat, count, prm, row0=0, nrows=nrows)
    t_asm = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    p = rng.standard_normal(count) + 1j * rng.standard_normal(count)
    q = rng.standard_normal(nrows) + 1j * rng.standard_normal(nrows)
    sweeps, t_sw = 0, 0.0
    while t_sw < budget_s * 0.5 and sweeps < 50:
        t0 = time.perf_counter()
        O.sweep_rows(are, aim, p, q)
        t_sw += time.perf_counter() - t0
        sweeps += 1
    per_row_asm = t_asm / nrows
    per_row_sweep = t_sw / sweeps / nrows
    t_solve = count * per_row_asm + iters * count * per_row_sweep
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": "oracle FP64 assembly of %d of %d rows (%.2f s) and %d reference-order sweeps over them "
                  "(%.2f s), extrapolated to one solve = full assembly + %d BiCG sweeps" %
                  (nrows, count, t_asm, sweeps, t_sw, iters),
        "seconds_per_solve": t_solve,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="torus128", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    w = WORKLOADS[args.workload]
    level = level_of(w)
    recv = shapes.shell_receivers(NRECV, 1.5, 3.0, seed=0)
    exact = shapes.point_source(recv, w["x0"], w["k"])
    print("[bench] workload %s: uploading level set" % args.workload, file=sys.stderr, flush=True)

    # FP64 reference density (untimed): dense operator, BiCG to REF_TOL
    p64 = make_problem(level, w)
    p64.geometry(5)
    count = p64.tube(params_of(w), ibim.TUBE_TRADITIONAL)
    p64.assemble(ibim.C128)
    p64.rhs(ibim.RHS_POINT_SOURCE, x0=w["x0"])
    info_ref, _ = p64.solve(ibim.SOLVE_BICG, maxit=4000, tol=REF_TOL)
    xr = p64.density()
    x_ref = xr[0] + 1j * xr[1]
    print("[bench] count=%d reference BiCG: %d iterations, status %d" % (count, info_ref.iterations, info_ref.status),
          file=sys.stderr, flush=True)

    t64, info64, x64, u64 = timed(p64, w, recv, False, args.steps, args.warmup)
    ms_mv, n_mv = p64.kernel_time("matvec")
    ms_asm, n_asm = p64.kernel_time("assemble")
    print("[bench] fp64 path: %.3f s/solve, %d iterations" % (t64, info64.iterations), file=sys.stderr, flush=True)
    p64.close()

    pm = make_problem(level, w)
    tm, infom, xm, um = timed(pm, w, recv, True, args.steps, args.warmup)
    stored = mixed_stores(count)
    ms_dom, n_dom = pm.kernel_time("matvec" if stored else "sweep_free32")
    ms_rep, n_rep = pm.kernel_time("sweep_free64")
    print("[bench] mixed path: %.3f s/solve, %d iterations, %d FP64 residual replacements" %
          (tm, infom.iterations, infom.outer), file=sys.stderr, flush=True)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    err_mixed = rel(xm[0] + 1j * xm[1], x_ref)
    err_fp64 = rel(x64[0] + 1j * x64[1], x_ref)
    field_err = rel(um[0] + 1j * um[1], exact)

    n2 = float(count) * count
    vec_bytes = 4 * 16 * count  # p, p~ in; A p, A^H p~ out (complex128)
    avg_dom = ms_dom / max(n_dom, 1) * 1e-3
    if stored:
        # dominant kernel: the fused sweep over the complex64 operator (HBM-bound)
        dom_bytes = n2 * 8 + vec_bytes
        achieved = dom_bytes / avg_dom / 1e9 if n_dom else 0.0
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_sweep<float2>: 8 B/entry x %d^2 entries + 64 B/point vectors per launch" % count}
    else:
        flops = n2 * (C_EVAL + 8 * 2)
        achieved = flops / avg_dom / 1e12 if n_dom else 0.0
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                "kernel": "k_sweep_free32: %d flops/entry x %d^2 entries per launch" % (C_EVAL + 16, count)}
    # the FP64 path's dense fused sweep (HBM-bound): one pass over the c128 operator
    avg_mv = ms_mv / max(n_mv, 1) * 1e-3
    mv_gbs = (n2 * 16 + vec_bytes) / avg_mv / 1e9 if n_mv else 0.0

    line = {
        "metric": "solves/s + matvec GB/s (%peak), mixed-prec vs FP64 GPU at rel err <=1e-6",
        "value": 1.0 / tm,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tm * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 operator (fp64 phase, fp64 sums) + fp64 residual replacement",
        "data": "synthetic: %s signed-distance level set, point-source Dirichlet data" % w["shape"],
        "config": {"workload": "%s: %s N=%d, k=%g, eps=%gh, tau=%gh, point source %s, %d receivers, tol %g" % (
            args.workload, w["shape"], w["n"], w["k"], w["eps_h"], w["tau_h"], list(w["x0"]), NRECV, TOL),
            "tube_points": count, "parallelism": "single GPU",
            "mixed_operator": "stored complex64" if stored else "recomputed FP32 entries"},
        "mixed": {"solves_per_s": 1.0 / tm, "iterations": infom.iterations, "replacements": infom.outer,
                  "rel_err_vs_fp64": err_mixed, "field_rel_err_vs_analytic": field_err,
                  "sweep_ms": avg_dom * 1e3, "replacement_sweep_ms": ms_rep / max(n_rep, 1)},
        "fp64": {"solves_per_s": 1.0 / t64, "iterations": info64.iterations, "rel_err_vs_fp64": err_fp64,
                 "assemble_ms": ms_asm / max(n_asm, 1), "matvec_ms": avg_mv * 1e3,
                 "matvec_GBs": mv_gbs, "matvec_frac_of_peak": mv_gbs / HBM_PEAK_GBS},
        "speedup_mixed_vs_fp64": t64 / tm,
        "roofline": roof,
    }
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, bounded sample)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(w, count, info64.iterations)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
