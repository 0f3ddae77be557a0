#!/usr/bin/env python
"""IBIM exterior Helmholtz solve throughput on one MI355X.

A step is one full solve of the hot path on a level set already resident in
HBM: WENO gradient + curvatures, tube extraction / projection / weights, the
right-hand side, the solve, and the field at 500 exterior receivers.

Two paths are timed on the same workload:
  mixed  complex BiCG on the FP32 operator -- its radial factors stored once per
         unordered pair as complex64, with a complex128 copy for the FP64
         residual replacements (IBIM_MIXED_PAIR), while that fits in HBM, else
         the factors recomputed per pair and sweep with FP32 sine and cosine
         of an FP64 phase -- with FP64 residual replacement by the FP64
         operator (method IBIM_SOLVE_RELIABLE);
  fp64   the FP64 GPU path: the operator's radial factors stored once per pair
         as complex128 (IBIM_C128_PAIR), the reference's complex BiCG.
--matrix-free stores nothing on either path: both recompute the factors per
pair and sweep (FP64 path: bitwise the complex128-pair operator).
Both run with one host check per iteration (chunk 1), so every sweep launch in
the timed region does a full pass.
`value` is the mixed path's solves/s; both paths' densities are checked against
an FP64 solve to tol 1e-11 (rel err must be <= 1e-6) and the field against the
analytic point-source field.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload torus128] [--matrix-free]
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
FP64_PEAK_TFLOPS = 78.6
FP32_MATRIX_PEAK_TFLOPS = 157.3  # dense FP32 matrix cores
C_EVAL = 26  # flops per entry evaluation (fixed convention)

WORKLOADS = {
    # Torus (R=1, r=0.4) on 128^3 over [-2,2]^3, eps = tau = 2h, k = 10, interior
    # point source at (1, 0, 0.05): exterior Dirichlet problem, double layer.
    # mixed path: enlarged GMRES with the 16-column passes on the FP32 matrix cores
    # (23 passes; GMRES column by column: 107 single-vector A p sweeps, 0.356 s against 0.253 s)
    "torus128": dict(shape="torus", n=128, k=10.0, eps_h=2.0, tau_h=2.0, x0=(1.0, 0.0, 0.05), tol=1e-8,
                     mixed_solver="enlarged32"),
    # Unit sphere on 64^3, eps = 3h, k = 2, source at (0.1, 0.2, 0.3), BiCG tol 1e-10;
    # 9 iterations: replacements every 1e-5 of residual drop (2, and the FP64 path's
    # 9 iterations: 8.9 ms per solve; at 1e-3 3 replacements and 11 iterations, 10.6 ms)
    # The mixed path keeps reliable BiCG here: 9.07 ms against 9.46 ms for GMRES
    # refinement, whose host check per iteration weighs on a 10-iteration solve.
    "sphere64": dict(shape="sphere", n=64, k=2.0, eps_h=3.0, tau_h=2.0, x0=(0.1, 0.2, 0.3), tol=1e-10,
                     replace=1e-5, mixed_solver="reliable"),
    # small case for quick checks
    "sphere32": dict(shape="sphere", n=32, k=2.0, eps_h=3.0, tau_h=2.0, x0=(0.1, 0.2, 0.3), tol=1e-8),
}
# workloads with their own drivers (below): many right-hand sides and surfaces,
# the misfit gradient, and the 256^3 matrix-free case
EXTRA = ("batch", "adjoint96", "star256")
TOL = 1e-8          # default relative residual target of both timed paths
REF_TOL = 1e-11     # the FP64 solve the densities are compared against
REPLACE = 1e-2      # residual drop between FP64 residual replacements (torus128: 127 iterations, 4 replacements)
# the same for the batched solvers (at 1e-1 the torus128 columns stagnated): batch workload
# 15.71 s at 1e-4, 16.68 s at 1e-3, 17.02 s at 1e-2, 15.85 s at 1e-5
BATCH_REPLACE = 1e-4
# star256 mixed path (FP32 products): 4 replacement passes at 1e-4 (9.13 s) against 5 at
# 1e-3 (9.70 s); at 1e-5 the FP32 operator's recursive residual drifted from the true one
# and the columns stalled near 1e-5
STAR_REPLACE = 1e-4
NRECV = 500
STORE_MAX_BYTES = 200e9  # store an operator in HBM up to this size (288 GB per GPU)
MATRIX_FREE = False  # --matrix-free: both timed paths recompute the operator per sweep
REPLACE_OVERRIDE = None  # --replace: the mixed path's replacement drop for every single-vector workload
SOLVER_OVERRIDE = None  # --mixed-solver: the mixed path's solver for every single-vector workload
# the 256^3 star's FP64 path keeps no complex128 pairs by default: caching the
# first row blocks' pairs (249 of 491 GB) made its pass slower, 492.6 against
# 478.8 ms (reading complex128 pairs at both visits costs what computing them does)
STAR_FP64_CACHE_BYTES = 0.0
STAR_STORE_MAX_BYTES = 260e9  # the 256^3 star's complex64 pairs (~251 GB) alone beside its vectors
# star256 BiCG tolerance: a decade below the 1e-6 asked of the densities (at
# tol 1e-6 the reference's stagnation rule, |change| < 1e-3 tol, stopped BiCG
# at 1.02e-6)
STAR_TOL = 1e-7
PAIR_ROWS, PAIR_COLS = 512, 1024  # tile shape of the pair storage


def pair_bytes(count, esz):
    """HBM bytes of the pair storage: row blocks of 512 rows, tiles of 512 x 1024
    values from the diagonal block to the last column."""
    tiles = sum((count - b * PAIR_ROWS + PAIR_COLS - 1) // PAIR_COLS for b in range((count + PAIR_ROWS - 1) // PAIR_ROWS))
    return tiles * PAIR_ROWS * PAIR_COLS * esz


def cache_bytes(count, budget):
    """Bytes of complex128 pair tiles the FP64 matrix-free cache keeps for a
    budget: whole row blocks of 512 rows, 512 x 1024-value tiles, from the first."""
    tile, total = 512 * 1024 * 16, 0
    for blk in range((count + 511) // 512):
        t = (count - 512 * blk + 1023) // 1024
        if (total + t) * tile > budget:
            break
        total += t
    return total * tile


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
    return not MATRIX_FREE and pair_bytes(count, 8) <= STORE_MAX_BYTES


def mixed_storage(count):
    """complex64 pairs + their complex128 copy when both fit, else complex64 pairs
    (FP64 replacements recomputed), else nothing stored."""
    if MATRIX_FREE:
        return None
    if pair_bytes(count, 8) + pair_bytes(count, 16) <= STORE_MAX_BYTES:
        return ibim.MIXED_PAIR
    return ibim.C64_PAIR if mixed_stores(count) else None


def mixed_solver_of(w):
    s = SOLVER_OVERRIDE or w.get("mixed_solver", "gmres")
    # FP32 matrix-core products exist for stored complex64 pairs only
    return "enlarged" if s == "enlarged32" and MATRIX_FREE else s


def step(pb, w, recv, path):
    """One solve of a single-right-hand-side workload. path: "fp64" (the
    reference's BiCG on complex128 pairs), "fp64_gmres" (FP64 GMRES on the same
    pairs), "mixed" (FP64 refinement around GMRES on complex64 pairs; with
    nothing stored, on the mixed matrix-free operator with FP64 matrix-free
    residuals) or "mixed_reliable" (BiCG on the same operator with FP64
    residual replacement)."""
    mixed = path.startswith("mixed")
    pb.geometry(5)
    count = pb.tube(params_of(w), ibim.TUBE_TRADITIONAL)
    if not mixed and not MATRIX_FREE:
        pb.assemble(ibim.C128_PAIR if pair_bytes(count, 16) <= STORE_MAX_BYTES else ibim.C128)
    elif mixed and mixed_storage(count) is not None:
        pb.assemble(mixed_storage(count))
    pb.rhs(ibim.RHS_POINT_SOURCE, x0=w["x0"])
    if path == "fp64":
        info, _ = pb.solve(ibim.SOLVE_BICG, maxit=3000, tol=w["tol"], chunk=1)
    elif path in ("fp64_gmres", "mixed_gmres") or (path == "mixed" and mixed_solver_of(w) == "gmres"):
        info, _ = pb.solve(ibim.SOLVE_GMRES, maxit=3000, tol=w["tol"])
    elif path == "mixed" and mixed_solver_of(w) in ("enlarged", "enlarged32"):
        pb.set_fp32_products(mixed_solver_of(w) == "enlarged32")
        info, _ = pb.solve(ibim.SOLVE_BLOCK_GMRES, maxit=3000, tol=w["tol"])
    else:
        info, _ = pb.solve(ibim.SOLVE_RELIABLE, maxit=3000, tol=w["tol"],
                           inner_tol=REPLACE_OVERRIDE or w.get("replace", REPLACE), chunk=1)
    u = pb.evaluate(recv, ibim.EVAL_DOUBLE)
    x = pb.density()
    return info, x, u


def timed(pb, w, recv, path, steps, warmup):
    for _ in range(warmup):
        step(pb, w, recv, path)
    pb.sync()
    pb.reset_timers()
    pb.set_profiling(True)
    pb.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        info, x, u = step(pb, w, recv, path)
    pb.sync()
    t = time.perf_counter() - t0
    pb.set_profiling(False)
    return t / steps, info, x, u


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's processor string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


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
        "value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
        "sample": "oracle FP64 assembly of %d of %d rows (%.2f s) and %d reference-order sweeps over them "
                  "(%.2f s), extrapolated to one solve = full assembly + %d BiCG sweeps" %
                  (nrows, count, t_asm, sweeps, t_sw, iters),
        "seconds_per_solve": t_solve,
    }


def fibonacci_directions(m):
    """Unit vectors on a Fibonacci sphere: z = 1 - (2i+1)/m, phi = i pi (3 - sqrt 5)."""
    i = np.arange(m)
    z = 1 - (2 * i + 1) / m
    ph = i * np.pi * (3 - np.sqrt(5))
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(ph), s * np.sin(ph), z], -1)


def emit(line):
    print(json.dumps(line), flush=True)


def progress(pb):
    pb.set_progress(lambda what, it, res: print("[bench]   %s iteration %d, largest residual %.2e" % (what, it, res),
                                                file=sys.stderr, flush=True))


def run_batch(args):
    """Many right-hand sides and surfaces: the torus (128^3, k = 10) with 64 plane
    waves on a Fibonacci sphere, plus 8 spheres / ellipsoids with seed-random
    radii in [0.6, 1.2] on 48^3-96^3 grids with 16 plane waves each; sound-soft
    data, tol 1e-7. All 9 surfaces (12 batches of 16 right-hand sides) are
    solved together by ibim_solve_group: every BiCG iteration applies all of
    their operators in one persistent launch (a queue of 64-row items, the
    torus first). FP64 path: batched BiCG on complex128 pair storage, products
    on the FP64 matrix cores. Mixed path (IBIM_MIXED_PAIR): complex64 pairs,
    FP64 iterative refinement around grouped block GMRES, one pass Y = A P per
    iteration, its FP64 residuals over the complex128 copy of the same pairs
    (--batch-solver gmres: GMRES per column; reliable: grouped BiCG with FP64
    residual replacement). The passes run on the FP32 matrix cores for block
    GMRES and on the FP64 ones for the per-column solvers (--batch-products
    overrides)."""
    tol, k = 1e-7, 10.0
    rng = np.random.default_rng(0)
    torus = ("torus128", ibim.shape(ibim.SHAPE_TORUS, a=1.0, b=0.4), 128, 64)
    small = []
    grids = [48, 56, 64, 72, 80, 88, 96, 96]
    for i in range(8):
        r = rng.uniform(0.6, 1.2, 3)
        if i % 2 == 0:
            shp = ibim.shape(ibim.SHAPE_SPHERE, a=float(r[0]))
            name = "sphere R=%.2f" % r[0]
        else:
            shp = ibim.shape(ibim.SHAPE_ELLIPSOID, a=float(r[0]), b=float(r[1]), c=float(r[2]))
            name = "ellipsoid %.2f/%.2f/%.2f" % tuple(r)
        small.append((name, shp, grids[i], 16))
    surfaces = [torus] + small
    dirs = fibonacci_directions(64)

    def run(path):
        pbs = []
        for name, shp, n, m in surfaces:
            pb = ibim.Problem()
            pb.set_level_shape(shp, n)
            pb.sync()
            pbs.append(pb)
        progress(pbs[0])
        t0 = time.perf_counter()
        counts = []
        for (name, shp, n, m), pb in zip(surfaces, pbs):
            h = 4.0 / n
            pb.geometry(5)
            counts.append(pb.tube(ibim.params(k, 2 * h, 2 * h, dirichlet=1)))
            if path.startswith("fp64"):
                pb.assemble(ibim.C128_PAIR)
            else:
                pb.assemble(ibim.MIXED_PAIR)
                pb.set_fp32_products(batch_fp32(args))
            pb.set_nrhs(m)
            for c in range(m):
                pb.rhs(ibim.RHS_PLANE_DIRICHLET, amp=1.0, direction=dirs[c], col=c)
        pbs[0].reset_timers()
        if args.batch_graphs:
            pbs[0].set_graphs(True)  # replay captured chunks (graphs are off while profiling)
        else:
            pbs[0].set_profiling(True)
        t1 = time.perf_counter()
        if path == "fp64":
            infos = ibim.solve_group(pbs, ibim.SOLVE_BICG, maxit=3000, tol=tol)
        elif path == "fp64_block":
            infos = ibim.solve_group(pbs, ibim.SOLVE_BLOCK_GMRES, maxit=3000, tol=tol)
        elif args.batch_solver in ("gmres", "block"):
            infos = ibim.solve_group(pbs, ibim.SOLVE_GMRES if args.batch_solver == "gmres" else
                                     ibim.SOLVE_BLOCK_GMRES, maxit=3000, tol=tol)
        else:
            infos = ibim.solve_group(pbs, ibim.SOLVE_RELIABLE, maxit=3000, tol=tol, inner_tol=args.batch_replace)
        dens = []
        for (name, shp, n, m), pb in zip(surfaces, pbs):
            x = []
            for c in range(m):
                d = pb.density(c)
                x.append(d[0] + 1j * d[1])
            dens.append(np.stack(x))
        for pb in pbs:
            pb.sync()
        el = time.perf_counter() - t0
        solve_s = time.perf_counter() - t1
        pbs[0].set_profiling(False)
        ms, launches = pbs[0].kernel_time("batch_group")
        for pb in pbs:
            pb.close()
        print("[bench] batch %s: %.2f s (solve %.2f s), %d grouped passes of %.1f ms" %
              (path, el, solve_s, launches, ms / max(launches, 1)), file=sys.stderr, flush=True)
        return el, infos, dens, counts, dict(solve_s=solve_s, launches=launches, pass_ms=ms / max(launches, 1))

    # beside the two paths (not with --no-extra): block GMRES over the complex128
    # pairs, which separates the solver's gain from the precision's
    res = {path: run(path) for path in ("fp64", "mixed") + (() if args.no_extra else ("fp64_block",))}
    rows, n_rhs, worst_err = [], 0, 0.0
    for s_i, (name, _, n, m) in enumerate(surfaces):
        f, mx = res["fp64"], res["mixed"]
        err = float(max(np.linalg.norm(mx[2][s_i][c] - f[2][s_i][c]) / np.linalg.norm(f[2][s_i][c])
                        for c in range(m)))
        worst_err = max(worst_err, err)
        n_rhs += m
        rows.append({"surface": name, "n": n, "tube_points": f[3][s_i], "rhs": m,
                     "fp64_iterations": f[1][s_i].iterations, "fp64_status": f[1][s_i].status,
                     "mixed_iterations": mx[1][s_i].iterations, "mixed_status": mx[1][s_i].status,
                     "mixed_fp64_residual_steps": mx[1][s_i].outer, "mixed_rel_err_vs_fp64": err})
        print("[bench] %s: N_t=%d, %d rhs, iterations fp64 %d / mixed %d, rel err %.1e" %
              (name, f[3][s_i], m, f[1][s_i].iterations, mx[1][s_i].iterations, err), file=sys.stderr, flush=True)
    total = {p: res[p][0] for p in res}
    extra = {}
    if "fp64_block" in res:
        b, f = res["fp64_block"], res["fp64"]
        extra["fp64_block_gmres"] = {
            "solves_per_s": n_rhs / total["fp64_block"], "speedup_mixed_vs_it": total["fp64_block"] / total["mixed"],
            "grouped_passes": b[4]["launches"],
            "worst_rel_err_vs_fp64": float(max(np.linalg.norm(b[2][s_i][c] - f[2][s_i][c]) /
                                               np.linalg.norm(f[2][s_i][c])
                                               for s_i, (_, _, _, m) in enumerate(surfaces) for c in range(m)))}
    return ({"metric": "solves/s (batched right-hand sides, all surfaces)", "value": n_rhs / total["mixed"],
          "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": total["mixed"] * 1e3,
          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
          "dtype": "complex64 pair storage, products on the %s matrix cores, %s over the complex128 copy of the "
                   "pairs" % ("fp32" if batch_fp32(args) else "fp64",
                              {"gmres": "GMRES with fp64 iterative refinement",
                               "block": "block GMRES with fp64 iterative refinement"}.get(
                                  args.batch_solver, "BiCG with fp64 residual replacement")),
          "data": "synthetic: torus + 8 seed-random spheres/ellipsoids, Fibonacci plane waves",
          "config": {"workload": "batch: torus 128^3 x 64 plane waves + 8 surfaces on 48^3-96^3 x 16 plane waves, "
                                 "k=10, eps=tau=2h, tol %g" % tol, "rhs": n_rhs, "parallelism": "single GPU",
                     "schedule": "all 9 surfaces (12 batches of 16 right-hand sides) in one grouped solve: one "
                                 "persistent launch of every operator per iteration (ibim_solve_group)",
                     "mixed_solver": args.batch_solver},
          "fp64_solves_per_s": n_rhs / total["fp64"], "speedup_mixed_vs_fp64": total["fp64"] / total["mixed"],
          "worst_rel_err_mixed_vs_fp64": worst_err,
          "grouped_pass": {p: res[p][4] for p in res}, "surfaces": rows, **extra})


def run_adjoint(args):
    """Misfit gradient: perturbed sphere r = 1 + 0.15 sin3t cos2p on 96^3, k = 8,
    point source, 200 seed-random receivers on |x| = 2.5, target field from
    r = 1 + 0.1 sin3t cos2p. Times the forward solve and the gradient with
    respect to every level-set value and k (adjoint solve + adjoint kernels) with
    enlarged GMRES for the forward and adjoint solves, the reference's BiCG timed
    beside it, and
    checks 50 seed-random entries against central differences (step 1e-8, the
    forward path re-solved by FP64 refinement to the FP64 floor; the same
    entries at step 1e-6 are reported beside)."""
    n, k = 96, 8.0
    h = 4.0 / n
    ctr = (0.0137, -0.0219, 0.0071)  # off the grid's symmetry planes
    v = np.random.default_rng(0).standard_normal((200, 3))
    recv = 2.5 * v / np.linalg.norm(v, axis=1, keepdims=True)
    x0 = (0.1, 0.2, 0.3)
    pb = ibim.Problem()
    progress(pb)

    def prepare(level, kk, storage=ibim.C128_PAIR):
        pb.set_level(level)
        pb.geometry(5)
        pb.tube(ibim.params(kk, 2 * h, 2 * h, dirichlet=1))
        pb.assemble(storage)
        pb.rhs(ibim.RHS_POINT_SOURCE, x0=x0)

    def field(level, kk, tol_refined=True):
        # the finite-difference fields: FP64 refinement to the FP64 floor around
        # BiCG on the complex64 copy of the pairs, residuals on their complex128
        # copy (the FP64 path's operator, so the same J)
        prepare(level, kk, ibim.MIXED_PAIR)
        if tol_refined:
            pb.solve(ibim.SOLVE_REFINED, tol=1e-15, inner_tol=1e-5, max_outer=10)
        re, im = pb.evaluate(recv, ibim.EVAL_DOUBLE)
        return re + 1j * im

    pb.set_level_shape(ibim.shape(ibim.SHAPE_STAR, a=1.0, b=0.10, m=3, n=2, variant=0, center=ctr), n)
    target = field(pb.level(), k)
    pb.set_level_shape(ibim.shape(ibim.SHAPE_STAR, a=1.0, b=0.15, m=3, n=2, variant=0, center=ctr), n)
    level = pb.level()
    def timed_gradient(method):
        times = []
        for _ in range(2):
            pb.sync()
            t0 = time.perf_counter()
            prepare(level, k)
            info, _ = pb.solve(method, maxit=3000, tol=1e-12)
            t1 = time.perf_counter()
            pb.set_receivers(recv, target)
            J, g, dk = pb.gradient(mode=ibim.EVAL_DOUBLE, tol=1e-12, maxit=3000, method=method)
            pb.sync()
            times.append((t1 - t0, time.perf_counter() - t1))
        return times[-1], info, J, g, dk

    # the reference's BiCG for the forward and adjoint solves, then (timed as the
    # record's value) enlarged GMRES on the 16-column passes for both
    (t_fwd_bicg, t_adj_bicg), info_bicg, _, g_bicg, _ = timed_gradient(ibim.SOLVE_BICG)
    (t_fwd, t_adj), info, J, g, dk = timed_gradient(ibim.SOLVE_BLOCK_GMRES)
    count = pb.count

    def J_of(lv, kk):
        return float(np.sum(np.abs(field(lv, kk) - target) ** 2))

    gmax = np.abs(g).max()
    live = np.argwhere(np.abs(g) > 1e-2 * gmax)
    picks = live[np.random.default_rng(1).choice(len(live), 50, replace=False)]
    errs, resolved = [], 0
    for idx in picks:
        idx = tuple(idx)
        f = []
        for st in (1e-8, 2e-8):
            lp, lm = level.copy(), level.copy()
            lp[idx] += st
            lm[idx] -= st
            f.append((J_of(lp, k) - J_of(lm, k)) / (2 * st))
        if abs(f[0] - f[1]) <= 5e-6 * abs(g[idx]) + 3e-8 * gmax:
            resolved += 1
            errs.append(abs(f[0] - g[idx]) / abs(g[idx]))
    # the same 50 entries at step 1e-6: a +-1e-6 change of one level value moves
    # the WENO weights of its neighbours (smoothness indicators near 1e-12 along a
    # distance function's normals) or crosses a limiter branch, so J is far from
    # linear over the step; the count within 1e-5 and the median are reported
    errs6 = []
    for idx in picks:
        idx = tuple(idx)
        lp, lm = level.copy(), level.copy()
        lp[idx] += 1e-6
        lm[idx] -= 1e-6
        errs6.append(abs((J_of(lp, k) - J_of(lm, k)) / 2e-6 - g[idx]) / abs(g[idx]))
    errs6 = np.array(errs6)
    fk = (J_of(level, k + 1e-6) - J_of(level, k - 1e-6)) / 2e-6
    print("[bench] adjoint96: N_t=%d forward %.3f s gradient %.3f s; %d/50 entries resolved, worst rel err %.1e; "
          "step 1e-6: %d/50 within 1e-5" % (count, t_fwd, t_adj, resolved, max(errs) if errs else float("nan"),
                                           int(np.sum(errs6 <= 1e-5))), file=sys.stderr, flush=True)
    return ({"metric": "misfit gradients/s (forward solve + adjoint)", "value": 1.0 / (t_fwd + t_adj),
          "unit": "gradients/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": (t_fwd + t_adj) * 1e3,
          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
          "data": "synthetic: perturbed-sphere signed distance, point source, 200 receivers",
          "config": {"workload": "adjoint96: r=1+0.15 sin3t cos2p on 96^3, k=8, eps=tau=2h, target from b=0.1, "
                                 "200 receivers on |x|=2.5", "tube_points": count, "parallelism": "single GPU"},
          "solver": "enlarged GMRES (forward and adjoint, 16-column passes on the complex128 pairs)",
          "forward_s": t_fwd, "gradient_s": t_adj, "forward_iterations": info.iterations, "J": J,
          "reference_bicg": {"forward_s": t_fwd_bicg, "gradient_s": t_adj_bicg,
                             "gradients_per_s": 1.0 / (t_fwd_bicg + t_adj_bicg),
                             "forward_iterations": info_bicg.iterations,
                             "gradient_rel_diff": float(np.linalg.norm(g - g_bicg) / np.linalg.norm(g_bicg))},
          "fd_entries_resolved": resolved, "fd_worst_rel_err": max(errs) if errs else None,
          "fd_step_1e-6": {"within_1e-5": int(np.sum(errs6 <= 1e-5)), "median_rel_err": float(np.median(errs6)),
                           "worst_rel_err": float(errs6.max())},
          "dk": dk, "dk_fd": fk, "dk_rel_err": abs(fk - dk) / abs(dk)})


def run_star(args):
    """The 256^3 case: star r = 1 + 0.2 cos4t sin3p, k = 40 (--star-k), eps = tau
    = 2h, 16 plane waves (Fibonacci), sound-soft data, BiCG to tol 1e-7 (the
    relative residual; the mixed path's is that of the FP64 operator).
    Formulation: the combined-field operator 1/2 I + K - i eta S with eta = k
    (--star-formulation double: the reference's double layer 1/2 I + K alone,
    singular near the interior Neumann eigenvalues, which lie ~0.003 apart in k
    here; BiCG does not converge on it at k = 40).
    FP64 path: batched BiCG on the FP64 matrix-free operator, the fused pair
    pass computing each factor where complex128 pairs (which would need
    ~490 GB) would be read. Mixed path: complex64 pair storage (~246 GB,
    assembled inside the timed region into memory allocated before it), its
    16-column passes on the FP32 matrix cores (--star-fp64-products: on the
    FP64 ones); FP64 iterative refinement around batched GMRES, one pass
    Y = A P per iteration, the FP64 residuals by the FP64 matrix-free operator
    (--star-solver reliable: batched BiCG, A P and A^H Q per pass, with FP64
    residual replacement). Reports solves/s for both paths when every
    column converged (else batched iterations/s and the residuals reached) and
    the mixed densities' largest relative difference from the FP64 ones."""
    n, k, m, its = 256, args.star_k, 16, args.star_iterations
    combined = args.star_formulation == "combined"
    h = 4.0 / n
    dirs = fibonacci_directions(m)
    out, dens = {}, {}
    for path in ("mixed",) if args.star_mixed_only else ("mixed", "fp64"):
        pb = ibim.Problem()
        progress(pb)
        pb.set_level_shape(ibim.shape(ibim.SHAPE_STAR, a=1.0, b=0.2, m=4, n=3, variant=1), n)
        pb.sync()
        t0 = time.perf_counter()
        pb.geometry(5)
        count = pb.tube(ibim.params(k, 2 * h, 2 * h, dirichlet=1, combined=1 if combined else 0, eta=k))
        pb.set_nrhs(m)
        for c in range(m):
            pb.rhs(ibim.RHS_PLANE_DIRICHLET, amp=1.0, direction=dirs[c], col=c)
        pb.sync()
        t_setup = time.perf_counter() - t0
        store = "matrix-free"
        if path == "mixed" and pair_bytes(count, 8) <= STAR_STORE_MAX_BYTES:
            # the 246 GB allocation is made once, untimed (a fresh allocation of
            # that size right after another process freed one took seconds);
            # the timed region assembles again into it
            try:
                pb.assemble(ibim.C64_PAIR)
                store = "complex64 pairs (%.0f GB)" % (pair_bytes(count, 8) / 1e9)
            except ibim.IbimError as e:
                print("[bench] star256: complex64 pairs not stored (%s); matrix-free" % e, file=sys.stderr,
                      flush=True)
        if path == "fp64" and args.star_fp64_cache > 0:
            # the cache of complex128 pairs: allocated here, untimed (as the
            # mixed path's complex64 pairs), filled again inside the timed solve
            pb.set_free_cache(args.star_fp64_cache)
            pb.apply_batch(np.zeros((count, 1)), ibim.OP_FREE64)
            pb.set_free_cache(args.star_fp64_cache)
        pb.sync()
        pb.reset_timers()
        pb.set_profiling(True)
        t0 = time.perf_counter()
        if path == "mixed":
            if store != "matrix-free":
                pb.assemble(ibim.C64_PAIR)
            out["mixed_operator"] = store
            if not args.star_fp64_products:
                pb.set_fp32_products(True)
                out["mixed_operator"] += ", FP32 matrix-core products"
            if args.star_solver in ("gmres", "block"):
                info, _ = pb.solve(ibim.SOLVE_GMRES if args.star_solver == "gmres" else ibim.SOLVE_BLOCK_GMRES,
                                   maxit=its, tol=STAR_TOL, batch=0)
            else:
                info, _ = pb.solve(ibim.SOLVE_RELIABLE, maxit=its, tol=STAR_TOL, inner_tol=args.star_replace,
                                   batch=0)
            out["mixed_solver"] = args.star_solver
        else:
            info, _ = pb.solve(ibim.SOLVE_BICG, maxit=its, tol=STAR_TOL, batch=0)
        pb.sync()
        el = time.perf_counter() - t0
        pb.set_profiling(False)
        kern = {"mixed": ("batch_pair" if args.star_fp64_products else "batch_pair32")
                         if not out["mixed_operator"].startswith("matrix-free") else "batch_free32",
                "fp64": "batch_free64"}[path]
        ms, nl = pb.kernel_time(kern)
        ms_asm, _ = pb.kernel_time("assemble")
        ms_rep, n_rep = pb.kernel_time("batch_free64") if path == "mixed" else (0.0, 0)
        d = [pb.density(c) for c in range(m)]
        dens[path] = [a + 1j * b for a, b in d]
        out[path] = dict(seconds=el, info=info, count=count, setup=t_setup, pass_ms=ms / max(nl, 1), passes=nl,
                         assemble_ms=ms_asm, replacement_passes=n_rep,
                         cache_bytes=cache_bytes(count, args.star_fp64_cache) if path == "fp64" else 0)
        print("[bench] star256 %s: N_t=%d setup %.2f s, %d iterations (budget %d) %.1f s, residual %.2e, status %d, "
              "%d passes of %.1f ms" % (path, count, t_setup, info.iterations, its, el, info.rel_residual,
                                         info.status, nl, ms / max(nl, 1)), file=sys.stderr, flush=True)
        pb.close()
    mixed_operator = out.pop("mixed_operator")
    mixed_solver = out.pop("mixed_solver")
    if args.star_mixed_only:
        r = out["mixed"]
        return {"workload": "star256 mixed path only (k=%g)" % k, "mixed_operator": mixed_operator,
                "mixed_solver": mixed_solver,
                "seconds": r["seconds"], "iterations_summed_over_columns": r["info"].iterations,
                "residual": r["info"].rel_residual, "status": r["info"].status, "pass_ms": r["pass_ms"],
                "passes": r["passes"], "fp64_residual_passes": r["replacement_passes"]}
    converged = {p: out[p]["info"].status == 0 for p in out}
    rel_err = float(max(np.linalg.norm(dens["mixed"][c] - dens["fp64"][c]) / np.linalg.norm(dens["fp64"][c])
                        for c in range(m)))
    if all(converged.values()):
        metric, unit, value = "solves/s (16 right-hand sides)", "solves/s", m / out["mixed"]["seconds"]
        ms = out["mixed"]["seconds"] * 1e3
        fp64_rate = m / out["fp64"]["seconds"]
    else:
        # iteration budget exhausted: batched iterations run per second
        rate = {p: out[p]["info"].iterations / m / out[p]["seconds"] for p in out}
        metric, unit, value = "batched BiCG iterations/s (16 right-hand sides)", "iterations/s", rate["mixed"]
        ms = 1e3 / rate["mixed"]
        fp64_rate = rate["fp64"]
    npairs = float(out["mixed"]["count"]) * (out["mixed"]["count"] - 1) / 2
    # complex products per pass: A P and A^H Q (BiCG; GMRES: A P alone), 16
    # columns, every pair visited from both of its rows (2 N^2 entries), 8 real
    # flops per complex multiply-add
    pass_flops = {p: (1 if p == "mixed" and mixed_solver in ("gmres", "block") else 2) * 2.0 * npairs * 16 * 8
                  for p in out}
    return ({"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": ms,
          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
          "dtype": "mixed: complex64 pair storage (else fp32 sine / cosine of an fp64 phase), entries and products "
                   "on the fp32 matrix cores summed in fp64 per 64 columns, %s; fp64: fp64 "
                   "matrix-free factors, fp64 matrix-core products, the reference's BiCG" % (
                       {"gmres": "GMRES with fp64 iterative refinement",
                        "block": "block GMRES with fp64 iterative refinement"}.get(
                           mixed_solver, "BiCG with fp64 residual replacement")),
          "data": "synthetic: star signed distance 256^3, 16 Fibonacci plane waves",
          "config": {"workload": "star256: r=1+0.2 cos4t sin3p on 256^3, k=%g, eps=tau=2h, 16 plane waves, tol 1e-7, "
                                 "%s, %d-iteration budget" % (
                                     k, "combined field 1/2 I + K - i eta S, eta = k" if combined else
                                     "double layer 1/2 I + K", its),
                     "tube_points": out["mixed"]["count"], "parallelism": "single GPU",
                     "mixed_operator": mixed_operator, "mixed_solver": mixed_solver,
                     "fp64_operator": "matrix-free (fused pair pass, factors computed)%s" % (
                         ", complex128 pairs of the first row blocks cached (%.0f GB of %.0f GB)" % (
                             out["fp64"]["cache_bytes"] / 1e9, pair_bytes(out["fp64"]["count"], 16) / 1e9)
                         if out.get("fp64", {}).get("cache_bytes") else "")},
          "fp64_rate": fp64_rate, "speedup_mixed_vs_fp64": value / fp64_rate,
          "mixed_rel_err_vs_fp64_density": rel_err,
          "seconds": {p: out[p]["seconds"] for p in out},
          "iterations_summed_over_columns": {p: out[p]["info"].iterations for p in out},
          "residual": {p: out[p]["info"].rel_residual for p in out}, "converged": converged,
          "pass_ms": {p: out[p]["pass_ms"] for p in out},
          "pass_tflops": {p: pass_flops[p] / (out[p]["pass_ms"] * 1e-3) / 1e12 for p in out
                          if out[p]["pass_ms"] > 0},
          "mixed_assemble_ms": out["mixed"]["assemble_ms"],
          "mixed_fp64_residual_passes": out["mixed"]["replacement_passes"],
          "setup_s": out["mixed"]["setup"]})


def batch_fp32(args):
    """Whether the batch workload's mixed passes run on the FP32 matrix cores:
    block GMRES gains from them (one Krylov space absorbs the products' rounding
    in a few more passes), GMRES and BiCG per column lose more iterations near
    resonance than the cheaper passes save."""
    if args.batch_products != "auto":
        return args.batch_products == "fp32"
    return args.batch_solver == "block"


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), or None."""
    import re
    path = os.path.join(REPO, "profiles", "pmc_sweeps_%s.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*traffic (\S+) (\S+) = ([0-9.e+-]+)", line)
        if m and m.group(1) == kernel and m.group(2) == workload:
            return float(m.group(3))
    return None


def calibrate_copy(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size (the runtime's blit kernel),
    the reference point for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 0, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(reps):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    emit({"calibration": "device-to-device copy", "bytes_read": nbytes, "bytes_written": nbytes, "reps": reps})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="torus128", choices=sorted(WORKLOADS) + list(EXTRA))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="single-vector workloads: skip the reliable-BiCG and FP64-GMRES records beside the two paths")
    ap.add_argument("--no-configs", action="store_true",
                    help="default run: only the torus128 line, without the other workloads' records")
    ap.add_argument("--no-star", action="store_true", help="default run: leave out the star256 record")
    ap.add_argument("--star-iterations", type=int, default=1000, help="iteration budget of the star256 workload")
    ap.add_argument("--star-formulation", choices=("combined", "double"), default="combined",
                    help="star256: combined-field operator (eta = k) or the double layer alone")
    ap.add_argument("--star-k", type=float, default=40.0, help="wavenumber of the star256 workload")
    ap.add_argument("--star-mixed-only", action="store_true",
                    help="star256: time the mixed path alone (no FP64 path, no comparison)")
    ap.add_argument("--star-fp64-products", action="store_true",
                    help="star256 mixed path: 16-column passes on the FP64 matrix cores (default: FP32)")
    ap.add_argument("--star-fp64-cache", type=float, default=STAR_FP64_CACHE_BYTES,
                    help="star256 FP64 path: bytes of complex128 pairs the matrix-free operator keeps (0: none)")
    ap.add_argument("--star-replace", type=float, default=STAR_REPLACE,
                    help="star256 mixed path with --star-solver reliable: residual drop between FP64 residual "
                         "replacements")
    ap.add_argument("--star-solver", choices=("block", "gmres", "reliable"), default="block",
                    help="star256 mixed path: FP64 refinement around block GMRES (one Krylov space for the 16 "
                         "columns) or GMRES per column (one pass A P per iteration), or batched BiCG with FP64 "
                         "residual replacement")
    ap.add_argument("--batch-products", choices=("auto", "fp32", "fp64"), default="auto",
                    help="batch mixed path: matrix cores of the 16-column passes (auto: FP32 for block GMRES, "
                         "FP64 for the per-column solvers)")
    ap.add_argument("--batch-graphs", action="store_true",
                    help="batch workload: replay the grouped iterations as HIP graphs (no per-pass timing)")
    ap.add_argument("--batch-replace", type=float, default=BATCH_REPLACE,
                    help="batch mixed path with --batch-solver reliable: residual drop between FP64 residual "
                         "replacements")
    ap.add_argument("--batch-solver", choices=("block", "gmres", "reliable"), default="block",
                    help="batch mixed path: FP64 refinement around grouped block GMRES (one Krylov space per "
                         "batch of 16) or GMRES per column (one pass Y = A P per iteration), or grouped BiCG "
                         "with FP64 residual replacement")
    ap.add_argument("--replace", type=float, default=None,
                    help="single-vector workloads: residual drop between the mixed path's FP64 replacements")
    ap.add_argument("--mixed-solver", choices=("enlarged", "enlarged32", "gmres", "reliable"), default=None,
                    help="single-vector workloads: the mixed path's solver (default: the workload's own)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (reference for the HBM traffic counters)")
    ap.add_argument("--matrix-free", action="store_true",
                    help="store no operator: both paths recompute the radial factors per pair and sweep")
    args = ap.parse_args()
    global MATRIX_FREE, REPLACE_OVERRIDE, SOLVER_OVERRIDE
    MATRIX_FREE = args.matrix_free
    REPLACE_OVERRIDE = args.replace
    SOLVER_OVERRIDE = args.mixed_solver
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload in EXTRA:
        emit({"batch": run_batch, "adjoint96": run_adjoint, "star256": run_star}[args.workload](args))
        return
    line = run_standard(args, args.workload)
    if args.workload == "torus128" and not args.matrix_free and not args.no_configs:
        line["configs"] = other_configs(args)
    print(json.dumps(line), flush=True)


def run_standard(args, workload):
    """One of the single-right-hand-side workloads (torus128, sphere64, ...):
    the FP64 and mixed paths timed, checked against an FP64 solve and the
    analytic field. Returns the JSON record."""
    w = WORKLOADS[workload]
    level = level_of(w)
    recv = shapes.shell_receivers(NRECV, 1.5, 3.0, seed=0)
    exact = shapes.point_source(recv, w["x0"], w["k"])
    print("[bench] workload %s: uploading level set" % workload, file=sys.stderr, flush=True)

    # FP64 reference density (untimed): the dense complex128 operator (the
    # reference's entries), BiCG to REF_TOL; its sweep time is reported beside
    p64 = make_problem(level, w)
    p64.geometry(5)
    count = p64.tube(params_of(w), ibim.TUBE_TRADITIONAL)
    p64.assemble(ibim.C128)
    p64.rhs(ibim.RHS_POINT_SOURCE, x0=w["x0"])
    p64.set_profiling(True)
    info_ref, _ = p64.solve(ibim.SOLVE_BICG, maxit=4000, tol=REF_TOL)
    p64.sync()
    ms_dense, n_dense = p64.kernel_time("matvec")
    p64.set_profiling(False)
    p64.reset_timers()
    xr = p64.density()
    x_ref = xr[0] + 1j * xr[1]
    print("[bench] count=%d reference BiCG (dense complex128): %d iterations, status %d" %
          (count, info_ref.iterations, info_ref.status), file=sys.stderr, flush=True)

    t64, info64, x64, u64 = timed(p64, w, recv, "fp64", args.steps, args.warmup)
    ms_mv, n_mv = p64.kernel_time("sweep_free64" if MATRIX_FREE else "matvec")
    ms_asm, n_asm = p64.kernel_time("assemble")
    print("[bench] fp64 path: %.3f s/solve, %d iterations" % (t64, info64.iterations), file=sys.stderr, flush=True)
    p64.close()

    pm = make_problem(level, w)
    tm, infom, xm, um = timed(pm, w, recv, "mixed", args.steps, args.warmup)
    stored = mixed_stores(count)
    enlarged = mixed_solver_of(w) in ("enlarged", "enlarged32")
    dom_timer = ({"enlarged32": "batch_pair32", "enlarged": "batch_pair" if stored else "batch_free32"}
                 [mixed_solver_of(w)] if enlarged else "matvec" if stored else "sweep_free32")
    ms_dom, n_dom = pm.kernel_time(dom_timer)
    ms_rep, n_rep = pm.kernel_time("matvec64" if mixed_storage(count) == ibim.MIXED_PAIR else "sweep_free64")
    msolver = mixed_solver_of(w)
    msolver_text = {
        "enlarged32": "entries and products on the fp32 matrix cores (16-column passes, fp64 sums per 64 columns), "
                      "enlarged GMRES, fp64 iterative refinement",
        "enlarged": "fp64 products (16-column passes), enlarged GMRES, fp64 iterative refinement",
        "gmres": "fp64 point factors, products and sums, GMRES, fp64 iterative refinement"}.get(
        msolver, "fp64 point factors, products and sums, BiCG, fp64 residual replacement")
    gmres = msolver in ("gmres", "enlarged", "enlarged32")
    solver = {"gmres": "FP64 refinement around GMRES", "enlarged": "FP64 refinement around enlarged GMRES",
              "enlarged32": "FP64 refinement around enlarged GMRES, FP32 matrix-core products"}.get(
        msolver, "BiCG with FP64 residual replacement")
    print("[bench] mixed path (%s): %.3f s/solve, %d iterations, %d FP64 residuals" %
          (solver, tm, infom.iterations, infom.outer), file=sys.stderr, flush=True)
    # beside them, fewer steps: the former mixed solver (reliable BiCG on the same
    # complex64 pairs) and FP64 GMRES on the complex128 pairs, which separate the
    # precision's gain from the solver's
    extra = {}
    if not args.no_extra:
        xs, ws = min(args.steps, 5), min(args.warmup, 1)
        # FP64 GMRES needs a stored complex128 operator; mixed_gmres: GMRES column by
        # column on the mixed operator, when the mixed path is enlarged GMRES
        names = ("mixed_reliable", "fp64_gmres") if stored else ("mixed_reliable",)
        if enlarged:
            names = ("mixed_gmres",) + names
        for name in names:
            pe = make_problem(level, w)
            te, infoe, xe, _ = timed(pe, w, recv, name, xs, ws)
            extra[name] = (te, infoe, xe)
            pe.close()
            print("[bench] %s: %.3f s/solve, %d iterations" % (name, te, infoe.iterations), file=sys.stderr,
                  flush=True)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    err_mixed = rel(xm[0] + 1j * xm[1], x_ref)
    err_fp64 = rel(x64[0] + 1j * x64[1], x_ref)
    field_err = rel(um[0] + 1j * um[1], exact)

    n2 = float(count) * count
    npairs = float(count) * (count - 1) / 2
    vec_bytes = 4 * 16 * count  # p, p~ in; A p, A^H p~ out (complex128)
    avg_dom = ms_dom / max(n_dom, 1) * 1e-3
    if enlarged:
        # dominant kernel: the 16-column pass Y = A P over the operator (every
        # ordered entry times 16 columns, 8 flops per complex multiply-add) on the
        # FP32 (stored complex64 pairs, FP32 products) or FP64 matrix cores
        flops = float(count) * count * 16 * 8
        peak = FP32_MATRIX_PEAK_TFLOPS if mixed_solver_of(w) == "enlarged32" else FP64_PEAK_TFLOPS
        achieved = flops / avg_dom / 1e12 if n_dom else 0.0
        kname = {"batch_pair32": "k_pair_batch32<1, A P>", "batch_pair": "k_pair_batch<float2, A P>",
                 "batch_free32": "k_pair_batch<double2, fast factors, A P>"}[dom_timer]
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": measured_traffic(kname.replace(" ", ""), workload),
                "kernel": "%s: 16 columns x %d^2 entries x 8 flops per launch (enlarged GMRES)" % (kname, count)}
    elif stored and gmres:
        # dominant kernel: GMRES's pair sweep over the complex64 radial factors,
        # A p alone (p in, A p out: 32 B per point)
        dom_bytes = npairs * 8 + 32 * count
        achieved = dom_bytes / avg_dom / 1e9 if n_dom else 0.0
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_pair_sweep<float2,A_p>", workload),
                "kernel": "k_pair_sweep<float2, A p only>: 8 B/pair x %d*%d/2 pairs + 32 B/point vectors per launch"
                          % (count, count - 1)}
    elif stored:
        # dominant kernel: BiCG's fused pair sweep over the complex64 radial factors
        dom_bytes = npairs * 8 + vec_bytes
        achieved = dom_bytes / avg_dom / 1e9 if n_dom else 0.0
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_pair_sweep<float2>", workload),
                "kernel": "k_pair_sweep<float2>: 8 B/pair x %d*%d/2 pairs + 64 B/point vectors per launch" %
                          (count, count - 1)}
    else:
        # dominant kernel: the pair sweep computing the mixed-precision factors,
        # one evaluation per pair (C_EVAL) + point factors alpha, beta (12) + four
        # complex multiply-adds with their scalings (36), FP64 on the vector
        # units, whose peak equals the FP64 matrix-core figure on this GPU
        # (GMRES: A p alone, two complex multiply-adds with their scalings, 18)
        per_pair = C_EVAL + (30 if gmres else 48)
        flops = npairs * per_pair
        achieved = flops / avg_dom / 1e12 if n_dom else 0.0
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                "kernel": "k_pair_sweep<double2, fast factors%s>: %d flops/pair x %d*%d/2 pairs per launch" %
                          (", A p only" if gmres else "", per_pair, count, count - 1)}
    # the FP64 path's fused sweep (HBM-bound): one pass over the complex128 pairs
    avg_mv = ms_mv / max(n_mv, 1) * 1e-3
    mv_gbs = (npairs * 16 + vec_bytes) / avg_mv / 1e9 if n_mv and not MATRIX_FREE else None
    avg_dense = ms_dense / max(n_dense, 1) * 1e-3

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
        "dtype": ("fp64 with fp32 sine / cosine in the recomputed radial factors, %s" % msolver_text
                  if MATRIX_FREE else
                  "complex64 radial factors per pair, %s" % msolver_text),
        "data": "synthetic: %s signed-distance level set, point-source Dirichlet data" % w["shape"],
        "config": {"workload": "%s: %s N=%d, k=%g, eps=%gh, tau=%gh, point source %s, %d receivers, tol %g%s" % (
            workload, w["shape"], w["n"], w["k"], w["eps_h"], w["tau_h"], list(w["x0"]), NRECV, w["tol"],
            ", matrix-free" if MATRIX_FREE else ""),
            "tube_points": count, "parallelism": "single GPU",
            "mixed_operator": {ibim.MIXED_PAIR: "complex64 pair storage + complex128 copy for the FP64 residuals",
                               ibim.C64_PAIR: "complex64 pair storage"}.get(mixed_storage(count),
                                                                            "radial factors recomputed per pair"),
            "fp64_operator": ("FP64 radial factors recomputed per pair" if MATRIX_FREE else
                              "complex128 pair storage" if pair_bytes(count, 16) <= STORE_MAX_BYTES else
                              "dense complex128")},
        "mixed": {"solver": solver, "solves_per_s": 1.0 / tm, "iterations": infom.iterations,
                  "fp64_residuals": infom.outer,
                  "rel_err_vs_fp64": err_mixed, "field_rel_err_vs_analytic": field_err,
                  "sweep_ms": avg_dom * 1e3, "replacement_sweep_ms": ms_rep / max(n_rep, 1)},
        "fp64": {"solves_per_s": 1.0 / t64, "iterations": info64.iterations, "rel_err_vs_fp64": err_fp64,
                 "assemble_ms": ms_asm / max(n_asm, 1), "matvec_ms": avg_mv * 1e3,
                 "matvec_GBs": mv_gbs, "matvec_frac_of_peak": mv_gbs / HBM_PEAK_GBS if mv_gbs else None,
                 "matvec_traffic": None if MATRIX_FREE else measured_traffic("k_pair_sweep<double2>", workload)},
        "dense_c128_reference": {"iterations": info_ref.iterations, "tol": REF_TOL, "sweep_ms": avg_dense * 1e3,
                                 "sweep_GBs": (n2 * 16 + vec_bytes) / avg_dense / 1e9 if n_dense else 0.0},
        "speedup_mixed_vs_fp64": t64 / tm,
        "roofline": roof,
    }
    for name, (te, infoe, xe) in extra.items():
        line[name] = {"solves_per_s": 1.0 / te, "iterations": infoe.iterations,
                      "rel_err_vs_fp64": rel(xe[0] + 1j * xe[1], x_ref)}
    if "fp64_gmres" in extra:
        line["speedup_mixed_vs_fp64_gmres"] = extra["fp64_gmres"][0] / tm
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, bounded sample)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(w, count, info64.iterations)
    return line


def other_configs(args):
    """The other workloads of the default run, one compact record each: solves
    (or gradients) per second of both paths, the mixed path's error against
    the FP64 one, and the checks each workload carries."""
    global MATRIX_FREE
    out = {}
    sub = argparse.Namespace(**vars(args))
    sub.no_cpu = True
    sub.steps, sub.warmup = min(args.steps, 5), min(args.warmup, 1)
    for name, fn in (("sphere64", lambda: run_standard(sub, "sphere64")),
                     ("torus128_matrix_free", None),
                     ("batch", lambda: run_batch(sub)),
                     ("adjoint96", lambda: run_adjoint(sub)),
                     ("star256", lambda: run_star(sub))):
        if name == "star256" and args.no_star:
            continue
        t0 = time.perf_counter()
        if name == "torus128_matrix_free":
            MATRIX_FREE = True
            try:
                rec = run_standard(sub, "torus128")
            finally:
                MATRIX_FREE = False
        else:
            rec = fn()
        out[name] = compact(name, rec)
        out[name]["wall_s"] = time.perf_counter() - t0
        print("[bench] %s: %s" % (name, json.dumps(out[name])), file=sys.stderr, flush=True)
    return out


def compact(name, r):
    """The figures of a workload record that the default run reports."""
    if name in ("sphere64", "torus128_matrix_free"):
        d = {"workload": r["config"]["workload"], "tube_points": r["config"]["tube_points"],
                "mixed_solver": r["mixed"]["solver"],
                "solves_per_s": r["mixed"]["solves_per_s"], "fp64_solves_per_s": r["fp64"]["solves_per_s"],
                "speedup_mixed_vs_fp64": r["speedup_mixed_vs_fp64"],
                "rel_err_vs_fp64": {"mixed": r["mixed"]["rel_err_vs_fp64"], "fp64": r["fp64"]["rel_err_vs_fp64"]},
                "field_rel_err_vs_analytic": r["mixed"]["field_rel_err_vs_analytic"],
                "iterations": {"mixed": r["mixed"]["iterations"], "fp64": r["fp64"]["iterations"]},
                "mixed_operator": r["config"]["mixed_operator"], "fp64_operator": r["config"]["fp64_operator"],
                "roofline": r["roofline"]}
        for k in ("mixed_gmres", "mixed_reliable", "fp64_gmres"):
            if k in r:
                d[k + "_solves_per_s"] = r[k]["solves_per_s"]
        return d
    if name == "batch":
        return {"workload": r["config"]["workload"], "rhs": r["config"]["rhs"], "solves_per_s": r["value"],
                "mixed_solver": r["config"]["mixed_solver"],
                "fp64_solves_per_s": r["fp64_solves_per_s"], "speedup_mixed_vs_fp64": r["speedup_mixed_vs_fp64"],
                "worst_rel_err_mixed_vs_fp64": r["worst_rel_err_mixed_vs_fp64"],
                "schedule": r["config"]["schedule"],
                **({"fp64_block_gmres_solves_per_s": r["fp64_block_gmres"]["solves_per_s"]}
                   if "fp64_block_gmres" in r else {})}
    if name == "adjoint96":
        keys = ("value", "unit", "forward_s", "gradient_s", "forward_iterations", "reference_bicg",
                "fd_entries_resolved", "fd_worst_rel_err", "fd_step_1e-6", "dk_rel_err")
        d = {k: r[k] for k in keys}
        d["workload"] = r["config"]["workload"]
        d["tube_points"] = r["config"]["tube_points"]
        return d
    keys = ("metric", "value", "unit", "fp64_rate", "speedup_mixed_vs_fp64", "mixed_rel_err_vs_fp64_density",
            "seconds", "iterations_summed_over_columns", "residual", "converged", "pass_ms", "pass_tflops")
    d = {k: r[k] for k in keys}
    d["workload"] = r["config"]["workload"]
    d["tube_points"] = r["config"]["tube_points"]
    d["mixed_operator"] = r["config"]["mixed_operator"]
    d["mixed_solver"] = r["config"]["mixed_solver"]
    return d


if __name__ == "__main__":
    main()
