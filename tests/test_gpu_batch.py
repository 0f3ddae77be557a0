"""Batched right-hand sides: the 16-wide matrix-core products and the batched
BiCG against dense numpy products and the one-column-at-a-time solver, on dense,
pair (one fused pass for A P and A^H Q) and matrix-free operators."""
import numpy as np
import pytest

import shapes

pytestmark = pytest.mark.gpu

N = 28  # 3,7xx tube points: not a multiple of 16 or 64, so padding is exercised


def problem(gpu, k=4.0, dirichlet=1):
    pb = gpu.Problem()
    pb.set_level(shapes.sphere(N))
    pb.geometry(5)
    h = 4.0 / N
    pb.tube(gpu.params(k, 3 * h, 2 * h, dirichlet=dirichlet))
    return pb


def dense(pb):
    re, im = pb.rows(0, pb.count)
    return re + 1j * im


def relnorm(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nvec", [5, 16])
def test_batched_products_match_dense(gpu, nvec):
    pb = problem(gpu)
    rng = np.random.default_rng(nvec)
    P = rng.standard_normal((pb.count, nvec)) + 1j * rng.standard_normal((pb.count, nvec))
    pb.assemble(gpu.C128)
    A = dense(pb)
    for adjoint in (False, True):
        ref = (A.conj().T if adjoint else A) @ P
        assert relnorm(pb.apply_batch(P, gpu.OP_STORED, adjoint), ref) < 1e-13
        assert relnorm(pb.apply_batch(P, gpu.OP_FREE64, adjoint), ref) < 1e-13
        assert relnorm(pb.apply_batch(P, gpu.OP_FREE32, adjoint), ref) < 1e-6
    pb.assemble(gpu.C64)
    A32 = dense(pb)  # the stored complex64 values
    assert relnorm(A32, A) < 1e-7
    for adjoint in (False, True):
        ref = (A32.conj().T if adjoint else A32) @ P
        assert relnorm(pb.apply_batch(P, gpu.OP_STORED, adjoint), ref) < 1e-13


@pytest.mark.parametrize("dirichlet", [1, 0])
def test_batched_pair_products_match_their_rows(gpu, dirichlet):
    """Pair storage: A P and A^H P against numpy on the rows the storage
    rebuilds (complex128 pairs: also against the dense complex128 operator)."""
    pb = problem(gpu, dirichlet=dirichlet)
    rng = np.random.default_rng(3 + dirichlet)
    P = rng.standard_normal((pb.count, 16)) + 1j * rng.standard_normal((pb.count, 16))
    pb.assemble(gpu.C128)
    A = dense(pb)
    for store in ("C128_PAIR", "C64_PAIR", "MIXED_PAIR"):
        pb.assemble(getattr(gpu, store))
        Ap = dense(pb)
        if store == "C128_PAIR":
            assert relnorm(Ap, A) < 1e-14
        for adjoint in (False, True):
            ref = (Ap.conj().T if adjoint else Ap) @ P
            got = pb.apply_batch(P, gpu.OP_STORED, adjoint)
            assert relnorm(got, ref) < 1e-13, (store, adjoint)
            assert np.array_equal(got, pb.apply_batch(P, gpu.OP_STORED, adjoint))  # fixed order


def test_batched_pair_pass_shapes_give_the_same_bits(gpu, monkeypatch):
    """The fused pair pass runs 16, 32 or 64 rows per workgroup (IBIM_BATCH_WAVES
    forces one, else the fastest is timed on the first call): every wave walks
    its rows' columns in the same order, so products and solves are bitwise
    the same for every shape."""
    pb = problem(gpu)
    rng = np.random.default_rng(11)
    P = rng.standard_normal((pb.count, 16)) + 1j * rng.standard_normal((pb.count, 16))
    pb.assemble(gpu.C64_PAIR)
    fill_rhs(gpu, pb, 16)
    runs = []
    for w in ("1", "2", "4", None):
        if w is None:
            monkeypatch.delenv("IBIM_BATCH_WAVES", raising=False)
        else:
            monkeypatch.setenv("IBIM_BATCH_WAVES", w)
        y = pb.apply_batch(P, gpu.OP_STORED, False)
        z = pb.apply_batch(P, gpu.OP_STORED, True)
        info, _ = pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, batch=0)
        x = [pb.density(c) for c in range(16)]
        runs.append((w, y, z, info.iterations, x))
    for w, y, z, iters, x in runs[1:]:
        assert np.array_equal(y, runs[0][1]) and np.array_equal(z, runs[0][2]), w
        assert iters == runs[0][3], w
        for c in range(16):
            assert np.array_equal(x[c][0], runs[0][4][c][0]) and np.array_equal(x[c][1], runs[0][4][c][1]), (w, c)



@pytest.mark.parametrize("dirichlet,interior", [(1, 0), (0, 0), (1, 1)])
def test_matrix_free_batched_pass_gives_the_pair_storage_bits(gpu, dirichlet, interior):
    """The FP64 matrix-free batched products evaluate each pair's factor where
    the complex128 pair storage reads it, with the assembly's operations, so
    products and batched solves are that storage's bits. The matrix-free
    problem builds its point table and near lists alone, and again when its
    tube changes (a wider near field)."""
    def make(tau_h):
        pb = gpu.Problem()
        pb.set_level(shapes.sphere(N))
        pb.geometry(5)
        h = 4.0 / N
        pb.tube(gpu.params(4.0, 3 * h, tau_h * h, dirichlet=dirichlet, interior=interior))
        return pb

    rng = np.random.default_rng(21 + dirichlet + 2 * interior)
    free = make(2.0)
    for tau_h in (2.0, 2.5):
        if tau_h != 2.0:
            h = 4.0 / N
            free.tube(gpu.params(4.0, 3 * h, tau_h * h, dirichlet=dirichlet, interior=interior))
        stored = make(tau_h)
        stored.assemble(gpu.C128_PAIR)
        assert free.count == stored.count
        P = rng.standard_normal((free.count, 16)) + 1j * rng.standard_normal((free.count, 16))
        for adjoint in (False, True):
            got = free.apply_batch(P, gpu.OP_FREE64, adjoint)
            assert np.array_equal(got, stored.apply_batch(P, gpu.OP_STORED, adjoint)), (tau_h, adjoint)
            assert np.array_equal(got, free.apply_batch(P, gpu.OP_FREE64, adjoint))  # fixed order
            # the mixed-precision factors (FP32 sine / cosine): the same pass,
            # and the single-vector sweep's operator
            fast = free.apply_batch(P, gpu.OP_FREE32, adjoint)
            assert relnorm(fast, got) < 1e-7, (tau_h, adjoint, relnorm(fast, got))
            for c in (0, 9):
                y, z = free.apply(P[:, c], P[:, c], gpu.OP_FREE32)
                assert relnorm(fast[:, c], z if adjoint else y) < 1e-13, (tau_h, adjoint, c)
        fill_rhs(gpu, free, 19)
        fill_rhs(gpu, stored, 19)
        ia, _ = free.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, batch=0)
        ib, _ = stored.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, batch=0)
        assert ia.status == 0 and (ia.iterations, ia.status) == (ib.iterations, ib.status), tau_h
        for c in range(19):
            a, b = free.density(c), stored.density(c)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (tau_h, c)
        stored.close()
    free.close()

def fill_rhs(gpu, pb, n):
    pb.set_nrhs(n)
    rng = np.random.default_rng(7)
    for c in range(n):
        if c == 3:
            continue  # a zero column: zero density, no iterations
        if c % 2:
            d = rng.standard_normal(3)
            pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=d / np.linalg.norm(d), col=c)
        else:
            pb.rhs(gpu.RHS_POINT_SOURCE, x0=0.3 * rng.standard_normal(3), col=c)


STORES = {"c128": "C128", "c64": "C64", "c128_pair": "C128_PAIR", "c64_pair": "C64_PAIR",
          "mixed_pair": "MIXED_PAIR"}


@pytest.mark.parametrize("store", ["c128", "c64", "free", "c128_pair", "c64_pair"])
def test_batched_bicg_matches_column_solves(gpu, store):
    pb = problem(gpu)
    if store in STORES:
        pb.assemble(getattr(gpu, STORES[store]))
    n = 19  # one full batch of 16 and one of 3
    fill_rhs(gpu, pb, n)
    info_b, hist_b = pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, batch=0)
    xb = [pb.density(c) for c in range(n)]
    info_s, hist_s = pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, batch=1)
    xs = [pb.density(c) for c in range(n)]
    assert info_b.status == 0 and info_s.status == 0
    for c in range(n):
        b = xb[c][0] + 1j * xb[c][1]
        s = xs[c][0] + 1j * xs[c][1]
        if c == 3:
            assert not b.any() and not s.any()
        else:
            assert relnorm(b, s) < 1e-10, (c, relnorm(b, s))
    # the same recurrences: iteration counts agree up to rounding-order effects
    assert abs(info_b.iterations - info_s.iterations) <= 0.05 * info_s.iterations + n
    # column 0's residual history follows the single-column solver's
    m = min(len(hist_b), len(hist_s), 10)
    np.testing.assert_allclose(hist_b[:m], hist_s[:m], rtol=1e-6)


@pytest.mark.parametrize("store", ["c64", "free32", "mixed_pair"])
def test_batched_refinement_reaches_the_fp64_solution(gpu, store):
    pb = problem(gpu)
    n = 19
    fill_rhs(gpu, pb, n)
    pb.assemble(gpu.C128)
    info64, _ = pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-13, batch=0)
    ref = [pb.density(c) for c in range(n)]
    if store in STORES:
        # inner: stored complex64 (dense or pairs); outer: FP64 matrix-free or
        # the complex128 copy of the pairs
        pb.assemble(getattr(gpu, STORES[store]))
    info, _ = pb.solve(gpu.SOLVE_REFINED, maxit=500, tol=1e-10, inner_tol=1e-6, max_outer=10, batch=0)
    assert info.status == 0 and info.rel_residual < 1e-10 and info.outer >= 2
    for c in range(n):
        x = pb.density(c)
        x = x[0] + 1j * x[1]
        r = ref[c][0] + 1j * ref[c][1]
        if c == 3:
            assert not x.any()
        else:
            assert relnorm(x, r) < 1e-8, (c, relnorm(x, r))


@pytest.mark.parametrize("store", ["c64", "free32", "c64_pair", "mixed_pair"])
def test_batched_residual_replacement_reaches_the_fp64_solution(gpu, store):
    """Batched BiCG on the FP32 operator with FP64 residual replacement: every
    column reaches the FP64 operator's solution with a true residual below tol
    and about the FP64 iteration count (no restarts)."""
    pb = problem(gpu)
    n = 19
    fill_rhs(gpu, pb, n)
    pb.assemble(gpu.C128)
    info64, _ = pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-13, batch=0)
    ref = [pb.density(c) for c in range(n)]
    if store in STORES:
        # inner: stored complex64; replacements: FP64 matrix-free, or the
        # complex128 copy of the pairs (mixed_pair)
        pb.assemble(getattr(gpu, STORES[store]))
    pb.reset_timers()
    pb.set_profiling(True)
    info, hist = pb.solve(gpu.SOLVE_RELIABLE, maxit=500, tol=1e-10, inner_tol=1e-2, batch=0)
    pb.sync()
    pb.set_profiling(False)
    assert info.status == 0 and info.rel_residual < 1e-10 and info.outer >= 2
    if store.endswith("pair"):  # one fused pass per iteration over the pairs
        assert pb.kernel_time("batch_pair")[1] > 0 and pb.kernel_time("batch_rows")[1] == 0
        assert (pb.kernel_time("batch_pair64")[1] > 0) == (store == "mixed_pair")
    for c in range(n):
        x = pb.density(c)
        x = x[0] + 1j * x[1]
        r = ref[c][0] + 1j * ref[c][1]
        if c == 3:
            assert not x.any()
        else:
            assert relnorm(x, r) < 1e-8, (c, relnorm(x, r))
    # the same solve one column at a time
    info1, _ = pb.solve(gpu.SOLVE_RELIABLE, maxit=500, tol=1e-10, inner_tol=1e-2, batch=1)
    assert info1.status == 0
    assert info.iterations <= 1.3 * info1.iterations + n


def test_contexts_solve_concurrently_from_host_threads(gpu):
    """Independent handles (one stream each) solved from several host threads
    at once, as the batch workload runs its smaller surfaces: the same bits as
    one after another."""
    from concurrent.futures import ThreadPoolExecutor

    def solve(n):
        pb = gpu.Problem()
        pb.set_level(shapes.sphere(n))
        pb.geometry(5)
        h = 4.0 / n
        pb.tube(gpu.params(4.0, 3 * h, 2 * h, dirichlet=1))
        pb.assemble(gpu.C128_PAIR)
        fill_rhs(gpu, pb, 16)
        info, _ = pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, batch=0)
        x = [pb.density(c) for c in range(16)]
        pb.close()
        return info.iterations, x

    sizes = (20, 24, 28)
    seq = [solve(n) for n in sizes]
    with ThreadPoolExecutor(max_workers=len(sizes)) as pool:
        par = list(pool.map(solve, sizes))
    for (i0, x0), (i1, x1) in zip(seq, par):
        assert i0 == i1
        for c in range(16):
            assert np.array_equal(x0[c][0], x1[c][0]) and np.array_equal(x0[c][1], x1[c][1])


def sphere_problem(gpu, n, store, fp32, dirichlet=1, rhs=16):
    pb = gpu.Problem()
    pb.set_level(shapes.sphere(n))
    pb.geometry(5)
    h = 4.0 / n
    pb.tube(gpu.params(4.0, 3 * h, 2 * h, dirichlet=dirichlet))
    pb.assemble(getattr(gpu, STORES[store]))
    if fp32:
        pb.set_fp32_products(True)
    fill_rhs(gpu, pb, rhs)
    return pb


@pytest.mark.parametrize("store,method,fp32", [("c128_pair", "SOLVE_BICG", False),
                                               ("c64_pair", "SOLVE_BICG", True),
                                               ("mixed_pair", "SOLVE_RELIABLE", False),
                                               ("mixed_pair", "SOLVE_RELIABLE", True),
                                               ("c64_pair", "SOLVE_RELIABLE", True),
                                               ("c128_pair", "SOLVE_GMRES", False),
                                               ("mixed_pair", "SOLVE_GMRES", False),
                                               ("c64_pair", "SOLVE_GMRES", True),
                                               ("mixed_pair", "SOLVE_BLOCK_GMRES", False),
                                               ("c64_pair", "SOLVE_BLOCK_GMRES", True)])
def test_group_solve_gives_each_problem_its_own_bits(gpu, store, method, fp32):
    """Several problems of uneven sizes solved together (one persistent launch
    of every operator per iteration, a problem with two batches among them):
    each problem's densities, iteration count and status are those of its own
    batched solve, bit for bit; the operators run only in the grouped launch."""
    m = getattr(gpu, method)
    sizes, nrhs = (20, 28, 24), (16, 19, 5)
    # chunk: BiCG's host-check interval; GMRES: its basis length (0: the default)
    kw = dict(maxit=500, tol=1e-10, inner_tol=1e-2, chunk=0 if "GMRES" in method else 2)
    own = []
    for n, r in zip(sizes, nrhs):
        pb = sphere_problem(gpu, n, store, fp32, rhs=r)
        info, _ = pb.solve(m, batch=0, **kw)
        own.append((info, [pb.density(c) for c in range(r)]))
        pb.close()
    pbs = [sphere_problem(gpu, n, store, fp32, rhs=r) for n, r in zip(sizes, nrhs)]
    for pb in pbs:
        pb.reset_timers()
        pb.set_profiling(True)
    infos = gpu.solve_group(pbs, m, **kw)
    for pb in pbs:
        pb.sync()
    launches = pbs[0].kernel_time("batch_group")[1]
    for pb, info, (info0, x0), r in zip(pbs, infos, own, nrhs):
        assert info.status == 0 and info.status == info0.status
        assert info.iterations == info0.iterations and info.outer == info0.outer
        assert info.rel_residual == info0.rel_residual
        for c in range(r):
            x = pb.density(c)
            assert np.array_equal(x[0], x0[c][0]) and np.array_equal(x[1], x0[c][1]), c
        # the single-problem passes did not run (the FP64 passes of the replacements
        # and refinement residuals aside: on complex128 pairs GMRES takes them
        # from the stored pairs themselves, one per refinement step and batch)
        own_passes = (r + 15) // 16 * info.outer if (method, store) == ("SOLVE_GMRES", "c128_pair") else 0
        assert pb.kernel_time("batch_pair")[1] <= own_passes and pb.kernel_time("batch_pair32")[1] == 0
        pb.set_profiling(False)
    longest = max(i.iterations for i in infos)  # bounds every column's own count
    assert 0 < launches <= longest + 2
    for pb in pbs:
        pb.close()


def test_group_solve_refuses_what_cannot_share_a_launch(gpu):
    a = sphere_problem(gpu, 20, "c128_pair", False)
    b = sphere_problem(gpu, 24, "c128_pair", False, dirichlet=0)
    with pytest.raises(gpu.IbimError):
        gpu.solve_group([a, b], gpu.SOLVE_BICG)  # Dirichlet and Neumann kernels
    c = sphere_problem(gpu, 24, "c64_pair", False)
    with pytest.raises(gpu.IbimError):
        gpu.solve_group([a, c], gpu.SOLVE_BICG)  # complex128 and complex64 pairs
    with pytest.raises(gpu.IbimError):
        gpu.solve_group([a], gpu.SOLVE_RELIABLE)  # the reliable solver needs complex64 pairs
    with pytest.raises(gpu.IbimError):
        gpu.solve_group([a, a], gpu.SOLVE_BICG)  # one problem twice
    d = sphere_problem(gpu, 24, "c128", False)
    with pytest.raises(gpu.IbimError):
        gpu.solve_group([a, d], gpu.SOLVE_BICG)  # dense storage
    for pb in (a, b, c, d):
        pb.close()


@pytest.mark.parametrize("combined", [0, 1])
def test_matrix_free_cache_keeps_the_bits(gpu, combined):
    """ibim_set_free_cache: the FP64 matrix-free operator keeps the complex128
    pairs of its first row blocks (here 2 of 8) and computes the rest, giving
    the uncached operator's bits in single-vector and 16-column products and in
    batched BiCG; the cache is built once per tube, again after a tube change,
    and released by a zero budget."""
    h = 4.0 / N

    def make():
        pb = gpu.Problem()
        pb.set_level(shapes.sphere(N))
        pb.geometry(5)
        pb.tube(gpu.params(4.0, 3 * h, 2 * h, dirichlet=1, combined=combined, eta=4.0))
        return pb

    plain, cached = make(), make()
    assert cached.count > 7 * 512  # 8 row blocks of 512 rows
    cached.set_free_cache(10 * 512 * 1024 * 16)  # 10 tiles: the 8 of row blocks 0 and 1
    cached.set_profiling(True)
    rng = np.random.default_rng(5)
    p = rng.standard_normal(plain.count) + 1j * rng.standard_normal(plain.count)
    q = rng.standard_normal(plain.count) + 1j * rng.standard_normal(plain.count)
    P = rng.standard_normal((plain.count, 16)) + 1j * rng.standard_normal((plain.count, 16))
    for _ in range(2):
        y0, z0 = plain.apply(p, q, gpu.OP_FREE64)
        y1, z1 = cached.apply(p, q, gpu.OP_FREE64)
        assert np.array_equal(y0, y1) and np.array_equal(z0, z1)
        for adjoint in (False, True):
            assert np.array_equal(plain.apply_batch(P, gpu.OP_FREE64, adjoint),
                                  cached.apply_batch(P, gpu.OP_FREE64, adjoint)), adjoint
    cached.sync()
    assert cached.kernel_time("assemble")[1] == 1  # the cache fill, once
    for pb in (plain, cached):
        fill_rhs(gpu, pb, 16)
    ia, _ = plain.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, batch=0)
    ib, _ = cached.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, batch=0)
    assert ia.status == 0 and (ia.iterations, ia.status, ia.rel_residual) == (ib.iterations, ib.status,
                                                                                ib.rel_residual)
    for c in range(16):
        a, b = plain.density(c), cached.density(c)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), c
    # a new tube (wider near field): the cache is rebuilt for it
    for pb in (plain, cached):
        pb.tube(gpu.params(4.0, 3 * h, 2.5 * h, dirichlet=1, combined=combined, eta=4.0))
    y0, z0 = plain.apply(p, q, gpu.OP_FREE64)
    y1, z1 = cached.apply(p, q, gpu.OP_FREE64)
    assert np.array_equal(y0, y1) and np.array_equal(z0, z1)
    cached.sync()
    assert cached.kernel_time("assemble")[1] == 2
    cached.set_free_cache(0)
    y1, z1 = cached.apply(p, q, gpu.OP_FREE64)
    assert np.array_equal(y0, y1) and np.array_equal(z0, z1)
    cached.set_profiling(False)
    plain.close()
    cached.close()
