"""The combined-field exterior Dirichlet operator A = 1/2 I + K - i eta S
(ibim_params.combined) on the GPU.

Pinned chain: the oracle's combined rows are pinned by the single layer of a
constant density on the sphere, quadrature of the near-field disc mean and the
exterior point-source field (test_oracle_math); here the dense complex128 GPU
rows match the oracle's to 1e-13 of each entry, the pair storage (G = Phi per
unordered pair, entries rebuilt with u = 1 / rho from the points) matches the
dense rows to 1e-13 (complex128) and 1e-7 of the largest entry (complex64),
sweeps and 16-column passes match the rebuilt rows, the FP64 matrix-free
operator equals the complex128 pairs bit for bit, and solves reproduce the
analytic field through the combined potential (IBIM_EVAL_COMBINED).
"""
import numpy as np
import pytest

import oracle_ctypes as O
import shapes

pytestmark = pytest.mark.gpu


def relnorm(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def problem(gpu, n, k=4.0, eta=None, level=None):
    pb = gpu.Problem()
    lv = shapes.sphere(n) if level is None else level
    pb.set_level(lv)
    pb.geometry(5)
    h = 4.0 / n
    eps = (3.0 if n >= 16 else 1.5) * h
    tau = 1.5 * h if n < 16 else 2 * h
    prm = gpu.params(k, eps, tau, dirichlet=1, combined=1, eta=k if eta is None else eta)
    pb.tube(prm)
    return pb, lv, prm


def dense_rows(pb):
    re, im = pb.rows(0, pb.count)
    return re + 1j * im


@pytest.mark.parametrize("n,k", [(12, 3.0), (28, 6.0)])
def test_combined_dense_rows_match_the_oracle(gpu, n, k):
    pb, lv, prm = problem(gpu, n, k)
    pb.assemble(gpu.C128)
    A = dense_rows(pb)
    geo = O.geometry(lv, n)
    oprm = O.make_params(k=prm.k, eps=prm.eps, tau=prm.tau, dirichlet=1, combined=1, eta=prm.eta)
    node, flat = O.tube(lv, geo, n, oprm, 0)
    assert len(node) == pb.count
    are, aim = O.kernel_rows(flat, pb.count, oprm)
    Ao = are + 1j * aim
    assert np.abs(A - Ao).max() <= 1e-13 * np.abs(Ao).max()
    nz = np.abs(Ao) > 1e-3 * np.abs(Ao).max()
    assert (np.abs(A - Ao)[nz] / np.abs(Ao)[nz]).max() <= 1e-12
    # the coupling is really there: the rows differ from the double layer's
    pb0 = gpu.Problem()
    pb0.set_level(lv)
    pb0.geometry(5)
    pb0.tube(gpu.params(prm.k, prm.eps, prm.tau, dirichlet=1))
    pb0.assemble(gpu.C128)
    assert relnorm(dense_rows(pb0), A) > 0.05


@pytest.mark.parametrize("n", [12, 28])
def test_combined_pair_rows_match_dense(gpu, n):
    pb, _, _ = problem(gpu, n)
    pb.assemble(gpu.C128)
    A = dense_rows(pb)
    scale = np.abs(A).max()
    pb.assemble(gpu.C128_PAIR)
    Ap = dense_rows(pb)
    assert np.abs(Ap - A).max() <= 1e-13 * scale
    nz = np.abs(A) > 1e-3 * scale
    assert (np.abs(Ap - A)[nz] / np.abs(A)[nz]).max() <= 1e-12
    pb.assemble(gpu.C64_PAIR)
    A32 = dense_rows(pb)
    assert np.abs(A32 - A).max() <= 1e-7 * scale
    assert relnorm(A32, A) <= 1e-7


@pytest.mark.parametrize("n", [12, 28])
@pytest.mark.parametrize("precision", ["C128_PAIR", "C64_PAIR", "C128", "C64"])
def test_combined_sweeps_match_rebuilt_rows(gpu, n, precision):
    pb, _, _ = problem(gpu, n)
    pb.assemble(getattr(gpu, precision))
    A = dense_rows(pb)
    rng = np.random.default_rng(n)
    p = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    q = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    y, z = pb.apply(p, q, gpu.OP_STORED)
    assert relnorm(y, A @ p) <= 1e-13
    assert relnorm(z, A.conj().T @ q) <= 1e-13
    if precision.endswith("PAIR"):
        # 16 right-hand sides on the matrix cores, both products
        P = rng.standard_normal((pb.count, 16)) + 1j * rng.standard_normal((pb.count, 16))
        Y = pb.apply_batch(P, gpu.OP_STORED)
        Z = pb.apply_batch(P, gpu.OP_STORED, adjoint=True)
        assert relnorm(Y, A @ P) <= 1e-13
        assert relnorm(Z, A.conj().T @ P) <= 1e-13


@pytest.mark.parametrize("n", [12, 28])
def test_combined_matrix_free_is_the_complex128_pairs(gpu, n):
    pb, _, _ = problem(gpu, n)
    rng = np.random.default_rng(3 + n)
    p = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    q = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    P = rng.standard_normal((pb.count, 16)) + 1j * rng.standard_normal((pb.count, 16))
    yf, zf = pb.apply(p, q, gpu.OP_FREE64)
    Yf = pb.apply_batch(P, gpu.OP_FREE64)
    Zf = pb.apply_batch(P, gpu.OP_FREE64, adjoint=True)
    y32, z32 = pb.apply(p, q, gpu.OP_FREE32)
    Y32 = pb.apply_batch(P, gpu.OP_FREE32)
    pb.assemble(gpu.C128_PAIR)
    ys, zs = pb.apply(p, q, gpu.OP_STORED)
    Ys = pb.apply_batch(P, gpu.OP_STORED)
    Zs = pb.apply_batch(P, gpu.OP_STORED, adjoint=True)
    assert np.array_equal(yf, ys) and np.array_equal(zf, zs)
    assert np.array_equal(Yf, Ys) and np.array_equal(Zf, Zs)
    # the mixed-precision factors: FP32 sine and cosine of an FP64 phase
    assert relnorm(y32, ys) <= 1e-7 and relnorm(z32, zs) <= 1e-7
    assert relnorm(Y32, Ys) <= 1e-7


def test_combined_solve_reproduces_the_point_source_field(gpu):
    # k = 2 near the unit sphere's interior Neumann eigenvalue: the combined
    # potential gives the exterior field, as accurately as the double layer's
    n, k, x0 = 48, 2.0, [0.1, 0.2, 0.3]
    rec = shapes.shell_receivers(60, 1.5, 3.0, seed=21)
    exact = shapes.point_source(rec, x0, k)
    pb, _, prm = problem(gpu, n, k, eta=k)
    pb.assemble(gpu.C128_PAIR)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=x0)
    info, _ = pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-10)
    assert info.status == 0
    ure, uim = pb.evaluate(rec, gpu.EVAL_COMBINED)
    err = relnorm(ure + 1j * uim, exact)
    assert err < 0.005, err
    pb0 = gpu.Problem()
    pb0.set_level(shapes.sphere(n))
    pb0.geometry(5)
    pb0.tube(gpu.params(k, prm.eps, prm.tau, dirichlet=1))
    pb0.assemble(gpu.C128_PAIR)
    pb0.rhs(gpu.RHS_POINT_SOURCE, x0=x0)
    info0, _ = pb0.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-10)
    assert info0.status == 0
    u0re, u0im = pb0.evaluate(rec, gpu.EVAL_DOUBLE)
    assert err <= relnorm(u0re + 1j * u0im, exact)
    # the same density through the dense operator (the reference's entries plus
    # the coupling) and the mixed path (complex64 pairs + FP64 replacements)
    x = pb.density()
    x = x[0] + 1j * x[1]
    pb.assemble(gpu.C128)
    pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12)
    xd = pb.density()
    assert relnorm(xd[0] + 1j * xd[1], x) < 1e-9
    pb.assemble(gpu.MIXED_PAIR)
    infom, _ = pb.solve(gpu.SOLVE_RELIABLE, maxit=500, tol=1e-10, inner_tol=1e-3, chunk=1)
    xm = pb.density()
    assert infom.status == 0 and relnorm(xm[0] + 1j * xm[1], x) < 1e-8


def test_combined_batched_solvers_match_single_column(gpu):
    # 16 plane waves on a perturbed surface: batched BiCG on complex128 pairs,
    # the FP64 matrix-free operator (the same bits) and the mixed refinement
    n, k = 32, 6.0
    h = 4.0 / n
    pb = gpu.Problem()
    pb.set_level_shape(gpu.shape(gpu.SHAPE_STAR, a=1.0, b=0.2, m=4, n=3, variant=1), n)
    pb.geometry(5)
    pb.tube(gpu.params(k, 2 * h, 2 * h, dirichlet=1, combined=1, eta=k))
    m = 16
    pb.set_nrhs(m)
    rng = np.random.default_rng(5)
    dirs = rng.standard_normal((m, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for c in range(m):
        pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=dirs[c], col=c)
    pb.assemble(gpu.C128_PAIR)
    info_b, _ = pb.solve(gpu.SOLVE_BICG, maxit=400, tol=1e-9, batch=0)
    xb = [pb.density(c) for c in range(m)]
    assert info_b.status == 0
    pb.assemble(gpu.C128)  # dense: column by column
    info_s, _ = pb.solve(gpu.SOLVE_BICG, maxit=400, tol=1e-9, batch=1)
    assert info_s.status == 0
    for c in range(m):
        xs = pb.density(c)
        assert relnorm(xb[c][0] + 1j * xb[c][1], xs[0] + 1j * xs[1]) < 1e-7
    # nothing stored: the FP64 matrix-free batched solve is the complex128 pairs' bits
    pf = gpu.Problem()
    pf.set_level_shape(gpu.shape(gpu.SHAPE_STAR, a=1.0, b=0.2, m=4, n=3, variant=1), n)
    pf.geometry(5)
    pf.tube(gpu.params(k, 2 * h, 2 * h, dirichlet=1, combined=1, eta=k))
    pf.set_nrhs(m)
    for c in range(m):
        pf.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=dirs[c], col=c)
    info_f, _ = pf.solve(gpu.SOLVE_BICG, maxit=400, tol=1e-9, batch=0)
    assert info_f.iterations == info_b.iterations
    for c in range(m):
        xf = pf.density(c)
        assert np.array_equal(xf[0], xb[c][0]) and np.array_equal(xf[1], xb[c][1])
    # mixed: FP64 refinement around BiCG on the mixed matrix-free operator
    info_r, _ = pf.solve(gpu.SOLVE_REFINED, maxit=400, tol=1e-9, inner_tol=1e-4, max_outer=6, batch=0)
    assert info_r.status == 0
    for c in range(m):
        xr = pf.density(c)
        assert relnorm(xr[0] + 1j * xr[1], xb[c][0] + 1j * xb[c][1]) < 1e-7


def test_combined_needs_the_exterior_dirichlet_problem(gpu):
    pb = gpu.Problem()
    pb.set_level(shapes.sphere(12))
    pb.geometry(5)
    h = 4.0 / 12
    for kw in (dict(dirichlet=0), dict(dirichlet=1, interior=1)):
        with pytest.raises(gpu.IbimError):
            pb.tube(gpu.params(3.0, 1.5 * h, 1.5 * h, combined=1, eta=3.0, **kw))
    with pytest.raises(gpu.IbimError):
        pb.tube(gpu.params(3.0, 1.5 * h, 1.5 * h, combined=1, eta=3.0, dirichlet=1), gpu.TUBE_COMBO)


@pytest.mark.parametrize("combined", [0, 1])
def test_fp32_matrix_core_pass_is_the_complex64_operator_to_fp32_rounding(gpu, combined):
    # ibim_set_fp32_products: the 16-column pass over complex64 pairs on the FP32
    # matrix cores against the same pass with FP64 products (both products,
    # each alone), and the mixed reliable solver on it against the FP64 solution
    n, k = 28, 6.0
    h = 4.0 / n
    pb = gpu.Problem()
    pb.set_level_shape(gpu.shape(gpu.SHAPE_STAR, a=1.0, b=0.2, m=4, n=3, variant=1), n)
    pb.geometry(5)
    pb.tube(gpu.params(k, 2 * h, 2 * h, dirichlet=1, combined=combined, eta=k))
    pb.assemble(gpu.C64_PAIR)
    rng = np.random.default_rng(31 + combined)
    P = rng.standard_normal((pb.count, 16)) + 1j * rng.standard_normal((pb.count, 16))
    Y64 = pb.apply_batch(P, gpu.OP_STORED)
    Z64 = pb.apply_batch(P, gpu.OP_STORED, adjoint=True)
    pb.set_fp32_products(True)
    Y32 = pb.apply_batch(P, gpu.OP_STORED)
    Z32 = pb.apply_batch(P, gpu.OP_STORED, adjoint=True)
    assert relnorm(Y32, Y64) <= 3e-6 and relnorm(Z32, Z64) <= 3e-6, (relnorm(Y32, Y64), relnorm(Z32, Z64))
    assert relnorm(Y32, Y64) > 0  # really a different product
    m = 16
    pb.set_nrhs(m)
    dirs = rng.standard_normal((m, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for c in range(m):
        pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=dirs[c], col=c)
    info32, _ = pb.solve(gpu.SOLVE_RELIABLE, maxit=600, tol=1e-9, inner_tol=1e-3, batch=0)
    x32 = [pb.density(c) for c in range(m)]
    pb.set_fp32_products(False)
    pb.assemble(gpu.C128_PAIR)
    info64, _ = pb.solve(gpu.SOLVE_BICG, maxit=600, tol=1e-11, batch=0)
    assert info32.status == 0 and info64.status == 0 and info32.rel_residual < 1e-9
    for c in range(m):
        x64 = pb.density(c)
        assert relnorm(x32[c][0] + 1j * x32[c][1], x64[0] + 1j * x64[1]) < 1e-7


def test_star_surface_at_the_large_scale_wavenumber(gpu):
    """The large-scale workload's star r = 1 + 0.2 cos4t sin3p at its k h = 0.625
    (k = 40 on 256^3) on 64^3, k = 10, eps = tau = 2h, with an interior point
    source: the GPU's combined rows are the oracle's on the GPU's own tube, the
    solve converges fast and gives the exterior field to ~2 % (the oracle
    pipeline on the same surface, with a closest-point distance from a dense
    parametric sample, gives 2.3 % in 35 BiCG iterations), while the double
    layer alone, near its interior resonances, takes ~200 erratic iterations and
    gives 20 %; the mixed path reproduces the FP64 density."""
    n, k = 64, 10.0
    h = 4.0 / n
    x0 = [0.1, 0.2, 0.05]
    star = gpu.shape(gpu.SHAPE_STAR, a=1.0, b=0.2, m=4, n=3, variant=1)
    rec = shapes.shell_receivers(60, 1.5, 3.0, seed=21)
    exact = shapes.point_source(rec, x0, k)

    def run(combined):
        pb = gpu.Problem()
        pb.set_level_shape(star, n)
        pb.geometry(5)
        prm = gpu.params(k, 2 * h, 2 * h, dirichlet=1, combined=combined, eta=k)
        pb.tube(prm)
        pb.assemble(gpu.C128_PAIR)
        pb.rhs(gpu.RHS_POINT_SOURCE, x0=x0)
        info, _ = pb.solve(gpu.SOLVE_BICG, maxit=1000, tol=1e-10)
        ure, uim = pb.evaluate(rec, gpu.EVAL_COMBINED if combined else gpu.EVAL_DOUBLE)
        return pb, prm, info, relnorm(ure + 1j * uim, exact)

    pb, prm, info, err = run(1)
    node, flat = pb.tube_fields()
    oprm = O.make_params(k=k, eps=prm.eps, tau=prm.tau, dirichlet=1, combined=1, eta=k)
    for r0 in (0, pb.count // 2, pb.count - 64):
        are, aim = O.kernel_rows(flat, pb.count, oprm, r0, 64)
        gre, gim = pb.rows(r0, 64)
        Ao = are + 1j * aim
        assert np.abs(gre + 1j * gim - Ao).max() <= 1e-13 * np.abs(Ao).max(), r0
    assert info.status == 0 and info.iterations <= 60, info.iterations
    assert err < 0.035, err
    pb0, _, info0, err0 = run(0)
    assert info0.iterations > 2 * info.iterations and err0 > 3 * err, (info0.iterations, err0)
    pb0.close()
    x = pb.density()
    x = x[0] + 1j * x[1]
    pb.assemble(gpu.MIXED_PAIR)
    infom, _ = pb.solve(gpu.SOLVE_RELIABLE, maxit=1000, tol=1e-10, inner_tol=1e-3)
    xm = pb.density()
    assert infom.status == 0 and relnorm(xm[0] + 1j * xm[1], x) < 1e-8
    pb.close()
