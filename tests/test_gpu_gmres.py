"""IBIM_SOLVE_GMRES: FP64 iterative refinement around GMRES on the stored
operator (complex64 storage: the mixed-precision path; complex128: FP64 GMRES),
against the FP64 solution of the reference's BiCG on the dense complex128
operator: every storage, restarted and unrestarted, several right-hand sides,
repeatability and the zero right-hand side."""
import numpy as np
import pytest

import shapes

pytestmark = pytest.mark.gpu


def problem(gpu, n=28, k=4.0, dirichlet=1, combined=0):
    pb = gpu.Problem()
    pb.set_level(shapes.sphere(n))
    pb.geometry(5)
    h = 4.0 / n
    pb.tube(gpu.params(k, 3 * h, 2 * h, dirichlet=dirichlet, combined=combined, eta=k))
    return pb


def set_columns(gpu, pb, n):
    pb.set_nrhs(n)
    rng = np.random.default_rng(3)
    for c in range(n):
        if c % 2:
            d = rng.standard_normal(3)
            pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=d / np.linalg.norm(d), col=c)
        else:
            pb.rhs(gpu.RHS_POINT_SOURCE, x0=0.3 * rng.standard_normal(3), col=c)


def densities(pb, n):
    return [(lambda d: d[0] + 1j * d[1])(pb.density(c)) for c in range(n)]


def relnorm(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("store", ["C128_PAIR", "MIXED_PAIR", "C64_PAIR", "C128", "C64"])
@pytest.mark.parametrize("dirichlet,combined", [(1, 0), (0, 0), (1, 1)])
def test_gmres_reaches_the_fp64_solution(gpu, store, dirichlet, combined):
    pb = problem(gpu, dirichlet=dirichlet, combined=combined)
    set_columns(gpu, pb, 3)
    pb.assemble(gpu.C128)
    ref_info, _ = pb.solve(gpu.SOLVE_BICG, maxit=2000, tol=1e-13, batch=1)
    assert ref_info.status == 0
    ref = densities(pb, 3)
    pb.assemble(getattr(gpu, store))
    info, hist = pb.solve(gpu.SOLVE_GMRES, maxit=2000, tol=1e-10)
    assert info.status == 0 and info.rel_residual < 1e-10, (info.status, info.rel_residual)
    got = densities(pb, 3)
    for c in range(3):
        assert relnorm(got[c], ref[c]) < 1e-8, (c, relnorm(got[c], ref[c]))
    pb.close()


def test_gmres_restarted_and_repeatable(gpu):
    """A restart every 8 products reaches the same solution; two solves give the
    same bits; the unrestarted solve applies the operator less often than BiCG
    on the same stored operator (one product per GMRES iteration, A and A^H per
    BiCG iteration)."""
    pb = problem(gpu, k=6.0)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    pb.assemble(gpu.MIXED_PAIR)
    full, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-10)
    x_full = densities(pb, 1)[0]
    again, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-10)
    x_again = densities(pb, 1)[0]
    assert (again.iterations, again.outer) == (full.iterations, full.outer)
    assert np.array_equal(x_full, x_again)
    short, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-10, chunk=8)
    x_short = densities(pb, 1)[0]
    assert short.status == 0 and short.iterations > full.iterations
    assert relnorm(x_short, x_full) < 1e-8
    bicg, _ = pb.solve(gpu.SOLVE_BICG, maxit=3000, tol=1e-10)
    assert full.iterations <= 2 * bicg.iterations, (full.iterations, bicg.iterations)
    pb.close()


def test_gmres_zero_right_hand_side_and_budget(gpu):
    pb = problem(gpu)
    pb.set_nrhs(2)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3], col=0)
    pb.set_rhs(np.zeros(pb.count), np.zeros(pb.count), col=1)
    pb.assemble(gpu.C64_PAIR)
    info, _ = pb.solve(gpu.SOLVE_GMRES, maxit=2000, tol=1e-10)
    assert not densities(pb, 2)[1].any()
    capped, _ = pb.solve(gpu.SOLVE_GMRES, maxit=5, tol=1e-12)
    assert capped.status == 1 and capped.iterations <= 2 * 5
    pb.close()


@pytest.mark.parametrize("store", ["C128_PAIR", "C64_PAIR", "MIXED_PAIR", "C128"])
@pytest.mark.parametrize("dirichlet,combined", [(1, 0), (0, 0), (1, 1)])
def test_product_without_the_adjoint_is_the_same_bits(gpu, store, dirichlet, combined):
    """ibim_apply with no A^H outputs (the sweep GMRES applies): y = A p is the
    full sweep's y, bit for bit."""
    pb = problem(gpu, dirichlet=dirichlet, combined=combined)
    pb.assemble(getattr(gpu, store))
    rng = np.random.default_rng(11)
    p = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    q = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    y, _ = pb.apply(p, q, gpu.OP_STORED)
    assert np.array_equal(pb.apply_y(p, gpu.OP_STORED), y)
    pb.close()


def test_gmres_batched_columns_match_single_columns(gpu):
    """20 right-hand sides (a batch of 16 and a ragged one of 4, one column
    zero) in 16-column passes against the same refinement one column at a time:
    the same solutions to the tolerance, zero data gives a zero density and no
    iterations, and two batched solves give the same bits."""
    pb = problem(gpu, k=6.0)
    m = 20
    set_columns(gpu, pb, m)
    pb.set_rhs(np.zeros(pb.count), np.zeros(pb.count), col=5)
    pb.assemble(gpu.MIXED_PAIR)
    single, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-10, batch=1)
    xs = densities(pb, m)
    batched, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-10)
    xb = densities(pb, m)
    print("GMRES products: one column at a time", single.iterations, "16-column passes", batched.iterations)
    assert single.status == 0 and batched.status == 0, (single.status, batched.status)
    assert batched.rel_residual < 1e-10
    assert not xb[5].any()
    for c in range(m):
        if c != 5:
            assert relnorm(xb[c], xs[c]) < 1e-8, (c, relnorm(xb[c], xs[c]))
    # products counted per column: the batched passes add only rounding differences
    assert batched.iterations <= 1.2 * single.iterations + 2 * m, (batched.iterations, single.iterations)
    again, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-10)
    assert (again.iterations, again.outer) == (batched.iterations, batched.outer)
    for c, x in enumerate(densities(pb, m)):
        assert np.array_equal(x, xb[c]), c
    pb.close()


@pytest.mark.parametrize("combined", [0, 1])
def test_gmres_batched_on_fp32_matrix_cores(gpu, combined):
    """16 plane waves on complex64 pairs with the products on the FP32 matrix
    cores (ibim_set_fp32_products; the large-scale workload's mixed path): the
    refinement's FP64 residuals bring every column to the FP64 solution, and
    the FP32 rounding of the products, which depends on the vector, costs GMRES
    few products against the FP64-product pass over the same pairs."""
    n, k = 28, 6.0
    h = 4.0 / n
    pb = gpu.Problem()
    pb.set_level_shape(gpu.shape(gpu.SHAPE_STAR, a=1.0, b=0.2, m=4, n=3, variant=1), n)
    pb.geometry(5)
    pb.tube(gpu.params(k, 2 * h, 2 * h, dirichlet=1, combined=combined, eta=k))
    m = 16
    pb.set_nrhs(m)
    rng = np.random.default_rng(41 + combined)
    dirs = rng.standard_normal((m, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for c in range(m):
        pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=dirs[c], col=c)
    pb.assemble(gpu.C128_PAIR)
    ref_info, _ = pb.solve(gpu.SOLVE_BICG, maxit=3000, tol=1e-12, batch=0)
    assert ref_info.status == 0
    ref = densities(pb, m)
    pb.assemble(gpu.C64_PAIR)
    info64, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-9)
    pb.set_fp32_products(True)
    info32, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-9)
    x32 = densities(pb, m)
    pb.set_fp32_products(False)
    print("GMRES products (summed over columns): FP64 products", info64.iterations, info64.outer,
          "FP32 products", info32.iterations, info32.outer)
    assert info64.status == 0 and info32.status == 0 and info32.rel_residual < 1e-9
    for c in range(m):
        assert relnorm(x32[c], ref[c]) < 1e-7, (c, relnorm(x32[c], ref[c]))
    assert info32.iterations <= 1.5 * info64.iterations, (info32.iterations, info64.iterations)
    pb.close()


@pytest.mark.parametrize("store", ["C128", "C64"])
def test_gmres_on_the_combo_tube(gpu, store):
    """The Neumann combo kernel's tube (no pair form, no 16-column FP64
    matrix-free pass): dense storage, two right-hand sides, against FP64 BiCG."""
    pb = gpu.Problem().set_level(shapes.sphere(24)).geometry(5)
    h = 4.0 / 24
    pb.tube(gpu.params(2.0, 3 * h, 2 * h, delta_in=0.3 * h, dirichlet=1), gpu.TUBE_COMBO)
    set_columns(gpu, pb, 2)
    pb.assemble(gpu.C128)
    ref_info, _ = pb.solve(gpu.SOLVE_BICG, maxit=2000, tol=1e-13, batch=1)
    assert ref_info.status == 0
    ref = densities(pb, 2)
    pb.assemble(getattr(gpu, store))
    info, _ = pb.solve(gpu.SOLVE_GMRES, maxit=2000, tol=1e-10)
    assert info.status == 0 and info.rel_residual < 1e-10, (info.status, info.rel_residual)
    for c, x in enumerate(densities(pb, 2)):
        assert relnorm(x, ref[c]) < 1e-8, (c, relnorm(x, ref[c]))
    pb.close()


@pytest.mark.parametrize("nrhs", [1, 3])
def test_gmres_matrix_free(gpu, nrhs):
    """Nothing stored: GMRES on the mixed matrix-free operator, the FP64
    residuals from the FP64 matrix-free one (16-column passes for several
    right-hand sides), against FP64 BiCG on complex128 pairs."""
    ref_pb = problem(gpu, k=6.0)
    set_columns(gpu, ref_pb, nrhs)
    ref_pb.assemble(gpu.C128_PAIR)
    ref_info, _ = ref_pb.solve(gpu.SOLVE_BICG, maxit=2000, tol=1e-13, batch=1)
    assert ref_info.status == 0
    ref = densities(ref_pb, nrhs)
    ref_pb.close()
    pb = problem(gpu, k=6.0)
    set_columns(gpu, pb, nrhs)
    info, _ = pb.solve(gpu.SOLVE_GMRES, maxit=2000, tol=1e-10)
    assert info.status == 0 and info.rel_residual < 1e-10, (info.status, info.rel_residual)
    for c, x in enumerate(densities(pb, nrhs)):
        assert relnorm(x, ref[c]) < 1e-8, (c, relnorm(x, ref[c]))
    pb.close()


@pytest.mark.parametrize("op", ["OP_FREE64", "OP_FREE32"])
@pytest.mark.parametrize("dirichlet,combined", [(1, 0), (0, 0), (1, 1)])
def test_matrix_free_product_without_the_adjoint_is_the_same_bits(gpu, op, dirichlet, combined):
    """The matrix-free pair sweeps without their A^H half (the products GMRES
    applies with nothing stored) give the full sweep's y = A p bit for bit."""
    pb = problem(gpu, dirichlet=dirichlet, combined=combined)
    rng = np.random.default_rng(13)
    p = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    q = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    y, _ = pb.apply(p, q, getattr(gpu, op))
    assert np.array_equal(pb.apply_y(p, getattr(gpu, op)), y)
    pb.close()


@pytest.mark.parametrize("store", ["C128_PAIR", "MIXED_PAIR", "C64_PAIR", "C128", "C64"])
@pytest.mark.parametrize("dirichlet,combined", [(1, 0), (0, 0), (1, 1)])
def test_block_gmres_reaches_the_fp64_solution(gpu, store, dirichlet, combined):
    """IBIM_SOLVE_BLOCK_GMRES, one Krylov space per batch of 16 columns: 20
    right-hand sides (a batch of 16 and a ragged one of 4 with a zero column)
    reach the FP64 solution on every storage and kernel; two solves give the
    same bits."""
    pb = problem(gpu, dirichlet=dirichlet, combined=combined)
    m = 20
    set_columns(gpu, pb, m)
    pb.set_rhs(np.zeros(pb.count), np.zeros(pb.count), col=17)
    pb.assemble(gpu.C128)
    ref_info, _ = pb.solve(gpu.SOLVE_BICG, maxit=2000, tol=1e-13, batch=0)
    assert ref_info.status == 0
    ref = densities(pb, m)
    pb.assemble(getattr(gpu, store))
    info, _ = pb.solve(gpu.SOLVE_BLOCK_GMRES, maxit=2000, tol=1e-10)
    assert info.status == 0 and info.rel_residual < 1e-10, (info.status, info.rel_residual)
    got = densities(pb, m)
    assert not got[17].any()
    for c in range(m):
        if c != 17:
            assert relnorm(got[c], ref[c]) < 1e-8, (c, relnorm(got[c], ref[c]))
    again, _ = pb.solve(gpu.SOLVE_BLOCK_GMRES, maxit=2000, tol=1e-10)
    assert (again.iterations, again.outer) == (info.iterations, info.outer)
    for c, x in enumerate(densities(pb, m)):
        assert np.array_equal(x, got[c]), c
    pb.close()


def test_block_gmres_restarted_and_against_column_gmres(gpu):
    """16 plane waves on the double layer near the sphere's first interior
    resonances (k = 4.5): the block space needs fewer passes than GMRES column
    by column needs products per column; a restart every 3 block passes still
    reaches the same solution; FP32 matrix-core products too."""
    pb = problem(gpu, k=4.5)
    m = 16
    pb.set_nrhs(m)
    rng = np.random.default_rng(17)
    dirs = rng.standard_normal((m, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for c in range(m):
        pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=dirs[c], col=c)
    pb.assemble(gpu.C64_PAIR)
    col, _ = pb.solve(gpu.SOLVE_GMRES, maxit=3000, tol=1e-10)
    x_col = densities(pb, m)
    pb.reset_timers()
    pb.set_profiling(True)
    blk, _ = pb.solve(gpu.SOLVE_BLOCK_GMRES, maxit=3000, tol=1e-10)
    passes = pb.kernel_time("batch_pair")[1]
    pb.set_profiling(False)
    x_blk = densities(pb, m)
    short, _ = pb.solve(gpu.SOLVE_BLOCK_GMRES, maxit=3000, tol=1e-10, chunk=3)
    x_short = densities(pb, m)
    pb.set_fp32_products(True)
    b32, _ = pb.solve(gpu.SOLVE_BLOCK_GMRES, maxit=3000, tol=1e-9)
    x32 = densities(pb, m)
    pb.set_fp32_products(False)
    print("products per column: GMRES %.1f; block GMRES passes %d (%d column products), restarted %d, FP32 %d"
          % (col.iterations / m, passes, blk.iterations, short.iterations, b32.iterations))
    for info in (col, blk, short, b32):
        assert info.status == 0, info.status
    assert passes < col.iterations / m, (passes, col.iterations / m)
    for c in range(m):
        assert relnorm(x_blk[c], x_col[c]) < 1e-8, c
        assert relnorm(x_short[c], x_col[c]) < 1e-8, c
        assert relnorm(x32[c], x_col[c]) < 1e-7, c
    pb.close()


@pytest.mark.parametrize("store", ["C128_PAIR", "MIXED_PAIR", "C64_PAIR", "C128", "C64", "FREE"])
@pytest.mark.parametrize("dirichlet,combined", [(1, 0), (0, 0), (1, 1)])
def test_enlarged_gmres_reaches_the_fp64_solution(gpu, store, dirichlet, combined):
    """IBIM_SOLVE_BLOCK_GMRES with one right-hand side at a time (enlarged
    GMRES: the residual split over 16 pieces of the surface, one block space,
    the one residual minimized over it): two right-hand sides solved column by
    column reach the FP64 solution on every storage, kernel and with nothing
    stored; two solves give the same bits; fewer passes than GMRES products."""
    ref_pb = problem(gpu, dirichlet=dirichlet, combined=combined)
    set_columns(gpu, ref_pb, 2)
    ref_pb.assemble(gpu.C128)
    ref_info, _ = ref_pb.solve(gpu.SOLVE_BICG, maxit=2000, tol=1e-13, batch=1)
    assert ref_info.status == 0
    ref = densities(ref_pb, 2)
    ref_pb.close()
    pb = problem(gpu, dirichlet=dirichlet, combined=combined)
    set_columns(gpu, pb, 2)
    if store != "FREE":
        pb.assemble(getattr(gpu, store))
    info, _ = pb.solve(gpu.SOLVE_BLOCK_GMRES, maxit=2000, tol=1e-10, batch=1)
    assert info.status == 0 and info.rel_residual < 1e-10, (info.status, info.rel_residual)
    got = densities(pb, 2)
    for c in range(2):
        assert relnorm(got[c], ref[c]) < 1e-8, (c, relnorm(got[c], ref[c]))
    again, _ = pb.solve(gpu.SOLVE_BLOCK_GMRES, maxit=2000, tol=1e-10, batch=1)
    assert (again.iterations, again.outer) == (info.iterations, info.outer)
    for c, x in enumerate(densities(pb, 2)):
        assert np.array_equal(x, got[c]), c
    col, _ = pb.solve(gpu.SOLVE_GMRES, maxit=2000, tol=1e-10, batch=1)
    assert info.iterations < col.iterations, (info.iterations, col.iterations)
    pb.close()


def test_enlarged_gmres_with_fp32_products_and_a_zero_column(gpu):
    """Enlarged GMRES on complex64 pairs with the 16-column passes on the FP32
    matrix cores: the FP64 refinement brings the one right-hand side to the
    FP64 solution; a zero right-hand side gives a zero density."""
    pb = problem(gpu, k=4.5)
    pb.set_nrhs(2)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3], col=0)
    pb.set_rhs(np.zeros(pb.count), np.zeros(pb.count), col=1)
    pb.assemble(gpu.C128_PAIR)
    ref_info, _ = pb.solve(gpu.SOLVE_BICG, maxit=3000, tol=1e-13, batch=1)
    ref = densities(pb, 1)[0]
    pb.assemble(gpu.MIXED_PAIR)
    pb.set_fp32_products(True)
    info, _ = pb.solve(gpu.SOLVE_BLOCK_GMRES, maxit=3000, tol=1e-10, batch=1)
    pb.set_fp32_products(False)
    x = densities(pb, 2)
    # the reference BiCG reports the zero column's status 4 (zero data) beside column 0's residual
    assert ref_info.status in (0, 4) and ref_info.rel_residual < 1e-12, (ref_info.status, ref_info.rel_residual)
    assert info.status == 0, info.status
    assert relnorm(x[0], ref) < 1e-8, relnorm(x[0], ref)
    assert not x[1].any()
    pb.close()
