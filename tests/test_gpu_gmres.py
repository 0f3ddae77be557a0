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
