"""The benchmark's own size: torus R = 1, r = 0.4 on 128^3, k = 10, eps = tau = 2h
(64,676 tube points). The oracle is too slow here, so these tests check
properties that hold at any size, between kernels that share no code path:

* rows of the pair storage (complex128 and complex64 factors) against the dense
  complex128 operator, whose entries are pinned to the reference at small sizes
  (to 1e-14 and 1e-7 of the row's largest entry);
* <A p, q> = <p, A^H q> for the fused sweeps of every storage, and the 16-column
  fused pair pass against the single-vector sweep, column by column;
* the FP64 matrix-free sweep and 16-column pass (factors computed, tables built
  without an assembly) against the complex128 pair storage's, bit for bit, and the
  mixed-precision matrix-free ones against them (1e-7) and each other (1e-13);
* the mixed-precision solve's density against the dense FP64 operator: its true
  residual ||b - A x|| / ||b|| is below the requested tolerance.
"""
import numpy as np
import pytest

import shapes

pytestmark = pytest.mark.gpu

N, K = 128, 10.0


def make_torus(gpu):
    pb = gpu.Problem()
    pb.set_level(shapes.torus(N))
    pb.geometry(5)
    h = 4.0 / N
    count = pb.tube(gpu.params(K, 2 * h, 2 * h, dirichlet=1))
    assert count == 64676
    return pb


@pytest.fixture(scope="module")
def torus(gpu):
    pb = make_torus(gpu)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[1.0, 0.0, 0.05])
    yield pb
    pb.close()


def relnorm(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def test_pair_rows_match_the_dense_operator_at_full_size(gpu, torus):
    pb = torus
    rows = [0, 511, 512, 31337, 64675]
    pb.assemble(gpu.C128)
    dense = [pb.rows(i, 1) for i in rows]
    pb.assemble(gpu.C128_PAIR)
    pair = [pb.rows(i, 1) for i in rows]
    pb.assemble(gpu.MIXED_PAIR)  # complex64 pairs (the rows the mixed sweeps apply)
    pair64 = [pb.rows(i, 1) for i in rows]
    for i, d, p, q in zip(rows, dense, pair, pair64):
        a = d[0] + 1j * d[1]
        b = p[0] + 1j * p[1]
        c = q[0] + 1j * q[1]
        nz = np.abs(a) > 0
        assert np.array_equal(nz, np.abs(b) > 0), i
        # w_j (r.n_j) in the dense kernel, r.(w_j n_j) in the pairs: the same to
        # FP64 rounding of |r| |n|, which is large relative only to the nearly
        # tangential pairs' tiny entries -- so against the row's largest entry
        scale = np.abs(a).max()
        assert np.abs(b - a).max() <= 1e-14 * scale, i
        assert relnorm(b, a) <= 1e-14, i
        assert np.abs(c - a).max() <= 1e-7 * scale, i


@pytest.mark.parametrize("store", ["C128", "C128_PAIR", "MIXED_PAIR"])
def test_sweeps_are_adjoint_at_full_size(gpu, torus, store):
    pb = torus
    pb.assemble(getattr(gpu, store))
    rng = np.random.default_rng(5)
    n = pb.count
    p = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    q = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y, z = pb.apply(p, q, gpu.OP_STORED)  # A p and A^H q from one pass
    lhs = np.vdot(q, y)  # <A p, q>
    rhs = np.vdot(z, p)  # <p, A^H q>
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(y) * np.linalg.norm(q)


def test_fused_batched_pass_matches_single_sweeps_at_full_size(gpu, torus):
    pb = torus
    pb.assemble(gpu.MIXED_PAIR)
    rng = np.random.default_rng(6)
    n = pb.count
    P = rng.standard_normal((n, 16)) + 1j * rng.standard_normal((n, 16))
    Y = pb.apply_batch(P, gpu.OP_STORED, False)
    Z = pb.apply_batch(P, gpu.OP_STORED, True)
    for c in (0, 7, 15):
        y, z = pb.apply(P[:, c], P[:, c], gpu.OP_STORED)
        assert relnorm(Y[:, c], y) <= 1e-13, c
        assert relnorm(Z[:, c], z) <= 1e-13, c


def test_matrix_free_passes_are_the_pair_storage_at_full_size(gpu, torus):
    free = make_torus(gpu)  # never assembled: its own point table and near lists
    torus.assemble(gpu.C128_PAIR)
    rng = np.random.default_rng(8)
    P = rng.standard_normal((free.count, 16)) + 1j * rng.standard_normal((free.count, 16))
    for adjoint in (False, True):
        got = free.apply_batch(P, gpu.OP_FREE64, adjoint)
        assert np.array_equal(got, torus.apply_batch(P, gpu.OP_STORED, adjoint)), adjoint
    # the single-vector sweeps: FP64 bitwise; the mixed-precision factors within
    # their FP32 sine and cosine of the FP64 ones, and their batched pass equal
    # to their sweep up to summation order
    p, q = P[:, 3], P[:, 11]
    y, z = free.apply(p, q, gpu.OP_FREE64)
    ys, zs = torus.apply(p, q, gpu.OP_STORED)
    assert np.array_equal(y, ys) and np.array_equal(z, zs)
    y32, z32 = free.apply(p, q, gpu.OP_FREE32)
    assert relnorm(y32, y) <= 1e-7 and relnorm(z32, z) <= 1e-7
    Y32 = free.apply_batch(P, gpu.OP_FREE32, False)
    Z32 = free.apply_batch(P, gpu.OP_FREE32, True)
    assert relnorm(Y32[:, 3], y32) <= 1e-13 and relnorm(Z32[:, 11], z32) <= 1e-13
    free.close()


def test_mixed_solve_meets_its_tolerance_on_the_fp64_operator(gpu, torus):
    pb = torus
    pb.assemble(gpu.MIXED_PAIR)
    tol = 1e-8
    info, _ = pb.solve(gpu.SOLVE_RELIABLE, maxit=3000, tol=tol, inner_tol=1e-2, chunk=1)
    assert info.status == 0 and info.outer >= 1
    x = pb.density()
    x = x[0] + 1j * x[1]
    bre, bim = pb.get_rhs()
    b = bre + 1j * bim
    pb.assemble(gpu.C128)  # the dense FP64 operator, an independent kernel
    ax, _ = pb.apply(x, x, gpu.OP_STORED)
    assert np.linalg.norm(b - ax) / np.linalg.norm(b) <= 1.5 * tol
