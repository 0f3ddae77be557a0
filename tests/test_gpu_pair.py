"""Pair storage of the traditional operator (IBIM_C128_PAIR / IBIM_C64_PAIR):
the radial factor of each unordered pair stored once, entries rebuilt per sweep.

Checked against the dense complex128 operator, whose entries are pinned to the
reference (test_gpu_parity): rebuilt rows entry by entry to 1e-14 relative
(complex FP64 pairs: F (r.m_j) with m_j = w_j n_j, a different rounding order
than w_j (c_1 r.n_j)) and to 1e-7 of the largest entry (complex FP32 pairs,
whose point factors use the cheaper x_i.m_j - c_j); sweeps to 1e-13 in norm
against the rebuilt rows; solutions, residual-replacement solves and the misfit
gradient against the dense operator's. The FP64 matrix-free operator, which
sweeps as complex128 pairs with computed factors, against that storage bit for
bit.
"""
import numpy as np
import pytest

import shapes

pytestmark = pytest.mark.gpu


def relnorm(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def problem(gpu, n, k=4.0, dirichlet=1, interior=0):
    pb = gpu.Problem()
    pb.set_level(shapes.sphere(n))
    pb.geometry(5)
    h = 4.0 / n
    # on the coarse grid a 3h tube would reach the sphere's centre; 1.5h keeps
    # the count below one row block (512)
    eps = (3.0 if n >= 16 else 1.5) * h
    pb.tube(gpu.params(k, eps, 1.5 * h if n < 16 else 2 * h, dirichlet=dirichlet, interior=interior))
    return pb


def dense_rows(pb):
    re, im = pb.rows(0, pb.count)
    return re + 1j * im


# n = 12: one row block (count < 512); n = 28: several row blocks and column tiles
@pytest.mark.parametrize("n", [12, 28])
@pytest.mark.parametrize("dirichlet,interior", [(1, 0), (0, 0), (1, 1), (0, 1)])
def test_pair_rows_match_dense_operator(gpu, n, dirichlet, interior):
    pb = problem(gpu, n, dirichlet=dirichlet, interior=interior)
    pb.assemble(gpu.C128)
    A = dense_rows(pb)
    scale = np.abs(A).max()
    pb.assemble(gpu.C128_PAIR)
    Ap = dense_rows(pb)
    nz = np.abs(A) > 0
    assert np.array_equal(nz, np.abs(Ap) > 0)
    assert (np.abs(Ap - A)[nz] / np.abs(A)[nz]).max() <= 1e-14
    pb.assemble(gpu.C64_PAIR)
    A32 = dense_rows(pb)
    assert np.abs(A32 - A).max() <= 1e-7 * scale
    assert relnorm(A32, A) <= 1e-7


@pytest.mark.parametrize("n", [12, 28])
@pytest.mark.parametrize("dirichlet", [1, 0])
@pytest.mark.parametrize("precision", ["C128_PAIR", "C64_PAIR"])
def test_pair_sweep_matches_rebuilt_rows(gpu, n, dirichlet, precision):
    pb = problem(gpu, n, dirichlet=dirichlet)
    pb.assemble(getattr(gpu, precision))
    A = dense_rows(pb)
    rng = np.random.default_rng(n + 10 * dirichlet)
    p = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    q = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    y, z = pb.apply(p, q, gpu.OP_STORED)
    assert relnorm(y, A @ p) <= 1e-13
    assert relnorm(z, A.conj().T @ q) <= 1e-13
    # repeated sweeps are bitwise identical (fixed-order partial planes)
    y2, z2 = pb.apply(p, q, gpu.OP_STORED)
    assert np.array_equal(y, y2) and np.array_equal(z, z2)


def test_pair_bicg_matches_dense_solution(gpu):
    pb = problem(gpu, 28, k=5.0)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    pb.assemble(gpu.C128)
    info_d, hist_d = pb.solve(gpu.SOLVE_BICG, maxit=1000, tol=1e-12)
    xd = pb.density()
    pb.assemble(gpu.C128_PAIR)
    info_p, hist_p = pb.solve(gpu.SOLVE_BICG, maxit=1000, tol=1e-12)
    xp = pb.density()
    assert info_d.status == 0 and info_p.status == 0
    assert abs(info_p.iterations - info_d.iterations) <= 2
    assert relnorm(xp[0] + 1j * xp[1], xd[0] + 1j * xd[1]) <= 1e-10
    np.testing.assert_allclose(hist_p[:10], hist_d[:10], rtol=1e-8)
    pb.solve(gpu.SOLVE_BICG, maxit=1000, tol=1e-12)
    xp2 = pb.density()
    assert np.array_equal(xp[0], xp2[0]) and np.array_equal(xp[1], xp2[1])


@pytest.mark.parametrize("storage", ["C64_PAIR", "MIXED_PAIR"])
def test_pair_mixed_solvers_reach_the_fp64_solution(gpu, storage):
    pb = problem(gpu, 28, k=6.0)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    pb.assemble(gpu.C128)
    info64, _ = pb.solve(gpu.SOLVE_BICG, maxit=2000, tol=1e-12)
    x64 = pb.density()
    x64 = x64[0] + 1j * x64[1]
    # inner: complex64 pairs; replacements / corrections: FP64 matrix-free
    # (C64_PAIR) or the complex128 copy of the same factors (MIXED_PAIR)
    pb.assemble(getattr(gpu, storage))
    pb.reset_timers()
    pb.set_profiling(True)
    info, hist = pb.solve(gpu.SOLVE_RELIABLE, maxit=2000, tol=1e-10, inner_tol=1e-3, chunk=1)
    pb.sync()
    pb.set_profiling(False)
    x = pb.density()
    assert info.status == 0 and info.rel_residual < 1e-10 and info.outer >= 1
    assert info.iterations <= 1.5 * info64.iterations
    assert relnorm(x[0] + 1j * x[1], x64) <= 1e-8
    n64 = pb.kernel_time("matvec64")[1]
    nfree = pb.kernel_time("sweep_free64")[1]
    if storage == "MIXED_PAIR":
        assert n64 == info.outer and nfree == 0
    else:
        assert n64 == 0 and nfree == info.outer
    info, _ = pb.solve(gpu.SOLVE_REFINED, maxit=2000, tol=1e-10, inner_tol=1e-4, max_outer=10)
    x = pb.density()
    assert info.status == 0 and info.rel_residual < 1e-10
    assert relnorm(x[0] + 1j * x[1], x64) <= 1e-8


def test_pair_several_right_hand_sides_solve_column_by_column_or_batched(gpu):
    pb = problem(gpu, 28)
    pb.set_nrhs(3)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3], col=0)
    pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=[0.0, 0.6, 0.8], col=2)  # column 1 stays zero
    pb.assemble(gpu.C128)
    pb.solve(gpu.SOLVE_BICG, maxit=1000, tol=1e-12, batch=1)
    ref = [pb.density(c) for c in range(3)]
    pb.assemble(gpu.C128_PAIR)
    for batch in (1, 0):  # column by column (pair sweeps), then batched (fused 16-column passes)
        info, _ = pb.solve(gpu.SOLVE_BICG, maxit=1000, tol=1e-12, batch=batch)
        assert info.status == 0
        for c in range(3):
            x = pb.density(c)
            if c == 1:
                assert not x[0].any() and not x[1].any()
            else:
                assert relnorm(x[0] + 1j * x[1], ref[c][0] + 1j * ref[c][1]) <= 1e-10, (batch, c)


@pytest.mark.parametrize("dirichlet,interior,mode", [(1, 0, "EVAL_DOUBLE"), (0, 1, "EVAL_SINGLE")])
def test_pair_gradient_matches_dense_gradient(gpu, dirichlet, interior, mode):
    n, k = 24, 3.0
    h = 4.0 / n
    center = (0.0137, -0.0219, 0.0071)
    star = lambda b: gpu.shape(gpu.SHAPE_STAR, a=1.0, b=b, m=3, n=2, variant=0, center=center)
    v = np.random.default_rng(0).standard_normal((40, 3))
    recv = 2.5 * v / np.linalg.norm(v, axis=1, keepdims=True)
    pb = gpu.Problem()
    pb.set_level_shape(star(0.15), n)
    pb.geometry(5)
    pb.tube(gpu.params(k, 3 * h, 2 * h, dirichlet=dirichlet, interior=interior))
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=(0.1, 0.2, 0.3))
    target = np.zeros(len(recv), complex) + 0.01
    out = {}
    for prec in ("C128", "C128_PAIR", "MIXED_PAIR"):
        pb.assemble(getattr(gpu, prec))
        if prec == "MIXED_PAIR":  # forward: complex64 pairs + FP64 replacement; adjoint: the complex128 copy
            info, _ = pb.solve(gpu.SOLVE_RELIABLE, tol=1e-13, maxit=2000, inner_tol=1e-3)
        else:
            info, _ = pb.solve(tol=1e-13, maxit=2000)
        assert info.status == 0
        pb.set_receivers(recv, target)
        out[prec] = pb.gradient(mode=getattr(gpu, mode), tol=1e-13, maxit=2000)
    J0, g0, dk0 = out["C128"]
    for prec in ("C128_PAIR", "MIXED_PAIR"):
        J1, g1, dk1 = out[prec]
        assert abs(J1 - J0) <= 1e-10 * J0, prec
        assert np.abs(g1 - g0).max() <= 1e-8 * np.abs(g0).max(), prec
        assert abs(dk1 - dk0) <= 1e-8 * abs(dk0), prec



@pytest.mark.parametrize("dirichlet,interior", [(1, 0), (0, 0), (1, 1)])
def test_matrix_free_sweep_is_the_pair_storage(gpu, dirichlet, interior):
    """The FP64 matrix-free operator of a traditional tube sweeps as complex128
    pair storage with each factor computed where it would be read (once per
    pair): its products, BiCG and residual-replacement solves (FP32 inner
    sweeps, FP64 replacements) and the adjoint gradient are that storage's
    bits -- the gradient's grid scatter, made with atomics, to 1e-12."""
    n, k = 24, 3.0
    h = 4.0 / n
    center = (0.0137, -0.0219, 0.0071)
    star = gpu.shape(gpu.SHAPE_STAR, a=1.0, b=0.15, m=3, n=2, variant=0, center=center)
    v = np.random.default_rng(0).standard_normal((40, 3))
    recv = 2.5 * v / np.linalg.norm(v, axis=1, keepdims=True)
    target = np.zeros(len(recv), complex) + 0.01
    out = []
    for stored in (False, True):
        pb = gpu.Problem()
        pb.set_level_shape(star, n)
        pb.geometry(5)
        pb.tube(gpu.params(k, 3 * h, 2 * h, dirichlet=dirichlet, interior=interior))
        pb.rhs(gpu.RHS_POINT_SOURCE, x0=(0.1, 0.2, 0.3))
        if stored:
            pb.assemble(gpu.C128_PAIR)
        rng = np.random.default_rng(4)
        p = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
        q = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
        y, z = pb.apply(p, q, gpu.OP_STORED if stored else gpu.OP_FREE64)
        rinfo, rhist = pb.solve(gpu.SOLVE_RELIABLE, tol=1e-11, maxit=2000, inner_tol=1e-3)
        xr = pb.density()
        info, hist = pb.solve(tol=1e-13, maxit=2000)
        x = pb.density()
        assert info.status == 0 and rinfo.status == 0
        pb.set_receivers(recv, target)
        J, g, dk = pb.gradient(tol=1e-13, maxit=2000)
        out.append(dict(y=y, z=z, its=(info.iterations, rinfo.iterations, rinfo.outer), hist=np.asarray(hist),
                        rhist=np.asarray(rhist), x=x, xr=xr, J=J, g=g, dk=dk))
        pb.close()
    free, pair = out
    assert np.array_equal(free["y"], pair["y"]) and np.array_equal(free["z"], pair["z"])
    assert free["its"] == pair["its"]
    assert np.array_equal(free["hist"], pair["hist"]) and np.array_equal(free["rhist"], pair["rhist"])
    for key in ("x", "xr"):
        assert np.array_equal(free[key][0], pair[key][0]) and np.array_equal(free[key][1], pair[key][1]), key
    assert free["J"] == pair["J"] and free["dk"] == pair["dk"]
    assert np.abs(free["g"] - pair["g"]).max() <= 1e-12 * np.abs(pair["g"]).max()


@pytest.mark.parametrize("tau_h", [2.0, 0.75])
@pytest.mark.parametrize("dirichlet,interior", [(1, 0), (0, 0), (1, 1), (0, 1)])
def test_mixed_matrix_free_factors_follow_the_fp64_operator(gpu, tau_h, dirichlet, interior):
    """The mixed-precision matrix-free factors take the near/far decision from
    rho^2 away from tau^2 and the FP64 rule next to it: for a wide and a narrow
    near field, Dirichlet and Neumann, exterior and interior, the operator stays
    within its FP32 sine and cosine of the FP64 one, with no pair lost or
    counted twice, single-vector and 16 columns at a time."""
    n = 24
    h = 4.0 / n
    pb = gpu.Problem()
    pb.set_level(shapes.sphere(n))
    pb.geometry(5)
    pb.tube(gpu.params(4.0, 3 * h, tau_h * h, dirichlet=dirichlet, interior=interior))
    rng = np.random.default_rng(9)
    p = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    q = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
    y64, z64 = pb.apply(p, q, gpu.OP_FREE64)
    y32, z32 = pb.apply(p, q, gpu.OP_FREE32)
    assert relnorm(y32, y64) <= 1e-7 and relnorm(z32, z64) <= 1e-7, (relnorm(y32, y64), relnorm(z32, z64))
    # the batched passes of the same operators: FP64 against FP64 to rounding,
    # mixed against its own single-vector sweep to rounding
    P = np.stack([p, q] + [rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count) for _ in range(14)], 1)
    # column 0 of A P is A p, column 1 of A^H P is A^H q
    for adjoint, col, single64, single32 in ((False, 0, y64, y32), (True, 1, z64, z32)):
        Y64 = pb.apply_batch(P, gpu.OP_FREE64, adjoint=adjoint)
        Y32 = pb.apply_batch(P, gpu.OP_FREE32, adjoint=adjoint)
        assert relnorm(Y32, Y64) <= 1e-7
        assert relnorm(Y64[:, col], single64) <= 1e-13
        assert relnorm(Y32[:, col], single32) <= 1e-13
    pb.close()

def test_pair_storage_rejects_what_it_cannot_do(gpu):
    pb = gpu.Problem().set_level(shapes.sphere(16)).geometry(5)
    h = 4.0 / 16
    pb.tube(gpu.params(2.0, 3 * h, 2 * h, dirichlet=1), gpu.TUBE_COMBO)
    with pytest.raises(gpu.IbimError):
        pb.assemble(gpu.C128_PAIR)  # the combined-field kernel has no pair form
    pb2 = problem(gpu, 16)
    with pytest.raises(gpu.IbimError):
        pb2.assemble(7)
