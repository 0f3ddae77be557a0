"""HIP graph replay of BiCG iteration chunks (ibim_set_graphs): every solve
captures one chunk of iterations and replays it for each full chunk (batched
solves: after a first eager chunk). Checked
bitwise against the same solves launched one kernel at a time, over several
solves in a row with other device work between them, for plain BiCG on dense,
pair and mixed pair storage and the matrix-free pair sweeps, the
residual-replacement and refinement solvers
(their replacements, restarts and corrections run between replays) and the
adjoint solve of the misfit gradient (J and dJ/dk bitwise; the level-set
gradient, scattered with atomics, to 1e-12)."""
import numpy as np
import pytest

import shapes

pytestmark = pytest.mark.gpu


def problem(gpu, n=28, k=5.0):
    pb = gpu.Problem()
    pb.set_level(shapes.sphere(n))
    pb.geometry(5)
    h = 4.0 / n
    pb.tube(gpu.params(k, 3 * h, 2 * h, dirichlet=1))
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    return pb


def run(gpu, pb, method, chunk):
    if method == "bicg":
        info, hist = pb.solve(gpu.SOLVE_BICG, maxit=1000, tol=1e-12, chunk=chunk)
    elif method == "reliable":
        info, hist = pb.solve(gpu.SOLVE_RELIABLE, maxit=1000, tol=1e-11, inner_tol=1e-2, chunk=chunk)
    else:
        info, hist = pb.solve(gpu.SOLVE_REFINED, maxit=1000, tol=1e-11, inner_tol=1e-4, chunk=chunk)
    x = pb.density()
    return info.iterations, info.outer, info.status, np.asarray(hist), x[0], x[1]


@pytest.mark.parametrize("store,method", [("C128", "bicg"), ("C128_PAIR", "bicg"), ("MIXED_PAIR", "bicg"),
                                          ("MIXED_PAIR", "reliable"), ("C64_PAIR", "refined"),
                                          ("matrix-free", "bicg"), ("matrix-free", "reliable")])
@pytest.mark.parametrize("chunk", [1, 3])
def test_graph_replay_gives_the_eager_bits(gpu, store, method, chunk):
    pb = problem(gpu)
    if store != "matrix-free":  # else the pair sweeps computing their factors
        pb.assemble(getattr(gpu, store))
    eager = run(gpu, pb, method, chunk)
    assert eager[2] == 0
    rng = np.random.default_rng(chunk)
    pb.set_graphs(True)
    for rep in range(3):
        got = run(gpu, pb, method, chunk)
        assert got[:3] == eager[:3], (rep, got[:3], eager[:3])
        assert np.array_equal(got[3], eager[3]), rep
        assert np.array_equal(got[4], eager[4]) and np.array_equal(got[5], eager[5]), rep
        # other device work between the solves
        p = rng.standard_normal(pb.count) + 1j * rng.standard_normal(pb.count)
        pb.apply(p, p.conj(), gpu.OP_FREE64 if store == "matrix-free" else gpu.OP_STORED)
    pb.set_graphs(False)
    again = run(gpu, pb, method, chunk)
    assert np.array_equal(again[4], eager[4]) and np.array_equal(again[5], eager[5])


def test_graph_replay_adjoint_gradient(gpu):
    n, k = 24, 3.0
    h = 4.0 / n
    star = gpu.shape(gpu.SHAPE_STAR, a=1.0, b=0.15, m=3, n=2, variant=0, center=(0.0137, -0.0219, 0.0071))
    v = np.random.default_rng(0).standard_normal((40, 3))
    recv = 2.5 * v / np.linalg.norm(v, axis=1, keepdims=True)
    pb = gpu.Problem()
    pb.set_level_shape(star, n)
    pb.geometry(5)
    pb.tube(gpu.params(k, 3 * h, 2 * h, dirichlet=1))
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=(0.1, 0.2, 0.3))
    pb.assemble(gpu.C128_PAIR)
    out = []
    for graphs in (False, True, True):
        pb.set_graphs(graphs)
        info, _ = pb.solve(tol=1e-13, maxit=2000, chunk=4)
        assert info.status == 0
        pb.set_receivers(recv, np.zeros(len(recv), complex) + 0.01)
        out.append(pb.gradient(tol=1e-13, maxit=2000, chunk=4))
    # the forward and adjoint solves give the same bits (J and dJ/dk depend on
    # nothing else); the level-set gradient is scattered onto the grid with
    # atomics, whose order varies from run to run with or without graphs
    for J, g, dk in out[1:]:
        assert J == out[0][0] and dk == out[0][2]
        assert np.abs(g - out[0][1]).max() <= 1e-12 * np.abs(out[0][1]).max()


def fill_rhs(gpu, pb, n):
    pb.set_nrhs(n)
    rng = np.random.default_rng(7)
    for c in range(n):
        d = rng.standard_normal(3)
        pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=d / np.linalg.norm(d), col=c)


@pytest.mark.parametrize("store,method", [("C128", "bicg"), ("C128_PAIR", "bicg"), ("MIXED_PAIR", "reliable"),
                                          ("C64", "refined"), ("matrix-free", "bicg")])
@pytest.mark.parametrize("chunk", [2, 5])
def test_batched_graph_replay_gives_the_eager_bits(gpu, store, method, chunk):
    """16 right-hand sides per pass (19 columns: a batch of 16 and one of 3):
    the first chunk runs eagerly, later full chunks replay one captured graph
    (matrix-free: the pair pass computing its factors)."""
    pb = problem(gpu)
    if store != "matrix-free":
        pb.assemble(getattr(gpu, store))
    fill_rhs(gpu, pb, 19)

    def run_batched():
        if method == "bicg":
            info, _ = pb.solve(gpu.SOLVE_BICG, maxit=500, tol=1e-12, chunk=chunk, batch=0)
        elif method == "reliable":
            info, _ = pb.solve(gpu.SOLVE_RELIABLE, maxit=500, tol=1e-11, inner_tol=1e-2, chunk=chunk, batch=0)
        else:
            info, _ = pb.solve(gpu.SOLVE_REFINED, maxit=500, tol=1e-11, inner_tol=1e-4, chunk=chunk, batch=0)
        return info.iterations, info.outer, info.status, [pb.density(c) for c in range(19)]

    eager = run_batched()
    assert eager[2] == 0
    pb.set_graphs(True)
    for rep in range(2):
        got = run_batched()
        assert got[:3] == eager[:3], (rep, got[:3], eager[:3])
        for c in range(19):
            assert np.array_equal(got[3][c][0], eager[3][c][0]) and np.array_equal(got[3][c][1], eager[3][c][1])
        other = gpu.OP_FREE64 if store == "matrix-free" else gpu.OP_STORED
        pb.apply_batch(np.ones((pb.count, 3)), other)  # other device work between the solves
    pb.set_graphs(False)


@pytest.mark.parametrize("store,method,fp32", [("C128_PAIR", "bicg", False), ("MIXED_PAIR", "reliable", True)])
@pytest.mark.parametrize("chunk", [2, 3])
def test_grouped_graph_replay_gives_the_eager_bits(gpu, store, method, fp32, chunk):
    """ibim_solve_group with graphs on its first problem: full chunks after the
    first replay a captured graph of every running batch's iterations, captured
    again whenever a host check changes which batches run (here problems of
    3,7xx, 2,6xx and 1,3xx points, one with two batches, finish at different
    iterations). Twice in a row, with other device work between: the eager bits."""
    sizes, nrhs = (28, 24, 18), (19, 16, 5)
    pbs = []
    for n, r in zip(sizes, nrhs):
        pb = problem(gpu, n=n)
        pb.assemble(getattr(gpu, store))
        if fp32:
            pb.set_fp32_products(True)
        fill_rhs(gpu, pb, r)
        pbs.append(pb)
    m = gpu.SOLVE_BICG if method == "bicg" else gpu.SOLVE_RELIABLE

    def run_group():
        infos = gpu.solve_group(pbs, m, maxit=500, tol=1e-11, inner_tol=1e-2, chunk=chunk)
        return ([(i.iterations, i.outer, i.status) for i in infos],
                [[pb.density(c) for c in range(r)] for pb, r in zip(pbs, nrhs)])

    eager = run_group()
    assert all(st == 0 for _, _, st in eager[0])
    pbs[0].set_graphs(True)
    for rep in range(2):
        got = run_group()
        assert got[0] == eager[0], (rep, got[0], eager[0])
        for dg, de in zip(got[1], eager[1]):
            for a, b in zip(dg, de):
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        pbs[1].apply_batch(np.ones((pbs[1].count, 3)), gpu.OP_STORED)  # other device work between the solves
    for pb in pbs:
        pb.close()
