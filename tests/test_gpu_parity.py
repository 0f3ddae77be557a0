"""GPU parity: libibim.so (HIP, gfx950) against the CPU oracle and the golden
fixtures written from the reference. Every call goes through the C ABI.

Tolerances (FP64 path): geometry and tube fields 1e-13 relative (the GPU build
has no FMA contraction, so only libm-vs-ocml transcendental ulps differ);
operator entries 1e-13 of the largest entry; fused A p / A^H q 1e-13 relative
in norm; BiCG solutions to the oracle's within 1e-9 relative when both run to
tol 1e-12. FP32 matrix-free sweep: 2e-5 relative in norm. Refinement: within
1e-6 of the FP64 solution, the bar the mixed-precision path is held to.
"""
import os

import numpy as np
import pytest

import oracle_ctypes as O
import shapes

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return np.load(os.path.join(G, name))


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def relnorm(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


# ----------------------------------------------------------------- geometry --
@pytest.mark.parametrize("name", ["geom_sphere.npz", "geom_torus.npz", "geom_twin.npz", "geom3_sphere.npz"])
def test_geometry_matches_reference_fixture(gpu, name):
    g = golden(name)
    order = 3 if name.startswith("geom3") else 5
    pb = gpu.Problem().set_level(g["level"]).geometry(order)
    gx, gy, gz, mc, gc = pb.grid_fields()
    for a, k in zip((gx, gy, gz, mc, gc), ("gx", "gy", "gz", "mc", "gc")):
        assert relmax(a, g[k]) <= 1e-13, (k, relmax(a, g[k]))
    # the Godunov choices and limiter branches are discrete: count exact agreement
    assert np.mean(gx == g["gx"]) > 0.999 and np.mean(mc == g["mc"]) > 0.999


def _oracle_tube(level, n, prm_kw, kind):
    geo = O.geometry(level, n)
    prm = O.make_params(**prm_kw)
    node, flat = O.tube(level, geo, n, prm, kind)
    return geo, prm, node, flat


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("shape", ["sphere", "torus"])
def test_tube_matches_oracle(gpu, kind, shape):
    n = 24
    level = shapes.sphere(n) if shape == "sphere" else shapes.torus(n)
    h = 4.0 / n
    kw = dict(k=3.0, eps=3 * h, tau=2 * h, delta_in=0.3 * h if kind else 0.0, eta=-1.0)
    geo, prm, node, flat = _oracle_tube(level, n, kw, kind)
    pb = gpu.Problem().set_level(level).geometry(5)
    cnt = pb.tube(gpu.params(**kw), kind)
    gnode, gflat = pb.tube_fields()
    assert cnt == len(node)
    assert np.array_equal(gnode, node)  # raster order, bit-exact indices
    F, GF = O.tube_fields(flat, cnt), O.tube_fields(gflat, cnt)
    for f in O.FIELDS:
        assert relmax(GF[f], F[f]) <= 1e-13, (f, relmax(GF[f], F[f]))


# ----------------------------------------------------------------- operator --
def _sphere_problem(gpu, n=24, k=2.0, kind=0, dirichlet=1, assemble=True):
    level = shapes.sphere(n)
    h = 4.0 / n
    kw = dict(k=k, eps=3 * h, tau=2 * h, delta_in=0.3 * h if kind else 0.0, eta=-1.0, dirichlet=dirichlet)
    geo, prm, node, flat = _oracle_tube(level, n, kw, kind)
    pb = gpu.Problem().set_level(level).geometry(5)
    pb.tube(gpu.params(**kw), kind)
    if assemble:
        pb.assemble(gpu.C128)
    return pb, prm, flat, len(node)


@pytest.mark.parametrize("kind,dirichlet", [(0, 0), (0, 1), (1, 0)])
def test_assembled_operator_matches_oracle(gpu, kind, dirichlet):
    pb, prm, flat, cnt = _sphere_problem(gpu, kind=kind, dirichlet=dirichlet)
    res = O.kernel_rows(flat, cnt, prm, kind=kind)
    are, aim = res[0], res[1]
    gre, gim = pb.rows(0, cnt)
    scale = max(np.abs(are).max(), np.abs(aim).max())
    assert np.abs(gre - are).max() <= 1e-13 * scale
    assert np.abs(gim - aim).max() <= 1e-13 * scale
    if kind == 1:
        assert pb.regcount() == res[2]


def test_dirichlet_kernel_is_minus_transposed_neumann(gpu):
    # grad_x Phi(x_j, y_i) = -grad_x Phi(x_i, y_j): K^D_ij / w_j = -K^N_ji / w_i off the near field
    pbn, prm, flat, cnt = _sphere_problem(gpu, dirichlet=0)
    pbd, _, _, _ = _sphere_problem(gpu, dirichlet=1)
    F = O.tube_fields(flat, cnt)
    s = np.stack([F["sx"], F["sy"], F["sz"]], 1)
    w = F["w"]
    nre, nim = pbn.rows(0, cnt)
    dre, dim = pbd.rows(0, cnt)
    far = np.linalg.norm(s[:, None, :] - s[None, :, :], axis=2) > prm.tau
    # K^D_ij / w_j = -K^N_ji / w_i, multiplied out (w vanishes at the tube edge)
    kn = (nre + 1j * nim) * w[:, None]
    kd = (dre + 1j * dim) * w[:, None]
    assert np.abs((kd + kn.T)[far]).max() <= 1e-12 * np.abs(kn[far]).max()


@pytest.mark.parametrize("kind", [0, 1])
def test_fused_sweep_matches_numpy(gpu, kind):
    pb, prm, flat, cnt = _sphere_problem(gpu, kind=kind, dirichlet=0)
    res = O.kernel_rows(flat, cnt, prm, kind=kind)
    A = res[0] + 1j * res[1]
    rng = np.random.default_rng(11)
    p = rng.standard_normal(cnt) + 1j * rng.standard_normal(cnt)
    q = rng.standard_normal(cnt) + 1j * rng.standard_normal(cnt)
    y, z = pb.apply(p, q, gpu.OP_STORED)
    assert relnorm(y, A @ p) <= 1e-13
    assert relnorm(z, A.conj().T @ q) <= 1e-13
    y64, z64 = pb.apply(p, q, gpu.OP_FREE64)
    assert relnorm(y64, A @ p) <= 1e-13
    assert relnorm(z64, A.conj().T @ q) <= 1e-13


def test_fp32_matrix_free_sweep_accuracy(gpu):
    pb, prm, flat, cnt = _sphere_problem(gpu, n=32, k=10.0, dirichlet=1)
    rng = np.random.default_rng(12)
    p = rng.standard_normal(cnt) + 1j * rng.standard_normal(cnt)
    q = rng.standard_normal(cnt) + 1j * rng.standard_normal(cnt)
    y, z = pb.apply(p, q, gpu.OP_STORED)
    y32, z32 = pb.apply(p, q, gpu.OP_FREE32)
    assert relnorm(y32, y) <= 2e-5, relnorm(y32, y)
    assert relnorm(z32, z) <= 2e-5, relnorm(z32, z)


# -------------------------------------------------------------------- solver --
def test_bicg_on_reference_fixture(gpu):
    b = golden("bicg.npz")
    pb = gpu.Problem()
    pb.set_matrix(b["are"], b["aim"])
    pb.set_rhs(b["bre"], b["bim"])
    info, hist = pb.solve(maxit=1000, tol=1e-7)
    x, xi = pb.density()
    assert abs(info.iterations - int(b["serial_its"])) <= 1
    np.testing.assert_allclose(x, b["serial_x"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(xi, b["serial_xi"], rtol=0, atol=1e-9)
    # the reference prints its final error with %.10E
    assert abs(hist[-1] - b["serial_err"][0]) <= 1e-6 * b["serial_err"][0]


def test_bicg_residual_history_matches_oracle(gpu):
    pb, prm, flat, cnt = _sphere_problem(gpu, dirichlet=1)
    bre, bim = O.rhs_point_source(flat, cnt, prm.k, [0.1, 0.2, 0.3])
    are, aim = O.kernel_rows(flat, cnt, prm)
    x, xi, its, hist = O.bicg(are, aim, bre, bim, maxit=500, tol=1e-12, omp=True)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    gb = pb.get_rhs()
    assert relmax(gb[0], bre) <= 1e-14 and relmax(gb[1], bim) <= 1e-14
    info, ghist = pb.solve(maxit=500, tol=1e-12)
    gx, gxi = pb.density()
    assert abs(info.iterations - its) <= 2
    m = min(10, its)
    # relative residuals agree to reduction-order rounding (~1e-13 of ||b||)
    np.testing.assert_allclose(ghist[:m], hist[:m], rtol=1e-9, atol=1e-11)
    assert relnorm(gx + 1j * gxi, x + 1j * xi) <= 1e-9


def test_zero_rhs_gives_zero_density(gpu):
    pb, prm, flat, cnt = _sphere_problem(gpu)
    pb.set_rhs(np.zeros(cnt), np.zeros(cnt))
    info, _ = pb.solve()
    assert info.status == 4 and info.iterations == 0
    assert not np.any(pb.density()[0])


def test_solves_are_bitwise_repeatable(gpu):
    pb, prm, flat, cnt = _sphere_problem(gpu, dirichlet=1)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    pb.solve(tol=1e-10)
    a = pb.density()
    pb.solve(tol=1e-10)
    b = pb.density()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_point_source_exterior_dirichlet_end_to_end(gpu):
    # the exterior field of an interior point source is the exact solution of the
    # exterior Dirichlet problem with its own trace as data
    n, k, x0 = 48, 2.0, [0.1, 0.2, 0.3]
    pb, prm, flat, cnt = _sphere_problem(gpu, n=n, k=k, dirichlet=1)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=x0)
    info, _ = pb.solve(maxit=1000, tol=1e-10)
    assert info.status == 0
    rec = shapes.shell_receivers(64, 1.5, 3.0, seed=0)
    ure, uim = pb.evaluate(rec, gpu.EVAL_DOUBLE)
    u = ure + 1j * uim
    exact = shapes.point_source(rec, x0, k)
    assert relnorm(u, exact) < 2e-2, relnorm(u, exact)
    # and the evaluator agrees with the oracle's on the same density
    x, xi = pb.density()
    ore, oim = O.evaluate(flat, cnt, prm, x, xi, rec, 1)
    assert relnorm(u, ore + 1j * oim) <= 1e-12


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_evaluation_matches_oracle(gpu, mode):
    kind = 1 if mode == 2 else 0
    pb, prm, flat, cnt = _sphere_problem(gpu, kind=kind, dirichlet=0)
    rng = np.random.default_rng(13)
    phr, phi = rng.standard_normal(cnt), rng.standard_normal(cnt)
    pb.set_density(phr, phi)
    rec = np.concatenate([shapes.shell_receivers(20, 1.2, 2.5, seed=1), [[1.0, 0.0, 0.0]]])
    ure, uim = pb.evaluate(rec, mode)
    ore, oim = O.evaluate(flat, cnt, prm, phr, phi, rec, mode)
    assert relnorm(ure + 1j * uim, ore + 1j * oim) <= 1e-12


def test_plane_wave_rhs_matches_oracle(gpu):
    pb, prm, flat, cnt = _sphere_problem(gpu, dirichlet=0)
    F = O.tube_fields(flat, cnt)
    d = np.array([-1.0, 0.0, 0.0])
    pb.rhs(gpu.RHS_PLANE_NEUMANN, amp=10.0, direction=d)
    gre, gim = pb.get_rhs()
    bre, bim = O.rhs_plane_neumann(flat, cnt, np.stack([F["nx"], F["ny"], F["nz"]]).copy(), prm.k, 10.0, d)
    assert relmax(gre, bre) <= 1e-13 and relmax(gim, bim) <= 1e-13
    pb.rhs(gpu.RHS_PLANE_DIRICHLET, amp=1.0, direction=[0.0, 0.6, 0.8])
    gre, gim = pb.get_rhs()
    bre, bim = O.rhs_plane_dirichlet(flat, cnt, prm.k, 1.0, [0.0, 0.6, 0.8])
    assert relmax(gre, bre) <= 1e-13 and relmax(gim, bim) <= 1e-13


# ------------------------------------------------------------ mixed precision --
def test_refinement_matches_fp64_solution(gpu):
    pb, prm, flat, cnt = _sphere_problem(gpu, n=32, k=6.0, dirichlet=1)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    info64, _ = pb.solve(gpu.SOLVE_BICG, maxit=2000, tol=1e-11)
    x64 = pb.density()
    free = gpu.Problem().set_level(shapes.sphere(32)).geometry(5)
    free.tube(gpu.params(k=prm.k, eps=prm.eps, tau=prm.tau, dirichlet=1))
    free.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    info, hist = free.solve(gpu.SOLVE_REFINED, maxit=2000, tol=1e-8, inner_tol=1e-4, max_outer=20)
    xr = free.density()
    assert info.status == 0
    assert relnorm(xr[0] + 1j * xr[1], x64[0] + 1j * x64[1]) <= 1e-6


@pytest.mark.parametrize("store", ["c64", "free32"])
def test_residual_replacement_matches_fp64_solution(gpu, store):
    """BiCG on the mixed-precision operator with FP64 residual replacement reaches
    the FP64 operator's solution (true residual below tol) with an iteration count
    close to the FP64 solve's."""
    pb, prm, flat, cnt = _sphere_problem(gpu, n=32, k=6.0, dirichlet=1)
    pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.1, 0.2, 0.3])
    info64, _ = pb.solve(gpu.SOLVE_BICG, maxit=2000, tol=1e-12)
    x64 = pb.density()
    if store == "c64":
        pb.assemble(gpu.C64)  # the replacements then use the FP64 matrix-free operator
    seen = []
    pb.set_progress(lambda what, it, res: seen.append((what, it, res)))
    info, hist = pb.solve(gpu.SOLVE_RELIABLE, maxit=2000, tol=1e-10, inner_tol=1e-3, chunk=1)
    pb.set_progress(None)
    x = pb.density()
    assert info.status == 0 and info.rel_residual < 1e-10 and info.outer >= 1
    assert len(hist) == info.iterations and hist[-1] < 1e-10
    assert info.iterations <= 1.5 * info64.iterations
    assert relnorm(x[0] + 1j * x[1], x64[0] + 1j * x64[1]) <= 1e-8
    assert seen and seen[0][0] == "BiCG"  # the first host check always reports


def test_degenerate_inputs_are_errors_not_faults(gpu):
    """A level set whose tube is empty, solves before a right-hand side exists,
    and out-of-range columns come back as IbimError (the reference exits)."""
    n = 16
    far = np.full((n + 1,) * 3, 5.0)  # every node 5 inside: no |d| <= eps
    pb = gpu.Problem().set_level(far).geometry(5)
    with pytest.raises(gpu.IbimError):
        pb.tube(gpu.params(k=2.0, eps=0.75, tau=0.5, dirichlet=1))
    with pytest.raises(gpu.IbimError):
        pb.assemble(gpu.C128)
    with pytest.raises(gpu.IbimError):
        pb.solve(gpu.SOLVE_BICG, maxit=10, tol=1e-8)
    pb2, prm, flat, cnt = _sphere_problem(gpu, n=16, k=2.0, dirichlet=1)
    with pytest.raises(gpu.IbimError):
        pb2.rhs(gpu.RHS_POINT_SOURCE, x0=[0.0, 0.0, 0.0], col=1)  # one column only
    with pytest.raises(gpu.IbimError):
        pb2.set_nrhs(0)
    with pytest.raises(gpu.IbimError):
        pb2.apply_batch(np.zeros((cnt, 17)), gpu.OP_STORED)  # at most 16 vectors


def test_host_matrix_without_a_tube_refuses_point_operators(gpu):
    """An operator uploaded from the host has no points: the matrix-free
    operators, the mixed solvers' FP64 operator, right-hand-side generators and
    the evaluation all need a tube, so they are errors before any launch --
    on a fresh handle and after an earlier, shorter tube alike."""
    rng = np.random.default_rng(2)
    m = 600
    A = np.eye(m) * 2 + 0.01 * (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
    b = rng.standard_normal(m) + 0j
    fresh = gpu.Problem()
    h = 4.0 / 12
    earlier = gpu.Problem().set_level(shapes.sphere(12)).geometry(5)
    cnt = earlier.tube(gpu.params(k=2.0, eps=1.5 * h, tau=1.5 * h, dirichlet=1))
    earlier.assemble(gpu.C128_PAIR)
    assert 0 < cnt < m
    for pb in (fresh, earlier):
        pb.set_matrix(A.real.copy(), A.imag.copy(), gpu.C128)
        pb.set_rhs(b.real.copy(), b.imag.copy())
        for op in (gpu.OP_FREE64, gpu.OP_FREE32):
            with pytest.raises(gpu.IbimError):
                pb.apply(b, b, op)
            with pytest.raises(gpu.IbimError):
                pb.apply_batch(np.ones((m, 2)), op)
        for method in (gpu.SOLVE_RELIABLE, gpu.SOLVE_REFINED):
            with pytest.raises(gpu.IbimError):
                pb.solve(method, maxit=50, tol=1e-10)
        with pytest.raises(gpu.IbimError):
            pb.rhs(gpu.RHS_POINT_SOURCE, x0=[0.0, 0.0, 0.0])
        with pytest.raises(gpu.IbimError):
            pb.evaluate(np.array([[2.0, 0.0, 0.0]]), gpu.EVAL_DOUBLE)
        # BiCG on the stored operator itself still works
        info, _ = pb.solve(gpu.SOLVE_BICG, maxit=200, tol=1e-12)
        x = pb.density()
        assert info.status == 0
        assert relnorm(x[0] + 1j * x[1], np.linalg.solve(A, b)) < 1e-10


# ------------------------------------------------------------------ drop-ins --
def test_compbicg_dropin_matches_reference(gpu):
    b = golden("bicg.npz")
    x, xi = gpu.CompBiCG(b["are"], b["aim"], b["bre"], b["bim"])
    np.testing.assert_allclose(x, b["serial_x"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(xi, b["serial_xi"], rtol=0, atol=1e-9)
    x, xi = gpu.omp_CompBiCG(b["are"], b["aim"], b["bre"], b["bim"])
    np.testing.assert_allclose(x, b["omp_x"], rtol=0, atol=1e-9)


def test_assign_gradient_dropin_matches_reference(gpu):
    g = golden("geom_twin.npz")
    n = int(g["n"])
    gx, gy, gz = gpu.ASSIGN_GRADIENT(n, 4.0 / n, g["level"], 5)
    for a, k in zip((gx, gy, gz), ("gx", "gy", "gz")):
        assert relmax(a, g[k]) <= 1e-13


def test_combo_dropin_matches_reference(gpu):
    c = golden("combo_sphere.npz")
    geo = golden("geom_sphere.npz")
    n = int(c["n"])
    gpu.set_globals(n, float(c["k"]))
    regs, arrs, K, Ki = gpu.new_HelmholtzKernelCombo(int(c["size"]), float(c["delta"]), float(c["eps"]),
                                                     geo["level"], geo["gx"], geo["gy"], geo["gz"], geo["mc"],
                                                     geo["gc"], float(c["k"]), float(c["eta"]))
    assert regs == int(c["regcount"])
    for k in ("sx", "sy", "sz", "dist", "nx", "ny", "nz", "jac", "reg", "w"):
        src = "siner" if k == "w" else k
        assert relmax(arrs[src], c[k]) <= 1e-13, k
    scale = np.abs(c["row_re"]).max()
    assert np.abs(K[c["rows"]] - c["row_re"]).max() <= 1e-13 * scale
    assert np.abs(Ki[c["rows"]] - c["row_im"]).max() <= 1e-13 * scale
    assert np.abs(K[c["pi"], c["pj"]] - c["entry_re"]).max() <= 1e-13 * scale


def test_double_layer_dropin_matches_reference(gpu):
    t = golden("trad_sphere.npz")
    geo = golden("geom_sphere.npz")
    n = int(t["n"])
    gpu.set_globals(n, float(t["k"]))
    arrs, K, Ki = gpu.DoubleLayerKernel(int(t["size"]), float(t["eps"]), geo["level"], geo["gx"], geo["gy"],
                                        geo["gz"], geo["mc"], geo["gc"], float(t["k"]))
    for k in ("sx", "sy", "sz", "dist", "delta", "jac"):
        assert relmax(arrs[k], t[k]) <= 1e-13, k
    assert relmax(arrs["result"], t["w"]) <= 1e-13
    assert relmax(arrs["reg"], t["reg_interior"]) <= 1e-13
    scale = np.abs(t["dl_re"]).max()
    # finite differences of Phi with step H amplify the ulp differences of the
    # normal (normalised once in the reference, twice here) and of libm vs ocml
    assert np.abs(K[t["dl_rows"]] - t["dl_re"]).max() <= 1e-11 * scale
    assert np.abs(Ki[t["dl_rows"]] - t["dl_im"]).max() <= 1e-11 * scale


def test_new_kernel_dropin_matches_reference(gpu):
    g = golden("newkernel.npz")
    n = int(g["n"])
    gpu.set_globals(n, float(g["k"]))
    arrs, K, Ki = gpu.new_HelmholtzKernel(int(g["size"]), float(g["delta"]), float(g["eps"]))
    for k in ("sx", "sy", "sz", "dist", "nx", "ny", "nz", "curv", "jac", "sinew"):
        assert relmax(arrs[k], g[k]) <= 1e-13, k
    assert relmax(arrs["siner"], g["w"]) <= 1e-13
    scale = np.abs(g["row_re"]).max()
    assert np.abs(K[g["rows"]] - g["row_re"]).max() <= 1e-11 * scale
    assert np.abs(Ki[g["rows"]] - g["row_im"]).max() <= 1e-11 * scale
