"""The CPU oracle against what the mathematics fixes, for the parts the
reference's golden fixtures leave open (Dirichlet / interior variants, right-hand
sides, evaluators, end-to-end accuracy) and as a second check on the rest.

Each check is chosen so that a plausible slip (a dropped term, a sign, a
transposed operand, a wrong index) fails it.
"""
import numpy as np
import pytest
from scipy import special
from scipy.integrate import quad

import oracle_ctypes as O
import shapes


def relnorm(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


# ------------------------------------------------------------------- WENO --
def test_weno_exact_on_linear_data():
    for order in (3, 5):
        f = 0.3 + 1.7 * np.arange(-3, 4) * 0.05
        dp, dm = O.weno(f if order == 5 else f[1:6], 0.05, order)
        assert abs(dp - 1.7) < 1e-12 and abs(dm - 1.7) < 1e-12


def test_weno5_fifth_order_on_smooth_data():
    errs = []
    for dx in (0.02, 0.01):
        f = np.sin(0.3 + np.arange(-3, 4) * dx)
        dp, dm = O.weno(f, dx, 5)
        errs.append(max(abs(dp - np.cos(0.3)), abs(dm - np.cos(0.3))))
    assert errs[1] < errs[0] / 20  # ~2^5 = 32


def test_gradient_exact_for_a_plane():
    n = 16
    x, y, z = shapes.grid_coords(n)
    nrm = np.array([0.48, -0.6, 0.64])
    u = 0.3 - (x * nrm[0] + y * nrm[1] + z * nrm[2])
    gx, gy, gz, mc, gc = O.geometry(u, n)
    # away from the faces, where the mirror extension (weno.c:131-139) bends the data
    inner = (slice(3, -3),) * 3
    assert np.abs(gx[inner] + nrm[0]).max() < 1e-12
    assert np.abs(gy[inner] + nrm[1]).max() < 1e-12
    assert np.abs(gz[inner] + nrm[2]).max() < 1e-12
    # a plane has no curvature (interior nodes; the faces use one-sided formulas)
    assert np.abs(mc[3:-3, 3:-3, 3:-3]).max() < 1e-10 and np.abs(gc[3:-3, 3:-3, 3:-3]).max() < 1e-10


def test_sphere_curvatures_converge_at_second_order():
    errs = []
    for n in (32, 64):
        u = shapes.sphere(n)
        _, _, _, mc, gc = O.geometry(u, n)
        rho = 1.0 - u
        band = (rho > 0.8) & (rho < 1.2)
        errs.append((np.abs(mc[band] - 1 / rho[band]).max(), np.abs(gc[band] - 1 / rho[band] ** 2).max()))
    assert errs[1][0] < errs[0][0] / 3 and errs[1][1] < errs[0][1] / 3


# ------------------------------------------------------------- sampling --
def test_trilinear_interpolation_exact_for_trilinear_fields():
    n = 16
    x, y, z = shapes.grid_coords(n)
    f = lambda a, b, c: 0.2 + 0.3 * a - 0.5 * b + 0.7 * c + 0.1 * a * b - 0.2 * b * c + 0.3 * a * c + 0.05 * a * b * c
    gx = np.ascontiguousarray(f(x, y, z))
    gy = np.ascontiguousarray(np.ones_like(x))
    gz = np.ascontiguousarray(np.zeros_like(x))
    rng = np.random.default_rng(0)
    for p in rng.uniform(-1.9, 1.9, (200, 3)):
        code, v = O.interp_grad(n, (gx, gy, gz), p)
        val = f(*p)
        # returns the normalised vector (val, 1, 0) / |.|
        assert abs(v[0] / v[1] - val) < 1e-12


# ------------------------------------------------------------- weights --
def test_cosine_delta_sums_to_one_on_shifted_grids():
    h, eps = 0.1, 0.3
    L = O.lib()
    for s in (0.0, 0.013, 0.05, 0.0999):
        xs = np.arange(-10, 11) * h + s
        assert abs(sum(L.ibo_delta_cos(eps, x) for x in xs) * h - 1.0) < 1e-14


def test_sine_weight_unit_mass_zero_first_moment():
    L = O.lib()
    w = lambda t: L.ibo_sine_weight(t, 0.1, 0.1)
    assert abs(quad(w, 0.1, 1, limit=200)[0] - 1) < 1e-10
    assert abs(quad(lambda t: t * w(t), 0.1, 1, limit=200)[0]) < 1e-10
    p = lambda t: L.ibo_poly_weight(t, 0.1)
    assert abs(quad(p, 0.1, 1)[0] - 1) < 1e-12 and abs(quad(lambda t: t * p(t), 0.1, 1)[0]) < 1e-12


# --------------------------------------------------------------- kernels --
def _sphere(n=24, k=3.0, **kw):
    h = 4.0 / n
    level = shapes.sphere(n)
    geo = O.geometry(level, n)
    prm = O.make_params(k=k, eps=3 * h, tau=2 * h, **kw)
    node, flat = O.tube(level, geo, n, prm, 1 if kw.get("delta_in") else 0)
    return prm, flat, len(node)


def phi(x, y, k):
    r = np.linalg.norm(np.asarray(x) - np.asarray(y))
    return np.exp(1j * k * r) / (4 * np.pi * r)


@pytest.mark.parametrize("dirichlet,interior", [(0, 0), (1, 0), (0, 1)])
def test_traditional_entries_are_normal_derivatives_of_phi(dirichlet, interior):
    prm, flat, cnt = _sphere(dirichlet=dirichlet, interior=interior)
    F = O.tube_fields(flat, cnt)
    are, aim = O.kernel_rows(flat, cnt, prm, 0, 40)
    s = np.stack([F["sx"], F["sy"], F["sz"]], 1)
    nrm = np.stack([F["nx"], F["ny"], F["nz"]], 1)
    rng = np.random.default_rng(1)
    d = 1e-5
    for i in range(40):
        for j in rng.integers(0, cnt, 10):
            if np.linalg.norm(s[i] - s[j]) <= prm.tau:
                continue
            v = nrm[j] if dirichlet else nrm[i]
            # grad_x Phi(x, y) . v by central differences in x
            fd = (phi(s[i] + d * v, s[j], prm.k) - phi(s[i] - d * v, s[j], prm.k)) / (2 * d)
            if interior:
                fd = -fd
            got = are[i, j] + 1j * aim[i, j] - (0.5 if i == j else 0)
            assert abs(got - F["w"][j] * fd) <= 1e-7 * abs(F["w"][j] * fd) + 1e-300


def test_traditional_near_field_and_diagonal():
    prm, flat, cnt = _sphere()
    F = O.tube_fields(flat, cnt)
    are, aim = O.kernel_rows(flat, cnt, prm, 0, cnt)
    # diagonal: rho = 0 <= tau -> w_i reg_i + 1/2, real
    assert np.allclose(np.diag(are), F["w"] * F["reg"] + 0.5, rtol=0, atol=1e-15)
    assert not np.any(np.diag(aim))
    # the exterior regularisation of a unit sphere: H* = 1, K* = 1 -> a closed form
    tau, k = prm.tau, prm.k
    reg = -(1 / (4 * np.pi * tau) - 5 * tau * 2 * 3 / (512 * np.pi) - 25 * tau * 2 / (1536 * np.pi)
            + k * k * tau * 2 / (48 * np.pi))
    close = np.abs(F["hm"] - 1) < 0.02
    assert np.allclose(F["reg"][close], reg, rtol=0.05)


def test_combo_entries_match_finite_differences():
    n, k, eta = 24, 3.0, -1.0
    h = 4.0 / n
    prm, flat, cnt = _sphere(n, k, delta_in=0.3 * h, eta=eta)
    F = O.tube_fields(flat, cnt)
    are, aim, _ = O.kernel_rows(flat, cnt, prm, 0, 20, kind=1)
    s = np.stack([F["sx"], F["sy"], F["sz"]], 1)
    nrm = np.stack([F["nx"], F["ny"], F["nz"]], 1)
    rng = np.random.default_rng(2)
    d = 1e-4
    for i in range(20):
        for j in rng.integers(0, cnt, 8):
            if np.linalg.norm(s[i] - s[j]) <= prm.tau or i == j:
                continue
            R = F["w"][j]
            single = -1j * eta * (-(phi(s[i] + d * nrm[i], s[j], k) - phi(s[i] - d * nrm[i], s[j], k)) / (2 * d)) * R
            x = s[i] + abs(F["dist"][j]) * nrm[i]
            # d^2 Phi / dn_x dn_y at (x, y*_j) by nested central differences
            g = lambda a, b: phi(x + a * nrm[i], s[j] + b * nrm[j], k)
            dd = (g(d, d) - g(d, -d) - g(-d, d) + g(-d, -d)) / (4 * d * d)
            want = single - dd * R  # d/dn_y of Phi(x, y) is -grad_x Phi . n_y
            got = are[i, j] + 1j * aim[i, j]
            assert abs(got - want) <= 2e-5 * abs(want), (i, j, got, want)


# ---------------------------------------------------------------- solver --
def test_bicg_matches_direct_solve():
    rng = np.random.default_rng(3)
    m = 80
    A = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    A = A / np.sqrt(m) + 2 * np.eye(m)
    b = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    x, xi, its, hist = O.bicg(A.real.copy(), A.imag.copy(), b.real, b.imag, tol=1e-13)
    assert hist[-1] < 1e-13 or its == 1000
    assert relnorm(x + 1j * xi, np.linalg.solve(A, b)) < 1e-11


def test_sweep_is_a_and_its_adjoint():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((30, 50)) + 1j * rng.standard_normal((30, 50))
    p = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    q = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    v, w = O.sweep_rows(A.real.copy(), A.imag.copy(), p, q)
    assert relnorm(v, A @ p) < 1e-14 and relnorm(w, A.conj().T @ q) < 1e-14
    assert abs(np.vdot(q, v) - np.vdot(w, p)) < 1e-12 * abs(np.vdot(q, v))  # <A p, q> = <p, A^H q>


# ------------------------------------------------------ layer potentials --
def test_single_and_double_layer_of_constant_density_on_sphere():
    # S[1](x) = R sin(kR) e^{ikr} / (k r); D[1] = i k^2 R^2 j0'(kR) h0(kr), outward normal
    n, k, R = 48, 2.0, 1.0
    prm, flat, cnt = _sphere(n, k)
    one = np.ones(cnt)
    rec = shapes.shell_receivers(30, 1.6, 3.0, seed=5)
    r = np.linalg.norm(rec, axis=1)
    sre, sim = O.evaluate(flat, cnt, prm, one, np.zeros(cnt), rec, 0)
    s_exact = R * np.sin(k * R) * np.exp(1j * k * r) / (k * r)
    assert relnorm(-(sre + 1j * sim), s_exact) < 5e-3  # mode 0 sums -Phi
    dre, dim = O.evaluate(flat, cnt, prm, one, np.zeros(cnt), rec, 1)
    h0 = np.exp(1j * k * r) / (1j * k * r)
    d_exact = 1j * k * k * R * R * special.spherical_jn(0, k * R, derivative=True) * h0
    assert relnorm(dre + 1j * dim, d_exact) < 5e-3


def _solve_sphere(n, k, rhs):
    prm, flat, cnt = _sphere(n, k, dirichlet=1)
    are, aim = O.kernel_rows(flat, cnt, prm)
    bre, bim = rhs(flat, cnt, prm)
    x, xi, its, hist = O.bicg(are, aim, bre, bim, maxit=1000, tol=1e-10, omp=True)
    return prm, flat, cnt, x, xi


def test_point_source_exterior_dirichlet_converges():
    k, x0 = 2.0, [0.1, 0.2, 0.3]
    rec = shapes.shell_receivers(40, 1.5, 3.0, seed=6)
    exact = shapes.point_source(rec, x0, k)
    errs = []
    for n in (24, 48):
        prm, flat, cnt, x, xi = _solve_sphere(n, k, lambda f, c, p: O.rhs_point_source(f, c, p.k, x0))
        ure, uim = O.evaluate(flat, cnt, prm, x, xi, rec, 1)
        errs.append(relnorm(ure + 1j * uim, exact))
    assert errs[1] < 0.02 and errs[1] < errs[0] / 2.5, errs


def test_sound_soft_plane_wave_matches_mie_series():
    # scattered field of a sound-soft unit sphere, incident e^{ik z}:
    # u_s = -sum_n (2n+1) i^n j_n(kR) / h_n(kR) h_n(kr) P_n(cos theta)
    n, k, R = 48, 2.0, 1.0
    d = np.array([0.0, 0.0, 1.0])
    prm, flat, cnt, x, xi = _solve_sphere(n, k, lambda f, c, p: O.rhs_plane_dirichlet(f, c, p.k, 1.0, d))
    rec = shapes.shell_receivers(40, 1.5, 3.0, seed=7)
    ure, uim = O.evaluate(flat, cnt, prm, x, xi, rec, 1)
    r = np.linalg.norm(rec, axis=1)
    ct = rec[:, 2] / r
    hn = lambda m, z: special.spherical_jn(m, z) + 1j * special.spherical_yn(m, z)
    us = sum(-(2 * m + 1) * 1j ** m * special.spherical_jn(m, k * R) / hn(m, k * R) * hn(m, k * r)
             * special.eval_legendre(m, ct) for m in range(30))
    assert relnorm(ure + 1j * uim, us) < 0.03


def test_neumann_plane_wave_data_is_the_normal_derivative():
    # cal_partialu: amp d/dn e^{ik dir.x} with n the (inward) tube normal
    prm, flat, cnt = _sphere(24, 3.0)
    F = O.tube_fields(flat, cnt)
    d = np.array([-1.0, 0.0, 0.0])
    nn = np.stack([F["nx"], F["ny"], F["nz"]]).copy()
    bre, bim = O.rhs_plane_neumann(flat, cnt, nn, prm.k, 10.0, d)
    s = np.stack([F["sx"], F["sy"], F["sz"]], 1)
    eps = 1e-6
    u = lambda p: 10.0 * np.exp(1j * prm.k * p @ d)
    fd = (u(s + eps * nn.T) - u(s - eps * nn.T)) / (2 * eps)
    assert relnorm(bre + 1j * bim, fd) < 1e-8


def test_combo_evaluator_is_double_minus_i_eta_single():
    # mode 2 = (dPhi/dn_y (inward) - i eta Phi) phi R off the near field
    n, k, eta = 24, 3.0, -1.0
    h = 4.0 / n
    prm, flat, cnt = _sphere(n, k, delta_in=0.3 * h, eta=eta)
    F = O.tube_fields(flat, cnt)
    rng = np.random.default_rng(8)
    ph = rng.standard_normal(cnt) + 1j * rng.standard_normal(cnt)
    rec = shapes.shell_receivers(5, 1.5, 2.5, seed=8)
    ure, uim = O.evaluate(flat, cnt, prm, ph.real.copy(), ph.imag.copy(), rec, 2)
    s = np.stack([F["sx"], F["sy"], F["sz"]], 1)
    nrm = np.stack([F["nx"], F["ny"], F["nz"]], 1)
    e = 1e-6
    for q, x in enumerate(rec):
        tot = 0
        for j in range(cnt):
            dn = (phi(x, s[j] + e * nrm[j], k) - phi(x, s[j] - e * nrm[j], k)) / (2 * e)
            tot += (dn - 1j * eta * phi(x, s[j], k)) * ph[j] * F["w"][j]
        assert abs(ure[q] + 1j * uim[q] - tot) < 1e-7 * abs(tot)


# ------------------------------------------- combined-field Dirichlet operator --
def _combined_pair(n, k, eta):
    """Oracle rows of the double-layer Dirichlet operator and of the combined
    one on the unit sphere, with the tube fields."""
    prm0, flat, cnt = _sphere(n, k, dirichlet=1)
    prm1 = O.make_params(k=k, eps=prm0.eps, tau=prm0.tau, dirichlet=1, combined=1, eta=eta)
    a0 = O.kernel_rows(flat, cnt, prm0)
    a1 = O.kernel_rows(flat, cnt, prm1)
    return prm0, prm1, flat, cnt, a0[0] + 1j * a0[1], a1[0] + 1j * a1[1]


def test_combined_single_layer_of_constant_density_on_sphere():
    # the coupling term is -i eta S: its rows applied to 1 give the single layer
    # of a constant density on the unit sphere, S[1] = R sin(kR) e^{ikR} / k at
    # |x| = R, here with the near field as the disc mean of Phi (without it
    # the sums would miss ~tau / 2, 18% at 48^3)
    k, eta, R = 2.0, 2.0, 1.0
    exact = R * np.sin(k * R) * np.exp(1j * k * R) / k
    errs = []
    for n in (24, 48):
        _, _, _, _, A0, A1 = _combined_pair(n, k, eta)
        s1 = (1j * (A1 - A0) / eta).sum(axis=1)
        errs.append(np.abs(s1 - exact).max() / abs(exact))
    assert errs[1] < 0.015 and errs[1] < errs[0] / 1.8, errs


def test_combined_near_field_is_the_disc_mean_of_phi():
    # near entries (rho <= tau): -i eta w_j (1 / (pi tau^2)) int_disc Phi dA, the
    # integral by quadrature in polar coordinates
    k, eta = 5.0, 3.0
    prm0, prm1, flat, cnt, A0, A1 = _combined_pair(24, k, eta)
    F = O.tube_fields(flat, cnt)
    tau = prm0.tau
    g = lambda r: np.exp(1j * k * r) / 2  # Phi(r) 2 pi r
    mean = (quad(lambda r: g(r).real, 0, tau)[0] + 1j * quad(lambda r: g(r).imag, 0, tau)[0]) / (np.pi * tau ** 2)
    s = np.stack([F["sx"], F["sy"], F["sz"]], 1)
    for i in (0, cnt // 3, cnt - 1):
        near = np.linalg.norm(s - s[i], axis=1) <= tau
        assert near.sum() > 10
        np.testing.assert_allclose((A1 - A0)[i, near], -1j * eta * F["w"][near] * mean, rtol=1e-13)


def test_combined_far_entries_are_double_layer_minus_i_eta_phi():
    k, eta = 4.0, 4.0
    prm0, prm1, flat, cnt, A0, A1 = _combined_pair(24, k, eta)
    F = O.tube_fields(flat, cnt)
    s = np.stack([F["sx"], F["sy"], F["sz"]], 1)
    rng = np.random.default_rng(11)
    for i in rng.integers(0, cnt, 20):
        for j in rng.integers(0, cnt, 20):
            if np.linalg.norm(s[i] - s[j]) <= prm0.tau:
                continue
            want = -1j * eta * F["w"][j] * phi(s[i], s[j], k)
            assert abs((A1 - A0)[i, j] - want) <= 1e-13 * abs(want)


def test_combined_field_solves_the_point_source_problem_well_conditioned():
    # k = 2 lies near the unit sphere's interior Neumann eigenvalue 2.0816, where
    # the double layer alone is ill-conditioned; the combined operator is not,
    # and its potential (mode 3) reproduces the exterior point-source field
    k, eta, x0 = 2.0, 2.0, [0.1, 0.2, 0.3]
    rec = shapes.shell_receivers(40, 1.5, 3.0, seed=6)
    exact = shapes.point_source(rec, x0, k)
    errs = []
    for n in (24, 48):
        prm0, prm1, flat, cnt, A0, A1 = _combined_pair(n, k, eta)
        bre, bim = O.rhs_point_source(flat, cnt, k, x0)
        x = np.linalg.solve(A1, bre + 1j * bim)
        ure, uim = O.evaluate(flat, cnt, prm1, x.real.copy(), x.imag.copy(), rec, 3)
        errs.append(relnorm(ure + 1j * uim, exact))
        if n == 24:
            assert np.linalg.cond(A1) < 0.25 * np.linalg.cond(A0)
    assert errs[1] < 0.005 and errs[1] < errs[0] / 2, errs


def test_combined_evaluator_is_double_minus_i_eta_single():
    k, eta = 3.0, 3.0
    prm0, prm1, flat, cnt, _, _ = _combined_pair(16, k, eta)
    rng = np.random.default_rng(12)
    ph = rng.standard_normal(cnt) + 1j * rng.standard_normal(cnt)
    rec = shapes.shell_receivers(6, 1.5, 2.5, seed=12)
    u3 = O.evaluate(flat, cnt, prm1, ph.real.copy(), ph.imag.copy(), rec, 3)
    u1 = O.evaluate(flat, cnt, prm1, ph.real.copy(), ph.imag.copy(), rec, 1)
    u0 = O.evaluate(flat, cnt, prm1, ph.real.copy(), ph.imag.copy(), rec, 0)  # sums -Phi
    want = (u1[0] + 1j * u1[1]) + 1j * eta * (u0[0] + 1j * u0[1])
    assert relnorm(u3[0] + 1j * u3[1], want) < 1e-13
