"""The CPU oracle against the golden fixtures written from the compiled reference.

Fixtures: tests/golden/*.npz, written by tests/golden/make_golden.py from the
reference's own C sources (oracle/_ref). Bitwise equality is required wherever
the oracle keeps the reference's operation order; the two documented exceptions
are the double normalisation of the traditional normal (1 ulp) and the %.15lf
text quantisation of the reference's right-hand side.
"""
import os

import numpy as np
import pytest

import oracle_ctypes as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name))


GEOM = ["geom_sphere.npz", "geom_torus.npz", "geom_twin.npz"]


def test_weno_stencils_bitwise():
    g = load("weno.npz")
    for order, key in ((3, "out3"), (5, "out5")):
        r = np.array([O.weno(f if order == 5 else f[1:6], float(g["dx"]), order) for f in g["stencils"]])
        assert np.array_equal(r, g[key]), order


@pytest.mark.parametrize("name", GEOM + ["geom3_sphere.npz"])
def test_gradient_and_curvatures_bitwise(name):
    g = load(name)
    n = int(g["n"])
    order = 3 if name.startswith("geom3") else 5
    res = O.geometry(g["level"], n, order)
    for a, k in zip(res, ("gx", "gy", "gz", "mc", "gc")):
        assert np.array_equal(a, g[k]), k


@pytest.mark.parametrize("name", GEOM)
def test_weno_axis_and_second_derivatives_per_node(name):
    g = load(name)
    n = int(g["n"])
    u = np.ascontiguousarray(g["level"])
    for m, (i, j, k) in enumerate(g["nodes"]):
        for ax in range(3):
            assert O.weno_axis(u, n, 4.0 / n, int(i), int(j), int(k), ax, 5) == tuple(g["weno_axis"][m, ax])
        assert np.array_equal(O.second_derivs(u, n, int(i), int(j), int(k)), g["d2"][m])


@pytest.mark.parametrize("name", ["sphere", "twin"])
def test_trilinear_sampling_bitwise(name):
    g = load("interp_%s.npz" % name)
    geo = load("geom_%s.npz" % name)
    n = int(g["n"])
    for m, p in enumerate(g["points"]):
        code, v = O.interp_grad(n, (geo["gx"], geo["gy"], geo["gz"]), p)
        assert code == g["code"][m]
        assert np.array_equal(v, g["grad"][m])
        assert np.array_equal(O.interp_curv(n, geo["mc"], geo["gc"], p), g["curv"][m])
        assert O.interp_dist(geo["level"], n, p) == g["dist"][m]


def test_primitives_bitwise():
    g = load("prims.npz")
    L = O.lib()
    assert np.array_equal([L.ibo_delta_cos(0.5, x) for x in g["xs"]], g["delta"])
    assert np.array_equal([L.ibo_sine_weight(t, 0.1, 0.1) for t in g["ts"]], g["sine"])
    assert np.array_equal([L.ibo_sine_weight(t, 0.3, 0.1) for t in g["ts"]], g["sine_tau03"])
    assert np.array_equal([L.ibo_poly_weight(t, 0.1) for t in g["ts"]], g["poly"])


def _trad_setup():
    t = load("trad_sphere.npz")
    geo = load("geom_sphere.npz")
    n = int(t["n"])
    prm = O.make_params(float(t["k"]), float(t["eps"]), float(t["tau"]))
    node, flat = O.tube(geo["level"], [geo[k] for k in ("gx", "gy", "gz", "mc", "gc")], n, prm, 0)
    return t, geo, n, prm, node, flat


def test_traditional_tube_matches_reference():
    t, geo, n, prm, node, flat = _trad_setup()
    cnt = len(node)
    assert cnt == int(t["size"])
    F = O.tube_fields(flat, cnt)
    for k in ("zx", "zy", "zz", "sx", "sy", "sz", "dist", "delta", "jac", "w"):
        assert np.array_equal(F[k], t[k]), k
    # the reference's traditional path normalises the interpolated gradient twice
    # (interpolate3D.c:99-101 then large3Dexample2.c:4407-4410); DoubleLayerKernel,
    # which wrote this fixture, once -- one ulp apart
    for k in ("nx", "ny", "nz"):
        np.testing.assert_allclose(F[k], t[k], rtol=0, atol=4.5e-16)
    # DoubleLayerKernel stores the interior sign of the regularisation (HelmholtzKernel.c:710)
    assert np.array_equal(F["reg"], -t["reg_interior"])


def test_traditional_kernel_entries_bitwise():
    t, geo, n, prm, node, flat = _trad_setup()
    are, aim = O.kernel_rows(flat, len(node), prm)
    pi, pj = t["pi"], t["pj"]
    assert np.array_equal(are[pi, pj], t["entry_re"])
    assert np.array_equal(aim[pi, pj], t["entry_im"])


def test_neumann_plane_wave_rhs_matches_text_output():
    t, geo, n, prm, node, flat = _trad_setup()
    cnt = len(node)
    F = O.tube_fields(flat, cnt)
    r = load("rhs_sphere.npz")
    # cal_partialu normalises the interpolated gradient a second time
    # (large3Dexample2.c:4738-4742), which is the tube normal
    gi = np.stack([F["nx"], F["ny"], F["nz"]]).copy()
    bre, bim = O.rhs_plane_neumann(flat, cnt, gi, 2.0, 10.0, [-1.0, 0.0, 0.0])
    # print_f writes %.15lf (large3Dexample2.c:1333): half a unit in the 15th decimal
    # plus the rounding of reading the decimal back (|b| < 32: 1.8e-15)
    np.testing.assert_allclose(bre, r["bre"], rtol=0, atol=5e-16 + 1.8e-15)
    np.testing.assert_allclose(bim, r["bim"], rtol=0, atol=5e-16 + 1.8e-15)


def test_combo_tube_and_kernel_bitwise():
    c = load("combo_sphere.npz")
    geo = load("geom_sphere.npz")
    n = int(c["n"])
    prm = O.make_params(float(c["k"]), float(c["eps"]), 2 * 4.0 / n, delta_in=float(c["delta"]),
                        eta=float(c["eta"]))
    node, flat = O.tube(geo["level"], [geo[k] for k in ("gx", "gy", "gz", "mc", "gc")], n, prm, 1)
    cnt = len(node)
    assert cnt == int(c["size"])
    F = O.tube_fields(flat, cnt)
    for k in ("zx", "sx", "sy", "sz", "dist", "nx", "ny", "nz", "jac", "reg", "w"):
        assert np.array_equal(F[k], c[k]), k
    assert np.array_equal(F["delta"], c["sinew"])
    are, aim, regs = O.kernel_rows(flat, cnt, prm, kind=1)
    assert regs == int(c["regcount"])
    assert np.array_equal(are[c["rows"]], c["row_re"])
    assert np.array_equal(aim[c["rows"]], c["row_im"])
    assert np.array_equal(are[c["pi"], c["pj"]], c["entry_re"])
    assert np.array_equal(aim[c["pi"], c["pj"]], c["entry_im"])


def test_bicg_serial_bitwise_and_omp_close():
    b = load("bicg.npz")
    x, xi, its, hist = O.bicg(b["are"], b["aim"], b["bre"], b["bim"])
    assert its == int(b["serial_its"])
    assert np.array_equal(x, b["serial_x"]) and np.array_equal(xi, b["serial_xi"])
    # the reference prints its final error with %.10E
    np.testing.assert_allclose(hist[-1], b["serial_err"][0], rtol=1e-9)
    x, xi, its, hist = O.bicg(b["are"], b["aim"], b["bre"], b["bim"], omp=True)
    # omp_inner's reduction order depends on the thread count (tools.c:21-31)
    assert abs(its - int(b["omp_its"])) <= 1
    np.testing.assert_allclose(x, b["omp_x"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(xi, b["omp_xi"], rtol=0, atol=1e-12)


def test_bicg_zero_rhs_gives_zero_density():
    b = load("bicg.npz")
    z = np.zeros(len(b["bre"]))
    x, xi, its, _ = O.bicg(b["are"], b["aim"], z, z)
    assert its == int(b["zero_its"]) == 0
    assert np.array_equal(x, b["zero_x"] * 0) and not np.any(xi)
