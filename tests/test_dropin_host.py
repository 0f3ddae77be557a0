"""The per-node drop-in entries of libibim.so that run on the host (weno,
WENO_DX/DY/DZ, BothCurvs, MeanCurv, GaussianCurv) against the reference's golden
fixtures. These entry points make no HIP call, so this runs without a GPU."""
import os

import numpy as np
import pytest

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ib():
    import ibim
    ibim.lib()
    return ibim


def test_weno_dropin_bitwise(ib):
    g = np.load(os.path.join(G, "weno.npz"))
    for order, key in ((3, "out3"), (5, "out5")):
        r = np.array([ib.weno(f if order == 5 else f[1:6], float(g["dx"]), order) for f in g["stencils"]])
        assert np.array_equal(r, g[key])


@pytest.mark.parametrize("name", ["geom_sphere.npz", "geom_torus.npz", "geom_twin.npz"])
def test_weno_axis_and_curvature_dropins_bitwise(ib, name):
    g = np.load(os.path.join(G, name))
    n = int(g["n"])
    h = 4.0 / n
    grid = ib.Grid3(g["level"])
    L = ib.lib()
    for m, (i, j, k) in enumerate(g["nodes"]):
        i, j, k = int(i), int(j), int(k)
        for ax in range(3):
            assert ib.WENO_D(ax, grid, i, j, k, n, h, 5) == tuple(g["weno_axis"][m, ax])
        mc, gc = ib.BothCurvs(grid, i, j, k, n, h)
        assert mc == g["mc"][i, j, k] and gc == g["gc"][i, j, k]
        assert L.MeanCurv(grid.ptr, i, j, k, n, h) == mc
        assert L.GaussianCurv(grid.ptr, i, j, k, n, h) == gc
