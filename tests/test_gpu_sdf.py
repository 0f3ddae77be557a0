"""Signed-distance level sets built on the GPU (ibim_set_level_shape) against
closed forms (sphere, torus) and an independent CPU closest-point search
(ellipsoid, star-shaped surfaces): dense parametric sampling, then a bounded
quasi-Newton refinement with scipy. Tolerance 1e-8 on the distance at nodes
within 0.3 of the surface."""
import numpy as np
import pytest
from scipy.optimize import minimize

import shapes

pytestmark = pytest.mark.gpu


def surface(kind, th, ph, a, b, c, m, n, variant):
    st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    if kind == "ellipsoid":
        return np.stack([a * st * cp, b * st * sp, c * ct], -1)
    g = np.sin(m * th) * np.cos(n * ph) if variant == 0 else np.cos(m * th) * np.sin(n * ph)
    r = a + b * g
    return np.stack([r * st * cp, r * st * sp, r * ct], -1)


def brute_distance(x, kind, **kw):
    th = np.linspace(0, np.pi, 401)
    ph = np.linspace(0, 2 * np.pi, 801)
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    S = surface(kind, T, Pp, **kw)
    d2 = np.sum((S - x) ** 2, axis=-1)
    i = np.unravel_index(np.argmin(d2), d2.shape)
    f = lambda v: np.sum((surface(kind, v[0], v[1], **kw) - x) ** 2)
    res = minimize(f, [T[i], Pp[i]], method="L-BFGS-B", bounds=[(0, np.pi), (-np.pi, 3 * np.pi)],
                   options={"ftol": 1e-30, "gtol": 1e-14, "maxiter": 500})
    return np.sqrt(min(res.fun, d2[i]))


def test_sphere_and_torus_closed_forms(gpu):
    n = 32
    pb = gpu.Problem()
    pb.set_level_shape(gpu.shape(gpu.SHAPE_SPHERE, a=1.0), n)
    assert np.abs(pb.level() - shapes.sphere(n)).max() < 1e-15
    pb.set_level_shape(gpu.shape(gpu.SHAPE_TORUS, a=1.0, b=0.4), n)
    assert np.abs(pb.level() - shapes.torus(n)).max() < 1e-15


@pytest.mark.parametrize("kind,kw", [
    ("ellipsoid", dict(a=1.1, b=0.8, c=0.65, m=0, n=0, variant=0)),
    ("star", dict(a=1.0, b=0.15, c=0.0, m=3, n=2, variant=0)),   # perturbed sphere r = 1 + 0.15 sin3t cos2p
    ("star", dict(a=1.0, b=0.2, c=0.0, m=4, n=3, variant=1)),    # star r = 1 + 0.2 cos4t sin3p
])
def test_parametric_surfaces_match_brute_force(gpu, kind, kw):
    n = 40
    code = gpu.SHAPE_ELLIPSOID if kind == "ellipsoid" else gpu.SHAPE_STAR
    pb = gpu.Problem()
    pb.set_level_shape(gpu.shape(code, **kw), n)
    d = pb.level()
    x, y, z = shapes.grid_coords(n)
    near = np.argwhere(np.abs(d) < 0.3)
    rng = np.random.default_rng(0)
    for idx in near[rng.choice(len(near), 40, replace=False)]:
        p = np.array([x[tuple(idx)], y[tuple(idx)], z[tuple(idx)]])
        ref = brute_distance(p, kind, **kw)
        assert abs(abs(d[tuple(idx)]) - ref) < 1e-8, (idx, d[tuple(idx)], ref)
    # inside is positive: the centre is inside, a corner outside
    assert d[n // 2, n // 2, n // 2] > 0 and d[0, 0, 0] < 0
    # a signed distance has unit gradient near the surface
    pb.geometry(5)
    gx, gy, gz, mc, gc = pb.grid_fields()
    band = np.abs(d) < 2 * 4.0 / n
    gn = np.sqrt(gx ** 2 + gy ** 2 + gz ** 2)[band]
    assert np.median(np.abs(gn - 1)) < 0.01
