"""Synthetic signed-distance level sets (positive inside) on the (n+1)^3 grid over
[-2, 2]^3, layout [i=y][j=x][k=z], and receiver sets -- shared by tests, smoke and
the benchmark."""
import numpy as np

LO = -2.0


def grid_coords(n):
    h = 4.0 / n
    c = LO + h * np.arange(n + 1)
    y, x, z = np.meshgrid(c, c, c, indexing="ij")
    return x, y, z


def sphere(n, radius=1.0, centre=(0.0, 0.0, 0.0)):
    x, y, z = grid_coords(n)
    return radius - np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2)


def torus(n, R=1.0, r=0.4):
    x, y, z = grid_coords(n)
    return r - np.sqrt((np.sqrt(x * x + y * y) - R) ** 2 + z * z)


def twin(n):
    x, y, z = grid_coords(n)
    a = 0.7 - np.sqrt((x + 0.5) ** 2 + y * y + z * z)
    b = 0.6 - np.sqrt((x - 0.45) ** 2 + (y - 0.1) ** 2 + z * z)
    return np.maximum(a, b)


def shell_receivers(m, rmin, rmax, seed=0):
    """m points with uniform directions and |x| uniform in [rmin, rmax]."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((m, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v * rng.uniform(rmin, rmax, (m, 1))


def point_source(xyz, x0, k):
    """e^{ik|x-x0|} / (4 pi |x-x0|)."""
    r = np.linalg.norm(np.asarray(xyz) - np.asarray(x0), axis=-1)
    return np.exp(1j * k * r) / (4 * np.pi * r)
