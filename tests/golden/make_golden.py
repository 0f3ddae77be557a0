"""Write the golden fixtures of tests/golden/ from the compiled reference.

Run in the development container only (it needs oracle/_ref/libibimref.so,
built by `make -C oracle ref` from the reference's own C sources):

    python tests/golden/make_golden.py

Each fixture group runs in its own subprocess, because the reference keeps
static state between calls (omp_get_Ccomp builds its tube once per process,
large3Dexample2.c:4298-4459) and calls exit(0) on its error paths.

Fixtures (inputs and the reference's outputs, float64):
  weno.npz           weno() on random / polynomial / kinked stencils, orders 3 and 5
  geom_<name>.npz    ASSIGN_GRADIENT (WENO-5) + BothCurvs on a whole grid
  geom3_sphere.npz   the same with WENO-3
  interp_<name>.npz  cal_interpgrad / cal_BothCurvs / cal_dist at sampled points
  trad_sphere.npz    traditional tube (via DoubleLayerKernel) and omp_get_Ccomp
                     entries with rho > tau
  combo_sphere.npz   new_HelmholtzKernelCombo tube, three rows and sampled entries
  newkernel.npz      new_HelmholtzKernel rows (analytic sphere variant)
  bicg.npz           CompBiCG / omp_CompBiCG on a seeded complex matrix
  rhs_sphere.npz     print_f (Neumann plane-wave data, %.15lf text)
  prims.npz          cal_delta, SineWeight, PolyWeight, cal_phi, phiimg
"""
import ctypes as C
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(REPO, "oracle", "_ref", "libibimref.so")
LO = -2.0

D = C.c_double
I = C.c_int
P = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


def ptr(a):
    return a.ctypes.data_as(P)


def iptr(a):
    return a.ctypes.data_as(IP)


def level_set(name, n):
    """Signed distance (positive inside) on the (n+1)^3 grid, [i=y][j=x][k=z]."""
    h = 4.0 / n
    c = LO + h * np.arange(n + 1)
    y, x, z = np.meshgrid(c, c, c, indexing="ij")
    if name == "sphere":
        return 1.0 - np.sqrt(x * x + y * y + z * z)
    if name == "torus":
        return 0.4 - np.sqrt((np.sqrt(x * x + y * y) - 1.0) ** 2 + z * z)
    if name == "twin":
        a = 0.7 - np.sqrt((x + 0.5) ** 2 + y * y + z * z)
        b = 0.6 - np.sqrt((x - 0.45) ** 2 + (y - 0.1) ** 2 + z * z)
        return np.maximum(a, b)
    raise ValueError(name)


def load():
    lib = C.CDLL(LIB)
    lib.ref_setup.argtypes = [I, D, P]
    lib.ref_geometry.argtypes = [I, P, P, P, P, P]
    lib.ref_set_gradient.argtypes = [P, P, P]
    lib.ref_set_curvatures.argtypes = [P, P]
    lib.ref_second_derivs.argtypes = [I, I, I, P]
    lib.ref_weno.argtypes = [P, D, I, P]
    lib.ref_weno_axis.argtypes = [I, I, I, I, I, P]
    lib.ref_interpgrad.argtypes = [D, D, D, P]
    lib.ref_interpgrad.restype = I
    lib.ref_bothcurvs_at.argtypes = [D, D, D, P]
    lib.ref_cal_dist.argtypes = [D, D, D]
    lib.ref_cal_dist.restype = D
    for f in ("ref_delta", "ref_sine_weight", "ref_poly_weight"):
        getattr(lib, f).argtypes = [D, D]
        getattr(lib, f).restype = D
    lib.ref_cal_phi.argtypes = [P, P]
    lib.ref_cal_phi.restype = D
    lib.ref_phiimg.argtypes = [P, P]
    lib.ref_phiimg.restype = D
    lib.ref_trad_entries.argtypes = [I, D, I, IP, IP, P, P]
    lib.ref_double_layer.argtypes = [I, D] + [P] * 17
    lib.ref_double_layer.restype = I
    lib.ref_combo.argtypes = [I, D, D, D, D] + [P] * 17
    lib.ref_combo.restype = I
    lib.ref_new_kernel.argtypes = [I, D, D] + [P] * 16
    lib.ref_new_kernel.restype = I
    lib.ref_compbicg.argtypes = [I, P, P, P, P, P, P, I, P]
    lib.ref_compbicg.restype = I
    lib.ref_print_f.argtypes = [I, D, C.c_char_p, C.c_char_p]
    lib.ref_print_f.restype = I
    return lib


def setup_geometry(lib, name, n, k, order=5):
    u = np.ascontiguousarray(level_set(name, n))
    lib.ref_setup(n, k, ptr(u))
    g = [np.zeros_like(u) for _ in range(5)]
    lib.ref_geometry(order, *[ptr(a) for a in g])
    return u, g


def out(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name), **arrays)
    print("wrote", name, {k: v.shape for k, v in arrays.items()})


def g_weno():
    lib = load()
    rng = np.random.default_rng(1)
    stencils = []
    stencils += list(rng.standard_normal((80, 7)))
    xs = np.arange(-3, 4, dtype=float)
    for deg in range(5):
        for _ in range(8):
            c = rng.standard_normal(deg + 1)
            stencils.append(np.polyval(c, xs * 0.1))
    for _ in range(40):  # kinks
        a, b, s = rng.standard_normal(3)
        stencils.append(np.abs(xs * 0.1 - s * 0.2) * a + b)
    stencils.append(np.zeros(7))
    st = np.array(stencils)
    dx = 0.0625
    res = np.zeros((2, len(st), 2))
    for oi, order in enumerate((3, 5)):
        for m, f in enumerate(st):
            f7 = np.ascontiguousarray(f if order == 5 else f[1:6])
            o = np.zeros(2)
            lib.ref_weno(ptr(f7), dx, order, ptr(o))
            res[oi, m] = o
    out("weno.npz", stencils=st, dx=np.array(dx), out3=res[0], out5=res[1])


def g_geom(name, n, order=5):
    lib = load()
    u, (gx, gy, gz, mc, gc) = setup_geometry(lib, name, n, 2.0, order)
    # per-axis WENO at a few nodes including the mirror boundary
    rng = np.random.default_rng(2)
    nodes = np.concatenate([rng.integers(0, n + 1, size=(60, 3)), [[0, 0, 0], [n, n, n], [1, n - 1, 2]]])
    wa = np.zeros((len(nodes), 3, 2))
    d2 = np.zeros((len(nodes), 6))
    for m, (i, j, k) in enumerate(nodes):
        for ax in range(3):
            o = np.zeros(2)
            lib.ref_weno_axis(ax, int(i), int(j), int(k), order, ptr(o))
            wa[m, ax] = o
        o6 = np.zeros(6)
        lib.ref_second_derivs(int(i), int(j), int(k), ptr(o6))
        d2[m] = o6
    tag = "geom%s_%s.npz" % ("" if order == 5 else "3", name)
    out(tag, n=np.array(n), level=u, gx=gx, gy=gy, gz=gz, mc=mc, gc=gc, nodes=nodes, weno_axis=wa, d2=d2)


def g_interp(name, n):
    lib = load()
    u, _ = setup_geometry(lib, name, n, 2.0)
    rng = np.random.default_rng(3)
    h = 4.0 / n
    # points in the interior cells, away from the outer boundary
    pts = rng.uniform(LO + 3 * h, -LO - 3 * h, size=(1000, 3))
    gi = np.zeros((len(pts), 3))
    code = np.zeros(len(pts), dtype=np.int32)
    cv = np.zeros((len(pts), 2))
    dist = np.zeros(len(pts))
    for m, p in enumerate(pts):
        o = np.zeros(3)
        code[m] = lib.ref_interpgrad(p[0], p[1], p[2], ptr(o))
        gi[m] = o
        o2 = np.zeros(2)
        lib.ref_bothcurvs_at(p[0], p[1], p[2], ptr(o2))
        cv[m] = o2
        dist[m] = lib.ref_cal_dist(p[0], p[1], p[2])
    out("interp_%s.npz" % name, n=np.array(n), points=pts, grad=gi, code=code, curv=cv, dist=dist)


def tube_count(u, eps):
    return int(np.count_nonzero(~(np.abs(u) > eps)))


def g_trad():
    lib = load()
    n, k = 16, 2.0
    h = 4.0 / n
    eps = 2.0 * h  # FACTOR, PDE.h:16
    tau = 2.0 * h  # TAU_FACTOR, PDE.h:29
    u, _ = setup_geometry(lib, "sphere", n, k)
    size = tube_count(u, eps)
    arrs = [np.zeros(size) for _ in range(15)]
    kre = np.zeros((size, size))
    kim = np.zeros((size, size))
    lib.ref_double_layer(size, eps, *[ptr(a) for a in arrs], ptr(kre), ptr(kim))
    zx, zy, zz, sx, sy, sz, dist, gx, gy, gz, delta, jac, reg, regimg, result = arrs
    s = np.stack([sx, sy, sz], 1)
    rng = np.random.default_rng(4)
    pi = rng.integers(0, size, 3000)
    pj = rng.integers(0, size, 3000)
    rows = np.array([0, size // 2, size - 1])
    pi = np.concatenate([np.repeat(rows, size), pi])
    pj = np.concatenate([np.tile(np.arange(size), 3), pj])
    rho = np.linalg.norm(s[pi] - s[pj], axis=1)
    keep = rho > tau * (1 + 1e-9)  # near-field entries index a stale loop variable
    pi = np.ascontiguousarray(pi[keep], dtype=np.int32)
    pj = np.ascontiguousarray(pj[keep], dtype=np.int32)
    re = np.zeros(len(pi))
    im = np.zeros(len(pi))
    lib.ref_trad_entries(size, eps, len(pi), iptr(pi), iptr(pj), ptr(re), ptr(im))
    out("trad_sphere.npz", n=np.array(n), k=np.array(k), eps=np.array(eps), tau=np.array(tau), size=np.array(size),
        zx=zx, zy=zy, zz=zz, sx=sx, sy=sy, sz=sz, dist=dist, nx=gx, ny=gy, nz=gz, delta=delta, jac=jac,
        reg_interior=reg, w=result, pi=pi, pj=pj, entry_re=re, entry_im=im,
        dl_rows=rows, dl_re=kre[rows], dl_im=kim[rows])


def g_rhs():
    lib = load()
    n, k = 16, 2.0
    h = 4.0 / n
    eps = 2.0 * h
    u, _ = setup_geometry(lib, "sphere", n, k)
    size = tube_count(u, eps)
    with tempfile.TemporaryDirectory() as td:
        a, b = os.path.join(td, "re.txt"), os.path.join(td, "im.txt")
        lib.ref_print_f(size, eps, a.encode(), b.encode())
        re = np.loadtxt(a)
        im = np.loadtxt(b)
        # the reference's own output files, kept verbatim for the format tests
        for src, dst in ((a, "Rel_bN16k2.txt"), (b, "Rel_bimgN16k2.txt")):
            with open(src) as f_in, open(os.path.join(HERE, dst), "w") as f_out:
                f_out.write(f_in.read())
    assert int(re[0]) == size
    out("rhs_sphere.npz", n=np.array(n), k=np.array(k), eps=np.array(eps), bre=re[1:], bim=im[1:])


def g_combo():
    lib = load()
    n, k, eta = 16, 2.0, -1.0
    h = 4.0 / n
    eps = 3.0 * h
    delta = 0.1 * eps
    u, _ = setup_geometry(lib, "sphere", n, k)
    size = int(np.count_nonzero(~((np.abs(u) > eps) | (np.abs(u) < delta))))
    arrs = [np.zeros(size) for _ in range(15)]
    kre = np.zeros((size, size))
    kim = np.zeros((size, size))
    regs = lib.ref_combo(size, delta, eps, k, eta, *[ptr(a) for a in arrs], ptr(kre), ptr(kim))
    zx, zy, zz, sx, sy, sz, dist, gx, gy, gz, jac, reg, regimg, sinew, siner = arrs
    rows = np.array([0, size // 2, size - 1])
    rng = np.random.default_rng(5)
    pi = rng.integers(0, size, 4000)
    pj = rng.integers(0, size, 4000)
    out("combo_sphere.npz", n=np.array(n), k=np.array(k), eta=np.array(eta), eps=np.array(eps),
        delta=np.array(delta), size=np.array(size), regcount=np.array(regs),
        zx=zx, zy=zy, zz=zz, sx=sx, sy=sy, sz=sz, dist=dist, nx=gx, ny=gy, nz=gz, jac=jac, reg=reg,
        sinew=sinew, w=siner, rows=rows, row_re=kre[rows], row_im=kim[rows], pi=pi, pj=pj,
        entry_re=kre[pi, pj], entry_im=kim[pi, pj])


def g_newkernel():
    lib = load()
    n, k = 16, 2.0
    h = 4.0 / n
    eps, delta = 3.0 * h, 0.3 * h
    u = level_set("sphere", n)
    lib.ref_setup(n, k, ptr(np.ascontiguousarray(u)))
    size = int(np.count_nonzero(~((np.abs(u) > eps) | (np.abs(u) < delta))))
    arrs = [np.zeros(size) for _ in range(14)]
    kre = np.zeros((size, size))
    kim = np.zeros((size, size))
    lib.ref_new_kernel(size, delta, eps, *[ptr(a) for a in arrs], ptr(kre), ptr(kim))
    rows = np.array([0, size // 2, size - 1])
    names = ["zx", "zy", "zz", "sx", "sy", "sz", "dist", "nx", "ny", "nz", "curv", "jac", "sinew", "w"]
    out("newkernel.npz", n=np.array(n), k=np.array(k), eps=np.array(eps), delta=np.array(delta),
        size=np.array(size), rows=rows, row_re=kre[rows], row_im=kim[rows],
        **{nm: a for nm, a in zip(names, arrs)})


def g_bicg():
    lib = load()
    rng = np.random.default_rng(6)
    m = 120
    are = rng.standard_normal((m, m)) / np.sqrt(m)
    aim = rng.standard_normal((m, m)) / np.sqrt(m)
    are += np.eye(m) * 2.0
    bre = rng.standard_normal(m)
    bim = rng.standard_normal(m)
    res = {}
    for tag, omp in (("serial", 0), ("omp", 1)):
        x = np.zeros(m)
        xi = np.zeros(m)
        err = np.zeros(1)
        its = lib.ref_compbicg(m, ptr(are), ptr(aim), ptr(bre), ptr(bim), ptr(x), ptr(xi), omp, ptr(err))
        res[tag + "_x"] = x
        res[tag + "_xi"] = xi
        res[tag + "_its"] = np.array(its)
        res[tag + "_err"] = err
    # zero right-hand side -> zero density (BICGSTAB.c:147-156)
    z = np.zeros(m)
    x = np.ones(m)
    xi = np.ones(m)
    err = np.zeros(1)
    its0 = lib.ref_compbicg(m, ptr(are), ptr(aim), ptr(z), ptr(z), ptr(x), ptr(xi), 0, ptr(err))
    out("bicg.npz", are=are, aim=aim, bre=bre, bim=bim, zero_its=np.array(its0), zero_x=x, **res)


def g_prims():
    lib = load()
    lib.ref_setup(16, 3.0, ptr(np.zeros((17, 17, 17))))
    rng = np.random.default_rng(7)
    xs = rng.uniform(-1.2, 1.2, 400)
    ts = np.concatenate([rng.uniform(-0.1, 1.1, 400), [0.1, 0.2, 1.0, 0.0]])
    delta = np.array([lib.ref_delta(0.5, x) for x in xs])
    sine = np.array([lib.ref_sine_weight(t, 0.1) for t in ts])
    sine2 = np.array([lib.ref_sine_weight(t, 0.3) for t in ts])
    poly = np.array([lib.ref_poly_weight(t, 0.1) for t in ts])
    a = rng.uniform(-1, 1, (200, 3))
    b = rng.uniform(-1, 1, (200, 3))
    phi = np.array([lib.ref_cal_phi(ptr(np.ascontiguousarray(p)), ptr(np.ascontiguousarray(q))) for p, q in zip(a, b)])
    phim = np.array([lib.ref_phiimg(ptr(np.ascontiguousarray(p)), ptr(np.ascontiguousarray(q))) for p, q in zip(a, b)])
    out("prims.npz", xs=xs, ts=ts, delta=delta, sine=sine, sine_tau03=sine2, poly=poly, a=a, b=b, k=np.array(3.0),
        phi=phi, phiimg=phim)


GROUPS = {
    "weno": g_weno,
    "geom_sphere": lambda: g_geom("sphere", 16),
    "geom_torus": lambda: g_geom("torus", 20),
    "geom_twin": lambda: g_geom("twin", 16),
    "geom3_sphere": lambda: g_geom("sphere", 16, order=3),
    "interp_sphere": lambda: g_interp("sphere", 16),
    "interp_twin": lambda: g_interp("twin", 16),
    "trad": g_trad,
    "rhs": g_rhs,
    "combo": g_combo,
    "newkernel": g_newkernel,
    "bicg": g_bicg,
    "prims": g_prims,
}

if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--group":
        GROUPS[sys.argv[2]]()
        sys.exit(0)
    if not os.path.exists(LIB):
        sys.exit("build the reference first: make -C oracle ref")
    for g in (sys.argv[1:] or GROUPS):
        rc = subprocess.call([sys.executable, __file__, "--group", g])
        if rc:
            sys.exit("group %s failed (%d)" % (g, rc))
