"""ctypes binding of the CPU oracle (oracle/liboracle.so), which checks the HIP library.

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(REPO, "oracle", "liboracle.so")
# IBIM_ORACLE_LIB: load another build of the same restatement (the sanitizer
# build, oracle/liboracle_asan.so, with libasan preloaded)
if os.environ.get("IBIM_ORACLE_LIB"):
    LIB_PATH = os.environ["IBIM_ORACLE_LIB"]

D = C.c_double
I = C.c_int
P = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)

# tube field order (oracle/ibim_oracle.h)
FIELDS = ["zx", "zy", "zz", "sx", "sy", "sz", "dist", "nx", "ny", "nz", "delta", "jac", "w", "reg", "hm", "gk"]
NF = len(FIELDS)


class Params(C.Structure):
    _fields_ = [("k", D), ("eps", D), ("tau", D), ("delta_in", D), ("eta", D), ("sine_root", D),
                ("interior", I), ("dirichlet", I), ("one_sided", I), ("combined", I)]


def make_params(k, eps, tau, delta_in=0.0, eta=-1.0, sine_root=0.1, interior=0, dirichlet=0, one_sided=0,
                combined=0):
    return Params(k, eps, tau, delta_in, eta, sine_root, interior, dirichlet, one_sided, combined)


def _p(a):
    return a.ctypes.data_as(P)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(REPO, "oracle")])
        L = C.CDLL(LIB_PATH)
        L.ibo_weno.argtypes = [P, D, I, P, P]
        L.ibo_weno_axis.argtypes = [P, I, D, I, I, I, I, I, P, P]
        L.ibo_gradient.argtypes = [P, I, D, I, P, P, P]
        L.ibo_d2.argtypes = [P, I, D, I, I, I, I]
        L.ibo_d2.restype = D
        L.ibo_dmixed.argtypes = [P, I, D, I, I, I, I]
        L.ibo_dmixed.restype = D
        L.ibo_curvatures.argtypes = [P, I, D, P, P]
        L.ibo_interp_grad.argtypes = [I, D, D, P, P, P, P, P]
        L.ibo_interp_grad.restype = I
        L.ibo_interp_curv.argtypes = [I, D, D, P, P, P, P]
        L.ibo_interp_dist.argtypes = [P, I, D, D, P]
        L.ibo_interp_dist.restype = D
        L.ibo_delta_cos.argtypes = [D, D]
        L.ibo_delta_cos.restype = D
        L.ibo_sine_weight.argtypes = [D, D, D]
        L.ibo_sine_weight.restype = D
        L.ibo_poly_weight.argtypes = [D, D]
        L.ibo_poly_weight.restype = D
        L.ibo_tube_count.argtypes = [P, I, C.POINTER(Params), I]
        L.ibo_tube_count.restype = I
        L.ibo_tube_build.argtypes = [P, P, P, P, P, P, I, D, D, C.POINTER(Params), I, I, IP, P]
        L.ibo_tube_build.restype = I
        L.ibo_kernel_trad_rows.argtypes = [I, P, C.POINTER(Params), I, I, P, P]
        L.ibo_kernel_combo_rows.argtypes = [I, P, C.POINTER(Params), I, I, P, P]
        L.ibo_kernel_combo_rows.restype = I
        for f in (L.ibo_bicg, L.ibo_bicg_omp):
            f.argtypes = [I, P, P, P, P, P, P, I, D, P]
            f.restype = I
        L.ibo_sweep_rows.argtypes = [I, I, P, P, P, P, P, P, P, P, P, P]
        L.ibo_rhs_point_source.argtypes = [I, P, D, P, P, P]
        L.ibo_rhs_plane_neumann.argtypes = [I, P, P, D, D, P, P, P]
        L.ibo_rhs_plane_dirichlet.argtypes = [I, P, D, D, P, P, P]
        L.ibo_eval.argtypes = [I, P, C.POINTER(Params), P, P, I, P, I, P, P]
        _lib = L
    return _lib


def weno(f, dx, order):
    f = np.ascontiguousarray(f, dtype=np.float64)
    a, b = np.zeros(1), np.zeros(1)
    lib().ibo_weno(_p(f), dx, order, _p(a), _p(b))
    return a[0], b[0]


def weno_axis(u, n, h, i, j, k, axis, order):
    a, b = np.zeros(1), np.zeros(1)
    lib().ibo_weno_axis(_p(u), n, h, i, j, k, axis, order, _p(a), _p(b))
    return a[0], b[0]


def geometry(u, n, order=5):
    """WENO/Godunov gradient and curvatures on the whole grid."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    h = 4.0 / n
    gx, gy, gz, mc, gc = (np.zeros_like(u) for _ in range(5))
    lib().ibo_gradient(_p(u), n, h, order, _p(gx), _p(gy), _p(gz))
    lib().ibo_curvatures(_p(u), n, h, _p(mc), _p(gc))
    return gx, gy, gz, mc, gc


def second_derivs(u, n, i, j, k):
    h = 4.0 / n
    L = lib()
    return np.array([L.ibo_d2(_p(u), n, h, i, j, k, 0), L.ibo_d2(_p(u), n, h, i, j, k, 1),
                     L.ibo_d2(_p(u), n, h, i, j, k, 2), L.ibo_dmixed(_p(u), n, h, i, j, k, 0),
                     L.ibo_dmixed(_p(u), n, h, i, j, k, 1), L.ibo_dmixed(_p(u), n, h, i, j, k, 2)])


def interp_grad(n, fields, p, lo=-2.0):
    gx, gy, gz = fields
    o = np.zeros(3)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    code = lib().ibo_interp_grad(n, 4.0 / n, lo, _p(gx), _p(gy), _p(gz), _p(pp), _p(o))
    return code, o


def interp_curv(n, mc, gc, p, lo=-2.0):
    o = np.zeros(2)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    lib().ibo_interp_curv(n, 4.0 / n, lo, _p(mc), _p(gc), _p(pp), _p(o))
    return o


def interp_dist(u, n, p, lo=-2.0):
    pp = np.ascontiguousarray(p, dtype=np.float64)
    return lib().ibo_interp_dist(_p(u), n, 4.0 / n, lo, _p(pp))


def tube(u, geom, n, prm, kind=0, lo=-2.0):
    """Build a tube; returns (node index array, dict of fields)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    gx, gy, gz, mc, gc = geom
    count = lib().ibo_tube_count(_p(u), n, C.byref(prm), kind)
    node = np.zeros(count, dtype=np.int32)
    out = np.zeros(NF * count)
    rc = lib().ibo_tube_build(_p(u), _p(gx), _p(gy), _p(gz), _p(mc), _p(gc), n, 4.0 / n, lo, C.byref(prm), kind,
                              count, node.ctypes.data_as(IP), _p(out))
    if rc != count:
        raise RuntimeError("tube build failed (%d)" % rc)
    return node, out


def tube_fields(flat, count):
    return {f: flat[i * count:(i + 1) * count] for i, f in enumerate(FIELDS)}


def kernel_rows(flat, count, prm, row0=0, nrows=None, kind=0):
    nrows = count - row0 if nrows is None else nrows
    are = np.zeros((nrows, count))
    aim = np.zeros((nrows, count))
    f = lib().ibo_kernel_trad_rows if kind == 0 else lib().ibo_kernel_combo_rows
    r = f(count, _p(flat), C.byref(prm), row0, nrows, _p(are), _p(aim))
    return (are, aim) if kind == 0 else (are, aim, r)


def bicg(are, aim, bre, bim, maxit=1000, tol=1e-7, omp=False):
    n = len(bre)
    are = np.ascontiguousarray(are)
    aim = np.ascontiguousarray(aim)
    bre = np.ascontiguousarray(bre, dtype=np.float64)
    bim = np.ascontiguousarray(bim, dtype=np.float64)
    x, xi, hist = np.zeros(n), np.zeros(n), np.zeros(maxit)
    f = lib().ibo_bicg_omp if omp else lib().ibo_bicg
    its = f(n, _p(are), _p(aim), _p(bre), _p(bim), _p(x), _p(xi), maxit, tol, _p(hist))
    return x, xi, its, hist[:its]


def sweep_rows(are, aim, p, q):
    """v = A p over a row block, w = A^H q restricted to it (reference loop nests)."""
    nr, nc = are.shape
    p = np.asarray(p, dtype=np.complex128)
    q = np.asarray(q, dtype=np.complex128)
    ins = [np.ascontiguousarray(a) for a in (p.real, p.imag, q.real, q.imag)]
    vr, vi, wr, wi = np.zeros(nr), np.zeros(nr), np.zeros(nc), np.zeros(nc)
    lib().ibo_sweep_rows(nr, nc, _p(are), _p(aim), *[_p(a) for a in ins], _p(vr), _p(vi), _p(wr), _p(wi))
    return vr + 1j * vi, wr + 1j * wi


def rhs_point_source(flat, count, k, x0):
    bre, bim = np.zeros(count), np.zeros(count)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    lib().ibo_rhs_point_source(count, _p(flat), k, _p(x0), _p(bre), _p(bim))
    return bre, bim


def rhs_plane_neumann(flat, count, interp_n, k, amp, direction):
    bre, bim = np.zeros(count), np.zeros(count)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    nn = np.ascontiguousarray(interp_n, dtype=np.float64)
    lib().ibo_rhs_plane_neumann(count, _p(flat), _p(nn), k, amp, _p(d), _p(bre), _p(bim))
    return bre, bim


def rhs_plane_dirichlet(flat, count, k, amp, direction):
    bre, bim = np.zeros(count), np.zeros(count)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    lib().ibo_rhs_plane_dirichlet(count, _p(flat), k, amp, _p(d), _p(bre), _p(bim))
    return bre, bim


def evaluate(flat, count, prm, phre, phim, xyz, mode):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    m = len(xyz)
    ure, uim = np.zeros(m), np.zeros(m)
    lib().ibo_eval(count, _p(flat), C.byref(prm), _p(np.ascontiguousarray(phre)), _p(np.ascontiguousarray(phim)),
                   m, _p(xyz), mode, _p(ure), _p(uim))
    return ure, uim
