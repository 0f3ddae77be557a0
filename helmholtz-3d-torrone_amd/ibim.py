"""Python host side of libibim.so (the gfx950 IBIM library), over its C ABI.

Two layers, like include/ibim.h:

* ``Problem`` -- the device-resident pipeline: level set -> geometry -> tube ->
  operator -> right-hand sides -> solve -> field evaluation, every array kept in
  HBM between calls.
* Reference-named helpers (``CompBiCG``, ``omp_CompBiCG``, ``ASSIGN_GRADIENT``,
  ``BothCurvs``, ``weno``, ``new_HelmholtzKernelCombo``, ``DoubleLayerKernel``,
  ``new_HelmholtzKernel``) that call the drop-in entry points with the
  reference's own data structures (double*** grids, double** matrices) built
  from numpy arrays, and return numpy results.

Loading fails loudly when libibim.so has not been built: there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libibim.so")

D, I, L = C.c_double, C.c_int, C.c_long
P = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
PP = C.POINTER(P)
PPP = C.POINTER(PP)

FIELDS = ["zx", "zy", "zz", "sx", "sy", "sz", "dist", "nx", "ny", "nz", "delta", "jac", "w", "reg", "hm", "gk"]
NF = len(FIELDS)

TUBE_TRADITIONAL, TUBE_COMBO = 0, 1
C128, C64 = 0, 1
# pair storage of the traditional operator (half the bytes per sweep)
C128_PAIR, C64_PAIR = 2, 3
# complex64 pairs + a complex128 copy for the FP64 residuals of the mixed solvers
MIXED_PAIR = 4
RHS_POINT_SOURCE, RHS_PLANE_NEUMANN, RHS_PLANE_DIRICHLET = 0, 1, 2
EVAL_SINGLE, EVAL_DOUBLE, EVAL_COMBO, EVAL_COMBINED = 0, 1, 2, 3
SOLVE_BICG, SOLVE_REFINED, SOLVE_RELIABLE, SOLVE_GMRES, SOLVE_BLOCK_GMRES = 0, 1, 2, 3, 4
OP_STORED, OP_FREE64, OP_FREE32 = 0, 1, 2


PROGRESS_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_int, C.c_double, C.c_void_p)


class IbimError(RuntimeError):
    pass


class Params(C.Structure):
    """ibim_params: run-time form of the reference's PDE.h switches."""
    _fields_ = [("k", D), ("eps", D), ("tau", D), ("delta_in", D), ("eta", D), ("sine_root", D),
                ("interior", I), ("dirichlet", I), ("one_sided", I), ("combined", I)]


SHAPE_SPHERE, SHAPE_TORUS, SHAPE_ELLIPSOID, SHAPE_STAR = 0, 1, 2, 3


class Shape(C.Structure):
    """ibim_shape: an analytic surface whose signed distance the GPU builds."""
    _fields_ = [("kind", I), ("variant", I), ("m", I), ("n", I), ("center", D * 3), ("a", D), ("b", D), ("c", D)]


def shape(kind, a=1.0, b=0.0, c=0.0, m=0, n=0, variant=0, center=(0.0, 0.0, 0.0)):
    return Shape(kind, variant, m, n, (D * 3)(*center), a, b, c)


class SolveOpts(C.Structure):
    _fields_ = [("method", I), ("maxit", I), ("tol", D), ("inner_tol", D), ("max_outer", I), ("chunk", I),
                ("batch", I)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", I), ("outer", I), ("rel_residual", D), ("status", I)]


def params(k, eps, tau, delta_in=0.0, eta=-1.0, sine_root=0.1, interior=0, dirichlet=0, one_sided=0, combined=0):
    """combined=1: the combined-field exterior Dirichlet operator 1/2 I + K - i eta S
    (dirichlet=1, interior=0, traditional tube); eta = k is the usual coupling."""
    return Params(k, eps, tau, delta_in, eta, sine_root, interior, dirichlet, one_sided, combined)


_lib = None


def lib():
    """The loaded libibim.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IbimError("libibim.so is not built (make -C helmholtz-3d-torrone_amd)")
        L_ = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "ibim_device_count": ([], I), "ibim_last_error": ([], C.c_char_p),
            "ibim_create": ([C.POINTER(vp)], I), "ibim_destroy": ([vp], None), "ibim_sync": ([vp], I),
            "ibim_set_level_host": ([vp, I, D, D, P], I), "ibim_set_level_device": ([vp, I, D, D, vp], I),
            "ibim_set_level_shape": ([vp, I, D, D, C.POINTER(Shape)], I),
            "ibim_geometry": ([vp, I], I), "ibim_redistance": ([vp, D], I), "ibim_get_level": ([vp, P], I), "ibim_get_grid": ([vp, P, P, P, P, P], I),
            "ibim_set_grid_fields_host": ([vp, P, P, P, P, P], I),
            "ibim_tube": ([vp, C.POINTER(Params), I, IP], I), "ibim_get_tube": ([vp, IP, P], I),
            "ibim_tube_device": ([vp], vp),
            "ibim_assemble": ([vp, I], I), "ibim_get_rows": ([vp, I, I, P, P], I),
            "ibim_set_matrix_host": ([vp, I, P, P, I], I), "ibim_regcount": ([vp], I),
            "ibim_apply": ([vp, I, P, P, P, P, P, P, P, P], I),
            "ibim_apply_batch": ([vp, I, I, I, P, P, P, P], I),
            "ibim_apply_noalloc": ([vp, I, vp, vp, vp, vp], I),
            "ibim_set_nrhs": ([vp, I], I), "ibim_rhs": ([vp, I, P, D, P, I], I),
            "ibim_set_rhs_host": ([vp, I, P, P], I), "ibim_get_rhs": ([vp, I, P, P], I),
            "ibim_solve": ([vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo), P], I),
            "ibim_solve_group": ([C.POINTER(vp), I, C.POINTER(SolveOpts), C.POINTER(SolveInfo)], I),
            "ibim_get_density": ([vp, I, P, P], I), "ibim_set_density_host": ([vp, I, P, P], I),
            "ibim_evaluate": ([vp, I, I, I, P, P, P], I),
            "ibim_set_receivers": ([vp, I, P, P, P], I),
            "ibim_gradient": ([vp, I, I, C.POINTER(SolveOpts), P, P, P], I),
            "ibim_set_profiling": ([vp, I], I),
            "ibim_set_graphs": ([vp, I], I),
            "ibim_set_fp32_products": ([vp, I], I), "ibim_set_free_cache": ([vp, D], I),
            "ibim_set_progress": ([vp, PROGRESS_FN, vp], I), "ibim_kernel_time": ([vp, C.c_char_p, P, C.POINTER(L)], I),
            "ibim_reset_timers": ([vp], I),
            "ibim_read_distance_file": ([C.c_char_p, I, P], I), "ibim_write_distance_file": ([C.c_char_p, I, P], I),
            "ibim_read_rhs_text": ([C.c_char_p, C.c_char_p, I, P, P], I),
            "ibim_write_rhs_text": ([C.c_char_p, C.c_char_p, I, P, P], I),
            "ibim_read_density_text": ([C.c_char_p, C.c_char_p, I, P, P], I),
            "ibim_write_density_text": ([C.c_char_p, C.c_char_p, I, P, P], I),
            "weno": ([P, P, P, D, I], None),
            "WENO_DX": ([PPP, I, I, I, I, D, P, P, I], None),
            "WENO_DY": ([PPP, I, I, I, I, D, P, P, I], None),
            "WENO_DZ": ([PPP, I, I, I, I, D, P, P, I], None),
            "GaussianCurv": ([PPP, I, I, I, I, D], D), "MeanCurv": ([PPP, I, I, I, I, D], D),
            "BothCurvs": ([PPP, I, I, I, I, D, P, P], I),
            "ASSIGN_GRADIENT": ([I, D, PPP, PPP, PPP, PPP, I], I),
            "CompBiCG": ([I, P, P, PP, PP, P, P], I),
            "omp_CompBiCG": ([I, P, P, PP, PP, P, P, I], I),
            "new_HelmholtzKernel": ([I, D, D] + [P] * 16 + [PP] * 4, I),
            "new_HelmholtzKernelCombo": ([I, D, D] + [PPP] * 6 + [P] * 17 + [PP] * 4 + [D, D], I),
            "DoubleLayerKernel": ([I, D] + [PPP] * 6 + [P] * 15 + [PP] * 2 + [D], I),
        }
        for name, (args, res) in sig.items():
            f = getattr(L_, name)
            f.argtypes = args
            f.restype = res
        _lib = L_
    return _lib


def device_count():
    return lib().ibim_device_count()


def _p(a):
    return a.ctypes.data_as(P)


def _check(rc, what):
    if rc < 0:
        raise IbimError("%s failed (%d): %s" % (what, rc, lib().ibim_last_error().decode()))
    return rc


class Problem:
    """One device-resident IBIM problem (a handle of the C API)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().ibim_create(C.byref(h)), "ibim_create")
        self.h = h
        self.n = 0
        self.count = 0
        self.nrhs = 1

    def close(self):
        if self.h:
            lib().ibim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # geometry ---------------------------------------------------------------
    def set_level(self, level, lo=-2.0, h=None):
        level = np.ascontiguousarray(level, dtype=np.float64)
        n = level.shape[0] - 1
        h = (-2.0 * lo) / n if h is None else h
        _check(lib().ibim_set_level_host(self.h, n, lo, h, _p(level)), "ibim_set_level_host")
        self.n = n
        return self

    def set_level_shape(self, shp, n, lo=-2.0, h=None):
        """Build the signed distance of `shp` (ibim.shape(...)) on the GPU."""
        h = (-2.0 * lo) / n if h is None else h
        _check(lib().ibim_set_level_shape(self.h, n, lo, h, C.byref(shp)), "ibim_set_level_shape")
        self.n = n
        return self

    def level(self):
        """Host copy of the level set."""
        S = self.n + 1
        out = np.zeros((S, S, S))
        _check(lib().ibim_get_level(self.h, _p(out)), "ibim_get_level")
        return out

    def redistance(self, band=8.0):
        """Level set -> signed distance to its zero set within band grid spacings."""
        _check(lib().ibim_redistance(self.h, band), "ibim_redistance")
        return self

    def geometry(self, weno_order=5):
        _check(lib().ibim_geometry(self.h, weno_order), "ibim_geometry")
        return self

    def grid_fields(self):
        S = self.n + 1
        out = [np.zeros((S, S, S)) for _ in range(5)]
        _check(lib().ibim_get_grid(self.h, *[_p(a) for a in out]), "ibim_get_grid")
        return out

    def set_grid_fields(self, gx, gy, gz, mc, gc):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (gx, gy, gz, mc, gc)]
        _check(lib().ibim_set_grid_fields_host(self.h, *[_p(a) for a in arrs]), "ibim_set_grid_fields_host")
        return self

    # tube ---------------------------------------------------------------------
    def tube(self, prm, kind=TUBE_TRADITIONAL):
        c = C.c_int(0)
        _check(lib().ibim_tube(self.h, C.byref(prm), kind, C.byref(c)), "ibim_tube")
        self.count = c.value
        self.nrhs = 1
        return self.count

    def tube_fields(self):
        node = np.zeros(self.count, dtype=np.int32)
        flat = np.zeros(NF * self.count)
        _check(lib().ibim_get_tube(self.h, node.ctypes.data_as(IP), _p(flat)), "ibim_get_tube")
        return node, flat

    # operator -----------------------------------------------------------------
    def assemble(self, precision=C128):
        _check(lib().ibim_assemble(self.h, precision), "ibim_assemble")
        return self

    def rows(self, row0, nrows):
        re = np.zeros((nrows, self.count))
        im = np.zeros((nrows, self.count))
        _check(lib().ibim_get_rows(self.h, row0, nrows, _p(re), _p(im)), "ibim_get_rows")
        return re, im

    def set_matrix(self, are, aim, precision=C128):
        are = np.ascontiguousarray(are, dtype=np.float64)
        aim = np.ascontiguousarray(aim, dtype=np.float64)
        _check(lib().ibim_set_matrix_host(self.h, are.shape[0], _p(are), _p(aim), precision), "ibim_set_matrix_host")
        self.count = are.shape[0]
        self.nrhs = 1
        return self

    def apply(self, p, q, op=0):
        """y = A p, z = A^H q (complex numpy vectors) with operator op
        (0 stored, 1 recomputed FP64, 2 recomputed FP32)."""
        p = np.asarray(p, dtype=np.complex128)
        q = np.asarray(q, dtype=np.complex128)
        arrs = [np.ascontiguousarray(a) for a in (p.real, p.imag, q.real, q.imag)]
        out = [np.zeros(self.count) for _ in range(4)]
        _check(lib().ibim_apply(self.h, op, *[_p(a) for a in arrs], *[_p(a) for a in out]), "ibim_apply")
        return out[0] + 1j * out[1], out[2] + 1j * out[3]

    def apply_y(self, p, op=0):
        """y = A p alone (no A^H product)."""
        p = np.asarray(p, dtype=np.complex128)
        pre, pim = np.ascontiguousarray(p.real), np.ascontiguousarray(p.imag)
        yre, yim = np.zeros(self.count), np.zeros(self.count)
        _check(lib().ibim_apply(self.h, op, _p(pre), _p(pim), None, None, _p(yre), _p(yim), None, None),
               "ibim_apply")
        return yre + 1j * yim

    def apply_noalloc(self, p_dev, q_dev, y_dev, z_dev, op=0):
        """ibim_apply_noalloc: y = A p, z = A^H q on device buffers (addresses of
        count complex128 values on the handle's device); allocates nothing."""
        _check(lib().ibim_apply_noalloc(self.h, op, C.c_void_p(p_dev), C.c_void_p(q_dev), C.c_void_p(y_dev),
                                        C.c_void_p(z_dev)), "ibim_apply_noalloc")

    def apply_batch(self, P, op=0, adjoint=False):
        """op(A) P for a (count, m <= 16) complex array, on the matrix cores."""
        P = np.asarray(P, dtype=complex)
        m = P.shape[1]
        pre = np.ascontiguousarray(P.real.T)
        pim = np.ascontiguousarray(P.imag.T)
        yre, yim = np.zeros((m, self.count)), np.zeros((m, self.count))
        _check(lib().ibim_apply_batch(self.h, op, 1 if adjoint else 0, m, _p(pre), _p(pim), _p(yre), _p(yim)),
               "ibim_apply_batch")
        return (yre + 1j * yim).T

    def regcount(self):
        return lib().ibim_regcount(self.h)

    # right-hand sides / solve ---------------------------------------------------
    def set_nrhs(self, nrhs):
        _check(lib().ibim_set_nrhs(self.h, nrhs), "ibim_set_nrhs")
        self.nrhs = nrhs

    def rhs(self, kind, x0=None, amp=1.0, direction=None, col=0):
        x0p = _p(np.ascontiguousarray(x0, dtype=np.float64)) if x0 is not None else None
        dp = _p(np.ascontiguousarray(direction, dtype=np.float64)) if direction is not None else None
        _check(lib().ibim_rhs(self.h, kind, x0p, amp, dp, col), "ibim_rhs")

    def set_rhs(self, bre, bim, col=0):
        bre = np.ascontiguousarray(bre, dtype=np.float64)
        bim = np.ascontiguousarray(bim, dtype=np.float64)
        _check(lib().ibim_set_rhs_host(self.h, col, _p(bre), _p(bim)), "ibim_set_rhs_host")

    def get_rhs(self, col=0):
        a, b = np.zeros(self.count), np.zeros(self.count)
        _check(lib().ibim_get_rhs(self.h, col, _p(a), _p(b)), "ibim_get_rhs")
        return a, b

    def solve(self, method=SOLVE_BICG, maxit=1000, tol=1e-7, inner_tol=1e-4, max_outer=20, chunk=0, batch=0):
        """batch 0: several right-hand sides go 16 per pass over the operator
        (BiCG); 1: one column at a time."""
        o = SolveOpts(method, maxit, tol, inner_tol, max_outer, chunk, batch)
        info = SolveInfo()
        hist = np.zeros(max(maxit, max_outer))
        _check(lib().ibim_solve(self.h, C.byref(o), C.byref(info), _p(hist)), "ibim_solve")
        n = info.iterations if method in (SOLVE_BICG, SOLVE_RELIABLE, SOLVE_GMRES, SOLVE_BLOCK_GMRES) and self.nrhs == 1 else info.outer
        return info, hist[:n]

    def density(self, col=0):
        a, b = np.zeros(self.count), np.zeros(self.count)
        _check(lib().ibim_get_density(self.h, col, _p(a), _p(b)), "ibim_get_density")
        return a, b

    def set_density(self, re, im, col=0):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        _check(lib().ibim_set_density_host(self.h, col, _p(re), _p(im)), "ibim_set_density_host")

    # evaluation -------------------------------------------------------------------
    def evaluate(self, xyz, mode=EVAL_DOUBLE, col=0):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        m = xyz.shape[0]
        a, b = np.zeros(m), np.zeros(m)
        _check(lib().ibim_evaluate(self.h, mode, col, m, _p(xyz), _p(a), _p(b)), "ibim_evaluate")
        return a, b

    # misfit gradient (adjoint) ---------------------------------------------------
    def set_receivers(self, xyz, target=None):
        """Receivers (m, 3) and complex targets t of J = sum |u(x_r) - t_r|^2."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        m = xyz.shape[0]
        t = np.zeros(m, complex) if target is None else np.asarray(target, complex)
        tre, tim = np.ascontiguousarray(t.real), np.ascontiguousarray(t.imag)
        _check(lib().ibim_set_receivers(self.h, m, _p(xyz), _p(tre), _p(tim)), "ibim_set_receivers")
        self.nrecv = m

    def gradient(self, mode=EVAL_DOUBLE, col=0, maxit=1000, tol=1e-12, chunk=0, method=SOLVE_BICG):
        """J and dJ/d(level-set values) (shape (n+1)^3, [y][x][z]) and dJ/dk for
        the solved column `col` (adjoint solve to `tol`: BiCG, or with
        SOLVE_BLOCK_GMRES enlarged GMRES on the 16-column A^H passes)."""
        o = SolveOpts(method, maxit, tol, 1e-4, 20, chunk, 1)
        J, dk = np.zeros(1), np.zeros(1)
        S = self.n + 1
        g = np.zeros((S, S, S))
        _check(lib().ibim_gradient(self.h, col, mode, C.byref(o), _p(J), _p(g), _p(dk)), "ibim_gradient")
        return float(J[0]), g, float(dk[0])

    # profiling --------------------------------------------------------------------
    def set_progress(self, fn):
        """fn(what, iteration, residual) at most every 5 s during long solves (None: off)."""
        if fn is None:
            self._progress = None
            _check(lib().ibim_set_progress(self.h, PROGRESS_FN(), None), "ibim_set_progress")
            return
        self._progress = PROGRESS_FN(lambda what, it, res, user: fn(what.decode(), it, res))
        _check(lib().ibim_set_progress(self.h, self._progress, None), "ibim_set_progress")

    def set_profiling(self, on=True):
        _check(lib().ibim_set_profiling(self.h, 1 if on else 0), "ibim_set_profiling")

    def set_fp32_products(self, on=True):
        """16-column passes over complex64 pairs on the FP32 matrix cores."""
        _check(lib().ibim_set_fp32_products(self.h, 1 if on else 0), "ibim_set_fp32_products")

    def set_free_cache(self, max_bytes):
        """Keep up to max_bytes of the FP64 matrix-free operator's complex128
        pairs (its first row blocks) in HBM; 0 frees them."""
        _check(lib().ibim_set_free_cache(self.h, float(max_bytes)), "ibim_set_free_cache")

    def set_graphs(self, on=True):
        """Replay one captured chunk of BiCG iterations as a HIP graph."""
        _check(lib().ibim_set_graphs(self.h, 1 if on else 0), "ibim_set_graphs")
        return self

    def kernel_time(self, name):
        ms, n = C.c_double(0), C.c_long(0)
        _check(lib().ibim_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)), "ibim_kernel_time")
        return ms.value, n.value

    def reset_timers(self):
        _check(lib().ibim_reset_timers(self.h), "ibim_reset_timers")

    def sync(self):
        _check(lib().ibim_sync(self.h), "ibim_sync")


# ---------------------------------------------------------------------------
# The reference driver's text files.
# ---------------------------------------------------------------------------
def solve_group(problems, method=SOLVE_BICG, maxit=1000, tol=1e-7, inner_tol=1e-4, chunk=0):
    """Every right-hand side of several problems solved together
    (ibim_solve_group: one persistent launch per iteration for all of their
    operators). Returns one SolveInfo per problem."""
    n = len(problems)
    hs = (C.c_void_p * n)(*[p.h for p in problems])
    infos = (SolveInfo * n)()
    o = SolveOpts(method, maxit, tol, inner_tol, 20, chunk, 0)
    _check(lib().ibim_solve_group(hs, n, C.byref(o), infos), "ibim_solve_group")
    return list(infos)


def reference_names(n, k):
    """File names the reference driver uses for grid size n and wavenumber k
    (large3Dexample2.c:407, 607-608, 3001-3002)."""
    return {"distance": "distance%d.txt" % n,
            "rhs": ("Rel_bN%dk%d.txt" % (n, int(k)), "Rel_bimgN%dk%d.txt" % (n, int(k))),
            "density": ("Rel_DensityN%dk%d.txt" % (n, int(k)), "Rel_DensityimgN%dk%d.txt" % (n, int(k)))}


def read_distance_file(path, n):
    out = np.zeros((n + 1,) * 3)
    _check(lib().ibim_read_distance_file(path.encode(), n, _p(out)), "ibim_read_distance_file")
    return out


def write_distance_file(path, level):
    level = np.ascontiguousarray(level, dtype=np.float64)
    _check(lib().ibim_write_distance_file(path.encode(), level.shape[0] - 1, _p(level)), "ibim_write_distance_file")


def _text_pair(fn, paths, count, a=None, b=None):
    a = np.zeros(count) if a is None else np.ascontiguousarray(a, dtype=np.float64)
    b = np.zeros(count) if b is None else np.ascontiguousarray(b, dtype=np.float64)
    _check(fn(paths[0].encode(), paths[1].encode(), count, _p(a), _p(b)), fn.__name__)
    return a, b


def read_rhs_text(paths, count):
    return _text_pair(lib().ibim_read_rhs_text, paths, count)


def write_rhs_text(paths, bre, bim):
    _text_pair(lib().ibim_write_rhs_text, paths, len(bre), bre, bim)


def read_density_text(paths, count):
    return _text_pair(lib().ibim_read_density_text, paths, count)


def write_density_text(paths, re, im):
    _text_pair(lib().ibim_write_density_text, paths, len(re), re, im)


# ---------------------------------------------------------------------------
# The reference's data structures from numpy, and the reference-named calls.
# ---------------------------------------------------------------------------
class Grid3:
    """A double*** view of a C-contiguous (S, S, S) float64 array (rows are the
    array's own memory, so the callee writes straight into it)."""

    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr, dtype=np.float64)
        S0, S1, _ = self.arr.shape
        self._rows = (P * (S0 * S1))()
        for i in range(S0):
            for j in range(S1):
                self._rows[i * S1 + j] = self.arr[i, j].ctypes.data_as(P)
        self._planes = (PP * S0)()
        for i in range(S0):
            self._planes[i] = C.cast(C.byref(self._rows, i * S1 * C.sizeof(P)), PP)
        self.ptr = C.cast(self._planes, PPP)


class Rows2:
    """A double** view of a C-contiguous (n, m) float64 array."""

    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._rows = (P * self.arr.shape[0])()
        for i in range(self.arr.shape[0]):
            self._rows[i] = self.arr[i].ctypes.data_as(P)
        self.ptr = C.cast(self._rows, PP)


def weno(f, dx, order):
    """weno (weno.c:259-345): one-sided derivatives from a 5/7-point stencil."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    a, b = np.zeros(1), np.zeros(1)
    lib().weno(_p(a), _p(b), _p(f), dx, order)
    return a[0], b[0]


def WENO_D(axis, grid, i, j, k, n, h, order):
    """WENO_DX / WENO_DY / WENO_DZ (weno.c:117-254); axis 0, 1, 2."""
    g = grid if isinstance(grid, Grid3) else Grid3(grid)
    a, b = np.zeros(1), np.zeros(1)
    f = (lib().WENO_DX, lib().WENO_DY, lib().WENO_DZ)[axis]
    f(g.ptr, i, j, k, n, h, _p(a), _p(b), order)
    return a[0], b[0]


def BothCurvs(grid, i, j, k, n, h):
    """BothCurvs (distance_application3D.c:35-54): (mean, Gaussian) curvature."""
    g = grid if isinstance(grid, Grid3) else Grid3(grid)
    a, b = np.zeros(1), np.zeros(1)
    lib().BothCurvs(g.ptr, i, j, k, n, h, _p(a), _p(b))
    return a[0], b[0]


def ASSIGN_GRADIENT(n, h, dist, weno_mode=5):
    """ASSIGN_GRADIENT (distance_application3D.c:56-141) on the GPU; returns gx, gy, gz."""
    d = Grid3(dist)
    outs = [Grid3(np.zeros_like(d.arr)) for _ in range(3)]
    _check(lib().ASSIGN_GRADIENT(n, h, d.ptr, *[o.ptr for o in outs], weno_mode), "ASSIGN_GRADIENT")
    return [o.arr for o in outs]


def _bicg(fn, are, aim, bre, bim, *extra):
    m = are.shape[0]
    A, B = Rows2(are), Rows2(aim)
    x, xi = np.zeros(m), np.zeros(m)
    bre = np.ascontiguousarray(bre, dtype=np.float64)
    bim = np.ascontiguousarray(bim, dtype=np.float64)
    _check(fn(m, _p(x), _p(xi), A.ptr, B.ptr, _p(bre), _p(bim), *extra), fn.__name__)
    return x, xi


def CompBiCG(are, aim, bre, bim):
    """CompBiCG (BICGSTAB.c:5-339) on the GPU: density (re, im) for A x = b."""
    return _bicg(lib().CompBiCG, are, aim, bre, bim)


def omp_CompBiCG(are, aim, bre, bim, chunk=16):
    """omp_CompBiCG (BICGSTAB.c:344-765) on the GPU."""
    return _bicg(lib().omp_CompBiCG, are, aim, bre, bim, chunk)


def set_globals(n, k):
    """Set the caller-side globals N, SIZE, H, HCUBED, WAVE the kernel drop-ins
    read (HelmholtzKernel.h:16-18)."""
    L_ = lib()
    h = 4.0 / n
    C.c_int.in_dll(L_, "N").value = n
    C.c_int.in_dll(L_, "SIZE").value = n + 1
    C.c_double.in_dll(L_, "H").value = h
    C.c_double.in_dll(L_, "HCUBED").value = h * h * h
    C.c_double.in_dll(L_, "WAVE").value = k


def new_HelmholtzKernelCombo(size, delta, eps, level, gx, gy, gz, mc, gc, wavenum, eta):
    """new_HelmholtzKernelCombo (HelmholtzKernel.c:173-610) on the GPU.

    Returns (regcount, dict of point arrays, SineKernelH, SineKernelHimg)."""
    grids = [Grid3(a) for a in (level, gx, gy, gz, mc, gc)]
    names = ["zx", "zy", "zz", "sx", "sy", "sz", "dist", "nx", "ny", "nz", "jac", "reg", "regimg", "polyw",
             "sinew", "polyr", "siner"]
    arrs = {nm: np.zeros(size) for nm in names}
    K, Ki = Rows2(np.zeros((size, size))), Rows2(np.zeros((size, size)))
    dummy = Rows2(np.zeros((1, 1)))
    rc = lib().new_HelmholtzKernelCombo(size, delta, eps, *[g.ptr for g in grids], *[_p(arrs[nm]) for nm in names],
                                        dummy.ptr, dummy.ptr, K.ptr, Ki.ptr, wavenum, eta)
    _check(rc, "new_HelmholtzKernelCombo")
    return rc, arrs, K.arr, Ki.arr


def DoubleLayerKernel(size, eps, level, gx, gy, gz, mc, gc, wavenum):
    """DoubleLayerKernel (HelmholtzKernel.c:613-761) on the GPU."""
    grids = [Grid3(a) for a in (level, gx, gy, gz, mc, gc)]
    names = ["zx", "zy", "zz", "sx", "sy", "sz", "dist", "nx", "ny", "nz", "delta", "jac", "reg", "regimg",
             "result"]
    arrs = {nm: np.zeros(size) for nm in names}
    K, Ki = Rows2(np.zeros((size, size))), Rows2(np.zeros((size, size)))
    _check(lib().DoubleLayerKernel(size, eps, *[g.ptr for g in grids], *[_p(arrs[nm]) for nm in names], K.ptr,
                                   Ki.ptr, wavenum), "DoubleLayerKernel")
    return arrs, K.arr, Ki.arr


def new_HelmholtzKernel(size, delta, eps):
    """new_HelmholtzKernel (HelmholtzKernel.c:4-169) on the GPU (unit sphere)."""
    names = ["zx", "zy", "zz", "sx", "sy", "sz", "dist", "nx", "ny", "nz", "curv", "jac", "polyw", "sinew",
             "polyr", "siner"]
    arrs = {nm: np.zeros(size) for nm in names}
    K, Ki = Rows2(np.zeros((size, size))), Rows2(np.zeros((size, size)))
    dummy = Rows2(np.zeros((1, 1)))
    _check(lib().new_HelmholtzKernel(size, delta, eps, *[_p(arrs[nm]) for nm in names], dummy.ptr, dummy.ptr, K.ptr,
                                     Ki.ptr), "new_HelmholtzKernel")
    return arrs, K.arr, Ki.arr
