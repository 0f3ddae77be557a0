"""Operator transfers between host rows and HBM (ibim_set_matrix_host,
ibim_get_rows, the drop-ins' double** rows): sizes that span several pinned
staging blocks (32 MiB each), row ranges that start and end inside a block,
complex64 storage, and row-pointer arrays whose rows are separate allocations.

Transfers are exact for complex128 (the interleave only moves bits) and round
each part to FP32 once for complex64, so the checks are bitwise.
"""
import ctypes as C

import numpy as np
import pytest

import shapes

pytestmark = pytest.mark.gpu


def random_operator(n, seed):
    rng = np.random.default_rng(seed)
    re = rng.standard_normal((n, n)) * (0.2 / np.sqrt(n))
    im = rng.standard_normal((n, n)) * (0.2 / np.sqrt(n))
    re[np.diag_indices(n)] += 1.0  # well conditioned: BiCG converges in a few dozen steps
    return re, im


# 2500 rows: 40 KiB per complex128 row at ld 2560, so 819 rows a block and
# four blocks, the last partial; complex64: two blocks
@pytest.mark.parametrize("precision", ["C128", "C64"])
def test_matrix_round_trip_spans_staging_blocks(gpu, precision):
    n = 2500
    re, im = random_operator(n, 1)
    pb = gpu.Problem()
    pb.set_matrix(re, im, getattr(gpu, precision))
    r2, i2 = pb.rows(0, n)
    if precision == "C128":
        want_re, want_im = re, im
    else:
        want_re = re.astype(np.float32).astype(np.float64)
        want_im = im.astype(np.float32).astype(np.float64)
    assert np.array_equal(r2, want_re) and np.array_equal(i2, want_im)
    for r0, k in [(1234, 700), (818, 2), (n - 1, 1), (0, 0)]:
        a, b = pb.rows(r0, k)
        assert np.array_equal(a, want_re[r0:r0 + k]) and np.array_equal(b, want_im[r0:r0 + k]), (r0, k)
    with pytest.raises(gpu.IbimError):
        pb.rows(n - 1, 2)


def test_dropin_compbicg_reads_scattered_row_pointers(gpu):
    n = 2500
    re, im = random_operator(n, 2)
    rng = np.random.default_rng(3)
    bre, bim = rng.standard_normal(n), rng.standard_normal(n)
    # every row its own allocation, as the reference's malloc-per-row double**
    rows_re = [np.array(re[i]) for i in range(n)]
    rows_im = [np.array(im[i]) for i in range(n)]
    ptr_t = C.POINTER(C.c_double)
    pre = (ptr_t * n)(*[r.ctypes.data_as(ptr_t) for r in rows_re])
    pim = (ptr_t * n)(*[r.ctypes.data_as(ptr_t) for r in rows_im])
    x, xi = np.zeros(n), np.zeros(n)
    L = gpu.lib()
    rc = L.CompBiCG(n, x.ctypes.data_as(ptr_t), xi.ctypes.data_as(ptr_t), C.cast(pre, C.POINTER(ptr_t)),
                    C.cast(pim, C.POINTER(ptr_t)), bre.ctypes.data_as(ptr_t), bim.ctypes.data_as(ptr_t))
    assert rc == 0
    # the same operator through the contiguous host path and the handle API
    pb = gpu.Problem()
    pb.set_matrix(re, im)
    pb.set_rhs(bre, bim)
    info, _ = pb.solve(gpu.SOLVE_BICG, maxit=1000, tol=1e-7)
    assert info.status == 0
    dre, dim = pb.density()
    assert np.array_equal(x, dre) and np.array_equal(xi, dim)
    A = re + 1j * im
    res = np.linalg.norm(A @ (x + 1j * xi) - (bre + 1j * bim)) / np.linalg.norm(bre + 1j * bim)
    assert res <= 1e-7


def test_pair_rows_download_in_blocks_from_any_row(gpu):
    # pair storage rebuilds rows on the device, 32 MiB of complex128 rows a block
    n = 28
    h = 4.0 / n
    pb = gpu.Problem().set_level(shapes.sphere(n)).geometry(5)
    pb.tube(gpu.params(4.0, 3 * h, 2 * h, dirichlet=1))
    count = pb.count
    per = (32 << 20) // (16 * count)
    assert count > 2 * per  # at least three blocks
    pb.assemble(gpu.C128_PAIR)
    re, im = pb.rows(0, count)
    for r0, k in [(per - 3, 7), (count - per - 1, per + 1), (5, 1)]:
        a, b = pb.rows(r0, k)
        assert np.array_equal(a, re[r0:r0 + k]) and np.array_equal(b, im[r0:r0 + k]), (r0, k)
    pb.assemble(gpu.C128)
    dre, dim = pb.rows(0, count)
    nz = np.abs(dre + 1j * dim) > 0
    assert (np.abs((re + 1j * im) - (dre + 1j * dim))[nz] / np.abs(dre + 1j * dim)[nz]).max() <= 1e-14
