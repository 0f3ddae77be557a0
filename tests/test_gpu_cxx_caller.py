"""The C++ boundary: a program that declares the reference's prototypes
(BICGSTAB.h:15-17, distance_application3D.h:12, HelmholtzKernel.h:27-33)
without extern "C" -- so it binds to the C++-mangled symbols, as a program
built from the reference's own headers would -- and defines the reference
driver's globals, linked against libibim.so and run on the golden fixtures.
Its results must equal those of the reference the fixtures were written from,
to the tolerances of the ctypes drop-in tests (test_gpu_parity).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CALLER = os.path.join(HERE, "cxx", "reference_caller")


def golden(name):
    return np.load(os.path.join(HERE, "golden", name))


def relmax(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


def test_cxx_caller_binds_the_mangled_drop_ins(tmp_path):
    if not os.path.exists(CALLER):
        pytest.fail("tests/cxx/reference_caller is not built (make -C tests/cxx)")
    c = golden("combo_sphere.npz")
    geo = golden("geom_sphere.npz")
    b = golden("bicg.npz")
    twin = golden("geom_twin.npz")
    n, size, m = int(c["n"]), int(c["size"]), b["are"].shape[0]
    assert int(geo["n"]) == n and int(twin["n"]) == n
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([n, size, m], dtype="<i4").tofile(f)
        np.array([float(c["k"]), float(c["eta"]), float(c["delta"]), float(c["eps"])], dtype="<f8").tofile(f)
        # ASSIGN_GRADIENT on the kinked twin level set (limiter and Godunov
        # branches); the combo kernel on the sphere and its geometry grids
        for a in (twin["level"], geo["level"], geo["gx"], geo["gy"], geo["gz"], geo["mc"], geo["gc"]):
            np.ascontiguousarray(a, dtype="<f8").tofile(f)
        for a in (b["are"], b["aim"], b["bre"], b["bim"]):
            np.ascontiguousarray(a, dtype="<f8").tofile(f)
    out = tmp_path / "out.bin"
    r = subprocess.run([CALLER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    v = np.fromfile(out, dtype="<f8")
    S3 = (n + 1) ** 3
    pos = 0

    def take(cnt):
        nonlocal pos
        a = v[pos:pos + cnt]
        pos += cnt
        return a

    for k in ("gx", "gy", "gz"):
        assert relmax(take(S3).reshape((n + 1,) * 3), twin[k]) <= 1e-13, k
    regcount = int(take(1)[0])
    pts = {k: take(size) for k in ("sx", "sy", "sz", "nx", "ny", "nz", "w")}
    K = take(size * size).reshape(size, size)
    Ki = take(size * size).reshape(size, size)
    x, xi = take(m), take(m)
    assert pos == len(v)
    assert regcount == int(c["regcount"])
    for k in ("sx", "sy", "sz", "nx", "ny", "nz", "w"):
        assert relmax(pts[k], c[k]) <= 1e-13, k
    scale = np.abs(c["row_re"]).max()
    assert np.abs(K[c["rows"]] - c["row_re"]).max() <= 1e-13 * scale
    assert np.abs(Ki[c["rows"]] - c["row_im"]).max() <= 1e-13 * scale
    assert np.abs(K[c["pi"], c["pj"]] - c["entry_re"]).max() <= 1e-13 * scale
    np.testing.assert_allclose(x, b["serial_x"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(xi, b["serial_xi"], rtol=0, atol=1e-9)
