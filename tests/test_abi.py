"""The C ABI library loads and exports everything include/ibim.h declares (CPU only:
no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "ibim.h")
LIB = os.path.join(REPO, "helmholtz-3d-torrone_amd", "libibim.so")

# The reference compiles its .c files as C++, so a program built against its
# headers references these mangled names (BICGSTAB.h:15-17, HelmholtzKernel.h:20-39,
# weno.h:26-31, distance_application3D.h:9-12).
MANGLED = [
    "_Z8CompBiCGiPdS_PS_S0_S_S_",
    "_Z12omp_CompBiCGiPdS_PS_S0_S_S_i",
    "_Z19new_HelmholtzKerneliddPdS_S_S_S_S_S_S_S_S_S_S_S_S_S_S_PS_S0_S0_S0_",
    "_Z24new_HelmholtzKernelComboiddPPPdS1_S1_S1_S1_S1_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S0_S0_S0_S0_dd",
    "_Z17DoubleLayerKernelidPPPdS1_S1_S1_S1_S1_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S0_S0_d",
    "_Z7WENO_DXPPPdiiiidS_S_i",
    "_Z7WENO_DYPPPdiiiidS_S_i",
    "_Z7WENO_DZPPPdiiiidS_S_i",
    "_Z4wenoPdS_S_di",
    "_Z12GaussianCurvPPPdiiiid",
    "_Z8MeanCurvPPPdiiiid",
    "_Z9BothCurvsPPPdiiiidS_S_",
    "_Z15ASSIGN_GRADIENTidPPPdS1_S1_S1_i",
]


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", "", text)  # function-pointer types are not symbols
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libibim.so is not built: run __graft_entry__.build()")
    return ctypes.CDLL(LIB)


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for ref in ("CompBiCG", "omp_CompBiCG", "new_HelmholtzKernel", "new_HelmholtzKernelCombo", "DoubleLayerKernel",
                "WENO_DX", "WENO_DY", "WENO_DZ", "weno", "GaussianCurv", "MeanCurv", "BothCurvs",
                "ASSIGN_GRADIENT"):
        assert ref in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_cxx_linkage_twins_are_exported(lib):
    missing = [n for n in MANGLED if not hasattr(lib, n)]
    assert not missing, missing


def test_caller_globals_are_weak_definitions():
    out = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    for g in ("N", "SIZE", "H", "HCUBED", "WAVE"):
        assert re.search(r"\s[VW]\s%s$" % g, out, flags=re.M), g


def test_library_is_built_for_gfx950():
    out = subprocess.run(["strings", LIB], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_cxx_caller_references_the_mangled_names_and_defines_the_globals():
    """tests/cxx/reference_caller declares the reference's prototypes without
    extern "C": the symbols it imports from libibim.so are the mangled twins,
    and it defines the driver's globals itself (the GPU test runs it)."""
    caller = os.path.join(REPO, "tests", "cxx", "reference_caller")
    if not os.path.exists(caller):
        subprocess.run(["make", "-s", "-C", os.path.dirname(caller)], check=True)
    undef = subprocess.run(["nm", "-u", caller], capture_output=True, text=True).stdout.split()
    for name in ("_Z8CompBiCGiPdS_PS_S0_S_S_", "_Z15ASSIGN_GRADIENTidPPPdS1_S1_S1_i",
                 "_Z24new_HelmholtzKernelComboiddPPPdS1_S1_S1_S1_S1_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S_S0_S0_S0_S0_dd"):
        assert name in undef, name
    assert not any(u in ("CompBiCG", "ASSIGN_GRADIENT", "new_HelmholtzKernelCombo") for u in undef)
    defined = subprocess.run(["nm", caller], capture_output=True, text=True).stdout
    for g in ("N", "SIZE", "H", "HCUBED", "WAVE"):
        assert re.search(r"\s[BD]\s%s$" % g, defined, flags=re.M), g
    linked = subprocess.run(["ldd", caller], capture_output=True, text=True).stdout
    assert "libibim.so" in linked
