"""The reference driver's text formats (host-only code of libibim.so, no GPU)."""
import os

import numpy as np
import pytest

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ib():
    import ibim
    ibim.lib()
    return ibim


def test_reads_the_reference_rhs_files(ib):
    g = np.load(os.path.join(G, "rhs_sphere.npz"))
    paths = (os.path.join(G, "Rel_bN16k2.txt"), os.path.join(G, "Rel_bimgN16k2.txt"))
    bre, bim = ib.read_rhs_text(paths, len(g["bre"]))
    assert np.array_equal(bre, g["bre"]) and np.array_equal(bim, g["bim"])
    assert ib.reference_names(16, 2.0)["rhs"] == ("Rel_bN16k2.txt", "Rel_bimgN16k2.txt")


def test_writes_rhs_files_byte_identical_to_the_reference(ib, tmp_path):
    g = np.load(os.path.join(G, "rhs_sphere.npz"))
    out = (str(tmp_path / "re.txt"), str(tmp_path / "im.txt"))
    ib.write_rhs_text(out, g["bre"], g["bim"])
    for mine, ref in zip(out, ("Rel_bN16k2.txt", "Rel_bimgN16k2.txt")):
        assert open(mine).read() == open(os.path.join(G, ref)).read()


def test_rhs_size_mismatch_is_an_error(ib):
    paths = (os.path.join(G, "Rel_bN16k2.txt"), os.path.join(G, "Rel_bimgN16k2.txt"))
    with pytest.raises(ib.IbimError):
        ib.read_rhs_text(paths, 10)


def test_distance_and_density_round_trips(ib, tmp_path):
    rng = np.random.default_rng(0)
    level = rng.standard_normal((9, 9, 9))
    p = str(tmp_path / "distance8.txt")
    ib.write_distance_file(p, level)
    assert np.array_equal(ib.read_distance_file(p, 8), level)
    re, im = rng.standard_normal(50), rng.standard_normal(50)
    d = (str(tmp_path / "d.txt"), str(tmp_path / "di.txt"))
    ib.write_density_text(d, re, im)
    a, b = ib.read_density_text(d, 50)
    assert np.abs(a - re).max() <= 5e-16 + 1e-15 and np.abs(b - im).max() <= 5e-16 + 1e-15
