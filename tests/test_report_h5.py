"""output.h5 in the reference's layout (Simulator::report, report.cpp:9-50;
Estimator::report, Estimator.cpp:368-422), written through the C ABI and
read back with a small program on the HDF5 C++ API (tests/h5_dump.cpp).
Needs the HDF5 library and headers of the image (/opt/conda); no GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from mc_testlib import ROOT, decks, mc, shield_xs_dir

CONDA = "/opt/conda"
needs_hdf5 = pytest.mark.skipif(not os.path.exists(os.path.join(CONDA, "include", "H5Cpp.h")),
                                reason="HDF5 C++ headers absent")


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("h5") / "h5_dump")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + CONDA + "/include", os.path.join(ROOT, "tests", "h5_dump.cpp"),
                    "-o", exe, "-L" + CONDA + "/lib", "-lhdf5_cpp", "-lhdf5", "-Wl,-rpath," + CONDA + "/lib"],
                   check=True)
    return exe


def read(exe, path, *items):
    r = subprocess.run([exe, path, *items], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, rest = line.split(" ", 1)
        out[k] = rest
    return out


def fake_tally(p):
    nb = p.n_bins
    rng = np.random.default_rng(5)
    mean, uncer = rng.random(nb), rng.random(nb) * 1e-3
    keep = [mean, uncer, np.zeros(nb), np.zeros(nb)]
    t = mc.TallyOut(mc._dptr(keep[2]), mc._dptr(keep[3]), mc._dptr(mean), mc._dptr(uncer), 123456789, 1.5)
    return t, keep


@needs_hdf5
def test_shield_report_layout(dump, tmp_path):
    p = mc.Problem.from_string(decks.shield_deck(1000), shield_xs_dir())
    t, keep = fake_tally(p)
    path = str(tmp_path / "output.h5")
    assert mc.lib.mc_report_h5(p._h, C.byref(t), 1000, path.encode()) == mc.MC_OK, mc.lib.mc_last_error()
    got = read(dump, path, "u:/summary/Ncycle", "u:/summary/Nsample", "u:/summary/Npassive", "u:/summary/Ntrack",
               "s:/summary/mode", "d:/summary/survival_roulette/wr", "d:/summary/survival_roulette/ws",
               "g:/transmission:indexing", "v:/transmission/surface", "t:/transmission/surface:unit",
               "v:/transmission/cross/mean", "v:/transmission/cross/uncertainty",
               "g:/leakage:indexing", "v:/leakage/energy", "t:/leakage/energy:unit", "v:/leakage/cross/mean")
    assert got["u:/summary/Ncycle"] == "1" and got["u:/summary/Nsample"] == "1000"
    assert got["u:/summary/Npassive"] == "0" and got["u:/summary/Ntrack"] == "123456789"
    assert got["s:/summary/mode"] == "fixed source"
    assert float(got["d:/summary/survival_roulette/wr"]) == 0.001
    assert float(got["d:/summary/survival_roulette/ws"]) == 1.0
    assert got["g:/transmission:indexing"] == "[surface]"
    assert got["t:/transmission/surface:unit"] == "id#"
    assert got["g:/leakage:indexing"] == "[surface][energy]"
    assert got["t:/leakage/energy:unit"] == "eV"

    def arr(s):
        head, vals = s.split(" values ")
        return [int(x) for x in head.split()[1:]], np.array([float(x) for x in vals.split()])

    dims, v = arr(got["v:/transmission/surface"])
    assert dims == [1] and v[0] == 3.0                   # px3 is surface 3 in deck order
    dims, v = arr(got["v:/transmission/cross/mean"])
    assert dims == [1] and v[0] == keep[0][0]
    dims, v = arr(got["v:/transmission/cross/uncertainty"])
    assert v[0] == keep[1][0]
    dims, v = arr(got["v:/leakage/energy"])
    est = p.estimators()[1]
    grid = np.ctypeslib.as_array((C.c_double * (est.egrid_begin + est.egrid_count)).from_address(p.head.egrid))
    assert dims == [101] and (v == grid[est.egrid_begin:]).all()
    dims, v = arr(got["v:/leakage/cross/mean"])
    assert dims == [2, 100]                              # [surface][energy]
    assert (v == keep[0][1:201]).all()


@needs_hdf5
def test_several_scores_and_cells(dump, tmp_path):
    from mc_testlib import reaction_slab
    p = mc.Problem.from_string(reaction_slab(2.0, 0.5, 100))
    t, keep = fake_tally(p)
    path = str(tmp_path / "output.h5")
    assert mc.lib.mc_report_h5(p._h, C.byref(t), 100, path.encode()) == mc.MC_OK
    names = ["flux", "absorption", "scatter", "capture", "total"]
    got = read(dump, path, "g:/tl:indexing", "v:/tl/cell", *[f"v:/tl/{s}/mean" for s in names],
               *[f"v:/col/{s}/uncertainty" for s in names])
    assert got["g:/tl:indexing"] == "[cell]"
    for k, s in enumerate(names):
        v = float(got[f"v:/tl/{s}/mean"].split(" values ")[1])
        assert v == keep[0][2 + k]
        u = float(got[f"v:/col/{s}/uncertainty"].split(" values ")[1])
        assert u == keep[1][7 + k]
