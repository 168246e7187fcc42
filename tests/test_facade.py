"""The C++ Simulator facade's read-only object views (simulator.h:118-122:
Surfaces, Cells, Materials, Nuclides; Surface::eval/distance/reflect,
Cell::importance/material/surfaces, Material::nuclides, search_cell) on the
reference's own shielding deck, from a small C++ program linked against the
library.  Loading and inspecting a deck needs no GPU."""
import os
import subprocess

import pytest

from mc_testlib import REFERENCE_DECK, ROOT, mc, oracle

HOST = os.path.join(ROOT, "mc-old_amd", "host")


@pytest.fixture(scope="module")
def facade(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("facade") / "facade_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + HOST, os.path.join(ROOT, "tests", "facade_check.cpp"),
                    "-o", exe, "-L" + os.path.dirname(mc.LIB_PATH), "-lmcold",
                    "-Wl,-rpath," + os.path.dirname(mc.LIB_PATH)], check=True)
    return exe


def run(exe, *extra):
    r = subprocess.run([exe, REFERENCE_DECK, os.path.join(REFERENCE_DECK, "xs_library"), *extra],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [l.split() for l in r.stdout.splitlines()]


def test_object_views_of_the_reference_deck(facade):
    lines = run(facade)
    assert lines[0] == ["counts", "8", "10", "3", "6", "100000000"]
    surf = {l[1]: l[2:] for l in lines if l[0] == "surface"}
    assert surf["px1"] == ["1", "0"] and surf["py3"] == ["2", "0"] and surf["cz1"] == ["8", "0"]
    nucs = {l[1]: l[2:] for l in lines if l[0] == "nuclide"}
    assert nucs["H1"][0] == "1" and nucs["O16"][0] == "16" and nucs["B11"][0] == "11"
    assert nucs["H1"][2] == "1"                       # free gas, the reference's default
    mats = {l[1]: l[2:] for l in lines if l[0] == "material"}
    assert [(x.split(":")[0], float(x.split(":")[1])) for x in mats["H2O"]] == [("H1", 0.066733), ("O16", 0.033368)]
    cells = {l[1]: l[2:] for l in lines if l[0] == "cell"}
    assert cells["detector"] == ["1", "mat_detector", "cz1:-1"]
    assert cells["water4"][1] == "H2O" and "cz1:+1" in cells["water4"]
    assert [l for l in lines if l[0] == "search"] == [["search", "water1"]]


def test_surface_queries_equal_the_oracle(facade):
    """eval, distance and reflect through the facade against the oracle's
    restatement of Geometry.cpp on the same surface records."""
    import ctypes as C
    import numpy as np
    lines = run(facade)
    p = mc.Problem.from_xml(REFERENCE_DECK, os.path.join(REFERENCE_DECK, "xs_library"))
    surfs = (C.c_byte * (80 * p.head.n_surfaces)).from_address(p.head.surfaces)
    geoms = [l for l in lines if l[0] == "geom"]
    assert len(geoms) == 8
    V3 = C.c_double * 3
    for k, g in enumerate(geoms):
        sd = C.addressof(surfs) + 80 * k
        pos, u = V3(1.5, 1.5, 0.0), V3(0.6, 0.8, 0.0)
        assert float(g[2]) == oracle.mco_surface_eval(sd, pos)
        assert float(g[3]) == oracle.mco_surface_distance(sd, pos, u)
        oracle.mco_surface_reflect(sd, u)
        assert [float(x) for x in g[4:7]] == [u[0], u[1], u[2]]
