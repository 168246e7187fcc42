"""The drop-in front/back end on the GPU: `mc <io_dir>` (the reference's
MC.exe), the C++ mcold::Simulator behind it, and the Python Simulator mirror,
on a deck directory in the reference's layout (<dir>/input.xml + xs files).
The CLI writes output.h5 where an HDF5 library loads (read back here with
tests/h5_dump.cpp on the image's HDF5 C++ API), else output.txt."""
import os
import subprocess

import numpy as np
import pytest

from mc_testlib import ROOT, decks, mc, oracle_run

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "mc-old_amd", "mc")
CONDA = "/opt/conda"
KEYS = ("/summary/Nsample", "/summary/Ntrack", "/transmission/cross/mean", "/flux/flux/mean")


def parse_text(path):
    out = {}
    for line in open(path):
        key, *vals = line.split()
        out[key] = vals
    return out


def parse_h5(path, tmp):
    exe = str(tmp / "h5_dump")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + CONDA + "/include", os.path.join(ROOT, "tests", "h5_dump.cpp"),
                    "-o", exe, "-L" + CONDA + "/lib", "-lhdf5_cpp", "-lhdf5", "-Wl,-rpath," + CONDA + "/lib"],
                   check=True)
    items = ["u:" + KEYS[0], "u:" + KEYS[1], "v:" + KEYS[2], "v:" + KEYS[3]]
    r = subprocess.run([exe, str(path), *items], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, rest = line.split(" ", 1)
        key = k[2:]
        out[key] = rest.split(" values ")[1].split() if " values " in rest else rest.split()
    return out


def cli_report(d, tmp):
    """The report the CLI wrote: {key: [values as text]} for KEYS."""
    if (d / "output.h5").exists():
        return parse_h5(d / "output.h5", tmp)
    return parse_text(d / "output.txt")


def test_cli_and_simulator_mirror(tmp_path):
    n = 20000
    d = tmp_path / "slab"
    d.mkdir()
    (d / "input.xml").write_text(decks.slab_deck(4.0, 0.5, 10, n))
    r = subprocess.run([CLI, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rep = cli_report(d, tmp_path)
    assert rep["/summary/Nsample"] == [str(n)]
    T_cli = float(rep["/transmission/cross/mean"][0])
    flux_cli = np.array([float(v) for v in rep["/flux/flux/mean"]])
    assert flux_cli.size == 10
    # the Python mirror of Simulator gives the same numbers
    sim = mc.Simulator(str(d))
    sim.start()
    assert [e.name for e in sim.Estimators] == ["transmission", "reflection", "flux"]
    assert abs(sim.Estimators[0].tally(0).mean - T_cli) <= 1e-12 * abs(T_cli)
    assert sim.Ntrack == int(rep["/summary/Ntrack"][0])
    path = sim.report(str(d))
    assert os.path.exists(path)
    # and both equal the oracle's moments
    o = oracle_run(mc.Problem.from_xml(str(d)), n)
    assert abs(T_cli - o.mean[0]) <= 1e-12 * abs(o.mean[0])


def test_cli_reports_input_errors(tmp_path):
    d = tmp_path / "bad"
    d.mkdir()
    (d / "input.xml").write_text(decks.slab_deck(1.0, 0.5, 2, 100).replace('material="slab"', 'material="x"', 1))
    r = subprocess.run([CLI, str(d)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "[ERROR 2]" in r.stderr
