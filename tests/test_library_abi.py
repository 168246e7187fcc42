"""The C-ABI library loads and exports every entry point include/*.h declares;
host-side error paths return status codes (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from mc_testlib import ROOT, decks, mc, shield_xs_dir


def declared_symbols():
    syms = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms |= set(re.findall(r"\b(mc_[a-z_0-9]+)\s*\(", txt))
    return syms


def test_exports_every_declared_symbol():
    lib = C.CDLL(mc.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 19
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(mc.EXPORTS) <= syms


def test_version_and_last_error():
    assert b"gfx950" in mc.lib.mc_version()
    h = C.c_void_p()
    st = mc.lib.mc_problem_from_xml(b"/nonexistent/dir", None, C.byref(h))
    assert st == mc.MC_ERR_IO
    assert b"cannot read" in mc.lib.mc_last_error()


def expect_status(xml, status, xs_dir=None):
    with pytest.raises(mc.MCError) as e:
        mc.Problem.from_string(xml, xs_dir)
    assert e.value.status == status, str(e.value)


def test_input_errors():
    ok = decks.slab_deck(2.0, 0.5, 4, 100)
    mc.Problem.from_string(ok)
    expect_status(ok.replace('<surface name="px1" sense="+1"/>', '<surface name="nope" sense="+1"/>', 1),
                  mc.MC_ERR_INPUT)
    expect_status(ok.replace('material="slab"', 'material="nope"', 1), mc.MC_ERR_INPUT)
    expect_status(ok.replace('scores="cross"', 'scores="bogus"', 1), mc.MC_ERR_INPUT)
    expect_status(ok.replace('scores="cross"', 'scores="fission"', 1), mc.MC_ERR_UNSUPPORTED)
    expect_status(ok.replace('<point x="1e-09"', '<disk_z x="1e-09"', 1).replace("</point>", ""),
                  mc.MC_ERR_UNSUPPORTED)
    expect_status(ok.replace("</simulation>", "<ksearch/></simulation>"), mc.MC_ERR_UNSUPPORTED)
    expect_status(ok.replace("<sources>", "<sources>\n    <point x=\"-5\" y=\"0\" z=\"0\"/>")
                  .replace('<cell name="left outside" importance="0.0">\n        <surface name="px0" sense="-1"/>',
                           '<cell name="left outside" importance="0.0">\n        <surface name="px0" sense="-1"/>'
                           '<surface name="px1" sense="+1"/>'), mc.MC_ERR_LOST_PARTICLE)
    expect_status("<simulation><description samples='10'/></simulation><oops", mc.MC_ERR_INPUT)
    expect_status(ok.replace('<nuclide name="onespeed">', '<nuclide name="onespeed" ZAID="999999">'),
                  mc.MC_ERR_IO, "/nonexistent")


def test_slab_deck_structure():
    p = mc.Problem.from_string(decks.slab_deck(4.0, 0.5, 100, 1000))
    assert p.head.n_cells == 102 and p.head.n_surfaces == 101
    assert p.estimator_names() == ["transmission", "reflection", "flux"]
    assert [e.n_bins for e in p.estimators()] == [1, 1, 100]
    assert p.n_bins == 102 and p.n_histories == 1000


def test_grid_lethargy_matches_reference_construction():
    """setup.cpp:764-776: u by repeated addition of ln(b/a)/N, reversed, E = b e^{-u}."""
    import math
    import numpy as np
    p = mc.Problem.from_string(decks.shield_deck(100), shield_xs_dir())
    est = p.estimators()[1]
    grid = np.ctypeslib.as_array((C.c_double * (est.egrid_begin + est.egrid_count)).from_address(p.head.egrid))
    g = grid[est.egrid_begin:]
    step = math.log(2e7 / 1e-5) / 100.0
    u = [0.0]
    while len(u) != 101:
        u.append(u[-1] + step)
    expect = [2e7 * math.exp(-v) for v in reversed(u)]
    assert list(g) == expect


def test_xs_library_five_and_six_columns(tmp_path):
    # rows of 5 columns (as xs_library/001001.txt) and of 6 are both accepted
    d = tmp_path / "xs"
    d.mkdir()
    (d / "111111.txt").write_text("2.0\n0 0\n0 0\n0 0\n1e-5 10 1 0 0\n1.0 5 0.5 0 0 0\n2e7 4 0.1 0 0\n")
    (d / "222222.txt").write_text("2.0\n0 0\n0 0\n0 0\n1e-5 10 1 0.5 2.4 0\n2e7 4 0.1 0.1 2.4 0\n")
    xml = decks.slab_deck(2.0, 0.5, 1, 100).replace('<nuclide name="onespeed">',
                                                    '<nuclide name="tab" ZAID="111111"/>\n'
                                                    '    <nuclide name="onespeed">')
    p = mc.Problem.from_string(xml, str(d))
    assert p.head.n_nuclides == 2
    fiss = xml.replace('ZAID="111111"', 'ZAID="222222"')
    expect_status(fiss, mc.MC_ERR_UNSUPPORTED, str(d))


@pytest.mark.skipif(not os.path.exists("/root/reference/xs_library/001001.txt"), reason="reference absent")
def test_reads_reference_five_column_file(tmp_path):
    # the reference's own H-1 file has 5 columns; its reader misaligns rows (setup.cpp:366)
    xml = decks.slab_deck(2.0, 0.5, 1, 100).replace('<nuclide name="onespeed">',
                                                    '<nuclide name="H1" ZAID="001001"/>\n'
                                                    '    <nuclide name="onespeed">')
    p = mc.Problem.from_string(xml, "/root/reference/xs_library")
    assert p.head.n_nuclides == 2


def test_engine_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    p = mc.Problem.from_string(decks.slab_deck(2.0, 0.5, 4, 100))
    with pytest.raises(mc.MCError) as e:
        mc.Engine(p)
    assert e.value.status == mc.MC_ERR_DEVICE


def test_cli_usage_and_missing_device(tmp_path):
    import subprocess
    cli = os.path.join(ROOT, "mc-old_amd", "mc")
    r = subprocess.run([cli], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    d = tmp_path / "deck"
    d.mkdir()
    (d / "input.xml").write_text(decks.slab_deck(1.0, 0.5, 2, 100))
    r = subprocess.run([cli, str(d)], capture_output=True, text=True, timeout=60)
    import torch
    if not torch.cuda.is_available():
        assert r.returncode == 1 and "[ERROR 5]" in r.stderr


def test_overlapping_parameters_rejected():
    """One parameter per (cell, nuclide, reaction): a second scatter, or a
    total beside a scatter, would replace the first's scaled density while
    both claimed its gradient."""
    base = decks.slab_deck(2.0, 0.5, 2, 100, params=("scatter", "capture"))
    mc.Problem.from_string(base)
    for extra in ("scatter", "total", "capture"):
        bad = base.replace("</parameters>", f'    <parameter cell="bin0" nuclide="onespeed" reaction="{extra}"/>\n'
                                            "</parameters>")
        expect_status(bad, mc.MC_ERR_INPUT)
    # the same reaction on another cell is a parameter of its own
    other = base.replace("</parameters>", '    <parameter cell="bin1" nuclide="onespeed" reaction="total"/>\n'
                                          "</parameters>")
    assert mc.Problem.from_string(other).n_params == 3
