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


def _desc_arrays(p):
    """Raw bytes of every array of a problem's mc_problem_desc."""
    import numpy as np
    h = p.head
    sizes = {"surfaces": (h.n_surfaces, 80), "cells": (h.n_cells, 40), "materials": (h.n_materials, 16),
             "nuclides": (h.n_nuclides, 48), "sources": (h.n_sources, 88), "estimators": (h.n_estimators, 92),
             "params": (h.n_params, 16), "source_cdf": (h.n_sources + 1, 8)}
    out = {}
    for name, (n, sz) in sizes.items():
        ptr = getattr(h, name)
        out[name] = C.string_at(ptr, n * sz) if n and ptr else b""
    return out


@pytest.mark.parametrize("deck", ["slab", "shield24", "box", "reference"])
def test_problem_create_round_trip(deck):
    """mc_problem_create on another problem's flat arrays gives the same
    problem: the same records, and the oracle runs it to the same sums."""
    import numpy as np
    from mc_testlib import oracle_run, reference_shielding_deck
    if deck == "slab":
        p = mc.Problem.from_string(decks.slab_deck(3.0, 0.7, 10, 500, params=("scatter", "capture")))
    elif deck == "shield24":
        p = mc.Problem.from_string(decks.shield_deck(500, params="24"), shield_xs_dir())
    elif deck == "box":
        p = mc.Problem.from_string(decks.box_deck(500), shield_xs_dir())
    else:
        p = reference_shielding_deck()
    q = mc.Problem.from_desc(p.desc_ptr)
    assert _desc_arrays(p) == _desc_arrays(q)
    for f in ("n_bins", "n_params", "n_histories", "seed", "wr", "ws"):
        assert getattr(p.head, f) == getattr(q.head, f), f
    go = np.ones(p.n_bins) if p.n_params else None
    a, b = oracle_run(p, 300, grad_out=go), oracle_run(q, 300, grad_out=go)
    assert (a.sum == b.sum).all() and (a.sq == b.sq).all() and a.flights == b.flights
    assert (a.gsum == b.gsum).all()
    assert q.estimator_names()[0] == "Estimator #1"


def test_problem_create_checks_indices():
    import numpy as np
    p = mc.Problem.from_string(decks.slab_deck(2.0, 0.5, 4, 100))
    h = p.head

    def with_array(field, dtype, n, patch):
        arr = np.ctypeslib.as_array((C.c_byte * (n * np.dtype(dtype).itemsize)).from_address(getattr(h, field)))
        arr = arr.view(dtype).copy()
        patch(arr)
        d = mc.ProblemDesc.from_buffer_copy(h)
        setattr(d, field, arr.ctypes.data)
        out = C.c_void_p()
        st = mc.lib.mc_problem_create(C.byref(d), C.byref(out))
        if st == 0:
            mc.lib.mc_problem_destroy(out)
        return st, mc.lib.mc_last_error().decode()

    ncs = 2 * h.n_cells
    st, msg = with_array("cell_surf", np.int32, ncs, lambda a: a.__setitem__(3, 999))
    assert st == mc.MC_ERR_INPUT and "cell_surf[3]" in msg
    st, msg = with_array("cell_sense", np.int32, ncs, lambda a: a.__setitem__(0, 2))
    assert st == mc.MC_ERR_INPUT and "cell_sense[0]" in msg
    st, msg = with_array("mat_nuclide", np.int32, 1, lambda a: a.__setitem__(0, 5))
    assert st == mc.MC_ERR_INPUT and "mat_nuclide[0]" in msg
    d = mc.ProblemDesc.from_buffer_copy(h)
    d.n_bins = h.n_bins + 1
    out = C.c_void_p()
    assert mc.lib.mc_problem_create(C.byref(d), C.byref(out)) == mc.MC_ERR_INPUT
    d = mc.ProblemDesc.from_buffer_copy(h)
    d.seed = 2
    assert mc.lib.mc_problem_create(C.byref(d), C.byref(out)) == mc.MC_ERR_INPUT
    assert mc.lib.mc_problem_create(None, C.byref(out)) == mc.MC_ERR_ARGUMENT
