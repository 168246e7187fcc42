"""The reference's own cross-section files and fixed-source deck load and run.

xs_library (kept under tests/golden/shielding_vReduction/xs_library, data
files of the reference) holds tables that repeat an energy where a cross
section jumps (008016: 1363 repeats, 006000: 235, 005010: 53, 002003: 1),
one file whose energies decrease twice by rounding (005011 at 19900 ->
19862.03 eV and 2.41e6 -> 2407275 eV) and rows of 5 or 6 columns.  The
reference's binary_search (Algorithm.cpp:46-64) and XSTable::xs
(XSec.cpp:9-40) give every such table a value at every energy; the loader
keeps the tables as they stand and the oracle's lookup must equal the
reference's own code on them, bit for bit.
"""
import ctypes as C
import os

import numpy as np
import pytest

from mc_testlib import REF_LIB, REFERENCE_DECK, dptr, mc, oracle, oracle_run, ref_lib, reference_shielding_deck

XS_DIR = os.path.join(REFERENCE_DECK, "xs_library")


def read_table(zaid):
    rows = [l.split() for l in open(os.path.join(XS_DIR, zaid + ".txt")) if l.strip()][4:]
    t = np.array([[float(v) for v in r[:3]] for r in rows])
    return (np.ascontiguousarray(t[:, 0]), np.ascontiguousarray(t[:, 1]), np.ascontiguousarray(t[:, 2]))


def probe_energies(E, rng):
    """Every tabulated energy, its neighbours in ulps, midpoints and random
    energies over and beyond the grid."""
    E = np.asarray(E)
    mids = 0.5 * (E[:-1] + E[1:])
    near = np.concatenate([np.nextafter(E, 0), np.nextafter(E, np.inf)])
    return np.concatenate([E, near, mids, np.exp(rng.uniform(np.log(E.min()) - 1, np.log(E.max()) + 1, 4000))])


def test_reference_tables_have_repeats_and_decreases():
    E, _, _ = read_table("008016")
    assert (np.diff(E) == 0).sum() == 1363
    E, _, _ = read_table("005011")
    assert (np.diff(E) < 0).sum() == 2


def test_reference_deck_loads_and_runs():
    p = reference_shielding_deck()
    assert p.n_bins == 2 and p.n_histories == 10 ** 8
    r = oracle_run(p, 3000, trace=3000)
    assert r.flights > 3000 and np.isfinite(r.sum).all()
    assert (r.trace["score"] > 0).sum() > 0          # the detector was reached


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built")
@pytest.mark.parametrize("zaid", ["001001", "002003", "005010", "005011", "006000", "008016"])
def test_nuclide_xs_equals_reference_xstable(zaid):
    """The loaded problem's nuclide cross sections (oracle lookup on the
    loader's arrays) against the reference's XSTable::xs with its own
    binary_search on the same file, at every tabulated energy and around it."""
    ref = ref_lib()
    p = reference_shielding_deck()
    names = {"008016": 0, "001001": 1, "002003": 2, "006000": 3, "005010": 4, "005011": 5}
    k = names[zaid]
    E, s, c = read_table(zaid)
    rng = np.random.default_rng(int(zaid))
    for x in probe_energies(E, rng):
        want_s = ref.ref_xstable(dptr(E), dptr(s), E.size, x)
        want_c = ref.ref_xstable(dptr(E), dptr(c), E.size, x)
        assert oracle.mco_nuclide_xs(p.desc_ptr, k, 1, x) == want_s, (zaid, x)
        assert oracle.mco_nuclide_xs(p.desc_ptr, k, 2, x) == want_c, (zaid, x)


def test_five_and_six_column_rows():
    p = reference_shielding_deck()
    E1, s1, _ = read_table("001001")                 # 5 columns
    E2, s2, _ = read_table("005010")                 # 6 columns
    assert oracle.mco_nuclide_xs(p.desc_ptr, 1, 1, E1[10]) == s1[10]
    assert oracle.mco_nuclide_xs(p.desc_ptr, 4, 1, E2[10]) == s2[10]
