"""mc_xs_lookup_noalloc (no device allocation: staging in the library's fixed
per-device array) against mc_xs_lookup, bit for bit, through the C ABI: two
calls in a row on each engine, the second with a smaller input, so the
second reads nothing the first left behind; sizes above the array's maximum
are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from mc_testlib import decks, mc, reference_shielding_deck, shield_xs_dir

pytestmark = pytest.mark.gpu

MAX_LOOKUPS = 1 << 19


def call(fn, eng, E, M):
    out = np.full(E.size, np.nan)
    st = fn(eng._h, E.ctypes.data_as(C.POINTER(C.c_double)), M.ctypes.data_as(C.POINTER(C.c_int32)),
            out.ctypes.data_as(C.POINTER(C.c_double)), E.size)
    assert st == mc.MC_OK, mc.lib.mc_last_error()
    return out


@pytest.mark.parametrize("deck", ["shield", "reference"])
def test_noalloc_lookup_equals_lookup(deck):
    if deck == "shield":
        p = mc.Problem.from_string(decks.shield_deck(1000), shield_xs_dir())
        lo, hi = 1e-5, 2e7
    else:
        p = reference_shielding_deck()
        lo, hi = 1e-6, 3e7
    eng = mc.Engine(p)
    nmat = p.head.n_materials
    rng = np.random.default_rng(17)
    for n in (MAX_LOOKUPS, 3001):
        print(f"{deck}: {n} lookups", flush=True)
        E = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
        M = rng.integers(-1, nmat, n).astype(np.int32)
        want = call(mc.lib.mc_xs_lookup, eng, E, M)
        got = call(mc.lib.mc_xs_lookup_noalloc, eng, E, M)
        assert np.isfinite(got).all()
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (n, bad[:5], got[bad[:5]], want[bad[:5]])


def test_noalloc_lookup_refuses_oversized_input():
    p = mc.Problem.from_string(decks.shield_deck(1000), shield_xs_dir())
    eng = mc.Engine(p)
    n = MAX_LOOKUPS + 1
    E = np.ones(n)
    M = np.zeros(n, dtype=np.int32)
    out = np.zeros(n)
    st = mc.lib.mc_xs_lookup_noalloc(eng._h, E.ctypes.data_as(C.POINTER(C.c_double)),
                                     M.ctypes.data_as(C.POINTER(C.c_int32)),
                                     out.ctypes.data_as(C.POINTER(C.c_double)), n)
    assert st == mc.MC_ERR_CAPACITY
    assert b"at most 524288" in mc.lib.mc_last_error()
    assert (out == 0).all()
