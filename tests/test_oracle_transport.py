"""The oracle's whole random walk against closed forms and the analytic slab."""
import os

import numpy as np
import pytest

import slab_analytic as sa
from mc_testlib import decks, mc, oracle_run, shield_xs_dir

REF_SLAB = "/root/reference/examples/slab_analytic"


def test_slab_c0_small_vs_analytic():
    p = mc.Problem.from_string(decks.slab_deck(4.0, 0.5, 100, 20000))
    r = oracle_run(p, 20000)
    a = sa.slab(4.0, 0.5)
    zT = (r.mean[0] - a["T"]) / r.uncer[0]
    zR = (r.mean[1] - a["R"]) / r.uncer[1]
    assert abs(zT) < 4 and abs(zR) < 4
    z = (r.mean[2:] - a["flux"]) / r.uncer[2:]
    assert np.abs(z).max() < 4.5
    # balance holds per history in expectation: T + R + (1-c) sum(flux) = 1
    assert abs(r.mean[0] + r.mean[1] + 0.5 * r.mean[2:].sum() - 1.0) < 5 * r.uncer[2:].sum() * 0.5 + 0.01


def test_pure_absorber_two_slabs_known_answer():
    """Reference integration test: leakage exp(-4.2) within one sigma
    (test_integral_Simulator.cpp:20-28), on the deck of
    examples/slab_analytic/input.xml (kept as tests/golden/slab_analytic.xml)."""
    here = os.path.dirname(__file__)
    xml = open(os.path.join(here, "golden", "slab_analytic.xml")).read()
    p = mc.Problem.from_string(xml)
    n = 400000
    r = oracle_run(p, n)
    assert abs(r.mean[0] - 0.0149956) <= 2 * r.uncer[0]
    # pure absorber: each history scores 0 or 1
    assert abs(r.sq[0] - r.sum[0]) < 1e-9


def test_one_speed_c0_limit():
    p = mc.Problem.from_string(decks.slab_deck(3.0, 1e-12, 0, 50000))
    r = oracle_run(p, 50000)
    assert abs(r.mean[0] - np.exp(-3.0)) < 4 * r.uncer[0]
    assert r.mean[1] < 1e-9


def test_trace_is_deterministic_and_keyed():
    p = mc.Problem.from_string(decks.slab_deck(4.0, 0.7, 10, 1000))
    a = oracle_run(p, 50, key_offset=0, trace=50).trace
    b = oracle_run(p, 50, key_offset=0, trace=50).trace
    assert (a == b).all()
    c = oracle_run(p, 40, key_offset=10, trace=40).trace
    assert (c == a[10:50]).all()   # history k depends only on its key


def test_shield_energy_leakage_balance():
    """CE shield: every history ends by leakage or roulette; the energy-binned
    leakage summed over bins equals the two face leakages."""
    p = mc.Problem.from_string(decks.shield_deck(2000), shield_xs_dir())
    r = oracle_run(p, 2000)
    names = p.estimator_names()
    assert names == ["transmission", "leakage"]
    T = r.mean[0]
    leak = r.mean[1:].reshape(2, 100)
    assert abs(leak[1].sum() - T) < 1e-12   # all leakage energies lie inside the filter
    assert 0.0 < leak[0].sum() < 1.0
