"""GPU parity: the HIP engine through the C ABI against the oracle.

Bit-exact per history (integer state and every FP64 field of the trace) on
small seeded runs; tally moment sums agree to 1e-12 relative (the GPU adds
histories in a different order); full-size runs are checked against the
analytic slab within 3 sigma (sigma from both estimates combined).
"""
import math

import numpy as np
import pytest

import slab_analytic as sa
from mc_testlib import decks, mc, oracle, oracle_run, shield_xs_dir

pytestmark = pytest.mark.gpu

FIELDS = ("rng", "flights", "x", "y", "z", "E", "score")


def assert_traces_equal(g, o, what=""):
    for f in FIELDS:
        bad = np.nonzero(g[f] != o[f])[0]
        assert bad.size == 0, f"{what}: field {f} differs at histories {bad[:8].tolist()} " \
                              f"(gpu {g[f][bad[:3]]}, oracle {o[f][bad[:3]]})"


def test_device_math_bitexact():
    import ctypes as C
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(200000), np.exp(rng.uniform(-40, 40, 20000)),
                        [1.0, 0.5, 0.25, 0.75, 2.0 ** -63, 1e-300]])
    d = mc.selftest_math(x)
    c, s = C.c_double(), C.c_double()
    for i in range(0, x.size, 7):
        xi = float(x[i])
        assert d[i, 0] == oracle.mco_log(xi), (xi, d[i, 0])
        assert d[i, 3] == math.sqrt(xi), (xi, d[i, 3])
        if xi <= 1.0:
            oracle.mco_sincos_2pi(xi, C.byref(c), C.byref(s))
            assert d[i, 1] == c.value and d[i, 2] == s.value, xi


# drain_divisor 1: the run's tail goes to the drain kernel; -1: event iterations only
@pytest.mark.parametrize("drain", [1, -1])
def test_slab_histories_bitexact(drain):
    n = 20000
    p = mc.Problem.from_string(decks.slab_deck(4.0, 0.5, 100, n))
    g = mc.Engine(p, bank_capacity=1 << 13, drain_divisor=drain).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    assert_traces_equal(g.trace, o.trace, "slab")
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.sq, o.sq, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("graph", [True, False])
def test_slab_graph_and_plain_launches_agree(graph):
    n = 5000
    p = mc.Problem.from_string(decks.slab_deck(2.0, 0.9, 20, n))
    g = mc.Engine(p, bank_capacity=1 << 12, use_graph=graph, iters_per_graph=7).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    assert_traces_equal(g.trace, o.trace, f"graph={graph}")


def test_slab_full_size_vs_analytic():
    """1e6 histories of the 4 cm, c = 0.5 slab: T, R and all 100 flux bins
    within 3 sigma of the deterministic solution (family check on the bins)."""
    n = 1000000
    p = mc.Problem.from_string(decks.slab_deck(4.0, 0.5, 100, n))
    g = mc.Engine(p).forward()
    a = sa.slab(4.0, 0.5)
    assert abs(g.mean[0] - a["T"]) < 3 * g.uncer[0]
    assert abs(g.mean[1] - a["R"]) < 3 * g.uncer[1]
    z = (g.mean[2:] - a["flux"]) / g.uncer[2:]
    assert np.mean(np.abs(z) > 3) < 0.03 and np.abs(z).max() < 4.5
    # balance T + R + (1-c) sum(flux) = 1 holds in expectation
    assert abs(g.mean[0] + g.mean[1] + 0.5 * g.mean[2:].sum() - 1.0) < 3e-3


@pytest.mark.parametrize("drain", [1, -1])
def test_shield_histories_bitexact(drain):
    n = 6000
    p = mc.Problem.from_string(decks.shield_deck(n), shield_xs_dir())
    g = mc.Engine(p, bank_capacity=1 << 12, drain_divisor=drain).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    assert_traces_equal(g.trace, o.trace, "shield")
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-11, atol=1e-300)


def test_xs_lookup_bitexact():
    p = mc.Problem.from_string(decks.shield_deck(10), shield_xs_dir())
    eng = mc.Engine(p)
    rng = np.random.default_rng(4)
    n = 20000
    E = np.exp(rng.uniform(np.log(1e-7), np.log(5e7), n))
    E[:50] = decks.shield_grid()[rng.integers(0, decks.GRID_POINTS, 50)]   # exactly on grid points
    M = rng.integers(0, p.head.n_materials, n).astype(np.int32)
    got = eng.xs_lookup(E, M)
    for i in range(0, n, 3):
        assert got[i] == oracle.mco_material_xs(p.desc_ptr, None, int(M[i]), 0, float(E[i])), i


def test_batched_equals_plain_equals_oracle():
    L, c, N = decks.batched_slabs(24, seed=5)
    N = np.maximum(50, N // 200)
    probs = [mc.Problem.from_string(decks.slab_deck(float(l), float(cc), 0, int(k))) for l, cc, k in zip(L, c, N)]
    eng = mc.Engine(probs, bank_capacity=1 << 12)
    tc = int(N.min())
    b = eng.forward(trace=tc)
    pl = eng.forward(trace=tc, mode=mc.MC_RUN_PLAIN)
    assert_traces_equal(b.trace, pl.trace, "batched/plain")
    off = 0
    for i, p in enumerate(probs):
        o = oracle_run(p, int(N[i]), key_offset=off, trace=tc)
        assert_traces_equal(b.trace[i * tc:(i + 1) * tc], o.trace, f"problem {i}")
        sl = eng.bin_range(i)
        np.testing.assert_allclose(b.sum[sl], o.sum, rtol=1e-12, atol=1e-300)
        off += int(N[i])


# drain_divisor 1: the run's tail goes to the drain kernel; -1: event iterations only
@pytest.mark.parametrize("drain", [1, -1])
def test_free_gas_histories_bitexact(drain):
    """Free-gas scattering (Reaction.cpp:27-118): a 1 eV beam thermalizing in
    a hydrogen-carbon slab, every history bit-exact against the oracle."""
    from mc_testlib import free_gas_slab
    n = 5000
    p = mc.Problem.from_string(free_gas_slab(n))
    g = mc.Engine(p, bank_capacity=1 << 12, drain_divisor=drain).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    assert_traces_equal(g.trace, o.trace, "free gas")
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.sq, o.sq, rtol=1e-12, atol=1e-300)


def test_graph_replay_after_trace_count_changes():
    """A captured graph holds the trace record stride: a second forward with
    a smaller trace count on the same engine (two problems, graphs on) must
    recapture, and each problem's records equal the oracle's."""
    n = 3000
    probs = [mc.Problem.from_string(decks.slab_deck(2.0, c, 8, n)) for c in (0.3, 0.8)]
    eng = mc.Engine(probs, bank_capacity=1 << 11, use_graph=True)
    for tc in (100, 50, 120):
        g = eng.forward(trace=tc)
        for p in range(2):
            o = oracle_run(probs[p], tc, key_offset=p * n, trace=tc)
            assert_traces_equal(g.trace[p * tc:(p + 1) * tc], o.trace, f"trace_count={tc} problem {p}")
