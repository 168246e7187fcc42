"""Flights per particle per event iteration (forward runs).

The flight kernel of a forward run collides in the same thread and lets a
surviving particle fly again (continuous energy: after a lookup at its new
energy), up to events_per_launch flights before the bank is refilled.  Each particle still takes its own events in order, so
every history stays bit-identical to the oracle's and the flight count is
exact whatever the setting; only the grouping of events into iterations
changes.
"""
import numpy as np
import pytest

from mc_testlib import decks, free_gas_slab, importance_slab, mc, oracle_run, shield_xs_dir

pytestmark = pytest.mark.gpu

FIELDS = ("rng", "flights", "x", "y", "z", "E", "score")


def _traces_equal(g, o, what):
    for f in FIELDS:
        bad = np.nonzero(g[f] != o[f])[0]
        assert bad.size == 0, f"{what}: {f} differs at histories {bad[:8].tolist()}"


@pytest.mark.parametrize("k", [1, 2, 7])
@pytest.mark.parametrize("drain", [1, -1])
def test_slab_flux_bins_any_events_per_launch(k, drain):
    """100 flux cells: most flights end on a surface, which stops the
    particle's run of events until the workgroup has crossed it."""
    n = 8000
    p = mc.Problem.from_string(decks.slab_deck(4.0, 0.9, 100, n))
    g = mc.Engine(p, bank_capacity=1 << 12, drain_divisor=drain, events_per_launch=k).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    _traces_equal(g.trace, o.trace, f"k={k} drain={drain}")
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.sq, o.sq, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("k", [1, 3, 8])
def test_shield_any_events_per_launch(k):
    """Continuous-energy shield: the interval and Sigma_t a further flight
    uses come from the lookup after the collision, the same values the lookup
    kernel gives."""
    n = 4000
    p = mc.Problem.from_string(decks.shield_deck(n), shield_xs_dir())
    g = mc.Engine(p, bank_capacity=1 << 12, events_per_launch=k).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    _traces_equal(g.trace, o.trace, f"k={k}")
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)


def test_free_gas_several_events_per_launch():
    n = 3000
    p = mc.Problem.from_string(free_gas_slab(n))
    g = mc.Engine(p, bank_capacity=1 << 11, events_per_launch=5, drain_divisor=-1).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    _traces_equal(g.trace, o.trace, "free gas")
    assert g.flights == o.flights


@pytest.mark.parametrize("k", [1, 8])
def test_batched_slabs_events_per_launch(k):
    """24 slabs of uneven history counts in one bank, batched and plain, with
    plain launches (no graph replay) and event iterations only."""
    L, c, N = decks.batched_slabs(24, seed=9)
    N = np.maximum(50, N // 200)
    probs = [mc.Problem.from_string(decks.slab_deck(float(l), float(cc), 0, int(m))) for l, cc, m in zip(L, c, N)]
    eng = mc.Engine(probs, bank_capacity=1 << 11, use_graph=False, drain_divisor=-1, events_per_launch=k)
    tc = int(N.min())
    b = eng.forward(trace=tc)
    pl = eng.forward(trace=tc, mode=mc.MC_RUN_PLAIN)
    assert b.flights == pl.flights
    off, flights = 0, 0
    for i, p in enumerate(probs):
        o = oracle_run(p, int(N[i]), key_offset=off, trace=tc)
        _traces_equal(b.trace[i * tc:(i + 1) * tc], o.trace, f"problem {i}")
        np.testing.assert_allclose(b.sum[eng.bin_range(i)], o.sum, rtol=1e-12, atol=1e-300)
        flights += o.flights
        off += int(N[i])
    assert b.flights == flights


def test_splitting_with_several_events_per_launch():
    """Importance doubling per bin: a run that can split takes one flight per
    particle per iteration whatever the setting (its copies wait in the ring
    for free slots, which appear once per iteration).  c = 0.7 with 4096
    slots: large families in a small bank (the copy ring applies
    back-pressure when they outgrow it)."""
    n = 6000
    p = mc.Problem.from_string(importance_slab(4.0, 0.7, 8, n, 2.0))
    eng = mc.Engine(p, bank_capacity=1 << 12, events_per_launch=5)
    g = eng.forward()
    print("copy ring", eng.copy_ring(), flush=True)
    o = oracle_run(p, n)
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.sq, o.sq, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("k,live", [(1, 1), (8, 1), (24, 4)])
def test_gradient_any_events_per_launch(k, live):
    """Gradient runs: at one flight per iteration the collision kernel takes
    the collisions; at several the flight kernel folds them in the same pass;
    with a quarter of the bank live (the default) or all of it.  The
    likelihood-ratio sums equal the oracle's either way."""
    n = 4000
    p = mc.Problem.from_string(decks.shield_deck(n, params="24"), shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12, grad_events_per_launch=k, grad_live_divisor=live)
    rng = np.random.default_rng(4)
    go = rng.normal(size=p.n_bins)
    theta = rng.uniform(0.9, 1.1, p.n_params)
    g = eng.backward(go, theta=theta)
    o = oracle_run(p, n, theta=theta, grad_out=go)
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.grad, o.grad, rtol=1e-9, atol=1e-13)
    assert g.flights == o.flights


def test_events_per_launch_out_of_range():
    p = mc.Problem.from_string(decks.slab_deck(1.0, 0.5, 0, 100))
    for kw in (dict(events_per_launch=65), dict(grad_events_per_launch=65), dict(grad_live_divisor=-1)):
        with pytest.raises(mc.MCError) as e:
            mc.Engine(p, **kw)
        assert e.value.status == mc.MC_ERR_ARGUMENT
