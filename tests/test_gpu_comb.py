"""The particle comb on the GPU (population_control.cpp:55-84 after each
particle's walk, handler.cpp:19-29) against the oracle.

A comb history runs its bank of split copies one after another in its own
slot, last in first out, so every history is the oracle's bit for bit: its
final random-number state, flight count, last position and energy and its
score.  Moment sums agree to rounding (they add the histories in a
launch-dependent order)."""
import numpy as np
import pytest

from mc_testlib import decks, importance_slab, mc, oracle_run
from test_comb import comb_slab

pytestmark = pytest.mark.gpu

FIELDS = ("rng", "flights", "x", "y", "z", "E", "score")


def check(p, n, **kw):
    print(f"comb deck n={n} {kw}", flush=True)
    eng = mc.Engine(p, **kw)
    g = eng.forward(trace=n)
    o = oracle_run(p, n, trace=n)
    for f in FIELDS:
        bad = np.nonzero(g.trace[f] != o.trace[f])[0]
        assert bad.size == 0, f"{f} differs at histories {bad[:8].tolist()}"
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.sq, o.sq, rtol=1e-12, atol=1e-300)
    return g, o


@pytest.mark.parametrize("drain", [1, -1])
@pytest.mark.parametrize("bank_max,teeth", [(8, 1), (8, 2), (4, 4), (10, 3)])
def test_comb_histories_bitexact(bank_max, teeth, drain):
    n = 6000
    p = mc.Problem.from_string(comb_slab(n, bank_max, teeth))
    g, o = check(p, n, bank_capacity=1 << 12, drain_divisor=drain)
    plain = oracle_run(mc.Problem.from_string(importance_slab(4.0, 0.5, 20, n, 1.5)), n)
    assert o.flights != plain.flights                 # the comb acted


def test_comb_with_several_events_per_launch_and_plain_launches():
    n = 5000
    p = mc.Problem.from_string(comb_slab(n, 8, 2))
    check(p, n, bank_capacity=1 << 11, events_per_launch=8)
    check(p, n, bank_capacity=1 << 11, use_graph=False)


def test_comb_batched_with_ring_splitting_and_plain_problems():
    """A comb problem (own banks), a splitting problem without a comb (the
    copy ring and split families) and one that cannot split, in one engine."""
    n = 4000
    ps = [mc.Problem.from_string(comb_slab(n, 8, 2)),
          mc.Problem.from_string(importance_slab(4.0, 0.5, 20, n, 1.5)),
          mc.Problem.from_string(decks.slab_deck(4.0, 0.5, 20, n))]
    eng = mc.Engine(ps, bank_capacity=1 << 12)
    g = eng.forward(trace=n)
    off = 0
    for i, p in enumerate(ps):
        o = oracle_run(p, n, key_offset=off, trace=n)
        np.testing.assert_allclose(g.sum[eng.bin_range(i)], o.sum, rtol=1e-12, atol=1e-300)
        if i != 1:                  # ring families write no trace records
            for f in FIELDS:
                assert (g.trace[f][i * n:(i + 1) * n] == o.trace[f]).all(), (i, f)
        off += n


def test_comb_errors():
    n = 1000
    deck = decks.slab_deck(4.0, 0.5, 20, n, params=("capture",))
    for k in range(1, 20):
        deck = deck.replace(f'<cell name="bin{k}" material="slab">',
                            f'<cell name="bin{k}" material="slab" importance="{1.5 ** k!r}">')
    deck = deck.replace("</simulation>", "</simulation>\n<population_control>\n"
                        '    <particle_comb bank_max="8" teeth="2"/>\n</population_control>', 1)
    p = mc.Problem.from_string(deck)
    assert p.n_params == 1 and p.head.comb_teeth == 2
    eng = mc.Engine(p, bank_capacity=1 << 10)
    # no likelihood-ratio gradient through the comb's selection
    with pytest.raises(mc.MCError) as e:
        eng.backward(np.ones(p.n_bins))
    assert e.value.status == mc.MC_ERR_UNSUPPORTED
    # a split into 2000 at one surface overflows a bank of 8 - 1 + 512 entries:
    # the copies that do not fit are counted and the run fails
    deck = comb_slab(n, 8, 2, nb=2, ratio=2000.0)
    q = mc.Problem.from_string(deck)
    with pytest.raises(mc.MCError) as e:
        mc.Engine(q, bank_capacity=1 << 10).forward()
    assert e.value.status == mc.MC_ERR_CAPACITY
