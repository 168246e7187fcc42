"""GPU importance splitting (population_control.cpp:21-49) against the oracle.

Each particle's walk is bit-identical to the oracle's (copy i of a split
starts (i + 1) 2^50 draws ahead of the parent's stream), so the total flight
count matches exactly.  A split family shares one tally row and its members
add to it with atomics in whatever order they run, so per-history sums differ
from the oracle's by rounding only: moment sums agree to 1e-12 relative.
A full-size run with importances rising through the slab is checked against
the deterministic slab solution within 3 sigma.
"""
import numpy as np
import pytest

import slab_analytic as sa
from mc_testlib import decks, importance_slab, mc, oracle_run, shield_xs_dir

pytestmark = pytest.mark.gpu


def _check(p, n, **kw):
    g = mc.Engine(p, **kw).forward()
    o = oracle_run(p, n)
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.sq, o.sq, rtol=1e-12, atol=1e-300)
    return g, o


# drain_divisor 1: the run's tail goes to the drain kernel; -1: event iterations only
@pytest.mark.parametrize("drain", [1, -1])
def test_split_doubling_matches_oracle(drain):
    """Importance doubles per bin: every deeper crossing splits in two."""
    n = 20000
    p = mc.Problem.from_string(importance_slab(4.0, 0.5, 8, n, 2.0))
    g, o = _check(p, n, bank_capacity=1 << 14, drain_divisor=drain)
    assert o.flights > 5 * n       # the deck does split (12 flights per history without)


def test_split_fractional_ratio_small_bank():
    """Ratio 1.5 (splits into one or two), strong scattering, and a bank of
    1024 slots: copies wait in the ring and take slots whose row a family
    still holds."""
    n = 5000
    p = mc.Problem.from_string(importance_slab(4.0, 0.9, 8, n, 1.5))
    _check(p, n, bank_capacity=1 << 10, iters_per_graph=5)


def test_split_plain_launches_and_batch():
    """The same deck through ordinary launches, and batched with a deck that
    does not split (rows of both kinds in one bank)."""
    n = 4000
    ps = [mc.Problem.from_string(importance_slab(3.0, 0.7, 6, n, 2.0)),
          mc.Problem.from_string(decks.slab_deck(3.0, 0.7, 6, n))]
    _check(ps[0], n, bank_capacity=1 << 12, use_graph=False)
    eng = mc.Engine(ps, bank_capacity=1 << 12)
    b = eng.forward()
    off = 0
    for i, p in enumerate(ps):
        o = oracle_run(p, n, key_offset=off)
        np.testing.assert_allclose(b.sum[eng.bin_range(i)], o.sum, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(b.sq[eng.bin_range(i)], o.sq, rtol=1e-12, atol=1e-300)
        off += n


def test_split_full_size_vs_analytic():
    """1e6 histories of the 4 cm, c = 0.5 slab with importance 1.03^k in flux
    bin k: T, R and the flux bins within 3 sigma of the deterministic slab."""
    n = 1000000
    p = mc.Problem.from_string(importance_slab(4.0, 0.5, 100, n, 1.03))
    g = mc.Engine(p).forward()
    a = sa.slab(4.0, 0.5)
    assert abs(g.mean[0] - a["T"]) < 3 * g.uncer[0]
    assert abs(g.mean[1] - a["R"]) < 3 * g.uncer[1]
    z = (g.mean[2:] - a["flux"]) / g.uncer[2:]
    assert np.mean(np.abs(z) > 3) < 0.03 and np.abs(z).max() < 4.5


@pytest.mark.parametrize("drain", [1, -1])
def test_split_gradient_matches_oracle(drain):
    """Gradient runs split too: a copy carries the particle's score function S
    (its H, Q and pending coefficients stay with it) and the family's members
    add their sums into one, closed out as the history's gradient."""
    n = 20000
    deck = decks.slab_deck(2.0, 0.6, 4, n, params=("scatter", "capture"))
    for k, imp in ((1, 2.0), (2, 3.0), (3, 1.5)):
        deck = deck.replace(f'<cell name="bin{k}" material="slab">',
                            f'<cell name="bin{k}" material="slab" importance="{imp!r}">')
    p = mc.Problem.from_string(deck)
    eng = mc.Engine(p, bank_capacity=1 << 12, drain_divisor=drain)
    go = np.zeros(p.n_bins)
    go[0], go[1], go[2:] = 1.0, -0.5, 0.25
    theta = np.array([1.05, 0.97])
    g = eng.backward(go, theta=theta)
    o = oracle_run(p, n, theta=theta, grad_out=go)
    assert o.flights > 1.5 * n * 2      # the deck does split
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.grad, o.grad, rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(g.grad_sigma, o.grad_sigma, rtol=1e-8)


def test_split_gradient_shield_matches_oracle():
    """CE shield with importance 1, 2, 4 in its three layers and 24 parameters
    (scatter and capture of every nuclide of every layer): copies carry the
    set of materials the particle has visited as well as S."""
    n = 3000
    deck = decks.shield_deck(n, params="24")
    for (m, _, _), imp in zip(decks.SHIELD_LAYERS, (1.0, 2.0, 4.0)):
        deck = deck.replace(f'<cell name="{m}" material="{m}">',
                            f'<cell name="{m}" material="{m}" importance="{imp!r}">')
    p = mc.Problem.from_string(deck, shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12)
    rng = np.random.default_rng(5)
    go = rng.normal(size=p.n_bins)
    theta = rng.uniform(0.9, 1.1, p.n_params)
    g = eng.backward(go, theta=theta)
    o = oracle_run(p, n, theta=theta, grad_out=go)
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.grad, o.grad, rtol=1e-9, atol=1e-13)
    f = eng.forward(theta=theta)
    assert f.flights == o.flights


def test_split_with_collision_and_track_length_estimators():
    """Splitting in a slab scored by track-length and collision estimators of
    every reaction (bins 2-11) besides T and R: members of a family add into
    one row, closed out once."""
    from mc_testlib import REACTION_SCORES
    n = 10000
    sc = " ".join(REACTION_SCORES)
    deck = importance_slab(3.0, 0.8, 6, n, 2.0)
    extra = "".join(f'    <estimator name="{t}" type="{t}" scores="{sc}">\n' +
                    "".join(f'        <cell name="bin{k}"/>\n' for k in range(6)) + "    </estimator>\n"
                    for t in ("TL", "C"))
    deck = deck.replace("</estimators>", extra + "</estimators>")
    p = mc.Problem.from_string(deck)
    _check(p, n, bank_capacity=1 << 13)


def test_split_with_free_gas_and_energy_tally():
    """Importance rising through a thermalizing hydrogen-carbon slab
    (free-gas scattering, energy-filtered reflection)."""
    from mc_testlib import free_gas_slab
    n = 3000
    # the slab as two cells of importance 1 and 2
    deck = free_gas_slab(n)
    deck = deck.replace('<plane_x name="px1" x="5.0"/>', '<plane_x name="pxm" x="2.5"/>\n    <plane_x name="px1" x="5.0"/>')
    deck = deck.replace("""    <cell name="slab" material="hc">
        <surface name="px0" sense="+1"/>
        <surface name="px1" sense="-1"/>
    </cell>""", """    <cell name="front" material="hc">
        <surface name="px0" sense="+1"/>
        <surface name="pxm" sense="-1"/>
    </cell>
    <cell name="back" material="hc" importance="2.0">
        <surface name="pxm" sense="+1"/>
        <surface name="px1" sense="-1"/>
    </cell>""")
    p = mc.Problem.from_string(deck)
    assert p.n_bins > 2
    _check(p, n, bank_capacity=1 << 12)


# Back-pressure of the copy ring: a split whose copies do not all fit leaves
# its particle waiting at the surface (before its roulette) until they do, and
# the engine doubles the ring once particles wait.  Importance x8 per bin makes
# families of tens of thousands of members from one history; a bank of 256
# slots starts with a ring of 2048 entries.
@pytest.mark.parametrize("drain", [1, -1])
def test_ring_back_pressure_matches_oracle(drain):
    n = 300
    p = mc.Problem.from_string(importance_slab(4.0, 0.5, 8, n, 8.0))
    eng = mc.Engine(p, bank_capacity=256, drain_divisor=drain)
    assert eng.copy_ring()["capacity"] == 0     # allocated by the first run
    g = eng.forward()
    o = oracle_run(p, n)
    print(f"flights {o.flights} ring {eng.copy_ring()}", flush=True)
    assert o.flights > 1000 * n                 # families far larger than the bank
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.sq, o.sq, rtol=1e-12, atol=1e-300)
    r = eng.copy_ring()
    assert r["waits"] > 0 and r["growths"] > 0 and r["capacity"] > 8 * 256
    # the grown ring stays: the same run again needs fewer waits, same sums
    g2 = eng.forward()
    assert g2.flights == o.flights
    np.testing.assert_allclose(g2.sum, o.sum, rtol=1e-12, atol=1e-300)
    assert eng.copy_ring()["waits"] < r["waits"]


def test_ring_back_pressure_gradient_matches_oracle():
    """Gradient run whose importance rises x8 per bin with a 64-slot bank: a
    waiting particle's copies carry its score function S as it was at the
    split (nothing moves it while the particle waits)."""
    n = 3000
    deck = decks.slab_deck(2.0, 0.6, 4, n, params=("scatter", "capture"))
    for k in range(1, 4):
        deck = deck.replace(f'<cell name="bin{k}" material="slab">',
                            f'<cell name="bin{k}" material="slab" importance="{8.0 ** k!r}">')
    p = mc.Problem.from_string(deck)
    eng = mc.Engine(p, bank_capacity=64)
    go = np.zeros(p.n_bins)
    go[0], go[1], go[2:] = 1.0, -0.5, 0.25
    theta = np.array([1.05, 0.97])
    g = eng.backward(go, theta=theta)
    o = oracle_run(p, n, theta=theta, grad_out=go)
    assert eng.copy_ring()["waits"] > 0
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(g.grad, o.grad, rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(g.grad_sigma, o.grad_sigma, rtol=1e-8)
    assert eng.forward(theta=theta).flights == o.flights


def test_ring_does_not_grow_without_waits():
    n = 4000
    p = mc.Problem.from_string(importance_slab(4.0, 0.5, 8, n, 1.5))
    eng = mc.Engine(p, bank_capacity=1 << 12)
    g = eng.forward()
    o = oracle_run(p, n)
    assert g.flights == o.flights
    assert eng.copy_ring() == {"capacity": 8 << 12, "growths": 0, "waits": 0}
