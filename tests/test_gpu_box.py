"""3D box shield with a spherical detector: planes, a sphere, a void region
and overlapping importance-0 outside cells; design variants batched in one
launch with their 80-parameter likelihood-ratio gradients."""
import numpy as np
import pytest

from mc_testlib import decks, mc, oracle_run, shield_xs_dir

pytestmark = pytest.mark.gpu

# thin shield and a large near detector so that small runs score often
THIN = dict(box_x=10.0, detector=(40.0, 0.0, 0.0, 20.0))


@pytest.mark.parametrize("drain", [1, -1])
def test_box_histories_bitexact(drain):
    n = 6000
    p = mc.Problem.from_string(decks.box_deck(n, **THIN), shield_xs_dir())
    g = mc.Engine(p, bank_capacity=1 << 12, drain_divisor=drain).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    for f in ("rng", "flights", "x", "y", "z", "E", "score"):
        assert (g.trace[f] == o.trace[f]).all(), f
    assert (o.trace["score"] > 0).sum() > 100        # the detector was reached
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-11)


def test_box_variants_batched_forward_and_gradient():
    """8 design variants (seeded per-(cell, nuclide) scales in [0.8, 1.2]) in
    one batched engine; each variant's detector response and 80-component
    gradient equal the oracle's for the same history keys."""
    nv, n = 8, 3000
    thetas = decks.box_variant_thetas(nv)
    probs = [mc.Problem.from_string(decks.box_deck(n, **THIN), shield_xs_dir()) for _ in range(nv)]
    eng = mc.Engine(probs, bank_capacity=1 << 13)
    assert eng.n_params == 80 * nv and eng.n_bins == nv
    theta = thetas.ravel()
    fwd = eng.forward(theta=theta)
    bwd = eng.backward(np.ones(eng.n_bins), theta=theta)
    np.testing.assert_allclose(bwd.sum, fwd.sum, rtol=1e-12)
    off = 0
    for v in range(nv):
        o = oracle_run(probs[v], n, key_offset=off, theta=thetas[v], grad_out=np.ones(1))
        sl, ps = eng.bin_range(v), eng.param_range(v)
        np.testing.assert_allclose(fwd.sum[sl], o.sum, rtol=1e-11)
        np.testing.assert_allclose(bwd.grad[ps], o.grad, rtol=1e-8, atol=1e-12)
        off += n
    # variants differ: their responses are not all equal
    assert np.ptp(fwd.mean) > 0
