"""GPU likelihood-ratio backward against the oracle, finite differences and
the torch.autograd bridge."""
import numpy as np
import pytest

from mc_testlib import decks, mc, oracle_run, reaction_slab, shield_xs_dir

pytestmark = pytest.mark.gpu


def test_slab_backward_matches_oracle():
    n = 50000
    p = mc.Problem.from_string(decks.slab_deck(2.0, 0.6, 4, n, params=("scatter", "capture")))
    eng = mc.Engine(p, bank_capacity=1 << 14)
    go = np.zeros(p.n_bins)
    go[0], go[1], go[2:] = 1.0, -0.5, 0.25
    g = eng.backward(go)
    o = oracle_run(p, n, grad_out=go)
    np.testing.assert_allclose(g.grad, o.grad, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(g.grad_sigma, o.grad_sigma, rtol=1e-8)
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12)


def test_reaction_score_backward_matches_oracle():
    """Track-length and collision estimators of flux and every reaction rate
    (pending score coefficients folded at collisions and surfaces)."""
    n = 50000
    p = mc.Problem.from_string(reaction_slab(2.0, 0.6, n))
    eng = mc.Engine(p, bank_capacity=1 << 14)
    go = np.random.default_rng(3).normal(size=p.n_bins)
    theta = np.array([1.07, 0.95])
    g = eng.backward(go, theta=theta)
    o = oracle_run(p, n, theta=theta, grad_out=go)
    np.testing.assert_allclose(g.grad, o.grad, rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-12)


@pytest.mark.parametrize("drain", [1, -1])
def test_shield_backward_matches_oracle(drain):
    n = 6000
    p = mc.Problem.from_string(decks.shield_deck(n, params="24"), shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12, drain_divisor=drain)
    rng = np.random.default_rng(2)
    go = rng.normal(size=p.n_bins)
    theta = rng.uniform(0.9, 1.1, p.n_params)
    g = eng.backward(go, theta=theta)
    o = oracle_run(p, n, theta=theta, grad_out=go)
    np.testing.assert_allclose(g.grad, o.grad, rtol=1e-9, atol=1e-13)
    f = eng.forward(theta=theta, trace=n)
    ot = oracle_run(p, n, theta=theta, trace=n)
    for k in ("rng", "flights", "score", "E"):
        assert (f.trace[k] == ot.trace[k]).all(), k


def test_shield_gradient_vs_crn_finite_differences():
    """d T / d theta along a random direction v: likelihood ratio vs common-
    random-number central differences (h = 1e-3), per-history paired."""
    n, h = 2_000_000, 1e-3
    p = mc.Problem.from_string(decks.shield_deck(n, params="24", leakage=False), shield_xs_dir())
    assert p.n_bins == 1                              # total transmitted fraction only
    eng = mc.Engine(p)
    rng = np.random.default_rng(5)
    v = rng.normal(size=p.n_params)
    v /= np.linalg.norm(v)
    b = eng.backward(np.ones(1), theta=np.ones(p.n_params))
    # the projection's variance needs the per-history vector; the sum of
    # |v_k| sigma_k bounds its standard error
    lr = float(b.grad @ v)
    lr_sig = float(np.abs(v) @ b.grad_sigma)
    plus = eng.forward(theta=1 + h * v, trace=n).trace
    minus = eng.forward(theta=1 - h * v, trace=n).trace
    d = (plus["score"] - minus["score"]) / (2 * h)   # per-history transmission difference
    fd, fd_sig = d.mean(), d.std(ddof=1) / np.sqrt(n)
    assert abs(lr - fd) < 3 * np.hypot(lr_sig, fd_sig), (lr, lr_sig, fd, fd_sig)


def test_autograd_function_matches_backward():
    torch = pytest.importorskip("torch")
    n = 20000
    p = mc.Problem.from_string(decks.slab_deck(2.0, 0.6, 0, n, params=("scatter", "capture")))
    eng = mc.Engine(p)
    F = mc.transport_function()
    theta = torch.tensor([1.05, 0.97], dtype=torch.float64, requires_grad=True)
    out = F.apply(theta, eng)
    w = torch.tensor([2.0, -1.0], dtype=torch.float64)
    (out * w).sum().backward()
    ref = eng.backward(w.numpy(), theta=theta.detach().numpy())
    np.testing.assert_allclose(theta.grad.numpy(), ref.grad, rtol=1e-12)
    np.testing.assert_allclose(out.detach().numpy(), ref.mean, rtol=1e-12)
