"""Likelihood-ratio gradients of the oracle (the CPU reference gradient).

Pinned by mathematics independent of the Monte Carlo:
  * pure absorber: T = exp(-theta L Sigma_c), dT/dtheta = -L Sigma_c T;
  * one-speed slab: the deterministic slab solution T(tau, c') with optical
    thickness tau = Sigma_t(theta) L and c' = Sigma_s/Sigma_t, differentiated
    numerically (it is smooth and deterministic); track-length and collision
    flux and reaction-rate scores follow from absorptions = 1 - T - R;
  * common-random-number central differences of the oracle itself
    (per-history paired differences give their standard error);
  * exact linearity of the vector-Jacobian product in grad_out.
"""
import numpy as np
import pytest

import slab_analytic as sa
from mc_testlib import REACTION_SCORES, decks, mc, oracle_run, reaction_slab


def analytic_TR(L, sig_s, sig_c, n=256):
    tau, cp = (sig_s + sig_c) * L, sig_s / (sig_s + sig_c)
    hs, TR = [], []
    for m in (n, 2 * n, 4 * n):
        hs.append(tau / m)
        TR.append(sa._solve_nodes(tau, cp, m)[2:])
    return sa._extrapolate(hs, TR)


def analytic_grad(L, c, h=1e-2):
    """d(T, R)/d(theta_s, theta_c) at theta = 1 for Sigma_s = c, Sigma_c = 1 - c."""
    g = np.zeros((2, 2))
    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        p = analytic_TR(L, c * (1 + e[0]), (1 - c) * (1 + e[1]))
        q = analytic_TR(L, c * (1 - e[0]), (1 - c) * (1 - e[1]))
        g[:, k] = (np.asarray(p) - np.asarray(q)) / (2 * h)
    return g   # rows T, R; columns scatter, capture


def test_pure_absorber_closed_form():
    L, n = 2.0, 200000
    p = mc.Problem.from_string(decks.slab_deck(L, 0.0, 0, n, params=("capture",)))
    assert p.n_params == 1
    r = oracle_run(p, n, grad_out=np.array([1.0, 0.0]))
    T = r.mean[0]
    assert abs(T - np.exp(-L)) < 4 * r.uncer[0]
    assert abs(r.grad[0] - (-L * np.exp(-L))) < 4 * r.grad_sigma[0]
    # every transmitted history flew L - x0 (source at x0 = 1e-9) through
    # Sigma_c = 1: its score is -(L - x0), so the estimator is -(L - x0) * T
    assert abs(r.grad[0] - (-(L - 1e-9) * T)) < 1e-12


@pytest.mark.parametrize("L,c", [(2.0, 0.6), (1.0, 0.9)])
def test_one_speed_gradient_vs_deterministic(L, c):
    n = 200000
    p = mc.Problem.from_string(decks.slab_deck(L, c, 0, n, params=("scatter", "capture")))
    A = analytic_grad(L, c)
    for row, go in enumerate(([1.0, 0.0], [0.0, 1.0])):
        r = oracle_run(p, n, grad_out=np.array(go))
        z = (r.grad - A[row]) / r.grad_sigma
        assert np.all(np.abs(z) < 4), (row, r.grad, A[row], r.grad_sigma)


def test_vjp_is_linear_in_grad_out():
    n = 20000
    p = mc.Problem.from_string(decks.slab_deck(2.0, 0.7, 0, n, params=("scatter", "capture")))
    rT = oracle_run(p, n, grad_out=np.array([1.0, 0.0]))
    rR = oracle_run(p, n, grad_out=np.array([0.0, 1.0]))
    rM = oracle_run(p, n, grad_out=np.array([0.3, -1.7]))
    np.testing.assert_allclose(rM.grad, 0.3 * rT.grad - 1.7 * rR.grad, rtol=1e-9, atol=1e-12)


def test_gradient_vs_common_random_number_differences():
    """Central differences of T + R with the same history keys (h = 1e-3)."""
    n, h, L, c = 50000, 1e-3, 2.0, 0.6
    p = mc.Problem.from_string(decks.slab_deck(L, c, 0, n, params=("scatter", "capture")))
    r = oracle_run(p, n, grad_out=np.array([1.0, 1.0]))
    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        plus = oracle_run(p, n, theta=1 + e, trace=n).trace["score"]
        minus = oracle_run(p, n, theta=1 - e, trace=n).trace["score"]
        d = (plus - minus) / (2 * h)
        fd, fd_sig = d.mean(), d.std(ddof=1) / np.sqrt(n)
        assert abs(r.grad[k] - fd) < 3 * np.hypot(r.grad_sigma[k], fd_sig), (k, r.grad[k], fd, fd_sig)


def reaction_rates_and_grad(L, c):
    """Expected flux and reaction rates of the slab and their theta derivatives.

    With a unit source, absorptions A = 1 - T - R; flux = A / Sigma_c and a
    reaction rate is Sigma_x * flux.  Sigma_s = theta_s c, Sigma_c = theta_c (1 - c)."""
    T, R = analytic_TR(L, c, 1.0 - c)
    dTR = analytic_grad(L, c)                       # rows T, R; columns theta_s, theta_c
    sig_s, sig_c = c, 1.0 - c
    A = 1.0 - T - R
    dA = -(dTR[0] + dTR[1])
    F = A / sig_c
    dF = np.array([dA[0] / sig_c, (dA[1] - A) / sig_c])
    dsig = {"s": np.array([sig_s, 0.0]), "c": np.array([0.0, sig_c])}
    val = {"flux": F, "absorption": A, "capture": A, "scatter": sig_s * F, "total": (sig_s + sig_c) * F}
    der = {"flux": dF, "absorption": dA, "capture": dA,
           "scatter": dsig["s"] * F + sig_s * dF,
           "total": (dsig["s"] + dsig["c"]) * F + (sig_s + sig_c) * dF}
    return val, der


def test_track_length_and_collision_score_gradients_vs_deterministic():
    """Gradient terms of track-length scores (flight density, d ln Sigma_x) and
    collision scores (1/Sigma_t kernel) against the deterministic slab."""
    L, c, n = 2.0, 0.6, 200000
    p = mc.Problem.from_string(reaction_slab(L, c, n))
    assert p.n_bins == 12
    val, der = reaction_rates_and_grad(L, c)
    for e0 in (2, 7):
        for j, name in enumerate(REACTION_SCORES):
            go = np.zeros(p.n_bins)
            go[e0 + j] = 1.0
            r = oracle_run(p, n, grad_out=go)
            assert abs(r.mean[e0 + j] - val[name]) < 4 * r.uncer[e0 + j], (e0, name)
            z = (r.grad - der[name]) / r.grad_sigma
            assert np.all(np.abs(z) < 4), (e0, name, r.grad, der[name], r.grad_sigma)


def test_gradient_with_importance_splitting_vs_common_random_numbers():
    """Split copies carry the score function; their scores add to the history's
    gradient.  Checked against paired central differences of the sum of all
    bins (T, R and the two cells' track-length flux), the traced history score."""
    n, h, L, c = 40000, 1e-3, 2.0, 0.6
    deck = decks.slab_deck(L, c, 2, n, params=("scatter", "capture"))
    deck = deck.replace('<cell name="bin1" material="slab">', '<cell name="bin1" material="slab" importance="3.0">')
    p = mc.Problem.from_string(deck)
    r = oracle_run(p, n, grad_out=np.ones(p.n_bins))
    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        tp = oracle_run(p, n, theta=1 + e, trace=n).trace
        tm = oracle_run(p, n, theta=1 - e, trace=n).trace
        d = (tp["score"] - tm["score"]) / (2 * h)
        fd, fd_sig = d.mean(), d.std(ddof=1) / np.sqrt(n)
        assert abs(r.grad[k] - fd) < 3.5 * np.hypot(r.grad_sigma[k], fd_sig), (k, r.grad[k], fd, fd_sig)


def test_shield_params_and_materials():
    from mc_testlib import shield_xs_dir
    p = mc.Problem.from_string(decks.shield_deck(1000, params="24"), shield_xs_dir())
    assert p.n_params == 24
    # three layer materials cloned per cell (parameters act per cell)
    assert p.head.n_materials == 6
    r = oracle_run(p, 1000, grad_out=np.eye(p.n_bins)[0])
    assert np.all(np.isfinite(r.grad))
