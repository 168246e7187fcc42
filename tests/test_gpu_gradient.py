"""GPU likelihood-ratio backward against the oracle, finite differences and
the torch.autograd bridge."""
import numpy as np
import pytest

from mc_testlib import decks, mc, oracle_run, reaction_slab, shield_xs_dir

pytestmark = pytest.mark.gpu


# workgroups of 512 and 1024 threads: the surface-crossing kernel runs at 256
# threads per workgroup and finds each flight workgroup's queue segment from
# the flight's workgroup size
@pytest.mark.parametrize("block", [256, 512, 1024])
def test_slab_backward_matches_oracle(block):
    n = 50000
    p = mc.Problem.from_string(decks.slab_deck(2.0, 0.6, 4, n, params=("scatter", "capture")))
    eng = mc.Engine(p, bank_capacity=1 << 14, block=block)
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


@pytest.mark.parametrize("drain,block", [(1, 256), (-1, 256), (1, 512)])
def test_shield_backward_matches_oracle(drain, block):
    n = 6000
    p = mc.Problem.from_string(decks.shield_deck(n, params="24"), shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12, drain_divisor=drain, block=block)
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


def crn_central_differences(eng, n, P, h, theta0=None):
    """Per-component common-random-number central differences of the run's
    summed tallies: for component k, each history's score (its trace
    record: the sum of its tally bins) at theta0 + h e_k and theta0 - h e_k,
    same key, same random numbers.  Returns the mean and paired standard
    error per component and the number of histories whose score changed
    discontinuously (present at one side only, or by more than 10 %): the
    crossing, collision and roulette decisions the step flipped."""
    th0 = np.ones(P) if theta0 is None else theta0
    fd, sig, flips = np.empty(P), np.empty(P), np.empty(P, dtype=np.int64)
    for k in range(P):
        e = np.zeros(P)
        e[k] = h
        plus = eng.forward(theta=th0 + e, trace=n).trace["score"]
        minus = eng.forward(theta=th0 - e, trace=n).trace["score"]
        d = (plus - minus) / (2 * h)
        fd[k], sig[k] = d.mean(), d.std(ddof=1) / np.sqrt(n)
        big = np.abs(plus - minus) > 0.1 * np.maximum(np.abs(plus), np.abs(minus))
        flips[k] = np.count_nonzero(((plus == 0) != (minus == 0)) | big)
    return fd, sig, flips


MIN_FLIPS = 20


@pytest.mark.parametrize("response", ["transmission", "leakage"])
def test_shield_gradient_vs_crn_finite_differences(response):
    """The 24 components of d(response)/d(theta) of the differentiable shield
    (3 layers x 4 nuclides x {scatter, capture} scale factors, theta = 1)
    against common-random-number central differences, h = 1e-3, within
    2 sigma, sigma = sqrt(sigma_LR^2 + sigma_FD^2) with the FD's paired
    per-history standard error.

    response: the transmitted fraction, or the summed energy-binned leakage
    (both faces, all 100 bins).

    A CRN difference sees the part of a derivative that acts through
    discrete decisions (a flight that crosses a surface instead of
    colliding, a nuclide choice, a roulette survival) only in the histories
    whose decision the step flips.  A component whose flips the sample
    holds (MIN_FLIPS or more: the scatter factors of the nuclides that make
    up each layer) is held to the 2-sigma bar.  One whose step flips none
    (the capture factors: 1/v capture is ~1e-5 of the total cross section at
    the energies that transmit, and even a step of 0.5 flips nothing in
    1e5 histories; the scatter factors of trace nuclides) gets from the
    differences only the smooth part of its derivative: the part carried by
    the crossing probabilities exp(-Sigma_t d), which the likelihood ratio
    holds as -d dSigma_t/dtheta, is missing there (the differences of the
    water capture factors are 0.43 of the likelihood ratio's on the
    transmitted fraction), so the two are not compared; those
    likelihood-ratio values are held to the oracle's to 1e-9
    (test_shield_backward_matches_oracle), and capture factors to
    differences in test_capture_gradient_vs_crn_finite_differences, where
    capture decides many crossings."""
    n, h = 8_000_000, 1e-3
    deck = decks.shield_deck(n, params="24", leakage=response == "leakage",
                             transmission=response == "transmission")
    p = mc.Problem.from_string(deck, shield_xs_dir())
    eng = mc.Engine(p)
    P = p.n_params
    assert P == 24
    b = eng.backward(np.ones(p.n_bins), theta=np.ones(P))
    fd, fd_sig, flips = crn_central_differences(eng, n, P, h)
    sig = np.hypot(b.grad_sigma, fd_sig)
    z = (b.grad - fd) / np.where(sig > 0, sig, 1.0)
    checked = flips >= MIN_FLIPS
    for k in range(P):
        print(f"{response} theta_{k:02d}: LR {b.grad[k]:+.5e} +- {b.grad_sigma[k]:.2e}  "
              f"FD {fd[k]:+.5e} +- {fd_sig[k]:.2e}  flips {flips[k]:6d}  z {z[k]:+.2f}"
              f"{'' if checked[k] else '  (no flips: not compared)'}", flush=True)
    # the scatter factors of each layer's main nuclides (water H and O, iron
    # Fe, concrete H, C and O) change enough decisions to be compared
    assert checked[[0, 4, 14, 16, 18, 20]].all(), flips
    assert np.all(np.abs(z[checked]) <= 2.0), z[checked]


def capture_slab_deck(n):
    """A thermal beam (0.0253 eV) on a 1 cm boron-carbon slab of the
    reference's own B-10, B-11 and carbon data (xs_library 005010, 005011,
    006000; B-10 at ten times its density in the reference's B4C, so that
    capture is 40 % of the total cross section), target-at-rest scattering:
    B-10 capture decides many crossings, so CRN differences of its scale
    factor flip many decisions."""
    from mc_testlib import REFERENCE_DECK
    return f"""<?xml version='1.0' encoding='UTF-8'?>
<simulation>
    <description name="capture slab" samples="{n}" seed="1"/>
</simulation>
<nuclides>
    <nuclide name="B10" ZAID="005010" kinematics="target_at_rest"/>
    <nuclide name="B11" ZAID="005011" kinematics="target_at_rest"/>
    <nuclide name="C0" ZAID="006000" kinematics="target_at_rest"/>
</nuclides>
<materials>
    <material name="B4C">
        <nuclide name="B10" density="0.219716"/>
        <nuclide name="B11" density="0.0878864"/>
        <nuclide name="C0" density="0.027468"/>
    </material>
</materials>
<surfaces>
    <plane_x name="px0" x="0.0"/>
    <plane_x name="px1" x="1.0"/>
</surfaces>
<cells>
    <cell name="slab" material="B4C">
        <surface name="px0" sense="+1"/>
        <surface name="px1" sense="-1"/>
    </cell>
    <cell name="left" importance="0.0">
        <surface name="px0" sense="-1"/>
    </cell>
    <cell name="right" importance="0.0">
        <surface name="px1" sense="+1"/>
    </cell>
</cells>
<distributions>
    <delta name="beam" datatype="point" x="1.0" y="0.0" z="0.0"/>
    <delta name="e" datatype="double" val="0.0253"/>
</distributions>
<sources>
    <point x="1e-9" y="0.0" z="0.0" direction="beam" energy="e"/>
</sources>
<estimators>
    <estimator name="transmission" scores="cross">
        <surface name="px1"/>
    </estimator>
</estimators>
<parameters>
    <parameter cell="slab" nuclide="B10" reaction="scatter"/>
    <parameter cell="slab" nuclide="B10" reaction="capture"/>
    <parameter cell="slab" nuclide="B11" reaction="scatter"/>
    <parameter cell="slab" nuclide="B11" reaction="capture"/>
    <parameter cell="slab" nuclide="C0" reaction="scatter"/>
    <parameter cell="slab" nuclide="C0" reaction="capture"/>
</parameters>
"""


def test_capture_gradient_vs_crn_finite_differences():
    """Capture and scatter scale factors of tabulated nuclides where the
    differences flip many decisions: every component whose flips the sample
    holds (B-10 capture and every scatter factor among them) within 2 sigma
    of CRN central differences (h = 1e-3)."""
    from mc_testlib import REFERENCE_DECK
    import os
    n, h = 8_000_000, 1e-3
    p = mc.Problem.from_string(capture_slab_deck(n), os.path.join(REFERENCE_DECK, "xs_library"))
    eng = mc.Engine(p)
    P = p.n_params
    b = eng.backward(np.ones(p.n_bins), theta=np.ones(P))
    fd, fd_sig, flips = crn_central_differences(eng, n, P, h)
    sig = np.hypot(b.grad_sigma, fd_sig)
    z = (b.grad - fd) / sig
    for k in range(P):
        print(f"capture slab theta_{k}: LR {b.grad[k]:+.5e} +- {b.grad_sigma[k]:.2e}  "
              f"FD {fd[k]:+.5e} +- {fd_sig[k]:.2e}  flips {flips[k]}  z {z[k]:+.2f}", flush=True)
    checked = flips >= MIN_FLIPS
    assert checked[[0, 1, 2, 4]].all(), flips            # B-10 capture and every scatter factor
    assert abs(b.grad[1]) > 10 * b.grad_sigma[1]         # B-10 capture is resolved
    assert np.all(np.abs(z[checked]) <= 2.0), z


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
