"""Pins the deterministic one-speed slab solver by independent mathematics:
the c = 0 closed form, particle balance, the half-space albedo from
Chandrasekhar's H function, and the reference's own pure-absorber known
answer exp(-4.2) (test_integral_Simulator.cpp:28)."""
import numpy as np
import pytest

import slab_analytic as sa


def test_pure_absorber_closed_form():
    r = sa.slab(4.0, 0.0)
    assert abs(r["T"] - np.exp(-4.0)) < 1e-14
    assert abs(r["R"]) < 1e-14
    edges = np.linspace(0, 4, 101)
    assert np.allclose(r["flux"], np.exp(-edges[:-1]) - np.exp(-edges[1:]), atol=1e-13)


@pytest.mark.parametrize("L,c", [(4.0, 0.5), (1.0, 0.9), (10.0, 0.99), (2.0, 0.1)])
def test_balance(L, c):
    r = sa.slab(L, c, n_bins=20)
    assert abs(r["T"] + r["R"] + (1 - c) * r["flux"].sum() - 1.0) < 1e-7


@pytest.mark.parametrize("c", [0.3, 0.5, 0.9, 0.99])
def test_thick_slab_reflection_is_half_space_albedo(c):
    T, R = sa.slab_leakage(40.0, c)
    assert abs(R - sa.half_space_albedo(c)) < 5e-7


def test_leakage_matches_flux_solver():
    r = sa.slab(4.0, 0.5)
    T, R = sa.slab_leakage(4.0, 0.5)
    assert abs(T - r["T"]) < 1e-7 and abs(R - r["R"]) < 1e-7


def test_h_function_identity():
    # zeroth moment: int_0^1 H(mu) dmu = 2 (1 - sqrt(1-c)) / c
    g, w = np.polynomial.legendre.leggauss(400)
    mu, w = 0.5 * (g + 1), 0.5 * w
    for c in (0.5, 0.9):
        H = sa.h_function(c, mu)
        assert abs(np.sum(w * H) - 2 * (1 - np.sqrt(1 - c)) / c) < 1e-9
