"""Free-gas elastic scattering (Reaction.cpp:27-118) of the oracle, pinned by
closed forms that hold for any correct sampler of that model:

* a particle at rest (E << kT) struck by a gas target of mass ratio A: the
  target speeds it meets are weighted by the relative speed, V^3 exp(-b V^2),
  and it leaves isotropically in the centre of mass, so <E'> = 4 A kT / (A+1)^2
  (kT itself at A = 1);
* E >> kT, the target-at-rest limit: E'/E uniform on [alpha, 1] with
  alpha = ((A-1)/(A+1))^2, <ln(E'/E)> = -xi, xi = 1 + alpha ln(alpha)/(1-alpha),
  and the lab cosine <mu> = 2/(3A);
* |u'| = 1.
kT follows from the reference's own constants: E = 5.2270376e-13 v^2 (eV, cm/s)
and b^2 = 2.0659834e-11 A = A m / (2 kT), so kT = 5.2270376e-13 / 2.0659834e-11.
"""
import ctypes as C

import numpy as np
import pytest

from mc_testlib import decks, free_gas_slab, mc, oracle, oracle_run

KT = 5.2270376e-13 / 2.0659834e-11
oracle.mco_free_gas_scatter.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_uint64)]


def sample(A, E0, n, u0=(0.0, 0.6, 0.8), seed=12345):
    rng = C.c_uint64(oracle.mco_init_particle(seed, 7))
    E = np.empty(n)
    mu = np.empty(n)
    norm = np.empty(n)
    e = C.c_double()
    u = (C.c_double * 3)()
    for i in range(n):
        e.value = E0
        u[0], u[1], u[2] = u0
        oracle.mco_free_gas_scatter(A, C.byref(e), u, C.byref(rng))
        E[i] = e.value
        mu[i] = u[0] * u0[0] + u[1] * u0[1] + u[2] * u0[2]
        norm[i] = u[0] ** 2 + u[1] ** 2 + u[2] ** 2
    return E, mu, norm


def z(x, expect):
    return (x.mean() - expect) / (x.std(ddof=1) / np.sqrt(x.size))


@pytest.mark.parametrize("A", [1.0, 12.0, 56.0])
def test_particle_at_rest_gains_the_gas_mean_energy(A):
    E, _, norm = sample(A, 1e-9, 100000)
    assert abs(z(E, 4 * A * KT / (A + 1) ** 2)) < 4
    np.testing.assert_allclose(norm, 1.0, atol=1e-12)


@pytest.mark.parametrize("A", [1.0, 12.0])
def test_fast_particle_reduces_to_target_at_rest(A):
    E, mu, norm = sample(A, 1e6, 100000)
    r = E / 1e6
    alpha = ((A - 1) / (A + 1)) ** 2
    xi = 1.0 if A == 1.0 else 1 + alpha * np.log(alpha) / (1 - alpha)
    # the gas's thermal motion (V/v ~ 1e-4 at 1 MeV) blurs the edges slightly
    assert r.min() > alpha * (1 - 1e-3) and r.max() < 1 + 1e-3
    assert abs(z(r, (1 + alpha) / 2)) < 4
    assert abs(z(np.log(r), -xi)) < 4
    assert abs(z(mu, 2 / (3 * A))) < 4
    np.testing.assert_allclose(norm, 1.0, atol=1e-12)


def test_deck_kinematics_attribute():
    n = 2000
    fg = mc.Problem.from_string(free_gas_slab(n))
    ta = mc.Problem.from_string(free_gas_slab(n, kinematics="target_at_rest"))
    a, b = oracle_run(fg, n), oracle_run(ta, n)
    # thermal particles up-scatter off the gas: the spectra differ
    assert not np.allclose(a.sum, b.sum)
    with pytest.raises(mc.MCError):
        mc.Problem.from_string(free_gas_slab(n, kinematics="hot_gas"))
