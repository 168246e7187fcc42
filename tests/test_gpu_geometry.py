"""GPU histories against the oracle, bit for bit, on the reference's own deck
and data and on decks that reach every boundary condition and surface type:

* examples/shielding_vReduction with the reference's xs_library files
  (repeated energies, a grid whose energies decrease, 5- and 6-column rows,
  materials whose nuclides sit on different grids, cylinder_z, free gas);
* vacuum (general.cpp:98-101: killed before its tally scores, so the
  tally is 0) and reflective (general.cpp:102-106, Geometry.cpp:195-222)
  boundaries on plane_x and a generic plane;
* cylinder_x, cylinder_z and a void sphere;
* the 3D box shield at its real size (60 cm box, detector 1 m beyond).
"""
import numpy as np
import pytest

from mc_testlib import (bc_box_deck, decks, mc, oracle, oracle_run, quadric_deck, reference_shielding_deck,
                        shield_xs_dir)

pytestmark = pytest.mark.gpu

FIELDS = ("rng", "flights", "x", "y", "z", "E", "score")


def run_both(p, n, **kw):
    print(f"deck bins={p.n_bins} n={n} {kw}", flush=True)
    g = mc.Engine(p, **kw).forward(n_histories=n, trace=n)
    o = oracle_run(p, n, trace=n)
    for f in FIELDS:
        bad = np.nonzero(g.trace[f] != o.trace[f])[0]
        assert bad.size == 0, f"{f} differs at histories {bad[:8].tolist()}"
    assert g.flights == o.flights
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(g.sq, o.sq, rtol=1e-11, atol=1e-300)
    return g, o


@pytest.mark.parametrize("drain", [1, -1])
def test_reference_shielding_deck_bitexact(drain):
    p = reference_shielding_deck()
    g, o = run_both(p, 40000, bank_capacity=1 << 13, drain_divisor=drain)
    assert (o.trace["score"] > 0).sum() > 10            # the He-3 detector scored


def test_reference_deck_xs_lookup_bitexact():
    """Sigma_t of each material at energies over the whole tabulated range,
    on tabulated points and around the points where 005011's energies
    decrease: the GPU lookup (mixed-grid path, full search on the
    non-monotone grid) against the oracle's Material::SigmaT."""
    import ctypes as C
    p = reference_shielding_deck()
    eng = mc.Engine(p)
    rng = np.random.default_rng(8)
    E = np.concatenate([np.exp(rng.uniform(np.log(1e-6), np.log(3e7), 30000)),
                        np.linspace(19850.0, 19910.0, 400), np.linspace(2.4e6, 2.42e6, 400),
                        [1e-5, 1.06e-5, 2e6, 2e7]])
    for m in range(3):
        mats = np.full(E.size, m, dtype=np.int32)
        got = eng.xs_lookup(E, mats)
        want = np.array([oracle.mco_material_xs(p.desc_ptr, None, m, 0, float(e)) for e in E])
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (m, E[bad[:5]], got[bad[:5]], want[bad[:5]])


@pytest.mark.parametrize("k", [1, 6])
def test_vacuum_and_reflective_boundaries_bitexact(k):
    n = 20000
    p = mc.Problem.from_string(bc_box_deck(n))
    g, o = run_both(p, n, bank_capacity=1 << 12, events_per_launch=k)
    assert g.sum[0] == 0.0                               # the vacuum face scores nothing
    assert g.sum[1] > 0 and g.sum[2] > 0                 # reflections on plane_x and the generic plane


@pytest.mark.parametrize("drain", [1, -1])
def test_cylinders_and_void_sphere_bitexact(drain):
    n = 20000
    p = mc.Problem.from_string(quadric_deck(n))
    g, _ = run_both(p, n, bank_capacity=1 << 12, drain_divisor=drain)
    assert g.sum[14] > 0                                 # track length through the void sphere
    assert g.sum[16] > 0                                 # absorption in the cylinder_z rod


def box_deck_scored(n):
    """The box shield at its real size with, besides the detector, the
    track-length flux of its 20 cells and the leakage through its back face
    (small runs rarely reach the detector; these bins score every history)."""
    cells = "".join(f'        <cell name="c{i}{j}{k}"/>\n' for i in range(5) for j in range(2) for k in range(2))
    extra = (f'    <estimator name="cell flux" scores="flux">\n{cells}    </estimator>\n'
             '    <estimator name="back face" scores="cross">\n        <surface name="x5"/>\n    </estimator>\n')
    return decks.box_deck(n).replace("</estimators>", extra + "</estimators>")


def test_box_shield_full_geometry_bitexact():
    n = 8000
    p = mc.Problem.from_string(box_deck_scored(n), shield_xs_dir())
    assert p.n_bins == 22
    g, _ = run_both(p, n, bank_capacity=1 << 13)
    assert (g.sum[1:] > 0).all()


def test_box_variants_full_geometry_forward_and_gradient():
    """64 design variants of the real box (60 cm, detector r = 10 cm 1 m
    beyond) in one batched engine, each variant's tallies and the
    80-component gradient of their sum against the oracle on the same keys."""
    nv, n = 64, 300
    thetas = decks.box_variant_thetas(nv)
    probs = [mc.Problem.from_string(box_deck_scored(n), shield_xs_dir()) for _ in range(nv)]
    eng = mc.Engine(probs, bank_capacity=1 << 14)
    theta = thetas.ravel()
    fwd = eng.forward(theta=theta)
    bwd = eng.backward(np.ones(eng.n_bins), theta=theta)
    off = 0
    for v in range(nv):
        o = oracle_run(probs[v], n, key_offset=off, theta=thetas[v], grad_out=np.ones(22))
        sl, ps = eng.bin_range(v), eng.param_range(v)
        np.testing.assert_allclose(fwd.sum[sl], o.sum, rtol=1e-11, atol=1e-300)
        np.testing.assert_allclose(bwd.grad[ps], o.grad, rtol=1e-8, atol=1e-12)
        off += n
    assert np.abs(bwd.grad).max() > 0
