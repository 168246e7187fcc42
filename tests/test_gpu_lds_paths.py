"""The fused flight kernel's geometry and material paths against the
oracle, bit-exact per history: a geometry of at most 64 cells and 64
surfaces is read from an LDS copy; a larger one from per-cell flight records
in forward runs when no cell has more than 8 surfaces, else from the cell,
cell-surface and surface arrays; the first 32 materials' lookup and
collision records are read from LDS, later materials from the arrays.
Continuous-energy slabs of many one-material layers cross these limits."""
import numpy as np
import pytest

from mc_testlib import decks, mc, oracle_run, shield_xs_dir

pytestmark = pytest.mark.gpu

FIELDS = ("rng", "flights", "x", "y", "z", "E", "score")
KINDS = list(decks.SHIELD_MATERIALS)


def layered_deck(n_layers: int, samples: int, params: bool = False, thick: float = 0.75, far: int = 0) -> str:
    """n_layers slabs, layer k of its own material m{k} (the composition of
    water, iron and concrete in turn); a transmission tally on the far face;
    params: a scatter and a capture factor per layer and nuclide; far: planes
    well beyond the slabs that also bound the importance-0 cell on the right
    (a cell of far + 1 surfaces)."""
    nucs = "\n".join(f'    <nuclide name="{n}" ZAID="{decks.ZAIDS[n]}" kinematics="target_at_rest"/>'
                     for n, *_ in decks.SHIELD_NUCLIDES)
    mats = []
    for k in range(n_layers):
        comp = decks.SHIELD_MATERIALS[KINDS[k % len(KINDS)]]
        inner = "\n".join(f'        <nuclide name="{n}" density="{decks._f(d)}"/>' for n, d in comp.items())
        mats.append(f'    <material name="m{k}">\n{inner}\n    </material>')
    surfs = "\n".join([f'    <plane_x name="px{k}" x="{decks._f(k * thick)}"/>' for k in range(n_layers + 1)] +
                      [f'    <plane_x name="far{k}" x="{decks._f(1000.0 + k)}"/>' for k in range(far)])
    cells = [f'    <cell name="c{k}" material="m{k}">\n        <surface name="px{k}" sense="+1"/>\n'
             f'        <surface name="px{k + 1}" sense="-1"/>\n    </cell>' for k in range(n_layers)]
    cells.append('    <cell name="left" importance="0.0">\n        <surface name="px0" sense="-1"/>\n    </cell>')
    far_sides = "".join(f'        <surface name="far{k}" sense="-1"/>\n' for k in range(far))
    cells.append(f'    <cell name="right" importance="0.0">\n        <surface name="px{n_layers}" sense="+1"/>\n'
                 f'{far_sides}    </cell>')
    prm = ""
    if params:
        lines = [f'    <parameter cell="c{k}" nuclide="{n}" reaction="{r}"/>'
                 for k in range(n_layers) for n, *_ in decks.SHIELD_NUCLIDES for r in ("scatter", "capture")]
        prm = "<parameters>\n" + "\n".join(lines) + "\n</parameters>\n"
    return f"""<?xml version='1.0' encoding='UTF-8'?>
<simulation>
    <description name="layered shield" samples="{samples}" seed="3"/>
</simulation>
<nuclides>
{nucs}
</nuclides>
<materials>
{chr(10).join(mats)}
</materials>
<surfaces>
{surfs}
</surfaces>
<cells>
{chr(10).join(cells)}
</cells>
<sources>
    <point x="1e-9" y="0.0" z="0.0"/>
</sources>
<estimators>
    <estimator name="transmission" scores="cross">
        <surface name="px{n_layers}"/>
    </estimator>
</estimators>
{prm}"""


# 40 layers: 42 cells, 41 surfaces (LDS geometry), 40 materials (8 past the
# LDS records); 70 layers: 72 cells, 71 surfaces (flight records); 70 layers
# and a 9-surface cell (the arrays)
@pytest.mark.parametrize("n_layers,far", [(40, 0), (70, 0), (70, 8)])
def test_layered_histories_bitexact(n_layers, far):
    n = 3000
    p = mc.Problem.from_string(layered_deck(n_layers, n, far=far), shield_xs_dir())
    g = mc.Engine(p, bank_capacity=1 << 12, drain_divisor=-1).forward(trace=n)
    o = oracle_run(p, n, trace=n)
    for f in FIELDS:
        bad = np.nonzero(g.trace[f] != o.trace[f])[0]
        assert bad.size == 0, f"{n_layers} layers: field {f} differs at histories {bad[:8].tolist()}"
    assert g.flights == o.flights
    # histories end in or beyond the layers past the 32nd material
    assert (o.trace["x"] > 32 * 0.75).sum() > 50
    np.testing.assert_allclose(g.sum, o.sum, rtol=1e-11, atol=1e-300)


@pytest.mark.parametrize("n_layers", [40, 70])
def test_layered_gradient(n_layers):
    """Scaled materials past the LDS records in the fused gradient kernel (8
    parameters per layer; each layer's material is cloned for its cell, and
    the clones follow the deck's own materials, so every layer's clone lies
    past the 32nd): forward moment sums equal to the oracle's
    to 1e-11, the gradient to the oracle's within 1e-8."""
    n = 3000
    p = mc.Problem.from_string(layered_deck(n_layers, n, params=True), shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12)
    assert eng.n_params == 8 * n_layers
    fwd = eng.forward()
    bwd = eng.backward(np.ones(eng.n_bins))
    o = oracle_run(p, n, grad_out=np.ones(eng.n_bins))
    np.testing.assert_allclose(fwd.sum, o.sum, rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(bwd.sum, fwd.sum, rtol=1e-12)
    np.testing.assert_allclose(bwd.grad, o.grad, rtol=1e-8, atol=1e-12)
    assert np.count_nonzero(o.grad[8 * 33:]) > 0       # layers past the 32nd contribute
