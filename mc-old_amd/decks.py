"""Deck generators for the benchmark problems, in the reference's XML schema
and xs_library file format (xs_library/README.txt:1-7).

  slab_deck        one-speed slab 0 <= x <= L, Sigma_t = 1/cm, scattering ratio c,
                   normally incident pencil beam at x = 0+, transmission and
                   reflection tallies (+ n_bins track-length flux cells)
  write_shield_nuclides / shield_deck
                   continuous-energy three-layer shield (water-like 20 cm /
                   iron-like 10 cm / concrete-like 30 cm) over four synthetic
                   non-fissile nuclides (A = 1, 12, 16, 56) on a shared
                   16384-point log grid 1e-5 eV - 20 MeV; isotropic 2 MeV point
                   source; 100-bin log-energy leakage tally per face
  batched_slabs    4096 one-speed slabs: L log-uniform 1-30 mfp, c uniform
                   0.1-0.99, histories log-uniform 1e3-1e6 (seeded)
  box_deck         3D plane-bounded box shield of 20 cells with a spherical
                   detector cell 1 m beyond it
"""
from __future__ import annotations

import os

import numpy as np

# ------------------------------------------------------------------ one-speed


def _f(x: float) -> str:
    return repr(float(x))


def slab_deck(L: float = 4.0, c: float = 0.5, n_bins: int = 100, samples: float = 1e6,
              seed: int = 1, x0: float = 1e-9, params=()) -> str:
    """One-speed slab deck; with n_bins = 0 the slab is a single cell.
    params: reactions ("scatter", "capture", "total") of gradient parameters
    declared on the first cell."""
    nb = max(1, n_bins)
    edges = [k * (L / nb) for k in range(nb + 1)]
    edges[-1] = L
    surfaces = "\n".join(f'    <plane_x name="px{k}" x="{_f(x)}"/>' for k, x in enumerate(edges))
    cells = []
    for k in range(nb):
        cells.append(f'    <cell name="bin{k}" material="slab">\n'
                     f'        <surface name="px{k}" sense="+1"/>\n'
                     f'        <surface name="px{k + 1}" sense="-1"/>\n    </cell>')
    cells.append('    <cell name="left outside" importance="0.0">\n'
                 '        <surface name="px0" sense="-1"/>\n    </cell>')
    cells.append(f'    <cell name="right outside" importance="0.0">\n'
                 f'        <surface name="px{nb}" sense="+1"/>\n    </cell>')
    flux = ""
    if n_bins > 0:
        flux_cells = "\n".join(f'        <cell name="bin{k}"/>' for k in range(nb))
        flux = f'    <estimator name="flux" scores="flux">\n{flux_cells}\n    </estimator>\n'
    return f"""<?xml version='1.0' encoding='UTF-8'?>
<simulation>
    <description name="one-speed slab L={_f(L)} c={_f(c)}" samples="{int(samples)}" seed="{seed}"/>
</simulation>
<nuclides>
    <nuclide name="onespeed">
        <scatter xs="{_f(c)}"/>
        <capture xs="{_f(1.0 - c)}"/>
    </nuclide>
</nuclides>
<materials>
    <material name="slab">
        <nuclide name="onespeed" density="1.0"/>
    </material>
</materials>
<surfaces>
{surfaces}
</surfaces>
<cells>
{chr(10).join(cells)}
</cells>
<distributions>
    <delta name="beam" datatype="point" x="1.0" y="0.0" z="0.0"/>
</distributions>
<sources>
    <point x="{_f(x0)}" y="0.0" z="0.0" direction="beam"/>
</sources>
<estimators>
    <estimator name="transmission" scores="cross">
        <surface name="px{nb}"/>
    </estimator>
    <estimator name="reflection" scores="cross">
        <surface name="px0"/>
    </estimator>
{flux}</estimators>
""" + ("<parameters>\n" + "".join(f'    <parameter cell="bin0" nuclide="onespeed" reaction="{r}"/>\n'
                                   for r in params) + "</parameters>\n" if params else "")


def batched_slabs(n: int = 4096, seed: int = 2024):
    """(L, c, histories) for the batched sweep, from a seeded generator."""
    rng = np.random.default_rng(seed)
    L = np.exp(rng.uniform(np.log(1.0), np.log(30.0), n))
    c = rng.uniform(0.1, 0.99, n)
    N = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n)).astype(np.int64)
    return L, c, N


# --------------------------------------------------------- continuous energy
# (name, A, constant scatter [b], 1/v scatter at 0.0253 eV [b], 1/v capture at 0.0253 eV [b])
SHIELD_NUCLIDES = [
    ("H1x", 1.0, 3.0, 17.0, 0.332),
    ("C12x", 12.0, 1.7, 3.0, 0.0035),
    ("O16x", 16.0, 1.2, 2.6, 0.00019),
    ("Fe56x", 56.0, 3.0, 8.0, 2.56),
]
# atom densities (atoms / barn cm); every layer holds all four nuclides
SHIELD_MATERIALS = {
    "water": {"H1x": 0.0668, "C12x": 1.0e-5, "O16x": 0.0334, "Fe56x": 1.0e-6},
    "iron": {"H1x": 1.0e-5, "C12x": 0.0010, "O16x": 1.0e-5, "Fe56x": 0.0848},
    "concrete": {"H1x": 0.0080, "C12x": 0.0050, "O16x": 0.0450, "Fe56x": 0.0005},
}
SHIELD_LAYERS = [("water", 0.0, 20.0), ("iron", 20.0, 30.0), ("concrete", 30.0, 60.0)]
ZAIDS = {"H1x": "901001", "C12x": "906012", "O16x": "908016", "Fe56x": "926056"}
GRID_POINTS = 16384
E_MIN, E_MAX = 1e-5, 2e7


def shield_grid() -> np.ndarray:
    return np.exp(np.linspace(np.log(E_MIN), np.log(E_MAX), GRID_POINTS))


def shield_xs(k: int, seed: int = 7):
    """(A, E, sigma_s, sigma_c) of synthetic nuclide k.

    sigma_s = potential + 1/v term + 40 seeded single-level Breit-Wigner
    resonances (E_r log-uniform 1 eV - 100 keV, width 0.5-3 % of E_r, peak
    50-5000 b); sigma_c = 1/v.
    """
    name, A, pot, s1v, c1v = SHIELD_NUCLIDES[k]
    E = shield_grid()
    rng = np.random.default_rng(seed * 1000 + k)
    Er = np.exp(rng.uniform(np.log(1.0), np.log(1e5), 40))
    G = Er * rng.uniform(0.005, 0.03, 40)
    pk = np.exp(rng.uniform(np.log(50.0), np.log(5000.0), 40))
    th = np.sqrt(0.0253 / E)
    bw = np.zeros_like(E)
    for r, g, p in zip(Er, G, pk):
        bw += p * (0.5 * g) ** 2 / ((E - r) ** 2 + (0.5 * g) ** 2)
    s = pot + s1v * th + bw
    c = c1v * th
    return A, E, s, c


def write_shield_nuclides(xs_dir: str, seed: int = 7) -> None:
    """Writes the four nuclides as <ZAID>.txt in the xs_library format."""
    os.makedirs(xs_dir, exist_ok=True)
    for k, (name, *_rest) in enumerate(SHIELD_NUCLIDES):
        A, E, s, c = shield_xs(k, seed)
        path = os.path.join(xs_dir, ZAIDS[name] + ".txt")
        rows = np.column_stack([E, s, c, np.zeros_like(E), np.zeros_like(E), np.zeros_like(E)])
        with open(path, "w") as f:
            f.write(f"{A!r}\n0.0 0.0\n0.0 0.0\n0.0 0.0\n")
            np.savetxt(f, rows, fmt="%.17g")


def _shield_common(samples: float, seed: int) -> tuple[str, str]:
    # elastic scattering with the target at rest (the workloads' kinematics)
    nucs = "\n".join(f'    <nuclide name="{n}" ZAID="{ZAIDS[n]}" kinematics="target_at_rest"/>'
                     for n, *_ in SHIELD_NUCLIDES)
    mats = []
    for m, comp in SHIELD_MATERIALS.items():
        inner = "\n".join(f'        <nuclide name="{n}" density="{_f(d)}"/>' for n, d in comp.items())
        mats.append(f'    <material name="{m}">\n{inner}\n    </material>')
    head = f"""<?xml version='1.0' encoding='UTF-8'?>
<simulation>
    <description name="three-layer shield" samples="{int(samples)}" seed="{seed}"/>
</simulation>
<nuclides>
{nucs}
</nuclides>
<materials>
{chr(10).join(mats)}
</materials>
"""
    return head, ""


def shield_deck(samples: float = 1e8, seed: int = 1, params: str = "", leakage: bool = True,
                transmission: bool = True) -> str:
    """Three-layer CE shield.  params: "" or "24" (3 layers x 4 nuclides x
    {scatter, capture} scale factors).  leakage=False keeps only the
    transmission estimator, transmission=False only the leakage one."""
    head, _ = _shield_common(samples, seed)
    planes = [0.0, 20.0, 30.0, 60.0]
    surfs = "\n".join(f'    <plane_x name="px{k}" x="{_f(x)}"/>' for k, x in enumerate(planes))
    cells = []
    for k, (m, a, b) in enumerate(SHIELD_LAYERS):
        cells.append(f'    <cell name="{m}" material="{m}">\n'
                     f'        <surface name="px{k}" sense="+1"/>\n'
                     f'        <surface name="px{k + 1}" sense="-1"/>\n    </cell>')
    cells.append('    <cell name="left outside" importance="0.0">\n        <surface name="px0" sense="-1"/>\n    </cell>')
    cells.append('    <cell name="right outside" importance="0.0">\n        <surface name="px3" sense="+1"/>\n    </cell>')
    prm = ""
    if params:
        lines = []
        for m, _, _ in SHIELD_LAYERS:
            for n, *_r in SHIELD_NUCLIDES:
                for r in ("scatter", "capture"):
                    lines.append(f'    <parameter cell="{m}" nuclide="{n}" reaction="{r}"/>')
        prm = "<parameters>\n" + "\n".join(lines) + "\n</parameters>\n"
    return head + f"""<surfaces>
{surfs}
</surfaces>
<cells>
{chr(10).join(cells)}
</cells>
<sources>
    <point x="1e-9" y="0.0" z="0.0"/>
</sources>
<estimators>
{TRANSMISSION if transmission else ""}{LEAKAGE if leakage else ""}</estimators>
{prm}"""


TRANSMISSION = """    <estimator name="transmission" scores="cross">
        <surface name="px3"/>
    </estimator>
"""

LEAKAGE = """    <estimator name="leakage" scores="cross">
        <surface name="px0"/>
        <surface name="px3"/>
        <filter type="energy" grid_lethargy="1e-5 2e7 100"/>
    </estimator>
"""


# ------------------------------------------------------------------ 3D box
BOX_NX, BOX_NY, BOX_NZ = 5, 2, 2           # 20 shield cells
BOX_X, BOX_HALF = 60.0, 50.0                # shield 0 <= x <= 60, |y|,|z| <= 50
WORLD_X = 200.0
DETECTOR = (160.0, 0.0, 0.0, 10.0)          # default: sphere centre x, y, z and radius
BOX_MATERIAL_OF_COLUMN = ["water", "water", "iron", "concrete", "concrete"]


def box_deck(samples: float = 1.5625e7, seed: int = 1, params: bool = True, box_x: float = BOX_X,
             detector=None) -> str:
    """3D box shield: 20 cells (5 x 2 x 2) of 4 materials filling
    0 <= x <= box_x (60 cm), |y|, |z| <= 50, then void to x = 200 holding a
    spherical detector cell (r = 10 cm) centred 1 m beyond the back face;
    track-length flux in the detector is the response.  Everything outside
    the world box has importance 0.  params: one total-XS scale factor per
    (cell, nuclide) -> 80 parameters.  detector: (x, y, z, r) override."""
    head, _ = _shield_common(samples, seed)
    # the fourth material is a lighter concrete variant
    head = head.replace("</materials>", '    <material name="light">\n'
                        '        <nuclide name="H1x" density="0.004"/>\n'
                        '        <nuclide name="C12x" density="0.002"/>\n'
                        '        <nuclide name="O16x" density="0.030"/>\n'
                        '        <nuclide name="Fe56x" density="0.0002"/>\n    </material>\n</materials>')
    xs = np.linspace(0.0, box_x, BOX_NX + 1)
    ys = np.linspace(-BOX_HALF, BOX_HALF, BOX_NY + 1)
    zs = np.linspace(-BOX_HALF, BOX_HALF, BOX_NZ + 1)
    s = []
    for i, x in enumerate(xs):
        s.append(f'    <plane_x name="x{i}" x="{_f(x)}"/>')
    s.append(f'    <plane_x name="xw" x="{_f(WORLD_X)}"/>')
    for j, y in enumerate(ys):
        s.append(f'    <plane_y name="y{j}" y="{_f(y)}"/>')
    for k, z in enumerate(zs):
        s.append(f'    <plane_z name="z{k}" z="{_f(z)}"/>')
    cx, cy, cz, r = detector if detector is not None else (box_x + 100.0, 0.0, 0.0, 10.0)
    s.append(f'    <sphere name="det" x="{_f(cx)}" y="{_f(cy)}" z="{_f(cz)}" r="{_f(r)}"/>')
    mats = BOX_MATERIAL_OF_COLUMN
    cells = []
    names = []
    for i in range(BOX_NX):
        for j in range(BOX_NY):
            for k in range(BOX_NZ):
                m = mats[i] if not (i == 3 and j == 1) else "light"
                nm = f"c{i}{j}{k}"
                names.append((nm, m))
                cells.append(f'    <cell name="{nm}" material="{m}">\n'
                             f'        <surface name="x{i}" sense="+1"/><surface name="x{i + 1}" sense="-1"/>\n'
                             f'        <surface name="y{j}" sense="+1"/><surface name="y{j + 1}" sense="-1"/>\n'
                             f'        <surface name="z{k}" sense="+1"/><surface name="z{k + 1}" sense="-1"/>\n'
                             f'    </cell>')
    cells.append(f'    <cell name="void">\n'
                 f'        <surface name="x{BOX_NX}" sense="+1"/><surface name="xw" sense="-1"/>\n'
                 f'        <surface name="y0" sense="+1"/><surface name="y{BOX_NY}" sense="-1"/>\n'
                 f'        <surface name="z0" sense="+1"/><surface name="z{BOX_NZ}" sense="-1"/>\n'
                 f'        <surface name="det" sense="+1"/>\n    </cell>')
    cells.append('    <cell name="detector">\n        <surface name="det" sense="-1"/>\n    </cell>')
    outside = [("x0", -1), ("xw", +1), ("y0", -1), (f"y{BOX_NY}", +1), ("z0", -1), (f"z{BOX_NZ}", +1)]
    for q, (sn, sense) in enumerate(outside):
        cells.append(f'    <cell name="out{q}" importance="0.0">\n'
                     f'        <surface name="{sn}" sense="{sense:+d}"/>\n    </cell>')
    prm = ""
    if params:
        lines = []
        for nm, m in names:
            for n, *_r in SHIELD_NUCLIDES:
                lines.append(f'    <parameter cell="{nm}" nuclide="{n}" reaction="total"/>')
        prm = "<parameters>\n" + "\n".join(lines) + "\n</parameters>\n"
    return head + f"""<surfaces>
{chr(10).join(s)}
</surfaces>
<cells>
{chr(10).join(cells)}
</cells>
<sources>
    <point x="1e-9" y="0.0" z="0.0"/>
</sources>
<estimators>
    <estimator name="detector" scores="flux">
        <cell name="detector"/>
    </estimator>
</estimators>
{prm}"""


def box_variant_thetas(n_variants: int = 64, seed: int = 99) -> np.ndarray:
    """Seeded per-(cell, nuclide) scale factors, uniform in [0.8, 1.2]."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.8, 1.2, (n_variants, 80))
