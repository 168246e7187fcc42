"""Test helpers: the oracle (CPU restatement) through ctypes, and fixtures
shared by the CPU and GPU tests.  The oracle is the checker only."""
from __future__ import annotations

import ctypes as C
import importlib
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libmc_oracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_kernels.so")

mc = importlib.import_module("mc-old_amd")
decks = importlib.import_module("mc-old_amd.decks")


def _oracle():
    o = C.CDLL(ORACLE_LIB)
    o.mco_lcg_next.restype = C.c_uint64
    o.mco_lcg_next.argtypes = [C.c_uint64]
    o.mco_urand.restype = C.c_double
    o.mco_urand.argtypes = [C.POINTER(C.c_uint64)]
    o.mco_skip_ahead.restype = C.c_uint64
    o.mco_skip_ahead.argtypes = [C.c_uint64, C.c_int64]
    o.mco_init_particle.restype = C.c_uint64
    o.mco_init_particle.argtypes = [C.c_uint64, C.c_uint64]
    o.mco_binary_search.restype = C.c_int
    o.mco_binary_search.argtypes = [C.c_double, C.POINTER(C.c_double), C.c_int]
    o.mco_interpolate.restype = C.c_double
    o.mco_interpolate.argtypes = [C.c_double] * 5
    o.mco_geometry_quad.restype = C.c_double
    o.mco_geometry_quad.argtypes = [C.c_double] * 3
    o.mco_scatter_direction.argtypes = [C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_double)]
    o.mco_log.restype = C.c_double
    o.mco_log.argtypes = [C.c_double]
    o.mco_sincos_2pi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    o.mco_surface_eval.restype = C.c_double
    o.mco_surface_eval.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    o.mco_surface_distance.restype = C.c_double
    o.mco_surface_distance.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    o.mco_surface_reflect.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    o.mco_nuclide_xs.restype = C.c_double
    o.mco_nuclide_xs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
    o.mco_material_xs.restype = C.c_double
    o.mco_material_xs.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double]
    o.mco_run.restype = C.c_int
    o.mco_run.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                          C.POINTER(C.c_uint64), C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
    return o


oracle = _oracle()


def ref_lib():
    """The reference's own Random/Algorithm/XSec compiled into oracle/_ref (or None)."""
    if not os.path.exists(REF_LIB):
        return None
    r = C.CDLL(REF_LIB)
    r.ref_rn_init_problem.argtypes = [C.c_uint64]
    r.ref_rn_init_particle.argtypes = [C.c_uint64]
    r.ref_urand.restype = C.c_double
    r.ref_skip_ahead.restype = C.c_uint64
    r.ref_skip_ahead.argtypes = [C.c_uint64, C.c_int64]
    r.ref_binary_search.restype = C.c_int
    r.ref_binary_search.argtypes = [C.c_double, C.POINTER(C.c_double), C.c_int]
    r.ref_interpolate.restype = C.c_double
    r.ref_interpolate.argtypes = [C.c_double] * 5
    r.ref_geometry_quad.restype = C.c_double
    r.ref_geometry_quad.argtypes = [C.c_double] * 3
    r.ref_scatter_direction.argtypes = [C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]
    r.ref_exponential_sample.restype = C.c_double
    r.ref_exponential_sample.argtypes = [C.c_double]
    r.ref_xstable.restype = C.c_double
    r.ref_xstable.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double]
    return r


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleResult:
    def __init__(self, n, nb, npar):
        self.n = n
        self.sum = np.zeros(max(1, nb))
        self.sq = np.zeros(max(1, nb))
        self.gsum = np.zeros(max(1, npar))
        self.gsq = np.zeros(max(1, npar))
        self.flights = 0
        self.trace = None

    @property
    def mean(self):
        return self.sum / self.n

    @property
    def uncer(self):
        m = self.mean
        return np.sqrt((self.sq / self.n - m * m) / (self.n - 1.0))

    @property
    def grad(self):
        return self.gsum / self.n

    @property
    def grad_sigma(self):
        m = self.grad
        return np.sqrt((self.gsq / self.n - m * m) / (self.n - 1.0))


def oracle_run(problem, n: int, key_offset: int = 0, theta=None, trace: int = 0, grad_out=None) -> OracleResult:
    """Runs n histories of `problem` on the CPU restatement."""
    r = OracleResult(n, problem.n_bins, problem.n_params)
    fl = C.c_uint64(0)
    tr = None
    if trace:
        tr = np.zeros(trace, dtype=mc.TRACE_DTYPE)
    th = None if theta is None else np.ascontiguousarray(theta, dtype=np.float64)
    go = None if grad_out is None else np.ascontiguousarray(grad_out, dtype=np.float64)
    st = oracle.mco_run(problem.desc_ptr, None if th is None else th.ctypes.data, key_offset, n,
                        r.sum.ctypes.data, r.sq.ctypes.data, C.byref(fl),
                        None if tr is None else tr.ctypes.data, trace,
                        None if go is None else go.ctypes.data,
                        r.gsum.ctypes.data if go is not None else None,
                        r.gsq.ctypes.data if go is not None else None)
    if st != 0:
        raise RuntimeError(f"oracle failed with status {st}")
    r.flights = fl.value
    r.trace = tr
    r.sum = r.sum[:problem.n_bins]
    r.sq = r.sq[:problem.n_bins]
    r.gsum = r.gsum[:problem.n_params]
    r.gsq = r.gsq[:problem.n_params]
    return r


REACTION_SCORES = ("flux", "absorption", "scatter", "capture", "total")


def reaction_slab(L, c, n):
    """Single-cell slab with a track-length and a collision estimator of every
    reaction score on the cell (bins 2-6 and 7-11 after T and R)."""
    sc = " ".join(REACTION_SCORES)
    extra = (f'    <estimator name="tl" type="TL" scores="{sc}">\n        <cell name="bin0"/>\n    </estimator>\n'
             f'    <estimator name="col" type="C" scores="{sc}">\n        <cell name="bin0"/>\n    </estimator>\n')
    deck = decks.slab_deck(L, c, 0, n, params=("scatter", "capture"))
    return deck.replace("</estimators>", extra + "</estimators>")


_SHIELD_XS = os.path.join(ROOT, "tests", "_xs_cache")


def shield_xs_dir() -> str:
    """Directory holding the synthetic shield nuclides (written on first use)."""
    marker = os.path.join(_SHIELD_XS, "926056.txt")
    if not os.path.exists(marker):
        decks.write_shield_nuclides(_SHIELD_XS)
    return _SHIELD_XS


def importance_slab(L: float, c: float, nb: int, n: int, ratio: float) -> str:
    """One-speed slab whose flux bin k has importance ratio**k: every crossing
    deeper splits (ratio > 1), every crossing back plays roulette."""
    deck = decks.slab_deck(L, c, nb, n)
    for k in range(1, nb):
        deck = deck.replace(f'<cell name="bin{k}" material="slab">',
                            f'<cell name="bin{k}" material="slab" importance="{ratio ** k!r}">')
    return deck


def free_gas_slab(n: int, L: float = 5.0, E0: float = 1.0, kinematics: str = "free_gas") -> str:
    """A hydrogen-carbon slab (constant cross sections, A = 1 and 12) hit by a
    normally incident E0 eV beam: collisions thermalize the particles, so
    free-gas kinematics decides their energies.  Transmission, reflection
    and the reflected energy spectrum."""
    return f"""<?xml version='1.0' encoding='UTF-8'?>
<simulation>
    <description name="free-gas slab" samples="{n}" seed="1"/>
</simulation>
<nuclides>
    <nuclide name="h" A="1.0" kinematics="{kinematics}">
        <scatter xs="20.0"/>
        <capture xs="0.3"/>
    </nuclide>
    <nuclide name="c" A="12.0" kinematics="{kinematics}">
        <scatter xs="4.7"/>
        <capture xs="0.0035"/>
    </nuclide>
</nuclides>
<materials>
    <material name="hc">
        <nuclide name="h" density="0.06"/>
        <nuclide name="c" density="0.03"/>
    </material>
</materials>
<surfaces>
    <plane_x name="px0" x="0.0"/>
    <plane_x name="px1" x="{L!r}"/>
</surfaces>
<cells>
    <cell name="slab" material="hc">
        <surface name="px0" sense="+1"/>
        <surface name="px1" sense="-1"/>
    </cell>
    <cell name="left outside" importance="0.0">
        <surface name="px0" sense="-1"/>
    </cell>
    <cell name="right outside" importance="0.0">
        <surface name="px1" sense="+1"/>
    </cell>
</cells>
<distributions>
    <delta name="beam" datatype="point" x="1.0" y="0.0" z="0.0"/>
    <delta name="e0" datatype="double" val="{E0!r}"/>
</distributions>
<sources>
    <point x="1e-9" y="0.0" z="0.0" direction="beam" energy="e0"/>
</sources>
<estimators>
    <estimator name="transmission" scores="cross">
        <surface name="px1"/>
    </estimator>
    <estimator name="reflection" scores="cross">
        <surface name="px0"/>
        <filter type="energy" grid_lethargy="1e-5 10.0 40"/>
    </estimator>
</estimators>
"""


REFERENCE_DECK = os.path.join(ROOT, "tests", "golden", "shielding_vReduction")


def reference_shielding_deck():
    """The reference's own examples/shielding_vReduction deck with its
    xs_library data (O-16, H-1, He-3, C, B-10, B-11: repeated energies, one
    file whose energies decrease twice, 5- and 6-column rows), kept as a
    fixture: water, a B4C shield and a He-3 cylinder_z detector, 2 MeV point
    source, free-gas scattering (the reference's default)."""
    return mc.Problem.from_xml(REFERENCE_DECK, os.path.join(REFERENCE_DECK, "xs_library"))


def bc_box_deck(n: int) -> str:
    """Boundary conditions: a box whose x = 0 face is reflective (plane_x),
    whose top is a reflective generic plane 0.2x + 0.1y + z = 3, whose x = 6
    face is vacuum (kills before scoring: its tally stays 0) and whose y and
    z = -3 faces lead to importance-0 cells.  Target-at-rest scattering off
    A = 12, isotropic 1 keV point source; surface, track-length and collision
    estimators."""
    return f"""<?xml version='1.0' encoding='UTF-8'?>
<simulation>
    <description name="boundary conditions" samples="{n}" seed="3"/>
</simulation>
<nuclides>
    <nuclide name="m" A="12.0" kinematics="target_at_rest">
        <scatter xs="2.0"/>
        <capture xs="0.1"/>
    </nuclide>
</nuclides>
<materials>
    <material name="mat">
        <nuclide name="m" density="0.3"/>
    </material>
</materials>
<surfaces>
    <plane_x name="px0" x="0.0" bc="reflective"/>
    <plane_x name="px1" x="6.0" bc="vacuum"/>
    <plane_y name="py0" y="-3.0"/>
    <plane_y name="py1" y="3.0"/>
    <plane_z name="pz0" z="-3.0"/>
    <plane name="top" a="0.2" b="0.1" c="1.0" d="3.0" bc="reflective"/>
</surfaces>
<cells>
    <cell name="in" material="mat">
        <surface name="px0" sense="+1"/>
        <surface name="px1" sense="-1"/>
        <surface name="py0" sense="+1"/>
        <surface name="py1" sense="-1"/>
        <surface name="pz0" sense="+1"/>
        <surface name="top" sense="-1"/>
    </cell>
    <cell name="ylo" importance="0.0">
        <surface name="py0" sense="-1"/>
    </cell>
    <cell name="yhi" importance="0.0">
        <surface name="py1" sense="+1"/>
    </cell>
    <cell name="zlo" importance="0.0">
        <surface name="pz0" sense="-1"/>
    </cell>
</cells>
<distributions>
    <delta name="e" datatype="double" val="1000.0"/>
</distributions>
<sources>
    <point x="1.0" y="0.0" z="0.0" energy="e"/>
</sources>
<estimators>
    <estimator name="vacuum face" scores="cross">
        <surface name="px1"/>
    </estimator>
    <estimator name="reflective faces" scores="cross">
        <surface name="px0"/>
        <surface name="top"/>
    </estimator>
    <estimator name="leak" scores="cross">
        <surface name="py0"/>
        <surface name="py1"/>
        <surface name="pz0"/>
    </estimator>
    <estimator name="tl" type="TL" scores="flux absorption total">
        <cell name="in"/>
    </estimator>
    <estimator name="col" type="C" scores="flux scatter">
        <cell name="in"/>
    </estimator>
</estimators>
"""


def quadric_deck(n: int) -> str:
    """Quadrics: a cylinder_x (r = 4, 0 <= x <= 20) crossed by an absorbing
    cylinder_z rod (r = 1.5 at x = 10), with a void sphere (r = 1 at x = 4);
    free-gas hydrogen-like scattering, isotropic 100 eV source."""
    return f"""<?xml version='1.0' encoding='UTF-8'?>
<simulation>
    <description name="quadrics" samples="{n}" seed="5"/>
</simulation>
<nuclides>
    <nuclide name="h" A="1.0">
        <scatter xs="10.0"/>
        <capture xs="0.2"/>
    </nuclide>
    <nuclide name="b" A="10.0" kinematics="target_at_rest">
        <scatter xs="2.0"/>
        <capture xs="30.0"/>
    </nuclide>
</nuclides>
<materials>
    <material name="mod">
        <nuclide name="h" density="0.05"/>
    </material>
    <material name="abs">
        <nuclide name="b" density="0.02"/>
        <nuclide name="h" density="0.01"/>
    </material>
</materials>
<surfaces>
    <cylinder_x name="cx" y="0.0" z="0.0" r="4.0"/>
    <plane_x name="p0" x="0.0"/>
    <plane_x name="p1" x="20.0"/>
    <cylinder_z name="cz" x="10.0" y="0.0" r="1.5"/>
    <sphere name="sp" x="4.0" y="0.0" z="0.0" r="1.0"/>
</surfaces>
<cells>
    <cell name="rod" material="abs">
        <surface name="cz" sense="-1"/>
        <surface name="cx" sense="-1"/>
    </cell>
    <cell name="ball">
        <surface name="sp" sense="-1"/>
    </cell>
    <cell name="body" material="mod">
        <surface name="cx" sense="-1"/>
        <surface name="p0" sense="+1"/>
        <surface name="p1" sense="-1"/>
        <surface name="cz" sense="+1"/>
        <surface name="sp" sense="+1"/>
    </cell>
    <cell name="side" importance="0.0">
        <surface name="cx" sense="+1"/>
    </cell>
    <cell name="back" importance="0.0">
        <surface name="p0" sense="-1"/>
    </cell>
    <cell name="front" importance="0.0">
        <surface name="p1" sense="+1"/>
    </cell>
</cells>
<distributions>
    <delta name="e" datatype="double" val="100.0"/>
</distributions>
<sources>
    <point x="2.0" y="0.0" z="0.0" energy="e"/>
</sources>
<estimators>
    <estimator name="transmission" scores="cross">
        <surface name="p1"/>
    </estimator>
    <estimator name="side leak" scores="cross">
        <surface name="cx"/>
        <filter type="energy" grid_lethargy="1e-3 1e3 12"/>
    </estimator>
    <estimator name="flux" scores="flux">
        <cell name="rod"/>
        <cell name="ball"/>
        <cell name="body"/>
    </estimator>
    <estimator name="rod absorption" type="C" scores="absorption">
        <cell name="rod"/>
    </estimator>
</estimators>
"""


class HipRuntime:
    """The HIP runtime libmcold.so itself links (/opt/rocm), through ctypes:
    device buffers and streams for the device-pointer entry points.  (torch
    bundles a HIP runtime of its own; two runtimes in one process do not
    share the device, so tests take device memory from the library's.)"""

    H2D, D2H = 1, 2

    def __init__(self):
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line and "torch" not in line:
                    path = line.split()[-1]
                    break
        self.lib = C.CDLL(path or "/opt/rocm/lib/libamdhip64.so")
        L = self.lib
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]

    def _ok(self, st, what):
        if st != 0:
            raise RuntimeError(f"{what} failed with HIP error {st}")

    def malloc(self, nbytes):
        p = C.c_void_p()
        self._ok(self.lib.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p.value

    def free(self, p):
        self._ok(self.lib.hipFree(C.c_void_p(p)), "hipFree")

    def upload(self, arr):
        p = self.malloc(arr.nbytes)
        self._ok(self.lib.hipMemcpy(C.c_void_p(p), arr.ctypes.data, arr.nbytes, self.H2D), "hipMemcpy")
        return p

    def download(self, p, arr):
        self._ok(self.lib.hipMemcpy(arr.ctypes.data, C.c_void_p(p), arr.nbytes, self.D2H), "hipMemcpy")
        return arr

    def fill(self, p, byte, nbytes):
        self._ok(self.lib.hipMemset(C.c_void_p(p), byte, nbytes), "hipMemset")

    def stream(self):
        s = C.c_void_p()
        self._ok(self.lib.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s.value

    def stream_destroy(self, s):
        self._ok(self.lib.hipStreamDestroy(C.c_void_p(s)), "hipStreamDestroy")


def with_roulette(problem, wr: float, ws: float = 1.0):
    """The same problem built by mc_problem_create with weight-roulette
    threshold wr and survival weight ws (simulator.h:90-91 fixes 0.001 and
    1; wr = 0 switches roulette off)."""
    d = mc.ProblemDesc.from_buffer_copy(problem.head)
    d.wr, d.ws = wr, ws
    h = C.c_void_p()
    st = mc.lib.mc_problem_create(C.byref(d), C.byref(h))
    if st != 0:
        raise mc.MCError(st, mc.lib.mc_last_error().decode())
    return mc.Problem(h.value)
