"""Pins the oracle's elastic target-at-rest scatter, its isotropic source
direction and its free-gas scatter.

Target at rest (the T -> 0 limit of ReactionScatter::sample, Reaction.cpp:
27-118), from two-body kinematics with the centre-of-mass cosine uniform:
  * E'/E is uniform on [alpha, 1], alpha = ((A-1)/(A+1))^2;
  * <ln(E'/E)> = -xi, xi = 1 + alpha ln(alpha) / (1 - alpha)  (xi = 1 at A = 1);
  * <mu_lab> = 2 / (3A);
  * every sample obeys mu_lab = ((A+1) sqrt(r) - (A-1) / sqrt(r)) / 2, r = E'/E
    (energy and angle come from the same centre-of-mass cosine);
  * the azimuth about the incoming direction is uniform, |u'| = 1;
  * the reference's own free-gas sampler (compiled into oracle/_ref) at
    2 MeV, where the gas's motion is negligible, obeys the same closed forms.
Isotropic direction (DistributionIsotropicDirection::sample, Distribution.cpp:
78-92): every component has mean 0 and mean square 1/3, the cross moments
vanish, the polar cosine on x and the azimuth in the y-z plane are uniform;
against the reference's own sampler on the same stream to 1e-15.
Free gas: against the reference's own ReactionScatter::sample on the same
stream, to 1e-13 relative over masses and energies.

Each check is shown to fail on a planted mistake: the oracle's source is
compiled with one expression changed, and the same check must reject it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest
from scipy import stats

from mc_testlib import REF_LIB, ROOT, oracle, ref_lib

ORACLE_SRC = os.path.join(ROOT, "oracle", "mc_oracle.c")
V3 = C.c_double * 3


def _bind(lib):
    for f in ("mco_elastic_target_at_rest", "mco_free_gas_scatter"):
        getattr(lib, f).argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_uint64)]
    lib.mco_isotropic_direction.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    lib.mco_init_particle.restype = C.c_uint64
    lib.mco_init_particle.argtypes = [C.c_uint64, C.c_uint64]
    return lib


_bind(oracle)


def mutant(old: str, new: str):
    """The oracle compiled with one occurrence of `old` replaced by `new`."""
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"planted mistake needs exactly one '{old}'"
    d = tempfile.mkdtemp(prefix="mco_mut_")
    c_path, so = os.path.join(d, "mut.c"), os.path.join(d, "mut.so")
    with open(c_path, "w") as f:
        f.write(src.replace(old, new))
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared",
                    "-I" + os.path.join(ROOT, "oracle"), "-o", so, c_path, "-lm"], check=True)
    return _bind(C.CDLL(so))


def z(x, expect):
    return (x.mean() - expect) / (x.std(ddof=1) / np.sqrt(x.size))


def azimuth(u, u0):
    """Azimuth of directions u about u0 in a fixed frame perpendicular to u0."""
    u0 = np.asarray(u0, dtype=float)
    a = np.cross(u0, [1.0, 0.0, 0.0] if abs(u0[0]) < 0.9 else [0.0, 1.0, 0.0])
    a /= np.linalg.norm(a)
    b = np.cross(u0, a)
    return np.arctan2(u @ b, u @ a)


# ------------------------------------------------------------ target at rest
def sample_target_at_rest(lib, A, n, E0=1e6, u0=(0.0, 0.6, 0.8), key=11):
    rng = C.c_uint64(lib.mco_init_particle(1, key))
    r = np.empty(n)
    u = np.empty((n, 3))
    e, d = C.c_double(), V3()
    for i in range(n):
        e.value = E0
        d[0], d[1], d[2] = u0
        lib.mco_elastic_target_at_rest(A, C.byref(e), d, C.byref(rng))
        r[i] = e.value / E0
        u[i] = d[0], d[1], d[2]
    return r, u


def target_at_rest_failures(lib, A, n=20000):
    """The closed-form checks a correct target-at-rest sampler passes."""
    u0 = np.array([0.0, 0.6, 0.8])
    r, u = sample_target_at_rest(lib, A, n, u0=tuple(u0))
    mu = u @ u0
    alpha = ((A - 1.0) / (A + 1.0)) ** 2
    xi = 1.0 if A == 1.0 else 1.0 + alpha * math.log(alpha) / (1.0 - alpha)
    bad = []
    if not (np.isfinite(r).all() and np.isfinite(u).all()):
        return ["non-finite energy or direction"]
    if not (r.min() >= alpha * (1 - 1e-12) and r.max() <= 1 + 1e-12):
        bad.append("E'/E outside [alpha, 1]")
    if stats.kstest((r - alpha) / (1.0 - alpha), "uniform").pvalue < 1e-3:
        bad.append("E'/E not uniform")
    if abs(z(np.log(np.maximum(r, 1e-300)), -xi)) > 4:
        bad.append("<ln E'/E> != -xi")
    if abs(z(mu, 2.0 / (3.0 * A))) > 4:
        bad.append("<mu_lab> != 2/(3A)")
    sr = np.sqrt(np.maximum(r, 1e-300))
    mu_kin = 0.5 * ((A + 1.0) * sr - (A - 1.0) / sr)
    ok = r > 1e-6
    if np.max(np.abs(mu[ok] - mu_kin[ok])) > 1e-9:
        bad.append("mu_lab and E'/E break two-body kinematics")
    if stats.kstest((azimuth(u, u0) + np.pi) / (2 * np.pi), "uniform").pvalue < 1e-3:
        bad.append("azimuth not uniform")
    if np.max(np.abs(np.einsum("ij,ij->i", u, u) - 1.0)) > 1e-12:
        bad.append("|u'| != 1")
    return bad


@pytest.mark.parametrize("A", [1.0, 12.0, 56.0])
def test_target_at_rest_closed_forms(A):
    assert target_at_rest_failures(oracle, A) == []


def test_target_at_rest_infinite_mass_is_one_speed():
    r, u = sample_target_at_rest(oracle, 1.0e300, 20000, E0=2.0)
    assert (r == 1.0).all()
    mu = u @ np.array([0.0, 0.6, 0.8])
    assert stats.kstest((mu + 1.0) / 2.0, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("old,new,A", [
    # (A+1)^2 in the wrong place: the energy can grow
    ("*E = *E * q / ((A + 1.0) * (A + 1.0));", "*E = *E * q / ((A + 1.0) * A);", 12.0),
    # the centre-of-mass cosine used as the lab cosine
    ("mu = (1.0 + A * mu_cm) / sqrt(q);", "mu = mu_cm;", 12.0),
    # sign slip in the kinematics
    ("const double q = A * A + 2.0 * A * mu_cm + 1.0;", "const double q = A * A - 2.0 * A * mu_cm + 1.0;", 12.0),
    # azimuth from half the circle
    ("mco_sincos_2pi(mco_urand(rng), &cos_azi, &sin_azi);", "mco_sincos_2pi(0.5 * mco_urand(rng), &cos_azi, &sin_azi);", 1.0),
])
def test_target_at_rest_checks_reject_planted_mistakes(old, new, A):
    assert target_at_rest_failures(mutant(old, new), A) != []


# --------------------------------------------------------------- isotropic
def sample_isotropic(lib, n, key=5):
    rng = C.c_uint64(lib.mco_init_particle(1, key))
    u = np.empty((n, 3))
    d = V3()
    for i in range(n):
        lib.mco_isotropic_direction(C.byref(rng), d)
        u[i] = d[0], d[1], d[2]
    return u


def isotropic_failures(lib, n=40000):
    u = sample_isotropic(lib, n)
    if not np.isfinite(u).all():
        return ["non-finite direction"]
    bad = []
    for k in range(3):
        if abs(z(u[:, k], 0.0)) > 4:
            bad.append(f"<u{k}> != 0")
        if abs(z(u[:, k] ** 2, 1.0 / 3.0)) > 4:
            bad.append(f"<u{k}^2> != 1/3")
    for a, b in ((0, 1), (0, 2), (1, 2)):
        if abs(z(u[:, a] * u[:, b], 0.0)) > 4:
            bad.append(f"<u{a} u{b}> != 0")
    if stats.kstest((u[:, 0] + 1.0) / 2.0, "uniform").pvalue < 1e-3:
        bad.append("polar cosine not uniform")
    if stats.kstest((np.arctan2(u[:, 2], u[:, 1]) + np.pi) / (2 * np.pi), "uniform").pvalue < 1e-3:
        bad.append("azimuth not uniform")
    if np.max(np.abs(np.einsum("ij,ij->i", u, u) - 1.0)) > 1e-12:
        bad.append("|u| != 1")
    return bad


def test_isotropic_direction_closed_forms():
    assert isotropic_failures(oracle) == []


@pytest.mark.parametrize("old,new", [
    ("const double mu = 2.0 * mco_urand(rng) - 1.0;\n    double ca, sa;\n    mco_sincos_2pi(mco_urand(rng), &ca, &sa);\n    const double c = sqrt(1.0 - mu * mu);",
     "const double mu = mco_urand(rng);\n    double ca, sa;\n    mco_sincos_2pi(mco_urand(rng), &ca, &sa);\n    const double c = sqrt(1.0 - mu * mu);"),
    ("    u[1] = ca * c;\n    u[2] = sa * c;\n    u[0] = mu;", "    u[1] = ca * c;\n    u[2] = sa * mu;\n    u[0] = c;"),
])
def test_isotropic_checks_reject_planted_mistakes(old, new):
    assert isotropic_failures(mutant(old, new)) != []


# ------------------------------------------- against the reference's own code
needs_full_ref = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def ref():
    r = ref_lib()
    if not r.ref_full():
        pytest.skip("oracle/_ref built without Distribution/Reaction/Particle/Geometry")
    r.ref_isotropic_direction.argtypes = [C.POINTER(C.c_double)]
    r.ref_free_gas_scatter.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    r.ref_surface.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]
    return r


def isotropic_vs_reference(lib, ref, n=3000):
    """Largest difference between the two samplers on the same streams."""
    ref.ref_rn_init_problem(1)
    worst = 0.0
    for k in range(n):
        ref.ref_rn_init_particle(k)
        a = V3()
        ref.ref_isotropic_direction(a)
        s = C.c_uint64(lib.mco_init_particle(1, k))
        b = V3()
        lib.mco_isotropic_direction(C.byref(s), b)
        worst = max(worst, max(abs(a[i] - b[i]) for i in range(3)))
    return worst


@needs_full_ref
def test_isotropic_direction_equals_reference(ref):
    # libm's cos/sin against the oracle's own: a few ulp
    assert isotropic_vs_reference(oracle, ref) < 2e-15


@needs_full_ref
def test_reference_comparison_rejects_swapped_axes(ref):
    # a swapped pair of axes keeps the distribution isotropic: only the
    # comparison with the reference's own sampler sees it
    m = mutant("    u[1] = ca * c;\n    u[2] = sa * c;", "    u[1] = sa * c;\n    u[2] = ca * c;")
    assert isotropic_failures(m) == []
    assert isotropic_vs_reference(m, ref, 200) > 1e-3


def free_gas_vs_reference(lib, ref, keys=300):
    ref.ref_rn_init_problem(1)
    worst = 0.0
    for A in (1.0, 12.0, 56.0):
        for E0 in (1e-4, 0.0253, 1.0, 1e3, 2e6):
            for k in range(keys):
                ref.ref_rn_init_particle(k)
                e, u = C.c_double(E0), V3(0.0, 0.6, 0.8)
                ref.ref_free_gas_scatter(A, C.byref(e), u)
                s = C.c_uint64(lib.mco_init_particle(1, k))
                e2, u2 = C.c_double(E0), V3(0.0, 0.6, 0.8)
                lib.mco_free_gas_scatter(A, C.byref(e2), u2, C.byref(s))
                worst = max(worst, abs(e.value - e2.value) / e.value,
                            max(abs(u[i] - u2[i]) for i in range(3)))
    return worst


@needs_full_ref
def test_free_gas_equals_reference(ref):
    # same draws in the same order (the two log(U) of Eq. 38 left to right,
    # as the reference's build evaluates them); libm against own log/cos/sin
    assert free_gas_vs_reference(oracle, ref) < 1e-13


@needs_full_ref
def test_free_gas_reference_comparison_rejects_planted_mistake(ref):
    m = mutant("cm[k] = (vl[k] + A * (nd[k] * V)) / (1.0 + A);", "cm[k] = (vl[k] + nd[k] * V) / (1.0 + A);")
    assert free_gas_vs_reference(m, ref, keys=20) > 1e-6


@needs_full_ref
def test_reference_free_gas_reduces_to_target_at_rest_closed_forms(ref):
    """The reference's own model at 2 MeV (thermal motion ~1e-4 of the
    particle's speed) against the closed forms the oracle's target-at-rest
    sampler is held to."""
    n, A, E0 = 20000, 12.0, 2e6
    u0 = np.array([0.0, 0.6, 0.8])
    ref.ref_rn_init_problem(1)
    ref.ref_rn_init_particle(777)
    r = np.empty(n)
    mu = np.empty(n)
    for i in range(n):
        e, u = C.c_double(E0), V3(*u0)
        ref.ref_free_gas_scatter(A, C.byref(e), u)
        r[i] = e.value / E0
        mu[i] = u[0] * u0[0] + u[1] * u0[1] + u[2] * u0[2]
    alpha = ((A - 1) / (A + 1)) ** 2
    xi = 1 + alpha * math.log(alpha) / (1 - alpha)
    assert r.min() > alpha * (1 - 1e-3) and r.max() < 1 + 1e-3
    assert abs(z(r, (1 + alpha) / 2)) < 4
    assert abs(z(np.log(r), -xi)) < 4
    assert abs(z(mu, 2 / (3 * A))) < 4


@needs_full_ref
def test_surfaces_equal_reference(ref):
    """eval, distance and reflect of every surface type against the
    reference's Geometry.cpp, bit for bit (same operations in the same order)."""
    rng = np.random.default_rng(31)
    kinds = {
        1: lambda: [rng.uniform(-5, 5)],
        2: lambda: [rng.uniform(-5, 5)],
        3: lambda: [rng.uniform(-5, 5)],
        4: lambda: list(rng.normal(size=3)) + [rng.uniform(-3, 3)],
        5: lambda: list(rng.uniform(-2, 2, 3)) + [rng.uniform(0.5, 4)],
        6: lambda: list(rng.uniform(-2, 2, 2)) + [rng.uniform(0.5, 4)],
        8: lambda: list(rng.uniform(-2, 2, 2)) + [rng.uniform(0.5, 4)],
    }
    oracle.mco_surface_eval.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    for t, gen in kinds.items():
        for _ in range(400):
            p = gen()
            desc = (C.c_int32 * 4)(t, 0, 0, 0)
            c = np.zeros(8)
            c[:len(p)] = p
            # the loader's derived coefficients (SurfacePlane ctor, r^2)
            if t == 4:
                L = 2.0 / (c[0] * c[0] + c[1] * c[1] + c[2] * c[2])
                c[4:7] = L * c[0], L * c[1], L * c[2]
            elif t == 5:
                c[4] = c[3] * c[3]
            elif t in (6, 8):
                c[4] = c[2] * c[2]
            buf = bytes(desc) + c.tobytes()
            sd = C.create_string_buffer(buf, len(buf))
            x = rng.uniform(-5, 5, 3)
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            out = (C.c_double * 5)()
            ref.ref_surface(t, c.ctypes.data_as(C.POINTER(C.c_double)), V3(*x), V3(*u), out)
            assert oracle.mco_surface_eval(sd, V3(*x)) == out[0], t
            assert oracle.mco_surface_distance(sd, V3(*x), V3(*u)) == out[1], t
            d = V3(*u)
            oracle.mco_surface_reflect(sd, d)
            assert (d[0], d[1], d[2]) == (out[2], out[3], out[4]), t
