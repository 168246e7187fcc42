"""Pins the oracle against the reference's own known-answer values and, where
the reference's sources build here, against the reference code itself.

Golden values:
  Random.cpp:111-117       RN_CHECK seeds for draws 1-5 and 123456-123460
  test/test_Solver.cpp     interpolation, quadratic solver, binary search,
                           scatter_direction vectors (Catch Approx)
  test/test_Plane.cpp      generic plane eval / distance
  test/unit_test.cpp       plane-x distance and reflection
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from mc_testlib import REF_LIB, dptr, oracle, ref_lib

FLT_MAX = 3.4028234663852886e38
RN_CHECK = [3512401965023503517, 5461769869401032777, 1468184805722937541, 5160872062372652241,
            6637647758174943277, 794206257475890433, 4662153896835267997, 6075201270501039433,
            889694366662031813, 7299299962545529297]


def approx(a, b):
    # Catch 1.x Approx: |a-b| < eps*(scale + max(|a|,|b|)), eps = 100*FLT_EPSILON, scale = 1
    return abs(a - b) < 100 * 1.1920929e-07 * (1.0 + max(abs(a), abs(b)))


def test_rn_check_sequential():
    s, seeds = 1, []
    for i in range(123460):
        s = oracle.mco_lcg_next(s)
        if i < 5 or i >= 123455:
            seeds.append(s)
    assert seeds == RN_CHECK


def test_rn_skip_ahead_matches_check():
    for j, idx in enumerate([1, 2, 3, 4, 5, 123456, 123457, 123458, 123459, 123460]):
        assert oracle.mco_skip_ahead(1, idx) == RN_CHECK[j]


def test_first_uniforms():
    # test_Solver.cpp:114-115
    s = C.c_uint64(1)
    u = [round(oracle.mco_urand(C.byref(s)), 6) for _ in range(4)]
    assert u == [0.380815, 0.592166, 0.159181, 0.559543]


def test_init_particle_is_stride_skip():
    for nps in [0, 1, 7, 1000, 123457]:
        assert oracle.mco_init_particle(1, nps) == oracle.mco_skip_ahead(1, nps * 152917)
    # skip-ahead composes: skip(a) then skip(b) == skip(a+b)
    s = oracle.mco_skip_ahead(12345, 777)
    assert oracle.mco_skip_ahead(s, 1000) == oracle.mco_skip_ahead(12345, 1777)


def test_interpolate():
    # test_Solver.cpp:9-31
    assert oracle.mco_interpolate(1.5, 1.0, 2.0, 2.0, 4.0) == 3.0
    assert oracle.mco_interpolate(1.5, 1.0, 2.0, 4.0, 2.0) == 3.0
    assert oracle.mco_interpolate(1.5, 1.0, 2.0, 3.0, 3.0) == 3.0


@pytest.mark.parametrize("a,b,c,expect", [(10, 1, 10, FLT_MAX), (1, 2, 1, FLT_MAX), (1, 3, 2, FLT_MAX),
                                          (1, -3, 2, 1.0), (1, -1, -2, 2.0)])
def test_geometry_quad(a, b, c, expect):
    # test_Solver.cpp:33-76
    assert approx(oracle.mco_geometry_quad(a, b, c), expect)


def test_binary_search():
    # test_Solver.cpp:78-109
    vec = np.array([2.0, 3, 4, 5, 6, 7, 8, 9, 10])
    bs = lambda x, v: oracle.mco_binary_search(x, dptr(v), v.size)
    assert bs(4.3, vec) == 2 and bs(7.3, vec) == 5
    assert bs(4.0, vec) == 1 and bs(7.0, vec) == 4 and bs(2.0, vec) == -1 and bs(10.0, vec) == 7
    assert bs(1.0, vec) == -1 and bs(12.0, vec) == 8
    mg = np.array([0.0, 5, 5, 10])
    assert bs(1, mg) == 0 and bs(8, mg) == 2


def test_scatter_direction_golden():
    # test_Solver.cpp:111-165: the four sections draw LCG numbers 1-4 in turn
    cases = [((2, 2, 1), 0.4, (-0.332775973037166, 0.549648481770809, 0.766254982532715)),
             ((2, 2, 1), -0.4, (-0.092800252430585, -0.802140132078611, 0.589880769018391)),
             ((0, 0, 1), 0.4, (-0.771301959816257, -0.495068971743938, 0.4)),
             ((0, 0, 1), -0.4, (0.334943449022552, 0.853119502740898, -0.4))]
    s = C.c_uint64(1)
    for d, mu, expect in cases:
        n = math.sqrt(sum(x * x for x in d))
        din = (C.c_double * 3)(*[x / n for x in d])
        out = (C.c_double * 3)()
        oracle.mco_scatter_direction(din, mu, C.byref(s), out)
        for k in range(3):
            assert approx(out[k], expect[k]), (d, mu, k, out[k], expect[k])
        # invariants: unit length, cosine with the incoming direction is mu
        assert abs(sum(out[k] ** 2 for k in range(3)) - 1.0) < 1e-14
        assert abs(sum(out[k] * din[k] for k in range(3)) - mu) < 1e-14


class Surf(C.Structure):
    _fields_ = [("type", C.c_int32), ("bc", C.c_int32), ("est_begin", C.c_int32), ("est_count", C.c_int32),
                ("c", C.c_double * 8)]


def plane(a, b, c, d):
    L = 2.0 / (a * a + b * b + c * c)
    return Surf(4, 0, 0, 0, (C.c_double * 8)(a, b, c, d, L * a, L * b, L * c, 0))


def v3(*x):
    return (C.c_double * 3)(*x)


def test_generic_plane():
    # test_Plane.cpp:11-93 with a,b,c,d = 1,2,-3,4
    S = plane(1.0, 2.0, -3.0, 4.0)
    assert approx(oracle.mco_surface_eval(C.byref(S), v3(3, 2, -1)), 6.0)
    assert approx(oracle.mco_surface_eval(C.byref(S), v3(-3, -2, 1)), -14.0)
    assert approx(oracle.mco_surface_eval(C.byref(S), v3(3, 2, 1)), 0.0)
    assert approx(oracle.mco_surface_distance(C.byref(S), v3(-3, -2, 1), v3(1, 0, 0)), 14.0)
    assert oracle.mco_surface_distance(C.byref(S), v3(-3, -2, 1), v3(-1, 0, 0)) == FLT_MAX
    r3 = 1 / math.sqrt(3)
    assert oracle.mco_surface_distance(C.byref(S), v3(-3, -2, 1), v3(r3, r3, r3)) == FLT_MAX
    rng = np.random.default_rng(3)
    for p in rng.uniform(-10, 10, (1000, 3)):
        assert oracle.mco_surface_eval(C.byref(S), v3(*p)) == 1.0 * p[0] + 2.0 * p[1] + -3.0 * p[2] - 4.0


def test_plane_x_unit():
    # unit_test.cpp:105-126
    S = Surf(1, 0, 0, 0, (C.c_double * 8)(0.0))
    assert oracle.mco_surface_eval(C.byref(S), v3(1, 0, 0)) > 0
    assert oracle.mco_surface_eval(C.byref(S), v3(-1, 0, 0)) < 0
    assert oracle.mco_surface_distance(C.byref(S), v3(1, 2, 3), v3(-1, 0, 0)) == 1.0
    assert oracle.mco_surface_distance(C.byref(S), v3(1, 2, 3), v3(1, 0, 0)) == FLT_MAX
    assert oracle.mco_surface_distance(C.byref(S), v3(-1, 2, 3), v3(1, 0, 0)) == 1.0
    assert oracle.mco_surface_distance(C.byref(S), v3(-1, 2, 3), v3(-1, 0, 0)) == FLT_MAX
    assert oracle.mco_surface_distance(C.byref(S), v3(-1, 2, 3), v3(0, 1, 0)) == FLT_MAX
    u = v3(0.3, 0.1, 0.4)
    oracle.mco_surface_reflect(C.byref(S), u)
    assert list(u) == [-0.3, 0.1, 0.4]


def test_generic_plane_reflection_is_mirror():
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = rng.normal(size=3)
        S = plane(*n, 1.0)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        w = v3(*u)
        oracle.mco_surface_reflect(C.byref(S), w)
        nh = n / np.linalg.norm(n)
        expect = u - 2 * np.dot(u, nh) * nh
        assert np.allclose(list(w), expect, atol=1e-14)


def test_sphere_and_cylinder_distances():
    # closed forms: from the centre of a sphere of radius r the distance is r
    sph = Surf(5, 0, 0, 0, (C.c_double * 8)(1.0, 2.0, 3.0, 2.5, 6.25))
    rng = np.random.default_rng(9)
    for _ in range(100):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        assert abs(oracle.mco_surface_distance(C.byref(sph), v3(1, 2, 3), v3(*u)) - 2.5) < 1e-14
    cz = Surf(8, 0, 0, 0, (C.c_double * 8)(0.0, 0.0, 2.0, 0, 4.0))
    # along +x from the axis: 2; moving parallel to the axis: never
    assert abs(oracle.mco_surface_distance(C.byref(cz), v3(0, 0, 5), v3(1, 0, 0)) - 2.0) < 1e-15
    assert oracle.mco_surface_distance(C.byref(cz), v3(0, 0, 5), v3(0, 0, 1)) == FLT_MAX
    cx = Surf(6, 0, 0, 0, (C.c_double * 8)(1.0, 1.0, 1.0, 0, 1.0))
    assert abs(oracle.mco_surface_distance(C.byref(cx), v3(7, 1, 1), v3(0, 0, -1)) - 1.0) < 1e-15


def test_log_against_libm():
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.random(20000), np.exp(rng.uniform(-700, 700, 5000)), [1.0, 0.5, 2.0, 1e-300]])
    for x in xs:
        got, want = oracle.mco_log(float(x)), math.log(x)
        assert abs(got - want) <= 2.5 * np.spacing(abs(want)) + 1e-300, (x, got, want)
    assert oracle.mco_log(1.0) == 0.0


def test_sincos_against_libm():
    rng = np.random.default_rng(12)
    c, s = C.c_double(), C.c_double()
    for x in np.concatenate([rng.random(20000), [0.0, 0.125, 0.25, 0.5, 0.75, 1.0]]):
        oracle.mco_sincos_2pi(float(x), C.byref(c), C.byref(s))
        assert abs(c.value - math.cos(2 * math.pi * x)) < 1e-15
        assert abs(s.value - math.sin(2 * math.pi * x)) < 1e-15
        assert abs(c.value ** 2 + s.value ** 2 - 1.0) < 4e-16


# ------------------------------------------- against the reference's own code
class _LazyRef:
    """oracle/_ref, loaded on first use (so collecting this module, as a GPU
    test run does, maps nothing of it)."""

    _lib = None

    def __getattr__(self, name):
        if _LazyRef._lib is None:
            _LazyRef._lib = ref_lib()
        return getattr(_LazyRef._lib, name)


ref = _LazyRef()
needs_ref = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (reference sources absent)")


@needs_ref
def test_ref_lcg_stream_and_skip():
    ref.ref_rn_init_problem(1)
    s = C.c_uint64(1)
    for _ in range(1000):
        assert ref.ref_urand() == oracle.mco_urand(C.byref(s))
    for seed, n in [(1, 5), (1, 123456), (987654321, 10 ** 9), (3, 2 ** 40 + 17)]:
        assert ref.ref_skip_ahead(seed, n) == oracle.mco_skip_ahead(seed, n)
    for nps in [0, 1, 2, 99, 10 ** 6]:
        ref.ref_rn_init_problem(1)
        ref.ref_rn_init_particle(nps)
        s = C.c_uint64(oracle.mco_init_particle(1, nps))
        for _ in range(5):
            assert ref.ref_urand() == oracle.mco_urand(C.byref(s))


@needs_ref
def test_ref_binary_search_interpolate_quad():
    rng = np.random.default_rng(21)
    for _ in range(200):
        v = np.sort(rng.uniform(0, 10, rng.integers(1, 40)))
        if rng.random() < 0.3:
            v = np.repeat(v, 2)
        for x in np.concatenate([rng.uniform(-1, 11, 10), v[:3]]):
            assert ref.ref_binary_search(x, dptr(v), v.size) == oracle.mco_binary_search(x, dptr(v), v.size)
    for a in rng.uniform(-5, 5, (2000, 5)):
        assert ref.ref_interpolate(*a) == oracle.mco_interpolate(*a) or (
            math.isnan(ref.ref_interpolate(*a)) and math.isnan(oracle.mco_interpolate(*a)))
    for a, b, c in rng.uniform(-3, 3, (5000, 3)):
        assert ref.ref_geometry_quad(a, b, c) == oracle.mco_geometry_quad(a, b, c)


@needs_ref
def test_ref_xstable_bitwise():
    rng = np.random.default_rng(22)
    grid = np.sort(np.exp(rng.uniform(-11, 16, 300)))
    y = rng.uniform(0.1, 50, 300)
    for E in np.concatenate([np.exp(rng.uniform(-13, 18, 3000)), grid[:20], [grid[0], grid[-1]]]):
        r = ref.ref_xstable(dptr(grid), dptr(y), grid.size, E)
        k = oracle.mco_binary_search(E, dptr(grid), grid.size)
        if k == grid.size - 1:
            o = y[-1]
        elif k == -1:
            o = y[0]
        else:
            o = oracle.mco_interpolate(E, grid[k], grid[k + 1], y[k], y[k + 1])
        assert r == o


@needs_ref
def test_ref_scatter_direction_and_exponential():
    # same LCG draws; the reference evaluates cos/sin(PI_2*xi) with libm, the
    # oracle reduces exactly on xi: agreement to a few ulp
    rng = np.random.default_rng(23)
    ref.ref_rn_init_problem(1)
    s = C.c_uint64(1)
    for _ in range(2000):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        mu = rng.uniform(-1, 1)
        a, b = (C.c_double * 3)(), (C.c_double * 3)()
        ref.ref_scatter_direction(v3(*d), mu, a)
        oracle.mco_scatter_direction(v3(*d), mu, C.byref(s), b)
        assert max(abs(a[k] - b[k]) for k in range(3)) < 2e-14
    ref.ref_rn_init_problem(1)
    s = C.c_uint64(1)
    for sig in rng.uniform(0.01, 10, 2000):
        want = ref.ref_exponential_sample(sig)
        got = -oracle.mco_log(oracle.mco_urand(C.byref(s))) / sig
        assert abs(got - want) <= 4 * np.spacing(want)
