"""Oracle grid calculators against the reference's golden values (CPU only).

Pins the oracle's XsCalculator (log-uniform grid, linear interpolation in
energy, 1/E scaling at and above the prime index), RangeCalculator (sqrt(E)
scaling below the grid) and InverseRangeCalculator (quadratic below the
grid) with the cases of test/celeritas/grid/XsCalculator.test.cc:38-181,
RangeCalculator.test.cc:21-60 and InverseRangeCalculator.test.cc:21-70,
built the way CalculatorTestBase.cc:56-101 builds them. Each planted
mistake in the oracle (a dropped unscaling, log-space interpolation, the
wrong power below the grid) fails at least one of the golden cases.
"""

import ctypes
import math

import numpy as np
import pytest

from tests import oracle_bind as ob


def _eval(emin, emax, values, prime, kind, xs):
    lib = ob.load()
    f = lib.orc_grid_eval
    f.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32,
                  ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                  ctypes.c_void_p]
    v = np.ascontiguousarray(values, dtype=np.float64)
    x = np.ascontiguousarray(xs, dtype=np.float64)
    out = np.zeros(len(x))
    f(emin, emax, len(v), prime, v.ctypes.data, kind, x.ctypes.data, len(x), out.ctypes.data)
    return out


def _mutate(m):
    f = ob.load().orc_grid_set_mutation
    f.argtypes = [ctypes.c_int]
    f(m)


@pytest.fixture(autouse=True)
def _no_mutation():
    _mutate(0)
    yield
    _mutate(0)


NO_PRIME = 0xFFFFFFFF


def build(emin, emax, count, fn):
    """CalculatorTestBase::build: values of fn at exp(log-uniform points)."""
    loge = np.linspace(math.log(emin), math.log(emax), count)
    # UniformGridData::from_bounds: front + delta * i
    delta = (math.log(emax) - math.log(emin)) / (count - 1)
    e = np.exp(math.log(emin) + delta * np.arange(count))
    assert np.allclose(np.exp(loge), e)
    return e, np.array([fn(x) for x in e], dtype=np.float64)


def to_prime(e, values, prime):
    """CalculatorTestBase::convert_to_prime: scale values at i >= prime by E."""
    v = values.copy()
    v[prime:] *= e[prime:]
    return v


def soft_eq(expected, actual, rel=1e-12):
    return abs(expected - actual) <= rel * max(abs(expected), 1.0)


# (emin, emax, count, values builder, prime, [(energy, expected)])
def _cases():
    cases = []
    # simple: XS = E, no scaling
    e, v = build(1.0, 1e5, 6, lambda x: x)
    cases.append(("simple", 1.0, 1e5, v, NO_PRIME,
                  [(1, 1.0), (1e2, 1e2), (1e5 - 1e-6, 1e5 - 1e-6), (1e5, 1e5), (5, 5),
                   (1e-4, 1.0), (1e7, 1e5)]))
    # scaled_lowest: values of 1 converted to prime at index 0
    e, v = build(0.1, 1e4, 6, lambda x: 1.0)
    cases.append(("scaled_lowest", 0.1, 1e4, to_prime(e, v, 0), 0,
                  [(0.1, 1), (1e2, 1), (1e4 - 1e-6, 1), (1e4, 1), (0.2, 1), (5, 1),
                   (1e-4, 1000), (1e5, 0.1)]))
    # scaled_middle: values of 3, prime index 3
    e, v = build(0.1, 1e4, 6, lambda x: 3.0)
    cases.append(("scaled_middle", 0.1, 1e4, to_prime(e, v, 3), 3,
                  [(0.1, 3), (1e2, 3), (1e4 - 1e-6, 3), (1e4, 3), (0.2, 3), (5, 3),
                   (1e-4, 3), (1e5, 0.3)]))

    # scaled_linear: 100 + 10 E, divided by E above 1 MeV; prime index 3
    def ref_xs(x):
        r = 100 + x * 10
        return r / x if x > 1 else r
    e, v = build(1e-3, 1e3, 7, ref_xs)
    cases.append(("scaled_linear", 1e-3, 1e3, to_prime(e, v, 3), 3,
                  [(x, ref_xs(x)) for x in (1e-3, 1e-1, 0.5, 1.0, 1.5, 10.0, 12.5, 1e3)]))
    # scaled_highest: values 1, 10, 100 with prime index 2 -> xs {1, 10, 1}
    e, v = build(1, 100, 3, lambda x: x)
    cases.append(("scaled_highest", 1, 100, v, 2,
                  [(1e-4, 1), (1, 1), (10, 10), (90, 2.0), (100, 1), (1000, 0.1)]))
    return cases


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_xs_calculator_golden(case):
    name, emin, emax, values, prime, points = case
    got = _eval(emin, emax, values, prime, 0, [p[0] for p in points])
    for (energy, expected), g in zip(points, got):
        # EXPECT_SOFT_EQ (relative 1e-12)
        assert soft_eq(expected, g), f"{name}: xs({energy}) = {g!r}, expected {expected}"


#: Range = E / 20 on 10..1e4 MeV with 4 points (RangeCalculator.test.cc:24-35)
RANGE_E, RANGE_V = build(10, 1e4, 4, lambda x: x * 0.05)
RANGE_POINTS = [(1, 0.5 * math.sqrt(1 / 10)), (2, 0.5 * math.sqrt(2 / 10)), (10, 0.5),
                (20, 1.0), (100, 5.0), (1e4, 500), (1.001e4, 500)]


def test_range_calculator_golden():
    got = _eval(10, 1e4, RANGE_V, NO_PRIME, 1, [p[0] for p in RANGE_POINTS])
    for (energy, expected), g in zip(RANGE_POINTS, got):
        assert soft_eq(expected, g), f"range({energy}) = {g!r}, expected {expected}"


def test_inverse_range_calculator_golden():
    v = RANGE_V.copy()
    # Final point set exactly for the top-of-range test
    # (InverseRangeCalculator.test.cc:38-40)
    assert soft_eq(500, v[-1])
    v[-1] = 500
    points = [(0.5 * math.sqrt(1 / 10), 1.0), (0.5 * math.sqrt(2 / 10), 2.0), (0.5, 10.0),
              (1, 20.0), (5, 100.0), (500, 1e4)]
    got = _eval(10, 1e4, v, NO_PRIME, 2, [p[0] for p in points])
    for (rng, expected), g in zip(points, got):
        assert soft_eq(expected, g), f"inverse_range({rng}) = {g!r}, expected {expected}"


def test_range_inverse_roundtrip():
    """Inverse range of the range returns the energy on and between the
    grid points and below the grid (both calculators are piecewise linear
    in E on the same nodes, so they are exact inverses there)."""
    energies = np.geomspace(0.5, 1e4, 97)
    ranges = _eval(10, 1e4, RANGE_V, NO_PRIME, 1, energies)
    back = _eval(10, 1e4, RANGE_V, NO_PRIME, 2, ranges)
    assert np.allclose(back, energies, rtol=1e-12, atol=0)


def _golden_failures():
    failures = 0
    for name, emin, emax, values, prime, points in CASES:
        got = _eval(emin, emax, values, prime, 0, [p[0] for p in points])
        failures += sum(not soft_eq(exp, g) for (_, exp), g in zip(points, got))
    got = _eval(10, 1e4, RANGE_V, NO_PRIME, 1, [p[0] for p in RANGE_POINTS])
    failures += sum(not soft_eq(exp, g) for (_, exp), g in zip(RANGE_POINTS, got))
    v = RANGE_V.copy()
    v[-1] = 500
    pts = [(0.5 * math.sqrt(1 / 10), 1.0), (0.5 * math.sqrt(2 / 10), 2.0), (5, 100.0)]
    got = _eval(10, 1e4, v, NO_PRIME, 2, [p[0] for p in pts])
    failures += sum(not soft_eq(exp, g) for (_, exp), g in zip(pts, got))
    return failures


def test_unmutated_has_no_failures():
    assert _golden_failures() == 0


@pytest.mark.parametrize("mutation", [1, 2, 3, 4])
def test_planted_grid_mistake_fails(mutation):
    """Each planted mistake in the oracle's calculators breaks a golden case."""
    _mutate(mutation)
    assert _golden_failures() > 0
