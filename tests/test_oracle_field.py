"""Oracle uniform-field propagation against the reference's golden values
(test/celeritas/field/FieldPropagator.test.cc), CPU only.

The Dormand-Prince step counts pin the driver's step control (chord
search, accurate advance, substep limit); positions and directions pin
the stepper and the chord/boundary logic. Closed forms (helix radius
p/(qB), arc length) pin the field equation itself.
"""

import math

import numpy as np
import pytest

from tests import field_cases as fc
from tests import oracle_bind as ob

INNER, WORLD = 1, 2


@pytest.fixture(scope="module")
def problem():
    return ob.Problem(fc.two_boxes())


def prop(problem, pos, d, energy, charge, field, steps, max_substeps=0):
    return ob.field_propagate(problem, pos, d, fc.momentum(energy), charge,
                              [0, 0, field], steps, max_substeps)


def test_helix_radius_closed_form():
    """r = p / (q B c'): the golden radius of electron_interior."""
    p = fc.momentum(10.9181415106)
    assert p / (fc.LORENTZ * 1.0) == pytest.approx(3.8085385437789383, rel=1e-10)
    assert fc.momentum(10) / (fc.LORENTZ * fc.UNIT_RADIUS_FIELD) == pytest.approx(1, rel=1e-12)


def test_electron_interior(problem):
    radius, steps = fc.radius_10mev_step()
    steps = steps + [math.pi * radius - 8.7323805094658429, 1e-10]
    r = prop(problem, [radius, 0, 0], [0, 1, 0], 10.9181415106, -1, 1.0, steps)
    count = np.diff(np.concatenate([[0], r[:, 10]]))
    # Short step (:197-207)
    assert r[0, 0] == pytest.approx(1e-2, rel=1e-12)
    np.testing.assert_allclose(r[0, 3:6], [3.80852541539105, 0.0099999885096862, 0],
                               rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(r[0, 6:9], [-0.00262567606832303, 0.999996552906651, 0],
                               rtol=1e-12, atol=1e-15)
    assert count[0] == 1
    # Quarter turn in 25 steps, one stepper call each (:210-226)
    for i in range(1, 26):
        assert r[i, 0] == pytest.approx(steps[i], rel=1e-12)
        assert not r[i, 1]
        assert count[i] == 1
    assert np.linalg.norm(r[25, 3:6] - [0, radius, 0]) < fc.COARSE_EPS
    assert np.dot([-1, 0, 0], r[25, 6:9]) == pytest.approx(1.0, rel=1e-12)
    # Next quarter turn in one call: 21 stepper calls (:229-237)
    assert r[26, 0] == pytest.approx(0.5 * math.pi * radius, rel=1e-12)
    assert np.linalg.norm(r[26, 3:6] - [-radius, 0, 0]) < fc.COARSE_EPS
    assert np.dot([0, -1, 0], r[26, 6:9]) == pytest.approx(1.0, rel=1e-12)
    assert count[26] == 21
    # Half turn: substep limit reached first, then the rest (:240-255)
    assert r[27, 0] == pytest.approx(8.7323805094658429, rel=1e-12)
    assert r[28, 0] == pytest.approx(math.pi * radius - 8.7323805094658429, rel=1e-12)
    assert np.linalg.norm(r[28, 3:6] - [radius, 0, 0]) < 1e-5
    assert np.dot([0, 1, 0], r[28, 6:9]) == pytest.approx(1.0, rel=1e-12)
    assert count[27] + count[28] == 40
    # Step below the driver minimum (:258-268)
    assert r[29, 0] == 1e-10
    assert not r[29, 1]
    np.testing.assert_allclose(r[29, 3:6], [3.8085385881855, -2.381487075086e-07, 0],
                               atol=fc.COARSE_EPS)
    np.testing.assert_allclose(r[29, 6:9], [6.25302065531623e-08, 1, 0], atol=fc.COARSE_EPS)
    assert count[29] == 1


def test_positron_interior(problem):
    """Clockwise quarter turn of radius 1 (:271-293)."""
    r = prop(problem, [1, 0, 0], [0, -1, 0], 10, 1, fc.UNIT_RADIUS_FIELD, [0.5 * math.pi])
    assert r[0, 0] == pytest.approx(0.5 * math.pi, rel=1e-12)
    assert np.linalg.norm(r[0, 3:6] - [0, -1, 0]) < 1e-5
    assert np.dot([-1, 0, 0], r[0, 6:9]) == pytest.approx(1.0, rel=1e-12)


def test_gamma_interior(problem):
    """Neutral tracks go straight with one stepper call per chord (:296-355)."""
    r = prop(problem, [0, 0, 0], [0, 0, 1], 1, 0, 1234.5, [3.0, -1, 3.0, 0, 5.0])
    assert r[0, 0] == pytest.approx(3.0) and not r[0, 1]
    np.testing.assert_allclose(r[0, 3:6], [0, 0, 3], atol=1e-12)
    assert r[0, 10] == 1
    assert r[2, 0] == pytest.approx(2.0) and r[2, 1]
    np.testing.assert_allclose(r[2, 3:6], [0, 0, 5], atol=1e-12)
    assert r[2, 10] - r[0, 10] == 2
    assert r[2, 9] == INNER and r[3, 9] == WORLD
    assert r[4, 0] == pytest.approx(5.0) and not r[4, 1]
    np.testing.assert_allclose(r[4, 3:6], [0, 0, 10], atol=1e-12)
    assert r[4, 10] - r[3, 10] == 1


def test_gamma_exit(problem):
    """Exact boundary and a request just past it (:383-446)."""
    r = prop(problem, [2, 4.75, 0], [0, 1, 0], 1, 0, 12345.6, [0.25, 0])
    assert r[0, 0] == pytest.approx(0.25) and r[0, 1]
    assert np.linalg.norm(r[0, 3:6] - [2, 5, 0]) < 1e-5
    assert r[0, 10] == 1
    assert r[0, 9] == INNER and r[1, 9] == WORLD
    r = prop(problem, [2, 4.749, 0], [0, 1, 0], 1, 0, 12345.6, [0.251 + 1e-7])
    assert r[0, 0] == pytest.approx(0.251, rel=1e-12) and r[0, 1]
    assert np.linalg.norm(r[0, 3:6] - [2, 5, 0]) < 1e-5


def test_electron_super_small_step(problem):
    """Tiny requests: far from a boundary, and bumped off one (:448-492)."""
    got = []
    bump = 1e-5 * 0.1
    for delta in (1e-20, 1e-14, 1e-8, 1e-2, 0.1):
        r = prop(problem, [9.5, 9.5, 9.5], [1, 0, 0], 2, -1, 1.0, [delta])
        assert r[0, 0] == delta
        assert r[0, 10] == 1
        r = prop(problem, [5 + bump * 0.99, 0, 0], [-1, 0, 0], 2, -1, 1.0, [delta])
        got.append(r[0, 0])
        assert r[0, 10] == 1
    np.testing.assert_allclose(got, [1e-20, 1e-14, 1e-08, 9.9002453648129e-07,
                                     9.924578491937e-07], rtol=1e-12)


def test_electron_small_step(problem):
    """Small steps up to, onto and from the boundary (:495-583)."""
    delta = 1e-7
    b = fc.UNIT_RADIUS_FIELD
    r = prop(problem, [5 - delta - 1e-5, 0, 0], [1, 0, 0], 10, -1, b, [delta])
    assert r[0, 0] == pytest.approx(delta, rel=1e-12) and not r[0, 1]
    np.testing.assert_allclose(r[0, 3:6], [5 - 1e-5, 0, 0], atol=fc.COARSE_EPS)
    r = prop(problem, [5 - 2 * delta, 0, 0], [1, 0, 0], 10, -1, b, [delta])
    assert r[0, 0] == pytest.approx(delta, rel=1e-12) and not r[0, 1]
    np.testing.assert_allclose(r[0, 3:6], [4.9999999, 0, 0], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(r[0, 6:9], [1, delta, 0], rtol=1e-12, atol=1e-18)
    r = prop(problem, [5 - delta, 0, 0], [1, 0, 0], 10, -1, b, [2 * delta, 0, delta])
    assert r[0, 0] <= 2 * delta
    assert r[0, 0] == pytest.approx(1.0000000044408872e-07, rel=1e-12)
    assert r[0, 1]
    np.testing.assert_allclose(r[0, 3:6], [5, 0, 0], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(r[0, 6:9], [1, 2 * delta, 0], rtol=1e-12, atol=1e-18)
    assert r[0, 9] == INNER and r[1, 9] == WORLD
    assert r[2, 0] == delta and not r[2, 1]
    assert np.linalg.norm(r[2, 3:6] - [5 + delta, 0, 0]) < 1e-12
    assert np.linalg.norm(r[2, 6:9] - [1, 3 * delta, 0]) < 1e-12


def test_electron_tangent(problem):
    """Near-tangent path under the boundary (:586-622)."""
    q1, q2 = 0.49 * math.pi, 0.51 * math.pi
    r = prop(problem, [1, 4, 0], [0, 1, 0], 10, -1, fc.UNIT_RADIUS_FIELD,
             [q1, -1, 0.02 * math.pi])
    assert not r[0, 1] and r[0, 0] == pytest.approx(q1, rel=1e-12)
    assert np.linalg.norm(r[0, 3:6] - [math.cos(q1), 4 + math.sin(q1), 0]) < 2e-6
    assert not r[2, 1] and r[2, 0] == pytest.approx(0.02 * math.pi, rel=1e-12)
    assert np.linalg.norm(r[2, 3:6] - [math.cos(q2), 4 + math.sin(q2), 0]) < 2e-6


def test_electron_cross(problem):
    """Exit, reentry and return over a radius-2 circle (:625-695)."""
    b = 0.5 * fc.UNIT_RADIUS_FIELD
    circ = 2.0 * 2 * math.pi
    s3 = math.sqrt(3)
    r = prop(problem, [2, 4, 0], [0, 1, 0], 10, -1, b, [math.pi, 0, circ, 0])
    assert r[0, 0] / circ == pytest.approx(1 / 12, abs=1e-5) and r[0, 1]
    assert np.linalg.norm(r[0, 3:6] - [s3, 5, 0]) < 1e-5
    assert np.linalg.norm(r[0, 6:9] - [-0.5, s3 / 2, 0]) < 1e-5
    assert r[1, 9] == WORLD
    assert r[2, 0] / circ == pytest.approx(1 / 3, abs=1e-5) and r[2, 1]
    assert np.linalg.norm(r[2, 3:6] - [-s3, 5, 0]) < 1e-5
    assert np.linalg.norm(r[2, 6:9] - [-0.5, -s3 / 2, 0]) < 1e-5
    assert r[3, 9] == INNER
    # Return to the start with a larger substep limit
    pos, d = r[3, 3:6], r[3, 6:9]
    r2 = ob.field_propagate(problem, pos, d, fc.momentum(10), -1, [0, 0, b],
                            [7 / 12 * circ], max_substeps=100)
    assert r2[0, 0] / circ == pytest.approx(7 / 12, abs=1e-5) and not r2[0, 1]
    assert np.linalg.norm(r2[0, 3:6] - [2, 4, 0]) < 2e-5
    assert np.linalg.norm(r2[0, 6:9] - [0, 1, 0]) < 1e-5


def test_electron_tangent_cross(problem):
    """Barely hitting and barely missing the boundary (:698-751)."""
    b = fc.UNIT_RADIUS_FIELD
    circ = 2 * math.pi
    dy = 1.1 * 0.025
    r = prop(problem, [1, 4 + dy, 0], [0, 1, 0], 10, -1, b, [circ, 0])
    theta = math.asin(1 - dy)
    x = math.sqrt(2 * dy - dy * dy)
    assert r[0, 0] == pytest.approx(theta, rel=0.025) and r[0, 1]
    assert np.linalg.norm(r[0, 3:6] - [x, 5, 0]) < 2e-5
    assert np.linalg.norm(r[0, 6:9] - [dy - 1, x, 0]) < 2e-5
    assert r[1, 9] == WORLD
    dy = 0.9 * 0.025
    r = prop(problem, [1, 4 + dy, 0], [0, 1, 0], 10, -1, b, [circ], max_substeps=100)
    assert r[0, 0] == pytest.approx(circ, rel=1e-12) and not r[0, 1]
    assert np.linalg.norm(r[0, 3:6] - [1, 4 + dy, 0]) < 2e-5
    assert np.linalg.norm(r[0, 6:9] - [0, 1, 0]) < 1e-5


def test_planted_driver_mistake_changes_counts(problem):
    """A coarser chord tolerance is visible in the golden step counts: with
    the substep limit at 5 the half turn stops at a different distance."""
    radius, steps = fc.radius_10mev_step()
    r = prop(problem, [radius, 0, 0], [0, 1, 0], 10.9181415106, -1, 1.0,
             steps[:-1] + [math.pi * radius], max_substeps=5)
    assert r[-1, 0] != pytest.approx(8.7323805094658429, rel=1e-6)
