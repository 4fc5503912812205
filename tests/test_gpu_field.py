"""Uniform-field propagation on the device against the oracle and the
reference's golden values (test/celeritas/field/FieldPropagator.test.cc).

The device and the oracle evaluate the same FMA chains, so positions,
directions and distances agree to rounding (the step-size control's
exp/log may differ in the last bit between device and host libraries,
so the tolerance is 1e-10 rather than exact).
"""

import math

import numpy as np
import pytest

from tests import field_cases as fc
from tests import oracle_bind as ob

pytestmark = pytest.mark.gpu

# (pos, dir, energy, charge, Bz, requests, max_substeps)
CASES = {
    "electron_interior": ([3.8085385437789383, 0, 0], [0, 1, 0], 10.9181415106, -1, 1.0,
                          fc.radius_10mev_step()[1]
                          + [math.pi * 3.8085385437789383 - 8.7323805094658429, 1e-10], 0),
    "positron_interior": ([1, 0, 0], [0, -1, 0], 10, 1, fc.UNIT_RADIUS_FIELD,
                          [0.5 * math.pi], 0),
    "gamma_interior": ([0, 0, 0], [0, 0, 1], 1, 0, 1234.5, [3.0, -1, 3.0, 0, 5.0], 0),
    "electron_small_step": ([5 - 1e-7, 0, 0], [1, 0, 0], 10, -1, fc.UNIT_RADIUS_FIELD,
                            [2e-7, 0, 1e-7], 0),
    "electron_cross": ([2, 4, 0], [0, 1, 0], 10, -1, 0.5 * fc.UNIT_RADIUS_FIELD,
                       [math.pi, 0, 4 * math.pi, 0], 0),
    "electron_tangent_cross": ([1, 4 + 1.1 * 0.025, 0], [0, 1, 0], 10, -1,
                               fc.UNIT_RADIUS_FIELD, [2 * math.pi, 0], 0),
    "tangent_miss": ([1, 4 + 0.9 * 0.025, 0], [0, 1, 0], 10, -1, fc.UNIT_RADIUS_FIELD,
                     [2 * math.pi], 100),
}


@pytest.fixture(scope="module")
def params(amd):
    return amd.CoreParams(fc.two_boxes())


@pytest.fixture(scope="module")
def oproblem():
    return ob.Problem(fc.two_boxes())


def _device(amd, params, pos, d, energy, charge, bz, steps, nsub):
    t = np.array([list(pos) + list(d) + [fc.momentum(energy), charge, 0, 0, bz]])
    return amd.verify_field_propagate(params, t, np.array([steps]), nsub)[0]


@pytest.mark.parametrize("name", sorted(CASES))
def test_case_matches_oracle(amd, params, oproblem, name):
    pos, d, energy, charge, bz, steps, nsub = CASES[name]
    dev = _device(amd, params, pos, d, energy, charge, bz, steps, nsub)
    ora = ob.field_propagate(oproblem, pos, d, fc.momentum(energy), charge, [0, 0, bz],
                             steps, nsub)
    print(name, dev[:, 0], flush=True)
    np.testing.assert_array_equal(dev[:, 1:3], ora[:, 1:3])  # boundary, looping
    np.testing.assert_array_equal(dev[:, 9], ora[:, 9])  # volume
    np.testing.assert_allclose(dev[:, 0], ora[:, 0], rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(dev[:, 3:9], ora[:, 3:9], rtol=1e-10, atol=1e-12)


def test_electron_interior_golden(amd, params):
    """Substep-limited half turn: the reference's partial distance."""
    pos, d, energy, charge, bz, steps, nsub = CASES["electron_interior"]
    dev = _device(amd, params, pos, d, energy, charge, bz, steps, nsub)
    assert dev[27, 0] == pytest.approx(8.7323805094658429, rel=1e-12)
    assert dev[0, 3:6] == pytest.approx([3.80852541539105, 0.0099999885096862, 0],
                                        rel=1e-12, abs=1e-15)
    assert np.linalg.norm(dev[28, 3:6] - [3.8085385437789383, 0, 0]) < 1e-5


def test_random_tracks_match_oracle(amd, params, oproblem):
    """Many tracks in one launch (one lane each) against the oracle."""
    rng = np.random.default_rng(42)
    n, nreq = 200, 6
    pos = rng.uniform(-9, 9, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    energy = 10 ** rng.uniform(-1, 2, size=n)
    charge = rng.choice([-1.0, 1.0], size=n)
    bz = rng.uniform(0.5, 4.0, size=n)
    steps = 10 ** rng.uniform(-3, 1, size=(n, nreq))
    tracks = np.column_stack([pos, d, [fc.momentum(e) for e in energy], charge,
                              np.zeros(n), np.zeros(n), bz])
    dev = amd.verify_field_propagate(params, tracks, steps)
    mismatched = 0
    for i in range(n):
        ora = ob.field_propagate(oproblem, pos[i], d[i], tracks[i, 6], charge[i],
                                 [0, 0, bz[i]], steps[i])
        same = (np.array_equal(dev[i, :, 1:3], ora[:, 1:3])
                and np.allclose(dev[i, :, 0], ora[:, 0], rtol=1e-10, atol=1e-15)
                and np.allclose(dev[i, :, 3:9], ora[:, 3:9], rtol=1e-10, atol=1e-12))
        if not same:
            mismatched += 1
            print(f"track {i}: device {dev[i, :, :3].tolist()} oracle {ora[:, :3].tolist()}")
    assert mismatched == 0
