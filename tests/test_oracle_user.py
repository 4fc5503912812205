"""Oracle user observers pinned to the reference's golden values (CPU only).

- Diagnostic.test.cc:110-131: SimpleCompton with 256 slots and 256 10 MeV
  photons, 32 steps: nonzero post-step action counts by particle and the
  steps-per-track histogram (max step bin 20, so 22 bins per particle).
- StepCollector.test.cc:179-226: the KN MC-truth example (event, track and
  step count with the pre-step volume, position and direction) after one
  step of 8 tracks and two steps of 4 tracks, one event per photon.
- StepCollector.test.cc:229-243: KN calorimeter, one photon, 64 steps,
  energy deposited in the "inner" box.

Each planted mistake in the oracle's observers (killed tracks left out of
the action tally, step bins one low, the pre-step position taken at the
step's end) fails one of these.
"""

import ctypes

import numpy as np
import pytest

import celeritas_amd as amd
from tests import oracle_bind as ob
from tests import problems as pb

PARTICLES = ["gamma", "electron"]


@pytest.fixture(autouse=True)
def _no_mutation():
    _mutate(0)
    yield
    _mutate(0)


def _mutate(m):
    f = ob.load().orc_user_set_mutation
    f.argtypes = [ctypes.c_int]
    f(m)


@pytest.fixture(scope="module")
def compton():
    return ob.Problem(pb.simple_compton())


def _gammas(n, pos, event_per_track):
    return amd.make_primaries(
        {"particle_id": 0, "energy": 10.0, "pos": pos, "dir": [1.0, 0, 0],
         "event_id": i if event_per_track else 0} for i in range(n))


def _run(state, primaries, num_steps):
    """SimpleLoopTestBase::run_impl (SimpleLoopTestBase.cc:34-54)."""
    state.insert_primaries(primaries)
    r = ob.step_once(state)
    while (r["queued"] > 0 or r["alive"] > 0) and num_steps > 1:
        num_steps -= 1
        r = ob.step_once(state)


def diagnostic_result(problem):
    labels = problem.action_labels()
    st = ob.State(problem, 256)
    st.enable_user(num_actions=len(labels), step_bins=20 + 2)
    _run(st, _gammas(256, [-22.0, 0, 0], False), 32)
    counts = st.action_counts(2)
    keys = {f"{labels[a]} {PARTICLES[p]}": int(counts[p, a])
            for p in range(2) for a in range(counts.shape[1]) if counts[p, a]}
    keys = dict(sorted(keys.items()))
    return keys, st.step_counts(2).ravel().tolist()


EXPECTED_KEYS = {"geo-boundary electron": 3780, "geo-boundary gamma": 525,
                 "scat-klein-nishina gamma": 3887}
EXPECTED_STEPS = [0, 0, 0, 87, 30, 10, 2, 0, 1, 0, 0, 3, 0, 0, 0,
                  0, 0, 0, 0, 0, 0, 1, 0, 0, 1840, 0, 0, 0, 0, 0,
                  0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0]


def test_simple_compton_diagnostics(compton):
    keys, steps = diagnostic_result(compton)
    assert keys == EXPECTED_KEYS
    assert steps == EXPECTED_STEPS


def mctruth(problem, num_tracks, num_steps):
    st = ob.State(problem, num_tracks)
    st.enable_user(collect=True)
    _run(st, _gammas(num_tracks, [0.0, 0, 0], True), num_steps)
    h = st.take_hits()
    # ExampleMctruth::sort: by (event, track, step)
    order = np.lexsort((h["track_step_count"], h["track_id"], h["event_id"]))
    return {k: v[order] for k, v in h.items()}


def test_kn_mctruth_single_step(compton):
    h = mctruth(compton, 8, 1)
    np.testing.assert_array_equal(h["event_id"], range(8))
    np.testing.assert_array_equal(h["track_id"], [0] * 8)
    np.testing.assert_array_equal(h["track_step_count"], [1] * 8)
    np.testing.assert_array_equal(h["pre.volume_id"], [1] * 8)
    np.testing.assert_array_equal(h["pre.pos"].ravel(), [0] * 24)
    np.testing.assert_array_equal(h["pre.dir"].ravel(), [1, 0, 0] * 8)


TWO_STEP_POS = [0, 0, 0, 2.6999255778482, 0, 0, 0, 0, 0, 3.5717683161497, 0, 0,
                0, 0, 0, 5, 0, 0, 0, 0, 0, 5, 0, 0]
TWO_STEP_DIR = [1, 0, 0, 0.45619379667222, 0.14402721708137, -0.87814769863479, 1, 0, 0,
                0.8985574206844, -0.27508545475671, -0.34193940152356, 1, 0, 0, 1, 0, 0,
                1, 0, 0, 1, 0, 0]


def test_kn_mctruth_two_step(compton):
    h = mctruth(compton, 4, 2)
    np.testing.assert_array_equal(h["event_id"], [0, 0, 1, 1, 2, 2, 3, 3])
    np.testing.assert_array_equal(h["track_id"], [0] * 8)
    np.testing.assert_array_equal(h["track_step_count"], [1, 2] * 4)
    np.testing.assert_array_equal(h["pre.volume_id"], [1, 1, 1, 1, 1, 2, 1, 2])
    # EXPECT_VEC_SOFT_EQ: relative 1e-12 (absolute near zero)
    np.testing.assert_allclose(h["pre.pos"].ravel(), TWO_STEP_POS, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(h["pre.dir"].ravel(), TWO_STEP_DIR, rtol=1e-12, atol=1e-14)


def test_kn_calo_single_track():
    prob = pb.simple_compton()
    prob["volume_to_detector"] = [-1, 0, -1]  # "inner"
    st = ob.State(ob.Problem(prob), 1)
    _run(st, _gammas(1, [0.0, 0, 0], True), 64)
    np.testing.assert_allclose(st.get_f64("calo_edep"), [0.00043564799352598], rtol=1e-12)


def test_planted_user_mistakes_fail(compton):
    _mutate(1)
    keys, _ = diagnostic_result(compton)
    assert keys != EXPECTED_KEYS
    _mutate(2)
    _, steps = diagnostic_result(compton)
    assert steps != EXPECTED_STEPS
    _mutate(3)
    h = mctruth(compton, 4, 2)
    assert not np.allclose(h["pre.pos"].ravel(), TWO_STEP_POS, rtol=1e-12, atol=1e-14)
