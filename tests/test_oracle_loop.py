"""Whole-loop oracle anchors (CPU).

SimpleCompton (test/celeritas/global/Stepper.test.cc:194-225, seed
20220511 from GlobalTestBase.cc:100-103): 32 gammas of 100 MeV into the
two-boxes geometry with 64 track slots give 919 step iterations, 53.8125
steps per primary, a queue high-water mark of 6 at step 1 and the
emptying step 3 in the reference's default build.
"""

import numpy as np

from tests import oracle_bind as ob
from tests import problems


def test_simple_compton_anchor():
    prob = ob.Problem(problems.simple_compton())
    st = ob.State(prob, 64)
    active, queued = ob.run(st, problems.simple_compton_primaries(32))
    assert len(active) == 919
    assert active.sum() / 32 == 53.8125
    assert (int(np.argmax(queued)), int(queued.max())) == (1, 6)
    assert ob.emptying_step(active) == 3


def test_simple_compton_action_order():
    prob = ob.Problem(problems.simple_compton())
    labels = prob.action_labels()
    # Registration order of the reference (PhysicsParams.cc:97-143 then
    # CoreParams.cc:139-215)
    assert labels[:7] == ["pre-step", "msc-range", "eloss-range", "physics-discrete-select",
                          "physics-integral-rejected", "scat-klein-nishina", "physics-failure"]
    assert labels[7:] == ["along-step-neutral", "initialize-tracks", "geo-propagation-limit",
                          "geo-boundary", "tracking-cut", "extend-from-secondaries",
                          "extend-from-primaries"]
